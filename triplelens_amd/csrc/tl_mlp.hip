// Kernels of the surrogate network (tl_mlp.h): one BF16 matrix-core product
// with fused epilogues for the forward and backward passes, the output
// layer's loss gradient, bias gradients, Adam, and input packing.
//
// The products use v_mfma_f32_16x16x16_bf16.  Its operands, per lane l of the
// wavefront: A row (l % 16) with k = 4 (l / 16) + j, B column (l % 16) with the
// same k, j = 0..3; the accumulator holds column (l % 16), rows 4 (l / 16) + j.
// A workgroup of four wavefronts computes a 64 x 64 tile of C, each wavefront
// a 32 x 32 quarter as 2 x 2 products of 16 x 16.
#include <hip/hip_runtime.h>
#include <math.h>

#include <mutex>

#include "tl_mlp.h"

namespace tl {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

MlpLayout mlp_layout() {
  MlpLayout L;
  const int K[kMlpLayers] = {kMlpInPad, kMlpHidden, kMlpHidden, kMlpHidden, kMlpHidden};
  const int N[kMlpLayers] = {kMlpHidden, kMlpHidden, kMlpHidden, kMlpHidden, kMlpOutPad};
  size_t off = 0;
  for (int l = 0; l < kMlpLayers; l++) {
    L.K[l] = K[l];
    L.N[l] = N[l];
    L.w[l] = off;
    off += (size_t)K[l] * N[l];
    L.b[l] = off;
    off += (size_t)N[l];
  }
  L.total = off;
  return L;
}

__device__ __forceinline__ uint16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ __forceinline__ uint16_t ldA(const GemmArgs& g, int i, int k) {
  if (i >= g.M) return 0;
  return g.transA ? g.A[(size_t)k * g.lda + i] : g.A[(size_t)i * g.lda + k];
}
__device__ __forceinline__ uint16_t ldB(const GemmArgs& g, int k, int n) {
  if (n >= g.N) return 0;
  return g.transB ? g.B[(size_t)n * g.ldb + k] : g.B[(size_t)k * g.ldb + n];
}

__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
  const int ln = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row0 = blockIdx.y * 64 + (w >> 1) * 32;
  const int col0 = blockIdx.x * 64 + (w & 1) * 32;
  const int lr = ln & 15, lk = (ln >> 4) * 4;
  f32x4 acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    bf16x4 fa[2], fb[2];
    for (int t = 0; t < 2; t++) {
      const int i = row0 + 16 * t + lr, n = col0 + 16 * t + lr;
      for (int j = 0; j < 4; j++) {
        fa[t][j] = (short)ldA(g, i, k0 + lk + j);
        fb[t][j] = (short)ldB(g, k0 + lk + j, n);
      }
    }
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[a], fb[b], acc[a][b], 0, 0, 0);
  }
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      const int n = col0 + 16 * b + lr;
      if (n >= g.N) continue;
      for (int j = 0; j < 4; j++) {
        const int i = row0 + 16 * a + lk + j;
        if (i >= g.M) continue;
        const float v = acc[a][b][j];
        switch (g.epi) {
          case kEpiGelu: {
            const float z = v + g.bias[n];
            if (g.Z) g.Z[(size_t)i * g.ldz + n] = z;
            g.H[(size_t)i * g.ldh + n] = f2bf(gelu(z));
            break;
          }
          case kEpiLinear:
            g.C[(size_t)i * g.ldc + n] = v + g.bias[n];
            break;
          case kEpiDGelu:
            g.H[(size_t)i * g.ldh + n] = f2bf(v * gelu_d(g.Z[(size_t)i * g.ldz + n]));
            break;
          default:
            g.C[(size_t)i * g.ldc + n] = v;
        }
      }
    }
}

int launch_gemm(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.K <= 0 || (g.M | g.N | g.K) % 16) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(gemm_kernel, dim3((g.N + 63) / 64, (g.M + 63) / 64), dim3(256), 0, s, g);
  return (int)hipGetLastError();
}

// Gaussian NLL 0.5 ((y - mu) e^-s)^2 + s, mean over rows; s clamped to
// [-12, 12] (its gradient is zero outside).
__global__ void nll_grad_kernel(const float* out, const float* y, int M, uint16_t* dz,
                                float* loss) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (m < M) {
    const float mu = out[(size_t)m * kMlpOutPad], s0 = out[(size_t)m * kMlpOutPad + 1];
    const float s = fminf(12.0f, fmaxf(-12.0f, s0));
    const float r = y[m] - mu, e = expf(-2.0f * s);
    l = 0.5f * r * r * e + s;
    const float inv = 1.0f / (float)M;
    const float dmu = -r * e * inv;
    const float ds = (s0 == s) ? (1.0f - r * r * e) * inv : 0.0f;
    for (int n = 0; n < kMlpOutPad; n++)
      dz[(size_t)m * kMlpOutPad + n] = n == 0 ? f2bf(dmu) : (n == 1 ? f2bf(ds) : (uint16_t)0);
  }
  // wave sum, one atomic per wave
  for (int o = 32; o >= 1; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l / (float)M);
}

int launch_nll_grad(const float* out, const float* y, int M, uint16_t* dz, float* loss,
                    hipStream_t s) {
  hipLaunchKernelGGL(nll_grad_kernel, dim3((M + 255) / 256), dim3(256), 0, s, out, y, M, dz, loss);
  return (int)hipGetLastError();
}

__global__ void colsum_kernel(const uint16_t* dz, int M, int N, float* db) {
  const int n = blockIdx.x;
  float s = 0.0f;
  for (int m = threadIdx.x; m < M; m += blockDim.x) s += bf2f(dz[(size_t)m * N + n]);
  __shared__ float part[256];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[n] = part[0];
}

int launch_colsum(const uint16_t* dz, int M, int N, float* db, hipStream_t s) {
  hipLaunchKernelGGL(colsum_kernel, dim3(N), dim3(256), 0, s, dz, M, N, db);
  return (int)hipGetLastError();
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, uint16_t* pb, size_t n,
                            float lr, float b1, float b2, float eps, float c1, float c2) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float w = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = w;
    pb[i] = f2bf(w);
  }
}

int launch_adam(float* p, const float* g, float* m, float* v, uint16_t* pb, size_t n, float lr,
                float b1, float b2, float eps, int t, hipStream_t s) {
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  hipLaunchKernelGGL(adam_kernel, dim3(256), dim3(256), 0, s, p, g, m, v, pb, n, lr, b1, b2, eps,
                     c1, c2);
  return (int)hipGetLastError();
}

__global__ void pack_input_kernel(const float* x, int n, const float* mean, const float* scale,
                                  uint16_t* xb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int f = 0; f < kMlpInPad; f++)
    xb[(size_t)i * kMlpInPad + f] =
        f < kMlpIn ? f2bf((x[(size_t)i * kMlpIn + f] - mean[f]) / scale[f]) : (uint16_t)0;
}

int launch_pack_input(const float* x, int n, const float* mean, const float* scale, uint16_t* xb,
                      hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(pack_input_kernel, dim3((n + 255) / 256), dim3(256), 0, s, x, n, mean, scale,
                     xb);
  return (int)hipGetLastError();
}

__global__ void to_bf16_kernel(const float* x, size_t n, uint16_t* y) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

int launch_to_bf16(const float* x, size_t n, uint16_t* y, hipStream_t s) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(256), dim3(256), 0, s, x, n, y);
  return (int)hipGetLastError();
}

// ---- inference without allocation (tl_mlp_predict_noalloc) ---------------
// The activations of one inference of at most kNoAllocRows rows live in this
// file-scope array: the packed input, two hidden layers used in turn, and the
// padded output (byte offsets kNa*).  Every GPU the library runs on holds
// its own copy until the process ends.  Host code reaches it only through
// the address noalloc_scratch() resolves per device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

int noalloc_scratch(void** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    addr[dev] = p;
  }
  *out = addr[dev];
  return 0;
}

// Rows [0, n) packed from x as pack_input_kernel does, rows [n, nPad) zero.
__global__ void pack_input_padded_kernel(const float* x, int n, int nPad, const float* mean,
                                         const float* scale, uint16_t* xb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nPad) return;
  for (int f = 0; f < kMlpInPad; f++)
    xb[(size_t)i * kMlpInPad + f] = (i < n && f < kMlpIn)
                                        ? f2bf((x[(size_t)i * kMlpIn + f] - mean[f]) / scale[f])
                                        : (uint16_t)0;
}

int launch_pack_input_padded(const float* x, int n, int nPad, const float* mean,
                             const float* scale, uint16_t* xb, hipStream_t s) {
  if (nPad <= 0) return 0;
  hipLaunchKernelGGL(pack_input_padded_kernel, dim3((nPad + 255) / 256), dim3(256), 0, s, x, n,
                     nPad, mean, scale, xb);
  return (int)hipGetLastError();
}

// (mu, s) of rows [0, n) from the padded output (n x kMlpOutPad) into n x 2.
__global__ void unpack_output_kernel(const float* c, int n, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[2 * (size_t)i] = c[(size_t)i * kMlpOutPad];
  out[2 * (size_t)i + 1] = c[(size_t)i * kMlpOutPad + 1];
}

int launch_unpack_output(const float* c, int n, float* out, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(unpack_output_kernel, dim3((n + 255) / 256), dim3(256), 0, s, c, n, out);
  return (int)hipGetLastError();
}

}  // namespace tl
