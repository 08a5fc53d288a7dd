// Host side of the surrogate network (tl_mlp.h): parameters, training loop
// and inference on the GPU, behind the C ABI of include/tl_surrogate.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "tl_mlp.h"
#include "../../include/tl_surrogate.h"

using namespace tl;

namespace {

thread_local std::string g_mlp_err;

#define MCHK(x)                                                                    \
  do {                                                                             \
    hipError_t e_ = (hipError_t)(x);                                               \
    if (e_ != hipSuccess)                                                          \
      throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_));    \
  } while (0)

template <class F>
int mguard(F&& f) {
  g_mlp_err.clear();
  try {
    return f();
  } catch (const std::exception& e) {
    g_mlp_err = e.what();
    return -1;
  } catch (...) {
    g_mlp_err = "unknown error";
    return -1;
  }
}

template <class T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  void ensure(size_t k) {
    if (k <= n) return;
    if (p) MCHK(hipFree(p));
    p = nullptr;
    MCHK(hipMalloc(&p, std::max<size_t>(k, 1) * sizeof(T)));
    n = k;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct Mlp {
  std::mutex mu;
  hipStream_t stream = nullptr;
  MlpLayout L;
  Dev<float> p, g, m, v;   // FP32 master parameters, gradients, Adam moments
  Dev<uint16_t> pb;        // BF16 copy used by the products
  Dev<float> mean, scale;  // feature standardisation
  int t = 0;               // Adam steps taken
  // workspaces for a batch of rows
  int cap = 0;
  Dev<uint16_t> h[kMlpLayers];  // h[0]: packed input, h[1..4]: hidden activations
  Dev<float> z[kMlpLayers - 1];
  Dev<float> out, y, xin, loss;
  Dev<uint16_t> dzA, dzB, dz5;

  ~Mlp() {
    for (auto* d : {&p, &g, &m, &v, &mean, &scale, &out, &y, &xin, &loss}) d->release();
    for (auto& d : h) d.release();
    for (auto& d : z) d.release();
    pb.release();
    dzA.release();
    dzB.release();
    dz5.release();
    if (stream) (void)hipStreamDestroy(stream);
  }

  void reserve(int rows) {
    if (rows <= cap) return;
    h[0].ensure((size_t)rows * kMlpInPad);
    for (int l = 1; l < kMlpLayers; l++) h[l].ensure((size_t)rows * kMlpHidden);
    for (int l = 0; l < kMlpLayers - 1; l++) z[l].ensure((size_t)rows * kMlpHidden);
    out.ensure((size_t)rows * kMlpOutPad);
    y.ensure(rows);
    xin.ensure((size_t)rows * kMlpIn);
    dzA.ensure((size_t)rows * kMlpHidden);
    dzB.ensure((size_t)rows * kMlpHidden);
    dz5.ensure((size_t)rows * kMlpOutPad);
    loss.ensure(1);
    cap = rows;
  }

  // forward pass over `rows` packed rows in h[0]; keeps the pre-activations
  // when `keep`
  void forward(int rows, bool keep) {
    for (int l = 0; l < kMlpLayers; l++) {
      GemmArgs a;
      std::memset(&a, 0, sizeof(a));
      a.A = h[l].p;
      a.lda = L.K[l];
      a.B = pb.p + L.w[l];
      a.ldb = L.N[l];
      a.M = rows;
      a.N = L.N[l];
      a.K = L.K[l];
      a.bias = p.p + L.b[l];
      if (l < kMlpLayers - 1) {
        a.epi = kEpiGelu;
        a.Z = keep ? z[l].p : nullptr;
        a.ldz = kMlpHidden;
        a.H = h[l + 1].p;
        a.ldh = kMlpHidden;
      } else {
        a.epi = kEpiLinear;
        a.C = out.p;
        a.ldc = kMlpOutPad;
      }
      MCHK(launch_gemm(a, stream));
    }
  }

  // forward pass over `rows` packed rows in h0 with caller-given buffers:
  // the hidden layers alternate between hA and hB, the output goes to out
  // (no pre-activations kept)
  void forward_into(int rows, const uint16_t* h0, uint16_t* hA, uint16_t* hB, float* o) {
    const uint16_t* in = h0;
    uint16_t* bufs[2] = {hA, hB};
    for (int l = 0; l < kMlpLayers; l++) {
      GemmArgs a;
      std::memset(&a, 0, sizeof(a));
      a.A = in;
      a.lda = L.K[l];
      a.B = pb.p + L.w[l];
      a.ldb = L.N[l];
      a.M = rows;
      a.N = L.N[l];
      a.K = L.K[l];
      a.bias = p.p + L.b[l];
      if (l < kMlpLayers - 1) {
        a.epi = kEpiGelu;
        a.H = bufs[l & 1];
        a.ldh = kMlpHidden;
        in = a.H;
      } else {
        a.epi = kEpiLinear;
        a.C = o;
        a.ldc = kMlpOutPad;
      }
      MCHK(launch_gemm(a, stream));
    }
  }

  // backward pass from the loss of targets y; gradients into g
  void backward(int rows) {
    MCHK(hipMemsetAsync(loss.p, 0, sizeof(float), stream));
    MCHK(launch_nll_grad(out.p, y.p, rows, dz5.p, loss.p, stream));
    uint16_t* dz = dz5.p;  // gradient at layer l's pre-activation
    for (int l = kMlpLayers - 1; l >= 0; l--) {
      const int Kl = L.K[l], Nl = L.N[l];
      GemmArgs w;  // dW_l = H_{l-1}^T dZ_l
      std::memset(&w, 0, sizeof(w));
      w.A = h[l].p;
      w.lda = Kl;
      w.transA = 1;
      w.B = dz;
      w.ldb = Nl;
      w.M = Kl;
      w.N = Nl;
      w.K = rows;
      w.epi = kEpiStore;
      w.C = g.p + L.w[l];
      w.ldc = Nl;
      MCHK(launch_gemm(w, stream));
      MCHK(launch_colsum(dz, rows, Nl, g.p + L.b[l], stream));
      if (l == 0) break;
      GemmArgs d;  // dZ_{l-1} = (dZ_l W_l^T) * GELU'(Z_{l-1})
      std::memset(&d, 0, sizeof(d));
      d.A = dz;
      d.lda = Nl;
      d.B = pb.p + L.w[l];
      d.ldb = Nl;
      d.transB = 1;
      d.M = rows;
      d.N = Kl;
      d.K = Nl;
      d.epi = kEpiDGelu;
      d.Z = z[l - 1].p;
      d.ldz = kMlpHidden;
      uint16_t* next = (dz == dzA.p) ? dzB.p : dzA.p;
      d.H = next;
      d.ldh = kMlpHidden;
      MCHK(launch_gemm(d, stream));
      dz = next;
    }
  }

  void pack(const float* xdev, int rows) {
    MCHK(launch_pack_input(xdev, rows, mean.p, scale.p, h[0].p, stream));
  }
};

// logical parameter order: W1 (14 x 256), b1, W2, b2, W3, b3, W4, b4,
// W5 (256 x 2), b5 (2)
size_t logical_count() {
  return (size_t)kMlpIn * kMlpHidden + kMlpHidden +
         3 * ((size_t)kMlpHidden * kMlpHidden + kMlpHidden) + (size_t)kMlpHidden * kMlpOut + kMlpOut;
}

void to_logical(const MlpLayout& L, const std::vector<float>& pad, float* out) {
  size_t o = 0;
  for (int l = 0; l < kMlpLayers; l++) {
    const int Kr = l == 0 ? kMlpIn : L.K[l];
    const int Nr = l == kMlpLayers - 1 ? kMlpOut : L.N[l];
    for (int k = 0; k < Kr; k++)
      for (int n = 0; n < Nr; n++) out[o++] = pad[L.w[l] + (size_t)k * L.N[l] + n];
    for (int n = 0; n < Nr; n++) out[o++] = pad[L.b[l] + n];
  }
}

void from_logical(const MlpLayout& L, const float* in, std::vector<float>& pad) {
  pad.assign(L.total, 0.0f);
  size_t o = 0;
  for (int l = 0; l < kMlpLayers; l++) {
    const int Kr = l == 0 ? kMlpIn : L.K[l];
    const int Nr = l == kMlpLayers - 1 ? kMlpOut : L.N[l];
    for (int k = 0; k < Kr; k++)
      for (int n = 0; n < Nr; n++) pad[L.w[l] + (size_t)k * L.N[l] + n] = in[o++];
    for (int n = 0; n < Nr; n++) pad[L.b[l] + n] = in[o++];
  }
}

void upload_params(Mlp& M, const std::vector<float>& pad) {
  MCHK(hipMemcpy(M.p.p, pad.data(), pad.size() * sizeof(float), hipMemcpyHostToDevice));
  MCHK(launch_to_bf16(M.p.p, pad.size(), M.pb.p, M.stream));
  MCHK(hipStreamSynchronize(M.stream));
}

Mlp* checked(void* h) {
  if (!h) throw std::invalid_argument("null network");
  return static_cast<Mlp*>(h);
}

}  // namespace

extern "C" {

void* tl_mlp_new(unsigned long long seed) {
  Mlp* M = nullptr;
  mguard([&] {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
      throw std::runtime_error("no HIP device available: the surrogate needs a gfx950 GPU");
    M = new Mlp();
    MCHK(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
    // the allocation-free inference's scratch on this device
    void* scratchBase = nullptr;
    MCHK(noalloc_scratch(&scratchBase));
    M->L = mlp_layout();
    const size_t n = M->L.total;
    M->p.ensure(n);
    M->g.ensure(n);
    M->m.ensure(n);
    M->v.ensure(n);
    M->pb.ensure(n);
    M->mean.ensure(kMlpIn);
    M->scale.ensure(kMlpIn);
    MCHK(hipMemset(M->m.p, 0, n * sizeof(float)));
    MCHK(hipMemset(M->v.p, 0, n * sizeof(float)));
    MCHK(hipMemset(M->g.p, 0, n * sizeof(float)));
    std::vector<float> ones(kMlpIn, 1.0f), zeros(kMlpIn, 0.0f);
    MCHK(hipMemcpy(M->mean.p, zeros.data(), kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(M->scale.p, ones.data(), kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    // He initialisation of the real weights; padding and biases zero
    std::mt19937_64 rng(seed ? seed : 1);
    std::vector<float> pad(n, 0.0f);
    for (int l = 0; l < kMlpLayers; l++) {
      const int Kr = l == 0 ? kMlpIn : M->L.K[l];
      const int Nr = l == kMlpLayers - 1 ? kMlpOut : M->L.N[l];
      const float sd = std::sqrt(2.0f / (float)Kr) * (l == kMlpLayers - 1 ? 0.1f : 1.0f);
      std::normal_distribution<float> nd(0.0f, sd);
      for (int k = 0; k < Kr; k++)
        for (int c = 0; c < Nr; c++) pad[M->L.w[l] + (size_t)k * M->L.N[l] + c] = nd(rng);
    }
    upload_params(*M, pad);
    return 0;
  });
  return M;
}

void tl_mlp_delete(void* h) { delete static_cast<Mlp*>(h); }

long tl_mlp_param_count(void) { return (long)logical_count(); }

int tl_mlp_set_scaling(void* h, const float* mean, const float* scale) {
  return mguard([&] {
    Mlp* M = checked(h);
    std::lock_guard<std::mutex> lock(M->mu);
    for (int f = 0; f < kMlpIn; f++)
      if (!(scale[f] > 0.0f) || !std::isfinite(mean[f])) throw std::invalid_argument("bad scaling");
    MCHK(hipMemcpy(M->mean.p, mean, kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(M->scale.p, scale, kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    return 0;
  });
}

int tl_mlp_get_params(void* h, float* out) {
  return mguard([&] {
    Mlp* M = checked(h);
    std::lock_guard<std::mutex> lock(M->mu);
    std::vector<float> pad(M->L.total);
    MCHK(hipMemcpy(pad.data(), M->p.p, pad.size() * sizeof(float), hipMemcpyDeviceToHost));
    to_logical(M->L, pad, out);
    return 0;
  });
}

int tl_mlp_set_params(void* h, const float* in) {
  return mguard([&] {
    Mlp* M = checked(h);
    std::lock_guard<std::mutex> lock(M->mu);
    std::vector<float> pad;
    from_logical(M->L, in, pad);
    upload_params(*M, pad);
    return 0;
  });
}

int tl_mlp_predict(void* h, int n, const float* x, float* out) {
  return mguard([&] {
    Mlp* M = checked(h);
    if (n < 0 || (n > 0 && (!x || !out))) throw std::invalid_argument("bad arguments");
    std::lock_guard<std::mutex> lock(M->mu);
    const int chunk = 1 << 16;
    M->reserve(std::min(((n + 63) / 64) * 64, chunk));
    std::vector<float> o;
    for (int r0 = 0; r0 < n; r0 += chunk) {
      const int rows = std::min(chunk, n - r0);
      const int padded = ((rows + 15) / 16) * 16;
      MCHK(hipMemcpyAsync(M->xin.p, x + (size_t)r0 * kMlpIn, (size_t)rows * kMlpIn * sizeof(float),
                          hipMemcpyHostToDevice, M->stream));
      MCHK(launch_pack_input(M->xin.p, rows, M->mean.p, M->scale.p, M->h[0].p, M->stream));
      if (padded > rows)
        MCHK(hipMemsetAsync(M->h[0].p + (size_t)rows * kMlpInPad, 0,
                            (size_t)(padded - rows) * kMlpInPad * sizeof(uint16_t), M->stream));
      M->forward(padded, false);
      o.resize((size_t)rows * kMlpOutPad);
      MCHK(hipMemcpyAsync(o.data(), M->out.p, o.size() * sizeof(float), hipMemcpyDeviceToHost,
                          M->stream));
      MCHK(hipStreamSynchronize(M->stream));
      for (int i = 0; i < rows; i++) {
        out[2 * ((size_t)r0 + i)] = o[(size_t)i * kMlpOutPad];
        out[2 * ((size_t)r0 + i) + 1] = o[(size_t)i * kMlpOutPad + 1];
      }
    }
    return 0;
  });
}

int tl_mlp_predict_noalloc(void* h, int n, const float* x, float* out) {
  // one inference at a time per process: every call uses this GPU's single
  // scratch array (tl_mlp.hip)
  static std::mutex noallocMu;
  return mguard([&] {
    Mlp* M = checked(h);
    if (n < 0 || (size_t)n > kNoAllocRows)
      throw std::invalid_argument("tl_mlp_predict_noalloc: n = " + std::to_string(n) +
                                  " rows; at most " + std::to_string(kNoAllocRows) +
                                  " fit its scratch array");
    if (n > 0 && (!x || !out)) throw std::invalid_argument("bad arguments");
    std::lock_guard<std::mutex> lock(noallocMu);
    std::lock_guard<std::mutex> lockNet(M->mu);
    if (n == 0) return 0;
    void* base = nullptr;
    MCHK(noalloc_scratch(&base));
    unsigned char* b = static_cast<unsigned char*>(base);
    uint16_t* h0 = reinterpret_cast<uint16_t*>(b + kNaH0);
    uint16_t* hA = reinterpret_cast<uint16_t*>(b + kNaHA);
    uint16_t* hB = reinterpret_cast<uint16_t*>(b + kNaHB);
    float* o = reinterpret_cast<float*>(b + kNaOut);
    const int padded = ((n + 15) / 16) * 16;
    MCHK(launch_pack_input_padded(x, n, padded, M->mean.p, M->scale.p, h0, M->stream));
    M->forward_into(padded, h0, hA, hB, o);
    MCHK(launch_unpack_output(o, n, out, M->stream));
    MCHK(hipStreamSynchronize(M->stream));
    return 0;
  });
}

int tl_mlp_grads(void* h, int n, const float* x, const float* y, float* grads, float* loss) {
  return mguard([&] {
    Mlp* M = checked(h);
    if (n <= 0 || n % 64 || !x || !y) throw std::invalid_argument("n must be a positive multiple of 64");
    std::lock_guard<std::mutex> lock(M->mu);
    M->reserve(n);
    MCHK(hipMemcpyAsync(M->xin.p, x, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice,
                        M->stream));
    MCHK(hipMemcpyAsync(M->y.p, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice, M->stream));
    M->pack(M->xin.p, n);
    M->forward(n, true);
    M->backward(n);
    std::vector<float> pad(M->L.total);
    float l = 0.0f;
    MCHK(hipMemcpyAsync(pad.data(), M->g.p, pad.size() * sizeof(float), hipMemcpyDeviceToHost,
                        M->stream));
    MCHK(hipMemcpyAsync(&l, M->loss.p, sizeof(float), hipMemcpyDeviceToHost, M->stream));
    MCHK(hipStreamSynchronize(M->stream));
    if (grads) to_logical(M->L, pad, grads);
    if (loss) *loss = l;
    return 0;
  });
}

int tl_mlp_train(void* h, int n, const float* x, const float* y, int steps, int batch, float lr,
                 unsigned long long seed, float* losses) {
  return mguard([&] {
    Mlp* M = checked(h);
    if (n <= 0 || !x || !y || steps < 0 || batch <= 0 || batch % 64 || batch > n)
      throw std::invalid_argument("bad training arguments (batch: a multiple of 64, <= n)");
    std::lock_guard<std::mutex> lock(M->mu);
    M->reserve(batch);
    std::mt19937_64 rng(seed ? seed : 1);
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    size_t pos = n;
    std::vector<float> bx((size_t)batch * kMlpIn), by(batch);
    std::vector<float> hx((size_t)n * kMlpIn);
    std::memcpy(hx.data(), x, hx.size() * sizeof(float));
    for (int s = 0; s < steps; s++) {
      if (pos + batch > (size_t)n) {
        std::shuffle(perm.begin(), perm.end(), rng);
        pos = 0;
      }
      for (int i = 0; i < batch; i++) {
        const int r = perm[pos + i];
        std::memcpy(&bx[(size_t)i * kMlpIn], &hx[(size_t)r * kMlpIn], kMlpIn * sizeof(float));
        by[i] = y[r];
      }
      pos += batch;
      MCHK(hipMemcpyAsync(M->xin.p, bx.data(), bx.size() * sizeof(float), hipMemcpyHostToDevice,
                          M->stream));
      MCHK(hipMemcpyAsync(M->y.p, by.data(), by.size() * sizeof(float), hipMemcpyHostToDevice,
                          M->stream));
      M->pack(M->xin.p, batch);
      M->forward(batch, true);
      M->backward(batch);
      M->t++;
      MCHK(launch_adam(M->p.p, M->g.p, M->m.p, M->v.p, M->pb.p, M->L.total, lr, 0.9f, 0.999f,
                       1e-8f, M->t, M->stream));
      if (losses) {
        MCHK(hipMemcpyAsync(&losses[s], M->loss.p, sizeof(float), hipMemcpyDeviceToHost,
                            M->stream));
      }
    }
    MCHK(hipStreamSynchronize(M->stream));
    return 0;
  });
}

const char* tl_mlp_last_error(void) { return g_mlp_err.c_str(); }

}  // extern "C"
