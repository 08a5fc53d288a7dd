// Surrogate network: a 14-256-256-256-256-2 MLP with GELU, BF16 weights and
// activations on the matrix cores, FP32 accumulation and FP32 master weights.
//
// Layer l computes Z_l = H_{l-1} W_l + b_l and H_l = GELU(Z_l) (the last layer
// is linear).  W_l is stored K x N row-major.  The input is padded from 14 to
// 16 features and the output from 2 to 16 columns; the padding stays zero.
// Outputs: mu = log(A_finite / A_point) and s = log(sigma).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tl {

constexpr int kMlpIn = 14;     // features
constexpr int kMlpInPad = 16;
constexpr int kMlpHidden = 256;
constexpr int kMlpOut = 2;
constexpr int kMlpOutPad = 16;
constexpr int kMlpLayers = 5;

// Offsets of the layers' weights and biases in the flat (padded) parameter
// vector; dims of layer l: K_l x N_l.
struct MlpLayout {
  int K[kMlpLayers], N[kMlpLayers];
  size_t w[kMlpLayers], b[kMlpLayers];
  size_t total;
};
MlpLayout mlp_layout();

// One matrix product on the matrix cores: C = A B with A logically M x K and
// B logically K x N (both BF16, either stored transposed), FP32 accumulation,
// and one of these epilogues per element (i, n):
enum MlpEpilogue {
  kEpiGelu = 0,    // z = acc + bias[n]; Z[i][n] = z (if Z); H[i][n] = bf16(GELU(z))
  kEpiLinear = 1,  // C[i][n] = acc + bias[n]
  kEpiDGelu = 2,   // H[i][n] = bf16(acc * GELU'(Z[i][n]))
  kEpiStore = 3,   // C[i][n] = acc
};

struct GemmArgs {
  const uint16_t* A; int lda; int transA;  // A[i][k] = transA ? A[k lda + i] : A[i lda + k]
  const uint16_t* B; int ldb; int transB;  // B[k][n] = transB ? B[n ldb + k] : B[k ldb + n]
  int M, N, K;                             // multiples of 16
  int epi;
  const float* bias;
  float* Z; int ldz;
  uint16_t* H; int ldh;
  float* C; int ldc;
};

int launch_gemm(const GemmArgs& g, hipStream_t s);
// Gaussian negative log-likelihood of y under (mu, s) = out[m][0..1], mean over
// the M rows; writes dL/d(out) into dz (BF16, M x 16) and adds the batch loss
// to *loss.
int launch_nll_grad(const float* out, const float* y, int M, uint16_t* dz, float* loss,
                    hipStream_t s);
// db[n] = sum over rows of dz (M x N, BF16).
int launch_colsum(const uint16_t* dz, int M, int N, float* db, hipStream_t s);
// Adam step on n parameters; writes the BF16 copy of the updated weights.
int launch_adam(float* p, const float* g, float* m, float* v, uint16_t* pb, size_t n, float lr,
                float b1, float b2, float eps, int t, hipStream_t s);
// Rows of kMlpIn features to BF16 rows of 16: (x - mean) / scale, zero padded.
int launch_pack_input(const float* x, int n, const float* mean, const float* scale, uint16_t* xb,
                      hipStream_t s);
// FP32 to BF16 (round to nearest even).
int launch_to_bf16(const float* x, size_t n, uint16_t* y, hipStream_t s);

// ---- inference without allocation ----------------------------------------
// Largest batch of tl_mlp_predict_noalloc, and the byte layout of its
// scratch array (tl_mlp.hip): packed input, two hidden layers used in turn,
// padded output.
constexpr size_t kNoAllocRows = 4096;
constexpr size_t kNaH0 = 0;
constexpr size_t kNaHA = kNaH0 + kNoAllocRows * kMlpInPad * sizeof(uint16_t);
constexpr size_t kNaHB = kNaHA + kNoAllocRows * kMlpHidden * sizeof(uint16_t);
constexpr size_t kNaOut = kNaHB + kNoAllocRows * kMlpHidden * sizeof(uint16_t);
constexpr size_t kNoAllocScratchBytes = kNaOut + kNoAllocRows * kMlpOutPad * sizeof(float);
static_assert(kNaHA % 256 == 0 && kNaHB % 256 == 0 && kNaOut % 256 == 0, "aligned parts");
// Device address of the scratch array on the current device (resolved with
// hipGetSymbolAddress once per device); 0 or a hipError_t.
int noalloc_scratch(void** out);
int launch_pack_input_padded(const float* x, int n, int nPad, const float* mean,
                             const float* scale, uint16_t* xb, hipStream_t s);
int launch_unpack_output(const float* c, int n, float* out, hipStream_t s);

}  // namespace tl
