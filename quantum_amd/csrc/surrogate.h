// MLP surrogate (in -> H -> H -> H -> 1, SiLU) trained with a Sobolev loss on
// BF16 MFMA GEMMs; see surrogate_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace tfqx {

struct MlpLayout {
  int P, H, maxB;
  // Offsets (floats) into the flat parameter vector.
  size_t oW1, ob1, oW2, ob2, oW3, ob3, ow4, ob4, nparams;
  // Byte offsets into the workspace.
  size_t w, g, m, v, wb1, wb2, wb3, xb;
  // Transposed BF16 weights ([K x H] row-major) for the fused small-batch
  // predict's backward products.
  size_t wt1, wt2, wt3;
  size_t z[3], a[3], h[3], hb[3], u[3], ub[3], vv[3], t1[3], t1b[3], t2[3], t2b[3], dag[3];
  size_t E, gin, dE, dg, dgb, loss;
  size_t bytes;
};

MlpLayout MakeMlpLayout(int P, int H, int maxB);
void LaunchGemmBf16(int M, int N, int K, const void* A, int lda, int transA, const void* B, int ldb,
                    int transB, float* C, int ldc, float alpha, float beta, hipStream_t stream);
int MlpSetParams(const MlpLayout& L, void* ws, const float* host_params, hipStream_t s, char* err,
                 int64_t len);
int MlpPredict(const MlpLayout& L, void* ws, const float* X, int B, float* E, float* G, hipStream_t s,
               char* err, int64_t len);
int MlpLossGrad(const MlpLayout& L, void* ws, const float* X, const float* Et, const float* Gt, int B,
                float lam, hipStream_t s, char* err, int64_t len);
int MlpTrainStep(const MlpLayout& L, void* ws, const float* X, const float* Et, const float* Gt, int B,
                 float lam, float lr, int step, hipStream_t s, char* err, int64_t len);

}  // namespace tfqx
