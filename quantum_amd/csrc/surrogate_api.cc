// C-ABI of the MLP surrogate (include/tfqx.h, "surrogate" section).
#include <string.h>

#include <string>

#include <hip/hip_runtime.h>

#include "../../include/tfqx.h"
#include "surrogate.h"

using namespace tfqx;

namespace {
hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }
int Bad(char* err, int64_t len, const char* msg) {
  if (err && len > 0) {
    strncpy(err, msg, size_t(len - 1));
    err[len - 1] = '\0';
  }
  return TFQX_INVALID_ARGUMENT;
}
bool Ok(int P, int H, int B) { return P > 0 && H > 0 && B > 0; }
}  // namespace

extern "C" {

int tfqx_mlp_layout(int32_t in_dim, int32_t hidden, int32_t max_batch, int64_t* out, int64_t n) {
  if (!Ok(in_dim, hidden, max_batch) || !out) return TFQX_INVALID_ARGUMENT;
  const MlpLayout L = MakeMlpLayout(in_dim, hidden, max_batch);
  const int64_t v[6] = {int64_t(L.bytes), int64_t(L.nparams), int64_t(L.w), int64_t(L.g),
                        int64_t(L.loss), int64_t(L.m)};
  for (int64_t i = 0; i < n && i < 6; ++i) out[i] = v[i];
  return TFQX_OK;
}

int tfqx_mlp_set_params(void* ws, int32_t in_dim, int32_t hidden, int32_t max_batch,
                        const float* host_params, void* stream, char* err, int64_t len) {
  if (!Ok(in_dim, hidden, max_batch) || !ws || !host_params) return Bad(err, len, "bad arguments");
  return MlpSetParams(MakeMlpLayout(in_dim, hidden, max_batch), ws, host_params, S(stream), err, len);
}

int tfqx_mlp_predict(void* ws, int32_t in_dim, int32_t hidden, int32_t max_batch, const float* x,
                     int32_t batch, float* energy, float* grad, void* stream, char* err,
                     int64_t len) {
  if (!Ok(in_dim, hidden, max_batch) || !ws || !x || !energy || batch < 1)
    return Bad(err, len, "bad arguments");
  return MlpPredict(MakeMlpLayout(in_dim, hidden, max_batch), ws, x, batch, energy, grad, S(stream),
                    err, len);
}

int tfqx_mlp_loss_grad(void* ws, int32_t in_dim, int32_t hidden, int32_t max_batch, const float* x,
                       const float* energy, const float* grad, int32_t batch, float lam,
                       void* stream, char* err, int64_t len) {
  if (!Ok(in_dim, hidden, max_batch) || !ws || !x || !energy || !grad || batch < 1)
    return Bad(err, len, "bad arguments");
  return MlpLossGrad(MakeMlpLayout(in_dim, hidden, max_batch), ws, x, energy, grad, batch, lam,
                     S(stream), err, len);
}

int tfqx_mlp_train_step(void* ws, int32_t in_dim, int32_t hidden, int32_t max_batch, const float* x,
                        const float* energy, const float* grad, int32_t batch, float lam, float lr,
                        int32_t step, void* stream, char* err, int64_t len) {
  if (!Ok(in_dim, hidden, max_batch) || !ws || !x || !energy || !grad || batch < 1 || step < 1)
    return Bad(err, len, "bad arguments");
  return MlpTrainStep(MakeMlpLayout(in_dim, hidden, max_batch), ws, x, energy, grad, batch, lam, lr,
                      step, S(stream), err, len);
}

int tfqx_gemm_bf16(int32_t M, int32_t N, int32_t K, const void* A, int32_t lda, int32_t transA,
                   const void* B, int32_t ldb, int32_t transB, float* C, int32_t ldc, float alpha,
                   float beta, void* stream, char* err, int64_t len) {
  if (M < 0 || N < 0 || K < 0 || !A || !B || !C) return Bad(err, len, "bad arguments");
  LaunchGemmBf16(M, N, K, A, lda, transA, B, ldb, transB, C, ldc, alpha, beta, S(stream));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    std::string m = std::string("gemm: ") + hipGetErrorString(e);
    Bad(err, len, m.c_str());
    return TFQX_DEVICE_ERROR;
  }
  return TFQX_OK;
}

}  // extern "C"
