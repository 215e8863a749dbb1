// MLP surrogate for VQE on gfx950: energy E(x) and its input gradient dE/dx,
// trained with a Sobolev loss  mean (E-E*)^2 + lam * mean ||dE/dx - g*||^2.
//
// Network (in -> H -> H -> H -> 1, SiLU): weights and Adam state in FP32,
// BF16 copies feed the GEMMs, which run on v_mfma_f32_32x32x16_bf16 with FP32
// accumulation.  Training needs the double backward of the input-gradient
// branch; it is written out by hand below (see TrainStep).
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "surrogate.h"

namespace tfqx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ __bf16 to_bf16(float x) { return __float2bfloat16(x); }
__device__ __forceinline__ float from_bf16(__bf16 x) { return __bfloat162float(x); }

// ------------------------------------------------------------------ GEMM
// C[MxN] = alpha * op(A)[MxK] * op(B)[KxN] + beta * C, BF16 in, FP32 out.
// op(A)(m,k) = transA ? A[k*lda+m] : A[m*lda+k]
// op(B)(k,n) = transB ? B[n*ldb+k] : B[k*ldb+n]
// Workgroup: 4 wave64 computing a 64x64 tile as 2x2 waves of 32x32; K in
// steps of 32 through LDS (A stored [m][k], B stored [n][k], k contiguous so a
// lane's 8 k-values are one 16-byte LDS read).
constexpr int kGT = 64;   // output tile
constexpr int kGK = 32;   // K step
constexpr int kGThreads = 256;

// Stage one 64x32 operand tile into LDS as T[row][k] (k contiguous).  Each of
// the 256 threads moves one run of 8 elements: along k when the operand is
// k-contiguous in memory (kmajor), else along rows (then scattered into LDS).
// Full, 16-byte aligned runs use one vector load; edges fall back to scalars.
__device__ __forceinline__ void stage_tile(__bf16 (*T)[kGK + 8], const __bf16* __restrict__ X, int ld,
                                           bool kmajor, int r0, int rows, int k0, int K, int tid) {
  if (kmajor) {
    const int r = tid >> 2, kc = (tid & 3) * 8;
    const int gr = r0 + r, gk = k0 + kc;
    bf16x8 v;
    const __bf16* src = X + size_t(gr) * ld + gk;
    if (gr < rows && gk + 8 <= K && ((reinterpret_cast<uintptr_t>(src) & 15) == 0)) {
      v = *reinterpret_cast<const bf16x8*>(src);
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = (gr < rows && gk + i < K) ? src[i] : to_bf16(0.f);
    }
    *reinterpret_cast<bf16x8*>(&T[r][kc]) = v;
  } else {
    const int kk = tid >> 3, rc = (tid & 7) * 8;
    const int gk = k0 + kk, gr = r0 + rc;
    bf16x8 v;
    const __bf16* src = X + size_t(gk) * ld + gr;
    if (gk < K && gr + 8 <= rows && ((reinterpret_cast<uintptr_t>(src) & 15) == 0)) {
      v = *reinterpret_cast<const bf16x8*>(src);
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = (gk < K && gr + i < rows) ? src[i] : to_bf16(0.f);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) T[rc + i][kk] = v[i];
  }
}

__global__ __launch_bounds__(kGThreads) void gemm_bf16_kernel(
    int M, int N, int K, const __bf16* __restrict__ A, int lda, int transA,
    const __bf16* __restrict__ B, int ldb, int transB, float* __restrict__ C, int ldc,
    float alpha, float beta) {
  __shared__ __align__(16) __bf16 As[kGT][kGK + 8];
  __shared__ __align__(16) __bf16 Bs[kGT][kGK + 8];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
  f32x16 acc = {};
  for (int k0 = 0; k0 < K; k0 += kGK) {
    // op(A)(m,k): k-contiguous unless transposed; op(B)(k,n): k-contiguous when transposed.
    stage_tile(As, A, lda, !transA, m0, M, k0, K, tid);
    stage_tile(Bs, B, ldb, transB != 0, n0, N, k0, K, tid);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kGK; ks += 16) {
      const int kb = ks + 8 * (lane >> 5);
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm + (lane & 31)][kb]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn + (lane & 31)][kb]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // D layout: column = lane % 32, row = 8*(r/4) + 4*(lane/32) + r%4.
  const int col = n0 + wn + (lane & 31);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (row < M && col < N) {
      float v = alpha * acc[r];
      if (beta != 0.f) v += beta * C[size_t(row) * ldc + col];
      C[size_t(row) * ldc + col] = v;
    }
  }
}

// ------------------------------------------------------------------ elementwise
__device__ __forceinline__ float sigm(float a) { return 1.f / (1.f + __expf(-a)); }

// BF16 copies of W1..W3 ([H x K] row-major), as they are and transposed.
struct RefreshArgs {
  const float* W[3];
  __bf16* wb[3];
  __bf16* wt[3];
  int cols[3];
  size_t end[3];  // cumulative element counts
};
__global__ void refresh_bf16_kernel(RefreshArgs r) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < r.end[2];
       i += size_t(gridDim.x) * blockDim.x) {
    const int l = i < r.end[0] ? 0 : i < r.end[1] ? 1 : 2;
    const size_t j = i - (l == 0 ? 0 : r.end[l - 1]);
    const size_t cols = size_t(r.cols[l]), row = j / cols, col = j % cols;
    const size_t rows = (r.end[l] - (l == 0 ? 0 : r.end[l - 1])) / cols;
    const __bf16 v = to_bf16(r.W[l][j]);
    r.wb[l][j] = v;
    r.wt[l][col * rows + row] = v;
  }
}

// ------------------------------------------------------------------ fused small-batch predict
// Energies and input gradients of up to 32 rows in one launch: one workgroup
// of 16 wavefronts, wavefront w owning output columns 32w..32w+31 of every
// product (a 32 x 32 tile on v_mfma_f32_32x32x16_bf16, rows = batch).
// Activations pass between layers through LDS as BF16 (the rounding points of
// the layer-by-layer path, whose products run K in the same order); each
// thread keeps silu'(a) of its 16 (row, column) positions in registers for
// the backward products, which touch exactly those positions again.  The
// backward products read the transposed BF16 weights, so every product streams
// k-contiguous 16-byte weight rows.
constexpr int kSmallRows = 32;
constexpr int kSmallThreads = 1024;
constexpr int kSmallMaxDim = 16 * 32;

struct SmallArgs {
  const float* X;
  int B, P, H, ld;
  const __bf16* W[3];   // [H x K]
  const __bf16* WT[3];  // [K x H]
  const float* bias[3];
  const float* w4;
  const float* b4;
  float* E;
  float* G;
};

// 32 x 32 tile: act[32 x K] (LDS, row stride ld) times rows n0..n0+31 of
// Wm [N x K] (row-major); columns >= N contribute zero.
__device__ __forceinline__ f32x16 small_tile(const __bf16* act, int ld, const __bf16* __restrict__ Wm,
                                             int K, int n0, int N, int lane) {
  f32x16 acc = {};
  const int n = n0 + (lane & 31);
  const bool ok = n < N;
  const __bf16* wrow = Wm + size_t(ok ? n : 0) * size_t(K);
  const __bf16* arow = act + (lane & 31) * ld;
  const int half = 8 * (lane >> 5);
  // Eight k-steps' weight operands are loaded before their MFMAs, so a chunk
  // costs one memory round trip instead of eight.
  constexpr int kDepth = 8;
  for (int c0 = 0; c0 < K; c0 += 16 * kDepth) {
    bf16x8 bv[kDepth];
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
      const int k0 = c0 + 16 * u;
      bv[u] = (ok && k0 < K) ? *reinterpret_cast<const bf16x8*>(wrow + k0 + half) : bf16x8{};
    }
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
      const int k0 = c0 + 16 * u;
      if (k0 < K) {
        const bf16x8 av = *reinterpret_cast<const bf16x8*>(arow + k0 + half);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv[u], acc, 0, 0, 0);
      }
    }
  }
  return acc;
}

// D layout of the tile: column lane % 32, row 8 (r / 4) + 4 (lane / 32) + r % 4.
__device__ __forceinline__ int tile_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

__global__ __launch_bounds__(kSmallThreads) void mlp_small_kernel(SmallArgs a) {
  extern __shared__ __align__(16) __bf16 sbuf[];
  __shared__ float epart[kSmallThreads / 64][kSmallRows];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int ld = a.ld, H = a.H, P = a.P, B = a.B;
  __bf16* in = sbuf;
  __bf16* out = sbuf + kSmallRows * ld;
  for (int i = tid; i < kSmallRows * P; i += kSmallThreads) {
    const int b = i / P, k = i - b * P;
    in[b * ld + k] = to_bf16(b < B ? a.X[size_t(b) * P + k] : 0.f);
  }
  __syncthreads();
  const bool owner = wave * 32 < H;
  const int n = wave * 32 + (lane & 31);
  float d[3][16];
  float e[16];
  int K = P;
  for (int l = 0; l < 3; ++l) {
    if (owner) {
      const f32x16 acc = small_tile(in, ld, a.W[l], K, wave * 32, H, lane);
      const float bn = a.bias[l][n];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int b = tile_row(r, lane);
        const float x = acc[r] + bn, sg = sigm(x), h = x * sg;
        d[l][r] = sg * (1.f + x * (1.f - sg));
        if (l == 2) e[r] = h * a.w4[n];
        out[b * ld + n] = to_bf16(b < B ? h : 0.f);
      }
    }
    __syncthreads();
    __bf16* t = in;
    in = out;
    out = t;
    K = H;
  }
  // Energies: row sums of h3 * w4 over the 32 columns of each wavefront, then
  // over the wavefronts.
  if (owner) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      float v = e[r];
      for (int off = 1; off < 32; off <<= 1) v += __shfl_xor(v, off, 64);
      if ((lane & 31) == 0) epart[wave][tile_row(r, lane)] = v;
    }
  }
  // u3 = w4 * silu'(a3) into the free buffer.
  if (owner) {
    const float w4n = a.w4[n];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int b = tile_row(r, lane);
      out[b * ld + n] = to_bf16(b < B ? w4n * d[2][r] : 0.f);
    }
  }
  __syncthreads();
  if (tid < B) {
    float s = 0.f;
    for (int w = 0; w * 32 < H; ++w) s += epart[w][tid];
    a.E[tid] = s + a.b4[0];
  }
  {
    __bf16* t = in;
    in = out;
    out = t;
  }
  // v_{l-1} = u_l W_l = u_l (W_l^T)^T;  u_{l-1} = v_{l-1} * silu'(a_{l-1}).
  for (int l = 2; l >= 1; --l) {
    if (owner) {
      const f32x16 acc = small_tile(in, ld, a.WT[l], H, wave * 32, H, lane);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int b = tile_row(r, lane);
        out[b * ld + n] = to_bf16(b < B ? acc[r] * d[l - 1][r] : 0.f);
      }
    }
    __syncthreads();
    __bf16* t = in;
    in = out;
    out = t;
  }
  // dE/dx = u1 W1 = u1 (W1^T)^T: columns p < P.
  if (a.G && wave * 32 < P) {
    const f32x16 acc = small_tile(in, ld, a.WT[0], H, wave * 32, P, lane);
    const int p = wave * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int b = tile_row(r, lane);
      if (b < B && p < P) a.G[size_t(b) * P + p] = acc[r];
    }
  }
}

// a = z + b (row broadcast); h = silu(a); writes a, h (fp32) and h (bf16).
__global__ void bias_silu_kernel(const float* __restrict__ z, const float* __restrict__ bias, int rows,
                                 int cols, float* __restrict__ a_out, float* __restrict__ h_out,
                                 __bf16* __restrict__ hb_out) {
  const size_t n = size_t(rows) * cols;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const float a = z[i] + bias[i % cols];
    const float h = a * sigm(a);
    a_out[i] = a;
    h_out[i] = h;
    hb_out[i] = to_bf16(h);
  }
}

__global__ void to_bf16_kernel(const float* __restrict__ x, __bf16* __restrict__ y, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    y[i] = to_bf16(x[i]);
}

// E[b] = h3[b,:] . w4 + b4   (one wave per row)
__global__ void head_kernel(const float* __restrict__ h3, const float* __restrict__ w4,
                            const float* __restrict__ b4, int rows, int H, float* __restrict__ E) {
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= rows) return;
  float s = 0.f;
  for (int j = lane; j < H; j += 64) s += h3[size_t(row) * H + j] * w4[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) E[row] = s + b4[0];
}

// u = w (row vector or matrix) * silu'(a); u_bf16 copy.  w_row: broadcast w4.
__global__ void mul_dsilu_kernel(const float* __restrict__ w, int w_is_row, const float* __restrict__ a,
                                 int rows, int cols, float* __restrict__ u, __bf16* __restrict__ ub) {
  const size_t n = size_t(rows) * cols;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const float x = a[i], s = sigm(x);
    const float d = s * (1.f + x * (1.f - s));
    const float v = (w_is_row ? w[i % cols] : w[i]) * d;
    u[i] = v;
    ub[i] = to_bf16(v);
  }
}

// Loss terms and output deltas.
//   dE[b]    = 2 (E - E*) / B
//   dg[b,j]  = 2 lam (g - g*) / (B P)
// loss partials are accumulated into loss[0] (energy) and loss[1] (gradient).
__global__ void loss_kernel(const float* __restrict__ E, const float* __restrict__ Et,
                            const float* __restrict__ g, const float* __restrict__ gt, int rows, int P,
                            float lam, float* __restrict__ dE, float* __restrict__ dg,
                            __bf16* __restrict__ dgb, float* __restrict__ loss) {
  __shared__ float red[2][8];
  float le = 0.f, lg = 0.f;
  const size_t n = size_t(rows) * P;
  const float inv_b = 1.f / float(rows), inv_bp = 1.f / (float(rows) * float(P));
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const float d = g[i] - gt[i];
    lg += d * d * inv_bp;
    const float v = 2.f * lam * d * inv_bp;
    dg[i] = v;
    dgb[i] = to_bf16(v);
  }
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < rows; b += gridDim.x * blockDim.x) {
    const float d = E[b] - Et[b];
    le += d * d * inv_b;
    dE[b] = 2.f * d * inv_b;
  }
  for (int off = 32; off > 0; off >>= 1) {
    le += __shfl_xor(le, off, 64);
    lg += __shfl_xor(lg, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = le;
    red[1][threadIdx.x >> 6] = lg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = 0.f, c = 0.f;
    for (int w = 0; w < int(blockDim.x >> 6); ++w) {
      a += red[0][w];
      c += red[1][w];
    }
    atomicAdd(&loss[0], a);
    atomicAdd(&loss[1], lam * c);
  }
}

// Input-gradient branch backward through u = v * silu'(a):
//   dv = du * silu'(a)          (-> bf16 copy for the next GEMM)
//   da_g = du * v * silu''(a)
__global__ void dsilu_back_kernel(const float* __restrict__ du, const float* __restrict__ v,
                                  const float* __restrict__ a, int n_rows, int cols,
                                  float* __restrict__ dv, __bf16* __restrict__ dvb,
                                  float* __restrict__ da_g) {
  const size_t n = size_t(n_rows) * cols;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const float x = a[i], s = sigm(x);
    const float d1 = s * (1.f + x * (1.f - s));
    const float d2 = s * (1.f - s) * (2.f + x * (1.f - 2.f * s));
    const float g = du[i];
    dv[i] = g * d1;
    dvb[i] = to_bf16(g * d1);
    da_g[i] = g * v[i] * d2;
  }
}

// Column reductions run on a (H/256) x (rows/kRowChunk) grid: each thread
// sums its column over one chunk of rows and adds the partial to the
// parameter gradient with one atomic (the gradient is zeroed first).
constexpr int kRowChunk = 32;

// Last layer: u3 = w4 * silu'(a3).  Given du3:
//   dw4 += sum_b du3 * silu'(a3) + dE * h3;  da3 = du3 * w4 * silu''(a3) + dE * w4 * silu'(a3)
// Writes da3 (fp32 + bf16); db4 += sum_b dE.
__global__ void head_back_kernel(const float* __restrict__ du3, const float* __restrict__ a3,
                                 const float* __restrict__ h3, const float* __restrict__ w4,
                                 const float* __restrict__ dE, int rows, int H,
                                 float* __restrict__ da3, __bf16* __restrict__ da3b,
                                 float* __restrict__ dw4, float* __restrict__ db4) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int b0 = blockIdx.y * kRowChunk, b1 = min(rows, b0 + kRowChunk);
  if (j < H) {
    float acc = 0.f;
    const float wj = w4[j];
    for (int b = b0; b < b1; ++b) {
      const size_t i = size_t(b) * H + j;
      const float x = a3[i], s = sigm(x);
      const float d1 = s * (1.f + x * (1.f - s));
      const float d2 = s * (1.f - s) * (2.f + x * (1.f - 2.f * s));
      acc += du3[i] * d1 + dE[b] * h3[i];
      const float v = du3[i] * wj * d2 + dE[b] * wj * d1;
      da3[i] = v;
      da3b[i] = to_bf16(v);
    }
    atomicAdd(&dw4[j], acc);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float s = 0.f;
    for (int b = b0; b < b1; ++b) s += dE[b];
    atomicAdd(db4, s);
  }
}

// da = da_g + dh * silu'(a)  (fp32 + bf16), and column sums into db.
__global__ void combine_back_kernel(const float* __restrict__ da_g, const float* __restrict__ dh,
                                    const float* __restrict__ a, int rows, int H,
                                    float* __restrict__ da, __bf16* __restrict__ dab,
                                    float* __restrict__ db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= H) return;
  const int b0 = blockIdx.y * kRowChunk, b1 = min(rows, b0 + kRowChunk);
  float acc = 0.f;
  for (int b = b0; b < b1; ++b) {
    const size_t i = size_t(b) * H + j;
    const float x = a[i], s = sigm(x);
    const float v = da_g[i] + dh[i] * s * (1.f + x * (1.f - s));
    da[i] = v;
    dab[i] = to_bf16(v);
    acc += v;
  }
  atomicAdd(&db[j], acc);
}

// db[j] += sum_b d[b, j]
__global__ void colsum_kernel(const float* __restrict__ d, int rows, int H, float* __restrict__ db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= H) return;
  const int b0 = blockIdx.y * kRowChunk, b1 = min(rows, b0 + kRowChunk);
  float acc = 0.f;
  for (int b = b0; b < b1; ++b) acc += d[size_t(b) * H + j];
  atomicAdd(&db[j], acc);
}

__global__ void fill_kernel(float* p, size_t n, float v) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// Adam on the flat FP32 parameter vector, then refresh the BF16 weight copies.
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

int Err(char* err, int64_t len, const std::string& msg) {
  if (err && len > 0) {
    const size_t n = std::min<size_t>(size_t(len - 1), msg.size());
    memcpy(err, msg.data(), n);
    err[n] = '\0';
  }
  return 14;
}

size_t Al(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

void LaunchGemmBf16(int M, int N, int K, const void* A, int lda, int transA, const void* B,
                    int ldb, int transB, float* C, int ldc, float alpha, float beta,
                    hipStream_t stream) {
  if (M <= 0 || N <= 0) return;
  dim3 grid((N + kGT - 1) / kGT, (M + kGT - 1) / kGT);
  hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(kGThreads), 0, stream, M, N, K,
                     static_cast<const __bf16*>(A), lda, transA, static_cast<const __bf16*>(B), ldb,
                     transB, C, ldc, alpha, beta);
}

// ------------------------------------------------------------------ MLP layout
MlpLayout MakeMlpLayout(int P, int H, int maxB) {
  MlpLayout L{};
  L.P = P;
  L.H = H;
  L.maxB = maxB;
  // Flat FP32 parameter vector: W1[HxP] b1[H] W2[HxH] b2[H] W3[HxH] b3[H] w4[H] b4[1].
  L.oW1 = 0;
  L.ob1 = L.oW1 + size_t(H) * P;
  L.oW2 = L.ob1 + H;
  L.ob2 = L.oW2 + size_t(H) * H;
  L.oW3 = L.ob2 + H;
  L.ob3 = L.oW3 + size_t(H) * H;
  L.ow4 = L.ob3 + H;
  L.ob4 = L.ow4 + H;
  L.nparams = L.ob4 + 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = Al(off + bytes);
    return o;
  };
  const size_t np = L.nparams;
  L.w = take(np * 4);
  L.g = take(np * 4);
  L.m = take(np * 4);
  L.v = take(np * 4);
  L.wb1 = take(size_t(H) * P * 2);
  L.wb2 = take(size_t(H) * H * 2);
  L.wb3 = take(size_t(H) * H * 2);
  L.wt1 = take(size_t(H) * P * 2);
  L.wt2 = take(size_t(H) * H * 2);
  L.wt3 = take(size_t(H) * H * 2);
  const size_t BH = size_t(maxB) * H, BP = size_t(maxB) * P;
  L.xb = take(BP * 2);
  for (int i = 0; i < 3; ++i) {
    L.z[i] = take(BH * 4);
    L.a[i] = take(BH * 4);
    L.h[i] = take(BH * 4);
    L.hb[i] = take(BH * 2);
    L.u[i] = take(BH * 4);
    L.ub[i] = take(BH * 2);
    L.vv[i] = take(BH * 4);
    L.t1[i] = take(BH * 4);
    L.t1b[i] = take(BH * 2);
    L.t2[i] = take(BH * 4);
    L.t2b[i] = take(BH * 2);
    L.dag[i] = take(BH * 4);
  }
  L.E = take(size_t(maxB) * 4);
  L.gin = take(BP * 4);
  L.dE = take(size_t(maxB) * 4);
  L.dg = take(BP * 4);
  L.dgb = take(BP * 2);
  L.loss = take(256);
  L.bytes = off;
  return L;
}

namespace {

template <typename T>
T* at(void* ws, size_t off) {
  return reinterpret_cast<T*>(static_cast<uint8_t*>(ws) + off);
}

int Grid(size_t n) { return int(std::min<size_t>((n + 255) / 256, 4096)); }

void RefreshBf16(const MlpLayout& L, void* ws, hipStream_t s) {
  const float* w = at<float>(ws, L.w);
  RefreshArgs r{};
  const size_t oW[3] = {L.oW1, L.oW2, L.oW3}, wb[3] = {L.wb1, L.wb2, L.wb3},
               wt[3] = {L.wt1, L.wt2, L.wt3};
  size_t total = 0;
  for (int l = 0; l < 3; ++l) {
    r.W[l] = w + oW[l];
    r.wb[l] = at<__bf16>(ws, wb[l]);
    r.wt[l] = at<__bf16>(ws, wt[l]);
    r.cols[l] = l == 0 ? L.P : L.H;
    r.end[l] = total += size_t(L.H) * size_t(r.cols[l]);
  }
  hipLaunchKernelGGL(refresh_bf16_kernel, dim3(Grid(total)), dim3(256), 0, s, r);
}

// Forward + input gradient for B rows of X (device fp32 [B x P]).
void Forward(const MlpLayout& L, void* ws, const float* X, int B, hipStream_t s) {
  const int P = L.P, H = L.H;
  const float* w = at<float>(ws, L.w);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(Grid(size_t(B) * P)), dim3(256), 0, s, X,
                     at<__bf16>(ws, L.xb), size_t(B) * P);
  const void* in_b = at<__bf16>(ws, L.xb);
  int K = P;
  const size_t wbo[3] = {L.wb1, L.wb2, L.wb3};
  const size_t bo[3] = {L.ob1, L.ob2, L.ob3};
  for (int l = 0; l < 3; ++l) {
    // z = in * W^T  (W is [H x K] row-major -> transB)
    LaunchGemmBf16(B, H, K, in_b, K, 0, at<__bf16>(ws, wbo[l]), K, 1, at<float>(ws, L.z[l]), H, 1.f,
                   0.f, s);
    hipLaunchKernelGGL(bias_silu_kernel, dim3(Grid(size_t(B) * H)), dim3(256), 0, s,
                       at<float>(ws, L.z[l]), w + bo[l], B, H, at<float>(ws, L.a[l]),
                       at<float>(ws, L.h[l]), at<__bf16>(ws, L.hb[l]));
    in_b = at<__bf16>(ws, L.hb[l]);
    K = H;
  }
  hipLaunchKernelGGL(head_kernel, dim3((B + 3) / 4), dim3(256), 0, s, at<float>(ws, L.h[2]),
                     w + L.ow4, w + L.ob4, B, H, at<float>(ws, L.E));
  // Input gradient: u3 = w4*s'(a3); v2 = u3 W3; u2 = v2*s'(a2); v1 = u2 W2; u1 = v1*s'(a1); g = u1 W1
  hipLaunchKernelGGL(mul_dsilu_kernel, dim3(Grid(size_t(B) * H)), dim3(256), 0, s, w + L.ow4, 1,
                     at<float>(ws, L.a[2]), B, H, at<float>(ws, L.u[2]), at<__bf16>(ws, L.ub[2]));
  for (int l = 2; l >= 1; --l) {
    // v_{l-1} = u_l * W_l (W_l [H x H], no transpose): [B x H]
    LaunchGemmBf16(B, H, H, at<__bf16>(ws, L.ub[l]), H, 0, at<__bf16>(ws, wbo[l]), H, 0,
                   at<float>(ws, L.vv[l - 1]), H, 1.f, 0.f, s);
    hipLaunchKernelGGL(mul_dsilu_kernel, dim3(Grid(size_t(B) * H)), dim3(256), 0, s,
                       at<float>(ws, L.vv[l - 1]), 0, at<float>(ws, L.a[l - 1]), B, H,
                       at<float>(ws, L.u[l - 1]), at<__bf16>(ws, L.ub[l - 1]));
  }
  LaunchGemmBf16(B, P, H, at<__bf16>(ws, L.ub[0]), H, 0, at<__bf16>(ws, L.wb1), P, 0,
                 at<float>(ws, L.gin), P, 1.f, 0.f, s);
}

// Loss and parameter gradient (into L.g) for targets Et [B], Gt [B x P].
void LossGrad(const MlpLayout& L, void* ws, const float* X, const float* Et, const float* Gt, int B,
              float lam, hipStream_t s) {
  const int P = L.P, H = L.H;
  float* w = at<float>(ws, L.w);
  float* g = at<float>(ws, L.g);
  Forward(L, ws, X, B, s);
  hipLaunchKernelGGL(fill_kernel, dim3(Grid(L.nparams)), dim3(256), 0, s, g, L.nparams, 0.f);
  hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(64), 0, s, at<float>(ws, L.loss), size_t(2), 0.f);
  hipLaunchKernelGGL(loss_kernel, dim3(std::max(1, std::min(Grid(size_t(B) * P), 1024))), dim3(256),
                     0, s, at<float>(ws, L.E), Et, at<float>(ws, L.gin), Gt, B, P, lam,
                     at<float>(ws, L.dE), at<float>(ws, L.dg), at<__bf16>(ws, L.dgb),
                     at<float>(ws, L.loss));
  const size_t wbo[3] = {L.wb1, L.wb2, L.wb3};
  const size_t Wo[3] = {L.oW1, L.oW2, L.oW3};
  const size_t bo[3] = {L.ob1, L.ob2, L.ob3};
  // g = u1 W1: dW1 += u1^T dg ; du1 = dg W1^T
  LaunchGemmBf16(H, P, B, at<__bf16>(ws, L.ub[0]), H, 1, at<__bf16>(ws, L.dgb), P, 0, g + L.oW1, P,
                 1.f, 1.f, s);
  LaunchGemmBf16(B, H, P, at<__bf16>(ws, L.dgb), P, 0, at<__bf16>(ws, L.wb1), P, 1,
                 at<float>(ws, L.t1[0]), H, 1.f, 0.f, s);
  // Walk the input-gradient chain upward: layer l = 0,1 (u_l = v_l * s'(a_l), v_l = u_{l+1} W_{l+1}).
  for (int l = 0; l < 2; ++l) {
    // dv_l = du_l * s'(a_l); da_g(l) = du_l * v_l * s''(a_l)
    hipLaunchKernelGGL(dsilu_back_kernel, dim3(Grid(size_t(B) * H)), dim3(256), 0, s,
                       at<float>(ws, L.t1[l]), at<float>(ws, L.vv[l]), at<float>(ws, L.a[l]), B, H,
                       at<float>(ws, L.t2[l]), at<__bf16>(ws, L.t2b[l]), at<float>(ws, L.dag[l]));
    // v_l = u_{l+1} W_{l+1}: dW_{l+1} += u_{l+1}^T dv_l ; du_{l+1} = dv_l W_{l+1}^T
    LaunchGemmBf16(H, H, B, at<__bf16>(ws, L.ub[l + 1]), H, 1, at<__bf16>(ws, L.t2b[l]), H, 0,
                   g + Wo[l + 1], H, 1.f, 1.f, s);
    LaunchGemmBf16(B, H, H, at<__bf16>(ws, L.t2b[l]), H, 0, at<__bf16>(ws, wbo[l + 1]), H, 1,
                   at<float>(ws, L.t1[l + 1]), H, 1.f, 0.f, s);
  }
  // Head: u3 = w4 * s'(a3) with du3 = t1[2]; plus the energy branch.
  const dim3 red_grid((H + 255) / 256, (B + kRowChunk - 1) / kRowChunk);
  hipLaunchKernelGGL(head_back_kernel, red_grid, dim3(256), 0, s, at<float>(ws, L.t1[2]),
                     at<float>(ws, L.a[2]), at<float>(ws, L.h[2]), w + L.ow4, at<float>(ws, L.dE), B, H,
                     at<float>(ws, L.t2[2]), at<__bf16>(ws, L.t2b[2]), g + L.ow4, g + L.ob4);
  // Backward through the forward chain: da_l -> dW_l, db_l, dh_{l-1}.
  for (int l = 2; l >= 0; --l) {
    const void* in_b = l == 0 ? static_cast<const void*>(at<__bf16>(ws, L.xb))
                              : static_cast<const void*>(at<__bf16>(ws, L.hb[l - 1]));
    const int K = l == 0 ? P : H;
    if (l < 2) {
      // da_l = da_g(l) + dh_l * s'(a_l); dh_l is in t1[l] (reused buffer)
      hipLaunchKernelGGL(combine_back_kernel, red_grid, dim3(256), 0, s,
                         at<float>(ws, L.dag[l]), at<float>(ws, L.t1[l]), at<float>(ws, L.a[l]), B, H,
                         at<float>(ws, L.t2[l]), at<__bf16>(ws, L.t2b[l]), g + bo[l]);
    } else {
      // da3 (both branches) is already in t2[2]; db3 = its column sums.
      hipLaunchKernelGGL(colsum_kernel, red_grid, dim3(256), 0, s, at<float>(ws, L.t2[2]),
                         B, H, g + bo[2]);
    }
    // dW_l += da_l^T in
    LaunchGemmBf16(H, K, B, at<__bf16>(ws, L.t2b[l]), H, 1, in_b, K, 0, g + Wo[l], K, 1.f, 1.f, s);
    if (l > 0) {
      // dh_{l-1} = da_l W_l  -> t1[l-1]
      LaunchGemmBf16(B, H, H, at<__bf16>(ws, L.t2b[l]), H, 0, at<__bf16>(ws, wbo[l]), H, 0,
                     at<float>(ws, L.t1[l - 1]), H, 1.f, 0.f, s);
    }
  }
  (void)w;
}

}  // namespace

int MlpSetParams(const MlpLayout& L, void* ws, const float* host_params, hipStream_t s, char* err,
                 int64_t len) {
  hipError_t e = hipMemcpyAsync(at<float>(ws, L.w), host_params, L.nparams * 4, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return Err(err, len, std::string("set params: ") + hipGetErrorString(e));
  hipLaunchKernelGGL(fill_kernel, dim3(Grid(L.nparams)), dim3(256), 0, s, at<float>(ws, L.m), L.nparams, 0.f);
  hipLaunchKernelGGL(fill_kernel, dim3(Grid(L.nparams)), dim3(256), 0, s, at<float>(ws, L.v), L.nparams, 0.f);
  RefreshBf16(L, ws, s);
  e = hipGetLastError();
  return e == hipSuccess ? 0 : Err(err, len, std::string("set params: ") + hipGetErrorString(e));
}

int MlpPredict(const MlpLayout& L, void* ws, const float* X, int B, float* E, float* G, hipStream_t s,
               char* err, int64_t len) {
  if (B > L.maxB) return Err(err, len, "batch larger than the workspace was sized for");
  if (B <= kSmallRows && L.H % 32 == 0 && L.H <= kSmallMaxDim && L.P % 16 == 0 &&
      L.P <= kSmallMaxDim) {
    SmallArgs a{};
    a.X = X;
    a.B = B;
    a.P = L.P;
    a.H = L.H;
    a.ld = std::max(L.P, L.H) + 8;
    const float* w = at<float>(ws, L.w);
    const size_t wb[3] = {L.wb1, L.wb2, L.wb3}, wt[3] = {L.wt1, L.wt2, L.wt3},
                 ob[3] = {L.ob1, L.ob2, L.ob3};
    for (int l = 0; l < 3; ++l) {
      a.W[l] = at<__bf16>(ws, wb[l]);
      a.WT[l] = at<__bf16>(ws, wt[l]);
      a.bias[l] = w + ob[l];
    }
    a.w4 = w + L.ow4;
    a.b4 = w + L.ob4;
    a.E = E;
    a.G = G;
    const size_t lds = size_t(2) * kSmallRows * size_t(a.ld) * sizeof(__bf16);
    hipLaunchKernelGGL(mlp_small_kernel, dim3(1), dim3(kSmallThreads), lds, s, a);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : Err(err, len, std::string("predict: ") + hipGetErrorString(e));
  }
  Forward(L, ws, X, B, s);
  hipError_t e = hipMemcpyAsync(E, at<float>(ws, L.E), size_t(B) * 4, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess && G)
    e = hipMemcpyAsync(G, at<float>(ws, L.gin), size_t(B) * L.P * 4, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipGetLastError();
  return e == hipSuccess ? 0 : Err(err, len, std::string("predict: ") + hipGetErrorString(e));
}

int MlpLossGrad(const MlpLayout& L, void* ws, const float* X, const float* Et, const float* Gt, int B,
                float lam, hipStream_t s, char* err, int64_t len) {
  if (B > L.maxB) return Err(err, len, "batch larger than the workspace was sized for");
  LossGrad(L, ws, X, Et, Gt, B, lam, s);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : Err(err, len, std::string("loss/grad: ") + hipGetErrorString(e));
}

int MlpTrainStep(const MlpLayout& L, void* ws, const float* X, const float* Et, const float* Gt, int B,
                 float lam, float lr, int step, hipStream_t s, char* err, int64_t len) {
  if (B > L.maxB) return Err(err, len, "batch larger than the workspace was sized for");
  LossGrad(L, ws, X, Et, Gt, B, lam, s);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - std::pow(b1, float(step)), c2 = 1.f - std::pow(b2, float(step));
  hipLaunchKernelGGL(adam_kernel, dim3(Grid(L.nparams)), dim3(256), 0, s, at<float>(ws, L.w),
                     at<float>(ws, L.g), at<float>(ws, L.m), at<float>(ws, L.v), L.nparams, lr, b1, b2,
                     eps, c1, c2);
  RefreshBf16(L, ws, s);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : Err(err, len, std::string("train step: ") + hipGetErrorString(e));
}

}  // namespace tfqx
