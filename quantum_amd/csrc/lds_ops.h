// Device routines shared by the LDS-resident batch kernel and the
// HBM-streaming tile-pass kernels: complex arithmetic, k-qubit gate
// application on a state tile held in LDS, block reductions/scans and the
// Philox4x32-10 sampler stream.
#pragma once

#include <hip/hip_runtime.h>

#include "tfqx_internal.h"

namespace tfqx {
namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
// float2 entries of one matrix staging buffer in LDS: a fused unitary of up
// to 4 qubits (16x16), or a parameterised gate (<= 2 qubits) with its two
// derivative matrices.
constexpr int kMatSlot = 256;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmul_conj_a(float2 a, float2 b) {  // conj(a) * b
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
// acc + m * v as four FMAs (a product followed by an add leaves separate
// multiplies and adds).
__device__ __forceinline__ float2 cmacf(float2 acc, float2 m, float2 v) {
  float re = __builtin_fmaf(m.x, v.x, acc.x);
  re = __builtin_fmaf(-m.y, v.y, re);
  float im = __builtin_fmaf(m.x, v.y, acc.y);
  im = __builtin_fmaf(m.y, v.x, im);
  return make_float2(re, im);
}

__device__ __forceinline__ uint32_t insert_zero(uint32_t x, uint32_t q) {
  return ((x >> q) << (q + 1)) | (x & ((1u << q) - 1u));
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Read-only plan data (op descriptors, gate matrices) is not written while a
// kernel runs, so it can be read through the constant address space: with a
// wave-uniform address the compiler then emits scalar loads into SGPRs
// instead of per-lane vector loads into VGPRs.
template <typename T>
using kconst = const __attribute__((address_space(4))) T;
template <typename T>
__device__ __forceinline__ kconst<T>* as_const(const T* p) {
  return (kconst<T>*)(p);
}

// Load an op descriptor into scalar registers (all lanes read the same op).
__device__ __forceinline__ DevOp load_op(const DevOp* p) {
  // Field-wise from 32-bit words so the descriptor stays in registers.
  kconst<uint32_t>* src = as_const(reinterpret_cast<const uint32_t*>(p));
  uint32_t w[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) w[i] = __builtin_amdgcn_readfirstlane(src[i]);
  DevOp o;
  o.type = uint8_t(w[0]);
  o.k = uint8_t(w[0] >> 8);
  o.nderiv = uint8_t(w[0] >> 16);
  o.pad0 = 0;
  o.q[0] = uint8_t(w[1]);
  o.q[1] = uint8_t(w[1] >> 8);
  o.q[2] = uint8_t(w[1] >> 16);
  o.q[3] = uint8_t(w[1] >> 24);
  o.cmask = w[2];
  o.cval = w[3];
  o.mat = w[4];
  o.dmat[0] = w[5];
  o.dmat[1] = w[6];
  o.col[0] = int32_t(w[7]);
  o.col[1] = int32_t(w[8]);
  return o;
}

// Copy an op's matrices into an LDS staging buffer: [G | dG_0 | dG_1].
__device__ __forceinline__ void stage_op(const DevOp& op, const float2* __restrict__ mats,
                                         float2* __restrict__ dst) {
  const int D = 1 << op.k;
  const int DD = op.type == kOpDiag ? D : D * D;
  for (int i = threadIdx.x; i < DD; i += kThreads) dst[i] = mats[op.mat + i];
  if (op.type == kOpParam) {
#pragma unroll
    for (int d = 0; d < kMaxDeriv; ++d)
      if (d < op.nderiv)
        for (int i = threadIdx.x; i < DD; i += kThreads) dst[DD * (1 + d) + i] = mats[op.dmat[d] + i];
  }
}

// Run ops [b, e): f(op, M) with M the op's matrices staged in LDS.  The next
// op's matrices are staged while the current op runs (two buffers), so each
// op costs one barrier.
template <typename F>
__device__ __forceinline__ void for_ops(const DevOp* ops, int b, int e, const float2* mats,
                                        float2* mbuf, F&& f) {
  if (b >= e) return;
  {
    const DevOp op0 = load_op(ops + b);
    stage_op(op0, mats, mbuf);
  }
  __syncthreads();
  for (int i = b; i < e; ++i) {
    const DevOp op = load_op(ops + i);
    if (i + 1 < e) {
      const DevOp nx = load_op(ops + i + 1);
      stage_op(nx, mats, mbuf + ((i + 1 - b) & 1) * kMatSlot);
    }
    f(op, mbuf + ((i - b) & 1) * kMatSlot);
    __syncthreads();
  }
}

// v + (v of the lane that DPP control CTRL selects within the lane's row of 16).
template <int CTRL>
__device__ __forceinline__ double dpp_add(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return v + __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_value(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}
// Sum over the 64 lanes of a wavefront (all lanes active), returned in every
// lane.  Within each row of 16 lanes: DPP quad permutations (lane ^ 1, lane ^
// 2), then the half-row mirror (lane 7 - i) and the row mirror (lane 15 - i),
// which pair each lane with lanes holding the other partial; the four row
// sums are then read lane by lane.  All steps are VALU operations, with no LDS
// round trip.
__device__ __forceinline__ double wave_sum(double v) {
  v = dpp_add<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v = dpp_add<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v = dpp_add<0x141>(v);  // row_half_mirror
  v = dpp_add<0x140>(v);  // row_mirror
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < kWaves; ++i) t += red[i];
  return t;
}

// Exclusive block scan of one double per thread.
__device__ __forceinline__ double block_exclusive_scan(double v, double* red, double* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double inc = v;
  for (int off = 1; off < 64; off <<= 1) {
    double o = __shfl_up(inc, off, 64);
    if (lane >= off) inc += o;
  }
  __syncthreads();
  if (lane == 63) red[w] = inc;
  __syncthreads();
  double base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kWaves; ++i) {
    if (i < w) base += red[i];
    tot += red[i];
  }
  *total = tot;
  return base + inc - v;
}

// ------------------------------------------------------------------ RNG
// Philox4x32-10 (Salmon et al., SC'11).  Stream layout: key = seed,
// counter = (shot lo, shot hi, stream lo, stream hi); u = top 53 bits of
// (x0:x1) / 2^53.
__device__ __forceinline__ double philox_uniform(unsigned long long seed, unsigned long long shot,
                                                 uint32_t s_lo, uint32_t s_hi) {
  uint32_t c0 = uint32_t(shot), c1 = uint32_t(shot >> 32), c2 = s_lo, c3 = s_hi;
  uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  const unsigned long long bits = (static_cast<unsigned long long>(c0) << 32) | c1;
  return double(bits >> 11) * (1.0 / 9007199254740992.0);
}

// ------------------------------------------------------------------ gate application
// s <- M s (or M^dagger s) on targets q[0..K-1] of an n-qubit state in LDS.
// For K >= 3 the matrix entries are re-read from LDS per group instead of
// being hoisted into registers (a 16x16 complex matrix would take 512 VGPRs).
template <int K, bool DAG>
__device__ __forceinline__ void apply_dense(float2* __restrict__ s, int n, const DevOp& op,
                            const float2* __restrict__ M) {
  constexpr int D = 1 << K;
  const uint32_t ngroups = 1u << (n - K);
  uint32_t dep[D];
#pragma unroll
  for (int m = 0; m < D; ++m) {
    uint32_t o = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) o |= uint32_t((m >> j) & 1) << op.q[j];
    dep[m] = o;
  }
  for (uint32_t g = threadIdx.x; g < ngroups; g += kThreads) {
    uint32_t base = g;
#pragma unroll
    for (int j = 0; j < K; ++j) base = insert_zero(base, op.q[j]);
    if ((base & op.cmask) != op.cval) continue;
    if constexpr (K >= 3) __asm__ volatile("" ::: "memory");  // keep M in LDS, not VGPRs
    float2 v[D];
#pragma unroll
    for (int m = 0; m < D; ++m) v[m] = s[base | dep[m]];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int c = 0; c < D; ++c) {
        float2 m = DAG ? M[c * D + r] : M[r * D + c];
        if (DAG) m.y = -m.y;
        acc = c == 0 ? cmul(m, v[c]) : cmacf(acc, m, v[c]);
      }
      s[base | dep[r]] = acc;
    }
  }
}

template <bool DAG>
__device__ __forceinline__ void apply_diag(float2* __restrict__ s, int n, const DevOp& op,
                           const float2* __restrict__ Dg) {
  const uint32_t N = 1u << n;
  const int k = op.k;
  for (uint32_t x = threadIdx.x; x < N; x += kThreads) {
    if ((x & op.cmask) != op.cval) continue;
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < kMaxFuse; ++j)
      if (j < k) m |= ((x >> op.q[j]) & 1u) << j;
    float2 d = Dg[m];
    if (DAG) d.y = -d.y;
    s[x] = cmul(d, s[x]);
  }
}

// KMAX bounds the fused width the caller may see (fewer template cases keep
// register pressure low when all ops are 1-2 qubit).
template <bool DAG, int KMAX = kMaxFuse>
__device__ __forceinline__ void apply_op(float2* s, int n, const DevOp& op, const float2* M) {
  if (op.type == kOpDiag) {
    apply_diag<DAG>(s, n, op, M);
  } else if (KMAX <= 2 || op.k <= 2) {
    if (op.k == 1) apply_dense<1, DAG>(s, n, op, M);
    else apply_dense<2, DAG>(s, n, op, M);
  } else {
    if (op.k == 3) apply_dense<(KMAX > 2 ? 3 : 2), DAG>(s, n, op, M);
    else apply_dense<(KMAX > 2 ? 4 : 2), DAG>(s, n, op, M);
  }
}

// One-qubit matrix on bit q (Z-basis rotations of the sampled expectation).
__device__ __forceinline__ void apply_1q(float2* __restrict__ s, int n, int q, const float2* M) {
  const uint32_t ngroups = 1u << (n - 1);
  const float2 m00 = M[0], m01 = M[1], m10 = M[2], m11 = M[3];
  for (uint32_t g = threadIdx.x; g < ngroups; g += kThreads) {
    const uint32_t b = insert_zero(g, q);
    const float2 a0 = s[b], a1 = s[b | (1u << q)];
    s[b] = cmacf(cmul(m00, a0), m01, a1);
    s[b | (1u << q)] = cmacf(cmul(m10, a0), m11, a1);
  }
}

// Adjoint step for one parameterised gate G (targets q, controls cmask):
//   psi <- G^dag psi;  acc_d += <dG_d psi | lam> (real part in acc[d], imaginary
//   part in acc[kMaxDeriv + d]);  lam <- G^dag lam.
// Control-mismatched amplitudes are untouched and contribute nothing
// (the BulkSetAmpl projection of tfq_adj_grad_op.cc:241-260).
template <int K>
__device__ __forceinline__ void apply_param(float2* __restrict__ psi, float2* __restrict__ lam, int n,
                            const DevOp& op, const float2* __restrict__ M, double* acc) {
  constexpr int D = 1 << K;
  const float2* G = M;
  const uint32_t ngroups = 1u << (n - K);
  const int nd = op.nderiv;
  uint32_t dep[D];
#pragma unroll
  for (int m = 0; m < D; ++m) {
    uint32_t o = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) o |= uint32_t((m >> j) & 1) << op.q[j];
    dep[m] = o;
  }
  for (uint32_t g = threadIdx.x; g < ngroups; g += kThreads) {
    uint32_t base = g;
#pragma unroll
    for (int j = 0; j < K; ++j) base = insert_zero(base, op.q[j]);
    if ((base & op.cmask) != op.cval) continue;
    float2 a[D], b[D], t[D];
#pragma unroll
    for (int m = 0; m < D; ++m) {
      a[m] = psi[base | dep[m]];
      b[m] = lam[base | dep[m]];
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float2 s = make_float2(0.f, 0.f);
#pragma unroll
      for (int c = 0; c < D; ++c) {
        float2 g = G[c * D + r];
        g.y = -g.y;
        s = c == 0 ? cmul(g, a[c]) : cmacf(s, g, a[c]);
      }
      t[r] = s;
    }
#pragma unroll
    for (int m = 0; m < D; ++m) psi[base | dep[m]] = t[m];
#pragma unroll
    for (int d = 0; d < kMaxDeriv; ++d) {
      if (d >= nd) break;
      const float2* dG = M + D * D * (1 + d);
      float part = 0.f, part_im = 0.f;
#pragma unroll
      for (int r = 0; r < D; ++r) {
        float2 s = make_float2(0.f, 0.f);
#pragma unroll
        for (int c = 0; c < D; ++c) s = c == 0 ? cmul(dG[r * D], t[0]) : cmacf(s, dG[r * D + c], t[c]);
        part += s.x * b[r].x + s.y * b[r].y;     // Re(conj(s) * lam)
        part_im += s.x * b[r].y - s.y * b[r].x;  // Im(conj(s) * lam)
      }
      acc[d] += double(part);
      acc[kMaxDeriv + d] += double(part_im);
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float2 s = make_float2(0.f, 0.f);
#pragma unroll
      for (int c = 0; c < D; ++c) {
        float2 g = G[c * D + r];
        g.y = -g.y;
        s = c == 0 ? cmul(g, b[c]) : cmacf(s, g, b[c]);
      }
      lam[base | dep[r]] = s;
    }
  }
}


// Tile addressing of a pass: tile t's base index (t's bits deposited into the
// non-tile bits) and the global index of tile-local amplitude j.
__device__ __forceinline__ uint64_t tile_base(uint64_t t, uint64_t tm, int n) {
  uint64_t base = 0;
  for (int b = 0; b < n; ++b) {
    if (!((tm >> b) & 1)) {
      base |= (t & 1) << b;
      t >>= 1;
    }
  }
  return base;
}

// Per-workgroup table of the deposit of j's high bits (j >> nlow) into the
// pass's high tile bits; kTileTable entries cover passes with L - nlow <= 8.
constexpr int kTileTable = 256;

__device__ __forceinline__ uint64_t deposit_high(uint32_t jh, const DevPass& p) {
  // Bits of jh go to tbit[nlow], tbit[nlow+1], ...; the loop is unrolled so
  // tbit is indexed statically (no private-memory copy of the pass).
  uint64_t g = 0;
#pragma unroll
  for (int h = 0; h < kMaxTile; ++h) {
    if (h >= p.nlow && h < p.L) g |= uint64_t((jh >> (h - p.nlow)) & 1u) << p.tbit[h];
  }
  return g;
}

__device__ __forceinline__ void build_tile_table(const DevPass& p, uint64_t* tab) {
  const int hb = p.L - p.nlow;
  if (hb > 8) return;
  for (int jh = threadIdx.x; jh < (1 << hb); jh += kThreads) tab[jh] = deposit_high(jh, p);
}

__device__ __forceinline__ uint64_t tile_index_tab(uint64_t base, uint32_t j, int nlow,
                                                   const uint64_t* tab) {
  return base | (j & ((1u << nlow) - 1u)) | tab[j >> nlow];
}

__device__ __forceinline__ uint64_t tile_index(uint64_t base, uint32_t j, const DevPass& p) {
  return base | (j & ((1u << p.nlow) - 1u)) | deposit_high(j >> p.nlow, p);
}

// ------------------------------------------------------------------ tile transfer
// Global <-> LDS tile copies with kGatherBatch independent global accesses in
// flight per thread (a one-at-a-time loop leaves the copy latency-bound).
// idx(j) maps tile-local amplitude j to its global index.
constexpr int kGatherBatch = 8;

template <typename IndexFn>
__device__ __forceinline__ void tile_gather(float2* __restrict__ A, const float2* __restrict__ G,
                                            uint32_t T, IndexFn&& idx) {
  for (uint32_t j0 = threadIdx.x; j0 < T; j0 += kThreads * kGatherBatch) {
    float v[2 * kGatherBatch];
#pragma unroll
    for (int i = 0; i < kGatherBatch; ++i) {
      const uint32_t j = j0 + uint32_t(i) * kThreads;
      if (j < T) {
        const float2 x = G[idx(j)];
        v[2 * i] = x.x;
        v[2 * i + 1] = x.y;
      }
    }
#pragma unroll
    for (int i = 0; i < kGatherBatch; ++i) {
      const uint32_t j = j0 + uint32_t(i) * kThreads;
      if (j < T) A[j] = make_float2(v[2 * i], v[2 * i + 1]);
    }
  }
}

// Two states through the same index map (psi and lambda of the adjoint).
template <typename IndexFn>
__device__ __forceinline__ void tile_gather2(float2* __restrict__ A, float2* __restrict__ B,
                                             const float2* __restrict__ GA,
                                             const float2* __restrict__ GB, uint32_t T,
                                             IndexFn&& idx) {
  for (uint32_t j0 = threadIdx.x; j0 < T; j0 += kThreads * kGatherBatch) {
    float va[2 * kGatherBatch], vb[2 * kGatherBatch];
#pragma unroll
    for (int i = 0; i < kGatherBatch; ++i) {
      const uint32_t j = j0 + uint32_t(i) * kThreads;
      if (j < T) {
        const uint64_t g = idx(j);
        const float2 x = GA[g], y = GB[g];
        va[2 * i] = x.x;
        va[2 * i + 1] = x.y;
        vb[2 * i] = y.x;
        vb[2 * i + 1] = y.y;
      }
    }
#pragma unroll
    for (int i = 0; i < kGatherBatch; ++i) {
      const uint32_t j = j0 + uint32_t(i) * kThreads;
      if (j < T) {
        A[j] = make_float2(va[2 * i], va[2 * i + 1]);
        B[j] = make_float2(vb[2 * i], vb[2 * i + 1]);
      }
    }
  }
}

template <typename IndexFn>
__device__ __forceinline__ void tile_scatter(float2* __restrict__ G, const float2* __restrict__ A,
                                             uint32_t T, IndexFn&& idx) {
  for (uint32_t j0 = threadIdx.x; j0 < T; j0 += kThreads * kGatherBatch) {
    float v[2 * kGatherBatch];
#pragma unroll
    for (int i = 0; i < kGatherBatch; ++i) {
      const uint32_t j = j0 + uint32_t(i) * kThreads;
      if (j < T) {
        const float2 x = A[j];
        v[2 * i] = x.x;
        v[2 * i + 1] = x.y;
      }
    }
#pragma unroll
    for (int i = 0; i < kGatherBatch; ++i) {
      const uint32_t j = j0 + uint32_t(i) * kThreads;
      if (j < T) G[idx(j)] = make_float2(v[2 * i], v[2 * i + 1]);
    }
  }
}

}  // namespace
}  // namespace tfqx
