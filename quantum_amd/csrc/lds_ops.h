// Device routines shared by the LDS-resident batch kernel and the
// HBM-streaming tile-pass kernels: complex arithmetic, k-qubit gate
// application on a state tile held in LDS, block reductions/scans and the
// Philox4x32-10 sampler stream.
#pragma once

#include <hip/hip_runtime.h>

#include "tfqx_internal.h"

namespace tfqx {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmul_conj_a(float2 a, float2 b) {  // conj(a) * b
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }

__device__ __forceinline__ uint32_t insert_zero(uint32_t x, uint32_t q) {
  return ((x >> q) << (q + 1)) | (x & ((1u << q) - 1u));
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < kWaves; ++i) t += red[i];
  return t;
}

// Exclusive block scan of one double per thread.
__device__ double block_exclusive_scan(double v, double* red, double* total) {
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
template <int K, bool DAG>
__device__ void apply_dense(float2* __restrict__ s, int n, const DevOp& op,
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
        acc = cadd(acc, cmul(m, v[c]));
      }
      s[base | dep[r]] = acc;
    }
  }
}

template <bool DAG>
__device__ void apply_diag(float2* __restrict__ s, int n, const DevOp& op,
                           const float2* __restrict__ Dg) {
  const uint32_t N = 1u << n;
  const int k = op.k;
  for (uint32_t x = threadIdx.x; x < N; x += kThreads) {
    if ((x & op.cmask) != op.cval) continue;
    uint32_t m = 0;
    for (int j = 0; j < k; ++j) m |= ((x >> op.q[j]) & 1u) << j;
    float2 d = Dg[m];
    if (DAG) d.y = -d.y;
    s[x] = cmul(d, s[x]);
  }
}

template <bool DAG>
__device__ void apply_op(float2* s, int n, const DevOp& op, const float2* mats) {
  const float2* M = mats + op.mat;
  if (op.type == kOpDiag) {
    apply_diag<DAG>(s, n, op, M);
  } else {
    switch (op.k) {
      case 1: apply_dense<1, DAG>(s, n, op, M); break;
      case 2: apply_dense<2, DAG>(s, n, op, M); break;
      case 3: apply_dense<3, DAG>(s, n, op, M); break;
      default: apply_dense<4, DAG>(s, n, op, M); break;
    }
  }
}

// One-qubit matrix on bit q (Z-basis rotations of the sampled expectation).
__device__ void apply_1q(float2* __restrict__ s, int n, int q, const float2* M) {
  const uint32_t ngroups = 1u << (n - 1);
  const float2 m00 = M[0], m01 = M[1], m10 = M[2], m11 = M[3];
  for (uint32_t g = threadIdx.x; g < ngroups; g += kThreads) {
    const uint32_t b = insert_zero(g, q);
    const float2 a0 = s[b], a1 = s[b | (1u << q)];
    s[b] = cadd(cmul(m00, a0), cmul(m01, a1));
    s[b | (1u << q)] = cadd(cmul(m10, a0), cmul(m11, a1));
  }
}

// Adjoint step for one parameterised gate G (targets q, controls cmask):
//   psi <- G^dag psi;  acc_d += Re<lam| dG_d psi>;  lam <- G^dag lam.
// Control-mismatched amplitudes are untouched and contribute nothing
// (the BulkSetAmpl projection of tfq_adj_grad_op.cc:241-260).
template <int K>
__device__ void apply_param(float2* __restrict__ psi, float2* __restrict__ lam, int n,
                            const DevOp& op, const float2* __restrict__ mats, double* acc) {
  constexpr int D = 1 << K;
  const float2* G = mats + op.mat;
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
      for (int c = 0; c < D; ++c) s = cadd(s, cmul_conj_a(G[c * D + r], a[c]));
      t[r] = s;
    }
#pragma unroll
    for (int m = 0; m < D; ++m) psi[base | dep[m]] = t[m];
    for (int d = 0; d < nd; ++d) {
      const float2* dG = mats + op.dmat[d];
      float part = 0.f;
#pragma unroll
      for (int r = 0; r < D; ++r) {
        float2 s = make_float2(0.f, 0.f);
#pragma unroll
        for (int c = 0; c < D; ++c) s = cadd(s, cmul(dG[r * D + c], t[c]));
        part += s.x * b[r].x + s.y * b[r].y;  // Re(conj(s) * lam)
      }
      acc[d] += double(part);
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float2 s = make_float2(0.f, 0.f);
#pragma unroll
      for (int c = 0; c < D; ++c) s = cadd(s, cmul_conj_a(G[c * D + r], b[c]));
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

__device__ __forceinline__ uint64_t tile_index(uint64_t base, uint32_t j, const DevPass& p) {
  uint64_t g = base | (j & ((1u << p.nlow) - 1u));
  uint32_t hj = j >> p.nlow;
  for (int h = p.nlow; h < p.L; ++h) {
    g |= uint64_t(hj & 1u) << p.tbit[h];
    hj >>= 1;
  }
  return g;
}

}  // namespace
}  // namespace tfqx
