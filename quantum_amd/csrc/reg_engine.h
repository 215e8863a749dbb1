// Register-resident gate engine for a 2^L-amplitude state tile held in LDS.
//
// A workgroup of 512 threads keeps the tile in VGPRs: each thread holds 2^R
// amplitudes (R = L - 9, 3..4) whose indices differ only in R "register
// qubits"; the thread index fills the other L - R bits.  Ops whose targets
// are register qubits (and diagonal ops on any qubits, with controls on any
// qubits) are applied in registers with no LDS traffic and no barrier.  The
// host planner splits the op list into segments whose non-diagonal targets
// fit R qubits and starts each with a kOpRemap op carrying the segment's
// register-qubit mask; a remap writes the registers back to LDS and reloads
// them in the new layout (two barriers).  Between remaps LDS holds the tile in
// a swizzled layout (x ^ ((x >> 4) & 15)) so that the strided register loads
// and stores of a wavefront are spread over the LDS banks; on entry and exit
// the layout is the natural one.
//
// Op matrices are read from an LDS window over the plan's matrix array,
// refilled when an op's matrices fall outside it.  Parameterised ops (reverse sweep) accumulate their FP64
// gradient partials per thread, reduce them across the wavefront and hand the
// sum to the caller's accumulator.
#pragma once

#include <type_traits>

#include "lds_ops.h"

namespace tfqx {
namespace {

static_assert(kThreads == (1 << kRegThreadBits), "register engine assumes 512 threads");

// Register-resident complex amplitude.  A plain two-float struct (not the HIP
// vector type, whose union members turn register arrays into scratch memory).
struct cf {
  float x, y;
};
__device__ __forceinline__ cf cmulr(cf a, cf b) {
  return cf{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cf caddr(cf a, cf b) { return cf{a.x + b.x, a.y + b.y}; }
// acc + u * v as four FMAs.
__device__ __forceinline__ cf cmac(cf acc, cf u, cf v) {
  acc.x = __builtin_fmaf(u.x, v.x, acc.x);
  acc.x = __builtin_fmaf(-u.y, v.y, acc.x);
  acc.y = __builtin_fmaf(u.x, v.y, acc.y);
  acc.y = __builtin_fmaf(u.y, v.x, acc.y);
  return acc;
}
__device__ __forceinline__ cf cconj(cf a) { return cf{a.x, -a.y}; }
__device__ __forceinline__ cf ldm(const float2* __restrict__ M, int i) {
  const float* f = reinterpret_cast<const float*>(M);
  return cf{f[2 * i], f[2 * i + 1]};
}

__device__ __forceinline__ uint32_t swz(uint32_t x) { return x ^ ((x >> 4) & 15u); }

// ------------------------------------------------------------------ MFMA 4x4 products
// v_mfma_f32_4x4x1f32 runs 16 independent 4x4x1 blocks per wavefront, block b
// on lanes 4b..4b+3: D_b[i][j] += A_b[i] * B_b[j] with A_b[i] taken from lane
// 4b + i, B_b[j] from lane 4b + j, and D_b[i][j] returned in register i of
// lane 4b + j.  With A in lane l holding G[l & 3][k] and B holding each
// lane's own amplitude of gate index k, four steps (k = 0..3) leave in
// register r of every lane sum_k G[r][k] * v_k of that lane's amplitudes: a
// 4x4 gate applied in place to the registers, at twice the FP32 rate of VALU
// FMAs.  Complex products split into real ones: Re = Gr vr - Gi vi,
// Im = Gi vr + Gr vi.  MFMA ignores the EXEC mask, so these paths run only in
// wave-uniform control flow (ops without per-lane control predicates).
typedef float mf4 __attribute__((ext_vector_type(4)));

// A operands of a 4x4 complex matrix (or its adjoint): row (lane & 3).
struct MfmaRows {
  float gr[4], gi[4], ngi[4];
};
template <bool DAG>
__device__ __forceinline__ void mfma_rows(const float2* __restrict__ M, MfmaRows& r) {
  const int row = int(threadIdx.x & 3u);
  const float* f = reinterpret_cast<const float*>(M);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    // (G^dag)[row][k] = conj(G[k][row]).
    const int e = DAG ? k * 4 + row : row * 4 + k;
    r.gr[k] = f[2 * e];
    r.gi[k] = DAG ? -f[2 * e + 1] : f[2 * e + 1];
    r.ngi[k] = -r.gi[k];
  }
}
__device__ __forceinline__ mf4 mfma4(float a, float b, mf4 c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}
// (dr, di) = G v for the four amplitudes v_k = a[idx[k]] of every lane.
template <int N>
__device__ __forceinline__ void mfma_apply(const cf (&a)[N], const int (&idx)[4],
                                           const MfmaRows& m, mf4& dr, mf4& di) {
  dr = mf4{0.f, 0.f, 0.f, 0.f};
  di = mf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    dr = mfma4(m.gr[k], a[idx[k]].x, dr);
    di = mfma4(m.gi[k], a[idx[k]].x, di);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    dr = mfma4(m.ngi[k], a[idx[k]].y, dr);
    di = mfma4(m.gr[k], a[idx[k]].y, di);
  }
}

__host__ __device__ constexpr int ins0(int g, int j) {
  return ((g >> j) << (j + 1)) | (g & ((1 << j) - 1));
}

template <int R>
struct RegMap {
  uint32_t rpos[R];  // register qubit positions, ascending (wave-uniform)
  uint32_t mask;     // wave-uniform
  uint32_t tbase;    // this thread's bits on the non-register positions
};

// Tile index of register i (compile-time i) relative to tbase.
template <int R>
__device__ __forceinline__ uint32_t rdep(const RegMap<R>& m, int i) {
  uint32_t o = 0;
#pragma unroll
  for (int j = 0; j < R; ++j)
    if ((i >> j) & 1) o |= 1u << m.rpos[j];
  return o;
}

template <int R, bool WAVE = false>
__device__ __forceinline__ void make_map(RegMap<R>& m, uint32_t mask, int L) {
  mask = uint32_t(uniform(int(mask)));
  m.mask = mask;
  uint32_t mm = mask;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    m.rpos[j] = mm ? uint32_t(__builtin_ctz(mm)) : 0u;  // the planner always sets R bits
    mm &= mm - 1;
  }
  // Wave mode: the lane id supplies the L - R <= 6 thread bits; lanes whose
  // higher bits are set replicate a lane below them.
  const uint32_t tid = WAVE ? (threadIdx.x & 63u) : threadIdx.x;
  uint32_t tb = 0;
  int k = 0;
  for (int p = 0; p < L; ++p) {
    if (!((mask >> p) & 1u)) {
      tb |= ((tid >> k) & 1u) << p;
      ++k;
    }
  }
  m.tbase = tb;
}

// Byte offsets in LDS of the 2^R registers of this thread.  Register i's
// index is tbase | rdep(i) = tbase ^ (XOR of bit(rpos[j]) over the set bits j
// of i), and the swizzle is XOR-linear (swz(a ^ b) = swz(a) ^ swz(b)), so
// offset(i) = offset(i without its highest bit h) ^ sb[h]: one XOR per
// register after R per-map terms.
template <int R>
__device__ __forceinline__ void reg_offsets(const RegMap<R>& m, bool swizzled,
                                            uint32_t (&off)[1 << R]) {
  uint32_t sb[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const uint32_t b = 1u << m.rpos[j];
    sb[j] = (swizzled ? swz(b) : b) << 3;
  }
  off[0] = (swizzled ? swz(m.tbase) : m.tbase) << 3;
#pragma unroll
  for (int i = 1; i < (1 << R); ++i) {
    const int h = 31 - __builtin_clz(uint32_t(i));
    off[i] = off[i ^ (1 << h)] ^ sb[h];
  }
}

template <int R>
__device__ __forceinline__ void reg_load(cf (&a)[1 << R], const float2* __restrict__ S,
                                         const RegMap<R>& m, bool swizzled) {
  uint32_t off[1 << R];
  reg_offsets<R>(m, swizzled, off);
  const char* base = reinterpret_cast<const char*>(S);
#pragma unroll
  for (int i = 0; i < (1 << R); ++i) {
    const float2 v = *reinterpret_cast<const float2*>(base + off[i]);
    a[i] = cf{v.x, v.y};
  }
}

template <int R>
__device__ __forceinline__ void reg_store(const cf (&a)[1 << R], float2* __restrict__ S,
                                          const RegMap<R>& m, bool swizzled) {
  uint32_t off[1 << R];
  reg_offsets<R>(m, swizzled, off);
  char* base = reinterpret_cast<char*>(S);
#pragma unroll
  for (int i = 0; i < (1 << R); ++i)
    *reinterpret_cast<float2*>(base + off[i]) = make_float2(a[i].x, a[i].y);
}

// ------------------------------------------------------------------ dense ops
template <int R, int J, bool DAG, bool BOTH, bool CTRL>
__device__ __forceinline__ void reg_dense1_c(cf (&a)[1 << R], cf (&b)[1 << R],
                                             const float2* __restrict__ M, const RegMap<R>& m,
                                             uint32_t cmask, uint32_t cval) {
  cf u00 = ldm(M, 0), u01 = ldm(M, 1), u10 = ldm(M, 2), u11 = ldm(M, 3);
  if (DAG) {
    const cf t = u01;
    u00 = cconj(u00);
    u01 = cconj(u10);
    u10 = cconj(t);
    u11 = cconj(u11);
  }
#pragma unroll
  for (int g = 0; g < (1 << (R - 1)); ++g) {
    constexpr int one = 1 << J;
    const int i0 = ins0(g, J);
    if (CTRL && (((m.tbase | rdep(m, i0)) & cmask) != cval)) continue;
    const cf v0 = a[i0], v1 = a[i0 | one];
    a[i0] = cmac(cmulr(u00, v0), u01, v1);
    a[i0 | one] = cmac(cmulr(u10, v0), u11, v1);
    if (BOTH) {
      const cf w0 = b[i0], w1 = b[i0 | one];
      b[i0] = cmac(cmulr(u00, w0), u01, w1);
      b[i0 | one] = cmac(cmulr(u10, w0), u11, w1);
    }
  }
}

template <int R, int J0, int J1>
__device__ __forceinline__ void group_idx(int g, int (&idx)[4]) {
  const int i0 = ins0(ins0(g, J0), J1);
#pragma unroll
  for (int s = 0; s < 4; ++s) idx[s] = i0 | ((s & 1) << J0) | ((s >> 1) << J1);
}

template <int R, int J0, int J1, bool DAG, bool BOTH, bool CTRL>
__device__ __forceinline__ void reg_dense2_c(cf (&a)[1 << R], cf (&b)[1 << R],
                                             const float2* __restrict__ M, const RegMap<R>& m,
                                             uint32_t cmask, uint32_t cval) {
  if constexpr (!CTRL) {
    MfmaRows rows;
    mfma_rows<DAG>(M, rows);
#pragma unroll
    for (int g = 0; g < (1 << (R - 2)); ++g) {
      int idx[4];
      group_idx<R, J0, J1>(g, idx);
      mf4 dr, di;
      mfma_apply(a, idx, rows, dr, di);
#pragma unroll
      for (int r = 0; r < 4; ++r) a[idx[r]] = cf{dr[r], di[r]};
      if (BOTH) {
        mfma_apply(b, idx, rows, dr, di);
#pragma unroll
        for (int r = 0; r < 4; ++r) b[idx[r]] = cf{dr[r], di[r]};
      }
    }
    return;
  }
  cf u[16];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) u[r * 4 + c] = DAG ? cconj(ldm(M, c * 4 + r)) : ldm(M, r * 4 + c);
#pragma unroll
  for (int g = 0; g < (1 << (R - 2)); ++g) {
    const int i0 = ins0(ins0(g, J0), J1);
    if (CTRL && (((m.tbase | rdep(m, i0)) & cmask) != cval)) continue;
    int idx[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) idx[s] = i0 | ((s & 1) << J0) | ((s >> 1) << J1);
    cf v[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) v[s] = a[idx[s]];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      cf acc = cmulr(u[r * 4], v[0]);
#pragma unroll
      for (int c = 1; c < 4; ++c) acc = cmac(acc, u[r * 4 + c], v[c]);
      a[idx[r]] = acc;
    }
    if (BOTH) {
#pragma unroll
      for (int s = 0; s < 4; ++s) v[s] = b[idx[s]];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        cf acc = cmulr(u[r * 4], v[0]);
#pragma unroll
        for (int c = 1; c < 4; ++c) acc = cmac(acc, u[r * 4 + c], v[c]);
        b[idx[r]] = acc;
      }
    }
  }
}

// Gates without control qubits (the common case) run a check-free loop: a
// per-group predicate makes the compiler copy every output register first.
template <int R, int J, bool DAG, bool BOTH>
__device__ __forceinline__ void reg_dense1(cf (&a)[1 << R], cf (&b)[1 << R],
                                           const float2* __restrict__ M, const RegMap<R>& m,
                                           uint32_t cmask, uint32_t cval) {
  if (cmask) reg_dense1_c<R, J, DAG, BOTH, true>(a, b, M, m, cmask, cval);
  else reg_dense1_c<R, J, DAG, BOTH, false>(a, b, M, m, cmask, cval);
}

template <int R, int J0, int J1, bool DAG, bool BOTH>
__device__ __forceinline__ void reg_dense2(cf (&a)[1 << R], cf (&b)[1 << R],
                                           const float2* __restrict__ M, const RegMap<R>& m,
                                           uint32_t cmask, uint32_t cval) {
  if (cmask) reg_dense2_c<R, J0, J1, DAG, BOTH, true>(a, b, M, m, cmask, cval);
  else reg_dense2_c<R, J0, J1, DAG, BOTH, false>(a, b, M, m, cmask, cval);
}

template <int R, bool DAG, bool BOTH, int J = 0>
__device__ __forceinline__ void dispatch1(int j, cf (&a)[1 << R], cf (&b)[1 << R], const float2* M,
                                          const RegMap<R>& m, uint32_t cmask, uint32_t cval) {
  if constexpr (J < R) {
    if (j == J) reg_dense1<R, J, DAG, BOTH>(a, b, M, m, cmask, cval);
    else dispatch1<R, DAG, BOTH, J + 1>(j, a, b, M, m, cmask, cval);
  }
}

template <int R, bool DAG, bool BOTH, int J0 = 0, int J1 = 1>
__device__ __forceinline__ void dispatch2(int j0, int j1, cf (&a)[1 << R], cf (&b)[1 << R],
                                          const float2* M, const RegMap<R>& m, uint32_t cmask,
                                          uint32_t cval) {
  if constexpr (J1 < R) {
    if (j0 == J0 && j1 == J1) {
      reg_dense2<R, J0, J1, DAG, BOTH>(a, b, M, m, cmask, cval);
    } else if constexpr (J1 + 1 < R) {
      dispatch2<R, DAG, BOTH, J0, J1 + 1>(j0, j1, a, b, M, m, cmask, cval);
    } else if constexpr (J0 + 2 < R) {
      dispatch2<R, DAG, BOTH, J0 + 1, J0 + 2>(j0, j1, a, b, M, m, cmask, cval);
    }
  }
}

// Tile positions of the wavefront-index bits of the thread index (bits
// 6..8) under register mask `mask`: make_map gives thread bit k the k-th
// lowest non-register position, so they are the non-register positions left
// after the six lowest (the lane bits).
__device__ __forceinline__ uint32_t wave_positions(uint32_t mask, int L) {
  uint32_t rest = ((1u << L) - 1u) & ~mask;
#pragma unroll
  for (int k = 0; k < 6; ++k) rest &= rest - 1u;
  return rest;
}

template <int R>
__device__ __forceinline__ int reg_slot(const RegMap<R>& m, uint32_t q) {
  return __builtin_popcount(m.mask & ((1u << q) - 1u));
}

// Dense op on a (and, BOTH, on b with the same matrix).
template <int R, bool DAG, bool BOTH>
__device__ __forceinline__ void reg_dense(cf (&a)[1 << R], cf (&b)[1 << R], const DevOp& op,
                                          const float2* M, const RegMap<R>& m) {
  if (op.k == 1) {
    dispatch1<R, DAG, BOTH>(reg_slot(m, op.q[0]), a, b, M, m, op.cmask, op.cval);
  } else {
    dispatch2<R, DAG, BOTH>(reg_slot(m, op.q[0]), reg_slot(m, op.q[1]), a, b, M, m, op.cmask,
                            op.cval);
  }
}

// ------------------------------------------------------------------ diagonal ops
// Targets and controls may be any tile qubits; k <= 2.
template <int R, bool DAG, bool BOTH>
__device__ __forceinline__ void reg_diag(cf (&a)[1 << R], cf (&b)[1 << R], const DevOp& op,
                                         const float2* __restrict__ Dg, const RegMap<R>& m) {
  const bool two = op.k == 2;
  cf d0 = ldm(Dg, 0), d1 = ldm(Dg, 1);
  cf d2 = two ? ldm(Dg, 2) : d0, d3 = two ? ldm(Dg, 3) : d1;
  if (DAG) {
    d0 = cconj(d0);
    d1 = cconj(d1);
    d2 = cconj(d2);
    d3 = cconj(d3);
  }
  const uint32_t q0 = op.q[0], q1 = two ? op.q[1] : op.q[0];
  const bool ctrl = op.cmask != 0;
#pragma unroll
  for (int i = 0; i < (1 << R); ++i) {
    const uint32_t x = m.tbase | rdep(m, i);
    if (ctrl && (x & op.cmask) != op.cval) continue;
    const uint32_t b0 = (x >> q0) & 1u, b1 = two ? ((x >> q1) & 1u) : 0u;
    const cf d = b1 ? (b0 ? d3 : d2) : (b0 ? d1 : d0);
    a[i] = cmulr(d, a[i]);
    if (BOTH) b[i] = cmulr(d, b[i]);
  }
}

// Uncontrolled diagonal ops with the position of each target resolved at
// compile time: P0 / P1 is the register slot of q[0] / q[1], or R when that
// target is a thread bit (fixed for the whole thread).  The phase of every
// register is then a compile-time choice among at most two per-thread values,
// so each amplitude costs one complex multiply.
template <int R, bool DAG, bool BOTH, bool TWO, int P0, int P1>
__device__ __forceinline__ void reg_diag_p(cf (&a)[1 << R], cf (&b)[1 << R], const DevOp& op,
                                           const float2* __restrict__ Dg, const RegMap<R>& m) {
  cf d[4];
  d[0] = ldm(Dg, 0);
  d[1] = ldm(Dg, 1);
  d[2] = TWO ? ldm(Dg, 2) : d[0];
  d[3] = TWO ? ldm(Dg, 3) : d[1];
  if (DAG) {
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = cconj(d[k]);
  }
  // Fold the thread-bit targets in: e[c] is the phase for register-bit
  // pattern c (bit 0: q0, bit 1: q1) of this thread.
  const bool t0 = P0 == R && ((m.tbase >> op.q[0]) & 1u);
  const bool t1 = TWO && P1 == R && ((m.tbase >> op.q[1]) & 1u);
  cf e[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int b0 = P0 == R ? -1 : (c & 1);
    const int b1 = (!TWO) ? 0 : (P1 == R ? -1 : (c >> 1));
    // Candidates for the thread bits, chosen per thread.
    const cf x0 = d[(b1 < 0 ? 0 : b1) * 2 + (b0 < 0 ? 0 : b0)];
    const cf x1 = d[(b1 < 0 ? 0 : b1) * 2 + (b0 < 0 ? 1 : b0)];
    const cf y0 = d[(b1 < 0 ? 1 : b1) * 2 + (b0 < 0 ? 0 : b0)];
    const cf y1 = d[(b1 < 0 ? 1 : b1) * 2 + (b0 < 0 ? 1 : b0)];
    const cf lo = (b0 < 0 && t0) ? x1 : x0;
    const cf hi = (b0 < 0 && t0) ? y1 : y0;
    e[c] = (b1 < 0 && t1) ? hi : lo;
  }
#pragma unroll
  for (int i = 0; i < (1 << R); ++i) {
    const int c = (P0 == R ? 0 : ((i >> P0) & 1)) | (TWO && P1 != R ? (((i >> P1) & 1) << 1) : 0);
    a[i] = cmulr(e[c], a[i]);
    if (BOTH) b[i] = cmulr(e[c], b[i]);
  }
}

template <int R, bool DAG, bool BOTH, int P0 = 0, int P1 = 0>
__device__ __forceinline__ void diag_dispatch2(int p0, int p1, cf (&a)[1 << R], cf (&b)[1 << R],
                                               const DevOp& op, const float2* Dg,
                                               const RegMap<R>& m) {
  if constexpr (P0 <= R) {
    if constexpr (P1 <= R) {
      // Two register targets sit in ascending slots (q[0] < q[1]); other
      // pairs cannot occur and get no code.
      if constexpr (P0 == R || P1 == R || P0 < P1) {
        if (p0 == P0 && p1 == P1) {
          reg_diag_p<R, DAG, BOTH, true, P0, P1>(a, b, op, Dg, m);
          return;
        }
      }
      diag_dispatch2<R, DAG, BOTH, P0, P1 + 1>(p0, p1, a, b, op, Dg, m);
    } else {
      diag_dispatch2<R, DAG, BOTH, P0 + 1, 0>(p0, p1, a, b, op, Dg, m);
    }
  }
}

template <int R, bool DAG, bool BOTH, int P0 = 0>
__device__ __forceinline__ void diag_dispatch1(int p0, cf (&a)[1 << R], cf (&b)[1 << R],
                                               const DevOp& op, const float2* Dg,
                                               const RegMap<R>& m) {
  if constexpr (P0 <= R) {
    if (p0 == P0) reg_diag_p<R, DAG, BOTH, false, P0, R>(a, b, op, Dg, m);
    else diag_dispatch1<R, DAG, BOTH, P0 + 1>(p0, a, b, op, Dg, m);
  }
}

// Diagonal op: position-resolved variants without controls, the generic
// per-amplitude selection with them.
template <int R, bool DAG, bool BOTH>
__device__ __forceinline__ void reg_diag_any(cf (&a)[1 << R], cf (&b)[1 << R], const DevOp& op,
                                             const float2* __restrict__ Dg, const RegMap<R>& m) {
  if (op.cmask) {
    reg_diag<R, DAG, BOTH>(a, b, op, Dg, m);
    return;
  }
  const int p0 = ((m.mask >> op.q[0]) & 1u) ? reg_slot(m, op.q[0]) : R;
  if (op.k == 2) {
    const int p1 = ((m.mask >> op.q[1]) & 1u) ? reg_slot(m, op.q[1]) : R;
    diag_dispatch2<R, DAG, BOTH>(p0, p1, a, b, op, Dg, m);
  } else {
    diag_dispatch1<R, DAG, BOTH>(p0, a, b, op, Dg, m);
  }
}

// ------------------------------------------------------------------ parameterised ops
// psi <- G^dag psi;  acc_d += <dG_d psi | lam> (real part in acc[d], imaginary
// part in acc[kMaxDeriv + d]);  lam <- G^dag lam  (see apply_param).
// Three sweeps over the registers so that only one matrix is live at a time.
template <int R, int K, int J0, int J1>
__device__ __forceinline__ int param_idx(int g, int s) {
  const int i0 = K == 1 ? ins0(g, J0) : ins0(ins0(g, J0), J1);
  return K == 1 ? (i0 | (s << J0)) : (i0 | ((s & 1) << J0) | ((s >> 1) << J1));
}

template <int R, int K, int J0, int J1, bool CTRL>
__device__ __forceinline__ void reg_dag_apply(cf (&a)[1 << R], const float2* __restrict__ G,
                                              const DevOp& op, const RegMap<R>& m) {
  constexpr int D = 1 << K;
  if constexpr (K == 2 && !CTRL) {
    MfmaRows rows;
    mfma_rows<true>(G, rows);
#pragma unroll
    for (int g = 0; g < (1 << (R - 2)); ++g) {
      int idx[4];
      group_idx<R, J0, J1>(g, idx);
      mf4 dr, di;
      mfma_apply(a, idx, rows, dr, di);
#pragma unroll
      for (int r = 0; r < 4; ++r) a[idx[r]] = cf{dr[r], di[r]};
    }
    return;
  }
  cf gd[D * D];
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int c = 0; c < D; ++c) gd[r * D + c] = cconj(ldm(G, c * D + r));
#pragma unroll
  for (int g = 0; g < (1 << (R - K)); ++g) {
    const int i0 = param_idx<R, K, J0, J1>(g, 0);
    if (CTRL && (((m.tbase | rdep(m, i0)) & op.cmask) != op.cval)) continue;
    cf v[D];
#pragma unroll
    for (int c = 0; c < D; ++c) v[c] = a[param_idx<R, K, J0, J1>(g, c)];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      cf s = cmulr(gd[r * D], v[0]);
#pragma unroll
      for (int c = 1; c < D; ++c) s = cmac(s, gd[r * D + c], v[c]);
      a[param_idx<R, K, J0, J1>(g, r)] = s;
    }
  }
}

template <int R, int K, int J0, int J1, bool CTRL>
__device__ __forceinline__ void reg_param_kc(cf (&a)[1 << R], cf (&b)[1 << R], const DevOp& op,
                                             const float2* __restrict__ mats, const RegMap<R>& m,
                                             double* acc) {
  constexpr int D = 1 << K;
  reg_dag_apply<R, K, J0, J1, CTRL>(a, mats + op.mat, op, m);
  for (int d = 0; d < op.nderiv; ++d) {
    const float2* dG = mats + (d == 0 ? op.dmat[0] : op.dmat[1]);  // static field access
    if constexpr (K == 2 && !CTRL) {
      MfmaRows rows;
      mfma_rows<false>(dG, rows);
      float part = 0.f, part_im = 0.f;
#pragma unroll
      for (int g = 0; g < (1 << (R - 2)); ++g) {
        int idx[4];
        group_idx<R, J0, J1>(g, idx);
        mf4 dr, di;
        mfma_apply(a, idx, rows, dr, di);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const cf l = b[idx[r]];
          part = __builtin_fmaf(dr[r], l.x, part);
          part = __builtin_fmaf(di[r], l.y, part);
          part_im = __builtin_fmaf(dr[r], l.y, part_im);
          part_im = __builtin_fmaf(-di[r], l.x, part_im);
        }
      }
      if (d == 0) {
        acc[0] += double(part);
        acc[kMaxDeriv] += double(part_im);
      } else {
        acc[1] += double(part);
        acc[kMaxDeriv + 1] += double(part_im);
      }
      continue;
    }
    cf dg[D * D];
#pragma unroll
    for (int i = 0; i < D * D; ++i) dg[i] = ldm(dG, i);
    float part = 0.f, part_im = 0.f;
#pragma unroll
    for (int g = 0; g < (1 << (R - K)); ++g) {
      const int i0 = param_idx<R, K, J0, J1>(g, 0);
      if (CTRL && (((m.tbase | rdep(m, i0)) & op.cmask) != op.cval)) continue;
#pragma unroll
      for (int r = 0; r < D; ++r) {
        cf s = cmulr(dg[r * D], a[param_idx<R, K, J0, J1>(g, 0)]);
#pragma unroll
        for (int c = 1; c < D; ++c) s = cmac(s, dg[r * D + c], a[param_idx<R, K, J0, J1>(g, c)]);
        const cf l = b[param_idx<R, K, J0, J1>(g, r)];
        part += s.x * l.x + s.y * l.y;
        part_im += s.x * l.y - s.y * l.x;
      }
    }
    if (d == 0) {
      acc[0] += double(part);
      acc[kMaxDeriv] += double(part_im);
    } else {
      acc[1] += double(part);
      acc[kMaxDeriv + 1] += double(part_im);
    }
  }
  reg_dag_apply<R, K, J0, J1, CTRL>(b, mats + op.mat, op, m);
}

template <int R, int K, int J0, int J1>
__device__ __forceinline__ void reg_param_k(cf (&a)[1 << R], cf (&b)[1 << R], const DevOp& op,
                                            const float2* __restrict__ mats, const RegMap<R>& m,
                                            double* acc) {
  if (op.cmask) reg_param_kc<R, K, J0, J1, true>(a, b, op, mats, m, acc);
  else reg_param_kc<R, K, J0, J1, false>(a, b, op, mats, m, acc);
}

template <int R, int J0 = 0>
__device__ __forceinline__ void param_dispatch1(int j, cf (&a)[1 << R], cf (&b)[1 << R],
                                                const DevOp& op, const float2* mats,
                                                const RegMap<R>& m, double* acc) {
  if constexpr (J0 < R) {
    if (j == J0) reg_param_k<R, 1, J0, 0>(a, b, op, mats, m, acc);
    else param_dispatch1<R, J0 + 1>(j, a, b, op, mats, m, acc);
  }
}

template <int R, int J0 = 0, int J1 = 1>
__device__ __forceinline__ void param_dispatch2(int j0, int j1, cf (&a)[1 << R],
                                                cf (&b)[1 << R], const DevOp& op,
                                                const float2* mats, const RegMap<R>& m,
                                                double* acc) {
  if constexpr (J1 < R) {
    if (j0 == J0 && j1 == J1) {
      reg_param_k<R, 2, J0, J1>(a, b, op, mats, m, acc);
    } else if constexpr (J1 + 1 < R) {
      param_dispatch2<R, J0, J1 + 1>(j0, j1, a, b, op, mats, m, acc);
    } else if constexpr (J0 + 2 < R) {
      param_dispatch2<R, J0 + 1, J0 + 2>(j0, j1, a, b, op, mats, m, acc);
    }
  }
}


// ------------------------------------------------------------------ noise channels
// Trajectory context of a noisy forward run, passed as the run's acc_fn:
// channel c of trajectory `traj` of output row `row` draws
// uniform(seed, shot = traj, stream = (row, c)).  In wave mode a wavefront
// holds 2^(6 - gbits) trajectories side by side, each on 2^gbits consecutive
// lanes, and traj is this lane's.
//
// The variates of a trajectory are drawn 2^gbits at a time: lane li of the
// trajectory's group computes the variate of channel base + li, and channel
// c (base <= c < base + 2^gbits) reads it from lane c - base of its group.
// Channel ordinals only grow along the op list and every lane of a wavefront
// meets the same channels, so `base` is wave-uniform.
struct NoiseDraw {
  unsigned long long seed;
  unsigned long long traj;
  uint32_t row;
  int gbits;
  int base = -1;
  double cached = 0.0;
  __device__ __forceinline__ double operator()(int ch) {
    const int width = 1 << gbits;
    if (base < 0 || ch < base || ch >= base + width) {
      base = ch;
      const uint32_t li = (threadIdx.x & 63u) & uint32_t(width - 1);
      cached = philox_uniform(seed, traj, row, uint32_t(ch) + li);
    }
    const int src = int((threadIdx.x & 63u) & ~uint32_t(width - 1)) + (ch - base);
    return __shfl(cached, src, 64);
  }
  // Gradient partials (reverse runs only; never called on a forward run).
  __device__ __forceinline__ void operator()(int, int, double, double) const {}
};
template <typename T>
struct IsNoise : std::false_type {};
template <>
struct IsNoise<NoiseDraw> : std::true_type {};

// Sum over the 2^gbits lanes of this lane's trajectory group (all of them
// end with the same bits: each butterfly step adds the same two values).
// The DPP steps pair each lane with one that holds the same partial as its
// butterfly partner (lane ^ 4, lane ^ 8), so the sums are the butterfly's.
__device__ __forceinline__ double group_sum(double v, int gbits) {
  if (gbits > 0) v = dpp_add<0xB1>(v);
  if (gbits > 1) v = dpp_add<0x4E>(v);
  if (gbits > 2) v = dpp_add<0x141>(v);
  if (gbits > 3) v = dpp_add<0x140>(v);
  if (gbits > 4) v += __shfl_xor(v, 16, 64);
  if (gbits > 5) v += __shfl_xor(v, 32, 64);
  return v;
}

// One channel on register slot J of this lane's trajectory.  Unitary
// mixtures pick unitary k with the first u < cum[k]; other channels take
// p_k = <psi|K_k^dag K_k|psi> from the qubit's reduced density matrix
// (group sums), pick the first k with u < p_0 + ... + p_k (the last k with
// p_k > 0 when none) and apply K_k / sqrt(p_k).  Lanes of different
// trajectories may pick different operators: the matrix is read per lane.
template <int R, int J>
__device__ __forceinline__ void reg_channel_j(cf (&a)[1 << R], const DevOp& op,
                                              const float2* __restrict__ mats,
                                              NoiseDraw& nd) {
  constexpr int G = 1 << (R - 1);
  const int nk = ChanKraus(op.cval);
  const double u = nd(op.col[0]);
  const float2* X = mats + op.dmat[0];
  int pick = nk - 1;
  double pp = 1.0;
  if (op.cval & kChanMixture) {
    for (int k = nk - 2; k >= 0; --k)
      if (u < double(X[k].x)) pick = k;
    if ((op.cval & kChanFirstIdentity) && pick == 0) return;
  } else {
    float s00 = 0.f, s11 = 0.f, sr = 0.f, si = 0.f;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int i0 = ((g >> J) << (J + 1)) | (g & ((1 << J) - 1));
      const cf x0 = a[i0], x1 = a[i0 | (1 << J)];
      s00 += x0.x * x0.x + x0.y * x0.y;
      s11 += x1.x * x1.x + x1.y * x1.y;
      sr += x0.x * x1.x + x0.y * x1.y;  // conj(x0) * x1
      si += x0.x * x1.y - x0.y * x1.x;
    }
    const double d00 = group_sum(double(s00), nd.gbits), d11 = group_sum(double(s11), nd.gbits);
    const double dr = group_sum(double(sr), nd.gbits), di = group_sum(double(si), nd.gbits);
    double cum = 0.0;
    int last = 0;
    pick = -1;
    for (int k = 0; k < nk; ++k) {
      const float2 e00 = X[4 * k], e01 = X[4 * k + 1], e11 = X[4 * k + 3];
      const double pk = double(e00.x) * d00 + double(e11.x) * d11 +
                        2.0 * (double(e01.x) * dr - double(e01.y) * di);
      if (pk > 0.0) last = k;
      cum += pk;
      if (pick < 0 && u < cum) {
        pick = k;
        pp = pk;
      }
    }
    if (pick < 0) {
      pick = last;
      const float2 e00 = X[4 * last], e01 = X[4 * last + 1], e11 = X[4 * last + 3];
      pp = double(e00.x) * d00 + double(e11.x) * d11 + 2.0 * (double(e01.x) * dr - double(e01.y) * di);
    }
  }
  const float2* M = mats + op.mat + 4 * pick;
  const float sc = (op.cval & kChanMixture) ? 1.0f : float(1.0 / sqrt(pp));
  const cf m00{M[0].x * sc, M[0].y * sc}, m01{M[1].x * sc, M[1].y * sc};
  const cf m10{M[2].x * sc, M[2].y * sc}, m11{M[3].x * sc, M[3].y * sc};
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int i0 = ((g >> J) << (J + 1)) | (g & ((1 << J) - 1));
    const cf x0 = a[i0], x1 = a[i0 | (1 << J)];
    a[i0] = cmac(cmulr(m00, x0), m01, x1);
    a[i0 | (1 << J)] = cmac(cmulr(m10, x0), m11, x1);
  }
}

template <int R, int J = 0>
__device__ __forceinline__ void reg_channel(int j, cf (&a)[1 << R], const DevOp& op,
                                            const float2* mats, NoiseDraw& nd) {
  if constexpr (J < R) {
    if (j == J) reg_channel_j<R, J>(a, op, mats, nd);
    else reg_channel<R, J + 1>(j, a, op, mats, nd);
  }
}

// ------------------------------------------------------------------ driver
// LDS window [lo, hi) of the plan's matrix array (global g).  All waves of a
// workgroup run the same op sequence, so a refill (two barriers) happens at
// the same op in every wave; windows stay valid for the whole launch because
// the matrices do not change while it runs.
struct MatCache {
  float2* lds;  // n entries
  const float2* g;
  uint32_t lo, hi;
  uint32_t n;
};
__device__ __forceinline__ MatCache mat_cache(float2* lds, const float2* g,
                                              uint32_t n = kMatCache) {
  return MatCache{lds, g, 0u, 0u, n};
}
// Make entries [off, off + len) resident in the window (every plan pads its
// matrix array by at least its kernel's window, so a refill stays inside).
__device__ __forceinline__ void cache_span(MatCache& c, uint32_t off, uint32_t len) {
  if (off >= c.lo && off + len <= c.hi) return;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < c.n; i += kThreads) c.lds[i] = c.g[off + i];
  __syncthreads();
  c.lo = off;
  c.hi = off + c.n;
}
// Matrix entries an op reads: its unitary (or diagonal) and derivatives,
// stored back to back.
__device__ __forceinline__ uint32_t op_mat_len(const DevOp& op) {
  const uint32_t D = 1u << op.k;
  const uint32_t DD = op.type == kOpDiag ? D : D * D;
  return DD * (1u + (op.type == kOpParam ? uint32_t(op.nderiv) : 0u));
}

// Matrices read straight from the plan's array (wave mode: waves run
// independent circuits, so no workgroup-wide refill is possible).
struct GlobalMats {
  const float2* g;
};
// Make an op's matrices readable and return the array its offsets index
// (offsets rebased to the LDS window when cached).
__device__ __forceinline__ const float2* mats_for(MatCache& c, DevOp& op) {
  cache_span(c, op.mat, op_mat_len(op));
  op.mat -= c.lo;
  op.dmat[0] -= c.lo;
  op.dmat[1] -= c.lo;
  return c.lds;
}
__device__ __forceinline__ const float2* mats_for(GlobalMats& c, DevOp&) { return c.g; }

// Engine synchronisation: the workgroup (one tile per workgroup) or, in wave
// mode, the wavefront (one circuit per wavefront; its LDS traffic is ordered
// by the wavefront's own in-order LDS queue).
template <bool WAVE>
__device__ __forceinline__ void esync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Tile I/O of the engine.  LdsIO: the tile enters and leaves in LDS buffers A
// (and B) in the natural layout; the caller synchronises before the run and
// the engine synchronises before returning.  TileIO: the first remap loads the
// registers straight from memory with load(x, a, b) and the run ends by
// storing them with store(x, a, b), x the tile index of each register (b only
// in the reverse sweep); LDS is then used only by the remaps in between, and
// no barrier surrounds the run (a remap's own barriers order its LDS use).
// With an index function (tile_io_indexed), load and store receive index(x)
// instead of x.  The index must be OR-linear over disjoint bits of x (a
// tile's bits deposited into the global index, plus the tile's base), so the
// engine computes it once per thread (its own bits) and once per register
// qubit, and each register's index is one OR of those.
struct LdsIO {};
struct IdentityIndex {
  __device__ __forceinline__ uint64_t operator()(uint32_t x) const { return x; }
};
template <typename LoadFn, typename StoreFn, typename IndexFn = IdentityIndex>
struct TileIO {
  LoadFn load;
  StoreFn store;
  IndexFn index;
};
template <typename LoadFn, typename StoreFn>
__device__ __forceinline__ TileIO<LoadFn, StoreFn> tile_io(LoadFn l, StoreFn s) {
  return TileIO<LoadFn, StoreFn>{l, s, IdentityIndex{}};
}
template <typename IndexFn, typename LoadFn, typename StoreFn>
__device__ __forceinline__ TileIO<LoadFn, StoreFn, IndexFn> tile_io_indexed(IndexFn i, LoadFn l,
                                                                            StoreFn s) {
  return TileIO<LoadFn, StoreFn, IndexFn>{l, s, i};
}

// index(tbase | rdep(m, r)) for every register r: the thread's part once, one
// term per register qubit, then one OR per register (f(r, index)).
template <int R, typename IO, typename F>
__device__ __forceinline__ void for_reg_index(const RegMap<R>& m, IO& io, F&& f) {
  uint64_t g[1 << R];
  g[0] = io.index(m.tbase);
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const uint64_t gj = io.index(1u << m.rpos[j]);
#pragma unroll
    for (int r = 0; r < (1 << j); ++r) g[r | (1 << j)] = g[r] | gj;
  }
#pragma unroll
  for (int r = 0; r < (1 << R); ++r) f(r, g[r]);
}
template <typename IO>
struct IsLdsIO {
  static constexpr bool value = false;
};
template <>
struct IsLdsIO<LdsIO> {
  static constexpr bool value = true;
};

// Apply ops [b, e) to a tile of 2^L amplitudes (A, and B in the reverse sweep).
// REV = false: apply each op's matrix to A.  REV = true: un-apply (M^dagger)
// non-parameterised ops on A and B; parameterised ops run the adjoint step
// and call acc_fn(slot, column, re, im) from lane 0 of each wavefront with the
// wavefront's partial <dG psi | lambda>
// per derivative, slot counting the derivative matrices of earlier
// parameterised ops and column being the derivative's symbol column.
// Ops [b, e) must start with a kOpRemap.  A and B also serve as the remap
// buffers.
template <int R, bool REV, bool WAVE, typename MS, typename AccFn, typename IO>
__device__ __forceinline__ void reg_run(float2* __restrict__ A, float2* __restrict__ B, int L,
                        const DevOp* __restrict__ ops, int b, int e,
                        MS& mc, AccFn&& acc_fn, IO& io) {
  // Wave mode: gradient partials count each amplitude once (replica lanes 0).
  const bool canon = !WAVE || ((threadIdx.x & 63u) >> (L - R)) == 0u;
  constexpr bool kLds = IsLdsIO<IO>::value;
  cf ra[1 << R];
  cf rb[1 << R];  // unused when !REV
  RegMap<R> m;
  bool have = false, swizzled = false, first_store = true;
  int slot = 0;
  for (int i = b; i < e; ++i) {
    DevOp op = load_op(ops + i);
    const float2* mats = op.type != kOpRemap ? mats_for(mc, op) : nullptr;
    if (op.type == kOpRemap) {
      if constexpr (!kLds) {
        if (!have) {
          make_map<R, WAVE>(m, op.cmask, L);
          for_reg_index<R>(m, io, [&](int r, uint64_t g) { io.load(g, ra[r], rb[r]); });
          have = true;
          continue;
        }
      }
      // A remap that leaves the wavefront-index qubits where they are moves
      // every amplitude within its own wavefront: wavefront w stores and
      // reloads exactly the (swizzled) LDS entries it read at the previous
      // remap, so it needs no workgroup barrier.  The first store of a run
      // still waits for the whole workgroup (entries last read in another
      // layout, or by other wavefronts on the previous tile).
      const uint32_t nmask = uint32_t(uniform(int(op.cmask)));
      const bool local = !WAVE && have && !first_store &&
                         wave_positions(m.mask, L) == wave_positions(nmask, L);
      if (have) {
        if (!local) esync<WAVE>();
        reg_store<R>(ra, A, m, true);
        if (REV) reg_store<R>(rb, B, m, true);
        swizzled = true;
        first_store = false;
      }
      if (local) esync<true>();
      else esync<WAVE>();
      make_map<R, WAVE>(m, nmask, L);
      reg_load<R>(ra, A, m, swizzled);
      if (REV) reg_load<R>(rb, B, m, swizzled);
      have = true;
      continue;
    }
    if (op.type == kOpDiag) {
      reg_diag_any<R, REV, REV>(ra, rb, op, mats + op.mat, m);
    } else if (REV && op.type == kOpParam) {
      double acc[2 * kMaxDeriv] = {0.0, 0.0, 0.0, 0.0};
      if (op.k == 1) param_dispatch1<R>(reg_slot(m, op.q[0]), ra, rb, op, mats, m, acc);
      else param_dispatch2<R>(reg_slot(m, op.q[0]), reg_slot(m, op.q[1]), ra, rb, op, mats, m, acc);
      if (!canon) acc[0] = acc[1] = acc[2] = acc[3] = 0.0;
      // The imaginary parts reach only the callers that use them (the
      // compiler drops them elsewhere).
      const double w0 = wave_sum(acc[0]), v0 = wave_sum(acc[kMaxDeriv]);
      if ((threadIdx.x & 63) == 0) acc_fn(slot, op.col[0], w0, v0);
      if (op.nderiv > 1) {
        const double w1 = wave_sum(acc[1]), v1 = wave_sum(acc[kMaxDeriv + 1]);
        if ((threadIdx.x & 63) == 0) acc_fn(slot + 1, op.col[1], w1, v1);
      }
      slot += op.nderiv;
    } else if (op.type == kOpChannel) {
      if constexpr (!REV && IsNoise<std::decay_t<AccFn>>::value)
        reg_channel<R>(reg_slot(m, op.q[0]), ra, op, mats, acc_fn);
    } else {
      reg_dense<R, REV, REV>(ra, rb, op, mats + op.mat, m);
    }
  }
  if (!have) return;
  if constexpr (kLds) {
    esync<WAVE>();
    reg_store<R>(ra, A, m, false);
    if (REV) reg_store<R>(rb, B, m, false);
    esync<WAVE>();
  } else {
    for_reg_index<R>(m, io, [&](int r, uint64_t g) { io.store(g, ra[r], rb[r]); });
  }
}

// Run ops [b, e) on a tile of L qubits with the register engine when the op
// list was planned for it (starts with a remap); returns false otherwise,
// before touching the tile.
// RMAX = 5 adds 14-qubit one-state tiles (32 registers per thread); only the
// streaming pass kernel's wide instantiation, whose plans use them, asks for it.
template <bool REV, int RMAX = 4, typename AccFn, typename IO>
__device__ __forceinline__ bool reg_run_any(float2* A, float2* B, int L, const DevOp* ops, int b,
                                            int e, MatCache& mc, AccFn&& acc_fn, IO& io) {
  if (b >= e) return false;
  if (uniform(int(ops[b].type)) != kOpRemap) return false;
  switch (RegBitsForTile(L, RMAX)) {
    case 3: reg_run<3, REV, false>(A, B, L, ops, b, e, mc, acc_fn, io); return true;
    case 4: reg_run<4, REV, false>(A, B, L, ops, b, e, mc, acc_fn, io); return true;
    case 5:
      if constexpr (RMAX >= 5 && !REV) {
        reg_run<5, REV, false>(A, B, L, ops, b, e, mc, acc_fn, io);
        return true;
      }
      return false;
    default: return false;
  }
}
template <bool REV, typename AccFn>
__device__ __forceinline__ bool reg_run_any(float2* A, float2* B, int L, const DevOp* ops, int b,
                                            int e, MatCache& mc, AccFn&& acc_fn) {
  LdsIO io;
  return reg_run_any<REV>(A, B, L, ops, b, e, mc, acc_fn, io);
}

// Wave mode: one wavefront runs a whole circuit of L = kWaveRegBits..
// kWaveMaxQubits qubits (ops planned with remaps for kWaveRegBits register
// qubits), A and B being the wavefront's own LDS buffers of 2^L amplitudes.
template <bool REV, typename AccFn, typename IO>
__device__ __forceinline__ void reg_run_wave(float2* A, float2* B, int L, const DevOp* ops, int b,
                                             int e, const float2* mats, AccFn&& acc_fn, IO& io) {
  if (b >= e) return;
  GlobalMats gm{mats};
  reg_run<kWaveRegBits, REV, true>(A, B, L, ops, b, e, gm, acc_fn, io);
}

}  // namespace
}  // namespace tfqx
