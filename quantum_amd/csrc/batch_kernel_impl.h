// Persistent LDS-resident circuit kernel for gfx950 (device code; instantiated
// per mode in batch_mode_*.hip so the modes compile in parallel).
//
// One workgroup (8 wave64) owns one circuit at a time; circuits are pulled from
// a device work queue in longest-first order, so thousands of circuits of
// uneven width and depth run in one launch.  The whole state (and, for the
// adjoint, lambda) stays in LDS for the full circuit.  Per mode:
//   expectation  fused forward ops (qsim::ApplyFusedGate,
//                tfq_simulate_expectation_op.cc:196-231), then a single-pass
//                XOR-mask reduction per Pauli string with FP64 partials
//                (replaces Copy + Pauli gates + RealInnerProduct, util_qsim.h:142-188)
//   adjoint      + lambda = sum_j g_j O_j psi in one pass (AccumulateOperators,
//                util_qsim.h:363-414) and the reverse sweep of
//                tfq_adj_grad_op.cc:214-276, with un-apply, derivative and inner
//                product fused into one LDS sweep per parameterised gate
//   state        copy-out with -2 padding (tfq_simulate_state_op.cc:152-170)
//   samples      FP64 in-place prefix scan of |psi|^2 + binary search per shot,
//                bits unpacked to int8 with -2 padding (tfq_simulate_samples_op.cc:157-188)
//   sampled exp. per term: Z-basis rotation, scan, shots, parity (util_qsim.h:199-270)
//   inner prod.  <psi_i | phi_ij> with FP64 accumulation (tfq_inner_product.cc:182-214)
#pragma once

#include <hip/hip_runtime.h>

#include "batch_kernels.h"
#include "lds_ops.h"
#include "reg_engine.h"

namespace tfqx {

namespace {

struct NoAcc {
  __device__ void operator()(int, int, double, double) const {}
};

// Run ops [b, e): f(op, M) with M the op's matrices (unitary or diagonal, then
// its derivatives, back to back) in the workgroup's LDS matrix window.  The
// window is refilled in one coalesced block load when an op's matrices fall
// outside it, so most ops start without waiting on global memory; the next
// op's descriptor is read while the current one runs.  Every op ends with a
// workgroup barrier.
template <typename F>
__device__ __forceinline__ void for_ops_cached(const DevOp* ops, int b, int e, MatCache& mc, F&& f) {
  if (b >= e) return;
  __syncthreads();
  DevOp nx = load_op(ops + b);
  for (int i = b; i < e; ++i) {
    DevOp op = nx;
    if (i + 1 < e) nx = load_op(ops + i + 1);
    const float2* W = mats_for(mc, op);
    f(op, W + op.mat);
    __syncthreads();
  }
}

// Forward ops [b, e) on an LDS tile of n qubits (register engine when planned).
template <int KMAX>
__device__ __forceinline__ void run_ops(float2* s, int n, const DevOp* ops, int b, int e, MatCache& mc) {
  if (reg_run_any<false>(s, s, n, ops, b, e, mc, NoAcc())) return;
  for_ops_cached(ops, b, e, mc,
                 [&](const DevOp& op, const float2* M) { apply_op<false, KMAX>(s, n, op, M); });
}

// Adjoint sweep of ops [b, e) on psi (A) and lambda (B), gradients into gacc.
template <int KMAX, bool CPLX = false>
__device__ __forceinline__ void run_reverse(float2* A, float2* B, int n, const DevOp* ops, int b, int e,
                                            double* gacc, MatCache& mc);

__device__ __forceinline__ void set_zero_state(float2* s, uint32_t N) {
  for (uint32_t x = threadIdx.x; x < N; x += kThreads) s[x] = make_float2(x == 0 ? 1.f : 0.f, 0.f);
  __syncthreads();
}

// In-place: float2 amplitudes -> FP64 inclusive prefix of |a|^2 (same bytes).
__device__ __forceinline__ double build_cdf(float2* s, uint32_t N, double* red) {
  double* cdf = reinterpret_cast<double*>(s);
  const uint32_t chunk = (N + kThreads - 1) / kThreads;
  const uint32_t b = threadIdx.x * chunk;
  const uint32_t e = min(N, b + chunk);
  double run = 0.0;
  for (uint32_t x = b; x < e; ++x) {
    const float2 v = s[x];
    run += double(v.x) * double(v.x) + double(v.y) * double(v.y);
    cdf[x] = run;
  }
  double total;
  const double off = block_exclusive_scan(run, red, &total);
  for (uint32_t x = b; x < e; ++x) cdf[x] += off;
  __syncthreads();
  return total;
}

// Smallest x with cdf[x] > target (clamped to N-1).
__device__ __forceinline__ uint32_t cdf_search(const double* cdf, uint32_t N, double target) {
  uint32_t lo = 0, hi = N - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cdf[mid] > target) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Pauli expectations of state S (LDS or global) for all slots of a circuit.
__device__ __forceinline__ void expectations(const float2* __restrict__ S, uint32_t N,
                                             const BatchArgs& a, const DevCircuit& c, int row,
                                             double* red) {
  const int tid = threadIdx.x;
  int t = uniform(c.term_begin);
  const int te = uniform(c.term_end);
  for (int j = 0; j < a.nslots; ++j) {
    double acc = 0.0;
    bool any = false;
    while (t < te && a.terms[t].slot == j) {
      const DevTerm T = a.terms[t];
      float part = 0.f;
      for (uint32_t x = tid; x < N; x += kThreads) {
        const float2 pr = cmul_conj_a(S[x ^ T.fmask], S[x]);
        const float v = T.re * pr.x - T.im * pr.y;
        part += (__popc(x & T.zmask) & 1) ? -v : v;
      }
      acc += double(part);
      ++t;
      any = true;
    }
    const double tot = any ? block_sum(acc, red) : 0.0;
    if (tid == 0)
      a.exp_out[size_t(row) * a.nslots + j] = float(tot + double(a.ident[size_t(row) * a.nslots + j]));
  }
}

// lambda = sum_t w_t P_t psi.
__device__ __forceinline__ void build_lambda(const float2* __restrict__ S, float2* __restrict__ Lm,
                                             uint32_t N, const BatchArgs& a, int lb, int le) {
  for (uint32_t y = threadIdx.x; y < N; y += kThreads) {
    float2 l = make_float2(0.f, 0.f);
    for (int t = lb; t < le; ++t) {
      const DevTerm T = a.terms[t];
      const uint32_t x = y ^ T.fmask;
      float2 w = cmul(make_float2(T.re, T.im), S[x]);
      if (__popc(x & T.zmask) & 1) w = make_float2(-w.x, -w.y);
      l = cadd(l, w);
    }
    Lm[y] = l;
  }
  __syncthreads();
}

// CPLX = false: gacc[col] += 2 Re<dG psi | lam> (adjoint gradient).
// CPLX = true: (gacc[2 col], gacc[2 col + 1]) += <dG psi | lam> (inner-product
// gradient).  Each wavefront adds its own partial with an FP64 atomic, so the
// step needs no workgroup barrier beyond the one that follows every op.
template <bool CPLX>
__device__ __forceinline__ void param_step(float2* A, float2* B, int n, const DevOp& op,
                                           const float2* M, double* gacc) {
  double acc[2 * kMaxDeriv] = {0.0, 0.0, 0.0, 0.0};
  switch (op.k) {
    case 1: apply_param<1>(A, B, n, op, M, acc); break;
    default: apply_param<2>(A, B, n, op, M, acc); break;
  }
  const bool lead = (threadIdx.x & 63) == 0;
#pragma unroll
  for (int d = 0; d < kMaxDeriv; ++d) {
    if (d >= op.nderiv) break;
    const double tot = wave_sum(acc[d]);
    if (CPLX) {
      const double tim = wave_sum(acc[kMaxDeriv + d]);
      if (lead) {
        atomicAdd(&gacc[2 * op.col[d]], tot);
        atomicAdd(&gacc[2 * op.col[d] + 1], tim);
      }
    } else if (lead) {
      atomicAdd(&gacc[op.col[d]], 2.0 * tot);
    }
  }
}

template <int KMAX, bool CPLX>
__device__ __forceinline__ void run_reverse(float2* A, float2* B, int n, const DevOp* ops, int b, int e,
                                            double* gacc, MatCache& mc) {
  auto acc = [gacc](int, int col, double w, double v) {
    if (CPLX) {
      atomicAdd(&gacc[2 * col], w);
      atomicAdd(&gacc[2 * col + 1], v);
    } else {
      atomicAdd(&gacc[col], 2.0 * w);
    }
  };
  if (reg_run_any<true>(A, B, n, ops, b, e, mc, acc)) return;
  for_ops_cached(ops, b, e, mc, [&](const DevOp& op, const float2* M) {
    if (op.type == kOpParam) {
      param_step<CPLX>(A, B, n, op, M, gacc);
    } else {
      apply_op<true, KMAX>(A, n, op, M);
      apply_op<true, KMAX>(B, n, op, M);
    }
  });
}

// ------------------------------------------------------------------ wave mode
// One wavefront per circuit of kWaveRegBits..kWaveMaxQubits qubits: the
// state stays in the lanes' registers (register engine, wave-synchronised
// remaps through the wavefront's own LDS buffers), expectations and lambda
// are wavefront reductions, and no workgroup barrier is involved.
__device__ __forceinline__ void wave_expectations(const float2* __restrict__ S, uint32_t N,
                                                  const BatchArgs& a, const DevCircuit& c,
                                                  int row) {
  const uint32_t lane = threadIdx.x & 63u;
  int t = uniform(c.term_begin);
  const int te = uniform(c.term_end);
  for (int j = 0; j < a.nslots; ++j) {
    double acc = 0.0;
    while (t < te && uniform(a.terms[t].slot) == j) {
      const DevTerm T = a.terms[t];
      float part = 0.f;
      for (uint32_t x = lane; x < N; x += 64u) {
        const float2 pr = cmul_conj_a(S[x ^ T.fmask], S[x]);
        const float v = T.re * pr.x - T.im * pr.y;
        part += (__popc(x & T.zmask) & 1) ? -v : v;
      }
      acc += double(part);
      ++t;
    }
    const double tot = wave_sum(acc);
    if (lane == 0)
      a.exp_out[size_t(row) * a.nslots + j] = float(tot + double(a.ident[size_t(row) * a.nslots + j]));
  }
}

__device__ __forceinline__ void wave_lambda(const float2* __restrict__ S, float2* __restrict__ Lm,
                                            uint32_t N, const BatchArgs& a, int lb, int le) {
  for (uint32_t y = threadIdx.x & 63u; y < N; y += 64u) {
    float2 l = make_float2(0.f, 0.f);
    for (int t = lb; t < le; ++t) {
      const DevTerm T = a.terms[t];
      const uint32_t x = y ^ T.fmask;
      float2 w = cmul(make_float2(T.re, T.im), S[x]);
      if (__popc(x & T.zmask) & 1) w = make_float2(-w.x, -w.y);
      l = cadd(l, w);
    }
    Lm[y] = l;
  }
  esync<true>();
}

static_assert(kLdsMaxQubitsOneBuffer >= kWaveMaxQubits + 4 &&
                  kLdsMaxQubitsTwoBuffers + 1 >= kWaveMaxQubits + 4,
              "wave mode needs two 2^kWaveMaxQubits buffers per wavefront (8 wavefronts)");

// Circuit ci (wave mode) on this wavefront with LDS buffers P and Q (2^n each).
template <int MODE>
__device__ __forceinline__ void wave_circuit(const BatchArgs& a, int ci, float2* P, float2* Q) {
  const DevCircuit c = a.circuits[ci];
  const int n = uniform(c.n);
  const int row = uniform(c.row);
  const uint32_t N = 1u << n;
  const uint32_t lane = threadIdx.x & 63u;
  // Forward from |0>, final state to P in the natural layout.
  auto fio = tile_io([](uint32_t x, cf& va, cf&) { va = cf{x == 0 ? 1.f : 0.f, 0.f}; },
                     [P](uint32_t x, const cf& va, const cf&) { P[x] = make_float2(va.x, va.y); });
  reg_run_wave<false>(P, P, n, a.ops, uniform(c.fwd_begin), uniform(c.fwd_end), a.mats, NoAcc(),
                      fio);
  esync<true>();
  if (MODE == kModeExpectation || (MODE == kModeAdjoint && a.exp_out))
    wave_expectations(P, N, a, c, row);
  if (MODE == kModeAdjoint) {
    double* gacc = a.grad_acc + size_t(row) * a.nsym;
    for (int j = int(lane); j < a.nsym; j += 64) gacc[j] = 0.0;
    __threadfence();
    wave_lambda(P, Q, N, a, uniform(c.lam_begin), uniform(c.lam_end));
    auto rio = tile_io(
        [P, Q](uint32_t x, cf& va, cf& vb) {
          const float2 u = P[x], v = Q[x];
          va = cf{u.x, u.y};
          vb = cf{v.x, v.y};
        },
        [](uint32_t, const cf&, const cf&) {});
    auto acc = [gacc](int, int col, double w, double) { atomicAdd(&gacc[col], 2.0 * w); };
    reg_run_wave<true>(P, Q, n, a.ops, uniform(c.rev_begin), uniform(c.rev_end), a.mats, acc, rio);
    __threadfence();
    for (int j = int(lane); j < a.nsym; j += 64)
      a.grad_out[size_t(row) * a.nsym + j] =
          float(__hip_atomic_load(&gacc[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  esync<true>();
}

// Tiled forward: the state lives in global memory G; each pass streams its
// tiles through LDS buffer A.  The first pass starts from |0>.
template <int KMAX, int RMAX = 4>
__device__ __forceinline__ void tiled_forward(float2* G, int n, const BatchArgs& a, int pb, int pe, float2* A,
                              uint64_t* tab, MatCache& mc) {
  const int tid = threadIdx.x;
  bool first = true;
  for (int p = pb; p < pe; ++p) {
    const DevPass ps = a.passes[p];
    const bool use_tab = ps.L - ps.nlow <= 8;
    build_tile_table(ps, tab);
    __syncthreads();
    const uint32_t T = 1u << ps.L;
    const uint32_t ntiles = 1u << (n - ps.L);
    for (uint32_t t = 0; t < ntiles; ++t) {
      const uint64_t base = tile_base(t, ps.tilemask, n);
      auto gidx = [&](uint32_t j) -> uint64_t {
        return use_tab ? tile_index_tab(base, j, ps.nlow, tab) : tile_index(base, j, ps);
      };
      // Register engine: the tile goes straight between G and the registers.
      auto io = tile_io_indexed(
          gidx,
          [&](uint64_t g, cf& va, cf&) {
            if (first) {
              va = cf{g == 0 ? 1.f : 0.f, 0.f};
            } else {
              const float2 v = G[g];
              va = cf{v.x, v.y};
            }
          },
          [&](uint64_t g, const cf& va, const cf&) { G[g] = make_float2(va.x, va.y); });
      if (reg_run_any<false, RMAX>(A, A, ps.L, a.ops, ps.op_begin, ps.op_end, mc, NoAcc(), io))
        continue;
      if (first) {
        for (uint32_t j = tid; j < T; j += kThreads) A[j] = make_float2(gidx(j) == 0 ? 1.f : 0.f, 0.f);
      } else {
        tile_gather(A, G, T, gidx);
      }
      __syncthreads();
      run_ops<KMAX>(A, ps.L, a.ops, ps.op_begin, ps.op_end, mc);
      tile_scatter(G, A, T, gidx);
      __syncthreads();
    }
    // The pass's stores to G precede the next pass's loads, and the tile table
    // is rebuilt only after every wave has finished with it.
    __syncthreads();
    first = false;
  }
  if (first) {
    for (uint32_t x = tid; x < (1u << n); x += kThreads) G[x] = make_float2(x == 0 ? 1.f : 0.f, 0.f);
    __syncthreads();
  }
}

// Tiled adjoint sweep on psi (G) and lambda (Lg) with LDS tiles A and B.
template <int KMAX>
__device__ __forceinline__ void tiled_reverse(float2* G, float2* Lg, int n, const BatchArgs& a, int pb, int pe,
                              float2* A, float2* B, double* gacc, uint64_t* tab,
                              MatCache& mc) {
  const int tid = threadIdx.x;
  for (int p = pb; p < pe; ++p) {
    const DevPass ps = a.passes[p];
    const bool use_tab = ps.L - ps.nlow <= 8;
    build_tile_table(ps, tab);
    __syncthreads();
    const uint32_t T = 1u << ps.L;
    const uint32_t ntiles = 1u << (n - ps.L);
    for (uint32_t t = 0; t < ntiles; ++t) {
      const uint64_t base = tile_base(t, ps.tilemask, n);
      auto gidx = [&](uint32_t j) -> uint64_t {
        return use_tab ? tile_index_tab(base, j, ps.nlow, tab) : tile_index(base, j, ps);
      };
      auto io = tile_io_indexed(
          gidx,
          [&](uint64_t g, cf& va, cf& vb) {
            const float2 u = G[g], v = Lg[g];
            va = cf{u.x, u.y};
            vb = cf{v.x, v.y};
          },
          [&](uint64_t g, const cf& va, const cf& vb) {
            G[g] = make_float2(va.x, va.y);
            Lg[g] = make_float2(vb.x, vb.y);
          });
      auto acc = [gacc](int, int col, double w, double) { atomicAdd(&gacc[col], 2.0 * w); };
      if (reg_run_any<true>(A, B, ps.L, a.ops, ps.op_begin, ps.op_end, mc, acc, io)) continue;
      tile_gather2(A, B, G, Lg, T, gidx);
      __syncthreads();
      run_reverse<KMAX>(A, B, ps.L, a.ops, ps.op_begin, ps.op_end, gacc, mc);
      tile_scatter(G, A, T, gidx);
      tile_scatter(Lg, B, T, gidx);
      __syncthreads();
    }
    __syncthreads();
  }
}

// Modes that keep one state in LDS (at most 64 KiB up to 13 qubits) cap their
// VGPRs at 128 so two workgroups share a CU (four wavefronts per SIMD): on
// 11-13-qubit expectation batches that runs 1.4-1.5x faster than 190 VGPRs
// at one workgroup per CU, the spills it costs included.
constexpr int BatchMinWaves(int mode) {
  return (mode == kModeExpectation || mode == kModeState || mode == kModeSamples) ? 4 : 1;
}

template <int MODE, int KMAX>
__global__ __launch_bounds__(kThreads, BatchMinWaves(MODE)) void batch_kernel(BatchArgs a) {
  extern __shared__ float2 lds[];
  __shared__ double red[kWaves];
  __shared__ int s_ci;
  __shared__ uint64_t tab[kTileTable];
  constexpr int kWindow = MODE == kModeAdjoint ? kMatCacheWide : kMatCache;
  __shared__ float2 mcache[kWindow];
  MatCache mc = mat_cache(mcache, a.mats, kWindow);
  __shared__ float2 rotS[8];
  if (MODE == kModeSampledExpectation && threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) rotS[i] = a.rot[i];
  }
  float2* psi = lds;
  float2* buf2 = lds + a.lds_amps;
  const int tid = threadIdx.x;
  float2* gpsi = a.scratch + size_t(blockIdx.x) * a.scratch_stride;
  float2* glam = gpsi + (a.scratch_stride >> 1);

  while (true) {
    if (tid == 0) s_ci = atomicAdd(a.queue, 1);
    __syncthreads();
    const int ci = uniform(s_ci);
    __syncthreads();
    if (ci >= a.nblock || ci >= a.ncirc) break;
    const DevCircuit c = a.circuits[ci];
    const int n = uniform(c.n);
    const int row = uniform(c.row);
    const uint32_t N = 1u << n;
    const bool empty = (c.flags & kFlagEmpty) != 0;
    const bool tiled = (c.flags & kFlagTiled) != 0;

    if (empty) {
      if (MODE == kModeExpectation || MODE == kModeSampledExpectation ||
          (MODE == kModeAdjoint && a.exp_out))
        for (int j = tid; j < a.nslots; j += kThreads) a.exp_out[size_t(row) * a.nslots + j] = -2.0f;
      if (MODE == kModeAdjoint)
        for (int j = tid; j < a.nsym; j += kThreads) a.grad_out[size_t(row) * a.nsym + j] = 0.0f;
      if (MODE == kModeInnerProduct)
        for (int j = tid; j < a.nslots; j += kThreads)
          a.ip_out[size_t(row) * a.nslots + j] = make_float2(1.f, 0.f);
      if (MODE == kModeInnerProductGrad)
        for (int j = tid; j < a.nsym; j += kThreads)
          a.ip_out[size_t(row) * a.nsym + j] = make_float2(0.f, 0.f);
      if (MODE == kModeState)
        for (long long x = tid; x < a.state_stride; x += kThreads)
          a.state_out[size_t(row) * a.state_stride + x] = make_float2(x == 0 ? 1.f : -2.f, 0.f);
      if (MODE == kModeSamples)
        for (long long x = tid; x < (long long)a.nshots * a.max_nq; x += kThreads)
          a.samples_out[size_t(row) * a.nshots * a.max_nq + x] = -2;
      continue;
    }

    if (MODE == kModeInnerProductGrad) {
      // lambda = sum_j g_j |phi_j> in buf2 (AccumulateFusedCircuits,
      // util_qsim.h:423-442), each phi_j simulated in the psi buffer first.
      float2* lam = buf2;
      for (uint32_t x = tid; x < N; x += kThreads) lam[x] = make_float2(0.f, 0.f);
      for (int o = uniform(c.other_begin); o < c.other_end; ++o) {
        const DevOther ot = a.others[o];
        set_zero_state(psi, N);
        run_ops<KMAX>(psi, n, a.ops, uniform(ot.op_begin), uniform(ot.op_end), mc);
        const float w = ot.weight;
        for (uint32_t x = tid; x < N; x += kThreads) {
          const float2 p = psi[x];
          lam[x] = make_float2(lam[x].x + w * p.x, lam[x].y + w * p.y);
        }
        __syncthreads();
      }
    }
    if (tiled) {
      // The adjoint's forward carries one state: its passes may use 14-qubit
      // tiles (the LDS of the two 13-qubit reverse tiles) and five register
      // qubits.
      tiled_forward<KMAX, MODE == kModeAdjoint ? 5 : 4>(gpsi, n, a, uniform(c.fwd_pass_begin),
                                                        uniform(c.fwd_pass_end), psi, tab, mc);
    } else {
      set_zero_state(psi, N);
      run_ops<KMAX>(psi, n, a.ops, uniform(c.fwd_begin), uniform(c.fwd_end), mc);
    }

    if (MODE == kModeExpectation || (MODE == kModeAdjoint && a.exp_out)) {
      if (tiled) expectations(gpsi, N, a, c, row, red);
      else expectations(psi, N, a, c, row, red);
    }

    if (MODE == kModeAdjoint) {
      double* gacc = a.grad_acc + size_t(row) * a.nsym;
      for (int j = tid; j < a.nsym; j += kThreads) gacc[j] = 0.0;
      __threadfence();  // the zeros reach L2 before any wavefront's atomics
      const int lb = uniform(c.lam_begin), le = uniform(c.lam_end);
      if (tiled) {
        build_lambda(gpsi, glam, N, a, lb, le);
        tiled_reverse<KMAX>(gpsi, glam, n, a, uniform(c.rev_pass_begin), uniform(c.rev_pass_end), psi,
                      buf2, gacc, tab, mc);
      } else {
        float2* lam = buf2;
        build_lambda(psi, lam, N, a, lb, le);
        run_reverse<KMAX>(psi, lam, n, a.ops, uniform(c.rev_begin), uniform(c.rev_end), gacc, mc);
      }
      // Gradients arrive by atomics; read them back from L2.
      __threadfence();
      __syncthreads();
      for (int j = tid; j < a.nsym; j += kThreads)
        a.grad_out[size_t(row) * a.nsym + j] =
            float(__hip_atomic_load(&gacc[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }

    if (MODE == kModeInnerProductGrad) {
      // Adjoint sweep with complex accumulation (tfq_inner_product_grad.cc:259-314).
      double* gacc = a.grad_acc + size_t(row) * a.nsym * 2;
      for (int j = tid; j < 2 * a.nsym; j += kThreads) gacc[j] = 0.0;
      __threadfence();  // the zeros reach L2 before any wavefront's atomics
      __syncthreads();
      run_reverse<KMAX, true>(psi, buf2, n, a.ops, uniform(c.rev_begin), uniform(c.rev_end), gacc,
                              mc);
      __threadfence();
      __syncthreads();
      for (int j = tid; j < a.nsym; j += kThreads)
        a.ip_out[size_t(row) * a.nsym + j] =
            make_float2(float(__hip_atomic_load(&gacc[2 * j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)),
                        float(__hip_atomic_load(&gacc[2 * j + 1], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT)));
    }

    if (MODE == kModeState) {
      float2* out = a.state_out + size_t(row) * a.state_stride;
      const float2* S = tiled ? gpsi : psi;
      for (long long x = tid; x < a.state_stride; x += kThreads)
        out[x] = x < (long long)N ? S[x] : make_float2(-2.f, 0.f);
    }

    if (MODE == kModeSamples) {
      const double total = build_cdf(psi, N, red);
      const double* cdf = reinterpret_cast<const double*>(psi);
      signed char* out = a.samples_out + size_t(row) * a.nshots * a.max_nq;
      const int mq = a.max_nq;
      for (int k = tid; k < a.nshots; k += kThreads) {
        const double u = philox_uniform(a.seed, (unsigned long long)k, uint32_t(row), 0xFFFFFFFFu);
        const uint32_t x = cdf_search(cdf, N, u * total);
        signed char* o = out + size_t(k) * mq;
        for (int q = 0; q < mq; ++q)
          o[mq - 1 - q] = q < n ? (signed char)((x >> q) & 1u) : (signed char)-2;
      }
    }

    if (MODE == kModeSampledExpectation) {
      int t = uniform(c.term_begin);
      const int te = uniform(c.term_end);
      float2* work = buf2;
      for (int j = 0; j < a.nslots; ++j) {
        double val = 0.0;
        const int ns = a.num_samples[size_t(row) * a.nslots + j];
        while (t < te && a.terms[t].slot == j) {
          const DevTerm T = a.terms[t];
          for (uint32_t x = tid; x < N; x += kThreads) work[x] = psi[x];
          __syncthreads();
          // Z-basis change: X -> Y^-0.5, Y -> X^0.5 (circuit_parser_qsim.cc:897-945).
          for (int q = 0; q < n; ++q) {
            if (!((T.fmask >> q) & 1u)) continue;
            const bool is_y = (T.zmask >> q) & 1u;
            apply_1q(work, n, q, is_y ? &rotS[4] : &rotS[0]);
            __syncthreads();
          }
          const double total = build_cdf(work, N, red);
          const double* cdf = reinterpret_cast<const double*>(work);
          const uint32_t pmask = T.fmask | T.zmask;
          int par = 0;
          for (int k = tid; k < ns; k += kThreads) {
            const double u = philox_uniform(a.seed, (unsigned long long)k, uint32_t(row),
                                            (uint32_t(j) << 16) | uint32_t(T.pad));
            const uint32_t x = cdf_search(cdf, N, u * total);
            par += (__popc(x & pmask) & 1) ? -1 : 1;
          }
          const double tot = block_sum(double(par), red);
          val += double(T.re) * tot / double(ns);
          ++t;
        }
        if (tid == 0)
          a.exp_out[size_t(row) * a.nslots + j] = float(val + double(a.ident[size_t(row) * a.nslots + j]));
        __syncthreads();
      }
    }

    if (MODE == kModeInnerProduct) {
      float2* phi = buf2;
      for (int o = uniform(c.other_begin); o < c.other_end; ++o) {
        const DevOther ot = a.others[o];
        set_zero_state(phi, N);
        run_ops<KMAX>(phi, n, a.ops, uniform(ot.op_begin), uniform(ot.op_end), mc);
        double re = 0.0, im = 0.0;
        for (uint32_t x = tid; x < N; x += kThreads) {
          const float2 p = cmul_conj_a(psi[x], phi[x]);
          re += double(p.x);
          im += double(p.y);
        }
        re = block_sum(re, red);
        im = block_sum(im, red);
        if (tid == 0) a.ip_out[size_t(row) * a.nslots + ot.col] = make_float2(float(re), float(im));
        __syncthreads();
      }
    }
    __syncthreads();
  }

  // Wave mode: each wavefront pulls small circuits of its own.
  if constexpr (MODE == kModeExpectation || MODE == kModeAdjoint) {
    if (a.nblock < a.ncirc) {
      const int nwave = a.ncirc - a.nblock;
      float2* P = lds + size_t(tid >> 6) * (2u << kWaveMaxQubits);
      float2* Q = P + (1u << kWaveMaxQubits);
      while (true) {
        int w = 0;
        if ((tid & 63) == 0) w = atomicAdd(a.wqueue, 1);
        w = uniform(__shfl(w, 0, 64));
        if (w >= nwave) break;
        wave_circuit<MODE>(a, a.nblock + w, P, Q);
      }
    }
    __syncthreads();
  }

  // The last workgroup out resets the queues so the launch can be replayed
  // (e.g. from a hipGraph) without a memset node.
  if (tid == 0) {
    __threadfence();
    const int done = atomicAdd(a.done, 1);
    if (done == int(gridDim.x) - 1) {
      atomicExch(a.queue, 0);
      atomicExch(a.wqueue, 0);
      atomicExch(a.done, 0);
    }
  }
}

}  // namespace
}  // namespace tfqx
