// HBM-streaming kernels for states wider than one LDS tile (gfx950).
//
//   pass_kernel      one pass of fused ops: every 2^L tile of the state is
//                    loaded into LDS (runs of >= 64 contiguous amplitudes),
//                    the pass's ops are applied there, and the tile is stored
//                    back: 16 bytes of HBM traffic per amplitude per pass.
//                    The two-buffer variant runs the adjoint reverse sweep on
//                    psi and lambda together and accumulates FP64 gradient
//                    partials per workgroup.
//   expect_kernel    XOR-mask Pauli reduction over the state, FP64 partials.
//   lambda_kernel    lambda = sum_t w_t P_t psi (one streaming pass).
//   ip_kernel        <psi|phi> with FP64 partials.
//   prob_blocks / shots_*  sampling without a full CDF: one coalesced read of
//                    the state gives, per 4096-amplitude block, the FP64
//                    inclusive prefix of its 128 sub-block masses (32
//                    amplitudes each) and the block total; a scan of the block
//                    totals follows.  Each shot draws a Philox uniform,
//                    binary-searches the blocks, then the block's sub-blocks,
//                    and scans the 32 amplitudes of one sub-block.
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <string>

#include "large.h"
#include "lds_ops.h"
#include "reg_engine.h"

namespace tfqx {
namespace {

constexpr int kMaxPassDerivs = 64;
constexpr int kCdfBlockLog = 12;
constexpr int kSubLog = 5;
constexpr int kMaxSub = 1 << (kCdfBlockLog - kSubLog);
// Terms of one measurement basis sampled by one launch (blockIdx.y).
constexpr int kParityJobs = 64;

// One term of a sampled expectation: its RNG stream, shot count and parity
// mask (X, Y and Z qubits after the basis change).
struct ParityJob {
  uint32_t s_lo, s_hi;
  int32_t nshots, pad;
  uint64_t pmask;
};

struct PassArgs {
  const float2* src;
  float2* dst;
  float2* psi;
  float2* lam;
  const DevOp* ops;
  const float2* mats;
  double* partials;
  DevPass pass;
  int n;
  int init;
};

// One-state passes use 64 KiB of LDS for 13-qubit tiles, so two workgroups
// share a CU when each wave keeps to 128 VGPRs (4 waves per SIMD).
// CPLX (two-buffer passes only): each derivative also accumulates the
// imaginary part of <dG psi|lambda>, written to the partial slot after the
// real part's.
// WIDE: one-state passes on 14-qubit tiles (128 KiB of LDS, one workgroup
// per CU) with the five-register-qubit engine and the VGPR budget of two
// waves per SIMD.
template <bool TWO, bool CPLX = false, bool WIDE = false>
__global__ __launch_bounds__(kThreads, (TWO || WIDE) ? 2 : 4) void pass_kernel(PassArgs a) {
  constexpr int kPacc = TWO ? (CPLX ? 2 : 1) * kMaxPassDerivs : 1;
  extern __shared__ float2 lds[];
  __shared__ double red[kWaves];
  __shared__ double pacc[kPacc];
  __shared__ float2 mbuf[2 * kMatSlot];
  __shared__ uint64_t tab[kTileTable];
  __shared__ float2 mcache[kMatCache];
  MatCache mc = mat_cache(mcache, a.mats);
  const DevPass ps = a.pass;
  const int L = ps.L;
  const uint32_t T = 1u << L;
  float2* A = lds;
  float2* Bf = lds + T;
  const int tid = threadIdx.x;
  const uint64_t ntiles = uint64_t(1) << (a.n - L);
  const bool use_tab = ps.L - ps.nlow <= 8;
  const int nlow = ps.nlow;
  build_tile_table(ps, tab);
  if (TWO)
    for (int k = tid; k < kPacc; k += kThreads) pacc[k] = 0.0;
  __syncthreads();
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint64_t base = tile_base(t, ps.tilemask, a.n);
    auto gidx = [&](uint32_t j) -> uint64_t {
      return use_tab ? tile_index_tab(base, j, nlow, tab) : tile_index(base, j, ps);
    };
    int pl = 0;
    auto acc = [](int slot, int, double w, double v) {
      atomicAdd(&pacc[slot], w);
      if (CPLX) atomicAdd(&pacc[kMaxPassDerivs + slot], v);
    };
    // Register engine: the tile goes straight between memory and registers.
    auto io = tile_io_indexed(
        gidx,
        [&](uint64_t g, cf& va, cf& vb) {
          if (TWO) {
            const float2 u = a.psi[g], v = a.lam[g];
            va = cf{u.x, u.y};
            vb = cf{v.x, v.y};
          } else if (a.init) {
            va = cf{g == 0 ? 1.f : 0.f, 0.f};
          } else {
            const float2 u = a.src[g];
            va = cf{u.x, u.y};
          }
        },
        [&](uint64_t g, const cf& va, const cf& vb) {
          if (TWO) {
            a.psi[g] = make_float2(va.x, va.y);
            a.lam[g] = make_float2(vb.x, vb.y);
          } else {
            a.dst[g] = make_float2(va.x, va.y);
          }
        });
    if (reg_run_any<TWO, WIDE ? 5 : 4>(A, Bf, L, a.ops, ps.op_begin, ps.op_end, mc, acc, io))
      continue;
    if (TWO) {
      tile_gather2(A, Bf, a.psi, a.lam, T, gidx);
    } else if (a.init) {
      for (uint32_t j = tid; j < T; j += kThreads) A[j] = make_float2(gidx(j) == 0 ? 1.f : 0.f, 0.f);
    } else {
      tile_gather(A, a.src, T, gidx);
    }
    __syncthreads();
    for_ops(a.ops, ps.op_begin, ps.op_end, a.mats, mbuf,
            [&](const DevOp& op, const float2* M) {
              if (!TWO) {
                apply_op<false>(A, L, op, M);
              } else if (op.type == kOpParam) {
                double acc[2 * kMaxDeriv] = {0.0, 0.0, 0.0, 0.0};
                if (op.k == 1) apply_param<1>(A, Bf, L, op, M, acc);
                else apply_param<2>(A, Bf, L, op, M, acc);
#pragma unroll
                for (int d = 0; d < kMaxDeriv; ++d) {
                  if (d >= op.nderiv) break;
                  const double tot = block_sum(acc[d], red);
                  if (tid == 0) pacc[pl + d] += tot;
                  if (CPLX) {
                    const double im = block_sum(acc[kMaxDeriv + d], red);
                    if (tid == 0) pacc[kMaxPassDerivs + pl + d] += im;
                  }
                }
                pl += op.nderiv;
              } else {
                apply_op<true>(A, L, op, M);
                apply_op<true>(Bf, L, op, M);
              }
            });
    if (TWO) {
      tile_scatter(a.psi, A, T, gidx);
      tile_scatter(a.lam, Bf, T, gidx);
    } else {
      tile_scatter(a.dst, A, T, gidx);
    }
    __syncthreads();
  }
  // Register-engine tiles end without a barrier: wait for every wave's
  // gradient partials before reading them.
  __syncthreads();
  if (TWO && tid == 0) {
    int pl = 0;
    for (int i = ps.op_begin; i < ps.op_end; ++i) {
      const DevOp op = load_op(a.ops + i);
      if (op.type != kOpParam) continue;
      for (int d = 0; d < op.nderiv; ++d) {
        a.partials[size_t(op.col[d]) * gridDim.x + blockIdx.x] = pacc[pl + d];
        if (CPLX)
          a.partials[size_t(op.col[d] + 1) * gridDim.x + blockIdx.x] = pacc[kMaxPassDerivs + pl + d];
      }
      pl += op.nderiv;
    }
  }
}

// ------------------------------------------------------------------ dense passes
// One fused block of K = 5 or 6 qubits applied to the whole state as a dense
// contraction on FP32 MFMA.  In real form the block is
//   [Yr; Yi] = [[Gr, -Gi]; [Gi, Gr]] [Xr; Xi],
// a (2D x 2D) . (2D x C) product per tile (D = 2^K target amplitudes, C =
// 2^(kDenseTile - K) columns: the other tile qubits), run as 16x16 output
// tiles of v_mfma_f32_16x16x4f32 (A: lane l holds row l%16, k-row l/16; B:
// k-row l/16, column l%16; D: rows 4(l/16)..+3 of column l%16).  There are 32
// output tiles; each wavefront owns four with one row tile, whose slice of the
// real matrix (2D/4 values per lane) stays in registers for the launch.  A
// tile moves global -> LDS -> MFMA operands -> LDS -> global; the LDS index of
// (target row r, column c) is rowoff[r] | coloff[c].  Each thread's share of
// the tile two ahead is loaded into registers while the current tile's MFMAs
// run, and its tile offsets are computed once per launch.  K = 5 fits 128
// VGPRs, so two workgroups share a CU (four tiles' loads in flight per CU);
// K = 6 holds a 32-value matrix slice and runs one workgroup per CU.
struct DenseArgs {
  float2* psi;
  const float2* G;  // D x D complex, row-major, index bit i <-> pos[i]
  DevPass pass;
  int n;
  int pos[kMaxFuseWide];
};

template <int K>
__global__ __launch_bounds__(kThreads, K == 5 ? 4 : 2) void dense_kernel(DenseArgs a) {
  constexpr int D = 1 << K, D2 = 2 * D, L = kDenseTile, C = 1 << (L - K);
  constexpr int RT = D2 / 16, CT = C / 16, KS = D2 / 4, TPW = 4;
  constexpr int PER = (1 << L) / kThreads;  // amplitudes each thread moves per tile
  static_assert(RT * CT == TPW * kWaves, "32 output tiles, four per wavefront");
  extern __shared__ float2 lds[];
  __shared__ uint32_t rowoff[D], coloff[C];
  __shared__ uint64_t tab[kTileTable];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int rt = (w * TPW) / CT, ct0 = (w * TPW) % CT;
  if (tid < D) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) o |= uint32_t((tid >> i) & 1) << a.pos[i];
    rowoff[tid] = o;
  }
  if (tid < C) {
    uint32_t tmask = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) tmask |= 1u << a.pos[i];
    uint32_t o = 0;
    int src = 0;
    for (int b = 0; b < L; ++b) {
      if ((tmask >> b) & 1u) continue;
      o |= uint32_t((tid >> src) & 1) << b;
      ++src;
    }
    coloff[tid] = o;
  }
  const DevPass ps = a.pass;
  build_tile_table(ps, tab);
  // This lane's A operands: real-form rows rt*16 + lane%16, k-rows 4s + lane/16.
  float areg[KS];
  {
    const int R = rt * 16 + (lane & 15), r = R & (D - 1);
    const bool rim = R >= D;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int kc = 4 * s + (lane >> 4), c = kc & (D - 1);
      const bool kim = kc >= D;
      const float2 g = a.G[r * D + c];
      areg[s] = !rim ? (!kim ? g.x : -g.y) : (!kim ? g.y : g.x);
    }
  }
  __syncthreads();
  // Tile-invariant part of the global index of this thread's amplitudes
  // j = tid + kThreads * i (tile index bits deposited onto the tile's global
  // bits); a tile adds its base, whose bits are all outside the tile.
  const bool use_tab = ps.L - ps.nlow <= 8;
  uint32_t off[PER];  // within-state offsets: states of up to 32 qubits
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const uint32_t j = uint32_t(tid + kThreads * i);
    off[i] = uint32_t(use_tab ? tile_index_tab(0, j, ps.nlow, tab) : tile_index(0, j, ps));
  }
  const uint64_t ntiles = uint64_t(1) << (a.n - L);
  float* Lf = reinterpret_cast<float*>(lds);
  const int cl = lane & 15, kq = lane >> 4;
  uint32_t co[TPW];
#pragma unroll
  for (int j = 0; j < TPW; ++j) co[j] = coloff[(ct0 + j) * 16 + cl];
  // Row offsets of this lane's k-rows 4s + kq (s < D/4): the real part of a
  // loaded amplitude is the operand of k-step s, its imaginary part that of
  // k-step s + D/4.
  uint32_t ro[D / 4];
#pragma unroll
  for (int s = 0; s < D / 4; ++s) ro[s] = rowoff[4 * s + kq];
  // Tile t's body: its amplitudes (loaded earlier into `pre`) go to LDS, the
  // loads of tile `tn` (two tiles ahead) are issued into `pre`, the MFMAs run,
  // and the result goes back to global memory.
  auto body = [&](uint64_t t, uint64_t tn, float2 (&pre)[PER]) {
    const uint64_t base = tile_base(t, ps.tilemask, a.n);
#pragma unroll
    for (int i = 0; i < PER; ++i) lds[tid + kThreads * i] = pre[i];
    __syncthreads();
    if (tn < ntiles) {
      const uint64_t nb = tile_base(tn, ps.tilemask, a.n);
#pragma unroll
      for (int i = 0; i < PER; ++i) pre[i] = a.psi[nb | off[i]];
    }
    mf4 acc[TPW];
#pragma unroll
    for (int j = 0; j < TPW; ++j) acc[j] = mf4{0.f, 0.f, 0.f, 0.f};
    // One column tile at a time: its operands are loaded before the MFMAs
    // that use them.
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
      float2 bv[D / 4];
#pragma unroll
      for (int s = 0; s < D / 4; ++s) bv[s] = lds[ro[s] | co[j]];
#pragma unroll
      for (int s = 0; s < D / 4; ++s)
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(areg[s], bv[s].x, acc[j], 0, 0, 0);
#pragma unroll
      for (int s = 0; s < D / 4; ++s)
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(areg[s + D / 4], bv[s].y, acc[j], 0, 0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int R = rt * 16 + 4 * kq + i;
      const uint32_t rw = rowoff[R & (D - 1)];
      const uint32_t part = R >= D ? 1u : 0u;
#pragma unroll
      for (int j = 0; j < TPW; ++j) Lf[2u * (rw | co[j]) + part] = acc[j][i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) a.psi[base | off[i]] = lds[tid + kThreads * i];
    // Every thread has read its LDS values before the next tile overwrites them.
    __syncthreads();
  };
  // Two tiles' loads in flight per workgroup: this workgroup's tiles alternate
  // between the register sets preA and preB.
  const uint64_t G = gridDim.x;
  float2 preA[PER], preB[PER];
  for (int h = 0; h < 2; ++h) {
    const uint64_t t = blockIdx.x + h * G;
    if (t < ntiles) {
      const uint64_t base = tile_base(t, ps.tilemask, a.n);
#pragma unroll
      for (int i = 0; i < PER; ++i) (h ? preB : preA)[i] = a.psi[base | off[i]];
    }
  }
  for (uint64_t t = blockIdx.x; t < ntiles; t += 2 * G) {
    body(t, t + 2 * G, preA);
    if (t + G < ntiles) body(t + G, t + 3 * G, preB);
  }
}

__global__ __launch_bounds__(kThreads) void expect_kernel(const float2* __restrict__ psi, int n,
                                                          const DevTermWide* __restrict__ terms,
                                                          int tb, int te, double* partials) {
  __shared__ double red[kWaves];
  const uint64_t N = uint64_t(1) << n;
  double acc = 0.0;
  for (uint64_t x = uint64_t(blockIdx.x) * kThreads + threadIdx.x; x < N;
       x += uint64_t(gridDim.x) * kThreads) {
    const float2 va = psi[x];
    float part = 0.f;
    for (int t = tb; t < te; ++t) {
      const DevTermWide T = terms[t];
      const float2 pr = cmul_conj_a(psi[x ^ T.fmask], va);
      const float v = T.re * pr.x - T.im * pr.y;
      part += (__popcll(x & T.zmask) & 1) ? -v : v;
    }
    acc += double(part);
  }
  const double tot = block_sum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kThreads) void lambda_kernel(const float2* __restrict__ psi,
                                                          float2* __restrict__ lam, int n,
                                                          const DevTermWide* __restrict__ terms,
                                                          int tb, int te) {
  const uint64_t N = uint64_t(1) << n;
  for (uint64_t y = uint64_t(blockIdx.x) * kThreads + threadIdx.x; y < N;
       y += uint64_t(gridDim.x) * kThreads) {
    float2 l = make_float2(0.f, 0.f);
    for (int t = tb; t < te; ++t) {
      const DevTermWide T = terms[t];
      const uint64_t x = y ^ T.fmask;
      float2 w = cmul(make_float2(T.re, T.im), psi[x]);
      if (__popcll(x & T.zmask) & 1) w = make_float2(-w.x, -w.y);
      l = cadd(l, w);
    }
    lam[y] = l;
  }
}

__global__ __launch_bounds__(kThreads) void ip_kernel(const float2* __restrict__ psi,
                                                      const float2* __restrict__ phi, int n,
                                                      double* partials) {
  __shared__ double red[kWaves];
  const uint64_t N = uint64_t(1) << n;
  double re = 0.0, im = 0.0;
  for (uint64_t x = uint64_t(blockIdx.x) * kThreads + threadIdx.x; x < N;
       x += uint64_t(gridDim.x) * kThreads) {
    const float2 p = cmul_conj_a(psi[x], phi[x]);
    re += double(p.x);
    im += double(p.y);
  }
  re = block_sum(re, red);
  im = block_sum(im, red);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = re;
    partials[2 * blockIdx.x + 1] = im;
  }
}

// out[r] = sum_{i<count} partials[r * row_stride + i * stride], r = blockIdx.x.
__global__ __launch_bounds__(kThreads) void reduce_kernel(const double* __restrict__ partials,
                                                          int count, int stride, int row_stride,
                                                          double* __restrict__ out) {
  __shared__ double red[kWaves];
  double s = 0.0;
  for (int i = threadIdx.x; i < count; i += kThreads)
    s += partials[size_t(blockIdx.x) * row_stride + size_t(i) * stride];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// Serial finishing steps (one thread): avoids host round trips between kernels.
__global__ void finish_expect(const double* sums, float ident, float* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = float(sums[0] + double(ident));
}

__global__ void finish_sampled(const double* sums, int count, const float* coefs,
                               const int* shots, float ident, float* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double v = double(ident);
  for (int i = 0; i < count; ++i) v += double(coefs[i]) * sums[i] / double(shots[i]);
  *out = float(v);
}

__global__ void finish_ip(const double* sums, float2* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = make_float2(float(sums[0]), float(sums[1]));
}

__global__ void accumulate_grads(const double* sums, const int* cols, int count, double scale,
                                 double* gacc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = 0; i < count; ++i) gacc[cols[i]] += scale * sums[i];
}

// dst = w src (first) or dst += w src: one term of lambda = sum_j g_j |phi_j>.
__global__ void axpy_c64(const float2* __restrict__ src, float2* __restrict__ dst, int64_t count,
                         float w, int first) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float2 v = src[i];
    const float2 d = first ? make_float2(0.f, 0.f) : dst[i];
    dst[i] = make_float2(d.x + w * v.x, d.y + w * v.y);
  }
}

__global__ void fill_f64(double* p, int64_t n, double v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void grads_to_float(const double* gacc, int S, float* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < S; i += gridDim.x * blockDim.x)
    out[i] = float(gacc[i]);
}

__global__ void fill_row_f32(float* out, int count, float v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
    out[i] = v;
}

__global__ void fill_row_c64(float2* out, int64_t count, float2 first, float2 rest) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += int64_t(gridDim.x) * blockDim.x)
    out[i] = i == 0 ? first : rest;
}

__global__ void fill_i8(signed char* out, int64_t count, signed char v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += int64_t(gridDim.x) * blockDim.x)
    out[i] = v;
}

__global__ void copy_state(const float2* __restrict__ psi, int n, float2* __restrict__ out,
                           int64_t stride) {
  const int64_t N = int64_t(1) << n;
  for (int64_t x = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; x < stride;
       x += int64_t(gridDim.x) * blockDim.x)
    out[x] = x < N ? psi[x] : make_float2(-2.f, 0.f);
}

__global__ void rot1q_kernel(const float2* __restrict__ src, float2* __restrict__ dst, int n, int q,
                             float2 m00, float2 m01, float2 m10, float2 m11) {
  const uint64_t half = uint64_t(1) << (n - 1);
  for (uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < half;
       g += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t b = ((g >> q) << (q + 1)) | (g & ((uint64_t(1) << q) - 1));
    const uint64_t b1 = b | (uint64_t(1) << q);
    const float2 a0 = src[b], a1 = src[b1];
    dst[b] = cadd(cmul(m00, a0), cmul(m01, a1));
    dst[b1] = cadd(cmul(m10, a0), cmul(m11, a1));
  }
}

__device__ __forceinline__ double prob(const float2 v) {
  return double(v.x) * double(v.x) + double(v.y) * double(v.y);
}

__device__ __forceinline__ int sampler_block_log(int n) { return min(n, kCdfBlockLog); }
__device__ __forceinline__ int sampler_sub_log(int n) { return min(sampler_block_log(n), kSubLog); }

// Per block of 2^blog amplitudes (blog = min(n, 12)): the inclusive FP64
// prefix of its sub-block masses (sub-blocks of 2^slog amplitudes, slog =
// min(blog, 5)) in sincl[block * nsub + j], and the block total in bsum.
// Threads read consecutive amplitudes (coalesced); a sub-block's lanes are
// adjacent in one wavefront and sum by shuffles.
__global__ __launch_bounds__(kThreads) void prob_blocks_kernel(const float2* __restrict__ s, int n,
                                                               double* __restrict__ sincl,
                                                               double* __restrict__ bsum) {
  __shared__ double sub[kMaxSub];
  __shared__ double wtot;
  const int blog = sampler_block_log(n), slog = sampler_sub_log(n);
  const uint32_t BLK = 1u << blog;
  const int nsub = 1 << (blog - slog);
  const uint32_t width = 1u << slog;
  const uint64_t nb = uint64_t(1) << (n - blog);
  const int tid = threadIdx.x, lane = tid & 63;
  for (uint64_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const float2* blk = s + b * BLK;
    for (uint32_t x0 = 0; x0 < BLK; x0 += kThreads) {
      const uint32_t x = x0 + uint32_t(tid);
      double p = x < BLK ? prob(blk[x]) : 0.0;
      for (uint32_t off = 1; off < width; off <<= 1) p += __shfl_xor(p, int(off), 64);
      if (x < BLK && (uint32_t(lane) & (width - 1)) == 0) sub[x >> slog] = p;
    }
    __syncthreads();
    // Inclusive scan of the nsub <= 128 sub-block masses by the first two waves.
    double v = tid < nsub ? sub[tid] : 0.0;
    if (tid < 128) {
      for (int off = 1; off < 64; off <<= 1) {
        const double o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
      }
      if (tid == 63) wtot = v;
    }
    __syncthreads();
    if (tid >= 64 && tid < 128) v += wtot;
    if (tid < nsub) sincl[b * uint64_t(nsub) + tid] = v;
    if (tid == nsub - 1) bsum[b] = v;
    __syncthreads();
  }
}

// Inclusive scan of block totals (single workgroup).
__global__ __launch_bounds__(kThreads) void cdf_scan_kernel(const double* __restrict__ bsum,
                                                            int64_t nb, double* __restrict__ bincl) {
  __shared__ double red[kWaves];
  const int64_t chunk = (nb + kThreads - 1) / kThreads;
  const int64_t lo = threadIdx.x * chunk, hi = min(nb, lo + chunk);
  double run = 0.0;
  for (int64_t i = lo; i < hi; ++i) run += bsum[i];
  double total;
  double pre = block_exclusive_scan(run, red, &total);
  for (int64_t i = lo; i < hi; ++i) {
    pre += bsum[i];
    bincl[i] = pre;
  }
}

// Smallest x whose inclusive prefix mass exceeds u * total: block by binary
// search of the block prefix, sub-block by binary search of the block's
// sub-block prefix, amplitude by a scan of the sub-block.
__device__ __forceinline__ uint64_t sample_index(const float2* __restrict__ s,
                                                 const double* __restrict__ sincl,
                                                 const double* __restrict__ bincl, int n,
                                                 double u) {
  const int blog = sampler_block_log(n), slog = sampler_sub_log(n);
  const uint64_t nb = uint64_t(1) << (n - blog);
  const int nsub = 1 << (blog - slog);
  const double t = u * bincl[nb - 1];
  uint64_t lo = 0, hi = nb - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (bincl[mid] > t) hi = mid;
    else lo = mid + 1;
  }
  const uint64_t b = lo;
  const double t2 = t - (b > 0 ? bincl[b - 1] : 0.0);
  const double* c = sincl + b * uint64_t(nsub);
  int l2 = 0, h2 = nsub - 1;
  while (l2 < h2) {
    const int mid = (l2 + h2) >> 1;
    if (c[mid] > t2) h2 = mid;
    else l2 = mid + 1;
  }
  double acc = l2 > 0 ? c[l2 - 1] : 0.0;
  const uint64_t x0 = (b << blog) + (uint64_t(l2) << slog);
  const uint32_t width = 1u << slog;
  for (uint32_t i = 0; i < width; ++i) {
    acc += prob(s[x0 + i]);
    if (acc > t2) return x0 + i;
  }
  return x0 + width - 1;
}

__global__ __launch_bounds__(kThreads) void shots_samples_kernel(
    const float2* __restrict__ psi, const double* __restrict__ sincl,
    const double* __restrict__ bincl, int n, unsigned long long seed, uint32_t row, int nshots,
    int max_nq, signed char* __restrict__ out) {
  for (int k = blockIdx.x * kThreads + threadIdx.x; k < nshots; k += gridDim.x * kThreads) {
    const double u = philox_uniform(seed, (unsigned long long)k, row, 0xFFFFFFFFu);
    const uint64_t x = sample_index(psi, sincl, bincl, n, u);
    signed char* o = out + size_t(k) * max_nq;
    for (int q = 0; q < max_nq; ++q)
      o[max_nq - 1 - q] = q < n ? (signed char)((x >> q) & 1u) : (signed char)-2;
  }
}

// Terms jobs[blockIdx.y] of one measurement basis: partials[y * gridDim.x + x]
// is workgroup x's sum of the +-1 parities of its share of the term's shots.
__global__ __launch_bounds__(kThreads) void shots_parity_kernel(
    const float2* __restrict__ psi, const double* __restrict__ sincl,
    const double* __restrict__ bincl, int n, unsigned long long seed,
    const ParityJob* __restrict__ jobs, double* partials) {
  __shared__ double red[kWaves];
  const ParityJob J = jobs[blockIdx.y];
  int par = 0;
  for (int k = blockIdx.x * kThreads + threadIdx.x; k < J.nshots; k += gridDim.x * kThreads) {
    const double u = philox_uniform(seed, (unsigned long long)k, J.s_lo, J.s_hi);
    const uint64_t x = sample_index(psi, sincl, bincl, n, u);
    par += (__popcll(x & J.pmask) & 1) ? -1 : 1;
  }
  const double tot = block_sum(double(par), red);
  if (threadIdx.x == 0) partials[size_t(blockIdx.y) * gridDim.x + blockIdx.x] = tot;
}

// The sampler's uniforms for shots [start, start + count) of one stream, as
// the sampling kernels draw them (exposed for a bit-exact check of the RNG).
__global__ __launch_bounds__(kThreads) void philox_uniform_kernel(
    unsigned long long seed, unsigned long long start, int64_t count, uint32_t s_lo,
    uint32_t s_hi, double* __restrict__ out) {
  for (int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x; k < count;
       k += int64_t(gridDim.x) * kThreads)
    out[k] = philox_uniform(seed, start + (unsigned long long)k, s_lo, s_hi);
}

size_t AlignUp(size_t x) { return (x + 255) & ~size_t(255); }

int Err(char* err, int64_t len, const std::string& msg) {
  if (err && len > 0) {
    const size_t n = std::min<size_t>(size_t(len - 1), msg.size());
    memcpy(err, msg.data(), n);
    err[n] = '\0';
  }
  return 14;
}

// Workgroups of the streaming kernels: 8 per CU of an MI355X (256 CUs).  A
// constant (not a device query) so plans can be built without a GPU.
int StreamGrid() { return 2048; }


}  // namespace

// Workspace layout: the image (ops, mats, terms, param columns, identity
// constants, per-term coefficients/shots) followed by scratch and state buffers.
static void LargeLayout(LargePlan* p, size_t* off_pcols, size_t* off_ident, size_t* off_coefs,
                        size_t* off_shots, size_t* off_jobs, size_t* off_red) {
  const int grid = StreamGrid();
  p->grid = grid;
  size_t off = 0;
  p->off_ops = off; off = AlignUp(off + p->ops.size() * sizeof(DevOp));
  p->off_mats = off; off = AlignUp(off + (p->mats.size() + kMatCache) * sizeof(cfloat));
  p->off_terms = off; off = AlignUp(off + p->terms.size() * sizeof(DevTermWide));
  *off_pcols = off; off = AlignUp(off + p->param_cols.size() * sizeof(int32_t));
  const size_t nslots = size_t(p->mode == kModeInnerProduct ? p->K : p->O);
  *off_ident = off; off = AlignUp(off + size_t(p->B) * nslots * sizeof(float));
  *off_coefs = off; off = AlignUp(off + p->terms.size() * sizeof(float));
  *off_shots = off; off = AlignUp(off + p->terms.size() * sizeof(int32_t));
  *off_jobs = off; off = AlignUp(off + p->terms.size() * sizeof(ParityJob));
  p->image_bytes = off;
  const size_t slots = std::max<size_t>({p->param_cols.size(), size_t(2), size_t(64)});
  p->partial_slots = slots;
  p->off_partials = off; off = AlignUp(off + slots * size_t(grid) * sizeof(double));
  *off_red = off; off = AlignUp(off + std::max<size_t>(slots, p->terms.size() + 8) * sizeof(double));
  const size_t gacc_n = size_t(std::max<int64_t>(p->S, 1)) * (p->mode == kModeInnerProductGrad ? 2 : 1);
  p->off_gacc = off; off = AlignUp(off + gacc_n * sizeof(double));
  const size_t amps = size_t(1) << p->max_nq;
  p->off_psi = off; off = AlignUp(off + amps * sizeof(float2));
  const bool two = p->mode == kModeAdjoint || p->mode == kModeInnerProduct ||
                   p->mode == kModeSampledExpectation || p->mode == kModeInnerProductGrad;
  p->off_buf2 = off;
  if (two) off = AlignUp(off + amps * sizeof(float2));
  // Sampling: the within-block prefix of sub-block masses (one FP64 per 32
  // amplitudes), then block totals and their prefix.
  p->off_cdf = off;
  const bool sampling = p->mode == kModeSamples || p->mode == kModeSampledExpectation;
  if (sampling) off = AlignUp(off + std::max<size_t>(1, amps >> kSubLog) * sizeof(double));
  p->off_blocks = off;
  if (sampling) {
    const size_t nb = std::max<size_t>(1, amps >> kCdfBlockLog);
    off = AlignUp(off + 2 * nb * sizeof(double));
  }
  p->workspace_bytes = off;
}

struct LargeOffsets {
  size_t pcols, ident, coefs, shots, jobs, red;
};

static LargeOffsets Offsets(LargePlan* p) {
  LargeOffsets o{};
  LargeLayout(p, &o.pcols, &o.ident, &o.coefs, &o.shots, &o.jobs, &o.red);
  return o;
}

size_t LargeWorkspaceBytes(LargePlan* p) {
  Offsets(p);
  return p->workspace_bytes;
}

int LargeUpload(LargePlan* p, void* ws, hipStream_t stream, char* err, int64_t len) {
  LargeOffsets o = Offsets(p);
  std::vector<uint8_t> img(p->image_bytes, 0);
  auto put = [&](size_t off, const void* src, size_t n) {
    if (n) memcpy(img.data() + off, src, n);
  };
  put(p->off_ops, p->ops.data(), p->ops.size() * sizeof(DevOp));
  put(p->off_mats, p->mats.data(), p->mats.size() * sizeof(cfloat));
  put(p->off_terms, p->terms.data(), p->terms.size() * sizeof(DevTermWide));
  put(o.pcols, p->param_cols.data(), p->param_cols.size() * sizeof(int32_t));
  const size_t nslots = size_t(p->mode == kModeInnerProduct ? p->K : p->O);
  std::vector<float> ident(size_t(p->B) * nslots, 0.0f), coefs(p->terms.size(), 0.0f);
  std::vector<int32_t> shots(p->terms.size(), 1);
  std::vector<ParityJob> jobs(p->terms.size());
  for (const auto& lc : p->circuits) {
    for (size_t j = 0; j < lc.ident.size(); ++j) ident[size_t(lc.row) * nslots + j] = lc.ident[j];
    for (int t = lc.term_begin; t < lc.term_end; ++t) {
      const DevTermWide& T = p->terms[t];
      coefs[t] = T.re;
      if (!p->num_samples.empty())
        shots[t] = p->num_samples[size_t(lc.row) * p->O + T.slot];
      // Stream (row, slot << 16 | term index), as the LDS-resident kernel's.
      jobs[t] = ParityJob{uint32_t(lc.row), (uint32_t(T.slot) << 16) | uint32_t(T.pad), shots[t], 0,
                          T.fmask | T.zmask};
    }
  }
  put(o.ident, ident.data(), ident.size() * sizeof(float));
  put(o.coefs, coefs.data(), coefs.size() * sizeof(float));
  put(o.shots, shots.data(), shots.size() * sizeof(int32_t));
  put(o.jobs, jobs.data(), jobs.size() * sizeof(ParityJob));
  p->staging.swap(img);
  hipError_t e = hipMemcpyAsync(ws, p->staging.data(), p->staging.size(), hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return Err(err, len, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
  return 0;
}

int LargeLaunch(LargePlan* p, void* ws, void* output, hipStream_t stream, char* err, int64_t len) {
  LargeOffsets o = Offsets(p);
  uint8_t* base = static_cast<uint8_t*>(ws);
  const DevOp* ops = reinterpret_cast<const DevOp*>(base + p->off_ops);
  const float2* mats = reinterpret_cast<const float2*>(base + p->off_mats);
  const DevTermWide* terms = reinterpret_cast<const DevTermWide*>(base + p->off_terms);
  const int* pcols = reinterpret_cast<const int*>(base + o.pcols);
  const float* coefs = reinterpret_cast<const float*>(base + o.coefs);
  const int* shots = reinterpret_cast<const int*>(base + o.shots);
  double* partials = reinterpret_cast<double*>(base + p->off_partials);
  double* redbuf = reinterpret_cast<double*>(base + o.red);
  double* gacc = reinterpret_cast<double*>(base + p->off_gacc);
  float2* psi = reinterpret_cast<float2*>(base + p->off_psi);
  float2* buf2 = reinterpret_cast<float2*>(base + p->off_buf2);
  double* sincl = reinterpret_cast<double*>(base + p->off_cdf);
  const ParityJob* jobs = reinterpret_cast<const ParityJob*>(base + o.jobs);
  const size_t amps_max = size_t(1) << p->max_nq;
  double* bsum = reinterpret_cast<double*>(base + p->off_blocks);
  double* bincl = bsum + std::max<size_t>(1, amps_max >> kCdfBlockLog);
  const int grid = p->grid;
  const int64_t nslots = p->mode == kModeInnerProduct ? p->K : p->O;

  auto run_passes = [&](const PassPlan& pp, int n, float2* dst, bool init_first) {
    bool first = init_first;
    for (const DevPass& ps : pp.passes) {
      if (ps.dense) {
        const DenseOp& d = p->dense[size_t(ps.dense - 1)];
        if (first)
          hipLaunchKernelGGL(fill_row_c64,
                             dim3(std::min<int64_t>(grid, ((int64_t(1) << n) + 255) / 256)),
                             dim3(256), 0, stream, dst, int64_t(1) << n, make_float2(1.f, 0.f),
                             make_float2(0.f, 0.f));
        DenseArgs da{};
        da.psi = dst;
        da.G = mats + d.mat;
        da.pass = ps;
        da.n = n;
        for (int i = 0; i < d.k; ++i) da.pos[i] = d.pos[i];
        const uint64_t ntiles = uint64_t(1) << (n - ps.L);
        const int g = int(std::min<uint64_t>(ntiles, uint64_t(grid)));
        const size_t lds = (size_t(1) << kDenseTile) * sizeof(float2);
        if (d.k == 5) hipLaunchKernelGGL(dense_kernel<5>, dim3(g), dim3(kThreads), lds, stream, da);
        else hipLaunchKernelGGL(dense_kernel<6>, dim3(g), dim3(kThreads), lds, stream, da);
        first = false;
        continue;
      }
      PassArgs a{};
      a.src = dst;
      a.dst = dst;
      a.ops = ops;
      a.mats = mats;
      a.pass = ps;
      a.n = n;
      a.init = first ? 1 : 0;
      const uint64_t ntiles = uint64_t(1) << (n - ps.L);
      const int g = int(std::min<uint64_t>(ntiles, uint64_t(grid)));
      const size_t lds = (size_t(1) << ps.L) * sizeof(float2);
      if (ps.L > 13) hipLaunchKernelGGL((pass_kernel<false, false, true>), dim3(g), dim3(kThreads), lds, stream, a);
      else hipLaunchKernelGGL(pass_kernel<false>, dim3(g), dim3(kThreads), lds, stream, a);
      first = false;
    }
    if (first) {  // no ops at all: still need |0>
      hipLaunchKernelGGL(fill_row_c64, dim3(std::min<int64_t>(grid, ((int64_t(1) << n) + 255) / 256)),
                         dim3(256), 0, stream, dst, int64_t(1) << n, make_float2(1.f, 0.f),
                         make_float2(0.f, 0.f));
    }
  };

  // The sampler's block sums of |s|^2 for an n-qubit state and their scan.
  auto prob_blocks = [&](const float2* s, int n) {
    const int blog = std::min(n, kCdfBlockLog);
    const uint64_t nb = uint64_t(1) << (n - blog);
    hipLaunchKernelGGL(prob_blocks_kernel, dim3(int(std::min<uint64_t>(nb, uint64_t(grid)))),
                       dim3(kThreads), 0, stream, s, n, sincl, bsum);
    hipLaunchKernelGGL(cdf_scan_kernel, dim3(1), dim3(kThreads), 0, stream, bsum, int64_t(nb), bincl);
  };

  for (const LargeCircuit& lc : p->circuits) {
    const int row = lc.row;
    const int n = lc.n;
    if (lc.empty) {
      switch (p->mode) {
        case kModeExpectation:
        case kModeSampledExpectation:
          hipLaunchKernelGGL(fill_row_f32, dim3(1), dim3(256), 0, stream,
                             static_cast<float*>(output) + size_t(row) * nslots, int(nslots), -2.0f);
          break;
        case kModeAdjoint: {
          float* gb = static_cast<float*>(output);
          if (p->with_expectation) {
            hipLaunchKernelGGL(fill_row_f32, dim3(1), dim3(256), 0, stream,
                               gb + size_t(row) * nslots, int(nslots), -2.0f);
            gb += size_t(p->B) * size_t(p->O);
          }
          hipLaunchKernelGGL(fill_row_f32, dim3(1), dim3(256), 0, stream, gb + size_t(row) * p->S,
                             int(p->S), 0.0f);
          break;
        }
        case kModeInnerProduct:
          hipLaunchKernelGGL(fill_row_c64, dim3(1), dim3(256), 0, stream,
                             static_cast<float2*>(output) + size_t(row) * p->K, int64_t(p->K),
                             make_float2(1.f, 0.f), make_float2(1.f, 0.f));
          break;
        case kModeInnerProductGrad:
          hipLaunchKernelGGL(fill_row_c64, dim3(1), dim3(256), 0, stream,
                             static_cast<float2*>(output) + size_t(row) * p->S, int64_t(p->S),
                             make_float2(0.f, 0.f), make_float2(0.f, 0.f));
          break;
        case kModeState:
          hipLaunchKernelGGL(fill_row_c64, dim3(grid), dim3(256), 0, stream,
                             static_cast<float2*>(output) + size_t(row) * amps_max, int64_t(amps_max),
                             make_float2(1.f, 0.f), make_float2(-2.f, 0.f));
          break;
        case kModeSamples:
          hipLaunchKernelGGL(fill_i8, dim3(grid), dim3(256), 0, stream,
                             static_cast<signed char*>(output) + size_t(row) * p->nshots * p->max_nq,
                             int64_t(p->nshots) * p->max_nq, (signed char)-2);
          break;
      }
      continue;
    }
    const int sgrid = int(std::min<uint64_t>(uint64_t(grid), ((uint64_t(1) << n) + 255) / 256));
    if (p->mode == kModeInnerProductGrad) {
      // lambda = sum_j g_j |phi_j> in buf2 (AccumulateFusedCircuits,
      // util_qsim.h:421-442), each phi_j simulated in psi's buffer first.
      const int64_t amps = int64_t(1) << n;
      const int agrid = int(std::min<int64_t>(grid, (amps + 255) / 256));
      for (size_t k = 0; k < lc.others.size(); ++k) {
        run_passes(lc.others[k], n, psi, true);
        hipLaunchKernelGGL(axpy_c64, dim3(agrid), dim3(256), 0, stream, psi, buf2, amps,
                           lc.other_weights[k], k == 0 ? 1 : 0);
      }
      if (lc.others.empty())
        hipLaunchKernelGGL(fill_row_c64, dim3(agrid), dim3(256), 0, stream, buf2, amps,
                           make_float2(0.f, 0.f), make_float2(0.f, 0.f));
    }
    run_passes(lc.fwd, n, psi, true);
    const float* ident = reinterpret_cast<const float*>(base + o.ident) + size_t(row) * nslots;
    if (p->mode == kModeExpectation || (p->mode == kModeAdjoint && p->with_expectation)) {
      float* exp_base = static_cast<float*>(output);
      int t = lc.term_begin;
      for (int64_t j = 0; j < nslots; ++j) {
        const int tb = t;
        while (t < lc.term_end && p->terms[t].slot == j) ++t;
        float* out = exp_base + size_t(row) * nslots + j;
        if (t == tb) {
          hipLaunchKernelGGL(fill_row_f32, dim3(1), dim3(64), 0, stream, out, 1, lc.ident[j]);
          continue;
        }
        hipLaunchKernelGGL(expect_kernel, dim3(sgrid), dim3(kThreads), 0, stream, psi, n, terms, tb,
                           t, partials);
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kThreads), 0, stream, partials, sgrid, 1, 0, redbuf);
        hipLaunchKernelGGL(finish_expect, dim3(1), dim3(64), 0, stream, redbuf, lc.ident[j], out);
      }
    }
    if (p->mode == kModeAdjoint || p->mode == kModeInnerProductGrad) {
      // Adjoint sweep.  Inner-product gradients accumulate the complex
      // <dG psi|lambda> (tfq_inner_product_grad.cc:259-314) into gacc[2 S];
      // expectation gradients 2 Re<dG psi|lambda> into gacc[S].
      const bool cplx = p->mode == kModeInnerProductGrad;
      const int width = cplx ? 2 : 1;
      float* grad_base = static_cast<float*>(output) +
                         (p->with_expectation ? size_t(p->B) * size_t(p->O) : 0);
      if (!cplx)
        hipLaunchKernelGGL(lambda_kernel, dim3(sgrid), dim3(kThreads), 0, stream, psi, buf2, n, terms,
                           lc.lam_begin, lc.lam_end);
      hipLaunchKernelGGL(fill_f64, dim3(1), dim3(256), 0, stream, gacc, int64_t(p->S) * width, 0.0);
      for (const DevPass& ps : lc.rev.passes) {
        PassArgs a{};
        a.psi = psi;
        a.lam = buf2;
        a.ops = ops;
        a.mats = mats;
        a.pass = ps;
        a.n = n;
        a.partials = partials;
        const uint64_t ntiles = uint64_t(1) << (n - ps.L);
        const int g = int(std::min<uint64_t>(ntiles, uint64_t(grid)));
        const size_t lds = 2 * (size_t(1) << ps.L) * sizeof(float2);
        if (cplx) hipLaunchKernelGGL((pass_kernel<true, true>), dim3(g), dim3(kThreads), lds, stream, a);
        else hipLaunchKernelGGL((pass_kernel<true, false>), dim3(g), dim3(kThreads), lds, stream, a);
        // Partial slots of a pass are contiguous [s0, s0 + nd) (numbered at planning).
        int s0 = -1, nd = 0;
        for (int i = ps.op_begin; i < ps.op_end; ++i) {
          const DevOp& op = p->ops[i];
          if (op.type != kOpParam) continue;
          if (s0 < 0) s0 = op.col[0];
          nd += op.nderiv;
        }
        if (nd == 0) continue;
        // partials[slot * g + wg]: row r = slot - s0 -> base + r * g.
        const int rows = nd * width;
        hipLaunchKernelGGL(reduce_kernel, dim3(rows), dim3(kThreads), 0, stream,
                           partials + size_t(s0) * g, g, 1, g, redbuf);
        hipLaunchKernelGGL(accumulate_grads, dim3(1), dim3(64), 0, stream, redbuf, pcols + s0, rows,
                           cplx ? 1.0 : 2.0, gacc);
      }
      const int64_t cnt = p->S * width;
      hipLaunchKernelGGL(grads_to_float, dim3(std::max<int>(1, int((cnt + 255) / 256))), dim3(256), 0,
                         stream, gacc, int(cnt), grad_base + size_t(row) * cnt);
    } else if (p->mode == kModeState) {
      hipLaunchKernelGGL(copy_state, dim3(grid), dim3(256), 0, stream, psi, n,
                         static_cast<float2*>(output) + size_t(row) * amps_max, int64_t(amps_max));
    } else if (p->mode == kModeSamples) {
      prob_blocks(psi, n);
      if (p->nshots > 0)
        hipLaunchKernelGGL(shots_samples_kernel,
                           dim3(std::max(1, std::min(grid, (p->nshots + kThreads - 1) / kThreads))),
                           dim3(kThreads), 0, stream, psi, sincl, bincl, n,
                           (unsigned long long)p->seed, uint32_t(row), p->nshots, p->max_nq,
                           static_cast<signed char*>(output) + size_t(row) * p->nshots * p->max_nq);
    } else if (p->mode == kModeSampledExpectation) {
      // Terms are ordered by slot, then by measurement basis: the sampler's
      // block sums are rebuilt only when the basis changes, and the terms of
      // one basis (up to kParityJobs) draw their shots in one launch, each
      // from its own stream.
      uint64_t cur_f = ~uint64_t(0), cur_y = ~uint64_t(0);
      const float2* src = psi;
      int t0 = lc.term_begin;
      while (t0 < lc.term_end) {
        const DevTermWide& T = p->terms[t0];
        const uint64_t yb = T.fmask & T.zmask;
        if (T.fmask != cur_f || yb != cur_y) {
          cur_f = T.fmask;
          cur_y = yb;
          src = psi;
          if (T.fmask) {
            bool first = true;
            for (int q = 0; q < n; ++q) {
              if (!((T.fmask >> q) & 1)) continue;
              float yp[5] = {-0.5f, 1.0f, 0.0f, 0, 0}, xp[5] = {0.5f, 1.0f, 0.0f, 0, 0};
              cdouble U[16];
              if ((yb >> q) & 1) GateUnitary(kGateX, xp, U);
              else GateUnitary(kGateY, yp, U);
              auto f2 = [&](int i) { return make_float2(float(U[i].real()), float(U[i].imag())); };
              hipLaunchKernelGGL(rot1q_kernel, dim3(grid), dim3(256), 0, stream, first ? psi : buf2,
                                 buf2, n, q, f2(0), f2(1), f2(2), f2(3));
              first = false;
            }
            src = buf2;
          }
          prob_blocks(src, n);
        }
        // Terms [t0, te): the same basis, at most kParityJobs of them.
        int te = t0 + 1;
        int max_ns = p->num_samples[size_t(row) * p->O + T.slot];
        while (te < lc.term_end && te - t0 < kParityJobs && p->terms[te].fmask == cur_f &&
               (p->terms[te].fmask & p->terms[te].zmask) == cur_y) {
          max_ns = std::max(max_ns, p->num_samples[size_t(row) * p->O + p->terms[te].slot]);
          ++te;
        }
        const int g = std::max(1, std::min(grid, (max_ns + kThreads - 1) / kThreads));
        hipLaunchKernelGGL(shots_parity_kernel, dim3(g, te - t0), dim3(kThreads), 0, stream, src,
                           sincl, bincl, n, (unsigned long long)p->seed, jobs + t0, partials);
        hipLaunchKernelGGL(reduce_kernel, dim3(te - t0), dim3(kThreads), 0, stream, partials, g, 1, g,
                           redbuf + (t0 - lc.term_begin));
        t0 = te;
      }
      int t = lc.term_begin;
      for (int64_t j = 0; j < nslots; ++j) {
        const int tb = t;
        while (t < lc.term_end && p->terms[t].slot == j) ++t;
        float* out = static_cast<float*>(output) + size_t(row) * nslots + j;
        hipLaunchKernelGGL(finish_sampled, dim3(1), dim3(64), 0, stream,
                           redbuf + (tb - lc.term_begin), t - tb, coefs + tb, shots + tb,
                           lc.ident[j], out);
      }
    } else if (p->mode == kModeInnerProduct) {
      for (size_t k = 0; k < lc.others.size(); ++k) {
        run_passes(lc.others[k], n, buf2, true);
        hipLaunchKernelGGL(ip_kernel, dim3(sgrid), dim3(kThreads), 0, stream, psi, buf2, n, partials);
        hipLaunchKernelGGL(reduce_kernel, dim3(2), dim3(kThreads), 0, stream, partials, sgrid, 2, 1, redbuf);
        hipLaunchKernelGGL(finish_ip, dim3(1), dim3(64), 0, stream, redbuf,
                           static_cast<float2*>(output) + size_t(row) * p->K + k);
      }
    }
    (void)ident;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return Err(err, len, std::string("large-state kernels: ") + hipGetErrorString(e));
  return 0;
}

int LaunchPhiloxUniforms(uint64_t seed, uint64_t start, int64_t count, uint32_t s_lo,
                         uint32_t s_hi, double* out, hipStream_t stream, char* err, int64_t len) {
  if (count <= 0) return 0;
  const int64_t blocks = std::min<int64_t>((count + kThreads - 1) / kThreads, 4096);
  hipLaunchKernelGGL(philox_uniform_kernel, dim3(unsigned(blocks)), dim3(kThreads), 0, stream,
                     (unsigned long long)seed, (unsigned long long)start, count, s_lo, s_hi, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return Err(err, len, std::string("philox kernel: ") + hipGetErrorString(e));
  return 0;
}

}  // namespace tfqx
