// Noisy trajectory kernel for gfx950 (see noisy.h).
//
// One persistent launch pulls work items (DevChunk) from two queues:
//   1. workgroup chunks: one trajectory of an 11..20-qubit circuit, channels
//      through workgroup reductions; the state is in LDS up to 14 qubits and
//      in the workgroup's own workspace buffer for 15..20 qubits;
//   2. wave chunks: up to 2^(10 - n) trajectories of an n <= 10 qubit circuit
//      side by side in one wavefront's registers (register engine with
//      L = 10: the bits above n index the trajectory), channels through
//      reductions over each trajectory's lanes.
// Expectation modes add each trajectory's values into FP64 accumulators per
// (row, slot); noisy_finish divides by the trajectory counts, adds identity
// terms and clears the accumulators, so a launch can be replayed.
#include <hip/hip_runtime.h>

#include <string.h>

#include <string>

#include "batch_kernels.h"
#include "noisy.h"
#include "reg_engine.h"

namespace tfqx {
namespace {

struct NoisyArgs {
  const DevCircuit* circuits;
  const DevOp* ops;
  const float2* mats;
  const DevTerm* terms;
  const DevChunk* chunks;
  const int32_t* num_samples;
  const float* ident;
  const int32_t* row_empty;
  double* nacc;
  float2* states;  // per-workgroup states of programs wider than LDS (or null)
  int state_log2;
  signed char* samples_out;
  float* exp_out;
  unsigned long long seed;
  int* queue;
  int* wqueue;
  int* done;
  int nchunks;
  int nblock;
  int nslots;
  int nshots;
  int max_nq;
  int B;
};

__device__ __forceinline__ uint32_t cdf_first_above(const double* cdf, uint32_t N, double target) {
  uint32_t lo = 0, hi = N - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cdf[mid] > target) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// The first x whose inclusive FP64 prefix of |psi|^2 exceeds u * total
// (the last index if none does), for one draw from a workgroup's state (LDS
// or global; N >= kThreads).  The state is split into kThreads contiguous
// chunks; wavefronts read each chunk with consecutive lanes on consecutive
// amplitudes (coalesced) and reduce its mass, a block scan of the chunk masses
// finds the chunk, and one wavefront scans that chunk 64 amplitudes at a time.
// Nothing is written to the state.
__device__ __forceinline__ double wave_inclusive_scan(double v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

__device__ __forceinline__ uint32_t sample_one(const float2* __restrict__ s, uint32_t N, double u,
                                               double* red) {
  __shared__ double mass[kThreads];
  __shared__ int s_chunk;
  __shared__ uint32_t s_x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t C = N / kThreads;
  for (int k = wave; k < kThreads; k += kWaves) {
    double m = 0.0;
    for (uint32_t i = uint32_t(lane); i < C; i += 64) {
      const float2 v = s[size_t(k) * C + i];
      m += double(v.x) * double(v.x) + double(v.y) * double(v.y);
    }
    m = wave_sum(m);
    if (lane == 0) mass[k] = m;
  }
  if (tid == 0) s_chunk = kThreads - 1;
  __syncthreads();
  const double mk = mass[tid];
  double total;
  const double off = block_exclusive_scan(mk, red, &total);
  const double target = u * total;
  if (off + mk > target) atomicMin(&s_chunk, tid);
  __syncthreads();
  const int k = s_chunk;
  if (tid == k) red[0] = off;
  __syncthreads();
  if (wave == 0) {
    const double goal = target - red[0];
    double run = 0.0;
    uint32_t x = uint32_t(k + 1) * C - 1u;
    for (uint32_t i0 = 0; i0 < C; i0 += 64) {
      const uint32_t i = i0 + uint32_t(lane);
      double p = 0.0;
      if (i < C) {
        const float2 v = s[size_t(k) * C + i];
        p = double(v.x) * double(v.x) + double(v.y) * double(v.y);
      }
      const double pre = run + wave_inclusive_scan(p);
      const unsigned long long hit = __ballot(i < C && pre > goal);
      if (hit) {
        x = uint32_t(k) * C + i0 + uint32_t(__ffsll((long long)hit) - 1);
        break;
      }
      run = __shfl(pre, 63, 64);
    }
    if (lane == 0) s_x = x;
  }
  __syncthreads();
  return s_x;
}

// ------------------------------------------------------------------ workgroup trajectories
// One channel on an LDS state (all threads; selection rule as reg_channel_j).
__device__ __forceinline__ void lds_channel(float2* s, int n, const DevOp& op, const float2* mats,
                                            double u, double* red) {
  const int nk = ChanKraus(op.cval);
  const uint32_t q = op.q[0];
  const float2* X = mats + op.dmat[0];
  int pick = nk - 1;
  double pp = 1.0;
  if (op.cval & kChanMixture) {
    for (int k = nk - 2; k >= 0; --k)
      if (u < double(X[k].x)) pick = k;
    if ((op.cval & kChanFirstIdentity) && pick == 0) return;
  } else {
    float s00 = 0.f, s11 = 0.f, sr = 0.f, si = 0.f;
    for (uint32_t g = threadIdx.x; g < (1u << (n - 1)); g += kThreads) {
      const uint32_t b = insert_zero(g, q);
      const float2 x0 = s[b], x1 = s[b | (1u << q)];
      s00 += x0.x * x0.x + x0.y * x0.y;
      s11 += x1.x * x1.x + x1.y * x1.y;
      sr += x0.x * x1.x + x0.y * x1.y;
      si += x0.x * x1.y - x0.y * x1.x;
    }
    const double d00 = block_sum(double(s00), red);
    const double d11 = block_sum(double(s11), red);
    const double dr = block_sum(double(sr), red);
    const double di = block_sum(double(si), red);
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
  const float2* K = mats + op.mat + 4 * pick;
  const float sc = (op.cval & kChanMixture) ? 1.0f : float(1.0 / sqrt(pp));
  const float2 M[4] = {make_float2(K[0].x * sc, K[0].y * sc), make_float2(K[1].x * sc, K[1].y * sc),
                       make_float2(K[2].x * sc, K[2].y * sc), make_float2(K[3].x * sc, K[3].y * sc)};
  apply_1q(s, n, int(q), M);
}

template <int MODE>
__device__ __forceinline__ void wg_trajectory(const NoisyArgs& a, const DevChunk& ch, float2* psi,
                                              float2* mbuf, double* red) {
  const DevCircuit c = a.circuits[ch.circuit];
  const int n = uniform(c.n);
  const int row = uniform(c.row);
  const uint32_t N = 1u << n;
  const int tid = threadIdx.x;
  const unsigned long long traj = (unsigned long long)ch.t0;
  for (uint32_t x = tid; x < N; x += kThreads) psi[x] = make_float2(x == 0 ? 1.f : 0.f, 0.f);
  __syncthreads();
  for_ops(a.ops, uniform(c.fwd_begin), uniform(c.fwd_end), a.mats, mbuf,
          [&](const DevOp& op, const float2* M) {
            if (op.type == kOpChannel) {
              const double u = philox_uniform(a.seed, traj, uint32_t(row), uint32_t(op.col[0]));
              lds_channel(psi, n, op, a.mats, u, red);
            } else {
              apply_op<false, 2>(psi, n, op, M);
            }
          });
  if (MODE == kModeNoisySamples) {
    const double u = philox_uniform(a.seed, traj, uint32_t(row), kStreamSample);
    const uint32_t x = sample_one(psi, N, u, red);
    signed char* o = a.samples_out + (size_t(row) * a.nshots + size_t(traj)) * a.max_nq;
    for (int q = tid; q < a.max_nq; q += kThreads)
      o[a.max_nq - 1 - q] = q < ch.nq ? (signed char)((x >> q) & 1u) : (signed char)-2;
    __syncthreads();
    return;
  }
  double norm2 = 1.0;
  if (MODE == kModeNoisySampledExpectation) {
    float part = 0.f;
    for (uint32_t x = tid; x < N; x += kThreads) {
      const float2 v = psi[x];
      part += v.x * v.x + v.y * v.y;
    }
    norm2 = block_sum(double(part), red);
  }
  int t = uniform(c.term_begin);
  const int te = uniform(c.term_end);
  for (int j = 0; j < a.nslots; ++j) {
    double val = 0.0;
    while (t < te && uniform(a.terms[t].slot) == j) {
      const DevTerm T = a.terms[t];
      float part = 0.f;
      for (uint32_t x = tid; x < N; x += kThreads) {
        const float2 pr = cmul_conj_a(psi[x ^ T.fmask], psi[x]);
        const float v = T.re * pr.x - T.im * pr.y;
        part += (__popc(x & T.zmask) & 1) ? -v : v;
      }
      const double e = block_sum(double(part), red);
      if (MODE == kModeNoisyExpectation) {
        val += e;
      } else {
        const float coef = __int_as_float(T.pad);
        const double u = philox_uniform(a.seed, traj, uint32_t(row),
                                        kStreamShot | uint32_t(t - c.term_begin));
        val += double(coef) * (u < 0.5 * (1.0 + e / norm2) ? 1.0 : -1.0);
      }
      ++t;
    }
    if (tid == 0 && (long long)traj < (long long)a.num_samples[size_t(row) * a.nslots + j])
      atomicAdd(&a.nacc[size_t(row) * a.nslots + j], val);
  }
  __syncthreads();
}

// ------------------------------------------------------------------ wave trajectories
template <int MODE>
__device__ __forceinline__ void wave_chunk(const NoisyArgs& a, const DevChunk& ch, float2* P) {
  const DevCircuit c = a.circuits[ch.circuit];
  const int n = uniform(c.n);
  const int row = uniform(c.row);
  const uint32_t lane = threadIdx.x & 63u;
  const int gbits = n - kWaveRegBits;         // lanes per trajectory: 2^gbits
  const uint32_t grp = lane >> gbits;          // trajectory slot of this lane
  const uint32_t li = lane & ((1u << gbits) - 1u);
  const unsigned long long traj = (unsigned long long)ch.t0 + grp;
  const bool live = traj < (unsigned long long)ch.t1;
  const uint32_t N = 1u << n;
  float2* S = P + (size_t(grp) << n);          // this trajectory's state after the run
  NoiseDraw nd{a.seed, traj, uint32_t(row), gbits};
  const int fb = uniform(c.fwd_begin), fe = uniform(c.fwd_end);
  if (fb < fe) {
    auto fio = tile_io(
        [N](uint32_t x, cf& va, cf&) { va = cf{(x & (N - 1u)) == 0u ? 1.f : 0.f, 0.f}; },
        [P](uint32_t x, const cf& va, const cf&) { P[x] = make_float2(va.x, va.y); });
    GlobalMats gm{a.mats};
    reg_run<kWaveRegBits, false, true>(P, P, kWaveMaxQubits, a.ops, fb, fe, gm, nd, fio);
  } else {
    for (uint32_t x = lane; x < (1u << kWaveMaxQubits); x += 64u)
      P[x] = make_float2((x & (N - 1u)) == 0u ? 1.f : 0.f, 0.f);
  }
  esync<true>();
  if (MODE == kModeNoisySamples) {
    // Lane li owns amplitudes [16 li, 16 li + 16) of its trajectory's state.
    const uint32_t y0 = li * (1u << kWaveRegBits);
    double own = 0.0;
#pragma unroll
    for (int r = 0; r < (1 << kWaveRegBits); ++r) {
      const float2 v = S[y0 + r];
      own += double(v.x) * double(v.x) + double(v.y) * double(v.y);
    }
    // Inclusive scan over the trajectory's lanes.
    double inc = own;
    for (int off = 1; off < (1 << gbits); off <<= 1) {
      const double o = __shfl_up(inc, off, 64);
      if (li >= uint32_t(off)) inc += o;
    }
    const double total = __shfl(inc, int((grp << gbits) | ((1u << gbits) - 1u)), 64);
    const double u = philox_uniform(a.seed, traj, uint32_t(row), kStreamSample);
    const double target = u * total;
    // The first index whose inclusive prefix exceeds the target.
    uint32_t x = N - 1u;
    double run = inc - own;
    bool found = false;
#pragma unroll
    for (int r = 0; r < (1 << kWaveRegBits); ++r) {
      const float2 v = S[y0 + r];
      run += double(v.x) * double(v.x) + double(v.y) * double(v.y);
      if (!found && run > target) {
        found = true;
        x = y0 + r;
      }
    }
    // Smallest found index over the trajectory's lanes.
    for (int off = 1; off < (1 << gbits); off <<= 1) x = min(x, uint32_t(__shfl_xor(int(x), off, 64)));
    if (live) {
      signed char* o = a.samples_out + (size_t(row) * a.nshots + size_t(traj)) * a.max_nq;
      for (int q = int(li); q < a.max_nq; q += (1 << gbits))
        o[a.max_nq - 1 - q] = q < ch.nq ? (signed char)((x >> q) & 1u) : (signed char)-2;
    }
    esync<true>();
    return;
  }
  double norm2 = 1.0;
  if (MODE == kModeNoisySampledExpectation) {
    float part = 0.f;
    for (uint32_t y = li; y < N; y += (1u << gbits)) {
      const float2 v = S[y];
      part += v.x * v.x + v.y * v.y;
    }
    norm2 = group_sum(double(part), gbits);
  }
  int t = uniform(c.term_begin);
  const int te = uniform(c.term_end);
  for (int j = 0; j < a.nslots; ++j) {
    double val = 0.0;
    while (t < te && uniform(a.terms[t].slot) == j) {
      const DevTerm T = a.terms[t];
      float part = 0.f;
      for (uint32_t y = li; y < N; y += (1u << gbits)) {
        const float2 pr = cmul_conj_a(S[y ^ T.fmask], S[y]);
        const float v = T.re * pr.x - T.im * pr.y;
        part += (__popc(y & T.zmask) & 1) ? -v : v;
      }
      const double e = group_sum(double(part), gbits);
      if (MODE == kModeNoisyExpectation) {
        val += e;
      } else {
        const float coef = __int_as_float(T.pad);
        const double u = philox_uniform(a.seed, traj, uint32_t(row),
                                        kStreamShot | uint32_t(t - c.term_begin));
        val += double(coef) * (u < 0.5 * (1.0 + e / norm2) ? 1.0 : -1.0);
      }
      ++t;
    }
    // One lane per trajectory contributes; the wavefront adds them up.
    const bool use = live && li == 0u &&
                     (long long)traj < (long long)a.num_samples[size_t(row) * a.nslots + j];
    const double tot = wave_sum(use ? val : 0.0);
    if (lane == 0u) atomicAdd(&a.nacc[size_t(row) * a.nslots + j], tot);
  }
  esync<true>();
}

static_assert(kLdsMaxQubitsOneBuffer >= kNoisyLdsMaxQubits, "workgroup trajectories keep the state in LDS");
static_assert(kLdsMaxQubitsOneBuffer >= kWaveMaxQubits + 3,
              "wave trajectories need one 2^kWaveMaxQubits buffer per wavefront (8 wavefronts)");

template <int MODE>
__global__ __launch_bounds__(kThreads, 4) void noisy_kernel(NoisyArgs a) {
  extern __shared__ float2 lds[];
  __shared__ double red[kWaves];
  __shared__ int s_ci;
  __shared__ float2 mbuf[2 * kMatSlot];
  const int tid = threadIdx.x;
  // Workgroup trajectories.
  while (true) {
    if (tid == 0) s_ci = atomicAdd(a.queue, 1);
    __syncthreads();
    const int ci = uniform(s_ci);
    __syncthreads();
    if (ci >= a.nblock) break;
    const DevChunk ch = a.chunks[ci];
    // Programs wider than LDS use this workgroup's own buffer in the
    // workspace; barriers order its global accesses as they do LDS ones.
    float2* psi = uniform(ch.nq) > kNoisyLdsMaxQubits
                      ? a.states + (size_t(blockIdx.x) << a.state_log2)
                      : lds;
    wg_trajectory<MODE>(a, ch, psi, mbuf, red);
  }
  // Wave trajectories: each wavefront pulls chunks of its own.
  if (a.nblock < a.nchunks) {
    const int nwave = a.nchunks - a.nblock;
    float2* P = lds + size_t(tid >> 6) * (1u << kWaveMaxQubits);
    while (true) {
      int w = 0;
      if ((tid & 63) == 0) w = atomicAdd(a.wqueue, 1);
      w = uniform(__shfl(w, 0, 64));
      if (w >= nwave) break;
      wave_chunk<MODE>(a, a.chunks[a.nblock + w], P);
    }
  }
  __syncthreads();
  // The last workgroup out resets the queues (replayable launches).
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

// Expectation modes: out = acc / trajectories + identity terms (-2 for rows
// of empty programs); the accumulators are cleared for the next launch.
// Samples mode: rows of empty programs are filled with -2.
__global__ void noisy_finish(NoisyArgs a, int mode) {
  if (mode == kModeNoisySamples) {
    const size_t per_row = size_t(a.nshots) * size_t(a.max_nq);
    const size_t total = size_t(a.B) * per_row;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += size_t(gridDim.x) * blockDim.x)
      if (a.row_empty[i / per_row]) a.samples_out[i] = -2;
    return;
  }
  const size_t total = size_t(a.B) * size_t(a.nslots);
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += size_t(gridDim.x) * blockDim.x) {
    const size_t row = i / size_t(a.nslots);
    const double acc = a.nacc[i];
    a.nacc[i] = 0.0;
    a.exp_out[i] = a.row_empty[row] ? -2.0f
                                    : float(acc / double(a.num_samples[i]) + double(a.ident[i]));
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

}  // namespace

int NoisyUpload(NoisyPlan* p, void* ws, hipStream_t stream, char* err, int64_t len) {
  // The image holds zeroed queue counters and accumulators.
  hipError_t e = hipMemcpyAsync(ws, p->staging.data(), p->staging.size(), hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return Err(err, len, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
  return 0;
}

int NoisyLaunch(NoisyPlan* p, void* ws, void* output, hipStream_t stream, char* err, int64_t len) {
  uint8_t* base = static_cast<uint8_t*>(ws);
  NoisyArgs a{};
  a.circuits = reinterpret_cast<const DevCircuit*>(base + p->off_circuits);
  a.ops = reinterpret_cast<const DevOp*>(base + p->off_ops);
  a.mats = reinterpret_cast<const float2*>(base + p->off_mats);
  a.terms = reinterpret_cast<const DevTerm*>(base + p->off_terms);
  a.chunks = reinterpret_cast<const DevChunk*>(base + p->off_chunks);
  a.ident = reinterpret_cast<const float*>(base + p->off_ident);
  a.num_samples = reinterpret_cast<const int32_t*>(base + p->off_nsamp);
  a.row_empty = reinterpret_cast<const int32_t*>(base + p->off_empty);
  a.nacc = reinterpret_cast<double*>(base + p->off_nacc);
  a.states = p->state_log2 ? reinterpret_cast<float2*>(base + p->off_states) : nullptr;
  a.state_log2 = p->state_log2;
  // Three queue counters, 64 bytes apart, inside the 256-byte counter block.
  int* counters = reinterpret_cast<int*>(base + p->off_counters);
  a.queue = counters;
  a.wqueue = counters + 16;
  a.done = counters + 32;
  a.nchunks = int(p->chunks.size());
  a.nblock = p->nblock;
  a.nslots = int(p->O);
  a.nshots = p->nshots;
  a.max_nq = p->max_nq;
  a.B = int(p->B);
  a.seed = (unsigned long long)p->seed;
  if (p->mode == kModeNoisySamples) a.samples_out = static_cast<signed char*>(output);
  else a.exp_out = static_cast<float*>(output);
  const size_t out_count = p->mode == kModeNoisySamples
                               ? size_t(p->B) * size_t(p->nshots) * size_t(p->max_nq)
                               : size_t(p->B) * size_t(p->O);
  if (out_count == 0) return 0;
  if (a.nchunks > 0) {
    // Dynamic LDS: 8 wavefronts of 2^10-amplitude buffers (64 KiB, two
    // workgroups per CU), or the widest workgroup state kept in LDS when that
    // is larger (a 14-qubit state takes 128 KiB, one workgroup per CU;
    // workgroups that find no room start when others end and then find the
    // queues empty).  Workspace trajectories (over 14 qubits) need no LDS
    // state.  The workspace holds one state buffer per launched workgroup.
    const int grid = p->grid;
    const size_t lds = (size_t(1) << p->lds_log2) * sizeof(float2);
    switch (p->mode) {
      case kModeNoisyExpectation:
        hipLaunchKernelGGL(noisy_kernel<kModeNoisyExpectation>, dim3(grid), dim3(kThreads), lds, stream, a);
        break;
      case kModeNoisySampledExpectation:
        hipLaunchKernelGGL(noisy_kernel<kModeNoisySampledExpectation>, dim3(grid), dim3(kThreads), lds,
                           stream, a);
        break;
      default:
        hipLaunchKernelGGL(noisy_kernel<kModeNoisySamples>, dim3(grid), dim3(kThreads), lds, stream, a);
        break;
    }
  }
  const int fgrid = int(std::min<size_t>(1024, (out_count + 255) / 256));
  hipLaunchKernelGGL(noisy_finish, dim3(std::max(1, fgrid)), dim3(256), 0, stream, a, p->mode);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return Err(err, len, std::string("noisy kernels: ") + hipGetErrorString(e));
  return 0;
}

}  // namespace tfqx
