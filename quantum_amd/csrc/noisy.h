// Noisy circuits: Monte-Carlo trajectories on the GPU.
//
// Replaces the reference's trajectory ops (core/ops/noise/tfq_noisy_expectation.cc,
// tfq_noisy_sampled_expectation.cc, tfq_noisy_samples.cc), which run qsim's
// QuantumTrajectorySimulator one trajectory per CPU thread.  Here every
// trajectory of every circuit is an independent work item of one persistent
// kernel:
//
//   * circuits of up to kWaveMaxQubits qubits run in wave mode: the register
//     engine holds 2^(kWaveMaxQubits - n) trajectories per wavefront, one per
//     group of 2^(n - kWaveRegBits) lanes (circuits of fewer than
//     kWaveRegBits qubits get idle qubits added above their own);
//   * circuits of kWaveMaxQubits+1 .. kNoisyLdsMaxQubits qubits run one
//     trajectory per workgroup with the state in LDS;
//   * circuits of kNoisyLdsMaxQubits+1 .. kNoisyMaxQubits qubits run one
//     trajectory per workgroup with the state in a per-workgroup buffer of
//     the plan's workspace (one buffer of 2^max_nq amplitudes per launched
//     workgroup: min(kNoisyGrid, work items)).
//
// Channel c of trajectory t of output row i draws the uniform variate
// philox(seed, shot = t, stream = (i, c)); results depend on the seed only,
// not on the launch shape.  Expectations accumulate per (row, slot) in FP64
// and a finishing kernel divides by the trajectory counts.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "circuit.h"
#include "plan.h"
#include "tfqx_internal.h"

namespace tfqx {

// Widest noisy program whose workgroup trajectories keep the state in LDS
// (2^14 complex64 amplitudes = 128 KiB of gfx950's 160 KiB).
constexpr int kNoisyLdsMaxQubits = 14;
// Widest noisy program: 20 qubits, 8 MiB per workgroup state, 4 GiB of
// workspace for the kNoisyGrid workgroups.
constexpr int kNoisyMaxQubits = 20;
// Workgroups of the noisy kernel's launch (at most).
constexpr int kNoisyGrid = 512;
// Random stream of the terminal measurement (samples) and of the one-shot
// term measurements (sampled expectation, ORed with the term's index).
constexpr uint32_t kStreamSample = 0xFFFFFFFFu;
constexpr uint32_t kStreamShot = 0x80000000u;

struct NoisyPlan {
  int mode = kModeNoisyExpectation;
  int64_t B = 0, O = 0;
  int max_nq = 0;
  int nshots = 0;                    // samples mode: trajectories per circuit
  uint64_t seed = 0;
  std::vector<DevCircuit> circuits;  // non-empty rows
  std::vector<DevOp> ops;
  std::vector<cfloat> mats;
  std::vector<DevTerm> terms;
  std::vector<DevChunk> chunks;      // workgroup chunks first, then wave chunks
  int nblock = 0;                    // chunks [0, nblock) run by whole workgroups
  std::vector<float> ident;          // [B, O] identity-term constants
  std::vector<int32_t> num_samples;  // [B, O] trajectories per (row, slot)
  std::vector<int32_t> row_empty;    // [B]
  int64_t trajectories = 0;
  // Device image layout (byte offsets).
  size_t off_counters = 0, off_circuits = 0, off_ops = 0, off_mats = 0, off_terms = 0,
         off_chunks = 0, off_ident = 0, off_nsamp = 0, off_empty = 0, off_nacc = 0, total = 0;
  // Workgroups of the launch: min(kNoisyGrid, chunks), at least 1.
  int grid = 1;
  // Dynamic LDS of the launch (log2 amplitudes): the widest workgroup
  // trajectory that keeps its state in LDS, at least the wave-mode buffers.
  int lds_log2 = 0;
  // Workgroup state buffers (programs wider than kNoisyLdsMaxQubits only):
  // `grid` buffers of 2^state_log2 amplitudes at off_states, else 0.
  size_t off_states = 0;
  int state_log2 = 0;
  size_t image = 0;
  std::vector<uint8_t> staging;
};

// num_samples: [B, O] (expectation modes) or null; nshots: samples mode.
Status BuildNoisyPlan(int mode, const std::vector<Circuit>& circuits,
                      const std::vector<std::vector<std::vector<Pauli>>>& sums,
                      const int32_t* num_samples, int nshots, int64_t B, int64_t O,
                      std::unique_ptr<NoisyPlan>* out);
void NoisyLayout(NoisyPlan* p);
int NoisyUpload(NoisyPlan* p, void* ws, hipStream_t stream, char* err, int64_t len);
int NoisyLaunch(NoisyPlan* p, void* ws, void* output, hipStream_t stream, char* err, int64_t len);

}  // namespace tfqx
