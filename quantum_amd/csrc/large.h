// HBM-streaming path for states wider than one LDS tile.
//
// Used when a batch holds a circuit wider than the LDS-resident kernel takes
// (e.g. 26- and 30-qubit states).  Circuits run one after another, each with
// the whole GPU: the fused ops are grouped into passes; a pass streams every
// 2^L-amplitude tile of the state through LDS once (coalesced runs of at
// least 2^6 contiguous amplitudes) and applies all of its ops there.
// Reductions (Pauli expectations, lambda, inner products, FP64 CDF for
// sampling) are separate streaming kernels with FP64 partials.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "circuit.h"
#include "plan.h"
#include "tfqx_internal.h"

namespace tfqx {

struct PassPlan {
  std::vector<DevPass> passes;
  int first_param = 0;   // index of this program's first kOpParam in the param table
};

struct LargeCircuit {
  int n = 0;
  int row = 0;
  bool empty = false;
  int L = 0;                          // tile qubits
  PassPlan fwd;                       // forward passes
  PassPlan rev;                       // adjoint reverse passes (two buffers)
  std::vector<PassPlan> others;       // inner product paired programs
  std::vector<float> other_weights;   // inner-product gradient: g[row, k] of each paired program
  int term_begin = 0, term_end = 0;   // DevTermWide ranges in the plan's term table
  int lam_begin = 0, lam_end = 0;
  std::vector<int> rot_groups;        // sampled expectation: term ranges per rotation pattern
  std::vector<float> ident;           // O identity constants
};

// A fused block of 5-6 qubits applied by one dense MFMA pass: matrix (2^k x
// 2^k complex64, row-major, index bit i <-> global bit bits[i]) at `mat` in the
// plan's matrix array; pos[i] is bits[i]'s position in the pass's tile.
struct DenseOp {
  int k = 0;
  int bits[kMaxFuseWide];
  int pos[kMaxFuseWide];
  uint32_t mat = 0;
};

struct LargePlan {
  int mode = 0;
  int64_t B = 0, S = 0, O = 0, K = 0;
  int max_nq = 0;
  uint64_t seed = 0;
  int nshots = 0;
  std::vector<int32_t> num_samples;
  std::vector<LargeCircuit> circuits;
  std::vector<DevOp> ops;
  std::vector<cfloat> mats;
  std::vector<DevPass> passes;
  std::vector<DevTermWide> terms;
  std::vector<int32_t> param_cols;    // symbol column of every (param op, derivative)
  std::vector<DenseOp> dense;         // dense passes (DevPass::dense - 1 indexes this)
  // Workspace layout.
  size_t off_ops = 0, off_mats = 0, off_passes = 0, off_terms = 0, off_partials = 0,
         off_gacc = 0, off_psi = 0, off_buf2 = 0, off_cdf = 0, off_blocks = 0,
         workspace_bytes = 0;
  size_t image_bytes = 0;
  std::vector<uint8_t> staging;
  int grid = 0;
  size_t partial_slots = 0;
  bool too_wide = false;
  bool with_expectation = false;  // adjoint mode also writes expectations
};

Status BuildLargePlan(int mode, const std::vector<Circuit>& circuits,
                      const std::vector<std::vector<std::vector<Pauli>>>& sums,
                      const float* downstream, const std::vector<std::vector<Circuit>>& others,
                      int64_t B, int64_t S, int64_t O, int64_t K, bool tfq_fd,
                      bool with_expectation, std::unique_ptr<LargePlan>* out,
                      int wide_fusion = 0);

// Dense passes use tiles of kDenseTile qubits: 2^(kDenseTile - k) columns of
// the fused block's 2^k target amplitudes per tile.
constexpr int kDenseTile = 12;

// Group fused ops into tile passes (exposed for tests of the scheduler).
// max_r: largest register-engine register count for the passes (0: LDS path
// only; 4: tiles of 12-13 qubits; 5: also 14-qubit one-state tiles).
// With `rec` (and the circuit `c` the ops come from), each op's matrix recipe
// is recorded for rebinding.
void SchedulePasses(const std::vector<FusedOp>& ops, int n, int L, std::vector<DevPass>* passes,
                    std::vector<DevOp>* dev_ops, std::vector<cfloat>* mats,
                    int max_r, const Circuit* c = nullptr,
                    SharedRecipe* rec = nullptr);

size_t LargeWorkspaceBytes(LargePlan* p);
int LargeUpload(LargePlan* p, void* ws, hipStream_t stream, char* err, int64_t len);
int LargeLaunch(LargePlan* p, void* ws, void* output, hipStream_t stream, char* err, int64_t len);
int LaunchPhiloxUniforms(uint64_t seed, uint64_t start, int64_t count, uint32_t s_lo,
                         uint32_t s_hi, double* out, hipStream_t stream, char* err, int64_t len);

}  // namespace tfqx
