// Host planner: parsed circuits -> device programs.
//
// The fuser and the op lists play the role of qsim's BasicGateFuser and of
// CreateGradientCircuit's partial fuses (core/src/adj_util.cc:155-172): the
// forward program is the whole circuit fused into k-qubit blocks; the adjoint
// program is the reversed list of fused segments between parameterised gates,
// with each parameterised gate kept as a kOpParam op carrying its derivative
// matrices and symbol columns.
#pragma once

#include <stdint.h>

#include <complex>
#include <vector>

#include "circuit.h"
#include "tfqx_internal.h"

namespace tfqx {

using cfloat = std::complex<float>;

// A fused block on global bit positions (ascending); matrix index bit j <-> bits[j].
struct FusedOp {
  int k = 0;
  int bits[kMaxFuse];
  uint64_t cmask = 0, cval = 0;   // controls kept outside the matrix
  std::vector<cdouble> M;
  bool diag = false;
  // kOpParam data (adjoint only)
  bool param = false;
  int nderiv = 0;
  std::vector<cdouble> dM[kMaxDeriv];
  int col[kMaxDeriv];
};

struct FuseOptions {
  int max_fuse = 2;   // qubits per fused block
};

// Fuse gates [begin, end) of a circuit into blocks (deterministic greedy
// frontier fuser: gates join the open block holding their qubits while the
// union stays within max_fuse qubits; disjoint open blocks merge when a gate
// spans them).  Controlled gates whose targets+controls exceed max_fuse stay
// unfused with a control mask.
void FuseGates(const Circuit& c, int begin, int end, const FuseOptions& opt,
               std::vector<FusedOp>* out);

// Expand one gate into kernel convention (sorted bits, little-endian index).
FusedOp GateToOp(const Gate& g, bool keep_controls_outside);

// A batch program for the LDS-resident kernel.
struct BatchProgram {
  std::vector<DevCircuit> circuits;   // in launch order (longest first)
  std::vector<DevOther> others;       // inner product paired programs
  std::vector<DevPass> passes;        // tiled circuits
  std::vector<DevOp> ops;
  std::vector<cfloat> mats;
  std::vector<DevTerm> terms;
  std::vector<float> ident;           // [B * O] identity-term constants
  int max_n = 0;
  int num_rows = 0, num_slots = 0, num_symbols = 0;
  bool adjoint = false;
  bool any_tiled = false;
};

struct CircuitJob {
  const Circuit* circuit;
  const std::vector<std::vector<Pauli>>* sums;  // O Pauli sums (or null)
  const float* downstream;                       // O weights (adjoint) or null
  const std::vector<Circuit>* others;            // K paired programs (inner product)
  int row;
};

struct PlanOptions {
  FuseOptions fuse;
  int mode = kModeExpectation;  // BatchMode
  bool adjoint = false;
  bool tfq_fd = true;     // adjoint derivative: TFQ's float32 FD (true) or exact
  int tile_log2 = 30;     // circuits wider than this run tiled (state in global scratch)
};

// Build the per-circuit program pieces and concatenate them.
void BuildBatchProgram(const std::vector<CircuitJob>& jobs, int num_slots, int num_symbols,
                       const PlanOptions& opt, int threads, BatchProgram* out);

// Cost estimate used for longest-first ordering and the autotuner.
double CircuitCost(const DevCircuit& c, const BatchProgram& p);

}  // namespace tfqx
