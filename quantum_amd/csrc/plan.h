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
#include "rebind_device.h"
#include "tfqx_internal.h"

namespace tfqx {

using cfloat = std::complex<float>;

// A fused block on global bit positions (ascending); matrix index bit j <-> bits[j].
struct FusedOp {
  int k = 0;
  int bits[kMaxFuseWide];
  uint64_t cmask = 0, cval = 0;   // controls kept outside the matrix
  std::vector<cdouble> M;
  bool diag = false;
  // kOpParam data (adjoint only)
  bool param = false;
  int nderiv = 0;
  std::vector<cdouble> dM[kMaxDeriv];
  int col[kMaxDeriv];
  // A parameterised block fused with its neighbours (controls inside the
  // matrix): dM[d] is the block's product with gate gates[dgate[d]] replaced
  // by its derivative for that gate's symbol dsym[d].
  bool fused_param = false;
  int dgate[kMaxDeriv] = {0, 0};
  int dsym[kMaxDeriv] = {0, 1};
  // Gates (circuit indices, application order) the matrix is the product of.
  std::vector<int> gates;
};

struct FuseOptions {
  int max_fuse = 2;   // qubits per fused block
  // Decide diagonal storage from the gate kinds only (Z, ZZ, CZ, I) instead of
  // the numeric matrix, so the op list does not depend on symbol values.
  bool structural_diag = false;
};

// How one device op's matrices are rebuilt from its gates for another row
// that shares the program (only symbol values differ).
struct OpRecipe {
  // 0 fused block, 1 single gate (controls outside), 2 parameterised gate
  // (controls outside), 3 fused block containing parameterised gates
  uint8_t kind = 0;
  uint8_t k = 0;
  bool diag = false;
  bool symbolic = false; // contains a symbolic gate (else the matrix is value-independent)
  int bits[kMaxFuse];
  int gate_begin = 0, gate_end = 0;  // range in SharedRecipe::gates
  uint32_t mat = 0;                  // offsets into the row's matrices
  uint32_t dmat[kMaxDeriv];
  int nderiv = 0;
  int8_t dgate[kMaxDeriv] = {0, 0};  // kind 3: block position of each derivative's gate
  int8_t dsym[kMaxDeriv] = {0, 1};   // ... and which of its symbols
};

struct SharedRecipe {
  std::vector<OpRecipe> ops;
  std::vector<int> gates;
  size_t mats_per_row = 0;
};

// Record how to rebuild the matrices of device op `d` (made from fused block
// `f` of circuit `c`, matrices at d.mat / d.dmat in the row's array).
void AppendRecipe(const FusedOp& f, const Circuit& c, const DevOp& d, SharedRecipe* rec);

// Fuse gates [begin, end) of a circuit into blocks (deterministic greedy
// frontier fuser: gates join the open block holding their qubits while the
// union stays within max_fuse qubits; disjoint open blocks merge when a gate
// spans them).  Controlled gates whose targets+controls exceed max_fuse stay
// unfused with a control mask.
void FuseGates(const Circuit& c, int begin, int end, const FuseOptions& opt,
               std::vector<FusedOp>* out);

// The adjoint sweep's op list for circuit `c`, in reverse application order.
// Reverse fusion on (the default): the whole circuit is fused as by
// FuseGates, except that a block holding parameterised gates spans at most
// two qubits and kMaxDeriv derivatives, and becomes a kOpParam op whose
// derivative matrices are the block's product with one gate replaced by its
// derivative (gradient = 2 Re <lambda_after| U_a dG U_b |psi_before>).
// Parameterised gates with controls stay alone with the controls outside.
// Reverse fusion off: fused segments between parameterised gates, each
// parameterised gate an op of its own (the reference's gradient circuit,
// core/src/adj_util.cc:155-172).
void BuildReverseOps(const Circuit& c, const FuseOptions& opt, bool tfq_fd,
                     std::vector<FusedOp>* rev);
// Process-wide switch for later plans (1 on, 0 off).
void SetReverseFusion(int on);
int GetReverseFusion();

// Register-engine segments: rewrite ops[begin, end) of an op list that runs on
// a tile of L qubits so that it starts with a kOpRemap and every non-diagonal
// op's targets lie in the register qubits of the latest remap (segments of at
// most RegBitsForTile(L, max_r) target qubits).  A segment may take an op ahead of
// earlier ops that commute with it (disjoint supports, or both diagonal), so
// the op order changes only between commuting ops.  No-op when the tile size
// has no register engine or an op is wider than 2 qubits.
void InsertRemaps(std::vector<DevOp>* ops, size_t begin, int L, int max_r = 4);
// The same with an explicit register-qubit count R (wave mode).
void InsertRemapsR(std::vector<DevOp>* ops, size_t begin, int L, int R);

// Planner tuning (process-wide; 0 = default): register engine on/off and the
// streaming path's tile qubits for one- and two-state passes.
struct Tuning {
  int reg_engine = 2;
  int tile_single = 0;
  int tile_two = 0;
};
Tuning GetTuning();
void SetTuning(const Tuning& t);

// Expand one gate into kernel convention (sorted bits, little-endian index).
FusedOp GateToOp(const Gate& g, bool keep_controls_outside, bool structural_diag = false);
// Gate kinds whose matrix is diagonal for every argument value.
bool KindDiagonal(GateKind kind);

// A batch program for the LDS-resident kernel.
struct BatchProgram {
  std::vector<DevCircuit> circuits;   // in launch order (longest first)
  std::vector<DevOther> others;       // inner product paired programs
  std::vector<DevPass> passes;        // tiled circuits
  std::vector<DevOp> ops;
  std::vector<cfloat> mats;
  std::vector<DevTerm> terms;
  std::vector<float> ident;           // [B * O] identity-term constants
  std::vector<uint32_t> row_mat_base; // shared programs: row -> first matrix of the row
  int max_n = 0;
  int num_block = 0;                  // circuits run by whole workgroups (the rest: wave mode)
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
  bool wave = false;      // small circuits (kWaveRegBits..kWaveMaxQubits qubits) in wave mode
  // Keep every adjoint lambda term, with weight 0 where TFQ drops it (|g c| <
  // 1e-5), so new downstream weights can be written into the same term list.
  bool keep_all_lambda = false;
};

// Build the per-circuit program pieces and concatenate them.  With `recipes`,
// row r's matrix recipe goes to (*recipes)[r] (relative to the row's first
// matrix, BatchProgram::row_mat_base), so RebindBatchProgram can re-evaluate
// the matrices for new symbol values; the caller then sets
// opt.fuse.structural_diag so the op list does not depend on the values.
void BuildBatchProgram(const std::vector<CircuitJob>& jobs, int num_slots, int num_symbols,
                       const PlanOptions& opt, int threads, BatchProgram* out,
                       std::vector<SharedRecipe>* recipes = nullptr);
// What re-evaluating one row's matrices needs: the recipes of its ops that
// contain symbolic gates, and those ops' gates in compact form (no matrix;
// recipe gate indices point into `gates`).
struct RebindGate {
  GateKind kind;
  int8_t nt, nsym;
  int8_t sym_pidx[2];
  int bits[2];
  uint64_t cmask, cval;
  float params[5];
  int sym_col[2];
};
struct RowRebind {
  std::vector<RebindGate> gates;
  SharedRecipe recipe;
};
// Keep the symbolic ops of `rec` (built from circuit `c`) with their gates.
RowRebind CompactRecipe(const Circuit& c, const SharedRecipe& rec);
// Re-evaluate every row's symbolic matrices for new values [rows, num_symbols],
// writing them into `mats` (an array laid out as prog.mats, e.g. the upload
// image's copy of it).
void RebindBatchProgram(const std::vector<RowRebind>& rows, const float* values, int num_symbols,
                        bool tfq_fd, int threads, const BatchProgram& prog, cfloat* mats);

// Flat tables for re-evaluating the symbolic matrices on the device
// (rebind_device.h).  Per-row plans: each row's compacted gates and symbolic
// recipes; shared plans: the template's gates and recipe, used by every row.
struct RebindTables {
  std::vector<RbGate> gates;
  std::vector<int32_t> gidx;
  std::vector<RbOp> ops;
  std::vector<RbRow> rows;
  int kmax = 1;
};
struct RebindTableSizes {
  size_t gates = 0, gidx = 0, ops = 0;
};
RebindTableSizes SizeRebindTables(const std::vector<RowRebind>& rows);
RebindTableSizes SizeSharedRebindTables(const Circuit& tmpl, const SharedRecipe& rec);
// false if a recipe falls outside what the device evaluates (a parameterised
// block wider than two qubits).
bool BuildRebindTables(const std::vector<RowRebind>& rows, const BatchProgram& prog,
                       RebindTables* t);
bool BuildSharedRebindTables(const Circuit& tmpl, const SharedRecipe& rec,
                             const BatchProgram& prog, RebindTables* t);

// Rows whose programs are byte-identical: `tmpl` is the program resolved with
// row 0's values; the op list is built once and each row re-evaluates only the
// matrices of ops that contain symbolic gates (from values[row * S ...]).
// Requires tmpl.num_qubits <= opt.tile_log2 (LDS-resident).
void BuildSharedBatchProgram(const Circuit& tmpl, const std::vector<CircuitJob>& jobs,
                             const float* values, int num_symbols, int num_slots,
                             const PlanOptions& opt, int threads, BatchProgram* out,
                             SharedRecipe* recipe);
// Re-evaluate every row's matrices of a shared program for new values, into
// `mats` (laid out as prog.mats).
void RebindSharedBatchProgram(const Circuit& tmpl, const SharedRecipe& recipe, const float* values,
                              int num_symbols, bool tfq_fd, int threads, const BatchProgram& prog,
                              cfloat* mats);

// A fused block as a device op (matrices appended to `mats`), and a Pauli term
// in XOR-mask form with weight w * i^{#Y}.
DevOp FusedToDevOp(const FusedOp& f, std::vector<cfloat>* mats);
DevTerm PauliToTerm(const Pauli& p, float w, int slot);

// Cost estimate used for longest-first ordering and the autotuner.
double CircuitCost(const DevCircuit& c, const BatchProgram& p);

}  // namespace tfqx
