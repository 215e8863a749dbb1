// Program -> gate list: qubit resolution, argument/symbol resolution and
// the gate matrices of TFQ's 16 gate ids.
//
// Restates (from the algorithm, not the code):
//   ResolveQubitIds            core/src/program_resolution.cc:49-186 (+ paired 188-312)
//   QsimCircuitFromProgram     core/src/circuit_parser_qsim.cc:828-861
//   gate builders              core/src/circuit_parser_qsim.cc:153-562
//   GateMetaData               core/src/circuit_parser_qsim.h:35-65
//   gradient gates (eps 5e-3)  core/src/adj_util.cc:32,175-302
// Gate matrices are Cirq's EigenGate definitions, U = sum_k e^{i pi t (l_k+s)} P_k,
// evaluated in double from float32 arguments (t = exp*exp_s computed in float32,
// circuit_parser_qsim.cc:229) and stored as complex64.
#pragma once

#include <stdint.h>

#include <complex>
#include <string>
#include <unordered_map>
#include <vector>

#include "wire.h"

namespace tfqx {

using cdouble = std::complex<double>;

enum GateKind : int {
  kGateI1 = 0, kGateI2, kGateH, kGateX, kGateY, kGateZ, kGateXX, kGateYY, kGateZZ,
  kGateCZ, kGateCX, kGateSwap, kGateISwap, kGatePhasedX, kGateFSim, kGatePhasedISwap
};

struct Status {
  int code = 0;  // 0 ok, otherwise a tfqx status code
  std::string msg;
  bool ok() const { return code == 0; }
  static Status Invalid(const std::string& m) { return Status{3, m}; }
};

// Symbol name -> (column, value), as GetSymbolMaps (parse_context.cc:291-347).
struct SymbolMap {
  const std::unordered_map<std::string, int>* columns = nullptr;
  const float* values = nullptr;  // row of symbol_values
  bool find(const std::string& name, int* col, float* val) const;
};

struct Gate {
  GateKind kind;
  int nt;            // target count (1 or 2)
  int bits[2];       // target bit positions; bits[0] is the matrix MSB
  uint64_t cmask = 0, cval = 0;
  int ncontrols = 0;
  float params[5];   // raw float32 arguments in GateMetaData::gate_params order
  int nsym = 0;
  int sym_pidx[2];   // which params[] entry is symbolic
  int sym_col[2];    // symbol column
  cdouble U[16];     // unitary, row-major (2^nt x 2^nt), bits[0] = MSB of the index
};

// A single-qubit noise channel of a noisy program (ParseAppendChannel,
// core/src/circuit_parser_qsim.cc:598-826; ids DP ADP GAD AD RST PD PF BF).
// Kraus operators follow Cirq's channel definitions (as qsim's
// lib/channels_cirq.h implements them), evaluated in double from the float32
// arguments.  Unitary mixtures (DP, ADP, PF, BF) keep each unitary with its
// probability; the other channels keep their Kraus matrices.
enum ChannelKind : int { kChanDP = 0, kChanADP, kChanGAD, kChanAD, kChanRST, kChanPD, kChanPF, kChanBF };
struct Channel {
  ChannelKind kind;
  int bit = 0;           // target bit position
  int gate_pos = 0;      // number of gates that precede it in program order
  bool mixture = false;
  int nk = 0;            // Kraus operators (2 or 4)
  double probs[4] = {0, 0, 0, 0};  // mixtures: probability of each unitary
  cdouble K[4][4];       // 2x2 row-major Kraus matrices (mixtures: the unitaries)
};

struct Circuit {
  int num_qubits = 0;
  bool no_moments = true;  // program had no moments: Pauli ids are not checked
  std::vector<Gate> gates;
  std::vector<Channel> channels;  // noisy programs only, in program order
  std::unordered_map<std::string, int> qubit_index;  // proto id -> sorted index
};

// Parse + resolve one program.  `sym` may have no columns (empty map).
Status BuildCircuit(const Program& prog, const SymbolMap& sym, Circuit* out);
// Noisy programs (NoisyQsimCircuitFromProgram, circuit_parser_qsim.cc:828-870):
// an operation whose id is not a gate id is parsed as a channel.
Status BuildNoisyCircuit(const Program& prog, const SymbolMap& sym, Circuit* out);
// `tmpl` with its symbolic arguments re-read from another row of values
// (same program bytes, so the same gates, qubits and symbol columns).
void RebindCircuit(const Circuit& tmpl, const float* values, Circuit* out);
// Paired resolution for inner products: other programs use the reference
// program's qubit map and must cover exactly its qubits.
Status BuildPairedCircuit(const Program& prog, const Circuit& reference, Circuit* out);

// Pauli term in XOR-mask form over n bits (bit = n-1-index).
struct Pauli {
  float coef;           // Re(coefficient); TFQ ignores the imaginary part
  uint64_t xmask = 0;   // X or Y
  uint64_t zmask = 0;   // Z or Y
  int ny = 0;           // number of Y
  std::vector<std::pair<int, char>> ops;  // (bit, 'X'|'Y'|'Z')
};
Status ResolvePauliSum(const PauliSum& ps, const Circuit& circ, std::vector<Pauli>* out);

// Unitary of a gate from its raw arguments (double precision).
void GateUnitary(GateKind kind, const float* params, cdouble* U);
// d U / d params[pidx].  tfq_fd=true: TFQ's float32 central difference with
// eps=5e-3 (adj_util.cc:175-302); false: the exact derivative.
void GateDerivative(GateKind kind, const float* params, int pidx, bool tfq_fd, cdouble* dU);
int GateTargets(GateKind kind);

// The eigen components of gate kinds 0..12 (I1 .. ISWAP) for the device
// evaluator: out[kind] = {count, lam[3], P[3] as (re, im) row-major}.
void EigenTableForDevice(double (*out)[100]);

}  // namespace tfqx
