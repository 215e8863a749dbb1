// Internal data structures shared by the host planner and the gfx950 kernels.
//
// A circuit reaches the device as a small "program": a list of ops applied to
// a state tile in LDS.  A tile is 2^L amplitudes; for circuits of L or fewer
// qubits the whole state is one tile that stays in LDS for the full circuit.
// Larger states are processed in passes; each pass covers a tile-qubit set and
// streams every tile of the state through LDS once.
#pragma once

#include <stdint.h>

namespace tfqx {

// Maximum target qubits of one fused op applied in LDS.
constexpr int kMaxFuse = 4;
// Widest fused block: the streaming path applies blocks of 5-6 qubits as dense
// MFMA contractions (every other kernel takes at most kMaxFuse).
constexpr int kMaxFuseWide = 6;
// Maximum derivative matrices attached to one parameterised gate (FSIM theta+phi,
// PhasedX / PhasedISwap phase+exponent).
constexpr int kMaxDeriv = 2;

enum OpType : uint8_t {
  kOpUnitary = 0,   // apply M (forward) or M^dagger (adjoint sweep) on targets
  kOpParam = 1,     // adjoint: psi <- G^dag psi; grad += 2Re<lam|dG psi>; lam <- G^dag lam
  kOpDiag = 2,      // diagonal M: only the 2^k diagonal entries are stored
  kOpRemap = 3,     // register engine: reload the tile with cmask = register-qubit mask
  kOpChannel = 4,   // noise channel on q[0] (noisy trajectories; layout below)
};

// kOpChannel layout.  cval = nk | flags << 8 (cmask 0); mat = nk 2x2 Kraus
// matrices (unitary mixtures: the unitaries); dmat[0] = unitary mixtures: nk
// entries whose .x is the float32 cumulative probability p_0 + ... + p_k;
// other channels: nk 2x2 matrices K_k^dag K_k.  nderiv = 1 (dmat[0] is
// rebased with the matrices).  col[0] = the channel's ordinal in its program
// (its random stream).
constexpr uint32_t kChanMixture = 1u << 8;    // unitary mixture
constexpr uint32_t kChanFirstIdentity = 1u << 9;  // unitary 0 is the identity
constexpr int ChanKraus(uint32_t cval) { return int(cval & 15u); }

// Register engine (reg_engine.h): 2^9 threads per workgroup, each holding
// 2^R amplitudes of a 2^L tile in registers, R = L - 9 for tiles of 12..13
// qubits; other tile sizes use the LDS path.  The planner inserts kOpRemap
// ops exactly when this returns nonzero and every op of the list has k <= 2.
constexpr int kRegThreadBits = 9;

// The register engine reads op matrices from an LDS window of this many
// complex entries over the plan's matrix array; the device copy of that array
// is padded by the same count so a window never reads past it.
constexpr int kMatCache = 768;
// The adjoint batch kernel keeps two 13-qubit tiles (128 KiB) and one
// workgroup per CU, so its window can take the LDS the others leave free.
constexpr int kMatCacheWide = 3072;
// Register qubits of the register engine for an L-qubit tile (0: no engine):
// L - 9 for tiles of 12..(9 + max_r) qubits.  Two-state runs and the batched
// kernel use at most 4; one-state streaming passes may use 5 (14-qubit tiles).
constexpr int RegBitsForTile(int L, int max_r = 4) {
  return (L >= 12 && L - kRegThreadBits <= max_r) ? L - kRegThreadBits : 0;
}

// Wave mode of the batched kernel: after the workgroup-sized circuits, each
// wavefront runs one small circuit (kWaveRegBits..kWaveMaxQubits qubits) in
// registers, kWaveRegBits register qubits per lane and the rest on lane bits.
constexpr int kWaveRegBits = 4;
constexpr int kWaveMaxQubits = 10;
// Batches with fewer wave-eligible circuits than this run them as workgroups.
constexpr int kWaveMinCircuits = 1024;

// One op on a tile.  Qubit positions are tile-local bit indices (ascending).
// Matrix layout: row-major 2^k x 2^k complex64, index bit j <-> q[j].
struct alignas(8) DevOp {
  uint8_t type;
  uint8_t k;
  uint8_t nderiv;
  uint8_t pad0;
  uint8_t q[kMaxFuse];
  uint32_t cmask;    // tile-local control bits
  uint32_t cval;     // required values of the control bits
  uint32_t mat;      // float2 offset of the unitary (or diagonal)
  uint32_t dmat[kMaxDeriv];  // float2 offsets of derivative matrices
  int32_t col[kMaxDeriv];    // symbol column each derivative accumulates into
};
static_assert(sizeof(DevOp) == 40, "DevOp layout");

// A Pauli string folded to XOR-mask form: P|x> = ph * (-1)^popc(x & zmask) |x ^ fmask>
// with ph = coef * i^{#Y}.  slot = output column (expectation) or 0 (lambda).
struct alignas(8) DevTerm {
  uint32_t fmask;
  uint32_t zmask;
  float re;
  float im;
  int32_t slot;
  int32_t pad;
};
static_assert(sizeof(DevTerm) == 24, "DevTerm layout");

// Wide-mask variant for states larger than one tile (masks over all n bits).
struct alignas(8) DevTermWide {
  uint64_t fmask;
  uint64_t zmask;
  float re;
  float im;
  int32_t slot;
  int32_t pad;
};

// One pass over a multi-tile state: tile-local bit j <-> global bit tbit[j].
constexpr int kMaxTile = 14;
struct alignas(8) DevPass {
  int32_t op_begin, op_end;
  int32_t L;          // tile qubits
  int32_t nlow;       // tile bits 0..nlow-1 are global bits 0..nlow-1
  uint8_t tbit[kMaxTile + 2];
  uint64_t tilemask;  // global bits covered by the tile
  // Streaming path: 1 + index of the dense (k >= 5 qubit) fused gate this pass
  // applies as an MFMA contraction instead of ops [op_begin, op_end); 0 if none.
  int32_t dense;
  int32_t pad;
};

// One circuit of a batched launch (LDS-resident: n <= L of the launch).
struct alignas(8) DevCircuit {
  int32_t n;
  int32_t row;            // output row
  int32_t fwd_begin, fwd_end;
  int32_t term_begin, term_end;   // expectation terms (slot = column)
  int32_t lam_begin, lam_end;     // adjoint lambda terms (weights folded in)
  int32_t rev_begin, rev_end;     // adjoint reverse ops
  int32_t flags;          // bit0: empty program, bit1: tiled
  int32_t other_begin;    // inner product: range into the `others` table
  int32_t other_end;
  // Tiled circuits (wider than the LDS buffer): the state lives in the
  // workgroup's global scratch and ops run pass by pass on LDS tiles.
  int32_t fwd_pass_begin, fwd_pass_end;
  int32_t rev_pass_begin, rev_pass_end;
};

// Inner product: one paired ("other") program of a row.
struct alignas(8) DevOther {
  int32_t op_begin, op_end;
  int32_t col;            // output column
  float weight;           // inner-product gradient: downstream weight g[row, col]
};

// Kernel modes of the persistent batch kernel.
enum BatchMode : int {
  kModeExpectation = 0,
  kModeAdjoint = 1,
  kModeState = 2,
  kModeSamples = 3,
  kModeSampledExpectation = 4,
  kModeInnerProduct = 5,
  kModeInnerProductGrad = 6,
  // Noisy trajectory kernel (noisy_kernels.hip).
  kModeNoisyExpectation = 7,
  kModeNoisySampledExpectation = 8,
  kModeNoisySamples = 9,
};

// Noisy trajectories: one work item of the trajectory kernel, trajectories
// [t0, t1) of circuit `circuit` (an index into the plan's DevCircuit table),
// whose program has `nq` qubits.
struct alignas(8) DevChunk {
  int32_t circuit;
  int32_t t0, t1;
  int32_t nq;
};

constexpr int kFlagEmpty = 1;
constexpr int kFlagTiled = 2;
constexpr int kFlagWave = 4;   // runs in wave mode (ops planned for kWaveRegBits)

}  // namespace tfqx
