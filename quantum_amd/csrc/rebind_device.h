// Device-side re-evaluation of a plan's symbolic matrices for new symbol
// values.  The host keeps, per plan, flat tables of the gates and op recipes
// that depend on symbols (the same recipes tfqx_plan_rebind evaluates on the
// host); a launch of rebind_kernel reads the symbol values from the workspace
// and rewrites those matrices in place, so a call with new values uploads only
// the values (B x S floats) instead of the plan's matrices.
//
// Arithmetic follows the host recipes operation for operation (double
// precision, products rounded one at a time, no fused multiply-add); only the
// device's double-precision sin/cos may differ from the host library's in the
// last bit, so the float32 matrices agree with the host's to float rounding.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tfqx {

// One gate of a recipe: kind, targets, controls and float32 arguments; the
// symbolic arguments are replaced by values[row * S + sym_col[j]].
struct RbGate {
  int32_t kind, nt, nsym, pad0;
  int32_t sym_pidx[2], sym_col[2];
  int32_t bits[2];
  uint64_t cmask, cval;
  float params[5];
  float pad1[3];
};

// One symbolic op: kind 0 fused block, 1 single gate (controls outside),
// 2 parameterised gate (+ derivative matrices), 3 fused block containing
// parameterised gates (+ derivative products).  Gates are
// gate_base(row) + gidx[gidx_base(row) + gate_begin .. gate_end).
struct RbOp {
  int32_t kind, k, diag, nderiv;
  int32_t bits[4];
  int32_t gate_begin, gate_end;
  uint32_t mat, dmat[2];
  int32_t dgate[2], dsym[2];
  int32_t row;  // per-row tables: the op's row; shared tables: unused
  int32_t pad;
};

// One batch row: its gate tables and matrix base in the plan's array (and,
// for shared tables, the op range every row evaluates).
struct RbRow {
  int32_t op_begin, op_end;
  int32_t gidx_base, gate_base;
  uint32_t mat_base;
  int32_t vrow;
  int32_t pad[2];
};

// Eigen components of the eigen gates (H .. ISWAP): U(t, gs) =
// sum_k e^{i pi t (lam_k + gs)} P_k.  P is row-major, (re, im) interleaved.
struct RbEigen {
  double count;
  double lam[3];
  double P[3][32];
};
constexpr int kRbEigenKinds = 13;

struct RebindArgs {
  const float* values;       // [rows, S]
  const RbEigen* eigen;      // [kRbEigenKinds]
  const RbRow* rows;
  const RbOp* ops;
  const int32_t* gidx;
  const RbGate* gates;
  float2* mats;              // the plan's matrix array in the workspace
  int num_rows;
  int num_ops;               // entries of `ops`
  int shared;                // 1: every row evaluates all `ops` (one program tiled)
  int num_symbols;
  int tfq_fd;                // derivative matrices: TFQ's float32 central difference
};

// One thread per (row, op); kmax is the widest recipe op (1..4).
hipError_t LaunchRebind(const RebindArgs& a, int kmax, hipStream_t stream);

}  // namespace tfqx
