// rebind_kernel: a plan's symbolic matrices re-evaluated on the device for
// new symbol values (tables described in rebind_device.h).
//
// Every function here is the device counterpart of a host function that
// builds the same matrices when a plan is created or rebound on the host
// (circuit.cc: UnitaryOrDeriv, EigenEval, GateDerivative; plan.cc: ApplyLeft,
// BlockMatrix, ParamBlockMatrices, Embed, WriteMatrix, RebindRow), with the
// same operations in the same order.  Contraction is off so every double
// product and sum is rounded on its own, as in the host build.
#include "rebind_device.h"

#pragma clang fp contract(off)

namespace tfqx {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr float kGradEps = 5e-3f;  // adj_util.cc:32
constexpr int kRbThreads = 64;

struct dc {
  double re, im;
};

__device__ inline dc Mul(dc a, dc b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline dc Scale(dc a, double s) { return {a.re * s, a.im * s}; }
__device__ inline dc Add(dc a, dc b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline dc Expi(double y) { return {cos(y), sin(y)}; }  // e^{iy}
__device__ inline double Mul32(float a, float b) { return double(a * b); }

__device__ inline int Targets(int kind) {
  return (kind == 0 || (kind >= 2 && kind <= 5) || kind == 13) ? 1 : 2;
}

// U(t, gs) = sum_k e^{i pi t (lam_k + gs)} P_k, or dU/dt (deriv); DIM is the
// gate's matrix dimension (2 or 4), fixed so U stays in registers.
template <int DIM>
__device__ void EigenEvalT(const RbEigen& e, double t, double gs, bool deriv, dc* U) {
#pragma unroll
  for (int i = 0; i < DIM * DIM; ++i) U[i] = {0.0, 0.0};
  const int count = int(e.count);
  for (int k = 0; k < count; ++k) {
    const double ph = kPi * t * (e.lam[k] + gs);
    dc w = Expi(ph);
    if (deriv) w = Mul(w, dc{0.0, kPi * (e.lam[k] + gs)});
#pragma unroll
    for (int i = 0; i < DIM * DIM; ++i)
      U[i] = Add(U[i], Mul(w, dc{e.P[k][2 * i], e.P[k][2 * i + 1]}));
  }
}

__device__ inline void EigenEval(const RbEigen* table, int kind, double t, double gs, bool deriv,
                                 dc* U) {
  if (Targets(kind) == 1)
    EigenEvalT<2>(table[kind], t, gs, deriv, U);
  else
    EigenEvalT<4>(table[kind], t, gs, deriv, U);
}

// pidx < 0: the gate's unitary; otherwise its exact derivative wrt p[pidx].
__device__ void UnitaryOrDeriv(const RbEigen* table, int kind, const float* p, int pidx, dc* U) {
  const dc negI{-0.0, -1.0};
  if (kind == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) U[i] = (pidx < 0 && i % 3 == 0) ? dc{1.0, 0.0} : dc{0.0, 0.0};
    return;
  }
  if (kind == 1) {
#pragma unroll
    for (int i = 0; i < 16; ++i) U[i] = (pidx < 0 && i % 5 == 0) ? dc{1.0, 0.0} : dc{0.0, 0.0};
    return;
  }
  if (kind >= 2 && kind <= 12) {
    const double t = Mul32(p[0], p[1]);
    EigenEval(table, kind, t, double(p[2]), pidx >= 0, U);
    if (pidx >= 0) {
#pragma unroll
      for (int i = 0; i < 16; ++i) U[i] = Scale(U[i], double(p[1]));
    }
    return;
  }
  if (kind == 13) {  // PhasedX
    const double ph = Mul32(p[0], p[1]);
    const double t = Mul32(p[2], p[3]);
    dc X[4];
    EigenEval(table, 3, t, double(p[4]), pidx == 2, X);
    const dc em = Expi(-kPi * ph), ep = Expi(kPi * ph);
    U[0] = X[0];
    U[1] = Mul(X[1], em);
    U[2] = Mul(X[2], ep);
    U[3] = X[3];
    if (pidx == 0) {
      U[0] = {0.0, 0.0};
      U[3] = {0.0, 0.0};
      U[1] = Mul(U[1], Scale(dc{0.0, -kPi}, double(p[1])));
      U[2] = Mul(U[2], Scale(dc{0.0, kPi}, double(p[1])));
    } else if (pidx == 2) {
#pragma unroll
      for (int i = 0; i < 4; ++i) U[i] = Scale(U[i], double(p[3]));
    }
    return;
  }
  if (kind == 14) {  // FSim
    const double th = Mul32(p[0], p[1]);
    const double phi = Mul32(p[2], p[3]);
#pragma unroll
    for (int i = 0; i < 16; ++i) U[i] = {0.0, 0.0};
    const double c = cos(th), s = sin(th);
    if (pidx < 0) {
      U[0] = {1.0, 0.0};
      U[5] = {c, 0.0};
      U[6] = Scale(negI, s);
      U[9] = Scale(negI, s);
      U[10] = {c, 0.0};
      U[15] = Expi(-phi);
    } else if (pidx == 0) {
      const double sc = p[1];
      U[5] = {-s * sc, 0.0};
      U[6] = Scale(Scale(negI, c), sc);
      U[9] = Scale(Scale(negI, c), sc);
      U[10] = {-s * sc, 0.0};
    } else {
      U[15] = Scale(Mul(negI, Expi(-phi)), double(p[3]));
    }
    return;
  }
  // kind 15: PhasedISwap
  const double ph = Mul32(p[0], p[1]);
  const double t = Mul32(p[2], p[3]);
  EigenEval(table, 12, t, 0.0, pidx == 2, U);
  const dc ep = Expi(2 * kPi * ph), em = Expi(-2 * kPi * ph);
  U[6] = Mul(U[6], ep);
  U[9] = Mul(U[9], em);
  if (pidx == 0) {
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if (i != 6 && i != 9) U[i] = {0.0, 0.0};
    U[6] = Mul(U[6], Scale(dc{0.0, 2 * kPi}, double(p[1])));
    U[9] = Mul(U[9], Scale(dc{0.0, -2 * kPi}, double(p[1])));
  } else if (pidx == 2) {
#pragma unroll
    for (int i = 0; i < 16; ++i) U[i] = Scale(U[i], double(p[3]));
  }
}

// TFQ's gradient gates (adj_util.cc:175-302) or the exact derivative.  The
// central difference keeps float(U(p + eps)) while U(p - eps) is evaluated
// into dU, then combines them in float as the host does.
__device__ void GateDerivative(const RbEigen* table, int kind, const float* params, int pidx,
                               bool tfq_fd, dc* dU) {
  if (!tfq_fd) {
    UnitaryOrDeriv(table, kind, params, pidx, dU);
    return;
  }
  float pp[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) pp[i] = i == pidx ? params[i] + kGradEps : params[i];
  UnitaryOrDeriv(table, kind, pp, -1, dU);
  float fr[16], fi[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    fr[i] = float(dU[i].re);
    fi[i] = float(dU[i].im);
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) pp[i] = i == pidx ? params[i] - kGradEps : params[i];
  UnitaryOrDeriv(table, kind, pp, -1, dU);
  const float scale = float(0.5 / double(kGradEps));
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const float re = fr[i] - float(dU[i].re);
    const float im = fi[i] - float(dU[i].im);
    dU[i] = {double(re * scale), double(im * scale)};
  }
}

struct LGate {
  int kind, nt;
  int bits[2];
  uint64_t cmask, cval;
  float params[5];
  int sym_pidx[2];
};

// The gate's arguments with the row's symbol values substituted.
__device__ void LoadGate(const RbGate& src, const float* v, LGate* g) {
  g->kind = src.kind;
  g->nt = src.nt;
  g->bits[0] = src.bits[0];
  g->bits[1] = src.bits[1];
  g->cmask = src.cmask;
  g->cval = src.cval;
  for (int j = 0; j < 2; ++j) g->sym_pidx[j] = src.sym_pidx[j];
  const float v0 = src.nsym > 0 ? v[src.sym_col[0]] : 0.0f;
  const float v1 = src.nsym > 1 ? v[src.sym_col[1]] : 0.0f;
#pragma unroll
  for (int j = 0; j < 5; ++j)
    g->params[j] = (src.nsym > 1 && src.sym_pidx[1] == j) ? v1
                 : (src.nsym > 0 && src.sym_pidx[0] == j) ? v0 : src.params[j];
}

// acc <- (gate U embedded over `bits`) * acc; acc is dim x dim, (re, im) interleaved.
__device__ void ApplyLeft(int nt, const int* gbits, uint64_t cmask, uint64_t cval, const dc* U,
                          const int* bits, int k, double* acc) {
  const int dim = 1 << k;
  int upos[2] = {0, 0};
  for (int j = 0; j < nt; ++j)
    for (int m = 0; m < k; ++m)
      if (bits[m] == gbits[j]) upos[j] = m;
  int cm = 0, cv = 0;
  for (int m = 0; m < k; ++m)
    if (cmask >> bits[m] & 1) {
      cm |= 1 << m;
      if (cval >> bits[m] & 1) cv |= 1 << m;
    }
  const int udim = 1 << nt;
  int gmask = 0;
  for (int j = 0; j < nt; ++j) gmask |= 1 << upos[j];
  int rows[4];
  for (int r0 = 0; r0 < dim; ++r0) {
    if (r0 & gmask) continue;
    if ((r0 & cm) != cv) continue;
    for (int sidx = 0; sidx < udim; ++sidx) {
      int r = r0;
      for (int j = 0; j < nt; ++j)
        if (sidx >> (nt - 1 - j) & 1) r |= 1 << upos[j];
      rows[sidx] = r;
    }
    for (int c = 0; c < dim; ++c) {
      double vr[4], vi[4];
      for (int sidx = 0; sidx < udim; ++sidx) {
        vr[sidx] = acc[2 * (rows[sidx] * dim + c)];
        vi[sidx] = acc[2 * (rows[sidx] * dim + c) + 1];
      }
      for (int e = 0; e < udim; ++e) {
        double sr = 0, si = 0;
        for (int b = 0; b < udim; ++b) {
          const double xr = U[e * udim + b].re, xi = U[e * udim + b].im;
          sr += xr * vr[b] - xi * vi[b];
          si += xr * vi[b] + xi * vr[b];
        }
        acc[2 * (rows[e] * dim + c)] = sr;
        acc[2 * (rows[e] * dim + c) + 1] = si;
      }
    }
  }
}

__device__ void Identity(double* acc, int dim) {
  for (int i = 0; i < 2 * dim * dim; ++i) acc[i] = 0;
  for (int i = 0; i < dim; ++i) acc[2 * (i * dim + i)] = 1.0;
}

__device__ void Write(float2* dst, const double* acc, bool diag_only, int dim) {
  if (diag_only) {
    for (int i = 0; i < dim; ++i)
      dst[i] = make_float2(float(acc[2 * (i * dim + i)]), float(acc[2 * (i * dim + i) + 1]));
  } else {
    for (int i = 0; i < dim * dim; ++i) dst[i] = make_float2(float(acc[2 * i]), float(acc[2 * i + 1]));
  }
}

// Single gate over the op's qubits (no controls inside): out = 0 + U placed.
__device__ void Embed(const dc* U, const int* ubits, int uk, const int* obits, int ok, double* out) {
  const int dim = 1 << ok;
  for (int i = 0; i < 2 * dim * dim; ++i) out[i] = 0;
  int upos[2] = {0, 0};
  for (int j = 0; j < uk; ++j)
    for (int m = 0; m < ok; ++m)
      if (obits[m] == ubits[j]) upos[j] = m;
  const int udim = 1 << uk;
  for (int c = 0; c < dim; ++c) {
    int sub = 0;
    for (int j = 0; j < uk; ++j) sub |= ((c >> upos[j]) & 1) << (uk - 1 - j);
    for (int r = 0; r < udim; ++r) {
      int row = c;
      for (int j = 0; j < uk; ++j) {
        const int bit = (r >> (uk - 1 - j)) & 1;
        row = (row & ~(1 << upos[j])) | (bit << upos[j]);
      }
      out[2 * (row * dim + c)] += U[r * udim + sub].re;
      out[2 * (row * dim + c) + 1] += U[r * udim + sub].im;
    }
  }
}

// ---------------------------------------------------------------- ops of <= 2 qubits
// The same recipes with every index fixed at compile time, so the block
// matrix and the gate's matrix stay in registers.

// ApplyLeft for a K-qubit block and an NT-qubit gate whose target j sits at
// block position Pj (P0 holds the gate's matrix MSB); controls inside the
// block (cm, cv) are runtime predicates.
template <int K, int NT, int P0, int P1>
__device__ void ApplyLeftT(const dc* U, int cm, int cv, double* acc) {
  constexpr int dim = 1 << K, udim = 1 << NT;
  constexpr int gmask = (1 << P0) | (NT == 2 ? (1 << P1) : 0);
#pragma unroll
  for (int r0 = 0; r0 < dim; ++r0) {
    if (r0 & gmask) continue;
    if ((r0 & cm) != cv) continue;
    int rows[udim];
#pragma unroll
    for (int sidx = 0; sidx < udim; ++sidx) {
      int r = r0;
      if ((sidx >> (NT - 1)) & 1) r |= 1 << P0;
      if (NT == 2 && (sidx & 1)) r |= 1 << P1;
      rows[sidx] = r;
    }
#pragma unroll
    for (int c = 0; c < dim; ++c) {
      double vr[udim], vi[udim];
#pragma unroll
      for (int sidx = 0; sidx < udim; ++sidx) {
        vr[sidx] = acc[2 * (rows[sidx] * dim + c)];
        vi[sidx] = acc[2 * (rows[sidx] * dim + c) + 1];
      }
#pragma unroll
      for (int e = 0; e < udim; ++e) {
        double sr = 0, si = 0;
#pragma unroll
        for (int b = 0; b < udim; ++b) {
          const double xr = U[e * udim + b].re, xi = U[e * udim + b].im;
          sr += xr * vr[b] - xi * vi[b];
          si += xr * vi[b] + xi * vr[b];
        }
        acc[2 * (rows[e] * dim + c)] = sr;
        acc[2 * (rows[e] * dim + c) + 1] = si;
      }
    }
  }
}

template <int K>
__device__ void ApplyLeftK(const LGate& g, const dc* U, const int* bits, double* acc) {
  int upos[2] = {0, 0};
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int m = 0; m < K; ++m)
      if (j < g.nt && bits[m] == g.bits[j]) upos[j] = m;
  int cm = 0, cv = 0;
#pragma unroll
  for (int m = 0; m < K; ++m)
    if (g.cmask >> bits[m] & 1) {
      cm |= 1 << m;
      if (g.cval >> bits[m] & 1) cv |= 1 << m;
    }
  if (K == 1) {
    ApplyLeftT<1, 1, 0, 0>(U, cm, cv, acc);
  } else if (g.nt == 1) {
    if (upos[0] == 0)
      ApplyLeftT<K, 1, 0, 0>(U, cm, cv, acc);
    else
      ApplyLeftT<K, 1, (K > 1 ? 1 : 0), 0>(U, cm, cv, acc);
  } else {
    if (upos[0] == 0)
      ApplyLeftT<K, 2, 0, (K > 1 ? 1 : 0)>(U, cm, cv, acc);
    else
      ApplyLeftT<K, 2, (K > 1 ? 1 : 0), 0>(U, cm, cv, acc);
  }
}

template <int K>
__device__ void IdentityK(double* acc) {
  constexpr int dim = 1 << K;
#pragma unroll
  for (int i = 0; i < 2 * dim * dim; ++i) acc[i] = 0;
#pragma unroll
  for (int i = 0; i < dim; ++i) acc[2 * (i * dim + i)] = 1.0;
}

template <int K>
__device__ void WriteK(float2* dst, const double* acc, bool diag_only) {
  constexpr int dim = 1 << K;
  if (diag_only) {
#pragma unroll
    for (int i = 0; i < dim; ++i)
      dst[i] = make_float2(float(acc[2 * (i * dim + i)]), float(acc[2 * (i * dim + i) + 1]));
  } else {
#pragma unroll
    for (int i = 0; i < dim * dim; ++i) dst[i] = make_float2(float(acc[2 * i]), float(acc[2 * i + 1]));
  }
}

// One symbolic op of <= 2 qubits.  Kind 3's derivative products recompute the
// running prefix up to their gate (the same operations the host's single
// sweep performs before copying it), so each needs one accumulator only.
// Kinds 1 and 2 place the gate by applying it to the identity: the entries
// equal the host's Embed (x - 0*y and 0 + x are x).
template <int K>
__device__ void EvalOpK(const RebindArgs& a, const RbOp& op, const RbRow& row, const float* v,
                        float2* mats) {
  const int32_t* gl = a.gidx + row.gidx_base + op.gate_begin;
  const int ng = op.gate_end - op.gate_begin;
  double acc[2 * (1 << K) * (1 << K)];
  dc U[16];
  LGate g;
  if (op.kind == 0 || op.kind == 3) {
    // Pass 0 builds the block's matrix; pass s + 1 its derivative product s.
    for (int pass = 0; pass <= (op.kind == 3 ? op.nderiv : 0); ++pass) {
      const int dg = pass == 0 ? -1 : op.dgate[pass - 1];
      IdentityK<K>(acc);
      for (int t = 0; t < ng; ++t) {
        LoadGate(a.gates[row.gate_base + gl[t]], v, &g);
        if (t == dg)
          GateDerivative(a.eigen, g.kind, g.params, g.sym_pidx[op.dsym[pass - 1]], a.tfq_fd != 0, U);
        else
          UnitaryOrDeriv(a.eigen, g.kind, g.params, -1, U);
        ApplyLeftK<K>(g, U, op.bits, acc);
      }
      WriteK<K>(mats + (pass == 0 ? op.mat : op.dmat[pass - 1]), acc,
                pass == 0 && op.kind == 0 && op.diag != 0);
    }
    return;
  }
  LoadGate(a.gates[row.gate_base + gl[0]], v, &g);
  g.cmask = g.cval = 0;  // recipes of kinds 1 and 2 keep controls outside the matrix
  for (int pass = 0; pass <= op.nderiv; ++pass) {
    if (pass == 0)
      UnitaryOrDeriv(a.eigen, g.kind, g.params, -1, U);
    else
      GateDerivative(a.eigen, g.kind, g.params, g.sym_pidx[pass - 1], a.tfq_fd != 0, U);
    IdentityK<K>(acc);
    ApplyLeftK<K>(g, U, op.bits, acc);
    WriteK<K>(mats + (pass == 0 ? op.mat : op.dmat[pass - 1]), acc, pass == 0 && op.diag != 0);
  }
}

// ---------------------------------------------------------------- ops of 3-4 qubits
template <int DIM>
__device__ void EvalOpWide(const RebindArgs& a, const RbOp& op, const RbRow& row, const float* v,
                           float2* mats) {
  double acc[2 * DIM * DIM];
  const int dim = 1 << op.k;
  const int32_t* gl = a.gidx + row.gidx_base + op.gate_begin;
  const int ng = op.gate_end - op.gate_begin;
  LGate g;
  dc U[16];
  if (op.kind == 0) {
    Identity(acc, dim);
    for (int t = 0; t < ng; ++t) {
      LoadGate(a.gates[row.gate_base + gl[t]], v, &g);
      UnitaryOrDeriv(a.eigen, g.kind, g.params, -1, U);
      ApplyLeft(g.nt, g.bits, g.cmask, g.cval, U, op.bits, op.k, acc);
    }
    Write(mats + op.mat, acc, op.diag != 0, dim);
    return;
  }
  // Kinds 1 and 2 (kind 3 blocks have at most two qubits).
  LoadGate(a.gates[row.gate_base + gl[0]], v, &g);
  UnitaryOrDeriv(a.eigen, g.kind, g.params, -1, U);
  Embed(U, g.bits, g.nt, op.bits, op.k, acc);
  Write(mats + op.mat, acc, op.diag != 0, dim);
  for (int d = 0; d < op.nderiv; ++d) {
    dc dU[16];
    GateDerivative(a.eigen, g.kind, g.params, g.sym_pidx[d], a.tfq_fd != 0, dU);
    Embed(dU, g.bits, g.nt, op.bits, op.k, acc);
    Write(mats + op.dmat[d], acc, false, dim);
  }
}

// One thread per (row, op): per-row tables list every row's ops (op.row
// names the row); shared tables hold one program's ops for every row.
template <int DIM>
__global__ __launch_bounds__(kRbThreads, 2) void rebind_kernel(RebindArgs a) {
  const long long i = (long long)blockIdx.x * kRbThreads + threadIdx.x;
  const long long items = a.shared ? (long long)a.num_ops * a.num_rows : a.num_ops;
  if (i >= items) return;
  const int oi = a.shared ? int(i % a.num_ops) : int(i);
  const RbOp op = a.ops[oi];
  const RbRow row = a.rows[a.shared ? int(i / a.num_ops) : op.row];
  const float* v = a.values + size_t(row.vrow) * a.num_symbols;
  float2* mats = a.mats + row.mat_base;
  if (op.k == 1)
    EvalOpK<1>(a, op, row, v, mats);
  else if (op.k == 2)
    EvalOpK<2>(a, op, row, v, mats);
  else if (DIM > 4)
    EvalOpWide<DIM>(a, op, row, v, mats);
}

}  // namespace

hipError_t LaunchRebind(const RebindArgs& a, int kmax, hipStream_t stream) {
  const long long items = a.shared ? (long long)a.num_ops * a.num_rows : a.num_ops;
  if (items <= 0) return hipSuccess;
  const dim3 grid(unsigned((items + kRbThreads - 1) / kRbThreads)), block(kRbThreads);
  if (kmax <= 2)
    hipLaunchKernelGGL(rebind_kernel<4>, grid, block, 0, stream, a);
  else if (kmax == 3)
    hipLaunchKernelGGL(rebind_kernel<8>, grid, block, 0, stream, a);
  else
    hipLaunchKernelGGL(rebind_kernel<16>, grid, block, 0, stream, a);
  return hipGetLastError();
}

}  // namespace tfqx
