// Host planner (see plan.h).
#include "plan.h"

#include "large.h"

#include <algorithm>
#include <tuple>
#include <atomic>
#include <cmath>
#include <numeric>
#include <thread>
#include <unordered_map>

namespace tfqx {
namespace {

inline int Popc(uint64_t x) { return __builtin_popcountll(x); }

// Embed a (controlled) gate given in MSB-first order on `ubits` into the
// little-endian matrix over ascending `obits`.  All control bits must be in obits.
void Embed(const cdouble* U, const int* ubits, int uk, uint64_t cmask, uint64_t cval,
           const int* obits, int ok, cdouble* out) {
  const int dim = 1 << ok;
  for (int i = 0; i < dim * dim; ++i) out[i] = 0;
  int upos[kMaxFuse + 4];  // position of ubits[j] inside obits
  for (int j = 0; j < uk; ++j)
    for (int m = 0; m < ok; ++m)
      if (obits[m] == ubits[j]) upos[j] = m;
  uint64_t cm_local = 0, cv_local = 0;
  for (int m = 0; m < ok; ++m) {
    if (cmask >> obits[m] & 1) {
      cm_local |= 1u << m;
      if (cval >> obits[m] & 1) cv_local |= 1u << m;
    }
  }
  const int udim = 1 << uk;
  for (int c = 0; c < dim; ++c) {
    if ((c & cm_local) != cv_local) {
      out[c * dim + c] += 1.0;
      continue;
    }
    int sub = 0;
    for (int j = 0; j < uk; ++j) sub |= ((c >> upos[j]) & 1) << (uk - 1 - j);
    for (int r = 0; r < udim; ++r) {
      int row = c;
      for (int j = 0; j < uk; ++j) {
        const int bit = (r >> (uk - 1 - j)) & 1;
        row = (row & ~(1 << upos[j])) | (bit << upos[j]);
      }
      out[row * dim + c] += U[r * udim + sub];
    }
  }
}

void MatMul(const cdouble* A, const cdouble* B, int dim, cdouble* C) {
  for (int r = 0; r < dim; ++r)
    for (int c = 0; c < dim; ++c) {
      cdouble s = 0;
      for (int k = 0; k < dim; ++k) s += A[r * dim + k] * B[k * dim + c];
      C[r * dim + c] = s;
    }
}

bool IsDiagonal(const std::vector<cdouble>& M, int dim) {
  for (int r = 0; r < dim; ++r)
    for (int c = 0; c < dim; ++c)
      if (r != c && M[r * dim + c] != cdouble(0, 0)) return false;
  return true;
}

uint64_t GateMask(const Gate& g) {
  uint64_t m = g.cmask;
  for (int j = 0; j < g.nt; ++j) m |= uint64_t(1) << g.bits[j];
  return m;
}

void MaskToBits(uint64_t m, int* bits, int* k) {
  *k = 0;
  for (int b = 0; b < 64; ++b)
    if (m >> b & 1) bits[(*k)++] = b;
}

bool GatesDiagonal(const Circuit& c, const std::vector<int>& gates) {
  for (int gi : gates)
    if (!KindDiagonal(c.gates[gi].kind)) return false;
  return true;
}

// acc <- G * acc for one (controlled) gate acting inside a k-bit block
// (`bits` ascending, little-endian block index; controls must lie in `bits`).
// Works on the udim-row groups the gate mixes, with explicit complex
// arithmetic (std::complex products take the slow C99 Annex G path).
void ApplyLeft(const Gate& g, const int* bits, int k, double* acc /* re,im interleaved */) {
  const int dim = 1 << k;
  int upos[2];
  for (int j = 0; j < g.nt; ++j)
    for (int m = 0; m < k; ++m)
      if (bits[m] == g.bits[j]) upos[j] = m;
  int cm = 0, cv = 0;
  for (int m = 0; m < k; ++m)
    if (g.cmask >> bits[m] & 1) {
      cm |= 1 << m;
      if (g.cval >> bits[m] & 1) cv |= 1 << m;
    }
  const int udim = 1 << g.nt;
  int gmask = 0;
  for (int j = 0; j < g.nt; ++j) gmask |= 1 << upos[j];
  int rows[4];
  double ur[16], ui[16];
  for (int i = 0; i < udim * udim; ++i) {
    ur[i] = g.U[i].real();
    ui[i] = g.U[i].imag();
  }
  for (int r0 = 0; r0 < dim; ++r0) {
    if (r0 & gmask) continue;
    if ((r0 & cm) != cv) continue;
    // rows[s]: sub-index s (MSB = g.bits[0]) placed on the gate's positions
    for (int sidx = 0; sidx < udim; ++sidx) {
      int r = r0;
      for (int j = 0; j < g.nt; ++j)
        if (sidx >> (g.nt - 1 - j) & 1) r |= 1 << upos[j];
      rows[sidx] = r;
    }
    for (int c = 0; c < dim; ++c) {
      double vr[4], vi[4];
      for (int sidx = 0; sidx < udim; ++sidx) {
        vr[sidx] = acc[2 * (rows[sidx] * dim + c)];
        vi[sidx] = acc[2 * (rows[sidx] * dim + c) + 1];
      }
      for (int a = 0; a < udim; ++a) {
        double sr = 0, si = 0;
        for (int b = 0; b < udim; ++b) {
          const double xr = ur[a * udim + b], xi = ui[a * udim + b];
          sr += xr * vr[b] - xi * vi[b];
          si += xr * vi[b] + xi * vr[b];
        }
        acc[2 * (rows[a] * dim + c)] = sr;
        acc[2 * (rows[a] * dim + c) + 1] = si;
      }
    }
  }
}

// Product (application order) of the gates' embedded matrices over `bits`.
void BlockMatrix(const std::vector<Gate>& gs, const int* gates, int ngates, const int* bits, int k,
                 cdouble* out) {
  const int dim = 1 << k;
  double small[2 * 256];
  std::vector<double> big;
  double* acc = small;
  if (dim > 16) {
    big.resize(size_t(2) * dim * dim);
    acc = big.data();
  }
  for (int i = 0; i < 2 * dim * dim; ++i) acc[i] = 0;
  for (int i = 0; i < dim; ++i) acc[2 * (i * dim + i)] = 1.0;
  for (int t = 0; t < ngates; ++t) ApplyLeft(gs[gates[t]], bits, k, acc);
  for (int i = 0; i < dim * dim; ++i) out[i] = cdouble(acc[2 * i], acc[2 * i + 1]);
}

// A parameterised block's product and derivative products in one sweep: M =
// U_{m-1} ... U_0 and, for derivative s of gate dgate[s] with derivative
// matrix dU[s] (the gate's layout), U_{m-1} ... dU[s] ... U_0.  Each
// derivative product copies the running prefix at its gate and then sees the
// same multiplications as a product built from the identity, so the results
// equal BlockMatrix's bit for bit.
void ParamBlockMatrices(const std::vector<Gate>& gs, const int* gates, int ngates, const int* bits,
                        int k, int nderiv, const int* dgate, const cdouble (*dU)[16],
                        cdouble* M, cdouble* const* dM) {
  const int dim = 1 << k, n2 = 2 * dim * dim;
  double acc[2 * 16], dacc[kMaxDeriv][2 * 16];
  for (int i = 0; i < n2; ++i) acc[i] = 0;
  for (int i = 0; i < dim; ++i) acc[2 * (i * dim + i)] = 1.0;
  bool live[kMaxDeriv] = {false, false};
  for (int t = 0; t < ngates; ++t) {
    const Gate& g = gs[size_t(gates[t])];
    for (int s = 0; s < nderiv; ++s) {
      if (dgate[s] != t) continue;
      for (int i = 0; i < n2; ++i) dacc[s][i] = acc[i];
      Gate dg = g;
      for (int e = 0; e < 16; ++e) dg.U[e] = dU[s][e];
      ApplyLeft(dg, bits, k, dacc[s]);
      live[s] = true;
    }
    for (int s = 0; s < nderiv; ++s)
      if (live[s] && dgate[s] != t) ApplyLeft(g, bits, k, dacc[s]);
    ApplyLeft(g, bits, k, acc);
  }
  for (int i = 0; i < dim * dim; ++i) M[i] = cdouble(acc[2 * i], acc[2 * i + 1]);
  for (int s = 0; s < nderiv; ++s)
    for (int i = 0; i < dim * dim; ++i) dM[s][i] = cdouble(dacc[s][2 * i], dacc[s][2 * i + 1]);
}

FusedOp BlockToOp(const Circuit& c, const std::vector<int>& gates, uint64_t qmask,
                  bool structural_diag) {
  FusedOp op;
  MaskToBits(qmask, op.bits, &op.k);
  const int dim = 1 << op.k;
  op.M.assign(size_t(dim) * dim, 0.0);
  BlockMatrix(c.gates, gates.data(), int(gates.size()), op.bits, op.k, op.M.data());
  op.diag = structural_diag ? GatesDiagonal(c, gates) : IsDiagonal(op.M, dim);
  op.gates = gates;
  return op;
}

}  // namespace

namespace {
std::atomic<int> g_reg_engine(2), g_tile_single(0), g_tile_two(0);
}  // namespace

Tuning GetTuning() {
  Tuning t;
  t.reg_engine = g_reg_engine.load();
  t.tile_single = g_tile_single.load();
  t.tile_two = g_tile_two.load();
  return t;
}

void SetTuning(const Tuning& t) {
  g_reg_engine.store(t.reg_engine);
  g_tile_single.store(t.tile_single);
  g_tile_two.store(t.tile_two);
}

void InsertRemaps(std::vector<DevOp>* ops, size_t begin, int L, int max_r) {
  if (!g_reg_engine.load()) return;
  InsertRemapsR(ops, begin, L, RegBitsForTile(L, max_r));
}

void InsertRemapsR(std::vector<DevOp>* ops, size_t begin, int L, int R) {
  if (R == 0 || begin >= ops->size()) return;
  for (size_t i = begin; i < ops->size(); ++i)
    if ((*ops)[i].k > 2) return;
  std::vector<DevOp> in(ops->begin() + long(begin), ops->end());
  ops->resize(begin);
  const size_t n = in.size();
  // Support (targets and controls) and register needs (targets of
  // non-diagonal ops) of every op.  Two ops commute when their supports are
  // disjoint or both are diagonal; a segment may take an op ahead of earlier
  // ops it commutes with.
  std::vector<uint32_t> sup(n), need(n);
  std::vector<char> diag(n), done(n, 0);
  // Noise channels keep their order among themselves (a channel's Kraus
  // choice depends on the state the earlier channels left): they share a
  // pseudo-qubit (bit 31, above any tile bit) in their support.
  constexpr uint32_t kChannelBit = 1u << 31;
  for (size_t i = 0; i < n; ++i) {
    const DevOp& o = in[i];
    uint32_t T = 0;
    for (int t = 0; t < o.k; ++t) T |= 1u << o.q[t];
    diag[i] = o.type == kOpDiag;
    sup[i] = T | o.cmask | (o.type == kOpChannel ? kChannelBit : 0u);
    need[i] = diag[i] ? 0u : T;
  }
  const uint32_t all = L >= 31 ? ~kChannelBit : (1u << L) - 1u;
  // Ops scanned past the first pending one per segment (bounds planning time).
  constexpr size_t kWindow = 96;
  // Scan pending ops in order from `first`: an op joins the segment when it
  // commutes with every op skipped before it and its targets fit the register
  // set (grown up to R qubits when `grow`).  Returns the register set and,
  // when `take` is given, the joined ops; `cost` counts joined non-diagonal ops.
  auto scan = [&](size_t first, uint32_t S, bool grow, std::vector<size_t>* take, int* cost) {
    uint32_t blocked = 0;  // supports of skipped ops (diagonal ones tracked apart)
    uint32_t blocked_diag = 0;
    int c = 0;
    size_t seen = 0;
    for (size_t j = first; j < n && seen < kWindow; ++j) {
      if (done[j]) continue;
      ++seen;
      // A diagonal op commutes with skipped diagonal ops.
      const uint32_t hit = diag[j] ? blocked : (blocked | blocked_diag);
      bool ok = (sup[j] & hit) == 0;
      if (ok && need[j]) ok = grow ? Popc(S | need[j]) <= R : (need[j] & ~S) == 0;
      if (!ok) {
        if (diag[j]) blocked_diag |= sup[j];
        else blocked |= sup[j];
        if (((blocked | blocked_diag) & all) == all && (blocked & all) == all) break;
        continue;
      }
      if (grow) S |= need[j];
      if (need[j]) ++c;
      if (take) take->push_back(j);
    }
    if (cost) *cost = c;
    return S;
  };
  size_t first = 0;
  std::vector<size_t> take;
  while (first < n) {
    // Seed the register set with the targets of one of the first few pending
    // non-diagonal ops (up to 8) and keep the seed whose segment takes the most ops.
    uint32_t best_S = 0;
    int best = -1;
    int seeds = 0;
    for (size_t j = first; j < n && seeds < 8; ++j) {
      if (done[j] || !need[j]) continue;
      ++seeds;
      int c = 0;
      const uint32_t S = scan(first, need[j], true, nullptr, &c);
      if (c > best) {
        best = c;
        best_S = S;
      }
    }
    uint32_t S = best_S;
    if (best < 0) S = scan(first, 0, true, nullptr, nullptr);  // only diagonal ops left
    // Pad with the targets of the next ops that did not fit, then low bits.
    for (size_t j = first; j < n && Popc(S) < R; ++j)
      if (!done[j] && (need[j] & ~S) && Popc(S | need[j]) <= R) S |= need[j];
    for (int b = 0; b < L && Popc(S) < R; ++b) S |= 1u << b;
    take.clear();
    scan(first, S, false, &take, nullptr);
    DevOp rm{};
    rm.type = kOpRemap;
    rm.cmask = S;
    ops->push_back(rm);
    for (size_t t : take) {
      ops->push_back(in[t]);
      done[t] = 1;
    }
    while (first < n && done[first]) ++first;
  }
}

bool KindDiagonal(GateKind kind) {
  return kind == kGateI1 || kind == kGateI2 || kind == kGateZ || kind == kGateZZ ||
         kind == kGateCZ;
}

FusedOp GateToOp(const Gate& g, bool keep_controls_outside, bool structural_diag) {
  FusedOp op;
  uint64_t m = 0;
  for (int j = 0; j < g.nt; ++j) m |= uint64_t(1) << g.bits[j];
  if (!keep_controls_outside) m |= g.cmask;
  MaskToBits(m, op.bits, &op.k);
  const int dim = 1 << op.k;
  op.M.assign(dim * dim, 0.0);
  if (keep_controls_outside) {
    Embed(g.U, g.bits, g.nt, 0, 0, op.bits, op.k, op.M.data());
    op.cmask = g.cmask;
    op.cval = g.cval;
  } else {
    Embed(g.U, g.bits, g.nt, g.cmask, g.cval, op.bits, op.k, op.M.data());
  }
  op.diag = structural_diag ? KindDiagonal(g.kind) : IsDiagonal(op.M, dim);
  return op;
}

void FuseGates(const Circuit& c, int begin, int end, const FuseOptions& opt,
               std::vector<FusedOp>* out) {
  struct Block {
    std::vector<int> gates;
    uint64_t qmask;
    int created;
    bool open;
  };
  std::vector<Block> blocks;
  std::vector<int> owner(64, -1);
  auto close = [&](int bi) {
    Block& b = blocks[bi];
    if (!b.open) return;
    b.open = false;
    for (int q = 0; q < 64; ++q)
      if (b.qmask >> q & 1) owner[q] = -1;
    out->push_back(BlockToOp(c, b.gates, b.qmask, opt.structural_diag));
  };
  for (int gi = begin; gi < end; ++gi) {
    const Gate& g = c.gates[gi];
    const uint64_t Q = GateMask(g);
    std::vector<int> touching;
    for (int q = 0; q < 64; ++q)
      if ((Q >> q & 1) && owner[q] >= 0 &&
          std::find(touching.begin(), touching.end(), owner[q]) == touching.end())
        touching.push_back(owner[q]);
    std::sort(touching.begin(), touching.end());
    if (Popc(Q) > opt.max_fuse) {
      // Unfusable: flush what it touches, emit alone with controls outside.
      for (int bi : touching) close(bi);
      out->push_back(GateToOp(g, true, opt.structural_diag));
      out->back().gates.assign(1, gi);
      continue;
    }
    uint64_t U = Q;
    for (int bi : touching) U |= blocks[bi].qmask;
    if (Popc(U) <= opt.max_fuse) {
      int keep;
      if (touching.empty()) {
        blocks.push_back(Block{{}, 0, int(blocks.size()), true});
        keep = int(blocks.size()) - 1;
      } else {
        keep = touching[0];
        for (size_t t = 1; t < touching.size(); ++t) {
          Block& o = blocks[touching[t]];
          blocks[keep].gates.insert(blocks[keep].gates.end(), o.gates.begin(), o.gates.end());
          blocks[keep].qmask |= o.qmask;
          o.open = false;
          o.gates.clear();
        }
      }
      blocks[keep].gates.push_back(gi);
      blocks[keep].qmask |= Q;
      for (int q = 0; q < 64; ++q)
        if (blocks[keep].qmask >> q & 1) owner[q] = keep;
    } else {
      // Open blocks within the gate's own qubits join the new block (their
      // gates first); the others close.  Open blocks are disjoint, so the
      // absorbed gates commute with those of the closed blocks.
      Block nb{{}, Q, int(blocks.size()), true};
      for (int bi : touching) {
        Block& o = blocks[bi];
        if ((o.qmask & ~Q) == 0) {
          nb.gates.insert(nb.gates.end(), o.gates.begin(), o.gates.end());
          o.open = false;
          o.gates.clear();
        } else {
          close(bi);
        }
      }
      nb.gates.push_back(gi);
      blocks.push_back(nb);
      for (int q = 0; q < 64; ++q)
        if (Q >> q & 1) owner[q] = int(blocks.size()) - 1;
    }
  }
  for (size_t bi = 0; bi < blocks.size(); ++bi) close(int(bi));
}

namespace {

std::atomic<int> g_rev_fuse(1);

// A parameterised gate as an op of its own, controls outside the matrix.
FusedOp ParamGateOp(const Circuit& c, int gi, const FuseOptions& opt, bool tfq_fd) {
  const Gate& g = c.gates[size_t(gi)];
  FusedOp f = GateToOp(g, true, opt.structural_diag);
  f.gates.assign(1, gi);
  f.param = true;
  f.nderiv = g.nsym;
  const int dim = 1 << f.k;
  for (int d = 0; d < g.nsym; ++d) {
    cdouble dU[16];
    GateDerivative(g.kind, g.params, g.sym_pidx[d], tfq_fd, dU);
    f.dM[d].assign(size_t(dim) * dim, 0.0);
    Embed(dU, g.bits, g.nt, 0, 0, f.bits, f.k, f.dM[d].data());
    f.col[d] = g.sym_col[d];
    f.dgate[d] = 0;
    f.dsym[d] = d;
  }
  return f;
}

// Derivative matrices of a fused block that holds parameterised gates (none
// of them controlled): for each symbol of each such gate, the block's product
// with that gate's matrix replaced by its derivative.
void MakeParamBlock(const Circuit& c, bool tfq_fd, FusedOp* f) {
  const int dim = 1 << f->k;
  f->param = true;
  f->fused_param = true;
  f->diag = false;
  f->nderiv = 0;
  int dgate[kMaxDeriv];
  cdouble dU[kMaxDeriv][16];
  for (size_t j = 0; j < f->gates.size(); ++j) {
    const Gate& g = c.gates[size_t(f->gates[j])];
    for (int d = 0; d < g.nsym; ++d) {
      const int s = f->nderiv++;
      GateDerivative(g.kind, g.params, g.sym_pidx[d], tfq_fd, dU[s]);
      dgate[s] = int(j);
      f->col[s] = g.sym_col[d];
      f->dgate[s] = int(j);
      f->dsym[s] = d;
    }
  }
  std::vector<cdouble> M(size_t(dim) * dim);
  cdouble* dM[kMaxDeriv];
  for (int s = 0; s < f->nderiv; ++s) {
    f->dM[s].assign(size_t(dim) * dim, 0.0);
    dM[s] = f->dM[s].data();
  }
  ParamBlockMatrices(c.gates, f->gates.data(), int(f->gates.size()), f->bits, f->k, f->nderiv,
                     dgate, dU, M.data(), dM);
  f->M.swap(M);
}

// The frontier fuser of FuseGates with parameterised gates inside blocks:
// blocks holding them stay within two qubits and kMaxDeriv derivatives.
// Blocks come out in a valid application order; the caller reverses them.
void FuseWithParams(const Circuit& c, const FuseOptions& opt, bool tfq_fd,
                    std::vector<FusedOp>* out) {
  struct Block {
    std::vector<int> gates;
    uint64_t qmask;
    int nderiv;
    bool open;
  };
  const int pmax = std::min(opt.max_fuse, 2);
  std::vector<Block> blocks;
  std::vector<int> owner(64, -1);
  auto close = [&](int bi) {
    Block& b = blocks[size_t(bi)];
    if (!b.open) return;
    b.open = false;
    for (int q = 0; q < 64; ++q)
      if (b.qmask >> q & 1) owner[size_t(q)] = -1;
    out->push_back(BlockToOp(c, b.gates, b.qmask, opt.structural_diag));
    if (b.nderiv > 0) MakeParamBlock(c, tfq_fd, &out->back());
  };
  auto claim = [&](int bi) {
    for (int q = 0; q < 64; ++q)
      if (blocks[size_t(bi)].qmask >> q & 1) owner[size_t(q)] = bi;
  };
  for (int gi = 0; gi < int(c.gates.size()); ++gi) {
    const Gate& g = c.gates[size_t(gi)];
    const uint64_t Q = GateMask(g);
    std::vector<int> touching;
    for (int q = 0; q < 64; ++q)
      if ((Q >> q & 1) && owner[size_t(q)] >= 0 &&
          std::find(touching.begin(), touching.end(), owner[size_t(q)]) == touching.end())
        touching.push_back(owner[size_t(q)]);
    std::sort(touching.begin(), touching.end());
    const bool param = g.nsym > 0;
    if ((param && (g.cmask != 0 || Popc(Q) > pmax)) || Popc(Q) > opt.max_fuse) {
      for (int bi : touching) close(bi);
      if (param) {
        out->push_back(ParamGateOp(c, gi, opt, tfq_fd));
      } else {
        out->push_back(GateToOp(g, true, opt.structural_diag));
        out->back().gates.assign(1, gi);
      }
      continue;
    }
    uint64_t U = Q;
    int nd = g.nsym;
    for (int bi : touching) {
      U |= blocks[size_t(bi)].qmask;
      nd += blocks[size_t(bi)].nderiv;
    }
    if (Popc(U) <= (nd > 0 ? pmax : opt.max_fuse) && nd <= kMaxDeriv) {
      int keep;
      if (touching.empty()) {
        blocks.push_back(Block{{}, 0, 0, true});
        keep = int(blocks.size()) - 1;
      } else {
        keep = touching[0];
        for (size_t t = 1; t < touching.size(); ++t) {
          Block& o = blocks[size_t(touching[t])];
          Block& k = blocks[size_t(keep)];
          k.gates.insert(k.gates.end(), o.gates.begin(), o.gates.end());
          k.qmask |= o.qmask;
          k.nderiv += o.nderiv;
          o.open = false;
          o.gates.clear();
        }
      }
      Block& k = blocks[size_t(keep)];
      k.gates.push_back(gi);
      k.qmask |= Q;
      k.nderiv += g.nsym;
      claim(keep);
    } else {
      // As FuseGates: open blocks within the gate's qubits join the new block
      // while the derivative count allows; the others close.
      Block nb{{}, Q, g.nsym, true};
      for (int bi : touching) {
        Block& o = blocks[size_t(bi)];
        if ((o.qmask & ~Q) == 0 && nb.nderiv + o.nderiv <= kMaxDeriv &&
            Popc(Q) <= (nb.nderiv + o.nderiv > 0 ? pmax : opt.max_fuse)) {
          nb.gates.insert(nb.gates.end(), o.gates.begin(), o.gates.end());
          nb.nderiv += o.nderiv;
          o.open = false;
          o.gates.clear();
        } else {
          close(bi);
        }
      }
      nb.gates.push_back(gi);
      blocks.push_back(nb);
      claim(int(blocks.size()) - 1);
    }
  }
  for (size_t bi = 0; bi < blocks.size(); ++bi) close(int(bi));
}

}  // namespace

void SetReverseFusion(int on) { g_rev_fuse.store(on ? 1 : 0); }
int GetReverseFusion() { return g_rev_fuse.load(); }

void BuildReverseOps(const Circuit& c, const FuseOptions& opt, bool tfq_fd,
                     std::vector<FusedOp>* rev) {
  rev->clear();
  if (g_rev_fuse.load()) {
    std::vector<FusedOp> fwd;
    FuseWithParams(c, opt, tfq_fd, &fwd);
    for (size_t i = fwd.size(); i-- > 0;) rev->push_back(std::move(fwd[i]));
    return;
  }
  std::vector<int> pg;
  for (int i = 0; i < int(c.gates.size()); ++i)
    if (c.gates[size_t(i)].nsym > 0) pg.push_back(i);
  int seg_end = int(c.gates.size());
  for (int p = int(pg.size()); p >= 0; --p) {
    const int seg_begin = p == 0 ? 0 : pg[size_t(p) - 1] + 1;
    std::vector<FusedOp> seg;
    FuseGates(c, seg_begin, seg_end, opt, &seg);
    for (int s = int(seg.size()) - 1; s >= 0; --s) rev->push_back(std::move(seg[size_t(s)]));
    if (p == 0) break;
    rev->push_back(ParamGateOp(c, pg[size_t(p) - 1], opt, tfq_fd));
    seg_end = pg[size_t(p) - 1];
  }
}

void AppendRecipe(const FusedOp& f, const Circuit& c, const DevOp& d, SharedRecipe* rec) {
  OpRecipe r;
  r.kind = f.fused_param ? 3 : f.param ? 2 : (f.cmask ? 1 : 0);
  r.k = uint8_t(f.k);
  r.diag = !f.param && f.diag;
  for (int j = 0; j < f.k; ++j) r.bits[j] = f.bits[j];
  r.gate_begin = int(rec->gates.size());
  for (int gi : f.gates) {
    rec->gates.push_back(gi);
    if (c.gates[gi].nsym > 0) r.symbolic = true;
  }
  r.gate_end = int(rec->gates.size());
  r.mat = d.mat;
  r.nderiv = f.param ? f.nderiv : 0;
  for (int i = 0; i < r.nderiv; ++i) {
    r.dmat[i] = d.dmat[i];
    r.dgate[i] = int8_t(f.dgate[i]);
    r.dsym[i] = int8_t(f.dsym[i]);
  }
  rec->ops.push_back(r);
}

namespace {

// One circuit's program, with offsets relative to its own arrays.
struct Piece {
  DevCircuit dc;
  std::vector<DevOther> others;
  std::vector<DevPass> passes;
  std::vector<DevOp> ops;
  std::vector<cfloat> mats;
  std::vector<DevTerm> terms;
  std::vector<float> ident;  // O entries
  double cost = 0;
};

uint32_t PushMatrix(std::vector<cfloat>* mats, const std::vector<cdouble>& M, bool diag_only,
                    int dim) {
  uint32_t off = uint32_t(mats->size());
  if (diag_only) {
    for (int i = 0; i < dim; ++i) mats->push_back(cfloat(float(M[i * dim + i].real()),
                                                         float(M[i * dim + i].imag())));
  } else {
    for (const auto& x : M) mats->push_back(cfloat(float(x.real()), float(x.imag())));
  }
  return off;
}

void WriteMatrix(cfloat* dst, const cdouble* M, bool diag_only, int dim) {
  if (diag_only) {
    for (int i = 0; i < dim; ++i)
      dst[i] = cfloat(float(M[i * dim + i].real()), float(M[i * dim + i].imag()));
  } else {
    for (int i = 0; i < dim * dim; ++i) dst[i] = cfloat(float(M[i].real()), float(M[i].imag()));
  }
}

// `c` and `rec` (both or neither): record how to rebuild this op's matrices.
DevOp MakeDevOp(const FusedOp& f, std::vector<cfloat>* mats, const Circuit* c = nullptr,
                SharedRecipe* rec = nullptr) {
  DevOp d{};
  d.k = uint8_t(f.k);
  for (int j = 0; j < f.k; ++j) d.q[j] = uint8_t(f.bits[j]);
  d.cmask = uint32_t(f.cmask);
  d.cval = uint32_t(f.cval);
  const int dim = 1 << f.k;
  if (f.param) {
    d.type = kOpParam;
    d.mat = PushMatrix(mats, f.M, false, dim);
    d.nderiv = uint8_t(f.nderiv);
    for (int i = 0; i < f.nderiv; ++i) {
      d.dmat[i] = PushMatrix(mats, f.dM[i], false, dim);
      d.col[i] = f.col[i];
    }
  } else {
    d.type = f.diag ? kOpDiag : kOpUnitary;
    d.mat = PushMatrix(mats, f.M, f.diag, dim);
  }
  if (rec) AppendRecipe(f, *c, d, rec);
  return d;
}

// XOR-mask term with weight w * i^{#Y} (phase_y) or the plain coefficient
// (sampled expectation, which measures in the rotated Z basis).  `index` is the
// term's position in its PauliSum (sampled-expectation RNG stream).
DevTerm MakeTerm(const Pauli& p, float w, int slot, bool phase_y = true, int index = 0) {
  DevTerm t{};
  t.fmask = uint32_t(p.xmask);
  t.zmask = uint32_t(p.zmask);
  float re = w, im = 0;
  for (int k = 0; phase_y && k < (p.ny & 3); ++k) {
    float nr = -im, ni = re;
    re = nr;
    im = ni;
  }
  t.re = re;
  t.im = im;
  t.slot = slot;
  t.pad = index;
  return t;
}

// Expectation terms (ComputeExpectationQsim, util_qsim.h:152-186; sampled:
// ComputeSampledExpectationQsim, util_qsim.h:199-270) and, for the adjoint, the
// lambda terms (AccumulateOperators, util_qsim.h:363-414: w = g_j * c_t in
// float32, terms with |w| < 1e-5 dropped).  Returns their cost.
double AddTerms(const CircuitJob& job, int num_slots, const PlanOptions& opt, Piece* pc) {
  DevCircuit& dc = pc->dc;
  const bool empty = (dc.flags & kFlagEmpty) != 0;
  pc->terms.clear();
  pc->ident.assign(num_slots, 0.0f);
  dc.term_begin = 0;
  const bool sampled = opt.mode == kModeSampledExpectation;
  if (!empty && job.sums) {
    for (int j = 0; j < num_slots; ++j) {
      int idx = 0;
      for (const auto& p : (*job.sums)[j]) {
        if (p.ops.empty()) {
          pc->ident[j] += p.coef;
        } else {
          pc->terms.push_back(MakeTerm(p, p.coef, j, !sampled, idx));
        }
        ++idx;
      }
    }
  }
  dc.term_end = int(pc->terms.size());
  dc.lam_begin = dc.lam_end = dc.term_end;
  // Adjoint lambda terms (the inner-product gradient builds lambda from its
  // paired programs instead and has no Pauli sums).
  if (opt.adjoint && !empty && job.sums) {
    for (int j = 0; j < num_slots; ++j) {
      const float gj = job.downstream[j];
      for (const auto& p : (*job.sums)[j]) {
        const float w = gj * p.coef;
        if (std::fabs(w) < 1e-5f) {
          if (opt.keep_all_lambda) pc->terms.push_back(MakeTerm(p, 0.0f, 0));
          continue;
        }
        pc->terms.push_back(MakeTerm(p, w, 0));
      }
    }
    dc.lam_end = int(pc->terms.size());
  }
  const double amps = std::ldexp(1.0, std::max(dc.n, 0));
  return amps * double(pc->terms.size());
}

// `rec`: record the matrix recipes (LDS-resident circuits only).
void BuildPiece(const CircuitJob& job, int num_slots, const PlanOptions& opt, Piece* pc,
                SharedRecipe* rec = nullptr) {
  const Circuit& c = *job.circuit;
  DevCircuit& dc = pc->dc;
  dc = DevCircuit{};
  dc.n = c.num_qubits;
  dc.row = job.row;
  const bool empty = c.gates.empty();
  if (empty) dc.flags |= kFlagEmpty;
  const Circuit* rc = rec ? &c : nullptr;
  // Forward program.
  std::vector<FusedOp> fused;
  if (!empty) FuseGates(c, 0, int(c.gates.size()), opt.fuse, &fused);
  const bool tiled = !empty && dc.n > opt.tile_log2;
  if (tiled) dc.flags |= kFlagTiled;
  // Wave mode: every fused op (forward and reverse segments alike, fused at
  // the same width) must fit the register engine's two-qubit ops.
  bool wave = opt.wave && g_reg_engine.load() && !empty && !tiled && dc.n >= kWaveRegBits &&
              dc.n <= kWaveMaxQubits && opt.fuse.max_fuse <= 2;
  for (const auto& f : fused) wave = wave && f.k <= 2;
  if (wave) dc.flags |= kFlagWave;
  auto remaps = [&](size_t begin) {
    if (wave) InsertRemapsR(&pc->ops, begin, dc.n, kWaveRegBits);
    else InsertRemaps(&pc->ops, begin, dc.n);
  };
  dc.fwd_begin = 0;
  dc.fwd_pass_begin = 0;
  if (tiled) {
    // Adjoint plans: the forward holds one state, so its tiles may take the
    // LDS of the reverse sweep's two tiles (one more qubit, five register
    // qubits).
    const bool wide = opt.mode == kModeAdjoint && g_reg_engine.load() >= 2;
    const int fwd_L = wide ? std::min(opt.tile_log2 + 1, dc.n) : opt.tile_log2;
    SchedulePasses(fused, dc.n, fwd_L, &pc->passes, &pc->ops, &pc->mats, wide ? 5 : 4, rc, rec);
  } else {
    for (const auto& f : fused) pc->ops.push_back(MakeDevOp(f, &pc->mats, rc, rec));
    remaps(size_t(dc.fwd_begin));
  }
  dc.fwd_end = int(pc->ops.size());
  dc.fwd_pass_end = int(pc->passes.size());
  double cost = AddTerms(job, num_slots, opt, pc);
  dc.rev_begin = dc.rev_end = dc.fwd_end;
  const double amps = std::ldexp(1.0, std::max(dc.n, 0));
  for (const auto& f : fused) cost += amps * (f.diag ? 1 : (1 << f.k)) + 2048.0;
  if (opt.adjoint && !empty) {
    // Reverse program (BuildReverseOps), last op first.
    dc.rev_begin = int(pc->ops.size());
    dc.rev_pass_begin = int(pc->passes.size());
    std::vector<FusedOp> rev_ops;
    BuildReverseOps(c, opt.fuse, opt.tfq_fd, &rev_ops);
    for (const auto& f : rev_ops) {
      if (f.param) cost += amps * (1 << f.k) * (3 + f.nderiv) + 4096.0;
      else cost += 2 * amps * (f.diag ? 1 : (1 << f.k)) + 2048.0;
    }
    if (tiled) {
      SchedulePasses(rev_ops, dc.n, opt.tile_log2, &pc->passes, &pc->ops, &pc->mats, 4, rc,
                     rec);
    } else {
      for (const auto& f : rev_ops) pc->ops.push_back(MakeDevOp(f, &pc->mats, rc, rec));
      remaps(size_t(dc.rev_begin));
    }
    dc.rev_end = int(pc->ops.size());
    dc.rev_pass_end = int(pc->passes.size());
  }
  if (tiled) cost *= 3.0;  // streams through L2 instead of staying in LDS
  if ((opt.mode == kModeInnerProduct || opt.mode == kModeInnerProductGrad) && !empty &&
      job.others) {
    for (size_t k = 0; k < job.others->size(); ++k) {
      const Circuit& oc = (*job.others)[k];
      std::vector<FusedOp> of;
      FuseGates(oc, 0, int(oc.gates.size()), opt.fuse, &of);
      DevOther o{};
      o.op_begin = int(pc->ops.size());
      for (const auto& f : of) {
        pc->ops.push_back(MakeDevOp(f, &pc->mats));
        cost += amps * (f.diag ? 1 : (1 << f.k)) + 2048.0;
      }
      InsertRemaps(&pc->ops, size_t(o.op_begin), dc.n);
      o.op_end = int(pc->ops.size());
      o.col = int(k);
      o.weight = opt.mode == kModeInnerProductGrad && job.downstream ? job.downstream[k] : 0.0f;
      pc->others.push_back(o);
    }
  }
  dc.other_begin = 0;
  dc.other_end = int(pc->others.size());
  pc->cost = cost;
}

// Gates of `tmpl` with one row's symbol values (ParseArg resolves a symbolic
// argument to the row's value; the unitary follows from the arguments).
void RowGates(const Circuit& tmpl, const float* values, std::vector<Gate>* gates) {
  *gates = tmpl.gates;
  for (auto& g : *gates) {
    if (g.nsym == 0) continue;
    for (int j = 0; j < g.nsym; ++j) g.params[g.sym_pidx[j]] = values[g.sym_col[j]];
    GateUnitary(g.kind, g.params, g.U);
  }
}

// Rewrite the matrices of symbolic ops for one row (others keep `base`'s).
void RebindRow(const std::vector<Gate>& gs, const SharedRecipe& rec, bool tfq_fd, cfloat* mats) {
  cdouble M[256];
  for (const auto& r : rec.ops) {
    if (!r.symbolic) continue;
    const int dim = 1 << r.k;
    const int* gl = rec.gates.data() + r.gate_begin;
    const int ng = r.gate_end - r.gate_begin;
    if (r.kind == 0) {
      BlockMatrix(gs, gl, ng, r.bits, r.k, M);
      WriteMatrix(mats + r.mat, M, r.diag, dim);
      continue;
    }
    if (r.kind == 3) {
      int dgate[kMaxDeriv];
      cdouble dU[kMaxDeriv][16], D[kMaxDeriv][16];
      cdouble* dM[kMaxDeriv] = {D[0], D[1]};
      for (int d = 0; d < r.nderiv; ++d) {
        const Gate& g = gs[size_t(gl[r.dgate[d]])];
        dgate[d] = r.dgate[d];
        GateDerivative(g.kind, g.params, g.sym_pidx[r.dsym[d]], tfq_fd, dU[d]);
      }
      ParamBlockMatrices(gs, gl, ng, r.bits, r.k, r.nderiv, dgate, dU, M, dM);
      WriteMatrix(mats + r.mat, M, false, dim);
      for (int d = 0; d < r.nderiv; ++d) WriteMatrix(mats + r.dmat[d], D[d], false, dim);
      continue;
    }
    const Gate& g = gs[gl[0]];
    Embed(g.U, g.bits, g.nt, 0, 0, r.bits, r.k, M);
    WriteMatrix(mats + r.mat, M, r.diag, dim);
    for (int d = 0; d < r.nderiv; ++d) {
      cdouble dU[16];
      GateDerivative(g.kind, g.params, g.sym_pidx[d], tfq_fd, dU);
      Embed(dU, g.bits, g.nt, 0, 0, r.bits, r.k, M);
      WriteMatrix(mats + r.dmat[d], M, false, dim);
    }
  }
}

template <typename F>
void RunParallel(int n, int threads, F f) {
  threads = std::max(1, std::min(threads, n));
  std::atomic<int> next(0);
  auto work = [&]() {
    while (true) {
      int i = next.fetch_add(1);
      if (i >= n) break;
      f(i);
    }
  };
  if (threads == 1) {
    work();
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) pool.emplace_back(work);
  for (auto& th : pool) th.join();
}

// Concatenate pieces in longest-first order (LPT) so the persistent kernel's
// tail is short.
void Concatenate(std::vector<Piece>& pieces, int num_slots, int num_symbols, const PlanOptions& opt,
                 BatchProgram* out) {
  const int B = int(pieces.size());
  std::vector<int> order(B);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return pieces[a].cost > pieces[b].cost; });
  // Wave-mode circuits run after every workgroup-sized one.
  const auto wave_begin = std::stable_partition(order.begin(), order.end(), [&](int i) {
    return (pieces[i].dc.flags & kFlagWave) == 0;
  });
  out->num_block = int(wave_begin - order.begin());
  out->circuits.clear();
  out->others.clear();
  out->passes.clear();
  out->any_tiled = false;
  out->ops.clear();
  out->mats.clear();
  out->terms.clear();
  out->row_mat_base.assign(size_t(B), 0);
  out->ident.assign(size_t(B) * num_slots, 0.0f);
  out->num_rows = B;
  out->num_slots = num_slots;
  out->num_symbols = num_symbols;
  out->adjoint = opt.adjoint;
  out->max_n = 0;
  size_t nops = 0, nmats = 0, nterms = 0;
  for (const auto& p : pieces) {
    nops += p.ops.size();
    nmats += p.mats.size();
    nterms += p.terms.size();
  }
  out->ops.reserve(nops);
  out->mats.reserve(nmats);
  out->terms.reserve(nterms);
  for (int oi : order) {
    Piece& p = pieces[oi];
    const int op_base = int(out->ops.size());
    const uint32_t mat_base = uint32_t(out->mats.size());
    const int term_base = int(out->terms.size());
    DevCircuit dc = p.dc;
    out->row_mat_base[size_t(dc.row)] = mat_base;
    const int pass_base = int(out->passes.size());
    for (DevPass ps : p.passes) {
      ps.op_begin += op_base;
      ps.op_end += op_base;
      out->passes.push_back(ps);
    }
    dc.fwd_pass_begin += pass_base; dc.fwd_pass_end += pass_base;
    dc.rev_pass_begin += pass_base; dc.rev_pass_end += pass_base;
    if (dc.flags & kFlagTiled) out->any_tiled = true;
    const int other_base = int(out->others.size());
    dc.other_begin += other_base; dc.other_end += other_base;
    for (DevOther o : p.others) {
      o.op_begin += op_base;
      o.op_end += op_base;
      out->others.push_back(o);
    }
    dc.fwd_begin += op_base; dc.fwd_end += op_base;
    dc.rev_begin += op_base; dc.rev_end += op_base;
    dc.term_begin += term_base; dc.term_end += term_base;
    dc.lam_begin += term_base; dc.lam_end += term_base;
    for (DevOp o : p.ops) {
      o.mat += mat_base;
      for (int d = 0; d < o.nderiv; ++d) o.dmat[d] += mat_base;
      out->ops.push_back(o);
    }
    out->mats.insert(out->mats.end(), p.mats.begin(), p.mats.end());
    out->terms.insert(out->terms.end(), p.terms.begin(), p.terms.end());
    for (int j = 0; j < num_slots; ++j) out->ident[size_t(dc.row) * num_slots + j] = p.ident[j];
    out->circuits.push_back(dc);
    out->max_n = std::max(out->max_n, dc.n);
  }
}

}  // namespace

DevOp FusedToDevOp(const FusedOp& f, std::vector<cfloat>* mats) { return MakeDevOp(f, mats); }

DevTerm PauliToTerm(const Pauli& p, float w, int slot) { return MakeTerm(p, w, slot, true, 0); }

double CircuitCost(const DevCircuit& c, const BatchProgram& p) {
  (void)p;
  return std::ldexp(1.0, c.n) * double((c.fwd_end - c.fwd_begin) + (c.rev_end - c.rev_begin));
}

void BuildBatchProgram(const std::vector<CircuitJob>& jobs, int num_slots, int num_symbols,
                       const PlanOptions& opt, int threads, BatchProgram* out,
                       std::vector<SharedRecipe>* recipes) {
  const int B = int(jobs.size());
  std::vector<Piece> pieces(B);
  if (recipes) recipes->assign(size_t(B), SharedRecipe());
  RunParallel(B, threads, [&](int i) {
    BuildPiece(jobs[i], num_slots, opt, &pieces[i],
               recipes ? &(*recipes)[size_t(jobs[i].row)] : nullptr);
  });
  Concatenate(pieces, num_slots, num_symbols, opt, out);
}

RowRebind CompactRecipe(const Circuit& c, const SharedRecipe& rec) {
  RowRebind out;
  std::unordered_map<int, int> where;
  for (const OpRecipe& r : rec.ops) {
    if (!r.symbolic) continue;
    OpRecipe q = r;
    q.gate_begin = int(out.recipe.gates.size());
    for (int i = r.gate_begin; i < r.gate_end; ++i) {
      const int gi = rec.gates[size_t(i)];
      auto it = where.find(gi);
      if (it == where.end()) {
        const Gate& g = c.gates[size_t(gi)];
        RebindGate rg{};
        rg.kind = g.kind;
        rg.nt = int8_t(g.nt);
        rg.nsym = int8_t(g.nsym);
        for (int j = 0; j < 2; ++j) {
          rg.bits[j] = g.bits[j];
          rg.sym_pidx[j] = int8_t(j < g.nsym ? g.sym_pidx[j] : 0);
          rg.sym_col[j] = j < g.nsym ? g.sym_col[j] : 0;
        }
        rg.cmask = g.cmask;
        rg.cval = g.cval;
        for (int j = 0; j < 5; ++j) rg.params[j] = g.params[j];
        it = where.emplace(gi, int(out.gates.size())).first;
        out.gates.push_back(rg);
      }
      out.recipe.gates.push_back(it->second);
    }
    q.gate_end = int(out.recipe.gates.size());
    out.recipe.ops.push_back(q);
  }
  return out;
}

void RebindBatchProgram(const std::vector<RowRebind>& rows, const float* values, int num_symbols,
                        bool tfq_fd, int threads, const BatchProgram& prog, cfloat* mats) {
  const int B = prog.num_rows;
  RunParallel(B, threads, [&](int row) {
    const RowRebind& rr = rows[size_t(row)];
    if (rr.recipe.ops.empty()) return;
    const float* v = values + size_t(row) * num_symbols;
    std::vector<Gate> gs(rr.gates.size());
    for (size_t i = 0; i < gs.size(); ++i) {
      const RebindGate& r = rr.gates[i];
      Gate& g = gs[i];
      g.kind = r.kind;
      g.nt = r.nt;
      g.nsym = r.nsym;
      g.bits[0] = r.bits[0];
      g.bits[1] = r.bits[1];
      g.cmask = r.cmask;
      g.cval = r.cval;
      for (int j = 0; j < 5; ++j) g.params[j] = r.params[j];
      for (int j = 0; j < 2; ++j) {
        g.sym_pidx[j] = r.sym_pidx[j];
        g.sym_col[j] = r.sym_col[j];
      }
      for (int j = 0; j < g.nsym; ++j) g.params[g.sym_pidx[j]] = v[g.sym_col[j]];
      GateUnitary(g.kind, g.params, g.U);
    }
    RebindRow(gs, rr.recipe, tfq_fd, mats + prog.row_mat_base[size_t(row)]);
  });
}

namespace {

RbGate ToRbGate(GateKind kind, int nt, int nsym, const int* sym_pidx, const int* sym_col,
                const int* bits, uint64_t cmask, uint64_t cval, const float* params) {
  RbGate g{};
  g.kind = int32_t(kind);
  g.nt = nt;
  g.nsym = nsym;
  for (int j = 0; j < 2; ++j) {
    g.sym_pidx[j] = j < nsym ? sym_pidx[j] : 0;
    g.sym_col[j] = j < nsym ? sym_col[j] : 0;
    g.bits[j] = j < nt ? bits[j] : 0;
  }
  g.cmask = cmask;
  g.cval = cval;
  for (int j = 0; j < 5; ++j) g.params[j] = params[j];
  return g;
}

bool ToRbOp(const OpRecipe& r, RbOp* o) {
  if (r.kind == 3 && r.k > 2) return false;
  *o = RbOp{};
  o->kind = r.kind;
  o->k = r.k;
  o->diag = r.diag;
  o->nderiv = r.nderiv;
  for (int j = 0; j < r.k; ++j) o->bits[j] = r.bits[j];
  o->gate_begin = r.gate_begin;
  o->gate_end = r.gate_end;
  o->mat = r.mat;
  for (int d = 0; d < r.nderiv; ++d) {
    o->dmat[d] = r.dmat[d];
    o->dgate[d] = r.dgate[d];
    o->dsym[d] = r.dsym[d];
  }
  return true;
}

}  // namespace

RebindTableSizes SizeRebindTables(const std::vector<RowRebind>& rows) {
  RebindTableSizes z;
  for (const RowRebind& rr : rows) {
    z.gates += rr.gates.size();
    z.gidx += rr.recipe.gates.size();
    z.ops += rr.recipe.ops.size();
  }
  return z;
}

RebindTableSizes SizeSharedRebindTables(const Circuit& tmpl, const SharedRecipe& rec) {
  RebindTableSizes z;
  z.gates = tmpl.gates.size();
  z.gidx = rec.gates.size();
  for (const OpRecipe& r : rec.ops) z.ops += r.symbolic;
  return z;
}

bool BuildRebindTables(const std::vector<RowRebind>& rows, const BatchProgram& prog,
                       RebindTables* t) {
  *t = RebindTables();
  const RebindTableSizes z = SizeRebindTables(rows);
  t->gates.reserve(z.gates);
  t->gidx.reserve(z.gidx);
  t->ops.reserve(z.ops);
  t->rows.resize(rows.size());
  for (size_t i = 0; i < rows.size(); ++i) {
    const RowRebind& rr = rows[i];
    RbRow& row = t->rows[i];
    row = RbRow{};
    row.op_begin = int32_t(t->ops.size());
    row.gidx_base = int32_t(t->gidx.size());
    row.gate_base = int32_t(t->gates.size());
    row.mat_base = uint32_t(prog.row_mat_base[i]);
    row.vrow = int32_t(i);
    for (const RebindGate& g : rr.gates) {
      const int pidx[2] = {g.sym_pidx[0], g.sym_pidx[1]};
      t->gates.push_back(ToRbGate(g.kind, g.nt, g.nsym, pidx, g.sym_col, g.bits, g.cmask,
                                  g.cval, g.params));
    }
    t->gidx.insert(t->gidx.end(), rr.recipe.gates.begin(), rr.recipe.gates.end());
    for (const OpRecipe& r : rr.recipe.ops) {
      if (!r.symbolic) continue;
      RbOp o;
      if (!ToRbOp(r, &o)) return false;
      o.row = int32_t(i);
      t->kmax = std::max(t->kmax, int(r.k));
      t->ops.push_back(o);
    }
    row.op_end = int32_t(t->ops.size());
  }
  // Neighbouring threads of rebind_kernel take neighbouring ops: group ops of
  // one shape (width, kind, derivative passes, gate count) so a wavefront's
  // threads follow the same path.  Each op names its row, so order is free.
  std::stable_sort(t->ops.begin(), t->ops.end(), [](const RbOp& x, const RbOp& y) {
    const auto key = [](const RbOp& o) {
      return std::make_tuple(o.k, o.kind, o.nderiv, o.gate_end - o.gate_begin);
    };
    return key(x) < key(y);
  });
  return true;
}

bool BuildSharedRebindTables(const Circuit& tmpl, const SharedRecipe& rec,
                             const BatchProgram& prog, RebindTables* t) {
  *t = RebindTables();
  for (const Gate& g : tmpl.gates)
    t->gates.push_back(ToRbGate(g.kind, g.nt, g.nsym, g.sym_pidx, g.sym_col, g.bits, g.cmask,
                                g.cval, g.params));
  t->gidx.assign(rec.gates.begin(), rec.gates.end());
  for (const OpRecipe& r : rec.ops) {
    if (!r.symbolic) continue;
    RbOp o;
    if (!ToRbOp(r, &o)) return false;
    t->kmax = std::max(t->kmax, int(r.k));
    t->ops.push_back(o);
  }
  t->rows.resize(size_t(prog.num_rows));
  for (int i = 0; i < prog.num_rows; ++i) {
    RbRow& row = t->rows[size_t(i)];
    row = RbRow{};
    row.op_begin = 0;
    row.op_end = int32_t(t->ops.size());
    row.mat_base = uint32_t(prog.row_mat_base[size_t(i)]);
    row.vrow = i;
  }
  return true;
}

void BuildSharedBatchProgram(const Circuit& tmpl, const std::vector<CircuitJob>& jobs,
                             const float* values, int num_symbols, int num_slots,
                             const PlanOptions& opt_in, int threads, BatchProgram* out,
                             SharedRecipe* recipe) {
  PlanOptions opt = opt_in;
  opt.fuse.structural_diag = true;
  const int B = int(jobs.size());
  *recipe = SharedRecipe();
  Piece base;
  CircuitJob j0 = jobs[0];
  j0.circuit = &tmpl;
  BuildPiece(j0, num_slots, opt, &base, recipe);
  recipe->mats_per_row = base.mats.size();
  std::vector<Piece> pieces(B);
  RunParallel(B, threads, [&](int i) {
    Piece& pc = pieces[i];
    pc.dc = base.dc;
    pc.dc.row = jobs[i].row;
    pc.ops = base.ops;
    pc.mats = base.mats;
    std::vector<Gate> gs;
    RowGates(tmpl, values + size_t(jobs[i].row) * num_symbols, &gs);
    RebindRow(gs, *recipe, opt.tfq_fd, pc.mats.data());
    const double term_cost = AddTerms(jobs[i], num_slots, opt, &pc);
    pc.cost = base.cost - std::ldexp(1.0, std::max(base.dc.n, 0)) * double(base.terms.size()) +
              term_cost;
  });
  Concatenate(pieces, num_slots, num_symbols, opt, out);
}

void RebindSharedBatchProgram(const Circuit& tmpl, const SharedRecipe& recipe, const float* values,
                              int num_symbols, bool tfq_fd, int threads, const BatchProgram& prog,
                              cfloat* mats) {
  const int B = prog.num_rows;
  RunParallel(B, threads, [&](int row) {
    std::vector<Gate> gs;
    RowGates(tmpl, values + size_t(row) * num_symbols, &gs);
    RebindRow(gs, recipe, tfq_fd, mats + prog.row_mat_base[size_t(row)]);
  });
}

}  // namespace tfqx
