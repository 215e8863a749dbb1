// Host planning of the HBM-streaming path (see large.h).
#include "large.h"

#include <algorithm>
#include <cmath>

namespace tfqx {
namespace {

constexpr int kMinLowRun = 6;     // >= 64 contiguous amplitudes (512 B) per tile row
constexpr int kMaxPassDerivs = 64;

inline int Popc(uint64_t x) { return __builtin_popcountll(x); }

uint64_t OpMask(const FusedOp& f) {
  uint64_t m = f.cmask;
  for (int j = 0; j < f.k; ++j) m |= uint64_t(1) << f.bits[j];
  return m;
}

uint32_t PushM(std::vector<cfloat>* mats, const std::vector<cdouble>& M, bool diag, int dim) {
  uint32_t off = uint32_t(mats->size());
  if (diag) {
    for (int i = 0; i < dim; ++i)
      mats->push_back(cfloat(float(M[i * dim + i].real()), float(M[i * dim + i].imag())));
  } else {
    for (const auto& x : M) mats->push_back(cfloat(float(x.real()), float(x.imag())));
  }
  return off;
}

DevTermWide WideTerm(const Pauli& p, float w, int slot, bool phase_y, int index) {
  DevTermWide t{};
  t.fmask = p.xmask;
  t.zmask = p.zmask;
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

}  // namespace

void SchedulePasses(const std::vector<FusedOp>& ops, int n, int L, std::vector<DevPass>* passes,
                    std::vector<DevOp>* dev_ops, std::vector<cfloat>* mats,
                    int max_r, const Circuit* c, SharedRecipe* rec) {
  L = std::min(L, n);
  const int low_run = std::min(kMinLowRun, L);
  const uint64_t lowmask = (uint64_t(1) << low_run) - 1;
  std::vector<int> remaining(ops.size());
  for (size_t i = 0; i < ops.size(); ++i) remaining[i] = int(i);
  // Ops a tile of qubits T can run now, in order: an op runs when its qubits
  // lie in T and it commutes with every op skipped before it (disjoint
  // qubits, or both diagonal), within the pass's derivative budget.
  auto take_ops = [&](uint64_t T, std::vector<int>* take, uint64_t* cand, int ncand) -> int {
    uint64_t blocked = 0, blocked_diag = 0;
    int derivs = 0, count = 0, nc = 0;
    for (int i : remaining) {
      const FusedOp& f = ops[i];
      const uint64_t Q = OpMask(f);
      const int nd = f.param ? f.nderiv : 0;
      const uint64_t hit = f.diag && !f.param ? blocked : (blocked | blocked_diag);
      bool ok = (Q & hit) == 0;
      if (ok && (Q & ~T)) {
        // Runnable but for missing qubits: a candidate extension of T.
        if (cand && nc < ncand && Popc(T | Q) <= L) {
          const uint64_t M = Q & ~T;
          bool dup = false;
          for (int c = 0; c < nc; ++c) dup = dup || cand[c] == M;
          if (!dup) cand[nc++] = M;
        }
        ok = false;
      }
      if (ok && derivs + nd > kMaxPassDerivs) ok = false;
      if (!ok) {
        if (f.diag && !f.param) blocked_diag |= Q;
        else blocked |= Q;
        continue;
      }
      derivs += nd;
      ++count;
      if (take) take->push_back(i);
    }
    if (cand) {
      for (int c = nc; c < ncand; ++c) cand[c] = 0;
    }
    return count;
  };
  while (!remaining.empty()) {
    // Grow the tile from the low run: each step adds the missing qubits of one
    // of the first few runnable-but-missing ops, choosing the extension that
    // lets the most ops run (ties: fewer qubits, then earlier op).
    constexpr int kCand = 8;
    uint64_t needed = Popc(lowmask) <= L ? lowmask : 0;
    while (Popc(needed) < L) {
      uint64_t cand[kCand];
      take_ops(needed, nullptr, cand, kCand);
      int best = -1;
      uint64_t best_T = 0;
      for (int c = 0; c < kCand && cand[c]; ++c) {
        const uint64_t T2 = needed | cand[c];
        const int v = take_ops(T2, nullptr, nullptr, 0);
        if (v > best || (v == best && Popc(T2) < Popc(best_T))) {
          best = v;
          best_T = T2;
        }
      }
      if (best < 0) break;
      needed = best_T;
    }
    std::vector<int> chosen, deferred;
    take_ops(needed, &chosen, nullptr, 0);
    {
      std::vector<char> in(ops.size(), 0);
      for (int i : chosen) in[size_t(i)] = 1;
      for (int i : remaining)
        if (!in[size_t(i)]) deferred.push_back(i);
    }
    if (chosen.empty()) {
      // An op wider than the low run allows: give it a tile of its own qubits
      // (it always fits: fused ops span at most kMaxFuse targets plus controls,
      // and the planner rejects ops wider than a tile).
      const int i = deferred.front();
      chosen.push_back(i);
      deferred.erase(deferred.begin());
      needed = OpMask(ops[i]);
    }
    uint64_t T = (Popc(needed | lowmask) <= L) ? (needed | lowmask) : needed;
    for (int b = 0; b < n && Popc(T) < L; ++b) T |= uint64_t(1) << b;
    DevPass p{};
    p.L = L;
    p.tilemask = T;
    int j = 0;
    int local[64];
    for (int b = 0; b < 64; ++b) {
      local[b] = -1;
      if (T >> b & 1) {
        p.tbit[j] = uint8_t(b);
        local[b] = j++;
      }
    }
    p.nlow = 0;
    while (p.nlow < L && p.tbit[p.nlow] == p.nlow) p.nlow++;
    p.op_begin = int(dev_ops->size());
    for (int i : chosen) {
      const FusedOp& f = ops[i];
      DevOp d{};
      d.k = uint8_t(f.k);
      for (int q = 0; q < f.k; ++q) d.q[q] = uint8_t(local[f.bits[q]]);
      for (int b = 0; b < 64; ++b) {
        if (f.cmask >> b & 1) {
          d.cmask |= 1u << local[b];
          if (f.cval >> b & 1) d.cval |= 1u << local[b];
        }
      }
      const int dim = 1 << f.k;
      if (f.param) {
        d.type = kOpParam;
        d.mat = PushM(mats, f.M, false, dim);
        d.nderiv = uint8_t(f.nderiv);
        for (int k = 0; k < f.nderiv; ++k) {
          d.dmat[k] = PushM(mats, f.dM[k], false, dim);
          d.col[k] = f.col[k];  // partial slot (remapped by the caller)
        }
      } else {
        d.type = f.diag ? kOpDiag : kOpUnitary;
        d.mat = PushM(mats, f.M, f.diag, dim);
      }
      if (rec && c) AppendRecipe(f, *c, d, rec);
      dev_ops->push_back(d);
    }
    if (max_r) InsertRemaps(dev_ops, size_t(p.op_begin), L, max_r);
    p.op_end = int(dev_ops->size());
    passes->push_back(p);
    remaining.swap(deferred);
  }
}

namespace {

int TileQubits(int mode, bool two_buffers) {
  (void)mode;
  const Tuning t = GetTuning();
  const int want = two_buffers ? t.tile_two : t.tile_single;
  if (want > 0) return std::min(want, two_buffers ? 13 : kMaxTile);
  // Measured on the 30-qubit grid circuit: 13-qubit tiles need 20 passes
  // instead of 28 at 12 qubits and finish the circuit 27% sooner.  Two-state
  // (adjoint) passes also take 13 qubits: the register engine needs tiles of
  // 12-13 qubits, and on the 26-qubit, 416-symbol adjoint 13-qubit tiles take
  // 48.7 ms per parameter set against 61.3 ms at 12 and 140 ms at 11 (LDS
  // path, one barrier-synchronised reduction per derivative per tile).
  return 13;
}

// A dense pass for one fused block of 5-6 qubits: the tile holds its targets
// and the lowest other qubits (kDenseTile in all, so rows stay contiguous).
void AddDensePass(LargePlan* lp, const FusedOp& f, int n, PassPlan* pp) {
  const int L = std::min(n, kDenseTile);
  uint64_t T = 0;
  for (int j = 0; j < f.k; ++j) T |= uint64_t(1) << f.bits[j];
  for (int b = 0; b < n && Popc(T) < L; ++b) T |= uint64_t(1) << b;
  DevPass p{};
  p.L = L;
  p.tilemask = T;
  int local[64], j = 0;
  for (int b = 0; b < 64; ++b) {
    local[b] = -1;
    if (T >> b & 1) {
      p.tbit[j] = uint8_t(b);
      local[b] = j++;
    }
  }
  p.nlow = 0;
  while (p.nlow < L && p.tbit[p.nlow] == p.nlow) p.nlow++;
  p.op_begin = p.op_end = int(lp->ops.size());
  DenseOp d;
  d.k = f.k;
  for (int i = 0; i < f.k; ++i) {
    d.bits[i] = f.bits[i];
    d.pos[i] = local[f.bits[i]];
  }
  d.mat = PushM(&lp->mats, f.M, false, 1 << f.k);
  lp->dense.push_back(d);
  p.dense = int(lp->dense.size());
  pp->passes.push_back(p);
}

void AddPasses(LargePlan* lp, const std::vector<FusedOp>& ops, int n, int L, PassPlan* pp) {
  for (const auto& f : ops)
    if (Popc(OpMask(f)) > std::min(L, n)) lp->too_wide = true;
  if (lp->too_wide) return;
  std::vector<DevPass> passes;
  // The streaming passes keep the LDS path unless tuning level 2 asks for the
  // register engine there too; one-state 14-qubit tiles get five register
  // qubits (pass_kernel's wide instantiation).
  SchedulePasses(ops, n, L, &passes, &lp->ops, &lp->mats, GetTuning().reg_engine >= 2 ? 5 : 0);
  pp->passes.insert(pp->passes.end(), passes.begin(), passes.end());
}

// Forward passes with wide fusion (`width` 5 or 6): blocks of 5-6 qubits
// become dense MFMA passes; the gates of the smaller blocks between them are
// fused again at width 2 and run as ordinary tile passes.  Blocks keep the
// fuser's order, which is a valid application order.
void AddPassesWide(LargePlan* lp, const Circuit& c, int width, int n, int L, PassPlan* pp) {
  FuseOptions fw;
  fw.max_fuse = width;
  std::vector<FusedOp> blocks;
  FuseGates(c, 0, int(c.gates.size()), fw, &blocks);
  FuseOptions f2;
  f2.max_fuse = 2;
  Circuit run;
  run.num_qubits = c.num_qubits;
  auto flush = [&]() {
    if (run.gates.empty()) return;
    std::vector<FusedOp> small;
    FuseGates(run, 0, int(run.gates.size()), f2, &small);
    AddPasses(lp, small, n, L, pp);
    run.gates.clear();
  };
  for (const FusedOp& b : blocks) {
    const bool dense = b.k >= 5 && b.cmask == 0 && n >= kDenseTile && !b.diag;
    if (!dense) {
      for (int gi : b.gates) run.gates.push_back(c.gates[gi]);
      continue;
    }
    flush();
    AddDensePass(lp, b, n, pp);
  }
  flush();
}

}  // namespace

Status BuildLargePlan(int mode, const std::vector<Circuit>& circuits,
                      const std::vector<std::vector<std::vector<Pauli>>>& sums,
                      const float* downstream, const std::vector<std::vector<Circuit>>& others,
                      int64_t B, int64_t S, int64_t O, int64_t K, bool tfq_fd,
                      bool with_expectation, std::unique_ptr<LargePlan>* out,
                      int wide_fusion) {
  auto lp = std::make_unique<LargePlan>();
  lp->mode = mode;
  lp->with_expectation = with_expectation;
  lp->B = B;
  lp->S = S;
  lp->O = O;
  lp->K = K;
  FuseOptions fo;
  fo.max_fuse = 2;
  for (const auto& c : circuits) lp->max_nq = std::max(lp->max_nq, c.num_qubits);
  lp->circuits.resize(size_t(B));
  for (int64_t i = 0; i < B; ++i) {
    const Circuit& c = circuits[i];
    LargeCircuit& lc = lp->circuits[i];
    lc.n = c.num_qubits;
    lc.row = int(i);
    lc.empty = c.gates.empty();
    lc.ident.assign(size_t(mode == kModeInnerProduct ? K : O), 0.0f);
    if (lc.empty) continue;
    const bool grad = mode == kModeAdjoint || mode == kModeInnerProductGrad;
    lc.L = std::min(lc.n, TileQubits(mode, grad));
    if (wide_fusion >= 5) {
      AddPassesWide(lp.get(), c, std::min(wide_fusion, kMaxFuseWide), lc.n,
                    std::min(lc.n, TileQubits(mode, false)), &lc.fwd);
    } else {
      std::vector<FusedOp> fused;
      FuseGates(c, 0, int(c.gates.size()), fo, &fused);
      AddPasses(lp.get(), fused, lc.n, std::min(lc.n, TileQubits(mode, false)), &lc.fwd);
    }
    // Terms.
    lc.term_begin = int(lp->terms.size());
    const bool sampled = mode == kModeSampledExpectation;
    if (mode == kModeExpectation || sampled || mode == kModeAdjoint) {
      for (int64_t j = 0; j < O; ++j) {
        int idx = 0;
        for (const auto& p : sums[i][j]) {
          if (p.ops.empty()) {
            lc.ident[j] += p.coef;
          } else if (mode != kModeAdjoint || with_expectation) {
            lp->terms.push_back(WideTerm(p, p.coef, int(j), !sampled, idx));
          }
          ++idx;
        }
      }
    }
    if (sampled) {
      // Group terms of each slot by rotation pattern (stable), so terms that
      // measure in the same basis share one CDF.
      std::stable_sort(lp->terms.begin() + lc.term_begin, lp->terms.end(),
                       [](const DevTermWide& a, const DevTermWide& b) {
                         if (a.slot != b.slot) return a.slot < b.slot;
                         if (a.fmask != b.fmask) return a.fmask < b.fmask;
                         return (a.fmask & a.zmask) < (b.fmask & b.zmask);
                       });
    }
    lc.term_end = int(lp->terms.size());
    if (mode == kModeAdjoint) {
      lc.lam_begin = int(lp->terms.size());
      for (int64_t j = 0; j < O; ++j) {
        const float gj = downstream[i * O + j];
        for (const auto& p : sums[i][j]) {
          const float w = gj * p.coef;
          if (std::fabs(w) < 1e-5f) continue;
          lp->terms.push_back(WideTerm(p, w, 0, true, 0));
        }
      }
      lc.lam_end = int(lp->terms.size());
    }
    if (grad) {
      // Reverse program (as the batch planner), param ops numbered by partial slot.
      std::vector<FusedOp> rev;
      BuildReverseOps(c, fo, tfq_fd, &rev);
      const size_t op0 = lp->ops.size();
      AddPasses(lp.get(), rev, lc.n, lc.L, &lc.rev);
      // Number gradient partial slots in pass order so each pass owns a
      // contiguous slot range; param_cols maps slot -> accumulator index.
      // Inner-product gradients take two slots per derivative (real part,
      // then imaginary part) accumulating into gacc[2 col] and gacc[2 col + 1].
      const bool cplx = mode == kModeInnerProductGrad;
      for (size_t i = op0; i < lp->ops.size(); ++i) {
        DevOp& d = lp->ops[i];
        if (d.type != kOpParam) continue;
        for (int k = 0; k < d.nderiv; ++k) {
          const int col = d.col[k];
          d.col[k] = int(lp->param_cols.size());
          if (cplx) {
            lp->param_cols.push_back(2 * col);
            lp->param_cols.push_back(2 * col + 1);
          } else {
            lp->param_cols.push_back(col);
          }
        }
      }
    }
    if (mode == kModeInnerProduct || mode == kModeInnerProductGrad) {
      for (int64_t k = 0; k < K; ++k) {
        const Circuit& oc = others[i][k];
        std::vector<FusedOp> of;
        FuseGates(oc, 0, int(oc.gates.size()), fo, &of);
        PassPlan pp;
        AddPasses(lp.get(), of, lc.n, std::min(lc.n, TileQubits(mode, false)), &pp);
        lc.others.push_back(pp);
        lc.other_weights.push_back(mode == kModeInnerProductGrad ? downstream[i * K + k] : 0.0f);
      }
    }
  }
  if (lp->too_wide)
    return Status::Invalid("A gate with its controls acts on more qubits than a state tile holds.");
  lp->workspace_bytes = LargeWorkspaceBytes(lp.get());
  *out = std::move(lp);
  return Status();
}

}  // namespace tfqx
