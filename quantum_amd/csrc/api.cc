// C-ABI of the op family (include/tfqx.h).
#include "../../include/tfqx.h"

#include <string.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include <hip/hip_runtime.h>

#include "batch_kernels.h"
#include "circuit.h"
#include "large.h"
#include "noisy.h"
#include "plan.h"
#include "wire.h"

using namespace tfqx;

namespace {

std::atomic<int> g_fusion_width(0);

void SetErr(char* err, int64_t len, const std::string& msg) {
  if (!err || len <= 0) return;
  const size_t n = std::min<size_t>(size_t(len - 1), msg.size());
  memcpy(err, msg.data(), n);
  err[n] = '\0';
}

int Fail(char* err, int64_t len, int code, const std::string& msg) {
  SetErr(err, len, msg);
  return code;
}

int HostThreads() {
  unsigned n = std::thread::hardware_concurrency();
  return int(std::max(1u, std::min(n, 16u)));  // one GPU's share of host cores
}

template <typename F>
void ParallelFor(int n, F f) {
  const int T = std::min(HostThreads(), n);
  if (T <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int> next(0);
  std::vector<std::thread> pool;
  for (int t = 0; t < T; ++t)
    pool.emplace_back([&]() {
      while (true) {
        int i = next.fetch_add(1);
        if (i >= n) break;
        f(i);
      }
    });
  for (auto& th : pool) th.join();
}

struct FirstError {
  std::atomic<bool> set{false};
  Status st;
  void record(const Status& s) {
    if (s.ok()) return;
    bool expected = false;
    if (set.compare_exchange_strong(expected, true)) st = s;
  }
};

std::string Bytes(const uint8_t* b, const int64_t* off, int64_t i) {
  return std::string(reinterpret_cast<const char*>(b) + off[i], size_t(off[i + 1] - off[i]));
}

size_t Align(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

struct tfqx_plan {
  int mode = kModeExpectation;
  int64_t B = 0, S = 0, O = 0, K = 0;
  int max_nq = 0;
  int nshots = 0;
  uint64_t seed = 0;
  BatchProgram prog;
  std::vector<int32_t> num_samples;
  // Workspace layout (bytes).
  size_t off_counters = 0, off_circuits = 0, off_others = 0, off_ops = 0, off_mats = 0,
         off_terms = 0, off_ident = 0, off_nsamp = 0, off_passes = 0, off_gacc = 0,
         off_scratch = 0, total = 0;
  std::vector<uint8_t> staging;
  // Device re-evaluation of the symbolic matrices (tfqx_plan_stage_values):
  // a region after the scratch holding the staged symbol values and the
  // rebind tables, whose host image rb_image is built on first use.
  RebindTableSizes rb_sizes;
  size_t off_rb = 0, rb_off_eigen = 0, rb_off_rows = 0, rb_off_ops = 0, rb_off_gidx = 0,
         rb_off_gates = 0, rb_bytes = 0;
  std::vector<uint8_t> rb_image;
  int rb_kmax = 1;
  bool rb_built = false;
  // Upload bookkeeping: the workspace of the last upload and what the host
  // changed since (tfqx_plan_upload_changed copies only that).
  void* last_ws = nullptr;
  bool mats_dirty = false, terms_dirty = false;
  bool rb_tables_dirty = false, values_staged = false;
  bool device_mats = false;  // the device holds matrices newer than staging's
  int lds_log2 = 0;
  int kmax = 2;
  int grid = 0;
  bool with_expectation = false;  // adjoint plan also writes expectations
  bool tfq_fd = true;
  // All rows share one program: the template circuit and matrix recipes let
  // tfqx_plan_rebind re-evaluate the matrices for new symbol values.
  bool shared = false;
  Circuit tmpl;
  SharedRecipe recipe;
  // Any other batch-kernel plan: per-row recipes of the ops with symbolic
  // gates, so tfqx_plan_rebind re-evaluates them for new values.
  bool rebindable = false;
  std::vector<RowRebind> rows_rebind;
  // Adjoint plans: every row's lambda terms in emission order (slot, Pauli
  // coefficient, number of Y factors), for tfqx_plan_rebind_downstream.
  struct LamTerm {
    int slot;
    float coef;
    int ny;
  };
  std::vector<std::vector<LamTerm>> lam_meta;
  // States wider than one LDS tile run on the HBM-streaming path.
  bool large = false;
  std::unique_ptr<LargePlan> large_plan;
  // Noisy programs run on the trajectory kernel.
  bool noisy = false;
  std::unique_ptr<NoisyPlan> noisy_plan;
};

struct tfqx_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

struct Inputs {
  std::vector<Program> programs;
  std::unordered_map<std::string, int> columns;
  std::vector<Circuit> circuits;
  std::vector<std::vector<std::vector<Pauli>>> sums;   // [B][O]
  std::vector<std::vector<Circuit>> others;            // [B][K]
  const float* values = nullptr;                       // [B, S]
  // All programs are byte-identical and LDS-resident: circuits[0] is the
  // program resolved with row 0's values and stands for every row.
  bool shared = false;
};

int LdsLimit(int mode) {
  return BatchBuffers(mode) == 2 ? kLdsMaxQubitsTwoBuffers : kLdsMaxQubitsOneBuffer;
}

bool SameBytes(const uint8_t* b, const int64_t* off, int64_t i, int64_t j) {
  const int64_t n = off[i + 1] - off[i];
  return n == off[j + 1] - off[j] && memcmp(b + off[i], b + off[j], size_t(n)) == 0;
}

Status ParsePrograms(const uint8_t* b, const int64_t* off, int64_t n, std::vector<Program>* out) {
  out->assign(size_t(n), Program());
  FirstError fe;
  ParallelFor(int(n), [&](int i) {
    if (!DecodeProgram(b + off[i], size_t(off[i + 1] - off[i]), &(*out)[i]))
      fe.record(Status::Invalid("Unparseable proto: " + Bytes(b, off, i)));
  });
  return fe.set ? fe.st : Status();
}

Status Prepare(const uint8_t* programs, const int64_t* poff, int64_t B, const uint8_t* names,
               const int64_t* noff, int64_t S, const float* values, const uint8_t* psums,
               const int64_t* psoff, int64_t O, const uint8_t* other, const int64_t* ooff,
               int64_t K, Inputs* in, int shared_max_n = -1, bool noisy = false) {
  if (B < 0 || S < 0 || O < 0 || K < 0) return Status::Invalid("negative dimension");
  in->values = values;
  // Byte-identical programs (a circuit tiled over a batch of parameter sets)
  // are parsed and resolved once.
  bool same = B >= 1 && !other && !noisy;
  for (int64_t i = 1; same && i < B; ++i) same = SameBytes(programs, poff, i, 0);
  Status s = ParsePrograms(programs, poff, same ? 1 : B, &in->programs);
  if (!s.ok()) return s;
  std::vector<PauliSum> ps;
  if (psums) {
    ps.assign(size_t(B * O), PauliSum());
    FirstError fe;
    ParallelFor(int(B * O), [&](int i) {
      if (!DecodePauliSum(psums + psoff[i], size_t(psoff[i + 1] - psoff[i]), &ps[i]))
        fe.record(Status::Invalid("Unparseable proto: " + Bytes(psums, psoff, i)));
    });
    if (fe.set) return fe.st;
  }
  std::vector<Program> others;
  if (other) {
    s = ParsePrograms(other, ooff, B * K, &others);
    if (!s.ok()) return s;
  }
  for (int64_t j = 0; j < S; ++j) in->columns[Bytes(names, noff, j)] = int(j);
  Circuit tmpl;
  if (same) {
    SymbolMap sm;
    sm.columns = &in->columns;
    sm.values = values;
    s = BuildCircuit(in->programs[0], sm, &tmpl);
    if (!s.ok()) return s;
    in->shared = tmpl.num_qubits <= shared_max_n;
  }
  in->circuits.assign(in->shared ? 1 : size_t(B), Circuit());
  if (in->shared) in->circuits[0] = tmpl;
  in->sums.assign(size_t(B), {});
  in->others.assign(size_t(B), {});
  FirstError fe;
  ParallelFor(int(B), [&](int i) {
    Status st;
    if (!in->shared) {
      if (same) {
        RebindCircuit(tmpl, values + size_t(i) * S, &in->circuits[i]);
      } else {
        SymbolMap sm;
        sm.columns = &in->columns;
        sm.values = values + size_t(i) * S;
        st = noisy ? BuildNoisyCircuit(in->programs[i], sm, &in->circuits[i])
                   : BuildCircuit(in->programs[i], sm, &in->circuits[i]);
        if (!st.ok()) return fe.record(st);
      }
    }
    const Circuit& circ = in->circuits[in->shared ? 0 : i];
    if (psums) {
      in->sums[i].resize(size_t(O));
      for (int64_t j = 0; j < O; ++j) {
        st = ResolvePauliSum(ps[size_t(i * O + j)], circ, &in->sums[i][j]);
        if (!st.ok()) return fe.record(st);
      }
    }
    if (other) {
      in->others[i].resize(size_t(K));
      for (int64_t k = 0; k < K; ++k) {
        st = BuildPairedCircuit(others[size_t(i * K + k)], circ, &in->others[i][k]);
        if (!st.ok()) return fe.record(st);
      }
    }
  });
  return fe.set ? fe.st : Status();
}

int AutoFusionWidth(int max_n) {
  const int w = g_fusion_width.load();
  if (w > 0) return std::min(w, kMaxFuse);
  (void)max_n;
  return 2;
}

void Layout(tfqx_plan* p) {
  const BatchProgram& g = p->prog;
  size_t off = 0;
  p->off_counters = off; off = Align(off + 256);
  p->off_circuits = off; off = Align(off + g.circuits.size() * sizeof(DevCircuit));
  p->off_others = off;   off = Align(off + g.others.size() * sizeof(DevOther));
  p->off_ops = off;      off = Align(off + g.ops.size() * sizeof(DevOp));
  p->off_mats = off;     off = Align(off + (g.mats.size() + kMatCacheWide) * sizeof(cfloat));
  p->off_terms = off;    off = Align(off + g.terms.size() * sizeof(DevTerm));
  p->off_ident = off;    off = Align(off + g.ident.size() * sizeof(float));
  p->off_nsamp = off;    off = Align(off + p->num_samples.size() * sizeof(int32_t));
  p->off_passes = off;   off = Align(off + g.passes.size() * sizeof(DevPass));
  p->off_gacc = off;
  const size_t image = off;
  if (p->mode == kModeAdjoint) off = Align(off + size_t(p->B) * size_t(p->S) * sizeof(double));
  if (p->mode == kModeInnerProductGrad)
    off = Align(off + 2 * size_t(p->B) * size_t(p->S) * sizeof(double));
  p->off_scratch = off;
  // One psi + lambda scratch pair per launched workgroup; the grid never
  // exceeds the number of circuits (tfqx_plan_launch, tfqx_graph_capture).
  const size_t tiled_wgs = std::min<size_t>(kTiledGridCap, std::max<size_t>(1, g.circuits.size()));
  if (g.any_tiled) off = Align(off + tiled_wgs * 2 * (size_t(1) << p->max_nq) * sizeof(float2));
  if (p->shared || p->rebindable) {
    p->rb_sizes = p->shared ? SizeSharedRebindTables(p->tmpl, p->recipe)
                            : SizeRebindTables(p->rows_rebind);
    size_t r = 0;
    r = Align(r + size_t(p->B) * size_t(p->S) * sizeof(float));
    p->rb_off_eigen = r;  r = Align(r + kRbEigenKinds * sizeof(RbEigen));
    p->rb_off_rows = r;   r = Align(r + size_t(p->B) * sizeof(RbRow));
    p->rb_off_ops = r;    r = Align(r + p->rb_sizes.ops * sizeof(RbOp));
    p->rb_off_gidx = r;   r = Align(r + p->rb_sizes.gidx * sizeof(int32_t));
    p->rb_off_gates = r;  r = Align(r + p->rb_sizes.gates * sizeof(RbGate));
    p->rb_bytes = r;
    p->off_rb = off;
    off += r;
  }
  p->total = off;
  p->staging.assign(image, 0);
  auto put = [&](size_t o, const void* src, size_t n) {
    if (n) memcpy(p->staging.data() + o, src, n);
  };
  put(p->off_circuits, g.circuits.data(), g.circuits.size() * sizeof(DevCircuit));
  put(p->off_others, g.others.data(), g.others.size() * sizeof(DevOther));
  put(p->off_ops, g.ops.data(), g.ops.size() * sizeof(DevOp));
  put(p->off_mats, g.mats.data(), g.mats.size() * sizeof(cfloat));
  put(p->off_terms, g.terms.data(), g.terms.size() * sizeof(DevTerm));
  put(p->off_ident, g.ident.data(), g.ident.size() * sizeof(float));
  put(p->off_nsamp, p->num_samples.data(), p->num_samples.size() * sizeof(int32_t));
  put(p->off_passes, g.passes.data(), g.passes.size() * sizeof(DevPass));
}

std::atomic<int> g_wave_min{kWaveMinCircuits};
std::atomic<int> g_stream_min{kTiledMaxQubits + 1};

int BuildPlan(int mode, Inputs& in, int64_t B, int64_t S, int64_t O, int64_t K,
              const float* downstream, bool tfq_fd, tfqx_plan** out, char* err, int64_t len,
              bool with_expectation = false) {
  auto p = std::make_unique<tfqx_plan>();
  p->mode = mode;
  p->with_expectation = with_expectation;
  p->tfq_fd = tfq_fd;
  p->B = B;
  p->S = S;
  p->O = O;
  p->K = K;
  for (const auto& c : in.circuits) p->max_nq = std::max(p->max_nq, c.num_qubits);
  PlanOptions opt;
  opt.mode = mode;
  opt.adjoint = mode == kModeAdjoint || mode == kModeInnerProductGrad;
  opt.tfq_fd = tfq_fd;
  opt.fuse.max_fuse = AutoFusionWidth(p->max_nq);
  const int resident = LdsLimit(mode);
  const bool tiled_ok = (mode == kModeExpectation || mode == kModeAdjoint || mode == kModeState) &&
                        p->max_nq <= kTiledMaxQubits && p->max_nq < g_stream_min.load();
  opt.tile_log2 = resident;
  // Wave mode gives each small circuit one wavefront.  It pays when there
  // are enough such circuits to occupy the chip (256 CUs x 4 SIMDs); fewer
  // run faster as whole workgroups, 8 wavefronts per circuit.
  int wave_eligible = 0;
  for (const auto& c : in.circuits)
    wave_eligible += !c.gates.empty() && c.num_qubits >= kWaveRegBits && c.num_qubits <= kWaveMaxQubits;
  if (in.shared) wave_eligible *= int(B);
  opt.wave = (mode == kModeExpectation || mode == kModeAdjoint) && wave_eligible >= g_wave_min.load();
  if (p->max_nq > resident && !tiled_ok) {
    p->large = true;
    Status s = BuildLargePlan(mode, in.circuits, in.sums, downstream, in.others, B, S, O, K,
                              tfq_fd, with_expectation, &p->large_plan, g_fusion_width.load());
    if (!s.ok()) return Fail(err, len, s.code, s.msg);
    p->total = p->large_plan->workspace_bytes;
    *out = p.release();
    return TFQX_OK;
  }
  std::vector<CircuitJob> jobs(B);
  for (int64_t i = 0; i < B; ++i) {
    jobs[i].circuit = &in.circuits[in.shared ? 0 : i];
    jobs[i].sums = in.sums[i].empty() ? nullptr : &in.sums[i];
    const int64_t dstride = mode == kModeInnerProductGrad ? K : O;
    jobs[i].downstream = downstream ? downstream + size_t(i) * dstride : nullptr;
    jobs[i].others = K > 0 ? &in.others[i] : nullptr;
    jobs[i].row = int(i);
  }
  const int slots = (mode == kModeInnerProduct || mode == kModeInnerProductGrad) ? int(K) : int(O);
  // Plans can be rebound to new symbol values (and, for the adjoint, new
  // downstream weights): op lists decide diagonal storage from gate kinds and
  // keep every lambda term.
  opt.fuse.structural_diag = true;
  opt.keep_all_lambda = mode == kModeAdjoint;
  if (in.shared) {
    p->shared = true;
    p->tmpl = in.circuits[0];
    BuildSharedBatchProgram(p->tmpl, jobs, in.values, int(S), slots, opt, HostThreads(), &p->prog,
                            &p->recipe);
  } else {
    std::vector<SharedRecipe> recipes;
    BuildBatchProgram(jobs, slots, int(S), opt, HostThreads(), &p->prog, &recipes);
    p->rows_rebind.resize(size_t(B));
    ParallelFor(int(B), [&](int i) {
      p->rows_rebind[size_t(i)] = CompactRecipe(in.circuits[size_t(i)], recipes[size_t(i)]);
    });
    p->rebindable = true;
  }
  if (mode == kModeAdjoint) {
    p->lam_meta.resize(size_t(B));
    for (int64_t i = 0; i < B; ++i) {
      const Circuit& c = in.circuits[in.shared ? 0 : size_t(i)];
      if (c.gates.empty() || in.sums[size_t(i)].empty()) continue;
      for (int64_t j = 0; j < O; ++j)
        for (const auto& pt : in.sums[size_t(i)][size_t(j)])
          p->lam_meta[size_t(i)].push_back({int(j), pt.coef, pt.ny});
    }
  }
  p->lds_log2 = std::max(1, std::min(p->max_nq, resident));
  // Wave mode carves the full LDS allocation into one pair of 2^kWaveMaxQubits
  // buffers per wavefront.
  if (p->prog.num_block < int(p->prog.circuits.size())) p->lds_log2 = resident;
  p->kmax = 2;
  for (const auto& o : p->prog.ops) p->kmax = std::max<int>(p->kmax, o.k);
  Layout(p.get());
  *out = p.release();
  return TFQX_OK;
}

hipStream_t AsStream(void* s) { return reinterpret_cast<hipStream_t>(s); }


// Real floating-point operations one launch of a batch program performs, by
// the op's algorithm (not the instruction count): a dense k-qubit op is a
// 2^k x 2^k complex matrix-vector product per 2^k amplitudes, i.e.
// (8 * 2^k - 2) flops per amplitude; a diagonal op 6; a Pauli term's
// XOR-mask reduction or lambda accumulation 8.  Adjoint sweeps un-apply each
// op on psi and lambda; a parameterised op also applies dG (per symbol) and
// accumulates <lambda|dG psi>.
double AlgorithmicFlops(const BatchProgram& g) {
  double f = 0;
  auto op_flops = [](const DevOp& o) {
    if (o.type == kOpRemap) return 0.0;
    return o.type == kOpDiag ? 6.0 : 8.0 * double(1 << o.k) - 2.0;
  };
  for (const auto& c : g.circuits) {
    if (c.flags & kFlagEmpty) continue;
    const double A = std::ldexp(1.0, c.n);
    for (int i = c.fwd_begin; i < c.fwd_end; ++i) f += A * op_flops(g.ops[size_t(i)]);
    for (int i = c.rev_begin; i < c.rev_end; ++i) {
      const DevOp& o = g.ops[size_t(i)];
      f += 2 * A * op_flops(o);
      if (o.type == kOpParam) f += o.nderiv * A * (op_flops(o) + 8.0);
    }
    f += A * 8.0 * double(c.term_end - c.term_begin);
    f += A * 8.0 * double(c.lam_end - c.lam_begin);
    for (int k = c.other_begin; k < c.other_end; ++k) {
      const DevOther& o = g.others[size_t(k)];
      for (int i = o.op_begin; i < o.op_end; ++i) f += A * op_flops(g.ops[size_t(i)]);
      f += A * 8.0;
    }
  }
  return f;
}

// Algorithmic flops of one launch that run on MFMA: uncontrolled two-qubit
// dense and parameterised ops of op ranges the register engine runs (ranges
// that start with a remap).  Per amplitude: 8*4 - 2 per product; the adjoint
// applies G^dag to psi and lambda and one dG product per derivative.
double MfmaFlopsOfRange(const std::vector<DevOp>& ops, int b, int e, double A, bool rev) {
  if (b >= e || ops[size_t(b)].type != kOpRemap) return 0.0;
  double f = 0;
  for (int i = b; i < e; ++i) {
    const DevOp& o = ops[size_t(i)];
    if (o.k != 2 || o.cmask != 0) continue;
    if (o.type == kOpUnitary) f += (rev ? 2.0 : 1.0) * A * 30.0;
    if (o.type == kOpParam) f += (2.0 + o.nderiv) * A * 30.0;
  }
  return f;
}

// Op mix of a batch program in amplitude-op units (2^n per op of an n-qubit
// circuit, / 1024): [0] forward ops, [1] reverse remaps, [2] reverse diagonal
// ops, [3] reverse dense one-qubit ops, [4] reverse dense two-qubit (or
// wider) ops, [5] reverse parameterised one-qubit ops, [6] reverse
// parameterised two-qubit ops, [7] reverse parameterised ops whose matrix is
// diagonal, [8] forward remaps, [9] forward diagonal ops, [10] forward dense
// one-qubit ops, [11] forward dense wider ops.
// Remaps of op ranges [b, e) on an L-qubit tile (register engine) that keep
// the wavefront-index positions of the previous map (the engine runs them
// without workgroup barriers), in amplitude-op units; [13] counts all
// remaps of those ranges.
void LocalRemaps(const std::vector<DevOp>& ops, int b, int e, int L, double A, double& local,
                 double& all) {
  if (b >= e || ops[size_t(b)].type != kOpRemap || RegBitsForTile(L) == 0) return;
  auto wp = [L](uint32_t mask) {
    uint32_t rest = ((1u << L) - 1u) & ~mask;
    for (int k = 0; k < 6; ++k) rest &= rest - 1u;
    return rest;
  };
  int seen = 0;
  uint32_t prev = 0;
  for (int i = b; i < e; ++i) {
    const DevOp& o = ops[size_t(i)];
    if (o.type != kOpRemap) continue;
    all += A;
    if (seen >= 2 && wp(o.cmask) == wp(prev)) local += A;
    prev = o.cmask;
    ++seen;
  }
}

void OpMix(const BatchProgram& g, double (&m)[14]) {
  for (double& x : m) x = 0;
  for (const auto& c : g.circuits) {
    if ((c.flags & kFlagEmpty) || (c.flags & kFlagWave)) continue;
    const double A = std::ldexp(1.0, c.n) / 1024.0;
    if (c.flags & kFlagTiled) {
      for (int p = c.fwd_pass_begin; p < c.rev_pass_end; ++p) {
        const DevPass& ps = g.passes[size_t(p)];
        LocalRemaps(g.ops, ps.op_begin, ps.op_end, ps.L, A, m[12], m[13]);
      }
    } else {
      LocalRemaps(g.ops, c.fwd_begin, c.fwd_end, c.n, A, m[12], m[13]);
      LocalRemaps(g.ops, c.rev_begin, c.rev_end, c.n, A, m[12], m[13]);
    }
  }
  auto diag_matrix = [&](const DevOp& o) {
    const int D = 1 << o.k;
    for (int r = 0; r < D; ++r)
      for (int c = 0; c < D; ++c)
        if (r != c && g.mats[size_t(o.mat) + size_t(r * D + c)] != cfloat(0.f, 0.f)) return false;
    return true;
  };
  for (const auto& c : g.circuits) {
    if (c.flags & kFlagEmpty) continue;
    const double A = std::ldexp(1.0, c.n) / 1024.0;
    for (int i = c.fwd_begin; i < c.fwd_end; ++i) {
      const DevOp& o = g.ops[size_t(i)];
      m[0] += A * (o.type != kOpRemap);
      if (o.type == kOpRemap) m[8] += A;
      else if (o.type == kOpDiag) m[9] += A;
      else m[o.k == 1 ? 10 : 11] += A;
    }
    for (int i = c.rev_begin; i < c.rev_end; ++i) {
      const DevOp& o = g.ops[size_t(i)];
      if (o.type == kOpRemap) m[1] += A;
      else if (o.type == kOpDiag) m[2] += A;
      else if (o.type == kOpParam) {
        m[o.k == 1 ? 5 : 6] += A;
        if (diag_matrix(o)) m[7] += A;
      } else {
        m[o.k == 1 ? 3 : 4] += A;
      }
    }
  }
}

double MfmaFlops(const BatchProgram& g) {
  double f = 0;
  for (const auto& c : g.circuits) {
    if (c.flags & kFlagEmpty) continue;
    const double A = std::ldexp(1.0, c.n);
    if (c.flags & kFlagTiled) {
      for (int k = c.fwd_pass_begin; k < c.fwd_pass_end; ++k)
        f += MfmaFlopsOfRange(g.ops, g.passes[size_t(k)].op_begin, g.passes[size_t(k)].op_end, A, false);
      for (int k = c.rev_pass_begin; k < c.rev_pass_end; ++k)
        f += MfmaFlopsOfRange(g.ops, g.passes[size_t(k)].op_begin, g.passes[size_t(k)].op_end, A, true);
    } else {
      f += MfmaFlopsOfRange(g.ops, c.fwd_begin, c.fwd_end, A, false);
      f += MfmaFlopsOfRange(g.ops, c.rev_begin, c.rev_end, A, true);
    }
    for (int k = c.other_begin; k < c.other_end; ++k)
      f += MfmaFlopsOfRange(g.ops, g.others[size_t(k)].op_begin, g.others[size_t(k)].op_end, A, false);
  }
  return f;
}

BatchArgs MakeArgs(tfqx_plan* p, void* ws, void* output) {
  uint8_t* base = static_cast<uint8_t*>(ws);
  BatchArgs a{};
  a.mode = p->mode;
  a.kmax = p->kmax;
  a.circuits = reinterpret_cast<const DevCircuit*>(base + p->off_circuits);
  a.ncirc = int(p->prog.circuits.size());
  a.nblock = p->prog.num_block;
  a.nslots = p->mode == kModeInnerProduct ? int(p->K) : int(p->O);
  a.nsym = int(p->S);
  a.lds_log2 = p->lds_log2;
  a.lds_amps = 1 << p->lds_log2;
  a.ops = reinterpret_cast<const DevOp*>(base + p->off_ops);
  a.mats = reinterpret_cast<const float2*>(base + p->off_mats);
  a.terms = reinterpret_cast<const DevTerm*>(base + p->off_terms);
  a.others = reinterpret_cast<const DevOther*>(base + p->off_others);
  a.passes = reinterpret_cast<const DevPass*>(base + p->off_passes);
  a.scratch = reinterpret_cast<float2*>(base + p->off_scratch);
  a.scratch_stride = 2ll << p->max_nq;
  a.ident = reinterpret_cast<const float*>(base + p->off_ident);
  a.num_samples = reinterpret_cast<const int*>(base + p->off_nsamp);
  a.grad_acc = reinterpret_cast<double*>(base + p->off_gacc);
  a.queue = reinterpret_cast<int*>(base + p->off_counters);
  a.done = reinterpret_cast<int*>(base + p->off_counters + 64);
  a.wqueue = reinterpret_cast<int*>(base + p->off_counters + 128);
  a.seed = p->seed;
  a.nshots = p->nshots;
  a.max_nq = p->max_nq;
  a.state_stride = 1ll << p->max_nq;
  switch (p->mode) {
    case kModeExpectation:
    case kModeSampledExpectation: a.exp_out = static_cast<float*>(output); break;
    case kModeAdjoint:
      if (p->with_expectation) {
        a.exp_out = static_cast<float*>(output);
        a.grad_out = static_cast<float*>(output) + size_t(p->B) * size_t(p->O);
      } else {
        a.grad_out = static_cast<float*>(output);
      }
      break;
    case kModeState: a.state_out = static_cast<float2*>(output); break;
    case kModeSamples: a.samples_out = static_cast<signed char*>(output); break;
    default: a.ip_out = static_cast<float2*>(output); break;
  }
  // Z-basis rotations: Y^-0.5 and X^0.5 (circuit_parser_qsim.cc:897-945).
  float yp[5] = {-0.5f, 1.0f, 0.0f, 0, 0}, xp[5] = {0.5f, 1.0f, 0.0f, 0, 0};
  cdouble Uy[16], Ux[16];
  GateUnitary(kGateY, yp, Uy);
  GateUnitary(kGateX, xp, Ux);
  for (int i = 0; i < 4; ++i) {
    a.rot[i] = make_float2(float(Uy[i].real()), float(Uy[i].imag()));
    a.rot[4 + i] = make_float2(float(Ux[i].real()), float(Ux[i].imag()));
  }
  return a;
}

}  // namespace

// Workspace bytes a launch touches: the device-evaluation region only after
// values have been staged (it sits at the end of the workspace).
namespace tfqx {
int64_t PlanLaunchBytes(const tfqx_plan* p) {
  return int64_t(p->rb_built ? p->total : p->total - p->rb_bytes);
}
}  // namespace tfqx

namespace {

int CheckHip(hipError_t e, char* err, int64_t len, const char* what) {
  if (e == hipSuccess) return TFQX_OK;
  return Fail(err, len, TFQX_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

// Re-evaluate the plan's symbolic matrices in `ws` from the staged values
// already copied there (rebind_kernel, one workgroup per row).
int RunDeviceRebind(tfqx_plan* p, void* ws, hipStream_t s, char* err, int64_t len) {
  p->device_mats = true;
  if (p->rb_sizes.ops == 0) return TFQX_OK;
  uint8_t* base = static_cast<uint8_t*>(ws);
  const uint8_t* rb = base + p->off_rb;
  RebindArgs a;
  a.values = reinterpret_cast<const float*>(rb);
  a.eigen = reinterpret_cast<const RbEigen*>(rb + p->rb_off_eigen);
  a.rows = reinterpret_cast<const RbRow*>(rb + p->rb_off_rows);
  a.ops = reinterpret_cast<const RbOp*>(rb + p->rb_off_ops);
  a.gidx = reinterpret_cast<const int32_t*>(rb + p->rb_off_gidx);
  a.gates = reinterpret_cast<const RbGate*>(rb + p->rb_off_gates);
  a.mats = reinterpret_cast<float2*>(base + p->off_mats);
  a.num_rows = int(p->B);
  a.num_ops = int(p->rb_sizes.ops);
  a.shared = p->shared ? 1 : 0;
  a.num_symbols = int(p->S);
  a.tfq_fd = p->tfq_fd ? 1 : 0;
  return CheckHip(LaunchRebind(a, p->rb_kmax, s), err, len, "rebind kernel launch");
}

}  // namespace

extern "C" {

const char* tfqx_version(void) { return "tfqx 0.1 (gfx950)"; }

void tfqx_set_fusion_width(int32_t width) { g_fusion_width.store(width); }

void tfqx_set_reverse_fusion(int32_t on) { SetReverseFusion(on != 0); }

void tfqx_set_streaming_min_qubits(int32_t num_qubits) {
  const bool valid = num_qubits >= kLdsMaxQubitsOneBuffer && num_qubits <= kTiledMaxQubits + 1;
  g_stream_min.store(valid ? int(num_qubits) : kTiledMaxQubits + 1);
}

void tfqx_set_wave_min_circuits(int32_t count) {
  g_wave_min.store(count < 0 ? kWaveMinCircuits : int(count));
}

void tfqx_set_tuning(int32_t register_engine, int32_t tile_single, int32_t tile_two) {
  Tuning t;
  t.reg_engine = std::max(0, std::min(2, int(register_engine)));
  t.tile_single = std::max(0, int(tile_single));
  t.tile_two = std::max(0, int(tile_two));
  SetTuning(t);
}

int tfqx_plan_simulate_expectation(const uint8_t* programs, const int64_t* poff, int64_t B,
                                   const uint8_t* names, const int64_t* noff, int64_t S,
                                   const float* values, const uint8_t* psums,
                                   const int64_t* psoff, int64_t O, tfqx_plan** plan, char* err,
                                   int64_t len) {
  Inputs in;
  Status s = Prepare(programs, poff, B, names, noff, S, values, psums, psoff, O, nullptr,
                     nullptr, 0, &in, LdsLimit(kModeExpectation));
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  return BuildPlan(kModeExpectation, in, B, S, O, 0, nullptr, true, plan, err, len);
}

int tfqx_plan_simulate_sampled_expectation(const uint8_t* programs, const int64_t* poff,
                                           int64_t B, const uint8_t* names, const int64_t* noff,
                                           int64_t S, const float* values, const uint8_t* psums,
                                           const int64_t* psoff, int64_t O,
                                           const int32_t* num_samples, uint64_t seed,
                                           tfqx_plan** plan, char* err, int64_t len) {
  for (int64_t i = 0; i < B * O; ++i)
    if (num_samples[i] < 1)
      return Fail(err, len, TFQX_INVALID_ARGUMENT,
                  "Each element of num_samples must be greater than 0.");
  Inputs in;
  Status s = Prepare(programs, poff, B, names, noff, S, values, psums, psoff, O, nullptr,
                     nullptr, 0, &in, LdsLimit(kModeSampledExpectation));
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  int rc = BuildPlan(kModeSampledExpectation, in, B, S, O, 0, nullptr, true, plan, err, len);
  if (rc != TFQX_OK) return rc;
  (*plan)->seed = seed;
  (*plan)->num_samples.assign(num_samples, num_samples + B * O);
  if ((*plan)->large) {
    (*plan)->large_plan->seed = seed;
    (*plan)->large_plan->num_samples = (*plan)->num_samples;
  } else {
    Layout(*plan);
  }
  return TFQX_OK;
}

int tfqx_plan_simulate_samples(const uint8_t* programs, const int64_t* poff, int64_t B,
                               const uint8_t* names, const int64_t* noff, int64_t S,
                               const float* values, int32_t num_samples, uint64_t seed,
                               tfqx_plan** plan, char* err, int64_t len) {
  if (num_samples < 0)
    return Fail(err, len, TFQX_INVALID_ARGUMENT, "num_samples must be non-negative.");
  Inputs in;
  Status s = Prepare(programs, poff, B, names, noff, S, values, nullptr, nullptr, 0, nullptr,
                     nullptr, 0, &in, LdsLimit(kModeSamples));
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  int rc = BuildPlan(kModeSamples, in, B, S, 0, 0, nullptr, true, plan, err, len);
  if (rc != TFQX_OK) return rc;
  (*plan)->seed = seed;
  (*plan)->nshots = num_samples;
  if ((*plan)->large) {
    (*plan)->large_plan->seed = seed;
    (*plan)->large_plan->nshots = num_samples;
  }
  return TFQX_OK;
}

namespace {

int BuildNoisy(int mode, Inputs& in, int64_t B, int64_t O, const int32_t* num_samples, int nshots,
               uint64_t seed, tfqx_plan** plan, char* err, int64_t len) {
  auto p = std::make_unique<tfqx_plan>();
  p->mode = mode;
  p->B = B;
  p->O = O;
  p->nshots = nshots;
  p->seed = seed;
  p->noisy = true;
  Status s = BuildNoisyPlan(mode, in.circuits, in.sums, num_samples, nshots, B, O, &p->noisy_plan);
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  p->noisy_plan->seed = seed;
  p->max_nq = p->noisy_plan->max_nq;
  p->total = p->noisy_plan->total;
  *plan = p.release();
  return TFQX_OK;
}

}  // namespace

int tfqx_plan_noisy_expectation(const uint8_t* programs, const int64_t* poff, int64_t B,
                                const uint8_t* names, const int64_t* noff, int64_t S,
                                const float* values, const uint8_t* psums, const int64_t* psoff,
                                int64_t O, const int32_t* num_samples, uint64_t seed,
                                tfqx_plan** plan, char* err, int64_t len) {
  for (int64_t i = 0; i < B * O; ++i)
    if (num_samples[i] < 1)
      return Fail(err, len, TFQX_INVALID_ARGUMENT,
                  "Each element of num_samples must be greater than 0.");
  Inputs in;
  Status s = Prepare(programs, poff, B, names, noff, S, values, psums, psoff, O, nullptr, nullptr,
                     0, &in, -1, true);
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  return BuildNoisy(kModeNoisyExpectation, in, B, O, num_samples, 0, seed, plan, err, len);
}

int tfqx_plan_noisy_sampled_expectation(const uint8_t* programs, const int64_t* poff, int64_t B,
                                        const uint8_t* names, const int64_t* noff, int64_t S,
                                        const float* values, const uint8_t* psums,
                                        const int64_t* psoff, int64_t O,
                                        const int32_t* num_samples, uint64_t seed,
                                        tfqx_plan** plan, char* err, int64_t len) {
  for (int64_t i = 0; i < B * O; ++i)
    if (num_samples[i] < 1)
      return Fail(err, len, TFQX_INVALID_ARGUMENT,
                  "Each element of num_samples must be greater than 0.");
  Inputs in;
  Status s = Prepare(programs, poff, B, names, noff, S, values, psums, psoff, O, nullptr, nullptr,
                     0, &in, -1, true);
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  return BuildNoisy(kModeNoisySampledExpectation, in, B, O, num_samples, 0, seed, plan, err, len);
}

int tfqx_plan_noisy_samples(const uint8_t* programs, const int64_t* poff, int64_t B,
                            const uint8_t* names, const int64_t* noff, int64_t S,
                            const float* values, int32_t num_samples, uint64_t seed,
                            tfqx_plan** plan, char* err, int64_t len) {
  if (num_samples < 0)
    return Fail(err, len, TFQX_INVALID_ARGUMENT, "num_samples must be non-negative.");
  Inputs in;
  Status s = Prepare(programs, poff, B, names, noff, S, values, nullptr, nullptr, 0, nullptr,
                     nullptr, 0, &in, -1, true);
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  return BuildNoisy(kModeNoisySamples, in, B, 0, nullptr, num_samples, seed, plan, err, len);
}

int tfqx_plan_simulate_state(const uint8_t* programs, const int64_t* poff, int64_t B,
                             const uint8_t* names, const int64_t* noff, int64_t S,
                             const float* values, tfqx_plan** plan, char* err, int64_t len) {
  Inputs in;
  Status s = Prepare(programs, poff, B, names, noff, S, values, nullptr, nullptr, 0, nullptr,
                     nullptr, 0, &in, LdsLimit(kModeState));
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  return BuildPlan(kModeState, in, B, S, 0, 0, nullptr, true, plan, err, len);
}

int tfqx_plan_inner_product(const uint8_t* programs, const int64_t* poff, int64_t B,
                            const uint8_t* names, const int64_t* noff, int64_t S,
                            const float* values, const uint8_t* other, const int64_t* ooff,
                            int64_t K, tfqx_plan** plan, char* err, int64_t len) {
  Inputs in;
  Status s = Prepare(programs, poff, B, names, noff, S, values, nullptr, nullptr, 0, other, ooff,
                     K, &in);
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  return BuildPlan(kModeInnerProduct, in, B, S, 0, K, nullptr, true, plan, err, len);
}

int tfqx_plan_inner_product_grad(const uint8_t* programs, const int64_t* poff, int64_t B,
                                 const uint8_t* names, const int64_t* noff, int64_t S,
                                 const float* values, const uint8_t* other, const int64_t* ooff,
                                 int64_t K, const float* downstream, int32_t derivative,
                                 tfqx_plan** plan, char* err, int64_t len) {
  if (S <= 0)
    return Fail(err, len, TFQX_INVALID_ARGUMENT,
                "The number of symbols must be a positive integer, got " + std::to_string(S) +
                    " symbols.");
  Inputs in;
  Status s = Prepare(programs, poff, B, names, noff, S, values, nullptr, nullptr, 0, other, ooff,
                     K, &in);
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  return BuildPlan(kModeInnerProductGrad, in, B, S, 0, K, downstream,
                   derivative == TFQX_GRAD_TFQ_FD, plan, err, len);
}

int tfqx_plan_adjoint_gradient(const uint8_t* programs, const int64_t* poff, int64_t B,
                               const uint8_t* names, const int64_t* noff, int64_t S,
                               const float* values, const uint8_t* psums, const int64_t* psoff,
                               int64_t O, const float* downstream, int32_t derivative,
                               tfqx_plan** plan, char* err, int64_t len) {
  Inputs in;
  Status s = Prepare(programs, poff, B, names, noff, S, values, psums, psoff, O, nullptr,
                     nullptr, 0, &in, LdsLimit(kModeAdjoint));
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  return BuildPlan(kModeAdjoint, in, B, S, O, 0, downstream, derivative == TFQX_GRAD_TFQ_FD,
                   plan, err, len);
}

int tfqx_plan_expectation_and_gradient(const uint8_t* programs, const int64_t* poff, int64_t B,
                                       const uint8_t* names, const int64_t* noff, int64_t S,
                                       const float* values, const uint8_t* psums,
                                       const int64_t* psoff, int64_t O, const float* downstream,
                                       int32_t derivative, tfqx_plan** plan, char* err,
                                       int64_t len) {
  Inputs in;
  Status s = Prepare(programs, poff, B, names, noff, S, values, psums, psoff, O, nullptr,
                     nullptr, 0, &in, LdsLimit(kModeAdjoint));
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  return BuildPlan(kModeAdjoint, in, B, S, O, 0, downstream, derivative == TFQX_GRAD_TFQ_FD,
                   plan, err, len, /*with_expectation=*/true);
}

int tfqx_debug_fuse(const uint8_t* program, int64_t program_len, const uint8_t* names,
                    const int64_t* noff, int64_t S, const float* values, int32_t width,
                    int32_t* ops_out, int64_t ops_cap, int32_t* gates_out, int64_t gates_cap,
                    int64_t* n_ops, int64_t* n_gates, char* err, int64_t len) {
  if (!program || !n_ops || !n_gates || width < 1 || width > kMaxFuse || S < 0)
    return Fail(err, len, TFQX_INVALID_ARGUMENT, "invalid argument");
  Program prog;
  if (!DecodeProgram(program, size_t(program_len), &prog))
    return Fail(err, len, TFQX_INVALID_ARGUMENT,
                "Unparseable proto: " + std::string(reinterpret_cast<const char*>(program),
                                                    size_t(program_len)));
  std::unordered_map<std::string, int> columns;
  for (int64_t j = 0; j < S; ++j) columns[Bytes(names, noff, j)] = int(j);
  SymbolMap sm;
  sm.columns = &columns;
  sm.values = values;
  Circuit c;
  Status s = BuildCircuit(prog, sm, &c);
  if (!s.ok()) return Fail(err, len, s.code, s.msg);
  std::vector<FusedOp> fused;
  FuseOptions fo;
  fo.max_fuse = width;
  if (!c.gates.empty()) FuseGates(c, 0, int(c.gates.size()), fo, &fused);
  *n_ops = int64_t(fused.size());
  int64_t ng = 0;
  for (const auto& f : fused) ng += int64_t(f.gates.size());
  *n_gates = ng;
  if (int64_t(fused.size()) > ops_cap || ng > gates_cap) return TFQX_OK;  // sizes only
  int64_t gpos = 0;
  for (size_t i = 0; i < fused.size(); ++i) {
    const FusedOp& f = fused[i];
    int32_t* r = ops_out + 8 * i;
    r[0] = f.k;
    for (int j = 0; j < 4; ++j) r[1 + j] = j < f.k ? f.bits[j] : -1;
    r[5] = f.cmask != 0;
    r[6] = int32_t(gpos);
    for (int g : f.gates) gates_out[gpos++] = g;
    r[7] = int32_t(gpos);
  }
  return TFQX_OK;
}

int tfqx_debug_philox_uniform(uint64_t seed, uint64_t start, int64_t count, uint32_t stream_lo,
                              uint32_t stream_hi, double* out, void* stream, char* err,
                              int64_t len) {
  if (count < 0 || (count > 0 && !out)) return Fail(err, len, TFQX_INVALID_ARGUMENT, "invalid argument");
  const int rc = LaunchPhiloxUniforms(seed, start, count, stream_lo, stream_hi, out,
                                      static_cast<hipStream_t>(stream), err, len);
  return rc ? TFQX_DEVICE_ERROR : TFQX_OK;
}

int tfqx_plan_stage_values(tfqx_plan* p, const float* values, char* err, int64_t len) {
  if (!p || (!values && p->B * p->S > 0))
    return Fail(err, len, TFQX_INVALID_ARGUMENT, "null argument");
  if (!p->shared && !p->rebindable)
    return Fail(err, len, TFQX_INVALID_ARGUMENT,
                "Plan cannot be rebound: it runs on the HBM-streaming path or is a noisy plan.");
  if (!p->rb_built) {
    RebindTables t;
    const bool ok = p->shared ? BuildSharedRebindTables(p->tmpl, p->recipe, p->prog, &t)
                              : BuildRebindTables(p->rows_rebind, p->prog, &t);
    if (!ok)
      return Fail(err, len, TFQX_INVALID_ARGUMENT,
                  "Plan cannot be rebound on the device: a parameterised block is wider than "
                  "two qubits; use tfqx_plan_rebind.");
    if (t.ops.size() != p->rb_sizes.ops || t.gidx.size() != p->rb_sizes.gidx ||
        t.gates.size() != p->rb_sizes.gates || t.rows.size() != size_t(p->B))
      return Fail(err, len, TFQX_INTERNAL, "rebind table sizes differ from the layout");
    p->rb_image.assign(p->rb_bytes, 0);
    uint8_t* img = p->rb_image.data();
    double eig[kRbEigenKinds][100];
    EigenTableForDevice(eig);
    memcpy(img + p->rb_off_eigen, eig, sizeof(eig));
    auto put = [&](size_t o, const void* src, size_t n) {
      if (n) memcpy(img + o, src, n);
    };
    put(p->rb_off_rows, t.rows.data(), t.rows.size() * sizeof(RbRow));
    put(p->rb_off_ops, t.ops.data(), t.ops.size() * sizeof(RbOp));
    put(p->rb_off_gidx, t.gidx.data(), t.gidx.size() * sizeof(int32_t));
    put(p->rb_off_gates, t.gates.data(), t.gates.size() * sizeof(RbGate));
    p->rb_kmax = t.kmax;
    p->rb_built = true;
    p->rb_tables_dirty = true;
  }
  if (p->B * p->S > 0)
    memcpy(p->rb_image.data(), values, size_t(p->B) * size_t(p->S) * sizeof(float));
  p->values_staged = true;
  return TFQX_OK;
}

int tfqx_plan_rebind(tfqx_plan* p, const float* values, char* err, int64_t len) {
  if (!p || (!values && p->S > 0)) return Fail(err, len, TFQX_INVALID_ARGUMENT, "null argument");
  // The matrices are written straight into the upload image (prog.mats keeps
  // the values the plan was built with; only Layout reads it).
  cfloat* mats = reinterpret_cast<cfloat*>(p->staging.data() + p->off_mats);
  if (p->shared) {
    RebindSharedBatchProgram(p->tmpl, p->recipe, values, int(p->S), p->tfq_fd, HostThreads(),
                             p->prog, mats);
  } else if (p->rebindable) {
    RebindBatchProgram(p->rows_rebind, values, int(p->S), p->tfq_fd, HostThreads(), p->prog, mats);
  } else {
    return Fail(err, len, TFQX_INVALID_ARGUMENT,
                "Plan cannot be rebound: it runs on the HBM-streaming path or is a noisy plan.");
  }
  p->mats_dirty = true;
  p->device_mats = false;
  p->values_staged = false;
  return TFQX_OK;
}

int tfqx_plan_rebind_downstream(tfqx_plan* p, const float* downstream, char* err, int64_t len) {
  if (!p || (!downstream && p->B * p->O > 0))
    return Fail(err, len, TFQX_INVALID_ARGUMENT, "null argument");
  if (p->mode != kModeAdjoint || p->large || p->noisy || p->lam_meta.size() != size_t(p->B))
    return Fail(err, len, TFQX_INVALID_ARGUMENT,
                "Plan cannot be rebound to new downstream gradients: it is not an adjoint plan "
                "of the batch kernel.");
  for (const DevCircuit& dc : p->prog.circuits) {
    const auto& meta = p->lam_meta[size_t(dc.row)];
    if (size_t(dc.lam_end - dc.lam_begin) != meta.size())
      return Fail(err, len, TFQX_INTERNAL, "lambda term count mismatch");
    for (int k = dc.lam_begin; k < dc.lam_end; ++k) {
      const auto& m = meta[size_t(k - dc.lam_begin)];
      float w = downstream[size_t(dc.row) * p->O + m.slot] * m.coef;
      // AccumulateOperators (util_qsim.h:379-383) skips |g c| < 1e-5.
      if (std::fabs(w) < 1e-5f) w = 0.0f;
      float re = w, im = 0.0f;
      for (int y = 0; y < (m.ny & 3); ++y) {
        const float nr = -im, ni = re;
        re = nr;
        im = ni;
      }
      DevTerm& t = p->prog.terms[size_t(k)];
      t.re = re;
      t.im = im;
      memcpy(p->staging.data() + p->off_terms + size_t(k) * sizeof(DevTerm), &t, sizeof(DevTerm));
    }
  }
  p->terms_dirty = true;
  return TFQX_OK;
}

int tfqx_plan_stats(const tfqx_plan* p, int64_t* out, int64_t n) {
  if (!p || !out) return TFQX_INVALID_ARGUMENT;
  int64_t v[24] = {0};
  if (!p->large && !p->noisy && n > 10) {
    double mix[14];
    OpMix(p->prog, mix);
    for (int i = 0; i < 14; ++i) v[10 + i] = int64_t(mix[i]);
  }
  v[3] = p->max_nq;
  if (p->large) {
    const LargePlan& lp = *p->large_plan;
    v[0] = int64_t(lp.circuits.size());
    auto real_ops = [&](const DevPass& ps) {
      int64_t k = 0;
      for (int i = ps.op_begin; i < ps.op_end; ++i) k += lp.ops[size_t(i)].type != kOpRemap;
      return k;
    };
    double mf = 0;
    for (const auto& c : lp.circuits) {
      v[2] += int64_t(c.fwd.passes.size());
      v[5] += int64_t(c.rev.passes.size());
      for (const auto& ps : c.fwd.passes) v[1] += real_ops(ps);
      for (const auto& ps : c.rev.passes) v[4] += real_ops(ps);
      const double A = std::ldexp(1.0, c.n);
      for (const auto& ps : c.fwd.passes) {
        mf += MfmaFlopsOfRange(lp.ops, ps.op_begin, ps.op_end, A, false);
        // A dense pass: 8 * 2^k real flops per amplitude (the 2^(k+1)-square
        // real form of the complex block times every column), all on MFMA.
        if (ps.dense) mf += 8.0 * A * double(1 << lp.dense[size_t(ps.dense - 1)].k);
      }
      for (const auto& ps : c.rev.passes) mf += MfmaFlopsOfRange(lp.ops, ps.op_begin, ps.op_end, A, true);
    }
    v[9] = int64_t(mf);
    v[6] = 1;
  } else {
    v[0] = int64_t(p->prog.circuits.size());
    const auto& ops = p->prog.ops;
    for (const auto& c : p->prog.circuits) {
      for (int i = c.fwd_begin; i < c.fwd_end; ++i) v[1] += ops[size_t(i)].type != kOpRemap;
      for (int i = c.rev_begin; i < c.rev_end; ++i) v[4] += ops[size_t(i)].type != kOpRemap;
    }
    v[8] = int64_t(AlgorithmicFlops(p->prog));
    v[9] = int64_t(MfmaFlops(p->prog));
  }
  if (p->noisy) {
    const NoisyPlan& np = *p->noisy_plan;
    v[0] = int64_t(np.circuits.size());
    v[1] = 0;
    for (const auto& op : np.ops) v[1] += op.type != kOpRemap;
    v[2] = np.trajectories;
    // Algorithmic real flops per trajectory of each circuit: a dense k-qubit
    // op 8 * 2^k per amplitude, a diagonal op 6, a channel its expected
    // application (16 per amplitude, times the non-identity probability of a
    // unitary mixture) plus 8 per amplitude for the reduced density matrix of
    // other channels; each Pauli term 8 per amplitude.
    std::vector<double> per(np.circuits.size(), 0.0);
    for (size_t ci = 0; ci < np.circuits.size(); ++ci) {
      const DevCircuit& c = np.circuits[ci];
      const double A = std::ldexp(1.0, c.n);
      double f = 8.0 * A * double(c.term_end - c.term_begin);
      for (int i = c.fwd_begin; i < c.fwd_end; ++i) {
        const DevOp& op = np.ops[size_t(i)];
        if (op.type == kOpUnitary) f += 8.0 * A * double(1 << op.k);
        else if (op.type == kOpDiag) f += 6.0 * A;
        else if (op.type == kOpChannel)
          f += (op.cval & kChanMixture) ? 16.0 * A * (1.0 - double(np.mats[op.dmat[0]].real()))
                                         : 24.0 * A;
      }
      per[ci] = f;
    }
    double total = 0.0;
    for (const auto& ch : np.chunks) total += per[size_t(ch.circuit)] * double(ch.t1 - ch.t0);
    v[8] = int64_t(total);
  }
  v[7] = int64_t(p->total);
  for (int64_t i = 0; i < n && i < 24; ++i) out[i] = v[i];
  return TFQX_OK;
}

int64_t tfqx_plan_max_num_qubits(const tfqx_plan* p) { return p ? p->max_nq : -1; }
int64_t tfqx_plan_workspace_bytes(const tfqx_plan* p) { return p ? int64_t(p->total) : -1; }

int64_t tfqx_plan_output_bytes(const tfqx_plan* p) {
  if (!p) return -1;
  switch (p->mode) {
    case kModeExpectation:
    case kModeSampledExpectation: return p->B * p->O * 4;
    case kModeAdjoint: return (p->B * p->S + (p->with_expectation ? p->B * p->O : 0)) * 4;
    case kModeState: return p->B * (int64_t(1) << p->max_nq) * 8;
    case kModeSamples: return p->B * int64_t(p->nshots) * p->max_nq;
    case kModeInnerProductGrad: return p->B * p->S * 8;
    case kModeNoisyExpectation:
    case kModeNoisySampledExpectation: return p->B * p->O * 4;
    case kModeNoisySamples: return p->B * int64_t(p->nshots) * p->max_nq;
    default: return p->B * p->K * 8;
  }
}

int tfqx_plan_upload(tfqx_plan* p, void* ws, void* stream, char* err, int64_t len) {
  if (!p || (!ws && p->total)) return Fail(err, len, TFQX_INVALID_ARGUMENT, "null plan or workspace");
  if (p->large) return LargeUpload(p->large_plan.get(), ws, AsStream(stream), err, len);
  if (p->noisy) return NoisyUpload(p->noisy_plan.get(), ws, AsStream(stream), err, len);
  if (p->staging.empty()) return TFQX_OK;
  const hipStream_t s = AsStream(stream);
  int rc = CheckHip(hipMemcpyAsync(ws, p->staging.data(), p->staging.size(), hipMemcpyHostToDevice, s),
                    err, len, "hipMemcpyAsync");
  if (rc) return rc;
  uint8_t* base = static_cast<uint8_t*>(ws);
  if (p->rb_built) {
    rc = CheckHip(hipMemcpyAsync(base + p->off_rb, p->rb_image.data(), p->rb_image.size(),
                                 hipMemcpyHostToDevice, s),
                  err, len, "hipMemcpyAsync");
    if (rc) return rc;
    p->rb_tables_dirty = false;
    // The image's matrices predate values evaluated (or staged) for the device.
    if (p->values_staged || p->device_mats) {
      rc = RunDeviceRebind(p, ws, s, err, len);
      if (rc) return rc;
    }
  }
  p->last_ws = ws;
  p->mats_dirty = p->terms_dirty = p->values_staged = false;
  return TFQX_OK;
}

int tfqx_plan_upload_changed(tfqx_plan* p, void* ws, void* stream, char* err, int64_t len) {
  if (!p || (!ws && p->total)) return Fail(err, len, TFQX_INVALID_ARGUMENT, "null plan or workspace");
  if (p->large || p->noisy || p->staging.empty() || ws != p->last_ws)
    return tfqx_plan_upload(p, ws, stream, err, len);
  const hipStream_t s = AsStream(stream);
  uint8_t* base = static_cast<uint8_t*>(ws);
  auto copy = [&](size_t dst, const uint8_t* src, size_t n) -> int {
    if (!n) return TFQX_OK;
    return CheckHip(hipMemcpyAsync(base + dst, src, n, hipMemcpyHostToDevice, s), err, len,
                    "hipMemcpyAsync");
  };
  int rc = TFQX_OK;
  if (p->mats_dirty)
    rc = copy(p->off_mats, p->staging.data() + p->off_mats, p->prog.mats.size() * sizeof(cfloat));
  if (!rc && p->terms_dirty)
    rc = copy(p->off_terms, p->staging.data() + p->off_terms,
              p->prog.terms.size() * sizeof(DevTerm));
  if (!rc && p->rb_built) {
    if (p->rb_tables_dirty)
      rc = copy(p->off_rb, p->rb_image.data(), p->rb_image.size());
    else if (p->values_staged)
      rc = copy(p->off_rb, p->rb_image.data(), size_t(p->B) * size_t(p->S) * sizeof(float));
    if (!rc) p->rb_tables_dirty = false;
    if (!rc && p->values_staged) rc = RunDeviceRebind(p, ws, s, err, len);
  }
  if (rc) return rc;
  p->mats_dirty = p->terms_dirty = p->values_staged = false;
  return TFQX_OK;
}

int tfqx_plan_launch(tfqx_plan* p, void* ws, void* output, void* stream, char* err, int64_t len) {
  if (!p) return Fail(err, len, TFQX_INVALID_ARGUMENT, "null plan");
  if (p->large) return LargeLaunch(p->large_plan.get(), ws, output, AsStream(stream), err, len);
  if (p->noisy) return NoisyLaunch(p->noisy_plan.get(), ws, output, AsStream(stream), err, len);
  if (p->prog.circuits.empty()) return TFQX_OK;
  if (p->grid == 0) {
    int maxg = BatchMaxGrid(p->lds_log2, p->mode, p->kmax);
    if (p->prog.any_tiled) maxg = std::min(maxg, kTiledGridCap);
    p->grid = std::max(1, std::min<int>(maxg, int(p->prog.circuits.size())));
  }
  BatchArgs a = MakeArgs(p, ws, output);
  return CheckHip(LaunchBatch(a, p->grid, AsStream(stream)), err, len, "batch kernel launch");
}

int tfqx_plan_launch_plain(tfqx_plan* p, void* ws, void* output, void* stream, char* err,
                           int64_t len) {
  if (!p) return Fail(err, len, TFQX_INVALID_ARGUMENT, "null plan");
  if (p->large) return LargeLaunch(p->large_plan.get(), ws, output, AsStream(stream), err, len);
  if (p->noisy) return NoisyLaunch(p->noisy_plan.get(), ws, output, AsStream(stream), err, len);
  BatchArgs a = MakeArgs(p, ws, output);
  const DevCircuit* all = a.circuits;
  for (size_t c = 0; c < p->prog.circuits.size(); ++c) {
    a.circuits = all + c;
    a.ncirc = 1;
    // One circuit: run by the workgroup, or by a wavefront in wave mode.
    a.nblock = int(c) < p->prog.num_block ? 1 : 0;
    hipError_t e = LaunchBatch(a, 1, AsStream(stream));
    if (e != hipSuccess) return CheckHip(e, err, len, "plain launch");
  }
  return TFQX_OK;
}

int tfqx_graph_capture(tfqx_plan* p, void* ws, void* output, void* stream, tfqx_graph** g,
                       char* err, int64_t len) {
  hipStream_t s = AsStream(stream);
  // Resolve occupancy and kernel attributes before capture.
  if (!p->large && p->grid == 0 && !p->prog.circuits.empty()) {
    int maxg = BatchMaxGrid(p->lds_log2, p->mode, p->kmax);
    if (p->prog.any_tiled) maxg = std::min(maxg, kTiledGridCap);
    p->grid = std::max(1, std::min<int>(maxg, int(p->prog.circuits.size())));
  }
  auto graph = std::make_unique<tfqx_graph>();
  int rc = CheckHip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), err, len,
                    "hipStreamBeginCapture");
  if (rc) return rc;
  rc = tfqx_plan_launch(p, ws, output, stream, err, len);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(s, &gr);
  if (rc) return rc;
  if (e != hipSuccess) return CheckHip(e, err, len, "hipStreamEndCapture");
  graph->graph = gr;
  rc = CheckHip(hipGraphInstantiate(&graph->exec, gr, nullptr, nullptr, 0), err, len,
                "hipGraphInstantiate");
  if (rc) return rc;
  *g = graph.release();
  return TFQX_OK;
}

int tfqx_graph_launch(tfqx_graph* g, void* stream, char* err, int64_t len) {
  if (!g) return Fail(err, len, TFQX_INVALID_ARGUMENT, "null graph");
  return CheckHip(hipGraphLaunch(g->exec, AsStream(stream)), err, len, "hipGraphLaunch");
}

void tfqx_graph_destroy(tfqx_graph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}

void tfqx_plan_destroy(tfqx_plan* p) { delete p; }

}  // extern "C"
