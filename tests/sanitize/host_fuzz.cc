// Mutation fuzz of the library's host code under AddressSanitizer and
// UndefinedBehaviorSanitizer (the analogue of the reference's sanitizer run
// over core/src, scripts/msan_test.sh:16-35).  Test infrastructure only.
//
// Valid seeds (programs of every gate id with symbols and controls, Pauli sums,
// their text-format renderings) are mutated (bit flips, byte overwrites,
// truncation, insertion, splicing) and fed through everything that parses
// untrusted bytes: the binary and text decoders, qubit resolution and gate
// building, Pauli-sum resolution, the fuser at widths 1-4, stream-pass
// scheduling, and the parameter-shift rewriting op.  Any sanitizer report
// aborts the process with a non-zero status.
//
//   host_fuzz <iterations> <seed>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <string>
#include <unordered_map>
#include <vector>

#include "circuit.h"
#include "large.h"
#include "plan.h"
#include "textproto.h"
#include "tfqx.h"
#include "wire.h"

using namespace tfqx;

// large.cc's plan builder sizes its workspace with a function that lives in
// the kernel translation unit (large_kernels.hip, not built here); the fuzz
// never builds a streaming plan, so this definition is never called.
size_t tfqx::LargeWorkspaceBytes(LargePlan*) {
  fprintf(stderr, "host_fuzz: LargeWorkspaceBytes is not part of the host build\n");
  abort();
}

namespace {

// ---------------------------------------------------------------- seed corpus
Arg Value(float f) {
  Arg a;
  a.has_value = true;
  a.f = f;
  return a;
}
Arg Sym(const std::string& s) {
  Arg a;
  a.has_symbol = true;
  a.symbol = s;
  return a;
}
Arg Str(const std::string& s) {
  Arg a;
  a.has_value = true;
  a.is_str = true;
  a.str = s;
  return a;
}

Operation Op(const std::string& id, std::vector<std::string> qubits,
             std::vector<std::pair<std::string, Arg>> args) {
  Operation o;
  o.gate_id = id;
  o.qubits = std::move(qubits);
  o.args = std::move(args);
  return o;
}

std::vector<std::pair<std::string, Arg>> Eigen(Arg exponent, float scalar = 1.0f, float gs = 0.0f) {
  return {{"exponent", exponent}, {"exponent_scalar", Value(scalar)}, {"global_shift", Value(gs)},
          {"control_qubits", Str("")}, {"control_values", Str("")}};
}

std::vector<Program> SeedPrograms() {
  std::vector<Program> out;
  Program p;
  p.moments.push_back({Op("HP", {"0_0"}, Eigen(Value(1.0f))),
                       Op("XP", {"0_1"}, Eigen(Sym("a"), 0.5f, -0.5f)),
                       Op("ZZP", {"1_0", "1_1"}, Eigen(Value(0.3f)))});
  p.moments.push_back({Op("CNP", {"0_0", "0_1"}, Eigen(Value(1.0f))),
                       Op("ISP", {"1_0", "1_1"}, Eigen(Sym("b"), -1.0f))});
  p.moments.push_back({Op("FSIM", {"0_1", "1_1"},
                          {{"theta", Sym("a")}, {"theta_scalar", Value(1.0f)},
                           {"phi", Value(0.25f)}, {"phi_scalar", Value(1.0f)},
                           {"control_qubits", Str("")}, {"control_values", Str("")}}),
                       Op("PXP", {"0_0"},
                          {{"phase_exponent", Sym("b")}, {"phase_exponent_scalar", Value(1.0f)},
                           {"exponent", Value(0.5f)}, {"exponent_scalar", Value(1.0f)},
                           {"global_shift", Value(0.0f)}, {"control_qubits", Str("1_0")},
                           {"control_values", Str("1")}})});
  out.push_back(p);
  Program q;
  q.moments.push_back({Op("YP", {"3"}, Eigen(Sym("c"))), Op("SP", {"0", "1"}, Eigen(Value(0.5f))),
                       Op("I", {"2"}, {{"unused", Value(1.0f)}})});
  q.moments.push_back({Op("PISP", {"0", "3"},
                          {{"phase_exponent", Value(0.25f)}, {"phase_exponent_scalar", Value(1.0f)},
                           {"exponent", Sym("a")}, {"exponent_scalar", Value(1.0f)},
                           {"control_qubits", Str("1,2")}, {"control_values", Str("1,0")}}),
                       Op("CZP", {"1", "2"}, Eigen(Value(0.7f))),
                       Op("XXP", {"4", "5"}, Eigen(Sym("b"))),
                       Op("YYP", {"6", "7"}, Eigen(Value(-0.2f)))});
  q.moments.push_back({Op("DP", {"0"}, {{"p", Value(0.1f)}}), Op("I2", {"1", "2"}, {})});
  out.push_back(q);
  out.push_back(Program());
  return out;
}

void Varint(std::string* s, uint64_t v) {
  while (v >= 0x80) {
    s->push_back(char(uint8_t(v) | 0x80));
    v >>= 7;
  }
  s->push_back(char(v));
}
void LenField(std::string* s, int field, const std::string& body) {
  Varint(s, uint64_t(field) << 3 | 2);
  Varint(s, body.size());
  *s += body;
}
void FloatField(std::string* s, int field, float f) {
  Varint(s, uint64_t(field) << 3 | 5);
  char b[4];
  memcpy(b, &f, 4);
  s->append(b, 4);
}

// PauliSum bytes (pauli_sum.proto: terms = 1; PauliTerm: coefficient_real = 1,
// coefficient_imag = 2, paulis = 3; PauliQubitPair: qubit_id = 1, pauli_type = 2).
std::string EncodeSum(const std::vector<std::vector<std::pair<std::string, std::string>>>& terms) {
  std::string s;
  float c = 0.5f;
  for (const auto& t : terms) {
    std::string term;
    FloatField(&term, 1, c);
    FloatField(&term, 2, 0.25f);
    c += 0.75f;
    for (const auto& pq : t) {
      std::string pair;
      LenField(&pair, 1, pq.first);
      LenField(&pair, 2, pq.second);
      LenField(&term, 3, pair);
    }
    LenField(&s, 1, term);
  }
  return s;
}

std::vector<std::string> SeedSums() {
  return {EncodeSum({{{"0_0", "Z"}, {"1_1", "X"}}, {{"0_1", "Y"}}, {}}),
          EncodeSum({{{"0", "X"}, {"3", "Z"}, {"7", "Y"}}, {{"2", "Z"}}}), EncodeSum({})};
}

const char* kTextSeeds[] = {
    "language { gate_set: \"tfq_gate_set\" } circuit { scheduling_strategy: MOMENT_BY_MOMENT "
    "moments { operations { gate { id: \"HP\" } args { key: \"exponent\" value { arg_value { "
    "float_value: 1.0 } } } args { key: \"exponent_scalar\" value { arg_value { float_value: 1 } } "
    "} args { key: \"global_shift\" value { arg_value { float_value: 0.0 } } } qubits { id: "
    "\"0_0\" } } } }",
    "terms { coefficient_real: 1.5 paulis { qubit_id: \"0_0\" pauli_type: \"Z\" } }",
    "circuit { moments { operations < gate < id: 'XP' > args { key: 'exponent' value { symbol: "
    "'x' } } qubits < id: '2' > > } } # comment",
};

// ---------------------------------------------------------------- mutations
std::string Mutate(std::string s, std::mt19937_64& rng, const std::vector<std::string>& pool) {
  const int n = 1 + int(rng() % 4);
  for (int k = 0; k < n; ++k) {
    const int kind = int(rng() % 6);
    const size_t len = s.size();
    switch (kind) {
      case 0:  // bit flip
        if (len) s[rng() % len] ^= char(1u << (rng() % 8));
        break;
      case 1:  // byte overwrite with an interesting value
        if (len) {
          static const uint8_t vals[] = {0x00, 0x01, 0x7f, 0x80, 0xff, 0x0a, 0x12, 0x1a, 0x22, 0x08};
          s[rng() % len] = char(vals[rng() % sizeof(vals)]);
        }
        break;
      case 2:  // truncate
        if (len) s.resize(rng() % len);
        break;
      case 3: {  // insert random bytes
        const size_t at = len ? rng() % (len + 1) : 0;
        std::string ins(1 + rng() % 8, '\0');
        for (auto& c : ins) c = char(rng());
        s.insert(at, ins);
        break;
      }
      case 4: {  // splice a piece of another seed
        const std::string& o = pool[rng() % pool.size()];
        if (!o.empty()) {
          const size_t a = rng() % o.size(), b = a + rng() % (o.size() - a + 1);
          s.insert(len ? rng() % (len + 1) : 0, o.substr(a, b - a));
        }
        break;
      }
      default:  // duplicate a range
        if (len) {
          const size_t a = rng() % len, b = a + rng() % (len - a + 1);
          s.insert(b, s.substr(a, b - a));
        }
    }
  }
  return s;
}

// ---------------------------------------------------------------- drive the host code
void ExercisePlan(const Circuit& c) {
  if (c.gates.empty() || c.num_qubits > 20) return;
  for (int w = 1; w <= kMaxFuse; ++w) {
    FuseOptions fo;
    fo.max_fuse = w;
    std::vector<FusedOp> fused;
    FuseGates(c, 0, int(c.gates.size()), fo, &fused);
    if (c.num_qubits >= 6 && w <= 2) {
      std::vector<DevPass> passes;
      std::vector<DevOp> ops;
      std::vector<cfloat> mats;
      SchedulePasses(fused, c.num_qubits, std::min(13 + (w == 2), c.num_qubits), &passes, &ops,
                     &mats, w == 2 ? 5 : 0);
    }
  }
}

// The device re-evaluation tables of one-row plans (rebind_kernel indexes
// them without bounds checks): every op's gates, matrices and qubits must lie
// inside what the plan holds; then the host re-evaluation runs on the same
// recipes.
uint64_t g_rebind_ops = 0;

void CheckRebindTables(const RebindTables& t, const BatchProgram& prog, int nq) {
  g_rebind_ops += t.ops.size();
  for (const RbOp& o : t.ops) {
    const RbRow& row = t.rows[size_t(o.row)];
    const int dim = 1 << o.k;
    if (o.k < 1 || o.k > kMaxFuse || (o.kind == 3 && o.k > 2)) abort();
    for (int j = 0; j < o.k; ++j)
      if (o.bits[j] < 0 || o.bits[j] >= nq) abort();
    if (o.gate_begin < 0 || o.gate_end <= o.gate_begin ||
        size_t(row.gidx_base + o.gate_end) > t.gidx.size())
      abort();
    for (int g = o.gate_begin; g < o.gate_end; ++g) {
      const int gi = row.gate_base + t.gidx[size_t(row.gidx_base + g)];
      if (gi < 0 || size_t(gi) >= t.gates.size()) abort();
      const RbGate& rg = t.gates[size_t(gi)];
      if (rg.nt < 1 || rg.nt > 2 || rg.nsym < 0 || rg.nsym > 2) abort();
      for (int j = 0; j < rg.nsym; ++j)
        if (rg.sym_pidx[j] < 0 || rg.sym_pidx[j] > 4 || rg.sym_col[j] < 0 || rg.sym_col[j] > 2) abort();
    }
    const size_t m = size_t(o.diag && o.kind != 3 ? dim : dim * dim);
    if (row.mat_base + size_t(o.mat) + m > prog.mats.size()) abort();
    for (int d = 0; d < o.nderiv; ++d) {
      if (o.kind == 3 && (o.dgate[d] < 0 || o.dgate[d] >= o.gate_end - o.gate_begin)) abort();
      if (row.mat_base + size_t(o.dmat[d]) + size_t(dim * dim) > prog.mats.size()) abort();
    }
  }
}

void ExerciseRebind(const Circuit& c, const float* vals) {
  if (c.gates.empty() || c.num_qubits > 13) return;
  for (int adj = 0; adj < 6; ++adj) {
    PlanOptions opt;
    opt.mode = adj & 1 ? kModeAdjoint : kModeExpectation;
    opt.adjoint = (adj & 1) != 0;
    opt.fuse.max_fuse = 2 + adj / 2;  // widths 2..4
    opt.fuse.structural_diag = true;
    opt.keep_all_lambda = opt.adjoint;
    std::vector<CircuitJob> jobs(1);
    jobs[0] = CircuitJob{&c, nullptr, nullptr, nullptr, 0};
    BatchProgram prog;
    std::vector<SharedRecipe> recipes;
    BuildBatchProgram(jobs, 1, 3, opt, 1, &prog, &recipes);
    std::vector<RowRebind> rows{CompactRecipe(c, recipes[0])};
    RebindTables t;
    if (BuildRebindTables(rows, prog, &t)) {
      const RebindTableSizes z = SizeRebindTables(rows);
      if (t.ops.size() != z.ops || t.gidx.size() != z.gidx || t.gates.size() != z.gates) abort();
      CheckRebindTables(t, prog, c.num_qubits);
    }
    std::vector<cfloat> mats(prog.mats);
    RebindBatchProgram(rows, vals, 3, opt.adjoint, 1, prog, mats.data());
    BatchProgram sprog;
    SharedRecipe srec;
    BuildSharedBatchProgram(c, jobs, vals, 3, 1, opt, 1, &sprog, &srec);
    RebindTables st;
    if (BuildSharedRebindTables(c, srec, sprog, &st)) {
      if (st.ops.size() != SizeSharedRebindTables(c, srec).ops) abort();
      for (RbOp& o : st.ops) o.row = 0;
      CheckRebindTables(st, sprog, c.num_qubits);
    }
  }
}

void OneProgram(const std::string& bytes, const std::vector<std::string>& sums, uint64_t* ok) {
  Program prog;
  if (!DecodeProgram(reinterpret_cast<const uint8_t*>(bytes.data()), bytes.size(), &prog)) return;
  std::unordered_map<std::string, int> cols{{"a", 0}, {"b", 1}, {"c", 2}, {"x", 0}};
  const float vals[3] = {0.25f, -1.5f, 3.0f};
  SymbolMap sm;
  sm.columns = &cols;
  sm.values = vals;
  Circuit c;
  if (BuildCircuit(prog, sm, &c).ok()) {
    ++*ok;
    ExercisePlan(c);
    ExerciseRebind(c, vals);
    Circuit r;
    RebindCircuit(c, vals, &r);
    for (const auto& sb : sums) {
      PauliSum ps;
      if (!DecodePauliSum(reinterpret_cast<const uint8_t*>(sb.data()), sb.size(), &ps)) continue;
      std::vector<Pauli> out;
      ResolvePauliSum(ps, c, &out);
    }
    Circuit paired;
    BuildPairedCircuit(prog, c, &paired);
  }
  Circuit noisy;
  BuildNoisyCircuit(prog, sm, &noisy);
  // Parameter-shift rewriting of the raw bytes (decode -> rewrite -> encode).
  const int64_t off[2] = {0, int64_t(bytes.size())};
  tfqx_strings* res = nullptr;
  char err[512];
  if (tfqx_ps_decompose(reinterpret_cast<const uint8_t*>(bytes.data()), off, 1, &res, err,
                        sizeof(err)) == TFQX_OK)
    tfqx_strings_destroy(res);
}

}  // namespace

int main(int argc, char** argv) {
  const long iters = argc > 1 ? atol(argv[1]) : 20000;
  std::mt19937_64 rng(argc > 2 ? strtoull(argv[2], nullptr, 10) : 1);
  std::vector<std::string> progs, sums = SeedSums(), texts;
  for (const auto& p : SeedPrograms()) progs.push_back(EncodeProgram(p));
  progs.push_back(EncodeEmptyProgram());
  for (const char* t : kTextSeeds) texts.push_back(t);
  std::vector<std::string> pool = progs;
  pool.insert(pool.end(), sums.begin(), sums.end());
  pool.insert(pool.end(), texts.begin(), texts.end());
  uint64_t built = 0, text_ok = 0;
  for (long i = 0; i < iters; ++i) {
    const int which = int(rng() % 4);
    if (which == 0) {
      std::string s = Mutate(sums[rng() % sums.size()], rng, pool);
      PauliSum ps;
      DecodePauliSum(reinterpret_cast<const uint8_t*>(s.data()), s.size(), &ps);
      continue;
    }
    if (which == 1) {
      std::string t = Mutate(texts[rng() % texts.size()], rng, pool), bin;
      text_ok += TextProgramToBinary(t, &bin);
      TextPauliSumToBinary(t, &bin);
      OneProgram(t, sums, &built);
      continue;
    }
    std::vector<std::string> ms;
    for (const auto& s : sums) ms.push_back(rng() % 2 ? s : Mutate(s, rng, pool));
    OneProgram(Mutate(progs[rng() % progs.size()], rng, pool), ms, &built);
  }
  // The unmutated seeds must build (the fuzz reaches the deep paths).
  for (const auto& p : progs) OneProgram(p, sums, &built);
  printf("host_fuzz: %ld iterations, %llu circuits built, %llu text programs parsed, %llu "
         "device rebind ops checked\n", iters, (unsigned long long)built,
         (unsigned long long)text_ok, (unsigned long long)g_rebind_ops);
  return built > 0 ? 0 : 2;
}
