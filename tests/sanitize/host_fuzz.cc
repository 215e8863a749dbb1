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
  printf("host_fuzz: %ld iterations, %llu circuits built, %llu text programs parsed\n", iters,
         (unsigned long long)built, (unsigned long long)text_ok);
  return built > 0 ? 0 : 2;
}
