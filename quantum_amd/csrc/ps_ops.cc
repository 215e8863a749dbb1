// Parameter-shift program rewriting (host ops behind the C-ABI).
//
// Restates, on the decoded Program structure, the three ops
// ParameterShift.get_gradient_circuits builds its shifted circuits from
// (python/differentiators/parameter_shift_util.py:parse_programs):
//   TfqPsDecompose            core/ops/tfq_ps_decompose_op.cc:39-160
//   TfqPsSymbolReplace        core/ops/tfq_ps_symbol_replace_op.cc:39-185
//   TfqPsWeightsFromSymbols   core/ops/tfq_ps_weights_from_symbols_op.cc:39-170
// Outputs are re-encoded programs (wire.cc:EncodeProgram).
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "tfqx.h"
#include "wire.h"

struct tfqx_strings {
  std::string data;
  std::vector<int64_t> offsets{0};
  std::vector<int64_t> shape;
  void push(const std::string& s) {
    data += s;
    offsets.push_back(int64_t(data.size()));
  }
};

namespace tfqx {
namespace {

int Fail(char* err, int64_t len, int code, const std::string& msg) {
  if (err && len > 0) {
    const size_t n = std::min<size_t>(size_t(len - 1), msg.size());
    memcpy(err, msg.data(), n);
    err[n] = '\0';
  }
  return code;
}

std::string Str(const uint8_t* b, const int64_t* off, int64_t i) {
  return std::string(reinterpret_cast<const char*>(b) + off[i], size_t(off[i + 1] - off[i]));
}

bool ParseAll(const uint8_t* progs, const int64_t* off, int64_t B, std::vector<Program>* out,
              std::string* bad) {
  out->resize(size_t(B));
  for (int64_t i = 0; i < B; ++i) {
    if (!DecodeProgram(progs + off[i], size_t(off[i + 1] - off[i]), &(*out)[size_t(i)])) {
      *bad = Str(progs, off, i);
      return false;
    }
  }
  return true;
}

Arg FloatArg(float v) {
  Arg a;
  a.has_value = true;
  a.f = v;
  return a;
}
Arg StrArg(const std::string& s) {
  Arg a;
  a.has_value = true;
  a.is_str = true;
  a.str = s;
  return a;
}
Arg SymArg(const std::string& s) {
  Arg a;
  a.has_symbol = true;
  a.symbol = s;
  return a;
}

// The reference reads missing map keys through operator[]: a default Arg.
Arg Get(const Operation& op, const char* key) {
  const Arg* a = op.find(key);
  return a ? *a : Arg();
}

void CopyControls(const Operation& from, Operation* to) {
  to->args.emplace_back("control_qubits", StrArg(Get(from, "control_qubits").str));
  to->args.emplace_back("control_values", StrArg(Get(from, "control_values").str));
}

// The new op's exponent from the source argument `target` (with scalar
// `scalar`): a symbol keeps the symbol with scalar sign * scalar / divisor; a
// value becomes sign * value / divisor with scalar 1.  Products in float, the
// division in double, stored as float (as the reference computes them).
void SetExponent(const Arg& target, float sign, float scalar, double divisor, Operation* op) {
  if (target.has_symbol) {
    op->args.emplace_back("exponent_scalar", FloatArg(float(double(sign * scalar) / divisor)));
    op->args.emplace_back("exponent", SymArg(target.symbol));
  } else if (target.has_value) {
    op->args.emplace_back("exponent_scalar", FloatArg(1.0f));
    op->args.emplace_back("exponent", FloatArg(float(double(sign * target.f) / divisor)));
  }
}

// XX / YY half of ISWAP^t: exponent_scalar * -0.5, global shift -0.5 (getOpForISP).
Operation IspPart(const Operation& op, const char* id, const std::string& symbol) {
  Operation o;
  o.gate_id = id;
  o.args.emplace_back("global_shift", FloatArg(-0.5f));
  o.args.emplace_back("exponent_scalar", FloatArg(float(double(Get(op, "exponent_scalar").f) * -0.5)));
  o.args.emplace_back("exponent", SymArg(symbol));
  CopyControls(op, &o);
  o.qubits = op.qubits;
  return o;
}

// Z^(+-p) of a PhasedISwap on its first (second) qubit (getOpForPISP).
Operation PispZ(const Operation& op, bool flip, bool second) {
  Operation o;
  o.gate_id = "ZP";
  o.args.emplace_back("global_shift", FloatArg(0.0f));
  const float sign = flip ? -1.0f : 1.0f;
  SetExponent(Get(op, "phase_exponent"), sign, Get(op, "phase_exponent_scalar").f, 1.0, &o);
  if (!op.qubits.empty()) {
    if (second) o.qubits.assign(op.qubits.begin() + 1, op.qubits.end());
    else o.qubits.assign(op.qubits.begin(), op.qubits.end() - 1);
  }
  CopyControls(op, &o);
  return o;
}

// Z^(-p) X^t Z^p parts of a PhasedX (getOpForPXP): global shift 0.
Operation PxpPart(const Operation& op, const char* id, const std::string& key, bool flip) {
  Operation o;
  o.gate_id = id;
  o.args.emplace_back("global_shift", FloatArg(0.0f));
  const float sign = flip ? -1.0f : 1.0f;
  SetExponent(Get(op, key.c_str()), sign, Get(op, (key + "_scalar").c_str()).f, 1.0, &o);
  o.qubits = op.qubits;
  CopyControls(op, &o);
  return o;
}

// XX^(theta/pi), YY^(theta/pi) (global shift -0.5) and CZ^(-phi/pi) parts of
// an FSim (getOpForFSIM; pi as the literal 3.14159265359).
Operation FsimPart(const Operation& op, const char* id, const std::string& key, bool shift) {
  Operation o;
  o.gate_id = id;
  o.args.emplace_back("global_shift", FloatArg(shift ? -0.5f : 0.0f));
  const float sign = key == "theta" ? 1.0f : -1.0f;
  SetExponent(Get(op, key.c_str()), sign, Get(op, (key + "_scalar").c_str()).f, 3.14159265359, &o);
  o.qubits = op.qubits;
  CopyControls(op, &o);
  return o;
}

bool Symbolic(const Operation& op, const char* key) {
  const Arg* a = op.find(key);
  return a && a->has_symbol;
}

Program Decompose(const Program& in) {
  Program out;
  constexpr int kBuffer = 5;
  for (const auto& moment : in.moments) {
    std::vector<Operation> cur = moment;
    std::vector<std::vector<Operation>> extra(kBuffer);
    int n_extra = 0;
    for (size_t k = 0; k < cur.size(); ++k) {
      const Operation op = cur[k];
      const std::string& id = op.gate_id;
      if (id == "PISP") {
        if (Symbolic(op, "exponent") || Symbolic(op, "phase_exponent")) {
          n_extra = 5;
          const std::string sym = Get(op, "exponent").symbol;
          cur[k] = PispZ(op, false, false);
          extra[0].push_back(PispZ(op, true, true));
          extra[1].push_back(IspPart(op, "XXP", sym));
          extra[2].push_back(IspPart(op, "YYP", sym));
          extra[3].push_back(PispZ(op, true, false));
          extra[4].push_back(PispZ(op, false, true));
        }
      } else if (id == "ISP") {
        if (Symbolic(op, "exponent")) {
          if (n_extra == 0) n_extra = 1;
          const std::string sym = Get(op, "exponent").symbol;
          cur[k] = IspPart(op, "XXP", sym);
          extra[0].push_back(IspPart(op, "YYP", sym));
        }
      } else if (id == "PXP") {
        if (Symbolic(op, "exponent") || Symbolic(op, "phase_exponent")) {
          n_extra = 2;
          cur[k] = PxpPart(op, "ZP", "phase_exponent", true);
          extra[0].push_back(PxpPart(op, "XP", "exponent", false));
          extra[1].push_back(PxpPart(op, "ZP", "phase_exponent", false));
        }
      } else if (id == "FSIM") {
        if (Symbolic(op, "theta") || Symbolic(op, "phi")) {
          n_extra = 2;
          cur[k] = FsimPart(op, "XXP", "theta", true);
          extra[0].push_back(FsimPart(op, "YYP", "theta", true));
          extra[1].push_back(FsimPart(op, "CZP", "phi", false));
        }
      }
    }
    // As in the reference, the extra-moment count is the last assignment in
    // the moment, and extra moments of different gates are shared.
    out.moments.push_back(std::move(cur));
    for (int l = 0; l < n_extra; ++l) out.moments.push_back(extra[size_t(l)]);
  }
  return out;
}

bool Ignored(const std::string& id) {
  static const char* kIgnore[] = {"I", "ISP", "PXP", "FSIM", "PISP", "AD", "ADP",
                                  "DP", "GAD", "BF", "PF", "PD", "RST"};
  for (const char* x : kIgnore)
    if (id == x) return true;
  return false;
}

}  // namespace
}  // namespace tfqx

using namespace tfqx;

extern "C" {

int tfqx_ps_decompose(const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
                      tfqx_strings** out, char* err, int64_t err_len) {
  if (!out || batch < 0) return Fail(err, err_len, TFQX_INVALID_ARGUMENT, "invalid arguments");
  std::vector<Program> progs;
  std::string bad;
  if (!ParseAll(programs, program_offsets, batch, &progs, &bad))
    return Fail(err, err_len, TFQX_INVALID_ARGUMENT, "Unparseable proto: " + bad);
  auto s = std::make_unique<tfqx_strings>();
  s->shape = {batch};
  for (const auto& p : progs) s->push(EncodeProgram(Decompose(p)));
  *out = s.release();
  return TFQX_OK;
}

int tfqx_ps_symbol_replace(const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
                           const uint8_t* symbols, const int64_t* symbol_offsets, int64_t num_symbols,
                           const uint8_t* replacements, const int64_t* replacement_offsets,
                           int64_t num_replacements, tfqx_strings** out, char* err,
                           int64_t err_len) {
  if (!out || batch < 0 || num_symbols < 0)
    return Fail(err, err_len, TFQX_INVALID_ARGUMENT, "invalid arguments");
  if (num_symbols != num_replacements)
    return Fail(err, err_len, TFQX_INVALID_ARGUMENT,
                "symbols.shape is not equal to replacement_symbols.shape: " +
                    std::to_string(num_symbols) + " != " + std::to_string(num_replacements));
  std::vector<Program> progs;
  std::string bad;
  if (!ParseAll(programs, program_offsets, batch, &progs, &bad))
    return Fail(err, err_len, TFQX_INVALID_ARGUMENT, "Unparseable proto: " + bad);
  // (i, j, k): the k-th program of symbols[j] in programs[i], one arg
  // occurrence of the symbol replaced, in moment / op / arg order.
  std::vector<std::vector<std::vector<std::string>>> res(
      static_cast<size_t>(batch),
      std::vector<std::vector<std::string>>(static_cast<size_t>(num_symbols)));
  size_t pad = 0;
  for (int64_t i = 0; i < batch; ++i) {
    const Program& p = progs[size_t(i)];
    for (int64_t j = 0; j < num_symbols; ++j) {
      const std::string sym = Str(symbols, symbol_offsets, j);
      const std::string rep = Str(replacements, replacement_offsets, j);
      auto& outs = res[size_t(i)][size_t(j)];
      for (size_t m = 0; m < p.moments.size(); ++m)
        for (size_t k = 0; k < p.moments[m].size(); ++k)
          for (size_t a = 0; a < p.moments[m][k].args.size(); ++a) {
            const Arg& arg = p.moments[m][k].args[a].second;
            const std::string cur = arg.has_symbol ? arg.symbol : std::string();
            if (cur != sym) continue;
            Program q = p;
            Arg& target = q.moments[m][k].args[a].second;
            target = Arg();
            target.has_symbol = true;
            target.symbol = rep;
            outs.push_back(EncodeProgram(q));
          }
      pad = std::max(pad, outs.size());
    }
  }
  auto s = std::make_unique<tfqx_strings>();
  s->shape = {batch, num_symbols, int64_t(pad)};
  const std::string empty = EncodeEmptyProgram();
  for (const auto& row : res)
    for (const auto& cell : row) {
      for (const auto& x : cell) s->push(x);
      for (size_t k = cell.size(); k < pad; ++k) s->push(empty);
    }
  *out = s.release();
  return TFQX_OK;
}

int tfqx_ps_weights_from_symbols(const uint8_t* programs, const int64_t* program_offsets,
                                 int64_t batch, const uint8_t* symbols,
                                 const int64_t* symbol_offsets, int64_t num_symbols, float* out,
                                 int64_t capacity, int64_t* pad, char* err, int64_t err_len) {
  if (!pad || batch < 0 || num_symbols < 0)
    return Fail(err, err_len, TFQX_INVALID_ARGUMENT, "invalid arguments");
  std::vector<Program> progs;
  std::string bad;
  if (!ParseAll(programs, program_offsets, batch, &progs, &bad))
    return Fail(err, err_len, TFQX_INVALID_ARGUMENT, "Unparseable proto: " + bad);
  std::unordered_map<std::string, int64_t> col;
  for (int64_t j = 0; j < num_symbols; ++j) col[Str(symbols, symbol_offsets, j)] = j;
  std::vector<std::vector<std::vector<float>>> res(
      static_cast<size_t>(batch), std::vector<std::vector<float>>(static_cast<size_t>(num_symbols)));
  size_t width = 0;
  for (int64_t i = 0; i < batch; ++i) {
    for (const auto& m : progs[size_t(i)].moments)
      for (const auto& op : m) {
        if (Ignored(op.gate_id)) continue;
        const Arg* e = op.find("exponent");
        if (!e)
          return Fail(err, err_len, TFQX_INVALID_ARGUMENT,
                      "Gate " + op.gate_id + " has no exponent argument.");
        if (!e->has_symbol) continue;
        auto it = col.find(e->symbol);
        if (it == col.end())
          return Fail(err, err_len, TFQX_INVALID_ARGUMENT,
                      "A circuit contains a sympy.Symbol not found in symbols!");
        res[size_t(i)][size_t(it->second)].push_back(Get(op, "exponent_scalar").f);
      }
    for (const auto& cell : res[size_t(i)]) width = std::max(width, cell.size());
  }
  *pad = int64_t(width);
  const int64_t need = batch * num_symbols * int64_t(width);
  if (!out || capacity < need) return TFQX_OK;  // size query
  for (int64_t i = 0; i < batch; ++i)
    for (int64_t j = 0; j < num_symbols; ++j) {
      const auto& cell = res[size_t(i)][size_t(j)];
      float* o = out + (i * num_symbols + j) * int64_t(width);
      for (size_t k = 0; k < width; ++k) o[k] = k < cell.size() ? cell[k] : 0.0f;
    }
  return TFQX_OK;
}

int64_t tfqx_strings_count(const tfqx_strings* s) {
  return s ? int64_t(s->offsets.size()) - 1 : -1;
}
const uint8_t* tfqx_strings_data(const tfqx_strings* s) {
  return s ? reinterpret_cast<const uint8_t*>(s->data.data()) : nullptr;
}
const int64_t* tfqx_strings_offsets(const tfqx_strings* s) { return s ? s->offsets.data() : nullptr; }
int64_t tfqx_strings_dim(const tfqx_strings* s, int64_t i) {
  return s && i >= 0 && i < int64_t(s->shape.size()) ? s->shape[size_t(i)] : -1;
}
void tfqx_strings_destroy(tfqx_strings* s) { delete s; }

}  // extern "C"
