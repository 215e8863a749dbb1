// Program -> gate list (see circuit.h for the reference correspondence).
#include "circuit.h"

#include <stdlib.h>

#include <algorithm>
#include <cmath>
#include <set>
#include <tuple>

namespace tfqx {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr float kGradEps = 5e-3f;  // adj_util.cc:32

// ------------------------------------------------------------------ eigen components
struct Eigen {
  int count;
  double lam[3];
  cdouble P[3][16];
};

void SetBlock(cdouble* P, int dim, std::initializer_list<int> idx,
              std::initializer_list<cdouble> vals) {
  // vals is a |idx| x |idx| row-major block placed at rows/cols idx.
  std::vector<int> ix(idx);
  std::vector<cdouble> v(vals);
  int m = int(ix.size());
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) P[ix[r] * dim + ix[c]] += v[r * m + c];
}

Eigen MakeEigen(GateKind kind) {
  Eigen e{};
  const cdouble I(0, 1);
  auto zero = [&](int k) {
    for (auto& x : e.P[k]) x = 0;
  };
  for (int k = 0; k < 3; ++k) zero(k);
  switch (kind) {
    case kGateX:
      e.count = 2; e.lam[0] = 0; e.lam[1] = 1;
      SetBlock(e.P[0], 2, {0, 1}, {0.5, 0.5, 0.5, 0.5});
      SetBlock(e.P[1], 2, {0, 1}, {0.5, -0.5, -0.5, 0.5});
      break;
    case kGateY:
      e.count = 2; e.lam[0] = 0; e.lam[1] = 1;
      SetBlock(e.P[0], 2, {0, 1}, {0.5, -0.5 * I, 0.5 * I, 0.5});
      SetBlock(e.P[1], 2, {0, 1}, {0.5, 0.5 * I, -0.5 * I, 0.5});
      break;
    case kGateZ:
      e.count = 2; e.lam[0] = 0; e.lam[1] = 1;
      e.P[0][0] = 1; e.P[1][3] = 1;
      break;
    case kGateH: {
      e.count = 2; e.lam[0] = 0; e.lam[1] = 1;
      const double r = 0.5 / std::sqrt(2.0);
      SetBlock(e.P[0], 2, {0, 1}, {0.5 + r, r, r, 0.5 - r});
      SetBlock(e.P[1], 2, {0, 1}, {0.5 - r, -r, -r, 0.5 + r});
      break;
    }
    case kGateXX:
    case kGateYY:
    case kGateZZ: {
      e.count = 2; e.lam[0] = 0; e.lam[1] = 1;
      // PP for P in {X, Y, Z}; (I +- PP)/2.
      cdouble PP[16] = {};
      if (kind == kGateXX) { PP[0 * 4 + 3] = PP[1 * 4 + 2] = PP[2 * 4 + 1] = PP[3 * 4 + 0] = 1; }
      if (kind == kGateYY) { PP[0 * 4 + 3] = -1; PP[1 * 4 + 2] = 1; PP[2 * 4 + 1] = 1; PP[3 * 4 + 0] = -1; }
      if (kind == kGateZZ) { PP[0] = 1; PP[5] = -1; PP[10] = -1; PP[15] = 1; }
      for (int i = 0; i < 16; ++i) {
        cdouble id = (i % 5 == 0) ? 1.0 : 0.0;
        e.P[0][i] = 0.5 * (id + PP[i]);
        e.P[1][i] = 0.5 * (id - PP[i]);
      }
      break;
    }
    case kGateCZ:
      e.count = 2; e.lam[0] = 0; e.lam[1] = 1;
      e.P[0][0] = e.P[0][5] = e.P[0][10] = 1; e.P[1][15] = 1;
      break;
    case kGateCX:
      e.count = 2; e.lam[0] = 0; e.lam[1] = 1;
      e.P[0][0] = e.P[0][5] = 1;
      SetBlock(e.P[0], 4, {2, 3}, {0.5, 0.5, 0.5, 0.5});
      SetBlock(e.P[1], 4, {2, 3}, {0.5, -0.5, -0.5, 0.5});
      break;
    case kGateSwap:
      e.count = 2; e.lam[0] = 0; e.lam[1] = 1;
      e.P[0][0] = e.P[0][15] = 1;
      SetBlock(e.P[0], 4, {1, 2}, {0.5, 0.5, 0.5, 0.5});
      SetBlock(e.P[1], 4, {1, 2}, {0.5, -0.5, -0.5, 0.5});
      break;
    case kGateISwap:
      e.count = 3; e.lam[0] = 0; e.lam[1] = 0.5; e.lam[2] = -0.5;
      e.P[0][0] = e.P[0][15] = 1;
      SetBlock(e.P[1], 4, {1, 2}, {0.5, 0.5, 0.5, 0.5});
      SetBlock(e.P[2], 4, {1, 2}, {0.5, -0.5, -0.5, 0.5});
      break;
    default:
      e.count = 0;
  }
  return e;
}

const Eigen& EigenOf(GateKind kind) {
  static const Eigen table[] = {
      MakeEigen(kGateI1), MakeEigen(kGateI2), MakeEigen(kGateH), MakeEigen(kGateX),
      MakeEigen(kGateY), MakeEigen(kGateZ), MakeEigen(kGateXX), MakeEigen(kGateYY),
      MakeEigen(kGateZZ), MakeEigen(kGateCZ), MakeEigen(kGateCX), MakeEigen(kGateSwap),
      MakeEigen(kGateISwap)};
  return table[int(kind)];
}

bool IsEigen(GateKind k) { return k >= kGateH && k <= kGateISwap; }

// U(t, gs) = sum_k e^{i pi t (l_k + gs)} P_k; dU/dt when deriv.
void EigenEval(GateKind kind, double t, double gs, bool deriv, cdouble* U) {
  const Eigen& e = EigenOf(kind);
  const int dim = GateTargets(kind) == 1 ? 2 : 4;
  for (int i = 0; i < dim * dim; ++i) U[i] = 0;
  for (int k = 0; k < e.count; ++k) {
    const double ph = kPi * t * (e.lam[k] + gs);
    cdouble w(std::cos(ph), std::sin(ph));
    if (deriv) w *= cdouble(0, kPi * (e.lam[k] + gs));
    for (int i = 0; i < dim * dim; ++i) U[i] += w * e.P[k][i];
  }
}

inline float Mul32(float a, float b) { return a * b; }  // float32 product, as circuit_parser_qsim.cc:229

void UnitaryOrDeriv(GateKind kind, const float* p, int pidx, cdouble* U) {
  // pidx < 0: unitary; otherwise the exact derivative wrt p[pidx].
  const cdouble I(0, 1);
  if (kind == kGateI1 || kind == kGateI2) {
    int dim = kind == kGateI1 ? 2 : 4;
    for (int i = 0; i < dim * dim; ++i) U[i] = (pidx < 0 && i % (dim + 1) == 0) ? 1.0 : 0.0;
    return;
  }
  if (IsEigen(kind)) {
    const double t = Mul32(p[0], p[1]);
    EigenEval(kind, t, p[2], pidx >= 0, U);
    if (pidx >= 0)
      for (int i = 0; i < 16; ++i) U[i] *= double(p[1]);
    return;
  }
  switch (kind) {
    case kGatePhasedX: {
      // U = Z^p X^t Z^-p  ->  off-diagonals pick up e^{-+i pi p}.
      const double ph = Mul32(p[0], p[1]);
      const double t = Mul32(p[2], p[3]);
      cdouble X[4];
      EigenEval(kGateX, t, p[4], pidx == 2, X);
      const cdouble em = std::exp(cdouble(0, -kPi * ph)), ep = std::exp(cdouble(0, kPi * ph));
      U[0] = X[0]; U[1] = X[1] * em; U[2] = X[2] * ep; U[3] = X[3];
      if (pidx == 0) {
        U[0] = 0; U[3] = 0;
        U[1] *= cdouble(0, -kPi) * double(p[1]);
        U[2] *= cdouble(0, kPi) * double(p[1]);
      } else if (pidx == 2) {
        for (int i = 0; i < 4; ++i) U[i] *= double(p[3]);
      }
      return;
    }
    case kGateFSim: {
      const double th = Mul32(p[0], p[1]);
      const double phi = Mul32(p[2], p[3]);
      for (int i = 0; i < 16; ++i) U[i] = 0;
      const double c = std::cos(th), s = std::sin(th);
      if (pidx < 0) {
        U[0] = 1; U[5] = c; U[6] = -I * s; U[9] = -I * s; U[10] = c;
        U[15] = std::exp(cdouble(0, -phi));
      } else if (pidx == 0) {
        const double sc = p[1];
        U[5] = -s * sc; U[6] = -I * c * sc; U[9] = -I * c * sc; U[10] = -s * sc;
      } else {
        U[15] = -I * std::exp(cdouble(0, -phi)) * double(p[3]);
      }
      return;
    }
    case kGatePhasedISwap: {
      // (Z^-p x Z^p) ISWAP^t (Z^p x Z^-p): U[01,10] gains e^{2i pi p}, U[10,01] e^{-2i pi p}.
      const double ph = Mul32(p[0], p[1]);
      const double t = Mul32(p[2], p[3]);
      EigenEval(kGateISwap, t, 0.0, pidx == 2, U);
      const cdouble ep = std::exp(cdouble(0, 2 * kPi * ph)), em = std::exp(cdouble(0, -2 * kPi * ph));
      U[6] *= ep;
      U[9] *= em;
      if (pidx == 0) {
        for (int i = 0; i < 16; ++i)
          if (i != 6 && i != 9) U[i] = 0;
        U[6] *= cdouble(0, 2 * kPi) * double(p[1]);
        U[9] *= cdouble(0, -2 * kPi) * double(p[1]);
      } else if (pidx == 2) {
        for (int i = 0; i < 16; ++i) U[i] *= double(p[3]);
      }
      return;
    }
    default:
      return;
  }
}

// ------------------------------------------------------------------ qubit ids
bool ParseQubitKey(const std::string& id, std::pair<int, int>* rc) {
  // "r_c" (GridQubit) or "x" (LineQubit, row padded with INT_MAX).
  auto parse_int = [](const std::string& s, int* v) {
    if (s.empty()) return false;
    char* end = nullptr;
    long x = strtol(s.c_str(), &end, 10);
    if (*end != '\0' || x > 2147483647L || x < -2147483647L - 1) return false;
    *v = int(x);
    return true;
  };
  size_t us = id.find('_');
  if (us == std::string::npos) {
    rc->first = 2147483647;
    return parse_int(id, &rc->second);
  }
  if (id.find('_', us + 1) != std::string::npos) return false;
  return parse_int(id.substr(0, us), &rc->first) && parse_int(id.substr(us + 1), &rc->second);
}

std::vector<std::string> SplitCsv(const std::string& s) {
  std::vector<std::string> out;
  size_t start = 0;
  while (true) {
    size_t c = s.find(',', start);
    out.push_back(s.substr(start, c == std::string::npos ? std::string::npos : c - start));
    if (c == std::string::npos) break;
    start = c + 1;
  }
  return out;
}

const std::string& ControlQubits(const Operation& op) {
  static const std::string empty;
  const Arg* a = op.find("control_qubits");
  return a ? a->str : empty;
}

Status RegisterQubits(const std::string& s,
                      std::set<std::tuple<int, int, std::string>>* ids) {
  if (s.empty()) return Status();
  for (const auto& q : SplitCsv(s)) {
    std::pair<int, int> rc;
    if (!ParseQubitKey(q, &rc)) return Status::Invalid("Unable to parse qubit: " + q);
    ids->emplace(rc.first, rc.second, q);
  }
  return Status();
}

Status ParseArg(const Operation& op, const char* name, const SymbolMap& sym, float* out,
                int* col) {
  const Arg* a = op.find(name);
  if (!a) return Status::Invalid(std::string("Could not find arg: ") + name + " in op.");
  *out = a->f;
  if (col) *col = -1;
  if (!a->symbol.empty()) {
    int c;
    float v;
    if (!sym.find(a->symbol, &c, &v))
      return Status::Invalid("Could not find symbol in parameter map: " + a->symbol);
    *out = v;
    if (col) *col = c;
  }
  return Status();
}

struct GateSpec {
  const char* id;
  GateKind kind;
};

const GateSpec kGateTable[] = {
    {"I", kGateI1},   {"I2", kGateI2},   {"HP", kGateH},    {"XP", kGateX},
    {"YP", kGateY},   {"ZP", kGateZ},    {"XXP", kGateXX},  {"YYP", kGateYY},
    {"ZZP", kGateZZ}, {"CZP", kGateCZ},  {"CNP", kGateCX},  {"SP", kGateSwap},
    {"ISP", kGateISwap}, {"PXP", kGatePhasedX}, {"FSIM", kGateFSim}, {"PISP", kGatePhasedISwap}};

Status ParseGate(const Operation& op, const SymbolMap& sym, const Circuit& circ, Gate* g) {
  const GateSpec* spec = nullptr;
  for (const auto& s : kGateTable)
    if (op.gate_id == s.id) spec = &s;
  if (!spec)
    return Status::Invalid("Could not parse gate id: " + op.gate_id +
                           ". This is likely because a cirq.Channel was used in an op that "
                           "does not support them.");
  g->kind = spec->kind;
  g->nt = GateTargets(spec->kind);
  const int n = circ.num_qubits;
  if (int(op.qubits.size()) != g->nt)
    return Status::Invalid("Gate " + op.gate_id + " expects " + std::to_string(g->nt) +
                           " qubits, got " + std::to_string(op.qubits.size()) + ".");
  uint64_t used = 0;
  for (int j = 0; j < g->nt; ++j) {
    auto it = circ.qubit_index.find(op.qubits[j]);
    if (it == circ.qubit_index.end()) return Status::Invalid("Unable to parse qubit: " + op.qubits[j]);
    g->bits[j] = n - 1 - it->second;
    if (used >> g->bits[j] & 1) return Status::Invalid("Operation acts on duplicate qubits.");
    used |= uint64_t(1) << g->bits[j];
  }
  // Controls (ParseProtoControls, circuit_parser_qsim.cc:84-124).
  const std::string& cq = ControlQubits(op);
  const Arg* cva = op.find("control_values");
  const std::string cv = cva ? cva->str : std::string();
  if (!(cq.empty() && cv.empty())) {
    auto qs = SplitCsv(cq);
    auto vs = SplitCsv(cv);
    if (qs.size() != vs.size())
      return Status::Invalid("Mistmatched number of control qubits and control values.");
    for (size_t k = 0; k < qs.size(); ++k) {
      auto it = circ.qubit_index.find(qs[k]);
      if (it == circ.qubit_index.end()) return Status::Invalid("Unable to parse qubit: " + qs[k]);
      char* end = nullptr;
      long val = strtol(vs[k].c_str(), &end, 10);
      if (vs[k].empty() || *end != '\0')
        return Status::Invalid("Unparseable control value: " + vs[k]);
      const int b = n - 1 - it->second;
      if (used >> b & 1) return Status::Invalid("Operation acts on duplicate qubits.");
      used |= uint64_t(1) << b;
      g->cmask |= uint64_t(1) << b;
      if (val & 1) g->cval |= uint64_t(1) << b;
      g->ncontrols++;
    }
  }
  // Arguments (builders, circuit_parser_qsim.cc:153-562).
  g->nsym = 0;
  for (float& x : g->params) x = 0.0f;
  auto add_sym = [&](int pidx, int col) {
    if (col >= 0) {
      g->sym_pidx[g->nsym] = pidx;
      g->sym_col[g->nsym] = col;
      g->nsym++;
    }
  };
  Status s;
  int c0 = -1, c1 = -1;
  switch (g->kind) {
    case kGateI1:
    case kGateI2:
      break;
    case kGatePhasedX:
      if (!(s = ParseArg(op, "exponent", sym, &g->params[2], &c1)).ok()) return s;
      if (!(s = ParseArg(op, "exponent_scalar", sym, &g->params[3], nullptr)).ok()) return s;
      if (!(s = ParseArg(op, "phase_exponent", sym, &g->params[0], &c0)).ok()) return s;
      if (!(s = ParseArg(op, "phase_exponent_scalar", sym, &g->params[1], nullptr)).ok()) return s;
      if (!(s = ParseArg(op, "global_shift", sym, &g->params[4], nullptr)).ok()) return s;
      add_sym(0, c0);
      add_sym(2, c1);
      break;
    case kGateFSim:
      if (!(s = ParseArg(op, "theta", sym, &g->params[0], &c0)).ok()) return s;
      if (!(s = ParseArg(op, "theta_scalar", sym, &g->params[1], nullptr)).ok()) return s;
      if (!(s = ParseArg(op, "phi", sym, &g->params[2], &c1)).ok()) return s;
      if (!(s = ParseArg(op, "phi_scalar", sym, &g->params[3], nullptr)).ok()) return s;
      add_sym(0, c0);
      add_sym(2, c1);
      break;
    case kGatePhasedISwap:
      if (!(s = ParseArg(op, "exponent", sym, &g->params[2], &c1)).ok()) return s;
      if (!(s = ParseArg(op, "exponent_scalar", sym, &g->params[3], nullptr)).ok()) return s;
      if (!(s = ParseArg(op, "phase_exponent", sym, &g->params[0], &c0)).ok()) return s;
      if (!(s = ParseArg(op, "phase_exponent_scalar", sym, &g->params[1], nullptr)).ok()) return s;
      add_sym(0, c0);
      add_sym(2, c1);
      break;
    default:  // single / two qubit eigen gates
      if (!(s = ParseArg(op, "exponent", sym, &g->params[0], &c0)).ok()) return s;
      if (!(s = ParseArg(op, "exponent_scalar", sym, &g->params[1], nullptr)).ok()) return s;
      if (!(s = ParseArg(op, "global_shift", sym, &g->params[2], nullptr)).ok()) return s;
      add_sym(0, c0);
      break;
  }
  GateUnitary(g->kind, g->params, g->U);
  return Status();
}

Status CollectQubits(const Program& prog, std::vector<std::string>* sorted_ids) {
  std::set<std::tuple<int, int, std::string>> ids;
  for (const auto& m : prog.moments) {
    for (const auto& op : m) {
      for (const auto& q : op.qubits) {
        Status s = RegisterQubits(q, &ids);
        if (!s.ok()) return s;
      }
      Status s = RegisterQubits(ControlQubits(op), &ids);
      if (!s.ok()) return s;
    }
  }
  sorted_ids->clear();
  for (const auto& t : ids) sorted_ids->push_back(std::get<2>(t));
  return Status();
}

}  // namespace

void EigenTableForDevice(double (*out)[100]) {
  for (int kind = 0; kind <= int(kGateISwap); ++kind) {
    const Eigen& e = EigenOf(GateKind(kind));
    double* o = out[kind];
    for (int i = 0; i < 100; ++i) o[i] = 0;
    o[0] = e.count;
    for (int k = 0; k < 3; ++k) o[1 + k] = e.lam[k];
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 16; ++i) {
        o[4 + 32 * k + 2 * i] = e.P[k][i].real();
        o[4 + 32 * k + 2 * i + 1] = e.P[k][i].imag();
      }
  }
}

int GateTargets(GateKind kind) {
  switch (kind) {
    case kGateI1: case kGateH: case kGateX: case kGateY: case kGateZ: case kGatePhasedX:
      return 1;
    default:
      return 2;
  }
}

bool SymbolMap::find(const std::string& name, int* col, float* val) const {
  if (!columns) return false;
  auto it = columns->find(name);
  if (it == columns->end()) return false;
  *col = it->second;
  *val = values[it->second];
  return true;
}

void GateUnitary(GateKind kind, const float* params, cdouble* U) {
  UnitaryOrDeriv(kind, params, -1, U);
}

void GateDerivative(GateKind kind, const float* params, int pidx, bool tfq_fd, cdouble* dU) {
  if (!tfq_fd) {
    UnitaryOrDeriv(kind, params, pidx, dU);
    return;
  }
  // PopulateGradient* (adj_util.cc:175-302): float32 gates at p +- eps,
  // float32 difference, scaled by float(0.5 / eps).
  float pp[5], pm[5];
  for (int i = 0; i < 5; ++i) pp[i] = pm[i] = params[i];
  pp[pidx] = params[pidx] + kGradEps;
  pm[pidx] = params[pidx] - kGradEps;
  cdouble Up[16], Um[16];
  GateUnitary(kind, pp, Up);
  GateUnitary(kind, pm, Um);
  const float scale = float(0.5 / double(kGradEps));
  const int dim = GateTargets(kind) == 1 ? 2 : 4;
  for (int i = 0; i < dim * dim; ++i) {
    float re = float(Up[i].real()) - float(Um[i].real());
    float im = float(Up[i].imag()) - float(Um[i].imag());
    dU[i] = cdouble(double(re * scale), double(im * scale));
  }
}

Status BuildCircuit(const Program& prog, const SymbolMap& sym, Circuit* out) {
  out->gates.clear();
  out->qubit_index.clear();
  out->num_qubits = 0;
  out->no_moments = prog.moments.empty();
  if (out->no_moments) return Status();  // (#679) empty program
  std::vector<std::string> ids;
  Status s = CollectQubits(prog, &ids);
  if (!s.ok()) return s;
  out->num_qubits = int(ids.size());
  for (size_t i = 0; i < ids.size(); ++i) out->qubit_index[ids[i]] = int(i);
  if (out->num_qubits == 0) return Status();
  if (out->num_qubits > 40) return Status::Invalid("Circuits of more than 40 qubits are not supported.");
  for (const auto& m : prog.moments) {
    for (const auto& op : m) {
      Gate g;
      s = ParseGate(op, sym, *out, &g);
      if (!s.ok()) return s;
      out->gates.push_back(g);
    }
  }
  return Status();
}

namespace {

void SetKraus(cdouble (&K)[4], cdouble a, cdouble b, cdouble c, cdouble d) {
  K[0] = a; K[1] = b; K[2] = c; K[3] = d;
}

// Channel builders (circuit_parser_qsim.cc:598-790): the target qubit and the
// float arguments; channel arguments cannot be symbols.
Status ParseChannel(const Operation& op, const Circuit& circ, Channel* ch) {
  static const char* kIds[] = {"DP", "ADP", "GAD", "AD", "RST", "PD", "PF", "BF"};
  int kind = -1;
  for (int i = 0; i < 8; ++i)
    if (op.gate_id == kIds[i]) kind = i;
  if (kind < 0) return Status::Invalid("Could not parse channel id: " + op.gate_id);
  ch->kind = ChannelKind(kind);
  if (op.qubits.size() != 1)
    return Status::Invalid("Channel " + op.gate_id + " expects 1 qubit, got " +
                           std::to_string(op.qubits.size()) + ".");
  auto it = circ.qubit_index.find(op.qubits[0]);
  if (it == circ.qubit_index.end()) return Status::Invalid("Unable to parse qubit: " + op.qubits[0]);
  ch->bit = circ.num_qubits - 1 - it->second;
  SymbolMap none;
  auto arg = [&](const char* name, double* v) {
    float f = 0.0f;
    Status s = ParseArg(op, name, none, &f, nullptr);
    *v = double(f);
    return s;
  };
  const cdouble o(0.0), l(1.0), im(0.0, 1.0);
  double p = 0, g = 0, px = 0, py = 0, pz = 0;
  Status s;
  switch (ch->kind) {
    case kChanDP:
      if (!(s = arg("p", &p)).ok()) return s;
      px = py = pz = p / 3.0;
      ch->mixture = true;
      break;
    case kChanADP:
      if (!(s = arg("p_x", &px)).ok()) return s;
      if (!(s = arg("p_y", &py)).ok()) return s;
      if (!(s = arg("p_z", &pz)).ok()) return s;
      ch->mixture = true;
      break;
    case kChanPF:
      if (!(s = arg("p", &p)).ok()) return s;
      ch->mixture = true;
      ch->nk = 2;
      ch->probs[0] = 1.0 - p;
      ch->probs[1] = p;
      SetKraus(ch->K[0], l, o, o, l);
      SetKraus(ch->K[1], l, o, o, -l);
      return Status();
    case kChanBF:
      if (!(s = arg("p", &p)).ok()) return s;
      ch->mixture = true;
      ch->nk = 2;
      ch->probs[0] = 1.0 - p;
      ch->probs[1] = p;
      SetKraus(ch->K[0], l, o, o, l);
      SetKraus(ch->K[1], o, l, l, o);
      return Status();
    case kChanAD:
      if (!(s = arg("gamma", &g)).ok()) return s;
      ch->nk = 2;
      SetKraus(ch->K[0], l, o, o, std::sqrt(1.0 - g));
      SetKraus(ch->K[1], o, std::sqrt(g), o, o);
      return Status();
    case kChanPD:
      if (!(s = arg("gamma", &g)).ok()) return s;
      ch->nk = 2;
      SetKraus(ch->K[0], l, o, o, std::sqrt(1.0 - g));
      SetKraus(ch->K[1], o, o, o, std::sqrt(g));
      return Status();
    case kChanGAD: {
      if (!(s = arg("p", &p)).ok()) return s;
      if (!(s = arg("gamma", &g)).ok()) return s;
      ch->nk = 4;
      const double a = std::sqrt(p), b = std::sqrt(1.0 - p);
      const double sg = std::sqrt(g), cg = std::sqrt(1.0 - g);
      SetKraus(ch->K[0], a, o, o, a * cg);
      SetKraus(ch->K[1], o, a * sg, o, o);
      SetKraus(ch->K[2], b * cg, o, o, b);
      SetKraus(ch->K[3], o, o, b * sg, o);
      return Status();
    }
    case kChanRST:
      ch->nk = 2;
      SetKraus(ch->K[0], l, o, o, o);
      SetKraus(ch->K[1], o, l, o, o);
      return Status();
  }
  // Depolarising mixtures: I, X, Y, Z.
  ch->nk = 4;
  ch->probs[0] = ch->kind == kChanDP ? 1.0 - p : 1.0 - px - py - pz;
  ch->probs[1] = px;
  ch->probs[2] = py;
  ch->probs[3] = pz;
  SetKraus(ch->K[0], l, o, o, l);
  SetKraus(ch->K[1], o, l, l, o);
  SetKraus(ch->K[2], o, -im, im, o);
  SetKraus(ch->K[3], l, o, o, -l);
  return Status();
}

}  // namespace

Status BuildNoisyCircuit(const Program& prog, const SymbolMap& sym, Circuit* out) {
  out->gates.clear();
  out->channels.clear();
  out->qubit_index.clear();
  out->num_qubits = 0;
  out->no_moments = prog.moments.empty();
  if (out->no_moments) return Status();
  std::vector<std::string> ids;
  Status s = CollectQubits(prog, &ids);
  if (!s.ok()) return s;
  out->num_qubits = int(ids.size());
  for (size_t i = 0; i < ids.size(); ++i) out->qubit_index[ids[i]] = int(i);
  if (out->num_qubits == 0) return Status();
  if (out->num_qubits > 40) return Status::Invalid("Circuits of more than 40 qubits are not supported.");
  for (const auto& m : prog.moments) {
    for (const auto& op : m) {
      Gate g;
      s = ParseGate(op, sym, *out, &g);
      if (s.ok()) {
        out->gates.push_back(g);
        continue;
      }
      // A gate id that parsed and failed keeps its error; an unknown id is
      // tried as a channel.
      if (s.msg.rfind("Could not parse gate id", 0) != 0) return s;
      Channel ch;
      s = ParseChannel(op, *out, &ch);
      if (!s.ok()) return s;
      ch.gate_pos = int(out->gates.size());
      out->channels.push_back(ch);
    }
  }
  return Status();
}

void RebindCircuit(const Circuit& tmpl, const float* values, Circuit* out) {
  *out = tmpl;
  for (auto& g : out->gates) {
    if (g.nsym == 0) continue;
    for (int j = 0; j < g.nsym; ++j) g.params[g.sym_pidx[j]] = values[g.sym_col[j]];
    GateUnitary(g.kind, g.params, g.U);
  }
}

Status BuildPairedCircuit(const Program& prog, const Circuit& reference, Circuit* out) {
  out->gates.clear();
  out->num_qubits = reference.num_qubits;
  out->no_moments = reference.no_moments;
  out->qubit_index = reference.qubit_index;
  std::set<std::string> unvisited;
  for (const auto& kv : reference.qubit_index) unvisited.insert(kv.first);
  for (const auto& m : prog.moments) {
    for (const auto& op : m) {
      std::vector<std::string> qs(op.qubits);
      const std::string& cq = ControlQubits(op);
      if (!cq.empty())
        for (const auto& q : SplitCsv(cq)) qs.push_back(q);
      for (const auto& q : qs) {
        unvisited.erase(q);
        if (!reference.qubit_index.count(q))
          return Status::Invalid("A paired circuit contains qubits not found in reference circuit.");
      }
    }
  }
  if (!unvisited.empty())
    return Status::Invalid("A reference circuit contains qubits not found in paired circuit.");
  SymbolMap none;
  for (const auto& m : prog.moments) {
    for (const auto& op : m) {
      Gate g;
      Status s = ParseGate(op, none, *out, &g);
      if (!s.ok()) {
        if (s.msg.rfind("Could not find symbol", 0) == 0)
          return Status::Invalid("Found symbols in other_programs.No symbols are allowed in these circuits.");
        return s;
      }
      out->gates.push_back(g);
    }
  }
  return Status();
}

Status ResolvePauliSum(const PauliSum& ps, const Circuit& circ, std::vector<Pauli>* out) {
  out->clear();
  // ResolveQubitIds returns before remapping Pauli ids when the program has
  // no moments (program_resolution.cc:93-98); such rows are sentinels anyway.
  if (circ.no_moments) return Status();
  const int n = circ.num_qubits;
  for (const auto& t : ps.terms) {
    Pauli p;
    p.coef = t.coef_real;
    for (const auto& pp : t.paulis) {
      auto it = circ.qubit_index.find(pp.qubit_id);
      if (it == circ.qubit_index.end())
        return Status::Invalid("Found a Pauli sum operating on qubits not found in circuit.");
      const int b = n - 1 - it->second;
      const uint64_t bit = uint64_t(1) << b;
      char c = pp.pauli_type.size() == 1 ? pp.pauli_type[0] : '?';
      if (c == 'X') p.xmask |= bit;
      else if (c == 'Z') p.zmask |= bit;
      else if (c == 'Y') { p.xmask |= bit; p.zmask |= bit; p.ny++; }
      else
        return Status::Invalid("Could not parse gate id: " + pp.pauli_type +
                               "P. This is likely because a cirq.Channel was used in an op "
                               "that does not support them.");
      p.ops.emplace_back(b, c);
    }
    out->push_back(std::move(p));
  }
  return Status();
}

}  // namespace tfqx
