// Protobuf text format for TFQ's Program and PauliSum messages.
//
// ParseProto (core/ops/parse_context.cc:41-56) first parses the binary wire
// format and, when that fails, the human-readable text format.  This file
// restates the second step for the two message types: the text is parsed
// against the schema of core/proto/program.proto:20-184 and
// core/proto/pauli_sum.proto:20-35 and re-encoded as binary, which the
// wire decoder then reads.  Accepted syntax: `name: value`, `name { ... }` or
// `name < ... >` (the colon is optional before a message), `[a, b]` lists,
// `#` comments, ',' / ';' separators, quoted strings with C escapes (adjacent
// strings concatenate), numbers, `true` / `false`, enum names.  Unknown field
// names fail the parse, as TextFormat::ParseFromString does.
#include "textproto.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <string>
#include <vector>

namespace tfqx {
namespace {

enum Kind { kMsg, kString, kFloat, kDouble, kBool, kEnum, kInt64 };

struct Schema;
struct FieldDef {
  const char* name;
  int num;
  Kind kind;
  const Schema* sub;  // messages
};
struct Schema {
  std::vector<FieldDef> fields;
  const FieldDef* find(const std::string& n) const {
    for (const auto& f : fields)
      if (n == f.name) return &f;
    return nullptr;
  }
};

// program.proto
extern const Schema kArg;
const Schema kBools{{{"values", 1, kBool, nullptr}}};
const Schema kArgValue{{{"float_value", 1, kFloat, nullptr}, {"bool_values", 2, kMsg, &kBools},
                        {"string_value", 3, kString, nullptr}, {"double_value", 4, kDouble, nullptr}}};
const Schema kArgFunction{{{"type", 1, kString, nullptr}, {"args", 2, kMsg, &kArg}}};
const Schema kArg{{{"arg_value", 1, kMsg, &kArgValue}, {"symbol", 2, kString, nullptr},
                   {"func", 3, kMsg, &kArgFunction}}};
const Schema kArgsEntry{{{"key", 1, kString, nullptr}, {"value", 2, kMsg, &kArg}}};
const Schema kGate{{{"id", 1, kString, nullptr}}};
const Schema kQubit{{{"id", 2, kString, nullptr}}};
const Schema kOperation{{{"gate", 1, kMsg, &kGate}, {"args", 2, kMsg, &kArgsEntry},
                         {"qubits", 3, kMsg, &kQubit}}};
const Schema kMoment{{{"operations", 1, kMsg, &kOperation}}};
const Schema kCircuit{{{"scheduling_strategy", 1, kEnum, nullptr}, {"moments", 2, kMsg, &kMoment}}};
const Schema kScheduled{{{"operation", 1, kMsg, &kOperation}, {"start_time_picos", 2, kInt64, nullptr}}};
const Schema kSchedule{{{"scheduled_operations", 3, kMsg, &kScheduled}}};
const Schema kLanguage{{{"gate_set", 1, kString, nullptr},
                        {"arg_function_language", 2, kString, nullptr}}};
const Schema kProgram{{{"language", 1, kMsg, &kLanguage}, {"circuit", 2, kMsg, &kCircuit},
                       {"schedule", 3, kMsg, &kSchedule}}};
// pauli_sum.proto
const Schema kPair{{{"qubit_id", 1, kString, nullptr}, {"pauli_type", 2, kString, nullptr}}};
const Schema kTerm{{{"coefficient_real", 1, kFloat, nullptr}, {"coefficient_imag", 2, kFloat, nullptr},
                    {"paulis", 3, kMsg, &kPair}}};
const Schema kPauliSum{{{"terms", 1, kMsg, &kTerm}}};

int EnumValue(const std::string& s, bool* ok) {
  *ok = true;
  if (s == "SCHEDULING_STRATEGY_UNSPECIFIED") return 0;
  if (s == "MOMENT_BY_MOMENT") return 1;
  char* end = nullptr;
  const long v = strtol(s.c_str(), &end, 10);
  if (!s.empty() && *end == '\0') return int(v);
  *ok = false;
  return 0;
}

void PutVarint(std::string* o, uint64_t v) {
  while (v >= 0x80) {
    o->push_back(char(uint8_t(v) | 0x80));
    v >>= 7;
  }
  o->push_back(char(uint8_t(v)));
}
void PutKey(std::string* o, int num, int wire) { PutVarint(o, (uint64_t(num) << 3) | uint64_t(wire)); }

struct Parser {
  const std::string& s;
  size_t i = 0;
  bool ok = true;

  void skip() {
    while (i < s.size()) {
      const char c = s[i];
      if (c == '#') {
        while (i < s.size() && s[i] != '\n') ++i;
      } else if (isspace(static_cast<unsigned char>(c))) {
        ++i;
      } else {
        break;
      }
    }
  }
  bool eat(char c) {
    skip();
    if (i < s.size() && s[i] == c) {
      ++i;
      return true;
    }
    return false;
  }
  bool ident(std::string* out) {
    skip();
    const size_t b = i;
    while (i < s.size() && (isalnum(static_cast<unsigned char>(s[i])) || s[i] == '_')) ++i;
    out->assign(s, b, i - b);
    return i > b;
  }
  // One quoted string (with escapes); adjacent ones concatenate.
  bool str(std::string* out) {
    out->clear();
    skip();
    if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) return false;
    while (i < s.size() && (s[i] == '"' || s[i] == '\'')) {
      const char q = s[i++];
      while (true) {
        if (i >= s.size()) return false;
        char c = s[i++];
        if (c == q) break;
        if (c == '\n') return false;
        if (c != '\\') {
          out->push_back(c);
          continue;
        }
        if (i >= s.size()) return false;
        c = s[i++];
        switch (c) {
          case 'n': out->push_back('\n'); break;
          case 't': out->push_back('\t'); break;
          case 'r': out->push_back('\r'); break;
          case 'a': out->push_back('\a'); break;
          case 'b': out->push_back('\b'); break;
          case 'f': out->push_back('\f'); break;
          case 'v': out->push_back('\v'); break;
          case 'x': {
            int v = 0, n = 0;
            while (n < 2 && i < s.size() && isxdigit(static_cast<unsigned char>(s[i]))) {
              v = v * 16 + (isdigit(static_cast<unsigned char>(s[i])) ? s[i] - '0' : (tolower(s[i]) - 'a' + 10));
              ++i;
              ++n;
            }
            if (n == 0) return false;
            out->push_back(char(v));
            break;
          }
          default:
            if (c >= '0' && c <= '7') {
              int v = c - '0', n = 1;
              while (n < 3 && i < s.size() && s[i] >= '0' && s[i] <= '7') v = v * 8 + (s[i++] - '0'), ++n;
              out->push_back(char(v));
            } else {
              out->push_back(c);  // \\ \" \' \?
            }
        }
      }
      skip();
    }
    return true;
  }
  // A number or identifier token (for scalars).
  bool scalar_token(std::string* out) {
    skip();
    const size_t b = i;
    while (i < s.size() && (isalnum(static_cast<unsigned char>(s[i])) || s[i] == '_' || s[i] == '.' ||
                            s[i] == '-' || s[i] == '+'))
      ++i;
    out->assign(s, b, i - b);
    return i > b;
  }

  bool number(const std::string& t, double* v) {
    std::string u = t;
    // Float suffixes (1.5f) are allowed by the text format.
    if (!u.empty() && (u.back() == 'f' || u.back() == 'F') && u.find_first_of("xX") == std::string::npos &&
        u != "inf" && u != "-inf")
      u.pop_back();
    if (u == "inf" || u == "infinity" || u == "Infinity") {
      *v = INFINITY;
      return true;
    }
    if (u == "-inf" || u == "-infinity" || u == "-Infinity") {
      *v = -INFINITY;
      return true;
    }
    if (u == "nan" || u == "NaN") {
      *v = NAN;
      return true;
    }
    char* end = nullptr;
    *v = strtod(u.c_str(), &end);
    return !u.empty() && *end == '\0';
  }

  bool scalar(const FieldDef& f, std::string* out) {
    if (f.kind == kString) {
      std::string v;
      if (!str(&v)) return false;
      PutKey(out, f.num, 2);
      PutVarint(out, v.size());
      out->append(v);
      return true;
    }
    std::string t;
    if (!scalar_token(&t)) return false;
    switch (f.kind) {
      case kFloat: {
        double d;
        if (!number(t, &d)) return false;
        const float x = float(d);
        PutKey(out, f.num, 5);
        uint32_t u;
        memcpy(&u, &x, 4);
        for (int k = 0; k < 4; ++k) out->push_back(char(uint8_t(u >> (8 * k))));
        return true;
      }
      case kDouble: {
        double d;
        if (!number(t, &d)) return false;
        PutKey(out, f.num, 1);
        uint64_t u;
        memcpy(&u, &d, 8);
        for (int k = 0; k < 8; ++k) out->push_back(char(uint8_t(u >> (8 * k))));
        return true;
      }
      case kBool: {
        int b;
        if (t == "true" || t == "True" || t == "t" || t == "1") b = 1;
        else if (t == "false" || t == "False" || t == "f" || t == "0") b = 0;
        else return false;
        PutKey(out, f.num, 0);
        PutVarint(out, uint64_t(b));
        return true;
      }
      case kEnum: {
        bool good;
        const int v = EnumValue(t, &good);
        if (!good) return false;
        PutKey(out, f.num, 0);
        PutVarint(out, uint64_t(int64_t(v)));
        return true;
      }
      case kInt64: {
        char* end = nullptr;
        const long long v = strtoll(t.c_str(), &end, 0);
        if (t.empty() || *end != '\0') return false;
        PutKey(out, f.num, 0);
        PutVarint(out, uint64_t(v));
        return true;
      }
      default:
        return false;
    }
  }

  // Fields until `close` (or the end of the input when close == 0).
  bool message(const Schema& sc, char close, std::string* out) {
    while (true) {
      skip();
      if (i >= s.size()) return close == 0;
      if (close && s[i] == close) {
        ++i;
        return true;
      }
      std::string name;
      if (!ident(&name)) return false;
      const FieldDef* f = sc.find(name);
      if (!f) return false;
      const bool colon = eat(':');
      if (f->kind == kMsg) {
        auto sub_message = [&]() {
          char c;
          if (eat('{')) c = '}';
          else if (eat('<')) c = '>';
          else return false;
          std::string body;
          if (!message(*f->sub, c, &body)) return false;
          PutKey(out, f->num, 2);
          PutVarint(out, body.size());
          out->append(body);
          return true;
        };
        (void)colon;  // optional before a message or a list of messages
        if (eat('[')) {
          if (!eat(']')) {
            do {
              if (!sub_message()) return false;
            } while (eat(','));
            if (!eat(']')) return false;
          }
        } else if (!sub_message()) {
          return false;
        }
      } else {
        if (!colon) return false;
        if (eat('[')) {
          if (!eat(']')) {
            do {
              if (!scalar(*f, out)) return false;
            } while (eat(','));
            if (!eat(']')) return false;
          }
        } else if (!scalar(*f, out)) {
          return false;
        }
      }
      if (!eat(',')) eat(';');
    }
  }
};

bool Convert(const Schema& sc, const std::string& text, std::string* out) {
  out->clear();
  Parser p{text};
  return p.message(sc, 0, out);
}

}  // namespace

bool TextProgramToBinary(const std::string& text, std::string* out) {
  return Convert(kProgram, text, out);
}

bool TextPauliSumToBinary(const std::string& text, std::string* out) {
  return Convert(kPauliSum, text, out);
}

}  // namespace tfqx
