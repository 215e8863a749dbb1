// proto3 wire decoding for Program / PauliSum (see wire.h).
#include "wire.h"

#include <string.h>

#include "textproto.h"

namespace tfqx {
namespace {

struct Reader {
  const uint8_t* p;
  const uint8_t* end;
  bool ok = true;

  bool done() const { return p >= end; }

  uint64_t varint() {
    uint64_t v = 0;
    int shift = 0;
    while (true) {
      if (p >= end || shift > 63) {
        ok = false;
        return 0;
      }
      uint8_t c = *p++;
      v |= uint64_t(c & 0x7f) << shift;
      if (!(c & 0x80)) return v;
      shift += 7;
    }
  }

  // Reads one field header and its payload.  For length-delimited fields the
  // payload is returned as [sub, sub_end); for fixed32 the raw bits in v32.
  bool field(uint32_t* num, uint32_t* wire, const uint8_t** sub, size_t* sub_len,
             uint64_t* v) {
    uint64_t key = varint();
    if (!ok) return false;
    if ((key >> 3) == 0) return ok = false;  // field number 0 is invalid
    *num = uint32_t(key >> 3);
    *wire = uint32_t(key & 7);
    switch (*wire) {
      case 0:
        *v = varint();
        return ok;
      case 1:
        if (end - p < 8) return ok = false;
        memcpy(v, p, 8);
        p += 8;
        return true;
      case 2: {
        uint64_t n = varint();
        if (!ok || uint64_t(end - p) < n) return ok = false;
        *sub = p;
        *sub_len = size_t(n);
        p += n;
        return true;
      }
      case 5: {
        if (end - p < 4) return ok = false;
        uint32_t x;
        memcpy(&x, p, 4);
        *v = x;
        p += 4;
        return true;
      }
      default:
        return ok = false;
    }
  }
};

#define FOR_FIELDS(buf, len)                                  \
  Reader rd{(buf), (buf) + (len)};                            \
  uint32_t num, wt;                                           \
  const uint8_t* sub = nullptr;                               \
  size_t sl = 0;                                              \
  uint64_t v = 0;                                             \
  while (!rd.done() && rd.field(&num, &wt, &sub, &sl, &v))

float AsFloat(uint64_t v) {
  uint32_t u = uint32_t(v);
  float f;
  memcpy(&f, &u, 4);
  return f;
}

bool DecodeArgValue(const uint8_t* b, size_t n, Arg* a) {
  FOR_FIELDS(b, n) {
    if (num == 1 && wt == 5) a->f = AsFloat(v);
    else if (num == 3 && wt == 2) {
      a->str.assign(reinterpret_cast<const char*>(sub), sl);
      a->is_str = true;
    }
    else if (num == 4 && wt == 1) {
      double d;
      memcpy(&d, &v, 8);
      a->f = float(d);
    }
  }
  return rd.ok;
}

bool DecodeArg(const uint8_t* b, size_t n, Arg* a) {
  FOR_FIELDS(b, n) {
    if (num == 1 && wt == 2) {
      a->has_value = true;
      if (!DecodeArgValue(sub, sl, a)) return false;
    } else if (num == 2 && wt == 2) {
      a->has_symbol = true;
      a->symbol.assign(reinterpret_cast<const char*>(sub), sl);
    }
  }
  return rd.ok;
}

bool DecodeOperation(const uint8_t* b, size_t n, Operation* op) {
  FOR_FIELDS(b, n) {
    if (num == 1 && wt == 2) {  // Gate
      Reader r2{sub, sub + sl};
      uint32_t n2, w2;
      const uint8_t* s2 = nullptr;
      size_t l2 = 0;
      uint64_t v2 = 0;
      while (!r2.done() && r2.field(&n2, &w2, &s2, &l2, &v2)) {
        if (n2 == 1 && w2 == 2) op->gate_id.assign(reinterpret_cast<const char*>(s2), l2);
      }
      if (!r2.ok) return false;
    } else if (num == 2 && wt == 2) {  // map<string, Arg> entry
      std::string key;
      Arg arg;
      Reader r2{sub, sub + sl};
      uint32_t n2, w2;
      const uint8_t* s2 = nullptr;
      size_t l2 = 0;
      uint64_t v2 = 0;
      while (!r2.done() && r2.field(&n2, &w2, &s2, &l2, &v2)) {
        if (n2 == 1 && w2 == 2) key.assign(reinterpret_cast<const char*>(s2), l2);
        else if (n2 == 2 && w2 == 2) {
          if (!DecodeArg(s2, l2, &arg)) return false;
        }
      }
      if (!r2.ok) return false;
      // proto map semantics: the last entry for a key wins.
      bool replaced = false;
      for (auto& kv : op->args) {
        if (kv.first == key) {
          kv.second = arg;
          replaced = true;
        }
      }
      if (!replaced) op->args.emplace_back(std::move(key), std::move(arg));
    } else if (num == 3 && wt == 2) {  // Qubit
      std::string id;
      Reader r2{sub, sub + sl};
      uint32_t n2, w2;
      const uint8_t* s2 = nullptr;
      size_t l2 = 0;
      uint64_t v2 = 0;
      while (!r2.done() && r2.field(&n2, &w2, &s2, &l2, &v2)) {
        if (n2 == 2 && w2 == 2) id.assign(reinterpret_cast<const char*>(s2), l2);
      }
      if (!r2.ok) return false;
      op->qubits.push_back(std::move(id));
    }
  }
  return rd.ok;
}

}  // namespace

const Arg* Operation::find(const char* name) const {
  for (const auto& kv : args)
    if (kv.first == name) return &kv.second;
  return nullptr;
}

bool DecodeProgramBinary(const uint8_t* data, size_t len, Program* out) {
  out->moments.clear();
  FOR_FIELDS(data, len) {
    if (num == 2 && wt == 2) {  // Circuit
      Reader r2{sub, sub + sl};
      uint32_t n2, w2;
      const uint8_t* s2 = nullptr;
      size_t l2 = 0;
      uint64_t v2 = 0;
      while (!r2.done() && r2.field(&n2, &w2, &s2, &l2, &v2)) {
        if (n2 == 2 && w2 == 2) {  // Moment
          out->moments.emplace_back();
          auto& ops = out->moments.back();
          Reader r3{s2, s2 + l2};
          uint32_t n3, w3;
          const uint8_t* s3 = nullptr;
          size_t l3 = 0;
          uint64_t v3 = 0;
          while (!r3.done() && r3.field(&n3, &w3, &s3, &l3, &v3)) {
            if (n3 == 1 && w3 == 2) {
              ops.emplace_back();
              if (!DecodeOperation(s3, l3, &ops.back())) return false;
            }
          }
          if (!r3.ok) return false;
        }
      }
      if (!r2.ok) return false;
    }
  }
  return rd.ok;
}

namespace {

void PutVarint(std::string* o, uint64_t v) {
  while (v >= 0x80) {
    o->push_back(char(uint8_t(v) | 0x80));
    v >>= 7;
  }
  o->push_back(char(uint8_t(v)));
}

void PutLen(std::string* o, uint32_t field, const std::string& payload) {
  PutVarint(o, (uint64_t(field) << 3) | 2);
  PutVarint(o, payload.size());
  o->append(payload);
}

std::string EncodeArg(const Arg& a) {
  std::string out;
  if (a.has_symbol) {
    PutLen(&out, 2, a.symbol);
  } else if (a.has_value) {
    std::string v;
    if (a.is_str) {
      PutLen(&v, 3, a.str);
    } else {
      PutVarint(&v, (1u << 3) | 5);
      uint32_t u;
      memcpy(&u, &a.f, 4);
      for (int i = 0; i < 4; ++i) v.push_back(char(uint8_t(u >> (8 * i))));
    }
    PutLen(&out, 1, v);
  }
  return out;
}

std::string EncodeOperation(const Operation& op) {
  std::string out, gate;
  PutLen(&gate, 1, op.gate_id);
  PutLen(&out, 1, gate);
  for (const auto& kv : op.args) {
    std::string entry;
    PutLen(&entry, 1, kv.first);
    PutLen(&entry, 2, EncodeArg(kv.second));
    PutLen(&out, 2, entry);
  }
  for (const auto& q : op.qubits) {
    std::string qb;
    PutLen(&qb, 2, q);
    PutLen(&out, 3, qb);
  }
  return out;
}

std::string Language() {
  std::string lang;
  PutLen(&lang, 1, "tfq_gate_set");
  return lang;
}

}  // namespace

std::string EncodeProgram(const Program& p) {
  std::string circ;
  PutVarint(&circ, (1u << 3) | 0);  // scheduling_strategy = MOMENT_BY_MOMENT
  PutVarint(&circ, 1);
  for (const auto& m : p.moments) {
    std::string mom;
    for (const auto& op : m) PutLen(&mom, 1, EncodeOperation(op));
    PutLen(&circ, 2, mom);
  }
  std::string out;
  PutLen(&out, 1, Language());
  PutLen(&out, 2, circ);
  return out;
}

std::string EncodeEmptyProgram() {
  std::string out;
  PutLen(&out, 1, Language());
  PutLen(&out, 2, std::string());
  return out;
}

bool DecodePauliSumBinary(const uint8_t* data, size_t len, PauliSum* out) {
  out->terms.clear();
  FOR_FIELDS(data, len) {
    if (num == 1 && wt == 2) {
      PauliTerm t;
      Reader r2{sub, sub + sl};
      uint32_t n2, w2;
      const uint8_t* s2 = nullptr;
      size_t l2 = 0;
      uint64_t v2 = 0;
      while (!r2.done() && r2.field(&n2, &w2, &s2, &l2, &v2)) {
        if (n2 == 1 && w2 == 5) t.coef_real = AsFloat(v2);
        else if (n2 == 2 && w2 == 5) t.coef_imag = AsFloat(v2);
        else if (n2 == 3 && w2 == 2) {
          PauliPair pp;
          Reader r3{s2, s2 + l2};
          uint32_t n3, w3;
          const uint8_t* s3 = nullptr;
          size_t l3 = 0;
          uint64_t v3 = 0;
          while (!r3.done() && r3.field(&n3, &w3, &s3, &l3, &v3)) {
            if (n3 == 1 && w3 == 2) pp.qubit_id.assign(reinterpret_cast<const char*>(s3), l3);
            else if (n3 == 2 && w3 == 2) pp.pauli_type.assign(reinterpret_cast<const char*>(s3), l3);
          }
          if (!r3.ok) return false;
          t.paulis.push_back(std::move(pp));
        }
      }
      if (!r2.ok) return false;
      out->terms.push_back(std::move(t));
    }
  }
  return rd.ok;
}

}  // namespace tfqx

namespace tfqx {

// ParseProto (core/ops/parse_context.cc:41-56): binary first, then text.
bool DecodeProgram(const uint8_t* data, size_t len, Program* out) {
  if (DecodeProgramBinary(data, len, out)) return true;
  std::string bin;
  if (!TextProgramToBinary(std::string(reinterpret_cast<const char*>(data), len), &bin)) return false;
  return DecodeProgramBinary(reinterpret_cast<const uint8_t*>(bin.data()), bin.size(), out);
}

bool DecodePauliSum(const uint8_t* data, size_t len, PauliSum* out) {
  if (DecodePauliSumBinary(data, len, out)) return true;
  std::string bin;
  if (!TextPauliSumToBinary(std::string(reinterpret_cast<const char*>(data), len), &bin)) return false;
  return DecodePauliSumBinary(reinterpret_cast<const uint8_t*>(bin.data()), bin.size(), out);
}

}  // namespace tfqx
