// proto3 decoding of TFQ's Program and PauliSum messages.
//
// Schema: core/proto/program.proto:20-184, core/proto/pauli_sum.proto:20-35.
// Replaces ParseProto / ParsePrograms / GetPauliSums
// (core/ops/parse_context.cc:41-56,60-92,253-289) without protobuf: only the
// fields the simulation path reads are kept.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

namespace tfqx {

struct Arg {
  bool has_value = false;   // arg_value set (oneof)
  bool has_symbol = false;  // symbol set (oneof)
  bool is_str = false;      // arg_value holds string_value
  float f = 0.0f;           // arg_value.float_value (or double_value narrowed)
  std::string str;          // arg_value.string_value
  std::string symbol;
};

struct Operation {
  std::string gate_id;
  std::vector<std::pair<std::string, Arg>> args;
  std::vector<std::string> qubits;
  const Arg* find(const char* name) const;
};

struct Program {
  std::vector<std::vector<Operation>> moments;
};

struct PauliPair {
  std::string qubit_id;
  std::string pauli_type;
};

struct PauliTerm {
  float coef_real = 0.0f;
  float coef_imag = 0.0f;
  std::vector<PauliPair> paulis;
};

struct PauliSum {
  std::vector<PauliTerm> terms;
};

// Return false on malformed input.  The binary wire format is tried first,
// then the protobuf text format (textproto.h), as TFQ's ParseProto does.
bool DecodeProgram(const uint8_t* data, size_t len, Program* out);
bool DecodeProgramBinary(const uint8_t* data, size_t len, Program* out);
bool DecodePauliSumBinary(const uint8_t* data, size_t len, PauliSum* out);
// Program -> proto3 bytes: language.gate_set "tfq_gate_set", scheduling
// MOMENT_BY_MOMENT, every kept field (float values as float_value).
std::string EncodeProgram(const Program& p);
// The empty program of TFQ's padding (gate set and an empty circuit).
std::string EncodeEmptyProgram();
bool DecodePauliSum(const uint8_t* data, size_t len, PauliSum* out);

}  // namespace tfqx
