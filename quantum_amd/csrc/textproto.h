// Protobuf text format -> binary wire format for TFQ's Program and PauliSum
// (the fallback of ParseProto, core/ops/parse_context.cc:41-56).
#pragma once

#include <string>

namespace tfqx {

// Return false when the text is not a valid text-format message of the type.
bool TextProgramToBinary(const std::string& text, std::string* out);
bool TextPauliSumToBinary(const std::string& text, std::string* out);

}  // namespace tfqx
