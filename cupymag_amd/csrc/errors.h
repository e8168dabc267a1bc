// Error reporting shared by every C-ABI entry point: each call returns an int
// status and leaves a message in a thread-local buffer (cpm_last_error).
#pragma once
#include <string>

#include "cupymag_hip.h"

namespace cpm {

void set_error(const std::string& msg);
inline int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

}  // namespace cpm
