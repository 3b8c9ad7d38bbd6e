// common.hpp -- error reporting shared by the ABI translation units.
#pragma once

#include <string>

namespace hlbm {

// Message of the calling thread's last failure (returned by hlbm_last_error).
inline std::string& last_error() {
  thread_local std::string e;
  return e;
}

inline int fail(int code, const std::string& msg) {
  last_error() = msg;
  return code;
}

}  // namespace hlbm
