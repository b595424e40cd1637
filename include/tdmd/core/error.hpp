// error.hpp -- error conventions of the reference (src/core/error.hpp:21-35,
// src/core/device_buffer.cuh:16-24) mapped onto the C-ABI status codes:
//   * bad input, bad files, capacity overflow, non-finite state, calls out
//     of order -> throw tdmd::Error (the reference throws or asserts);
//   * HIP runtime failure or device memory exhaustion -> print and abort
//     (the reference's TDMD_CUDA_CHECK).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "../../tdmd_hip.h"

namespace tdmd {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const noexcept { return status_; }

 private:
  int status_;
};

namespace hip {

inline void check(int status, const char* call) {
  if (status == TDMD_OK) return;
  const std::string msg = std::string(call) + ": " + tdmd_last_error();
  if (status == TDMD_E_HIP || status == TDMD_E_NOMEM) {
    std::fprintf(stderr, "tdmd: device failure: %s\n", msg.c_str());
    std::abort();
  }
  throw Error(status, msg);
}

}  // namespace hip
}  // namespace tdmd

#define TDMD_HIP_CALL(expr) ::tdmd::hip::check((expr), #expr)
