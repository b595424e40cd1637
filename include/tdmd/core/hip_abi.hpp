// hip_abi.hpp -- glue between the reference's C++ types and the C-ABI:
// Box -> tdmd_box, the build precision, stream handles.
#pragma once

#include <hip/hip_runtime_api.h>

#include "box.hpp"
#include "error.hpp"

namespace tdmd::hip {

inline tdmd_box to_c(const Box& b) {
  tdmd_box c{};
  c.lo[0] = b.lo.x;
  c.lo[1] = b.lo.y;
  c.lo[2] = b.lo.z;
  c.hi[0] = b.hi.x;
  c.hi[1] = b.hi.y;
  c.hi[2] = b.hi.z;
  for (int d = 0; d < 3; ++d) c.periodic[d] = b.periodic[static_cast<std::size_t>(d)] ? 1 : 0;
  return c;
}

// `real` = float (TDMD_PRECISION_MIXED) or double (TDMD_PRECISION_FP64).
constexpr int32_t kPrecision = sizeof(real) == 4 ? TDMD_MIXED : TDMD_FP64;

inline void* as_void(hipStream_t s) { return static_cast<void*>(s); }

}  // namespace tdmd::hip
