// types.hpp -- scalar and vector types of the TDMD step path
// (src/core/types.hpp:11-79): same names, same layouts.
//
// Exactly one of TDMD_PRECISION_MIXED / TDMD_PRECISION_FP64 selects `real`
// and ForceVec, as in the reference build. Positions and velocities are
// double in both modes; reductions are double.
#pragma once

#include <array>
#include <cstdint>

namespace tdmd {

#if defined(TDMD_PRECISION_FP64)
using real = double;
#elif defined(TDMD_PRECISION_MIXED)
using real = float;
#else
#error "define exactly one of TDMD_PRECISION_MIXED or TDMD_PRECISION_FP64"
#endif

using i32 = std::int32_t;
using i64 = std::int64_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;

template <class T>
struct Vec3T {
  T x;
  T y;
  T z;
};

using Vec3D = Vec3T<double>;
using Vec3F = Vec3T<float>;
using Vec3 = Vec3T<real>;

using pos_t = double;
using vel_t = double;
using force_t = real;
using accum_t = double;
using PositionVec = Vec3D;
using VelocityVec = Vec3D;
using ForceVec = Vec3T<force_t>;

static_assert(sizeof(Vec3D) == 24 && sizeof(Vec3F) == 12, "Vec3 must be packed AoS");

// Axis-aligned box (src/core/types.hpp:55-58): a zone's slab.
struct Aabb {
  Vec3D lo;
  Vec3D hi;
};

}  // namespace tdmd
