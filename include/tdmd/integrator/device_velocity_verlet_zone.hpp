// device_velocity_verlet_zone.hpp -- per-zone integration
// (src/integrator/device_velocity_verlet_zone.cuh:12-26): half kick, drift
// (single-image wrap) and force zeroing for atoms
// [first_atom, first_atom + atom_count). FP64 arithmetic.
#pragma once

#include "../core/hip_abi.hpp"

namespace tdmd::integrator {

inline void device_half_kick_zone(VelocityVec* d_velocities, const ForceVec* d_forces,
                                  const i32* d_types, const real* d_masses, i32 first_atom,
                                  i32 atom_count, real dt, hipStream_t stream = nullptr) {
  TDMD_HIP_CALL(tdmd_half_kick_zone(hip::kPrecision, reinterpret_cast<double*>(d_velocities),
                                    d_forces, d_types, d_masses, first_atom, atom_count,
                                    static_cast<double>(dt), hip::as_void(stream)));
}

inline void device_drift_zone(PositionVec* d_positions, const VelocityVec* d_velocities,
                              i32 first_atom, i32 atom_count, real dt, const Box& box,
                              hipStream_t stream = nullptr) {
  const tdmd_box b = hip::to_c(box);
  TDMD_HIP_CALL(tdmd_drift_zone(reinterpret_cast<double*>(d_positions),
                                reinterpret_cast<const double*>(d_velocities), first_atom,
                                atom_count, static_cast<double>(dt), &b, hip::as_void(stream)));
}

inline void device_zero_forces_zone(ForceVec* d_forces, i32 first_atom, i32 atom_count,
                                    hipStream_t stream = nullptr) {
  TDMD_HIP_CALL(
      tdmd_zero_forces_zone(hip::kPrecision, d_forces, first_atom, atom_count, hip::as_void(stream)));
}

}  // namespace tdmd::integrator
