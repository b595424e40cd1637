// device_velocity_verlet.hpp -- velocity-Verlet kernels
// (src/integrator/device_velocity_verlet.cuh:13-59). FP64 arithmetic;
// masses indexed by the 1-based type; single-image wrap on periodic axes.
#pragma once

#include "../core/hip_abi.hpp"

namespace tdmd::integrator {

inline void device_half_kick(VelocityVec* d_velocities, const ForceVec* d_forces,
                             const i32* d_types, const real* d_masses, i32 natoms, real dt,
                             hipStream_t stream = nullptr) {
  TDMD_HIP_CALL(tdmd_half_kick(hip::kPrecision, reinterpret_cast<double*>(d_velocities),
                               d_forces, d_types, d_masses, natoms, static_cast<double>(dt),
                               hip::as_void(stream)));
}

inline void device_drift(PositionVec* d_positions, const VelocityVec* d_velocities, i32 natoms,
                         real dt, const Box& box, hipStream_t stream = nullptr) {
  const tdmd_box b = hip::to_c(box);
  TDMD_HIP_CALL(tdmd_drift(reinterpret_cast<double*>(d_positions),
                           reinterpret_cast<const double*>(d_velocities), natoms,
                           static_cast<double>(dt), &b, hip::as_void(stream)));
}

inline void device_fused_half_kick_drift(VelocityVec* d_velocities, PositionVec* d_positions,
                                         const ForceVec* d_forces, const i32* d_types,
                                         const real* d_masses, i32 natoms, real dt,
                                         const Box& box, hipStream_t stream = nullptr) {
  const tdmd_box b = hip::to_c(box);
  TDMD_HIP_CALL(tdmd_fused_kick_drift(hip::kPrecision, 0, reinterpret_cast<double*>(d_velocities),
                                      reinterpret_cast<double*>(d_positions), d_forces, d_types,
                                      d_masses, natoms, static_cast<double>(dt), &b,
                                      hip::as_void(stream)));
}

inline void device_fused_full_kick_drift(VelocityVec* d_velocities, PositionVec* d_positions,
                                         const ForceVec* d_forces, const i32* d_types,
                                         const real* d_masses, i32 natoms, real dt,
                                         const Box& box, hipStream_t stream = nullptr) {
  const tdmd_box b = hip::to_c(box);
  TDMD_HIP_CALL(tdmd_fused_kick_drift(hip::kPrecision, 1, reinterpret_cast<double*>(d_velocities),
                                      reinterpret_cast<double*>(d_positions), d_forces, d_types,
                                      d_masses, natoms, static_cast<double>(dt), &b,
                                      hip::as_void(stream)));
}

inline void device_zero_forces(ForceVec* d_forces, i32 natoms, hipStream_t stream = nullptr) {
  TDMD_HIP_CALL(tdmd_zero_forces(hip::kPrecision, d_forces, natoms, hip::as_void(stream)));
}

}  // namespace tdmd::integrator
