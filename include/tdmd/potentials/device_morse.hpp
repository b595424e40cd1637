// device_morse.hpp -- MorseParams + compute_morse_gpu
// (src/potentials/device_morse.cuh:14-41). Forces are stored (the reference
// code stores, device_morse.cu:110-117); *d_energy is added to.
#pragma once

#include "../core/hip_abi.hpp"

namespace tdmd::potentials {

struct MorseParams {
  real d;
  real alpha;
  real r0;
  real rc;
  real rc_sq;
};

inline tdmd_morse_params to_c(const MorseParams& p) {
  return tdmd_morse_params{static_cast<double>(p.d), static_cast<double>(p.alpha),
                           static_cast<double>(p.r0), static_cast<double>(p.rc),
                           static_cast<double>(p.rc_sq)};
}

inline void compute_morse_gpu(const PositionVec* d_positions, ForceVec* d_forces,
                              const i32* d_neighbors, const i32* d_offsets, const i32* d_counts,
                              i32 natoms, const Box& box, const MorseParams& params,
                              accum_t* d_energy, hipStream_t stream = nullptr) {
  const tdmd_box b = hip::to_c(box);
  const tdmd_morse_params p = to_c(params);
  TDMD_HIP_CALL(tdmd_morse_force(hip::kPrecision, reinterpret_cast<const double*>(d_positions),
                                 d_forces, d_neighbors, d_offsets, d_counts, natoms, &b, &p,
                                 d_energy, nullptr, hip::as_void(stream)));
}

}  // namespace tdmd::potentials
