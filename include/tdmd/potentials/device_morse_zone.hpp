// device_morse_zone.hpp -- compute_morse_gpu_zone
// (src/potentials/device_morse_zone.cuh:14-22): Morse forces for atoms
// [first_atom, first_atom + atom_count) from the whole neighbour list and
// position array. Forces are ADDED (the caller zeroes the zone first);
// *d_energy is added to.
#pragma once

#include "device_morse.hpp"

namespace tdmd::potentials {

inline void compute_morse_gpu_zone(const PositionVec* d_positions, ForceVec* d_forces,
                                   const i32* d_neighbors, const i32* d_offsets,
                                   const i32* d_counts, i32 first_atom, i32 atom_count,
                                   const Box& box, const MorseParams& params, accum_t* d_energy,
                                   hipStream_t stream = nullptr) {
  const tdmd_box b = hip::to_c(box);
  const tdmd_morse_params p = to_c(params);
  TDMD_HIP_CALL(tdmd_morse_force_zone(hip::kPrecision,
                                      reinterpret_cast<const double*>(d_positions), d_forces,
                                      d_neighbors, d_offsets, d_counts, first_atom, atom_count,
                                      &b, &p, d_energy, nullptr, hip::as_void(stream)));
}

}  // namespace tdmd::potentials
