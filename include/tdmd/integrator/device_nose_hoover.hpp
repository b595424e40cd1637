// device_nose_hoover.hpp -- Nose-Hoover chain and its GPU reductions
// (src/integrator/device_nose_hoover.cuh:26-88). The chain update is host
// code (tdmd_nhc_half_step); KE and vmax are reduced on the device in a
// fixed order and read back, as the reference's functions return them.
#pragma once

#include <vector>

#include "../core/device_buffer.hpp"
#include "../core/hip_abi.hpp"

namespace tdmd::integrator {

struct NHCConfig {
  real t_target{300.0};
  real t_period{0.1};
  i32 chain_length{3};
};

class NoseHooverChain {
 public:
  NoseHooverChain(const NHCConfig& cfg, real dt, i32 n_dof)
      : cfg_(cfg), dt_(dt), n_dof_(n_dof),
        eta_(static_cast<std::size_t>(cfg.chain_length), 0.0),
        eta_v_(static_cast<std::size_t>(cfg.chain_length), 0.0),
        eta_m_(static_cast<std::size_t>(cfg.chain_length), 0.0) {
    if (cfg.chain_length < 1) throw Error(TDMD_E_INVALID, "NHC chain length must be >= 1");
  }

  // One half-step of the chain for the current KE; returns the velocity
  // scale factor (device_nose_hoover.cu:206-289).
  real half_step(accum_t ke_current) {
    return static_cast<real>(tdmd_nhc_half_step(
        cfg_.chain_length, eta_.data(), eta_v_.data(), eta_m_.data(),
        static_cast<double>(cfg_.t_target), static_cast<double>(cfg_.t_period),
        static_cast<double>(dt_), n_dof_, ke_current));
  }

  real t_target() const noexcept { return cfg_.t_target; }
  void set_t_target(real t) { cfg_.t_target = t; }

 private:
  NHCConfig cfg_;
  real dt_;
  i32 n_dof_;
  std::vector<double> eta_, eta_v_, eta_m_;
};

inline accum_t device_compute_ke(const VelocityVec* d_velocities, const i32* d_types,
                                 const real* d_masses, i32 natoms) {
  if (natoms == 0) return 0;
  std::size_t wb = 0;
  TDMD_HIP_CALL(tdmd_reduce_workspace_bytes(natoms, &wb));
  DeviceBuffer<unsigned char> work(wb);
  DeviceBuffer<double> out(1);
  TDMD_HIP_CALL(tdmd_compute_ke(hip::kPrecision, reinterpret_cast<const double*>(d_velocities),
                                d_types, d_masses, natoms, out.data(), work.data(), nullptr));
  double ke = 0;
  out.copy_to_host(&ke, 1);
  return ke;
}

inline void device_scale_velocities(VelocityVec* d_velocities, i32 natoms, real factor) {
  TDMD_HIP_CALL(tdmd_scale_velocities(reinterpret_cast<double*>(d_velocities), natoms,
                                      static_cast<double>(factor), nullptr));
}

inline accum_t device_compute_vmax(const VelocityVec* d_velocities, i32 natoms) {
  if (natoms == 0) return 0;
  std::size_t wb = 0;
  TDMD_HIP_CALL(tdmd_reduce_workspace_bytes(natoms, &wb));
  DeviceBuffer<unsigned char> work(wb);
  DeviceBuffer<double> out(1);
  TDMD_HIP_CALL(tdmd_compute_vmax(reinterpret_cast<const double*>(d_velocities), natoms,
                                  out.data(), work.data(), nullptr));
  double v = 0;
  out.copy_to_host(&v, 1);
  return v;
}

}  // namespace tdmd::integrator
