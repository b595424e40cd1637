// nhc.hpp -- Nose-Hoover chain half-step, host and device
// (NoseHooverChain::half_step, src/integrator/device_nose_hoover.cu:206-289).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace tdmd_hip {

// Single outer-to-inner sweep with half_dt increments on the chain state
// (eta, eta_v, eta_m), chain masses refreshed from t_target / t_period
// (Q_1 = n_dof kT tau^2, Q_k = kT tau^2); returns exp(-eta_v[0] half_dt).
__host__ __device__ inline double nhc_half_step(int M, double* eta, double* eta_v, double* eta_m,
                                                double t_target, double t_period, double dt,
                                                int n_dof, double ke) {
  if (M == 0) return 1.0;
  const double kB = 8.617333262e-5;
  const double half_dt = 0.5 * dt;
  const double ke_target = 0.5 * static_cast<double>(n_dof) * kB * t_target;
  const double kt = kB * t_target;
  const double tau2 = t_period * t_period;
  eta_m[0] = static_cast<double>(n_dof) * kt * tau2;
  for (int i = 1; i < M; ++i) eta_m[i] = kt * tau2;
  const int last = M - 1;
  if (M > 1) {
    const double g_last =
        (eta_m[last - 1] * eta_v[last - 1] * eta_v[last - 1] - kB * t_target) / eta_m[last];
    eta_v[last] += g_last * half_dt;
  }
  for (int k = M - 2; k >= 1; --k) {
    const double ef = exp(-eta_v[k + 1] * half_dt / 4.0);
    eta_v[k] *= ef;
    const double g_k = (eta_m[k - 1] * eta_v[k - 1] * eta_v[k - 1] - kB * t_target) / eta_m[k];
    eta_v[k] += g_k * half_dt;
    eta_v[k] *= ef;
  }
  {
    const double ef = (M > 1) ? exp(-eta_v[1] * half_dt / 4.0) : 1.0;
    eta_v[0] *= ef;
    const double g0 = (2.0 * ke - 2.0 * ke_target) / eta_m[0];
    eta_v[0] += g0 * half_dt;
    eta_v[0] *= ef;
  }
  const double scale = exp(-eta_v[0] * half_dt);
  for (int k = 0; k < M; ++k) eta[k] += eta_v[k] * half_dt;
  return scale;
}

}  // namespace tdmd_hip
