// integrate.cpp -- velocity Verlet and Nose-Hoover-chain support on the
// device (src/integrator/device_velocity_verlet.cu:14-198,
// src/integrator/device_nose_hoover.cu:24-289).
//
// All arithmetic is FP64 (the reference's integrator contract,
// device_velocity_verlet.cu:22-31). Forces and masses arrive in the
// boundary's `real` type. The KE and |v|max reductions write their result to
// device memory through fixed-order block partials: no cudaMalloc, no
// blocking copy, no host sum (the reference does all three per call,
// device_nose_hoover.cu:80-105).
#include <cmath>

#include "common.hpp"
#include "nhc.hpp"

namespace tdmd_hip {

namespace {

template <class R>
__global__ void half_kick_kernel(double* __restrict__ v, const R* __restrict__ f,
                                 const int32_t* __restrict__ types, const R* __restrict__ masses,
                                 int natoms, double half_dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const double mass = static_cast<double>(masses[types[i]]);
  const double factor = half_dt / (mass * kMvv2e);
  v[3 * i] += factor * static_cast<double>(f[3 * i]);
  v[3 * i + 1] += factor * static_cast<double>(f[3 * i + 1]);
  v[3 * i + 2] += factor * static_cast<double>(f[3 * i + 2]);
}

__device__ __forceinline__ double wrap_single(double x, double lo, double bs, int periodic) {
  // device_velocity_verlet.cu:49-61: one conditional +/- box.
  if (periodic) {
    if (x < lo)
      x += bs;
    else if (x >= lo + bs)
      x -= bs;
  }
  return x;
}

__global__ void drift_kernel(double* __restrict__ p, const double* __restrict__ v, int natoms,
                             double dt, DevBox box) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  for (int a = 0; a < 3; ++a)
    p[3 * i + a] = wrap_single(p[3 * i + a] + dt * v[3 * i + a], box.lo[a], box.size[a],
                               box.periodic[a]);
}

template <class R, bool kFull>
__global__ void fused_kick_drift_kernel(double* __restrict__ v, double* __restrict__ p,
                                        const R* __restrict__ f, const int32_t* __restrict__ types,
                                        const R* __restrict__ masses, int natoms, double dt,
                                        DevBox box) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const double kick_dt = kFull ? dt : 0.5 * dt;  // device_velocity_verlet.cu:97-101
  const double mass = static_cast<double>(masses[types[i]]);
  const double factor = kick_dt / (mass * kMvv2e);
  for (int a = 0; a < 3; ++a) {
    const double va = v[3 * i + a] + factor * static_cast<double>(f[3 * i + a]);
    v[3 * i + a] = va;
    p[3 * i + a] = wrap_single(p[3 * i + a] + dt * va, box.lo[a], box.size[a], box.periodic[a]);
  }
}

template <class R>
__global__ void zero_forces_kernel(R* __restrict__ f, int natoms) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * natoms) return;
  f[i] = static_cast<R>(0);
}

template <class R>
__global__ void __launch_bounds__(kBlock)
ke_partials_kernel(const double* __restrict__ v, const int32_t* __restrict__ types,
                   const R* __restrict__ masses, int natoms, double* __restrict__ partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (i < natoms) {
    // device_nose_hoover.cu:33-41: 0.5 m mvv2e |v|^2 in FP64.
    const double mass = static_cast<double>(masses[types[i]]);
    const double vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
    e = 0.5 * mass * kMvv2e * (vx * vx + vy * vy + vz * vz);
  }
  const double t = block_sum<kBlock>(e);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kBlock)
vmax_partials_kernel(const double* __restrict__ v, int natoms, double* __restrict__ partials) {
  __shared__ double s[kBlock];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0.0;
  if (i < natoms) m = v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2];
  s[threadIdx.x] = m;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w && s[threadIdx.x + w] > s[threadIdx.x])
      s[threadIdx.x] = s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

__global__ void vmax_final_kernel(const double* __restrict__ partials, int n,
                                  double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double m = 0.0;
  for (int k = 0; k < n; ++k) m = partials[k] > m ? partials[k] : m;
  *out = sqrt(m);
}

__global__ void scale_vel_kernel(double* __restrict__ v, int n3, double s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) v[i] *= s;
}

}  // namespace

size_t reduce_workspace_bytes(int64_t n) {
  return static_cast<size_t>(grid_for(n > 0 ? n : 1)) * sizeof(double);
}

int ke_launch(int precision, const double* v, const int32_t* types, const void* masses,
              int natoms, double* out, void* work, hipStream_t stream) {
  const int grid = grid_for(natoms > 0 ? natoms : 1);
  double* partials = static_cast<double*>(work);
  if (precision == TDMD_MIXED)
    ke_partials_kernel<float><<<grid, kBlock, 0, stream>>>(
        v, types, static_cast<const float*>(masses), natoms, partials);
  else
    ke_partials_kernel<double><<<grid, kBlock, 0, stream>>>(
        v, types, static_cast<const double*>(masses), natoms, partials);
  TDMD_LAUNCH_CHECK();
  return launch_sum_partials(partials, grid, out, 0, stream);
}

}  // namespace tdmd_hip

using namespace tdmd_hip;

extern "C" {

int tdmd_half_kick(int32_t precision, double* d_vel, const void* d_forces,
                   const int32_t* d_types, const void* d_masses, int32_t natoms, double dt,
                   void* stream) {
  if (natoms < 0) return fail(TDMD_E_INVALID, "half_kick: natoms < 0");
  if (natoms == 0) return TDMD_OK;
  if (!d_vel || !d_forces || !d_types || !d_masses)
    return fail(TDMD_E_INVALID, "half_kick: null pointer");
  // device_velocity_verlet.cu:138-140: half_dt = 0.5 * dt.
  const double half_dt = 0.5 * dt;
  if (precision == TDMD_MIXED)
    half_kick_kernel<float><<<grid_for(natoms), kBlock, 0, as_stream(stream)>>>(
        d_vel, static_cast<const float*>(d_forces), d_types, static_cast<const float*>(d_masses),
        natoms, half_dt);
  else
    half_kick_kernel<double><<<grid_for(natoms), kBlock, 0, as_stream(stream)>>>(
        d_vel, static_cast<const double*>(d_forces), d_types,
        static_cast<const double*>(d_masses), natoms, half_dt);
  TDMD_LAUNCH_CHECK();
  return TDMD_OK;
}

int tdmd_drift(double* d_pos, const double* d_vel, int32_t natoms, double dt,
               const tdmd_box* box, void* stream) {
  if (!box || natoms < 0) return fail(TDMD_E_INVALID, "drift: bad arguments");
  if (natoms == 0) return TDMD_OK;
  if (!d_pos || !d_vel) return fail(TDMD_E_INVALID, "drift: null pointer");
  drift_kernel<<<grid_for(natoms), kBlock, 0, as_stream(stream)>>>(d_pos, d_vel, natoms, dt,
                                                                   make_devbox(*box));
  TDMD_LAUNCH_CHECK();
  return TDMD_OK;
}

int tdmd_fused_kick_drift(int32_t precision, int32_t full_kick, double* d_vel, double* d_pos,
                          const void* d_forces, const int32_t* d_types, const void* d_masses,
                          int32_t natoms, double dt, const tdmd_box* box, void* stream) {
  if (!box || natoms < 0) return fail(TDMD_E_INVALID, "fused_kick_drift: bad arguments");
  if (natoms == 0) return TDMD_OK;
  if (!d_vel || !d_pos || !d_forces || !d_types || !d_masses)
    return fail(TDMD_E_INVALID, "fused_kick_drift: null pointer");
  const DevBox db = make_devbox(*box);
  hipStream_t s = as_stream(stream);
  const int g = grid_for(natoms);
  if (precision == TDMD_MIXED) {
    const float* f = static_cast<const float*>(d_forces);
    const float* m = static_cast<const float*>(d_masses);
    if (full_kick)
      fused_kick_drift_kernel<float, true><<<g, kBlock, 0, s>>>(d_vel, d_pos, f, d_types, m, natoms, dt, db);
    else
      fused_kick_drift_kernel<float, false><<<g, kBlock, 0, s>>>(d_vel, d_pos, f, d_types, m, natoms, dt, db);
  } else {
    const double* f = static_cast<const double*>(d_forces);
    const double* m = static_cast<const double*>(d_masses);
    if (full_kick)
      fused_kick_drift_kernel<double, true><<<g, kBlock, 0, s>>>(d_vel, d_pos, f, d_types, m, natoms, dt, db);
    else
      fused_kick_drift_kernel<double, false><<<g, kBlock, 0, s>>>(d_vel, d_pos, f, d_types, m, natoms, dt, db);
  }
  TDMD_LAUNCH_CHECK();
  return TDMD_OK;
}

int tdmd_zero_forces(int32_t precision, void* d_forces, int32_t natoms, void* stream) {
  if (natoms < 0) return fail(TDMD_E_INVALID, "zero_forces: natoms < 0");
  if (natoms == 0) return TDMD_OK;
  if (!d_forces) return fail(TDMD_E_INVALID, "zero_forces: null pointer");
  if (precision == TDMD_MIXED)
    zero_forces_kernel<float><<<grid_for(3 * static_cast<int64_t>(natoms)), kBlock, 0, as_stream(stream)>>>(
        static_cast<float*>(d_forces), natoms);
  else
    zero_forces_kernel<double><<<grid_for(3 * static_cast<int64_t>(natoms)), kBlock, 0, as_stream(stream)>>>(
        static_cast<double*>(d_forces), natoms);
  TDMD_LAUNCH_CHECK();
  return TDMD_OK;
}

int tdmd_reduce_workspace_bytes(int64_t n, size_t* bytes) {
  if (!bytes || n < 0) return fail(TDMD_E_INVALID, "reduce_workspace_bytes: bad args");
  *bytes = reduce_workspace_bytes(n);
  return TDMD_OK;
}

int tdmd_compute_ke(int32_t precision, const double* d_vel, const int32_t* d_types,
                    const void* d_masses, int32_t natoms, double* d_ke_out, void* d_work,
                    void* stream) {
  if (natoms < 0 || !d_ke_out || !d_work) return fail(TDMD_E_INVALID, "compute_ke: bad arguments");
  if (natoms > 0 && (!d_vel || !d_types || !d_masses))
    return fail(TDMD_E_INVALID, "compute_ke: null pointer");
  return ke_launch(precision, d_vel, d_types, d_masses, natoms, d_ke_out, d_work,
                   as_stream(stream));
}

int tdmd_scale_velocities(double* d_vel, int32_t natoms, double factor, void* stream) {
  if (natoms < 0) return fail(TDMD_E_INVALID, "scale_velocities: natoms < 0");
  if (natoms == 0) return TDMD_OK;
  if (!d_vel) return fail(TDMD_E_INVALID, "scale_velocities: null pointer");
  scale_vel_kernel<<<grid_for(3 * static_cast<int64_t>(natoms)), kBlock, 0, as_stream(stream)>>>(
      d_vel, 3 * natoms, factor);
  TDMD_LAUNCH_CHECK();
  return TDMD_OK;
}

int tdmd_compute_vmax(const double* d_vel, int32_t natoms, double* d_vmax_out, void* d_work,
                      void* stream) {
  if (natoms < 0 || !d_vmax_out || !d_work) return fail(TDMD_E_INVALID, "compute_vmax: bad arguments");
  const int grid = grid_for(natoms > 0 ? natoms : 1);
  double* partials = static_cast<double*>(d_work);
  vmax_partials_kernel<<<grid, kBlock, 0, as_stream(stream)>>>(d_vel, natoms, partials);
  TDMD_LAUNCH_CHECK();
  vmax_final_kernel<<<1, 64, 0, as_stream(stream)>>>(partials, grid, d_vmax_out);
  TDMD_LAUNCH_CHECK();
  return TDMD_OK;
}

double tdmd_nhc_half_step(int32_t M, double* eta, double* eta_v, double* eta_m, double t_target,
                          double t_period, double dt, int32_t n_dof, double ke) {
  return nhc_half_step(M, eta, eta_v, eta_m, t_target, t_period, dt, n_dof, ke);
}

// Zone variants (device_velocity_verlet_zone.cu:14-118): the same kernels
// over atoms [first_atom, first_atom + atom_count) of the arrays.
static size_t force_bytes(int32_t precision) { return precision == TDMD_MIXED ? 4 : 8; }

int tdmd_half_kick_zone(int32_t precision, double* d_vel, const void* d_forces,
                        const int32_t* d_types, const void* d_masses, int32_t first_atom,
                        int32_t atom_count, double dt, void* stream) {
  if (first_atom < 0 || atom_count < 0) return fail(TDMD_E_INVALID, "half_kick_zone: bad range");
  if (atom_count == 0) return TDMD_OK;
  if (!d_vel || !d_forces || !d_types || !d_masses)
    return fail(TDMD_E_INVALID, "half_kick_zone: null pointer");
  const size_t f0 = 3 * static_cast<size_t>(first_atom);
  return tdmd_half_kick(precision, d_vel + f0,
                        static_cast<const char*>(d_forces) + f0 * force_bytes(precision),
                        d_types + first_atom, d_masses, atom_count, dt, stream);
}

int tdmd_drift_zone(double* d_pos, const double* d_vel, int32_t first_atom, int32_t atom_count,
                    double dt, const tdmd_box* box, void* stream) {
  if (!box || first_atom < 0 || atom_count < 0)
    return fail(TDMD_E_INVALID, "drift_zone: bad arguments");
  if (atom_count == 0) return TDMD_OK;
  if (!d_pos || !d_vel) return fail(TDMD_E_INVALID, "drift_zone: null pointer");
  const size_t f0 = 3 * static_cast<size_t>(first_atom);
  return tdmd_drift(d_pos + f0, d_vel + f0, atom_count, dt, box, stream);
}

int tdmd_zero_forces_zone(int32_t precision, void* d_forces, int32_t first_atom,
                          int32_t atom_count, void* stream) {
  if (first_atom < 0 || atom_count < 0) return fail(TDMD_E_INVALID, "zero_forces_zone: bad range");
  if (atom_count == 0) return TDMD_OK;
  if (!d_forces) return fail(TDMD_E_INVALID, "zero_forces_zone: null pointer");
  const size_t f0 = 3 * static_cast<size_t>(first_atom);
  return tdmd_zero_forces(precision, static_cast<char*>(d_forces) + f0 * force_bytes(precision),
                          atom_count, stream);
}

}  // extern "C"
