// tables.hpp -- EAM/alloy spline tables: host construction and device layout.
#pragma once

#include <string>
#include <vector>

#include "common.hpp"

namespace tdmd_hip {

// FD-Hermite cubic coefficients of one tabulated function
// (SplineTable::build_splines, src/potentials/eam_alloy.cpp:19-50).
void build_fd_hermite(const double* v, int n, double dr, double* a, double* b, double* c,
                      double* d);

// Device spline descriptor (DeviceSpline, src/potentials/device_eam.cuh:14-20):
// spacing, its reciprocal (lookups multiply instead of divide) and origin in
// FP64 whatever the coefficient precision, so the interval offset is exact.
struct SplineMeta {
  double dr;
  double inv_dr;
  double rmin;
  int n;       // knots; intervals = n - 1
  int offset;  // first interval in the coefficient array
};

template <class R>
struct Quad;
template <>
struct Quad<float> {
  using type = float4;
};
template <>
struct Quad<double> {
  using type = double4;
};

// Host-side setfl content (EamAlloy after read_setfl, eam_alloy.cpp:90-192).
struct SetflData {
  int ntypes = 0, nrho = 0, nr = 0;
  double drho = 0, dr = 0, cutoff = 0;
  std::vector<double> F, rho, rphi;  // rphi blocks in i<=j row order
  bool phi_values = false;           // rphi already holds phi = (r phi) / r
};

int read_setfl(const std::string& path, SetflData* out);

// Index of the (ti, tj) pair spline, upper triangle (eam_alloy.hpp:88-92).
__host__ __device__ inline int phi_index(int ti, int tj, int ntypes) {
  if (ti > tj) {
    const int t = ti;
    ti = tj;
    tj = t;
  }
  return ti * ntypes - ti * (ti + 1) / 2 + tj;
}

}  // namespace tdmd_hip

// Device tables (DeviceEam private state, device_eam.cuh:55-74).
struct tdmd_eam_tables {
  int precision = TDMD_FP64;
  int ntypes = 0;
  int nrho = 0, nr = 0;
  double drho = 0, dr = 0;
  double cutoff = 0;
  // Spline order in the coefficient array: F[0..nt), rho[0..nt), phi[0..np).
  void* d_coeff = nullptr;  // float4 / double4 quads
  void* d_meta = nullptr;   // SplineMeta[2*nt + np]
  size_t coeff_bytes = 0;
  // FP64 coefficient copies on the host (for checks and the adjoint).
  std::vector<double> a, b, c, d;
  std::vector<int> offsets;
};

namespace tdmd_hip {
int eam_tables_create(tdmd_eam_tables** out, int precision, const SetflData& s);
// Same coefficients, uploaded again in `precision` (host FP64 copies).
int eam_tables_clone(const tdmd_eam_tables* src, int precision, tdmd_eam_tables** out);
void eam_tables_destroy(tdmd_eam_tables* t);
}  // namespace tdmd_hip
