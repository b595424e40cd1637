// device_eam.hpp -- DeviceEam (src/potentials/device_eam.cuh:14-75).
//
// upload_tables() takes any EamAlloy-like object with the reference's
// accessors (eam_alloy.hpp): ntypes(), cutoff(), embedding(), density(),
// phi(), each table with n, dr and values -- the reference's EamAlloy
// itself, or tdmd::potentials::SetflPotential from setfl_potential.hpp.
// compute() runs density + embedding + force over a full CSR list; forces
// stored, *d_energy added to.
#pragma once

#include <vector>

#include "../core/device_buffer.hpp"
#include "../core/hip_abi.hpp"

namespace tdmd::potentials {

struct DeviceSpline {
  real dr;
  real rmin;
  i32 n;
  i32 offset;
};

class DeviceEam {
 public:
  DeviceEam() = default;
  ~DeviceEam() {
    if (tables_) (void)tdmd_eam_tables_destroy(tables_);
  }
  DeviceEam(const DeviceEam&) = delete;
  DeviceEam& operator=(const DeviceEam&) = delete;

  template <class EamAlloyLike>
  void upload_tables(const EamAlloyLike& eam) {
    const i32 nt = eam.ntypes();
    const auto& F = eam.embedding();
    const auto& rho = eam.density();
    const auto& phi = eam.phi();
    const i32 np = nt * (nt + 1) / 2;
    if (nt < 1 || static_cast<i32>(F.size()) != nt || static_cast<i32>(rho.size()) != nt ||
        static_cast<i32>(phi.size()) != np)
      throw Error(TDMD_E_INVALID, "DeviceEam::upload_tables: inconsistent EamAlloy tables");
    const i32 nrho = F[0].n, nr = rho[0].n;
    std::vector<double> f, r, p;
    for (const auto& t : F) f.insert(f.end(), t.values.begin(), t.values.end());
    for (const auto& t : rho) r.insert(r.end(), t.values.begin(), t.values.end());
    for (const auto& t : phi) p.insert(p.end(), t.values.begin(), t.values.end());
    if (f.size() != static_cast<std::size_t>(nt) * nrho ||
        r.size() != static_cast<std::size_t>(nt) * nr ||
        p.size() != static_cast<std::size_t>(np) * nr)
      throw Error(TDMD_E_INVALID, "DeviceEam::upload_tables: tables of unequal length");
    if (tables_) (void)tdmd_eam_tables_destroy(tables_);
    tables_ = nullptr;
    TDMD_HIP_CALL(tdmd_eam_tables_create_values(
        &tables_, hip::kPrecision, nt, nrho, static_cast<double>(F[0].dr), nr,
        static_cast<double>(rho[0].dr), static_cast<double>(eam.cutoff()), f.data(), r.data(),
        p.data()));
    ntypes_ = nt;
    cutoff_ = eam.cutoff();
  }

  void compute(const PositionVec* d_positions, ForceVec* d_forces, const i32* d_types,
               const i32* d_neighbors, const i32* d_offsets, const i32* d_counts, i32 natoms,
               const Box& box, accum_t* d_energy, hipStream_t stream = nullptr) {
    if (!tables_) throw Error(TDMD_E_STATE, "DeviceEam::compute before upload_tables");
    // Per-atom scratch, as the reference allocates inside compute
    // (device_eam.cu:353-357).
    d_rho_.resize(static_cast<std::size_t>(natoms > 0 ? natoms : 1));
    d_fp_.resize(static_cast<std::size_t>(natoms > 0 ? natoms : 1));
    const tdmd_box b = hip::to_c(box);
    TDMD_HIP_CALL(tdmd_eam_compute(tables_, reinterpret_cast<const double*>(d_positions),
                                   d_forces, d_types, d_neighbors, d_offsets, d_counts, natoms,
                                   &b, d_rho_.data(), d_fp_.data(), d_energy, nullptr,
                                   hip::as_void(stream)));
  }

  real cutoff() const noexcept { return cutoff_; }
  i32 ntypes() const noexcept { return ntypes_; }
  const tdmd_eam_tables* handle() const noexcept { return tables_; }

 private:
  tdmd_eam_tables* tables_ = nullptr;
  i32 ntypes_{0};
  real cutoff_{0};
  DeviceBuffer<accum_t> d_rho_;
  DeviceBuffer<real> d_fp_;
};

}  // namespace tdmd::potentials
