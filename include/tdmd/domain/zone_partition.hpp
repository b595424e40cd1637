// zone_partition.hpp -- ZonePartition (src/domain/zone_partition.hpp,
// zone_partition.cpp:15-143): slabs along x for the zone pipeline. Host
// only; the same rules as the library's pipeline (tdmd_pipe_create), in
// `real` as the reference computes them.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "../core/box.hpp"
#include "../scheduler/zone.hpp"

namespace tdmd::domain {

class ZonePartition {
 public:
  // n_zones = 0: max(3, floor(Lx / r_cut)); fewer than 3 throws.
  void build(const Box& box, real r_cut, i32 n_zones = 0) {
    const Vec3D size = box.size();
    box_lo_x_ = static_cast<real>(box.lo.x);
    if (n_zones <= 0) n_zones = std::max(3, static_cast<i32>(std::floor(size.x / r_cut)));
    if (n_zones < 3) throw Error(TDMD_E_INVALID, "need at least 3 zones for PBC");
    zone_width_ = static_cast<real>(size.x / n_zones);
    zones_.assign(static_cast<std::size_t>(n_zones), scheduler::Zone{});
    for (i32 i = 0; i < n_zones; ++i) {
      auto& z = zones_[static_cast<std::size_t>(i)];
      z.id = i;
      z.lattice_index = {i, 0, 0};
      z.bbox.lo = {box.lo.x + static_cast<double>(i) * (size.x / n_zones), box.lo.y, box.lo.z};
      z.bbox.hi = {box.lo.x + static_cast<double>(i + 1) * (size.x / n_zones), box.hi.y,
                   box.hi.z};
      z.owner_rank = 0;
    }
  }

  i32 zone_of(real x) const noexcept {
    auto z = static_cast<i32>((x - box_lo_x_) / zone_width_);
    if (z < 0) z = 0;
    if (z >= n_zones()) z = n_zones() - 1;
    return z;
  }

  // Stable counting sort of the arrays by zone; sets each zone's range.
  void assign_atoms(PositionVec* positions, VelocityVec* velocities, ForceVec* forces,
                    i32* types, i32* ids, i64 natoms) {
    const auto n = static_cast<std::size_t>(natoms);
    const auto nz = static_cast<std::size_t>(n_zones());
    std::vector<i32> zone(n), count(nz, 0), offset(nz, 0), placed(nz, 0);
    for (std::size_t i = 0; i < n; ++i) {
      zone[i] = zone_of(static_cast<real>(positions[i].x));
      ++count[static_cast<std::size_t>(zone[i])];
    }
    for (std::size_t z = 1; z < nz; ++z) offset[z] = offset[z - 1] + count[z - 1];
    std::vector<std::size_t> perm(n);
    for (std::size_t i = 0; i < n; ++i) {
      const auto z = static_cast<std::size_t>(zone[i]);
      perm[static_cast<std::size_t>(offset[z] + placed[z]++)] = i;
    }
    permute(positions, perm);
    permute(velocities, perm);
    permute(forces, perm);
    permute(types, perm);
    permute(ids, perm);
    for (std::size_t z = 0; z < nz; ++z) {
      zones_[z].atom_offset = offset[z];
      zones_[z].natoms_in_zone = count[z];
    }
  }

  // Zones within ceil(r_list / width) slabs, periodic, sorted, unique.
  void build_zone_neighbors(real r_list) {
    const i32 nz = n_zones();
    const auto span = static_cast<i32>(std::ceil(r_list / zone_width_));
    zone_neighbors_.assign(static_cast<std::size_t>(nz), {});
    for (i32 z = 0; z < nz; ++z) {
      auto& v = zone_neighbors_[static_cast<std::size_t>(z)];
      for (i32 d = -span; d <= span; ++d) {
        i32 k = z + d;
        if (k < 0)
          k += nz;
        else if (k >= nz)
          k -= nz;
        v.push_back(k);
      }
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
    }
  }

  i32 n_zones() const noexcept { return static_cast<i32>(zones_.size()); }
  real zone_width() const noexcept { return zone_width_; }
  const std::vector<scheduler::Zone>& zones() const noexcept { return zones_; }
  std::vector<scheduler::Zone>& zones() noexcept { return zones_; }
  const std::vector<i32>& zone_neighbors(i32 z) const {
    return zone_neighbors_[static_cast<std::size_t>(z)];
  }

 private:
  template <class T>
  static void permute(T* a, const std::vector<std::size_t>& perm) {
    std::vector<T> tmp(perm.size());
    for (std::size_t k = 0; k < perm.size(); ++k) tmp[k] = a[perm[k]];
    std::copy(tmp.begin(), tmp.end(), a);
  }

  real zone_width_{0};
  real box_lo_x_{0};
  std::vector<scheduler::Zone> zones_;
  std::vector<std::vector<i32>> zone_neighbors_;
};

}  // namespace tdmd::domain
