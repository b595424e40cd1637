// device_cell_list.hpp -- DeviceCellList (src/domain/device_cell_list.cuh:
// 16-54) over tdmd_cell_build. Grid n = max(3, floor(L / r_list)) per axis,
// z-major cell ids, atoms of a cell in increasing index order.
#pragma once

#include "../core/device_buffer.hpp"
#include "../core/hip_abi.hpp"

namespace tdmd::domain {

class DeviceCellList {
 public:
  void build(const PositionVec* d_positions, i64 natoms, const Box& box, real r_list,
             hipStream_t stream = nullptr) {
    const tdmd_box b = hip::to_c(box);
    double cs[3];
    TDMD_HIP_CALL(tdmd_cell_grid(&b, static_cast<double>(r_list), n_, cs));
    cell_size_ = Vec3D{cs[0], cs[1], cs[2]};
    const auto nc = static_cast<std::size_t>(ncells_total());
    const auto na = static_cast<std::size_t>(natoms > 0 ? natoms : 1);
    cell_counts_.resize(nc);
    cell_offsets_.resize(nc);
    cell_atoms_.resize(na);
    atom_cells_.resize(na);
    std::size_t wb = 0;
    TDMD_HIP_CALL(tdmd_cell_workspace_bytes(natoms, ncells_total(), &wb));
    if (work_.size() < wb) work_.resize(wb);
    TDMD_HIP_CALL(tdmd_cell_build(reinterpret_cast<const double*>(d_positions), natoms, &b,
                                  static_cast<double>(r_list), cell_counts_.data(),
                                  cell_offsets_.data(), cell_atoms_.data(), atom_cells_.data(),
                                  work_.data(), hip::as_void(stream)));
  }

  i32 ncells_x() const noexcept { return n_[0]; }
  i32 ncells_y() const noexcept { return n_[1]; }
  i32 ncells_z() const noexcept { return n_[2]; }
  i32 ncells_total() const noexcept { return n_[0] * n_[1] * n_[2]; }
  const i32* ncells() const noexcept { return n_; }

  const i32* d_cell_offsets() const noexcept { return cell_offsets_.data(); }
  const i32* d_cell_counts() const noexcept { return cell_counts_.data(); }
  const i32* d_cell_atoms() const noexcept { return cell_atoms_.data(); }
  const i32* d_atom_cells() const noexcept { return atom_cells_.data(); }
  Vec3D cell_size() const noexcept { return cell_size_; }

 private:
  i32 n_[3] = {0, 0, 0};
  Vec3D cell_size_{0, 0, 0};
  DeviceBuffer<i32> cell_counts_, cell_offsets_, cell_atoms_, atom_cells_;
  DeviceBuffer<unsigned char> work_;
};

}  // namespace tdmd::domain
