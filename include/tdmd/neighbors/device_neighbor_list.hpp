// device_neighbor_list.hpp -- DeviceNeighborList (src/neighbors/
// device_neighbor_list.cuh:20-65): full CSR list, built as count pass ->
// atomics-free scan (total and max in one 8-byte read back, like the
// reference's single sync) -> fill. Inclusion is the CPU list's FP64 test,
// so each row is the symmetric closure of the CPU half list.
#pragma once

#include "../domain/device_cell_list.hpp"

namespace tdmd::neighbors {

class DeviceNeighborList {
 public:
  void build(const PositionVec* d_positions, i64 natoms, const Box& box, real r_cut,
             real r_skin, hipStream_t stream = nullptr) {
    r_cut_ = r_cut;
    r_skin_ = r_skin;
    const double r_list = static_cast<double>(r_cut) + static_cast<double>(r_skin);
    cell_list_.build(d_positions, natoms, box, static_cast<real>(r_list), stream);
    const auto n = static_cast<std::size_t>(natoms > 0 ? natoms : 1);
    counts_.resize(n);
    offsets_.resize(n);
    const tdmd_box b = hip::to_c(box);
    const auto* pos = reinterpret_cast<const double*>(d_positions);
    void* s = hip::as_void(stream);
    TDMD_HIP_CALL(tdmd_nlist_count(pos, natoms, &b, r_list, cell_list_.d_cell_counts(),
                                   cell_list_.d_cell_offsets(), cell_list_.d_cell_atoms(),
                                   cell_list_.d_atom_cells(), cell_list_.ncells(),
                                   counts_.data(), s));
    std::size_t wb = 0;
    TDMD_HIP_CALL(tdmd_scan_workspace_bytes(natoms, &wb));
    if (scan_work_.size() < wb) scan_work_.resize(wb);
    meta_.resize(2);
    TDMD_HIP_CALL(tdmd_scan_exclusive_i32(counts_.data(), offsets_.data(), natoms, meta_.data(),
                                          scan_work_.data(), s));
    i32 meta[2] = {0, 0};
    TDMD_HIP_CALL(tdmd_memcpy_dtoh(meta, meta_.data(), sizeof(meta), s));
    TDMD_HIP_CALL(tdmd_stream_synchronize(s));
    max_neighbors_ = meta[1];
    neighbors_.resize(static_cast<std::size_t>(meta[0] > 0 ? meta[0] : 1));
    overflow_.resize(1);
    overflow_.zero();
    TDMD_HIP_CALL(tdmd_nlist_fill(pos, natoms, &b, r_list, cell_list_.d_cell_counts(),
                                  cell_list_.d_cell_offsets(), cell_list_.d_cell_atoms(),
                                  cell_list_.d_atom_cells(), cell_list_.ncells(),
                                  offsets_.data(), neighbors_.data(), counts_.data(),
                                  max_neighbors_, overflow_.data(), s));
  }

  const i32* d_neighbors() const noexcept { return neighbors_.data(); }
  const i32* d_counts() const noexcept { return counts_.data(); }
  const i32* d_offsets() const noexcept { return offsets_.data(); }
  real cutoff() const noexcept { return r_cut_; }
  real skin() const noexcept { return r_skin_; }
  i32 max_neighbors() const noexcept { return max_neighbors_; }
  const domain::DeviceCellList& cell_list() const noexcept { return cell_list_; }

 private:
  real r_cut_{0};
  real r_skin_{0};
  i32 max_neighbors_{0};
  domain::DeviceCellList cell_list_;
  DeviceBuffer<i32> neighbors_, counts_, offsets_, meta_, overflow_;
  DeviceBuffer<unsigned char> scan_work_;
};

}  // namespace tdmd::neighbors
