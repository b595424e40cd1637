// nlist_csr.cpp -- full CSR Verlet list for the DeviceNeighborList boundary
// (src/neighbors/device_neighbor_list.cu:26-139,158-278).
//
// One thread per atom i walks the 27 stencil cells (z, y, x outer to inner,
// single periodic wrap -- cell_list.hpp:47-56) and tests every j != i in
// FP64 exactly as the CPU half list does (neighbor_list.cpp:58-62):
// d = x_j - x_i, one conditional minimum image per axis, r2 = dx*dx + dy*dy
// + dz*dz rounded term by term, include iff r2 < r_list^2. Because
// (x_j - x_i) == -(x_i - x_j) exactly in IEEE arithmetic and the minimum
// image is odd, the GPU full list is exactly the symmetric closure of the
// CPU half list. The cell of atom i is the one the cell build assigned
// (computed once, not re-derived in a second precision as the reference's
// kernel does at :43-58).
#include "common.hpp"

namespace tdmd_hip {

namespace {

template <bool kFill>
__global__ void __launch_bounds__(kBlock)
nlist_csr_kernel(const double* __restrict__ pos, int natoms, DevBox box, double r_list_sq,
                 const int32_t* __restrict__ cell_counts, const int32_t* __restrict__ cell_offsets,
                 const int32_t* __restrict__ cell_atoms, const int32_t* __restrict__ atom_cells,
                 int ncx, int ncy, int ncz, const int32_t* __restrict__ offsets,
                 int32_t* __restrict__ neighbors, int32_t* __restrict__ counts, int max_neighbors,
                 int32_t* __restrict__ overflow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  const double hx = 0.5 * box.size[0], hy = 0.5 * box.size[1], hz = 0.5 * box.size[2];
  const int ci = atom_cells[i];
  const int iz = ci / (ncx * ncy);
  const int iy = (ci - iz * ncx * ncy) / ncx;
  const int ix = ci - iz * ncx * ncy - iy * ncx;
  const int off = kFill ? offsets[i] : 0;
  int count = 0;
  for (int dz = -1; dz <= 1; ++dz) {
    int nz = iz + dz;
    nz = nz < 0 ? nz + ncz : (nz >= ncz ? nz - ncz : nz);
    for (int dy = -1; dy <= 1; ++dy) {
      int ny = iy + dy;
      ny = ny < 0 ? ny + ncy : (ny >= ncy ? ny - ncy : ny);
      for (int dx = -1; dx <= 1; ++dx) {
        int nx = ix + dx;
        nx = nx < 0 ? nx + ncx : (nx >= ncx ? nx - ncx : nx);
        const int c = nz * ncx * ncy + ny * ncx + nx;
        const int start = cell_offsets[c], cnt = cell_counts[c];
        for (int k = 0; k < cnt; ++k) {
          const int j = cell_atoms[start + k];
          if (j == i) continue;
          const double ddx = min_image(pos[3 * j] - xi, box.size[0], hx, box.periodic[0]);
          const double ddy = min_image(pos[3 * j + 1] - yi, box.size[1], hy, box.periodic[1]);
          const double ddz = min_image(pos[3 * j + 2] - zi, box.size[2], hz, box.periodic[2]);
          if (r2_unfused(ddx, ddy, ddz) < r_list_sq) {
            if (kFill) {
              if (count < max_neighbors)
                neighbors[off + count] = j;
              else
                atomicMax(overflow, count + 1);
            }
            ++count;
          }
        }
      }
    }
  }
  counts[i] = count;
}

}  // namespace

int nlist_csr(bool fill, const double* d_pos, int64_t natoms, const tdmd_box& box, double r_list,
              const int32_t* cc, const int32_t* co, const int32_t* ca, const int32_t* ac,
              const int32_t n[3], const int32_t* offsets, int32_t* nbrs, int32_t* counts,
              int32_t max_nb, int32_t* overflow, hipStream_t stream) {
  if (natoms == 0) return TDMD_OK;
  const DevBox db = make_devbox(box);
  const double r_list_sq = r_list * r_list;
  const int na = static_cast<int>(natoms);
  if (fill) {
    nlist_csr_kernel<true><<<grid_for(natoms), kBlock, 0, stream>>>(
        d_pos, na, db, r_list_sq, cc, co, ca, ac, n[0], n[1], n[2], offsets, nbrs, counts,
        max_nb, overflow);
  } else {
    nlist_csr_kernel<false><<<grid_for(natoms), kBlock, 0, stream>>>(
        d_pos, na, db, r_list_sq, cc, co, ca, ac, n[0], n[1], n[2], nullptr, nullptr, counts, 0,
        nullptr);
  }
  TDMD_LAUNCH_CHECK();
  return TDMD_OK;
}

}  // namespace tdmd_hip

using namespace tdmd_hip;

extern "C" {

int tdmd_nlist_count(const double* d_pos, int64_t natoms, const tdmd_box* box, double r_list,
                     const int32_t* d_cell_counts, const int32_t* d_cell_offsets,
                     const int32_t* d_cell_atoms, const int32_t* d_atom_cells,
                     const int32_t ncells[3], int32_t* d_counts, void* stream) {
  if (!box || !ncells || !(r_list > 0.0) || natoms < 0 || natoms > INT32_MAX)
    return fail(TDMD_E_INVALID, "nlist_count: bad arguments");
  if (natoms > 0 && (!d_pos || !d_cell_counts || !d_cell_offsets || !d_cell_atoms ||
                     !d_atom_cells || !d_counts))
    return fail(TDMD_E_INVALID, "nlist_count: null pointer");
  return nlist_csr(false, d_pos, natoms, *box, r_list, d_cell_counts, d_cell_offsets,
                   d_cell_atoms, d_atom_cells, ncells, nullptr, nullptr, d_counts, 0, nullptr,
                   as_stream(stream));
}

int tdmd_nlist_fill(const double* d_pos, int64_t natoms, const tdmd_box* box, double r_list,
                    const int32_t* d_cell_counts, const int32_t* d_cell_offsets,
                    const int32_t* d_cell_atoms, const int32_t* d_atom_cells,
                    const int32_t ncells[3], const int32_t* d_offsets, int32_t* d_neighbors,
                    int32_t* d_counts, int32_t max_neighbors, int32_t* d_overflow,
                    void* stream) {
  if (!box || !ncells || !(r_list > 0.0) || natoms < 0 || natoms > INT32_MAX ||
      max_neighbors < 0)
    return fail(TDMD_E_INVALID, "nlist_fill: bad arguments");
  if (natoms > 0 && (!d_pos || !d_cell_counts || !d_cell_offsets || !d_cell_atoms ||
                     !d_atom_cells || !d_offsets || !d_counts || !d_overflow ||
                     (max_neighbors > 0 && !d_neighbors)))
    return fail(TDMD_E_INVALID, "nlist_fill: null pointer");
  return nlist_csr(true, d_pos, natoms, *box, r_list, d_cell_counts, d_cell_offsets,
                   d_cell_atoms, d_atom_cells, ncells, d_offsets, d_neighbors, d_counts,
                   max_neighbors, d_overflow, as_stream(stream));
}

}  // extern "C"
