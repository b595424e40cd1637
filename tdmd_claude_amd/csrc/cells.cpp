// cells.cpp -- DeviceCellList build on gfx950
// (src/domain/device_cell_list.cu:59-150, CPU src/domain/cell_list.cpp:13-58).
//
// Pipeline (all on `stream`, no host round trip, unlike the reference's
// D2H-count / host-scan / H2D-offset / cudaDeviceSynchronize at :103-111,148):
//   assign   cell id per atom (the CPU cell_of formula: truncating cast of
//            (x - lo) / cell_size, clamped, cell_list.hpp:60-72), integer
//            atomic histogram (exact, order independent);
//   scan     atomics-free exclusive scan -> offsets;
//   scatter  slot reservation per cell;
//   order    each cell's atoms sorted by index, so the layout equals the
//            CPU counting sort's (index order within a cell) on every run.
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace tdmd_hip {

size_t scan_workspace_bytes(int64_t n);
int scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* meta,
                       void* work, hipStream_t stream);

namespace {

__global__ void assign_cells_kernel(const double* __restrict__ pos, int64_t natoms,
                                    DevBox box, double csx, double csy, double csz,
                                    int ncx, int ncy, int ncz,
                                    int32_t* __restrict__ atom_cells,
                                    int32_t* __restrict__ counts) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  int ix = static_cast<int>((pos[3 * i] - box.lo[0]) / csx);
  int iy = static_cast<int>((pos[3 * i + 1] - box.lo[1]) / csy);
  int iz = static_cast<int>((pos[3 * i + 2] - box.lo[2]) / csz);
  ix = ix >= ncx ? ncx - 1 : ix;
  iy = iy >= ncy ? ncy - 1 : iy;
  iz = iz >= ncz ? ncz - 1 : iz;
  ix = ix < 0 ? 0 : ix;
  iy = iy < 0 ? 0 : iy;
  iz = iz < 0 ? 0 : iz;
  const int c = iz * ncx * ncy + iy * ncx + ix;
  atom_cells[i] = c;
  atomicAdd(&counts[c], 1);
}

__global__ void scatter_cells_kernel(const int32_t* __restrict__ atom_cells, int64_t natoms,
                                     const int32_t* __restrict__ offsets,
                                     int32_t* __restrict__ cursor,
                                     int32_t* __restrict__ cell_atoms) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const int c = atom_cells[i];
  const int slot = offsets[c] + atomicAdd(&cursor[c], 1);
  cell_atoms[slot] = static_cast<int32_t>(i);
}

// Insertion sort of each cell's segment by atom index. Cells hold tens of
// atoms in a dense system, so one thread per cell is enough.
__global__ void order_cells_kernel(const int32_t* __restrict__ offsets,
                                   const int32_t* __restrict__ counts, int ncells,
                                   int32_t* __restrict__ cell_atoms) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  int32_t* seg = cell_atoms + offsets[c];
  const int n = counts[c];
  for (int a = 1; a < n; ++a) {
    const int32_t v = seg[a];
    int b = a - 1;
    while (b >= 0 && seg[b] > v) {
      seg[b + 1] = seg[b];
      --b;
    }
    seg[b + 1] = v;
  }
}

}  // namespace

void cell_grid(const tdmd_box& box, double r_list, int n[3], double cs[3]) {
  // device_cell_list.cu:65-75 == cell_list.cpp:22-27.
  for (int a = 0; a < 3; ++a) {
    const double L = box.hi[a] - box.lo[a];
    n[a] = std::max(3, static_cast<int>(std::floor(L / r_list)));
    cs[a] = L / static_cast<double>(n[a]);
  }
}

size_t cell_workspace_bytes(int64_t natoms, int ncells) {
  (void)natoms;
  const size_t cursor = ((static_cast<size_t>(ncells) * 4 + 255) / 256) * 256;
  return cursor + scan_workspace_bytes(ncells);
}

int cell_build(const double* d_pos, int64_t natoms, const tdmd_box& box, double r_list,
               int32_t* d_counts, int32_t* d_offsets, int32_t* d_cell_atoms,
               int32_t* d_atom_cells, void* d_work, hipStream_t stream) {
  int n[3];
  double cs[3];
  cell_grid(box, r_list, n, cs);
  const int ncells = n[0] * n[1] * n[2];
  char* work = static_cast<char*>(d_work);
  int32_t* cursor = reinterpret_cast<int32_t*>(work);
  void* scan_work = work + ((static_cast<size_t>(ncells) * 4 + 255) / 256) * 256;
  int rc = tdmd_fill_u32(d_counts, 0u, static_cast<size_t>(ncells), stream);
  if (rc) return rc;
  rc = tdmd_fill_u32(cursor, 0u, static_cast<size_t>(ncells), stream);
  if (rc) return rc;
  const DevBox db = make_devbox(box);
  if (natoms > 0) {
    assign_cells_kernel<<<grid_for(natoms), kBlock, 0, stream>>>(
        d_pos, natoms, db, cs[0], cs[1], cs[2], n[0], n[1], n[2], d_atom_cells, d_counts);
    TDMD_LAUNCH_CHECK();
  }
  rc = scan_exclusive_i32(d_counts, d_offsets, ncells, nullptr, scan_work, stream);
  if (rc) return rc;
  if (natoms > 0) {
    scatter_cells_kernel<<<grid_for(natoms), kBlock, 0, stream>>>(
        d_atom_cells, natoms, d_offsets, cursor, d_cell_atoms);
    TDMD_LAUNCH_CHECK();
    order_cells_kernel<<<grid_for(ncells), kBlock, 0, stream>>>(d_offsets, d_counts, ncells,
                                                                d_cell_atoms);
    TDMD_LAUNCH_CHECK();
  }
  return TDMD_OK;
}

}  // namespace tdmd_hip

using namespace tdmd_hip;

extern "C" {

int tdmd_cell_grid(const tdmd_box* box, double r_list, int32_t ncells[3],
                   double cell_size[3]) {
  if (!box || !(r_list > 0.0)) return fail(TDMD_E_INVALID, "cell_grid: r_list must be positive");
  int n[3];
  double cs[3];
  cell_grid(*box, r_list, n, cs);
  for (int a = 0; a < 3; ++a) {
    ncells[a] = n[a];
    if (cell_size) cell_size[a] = cs[a];
  }
  return TDMD_OK;
}

int tdmd_cell_workspace_bytes(int64_t natoms, int32_t ncells_total, size_t* bytes) {
  if (!bytes || natoms < 0 || ncells_total <= 0)
    return fail(TDMD_E_INVALID, "cell_workspace_bytes: bad args");
  *bytes = cell_workspace_bytes(natoms, ncells_total);
  return TDMD_OK;
}

int tdmd_cell_build(const double* d_pos, int64_t natoms, const tdmd_box* box,
                    double r_list, int32_t* d_cell_counts, int32_t* d_cell_offsets,
                    int32_t* d_cell_atoms, int32_t* d_atom_cells, void* d_work,
                    void* stream) {
  // device_cell_list.cu:60-62: r_list > 0, natoms >= 0.
  if (!box || !(r_list > 0.0)) return fail(TDMD_E_INVALID, "cell_build: r_list must be positive");
  if (natoms < 0) return fail(TDMD_E_INVALID, "cell_build: natoms must be non-negative");
  if (natoms > 0 && (!d_pos || !d_cell_atoms || !d_atom_cells))
    return fail(TDMD_E_INVALID, "cell_build: null pointer");
  if (!d_cell_counts || !d_cell_offsets || !d_work)
    return fail(TDMD_E_INVALID, "cell_build: null pointer");
  return cell_build(d_pos, natoms, *box, r_list, d_cell_counts, d_cell_offsets, d_cell_atoms,
                    d_atom_cells, d_work, as_stream(stream));
}

}  // extern "C"
