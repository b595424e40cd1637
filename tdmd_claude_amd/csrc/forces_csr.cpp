// forces_csr.cpp -- Morse and EAM forces over a full CSR list: the
// compute_morse_gpu / DeviceEam::compute boundary
// (src/potentials/device_morse.cu:32-170, src/potentials/device_eam.cu:72-421).
//
// One thread per atom. Pair geometry (difference, minimum image, r^2, cutoff,
// r) in FP64, pair functions in R (pair_math.hpp). Per-atom sums (force,
// density, energy) are FP64. Forces are stored, not accumulated
// (device_morse.cu:110-117); energies are added to *d_energy: one FP64
// atomic per block when no workspace is given, otherwise per-block partials
// summed in a fixed order.
#include "pair_math.hpp"

namespace tdmd_hip {

constexpr int kMaxSplines = 64;  // 2*ntypes + ntypes(ntypes+1)/2 for ntypes <= 8

namespace {

__device__ __forceinline__ double pair_geom(const double* __restrict__ pos, double xi, double yi,
                                            double zi, int j, const DevBox& box, double* dx,
                                            double* dy, double* dz) {
  *dx = min_image(xi - pos[3 * j], box.size[0], 0.5 * box.size[0], box.periodic[0]);
  *dy = min_image(yi - pos[3 * j + 1], box.size[1], 0.5 * box.size[1], box.periodic[1]);
  *dz = min_image(zi - pos[3 * j + 2], box.size[2], 0.5 * box.size[2], box.periodic[2]);
  return *dx * *dx + *dy * *dy + *dz * *dz;
}

// 0-based table index of a 1-based type. Types outside 1..ntypes (which the
// reference would use to index out of its tables) are clamped to the range.
__device__ __forceinline__ int clamp_type(int t, int ntypes) {
  t -= 1;
  return t < 0 ? 0 : (t >= ntypes ? ntypes - 1 : t);
}

__device__ __forceinline__ void emit_energy(double e, double* __restrict__ partials,
                                            double* __restrict__ d_energy) {
  const double t = block_sum<kBlock>(e);
  if (threadIdx.x == 0) {
    if (partials)
      partials[blockIdx.x] = t;
    else
      atomicAdd(d_energy, t);
  }
}

// Atoms [first, first + count); kAccum adds the forces to what `forces`
// holds (the zone pass, device_morse_zone.cu:29-117) instead of storing them.
template <class R, bool kEnergy, bool kAccum>
__global__ void __launch_bounds__(kBlock)
morse_csr_kernel(const double* __restrict__ pos, R* __restrict__ forces,
                 const int32_t* __restrict__ nbrs, const int32_t* __restrict__ offsets,
                 const int32_t* __restrict__ counts, int first, int count, DevBox box,
                 MorseCoef p, double* __restrict__ partials, double* __restrict__ d_energy) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = first + t;
  double fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0;
  if (t < count) {
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const int off = offsets[i], cnt = counts[i];
    for (int k = 0; k < cnt; ++k) {
      double dx, dy, dz;
      const double r2 = pair_geom(pos, xi, yi, zi, nbrs[off + k], box, &dx, &dy, &dz);
      if (r2 < p.rc_sq) {
        const PairR<R> pr = pair_r<R>(r2);
        R e;
        const double fpair = static_cast<double>(morse_fpair<R>(p, pr, &e));
        if (kEnergy) pe += 0.5 * static_cast<double>(e);
        fx = fma(fpair, dx, fx);
        fy = fma(fpair, dy, fy);
        fz = fma(fpair, dz, fz);
      }
    }
    if (kAccum) {
      forces[3 * i] += static_cast<R>(fx);
      forces[3 * i + 1] += static_cast<R>(fy);
      forces[3 * i + 2] += static_cast<R>(fz);
    } else {
      forces[3 * i] = static_cast<R>(fx);
      forces[3 * i + 1] = static_cast<R>(fy);
      forces[3 * i + 2] = static_cast<R>(fz);
    }
  }
  if (kEnergy) emit_energy(pe, partials, d_energy);
}

__device__ __forceinline__ void load_meta(SplineMeta* sm, const SplineMeta* meta, int ns) {
  for (int k = threadIdx.x; k < ns; k += blockDim.x) sm[k] = meta[k];
  __syncthreads();
}

// Pass 1: rho_i = sum_j rho_{t_j}(r_ij), FP64 sum (device_eam.cu:72-133).
template <class R>
__global__ void __launch_bounds__(kBlock)
eam_density_csr_kernel(const double* __restrict__ pos, const int32_t* __restrict__ types,
                       const int32_t* __restrict__ nbrs, const int32_t* __restrict__ offsets,
                       const int32_t* __restrict__ counts, int natoms, int ntypes, DevBox box,
                       double rc_sq, const SplineMeta* __restrict__ meta,
                       const typename Quad<R>::type* __restrict__ coeff,
                       double* __restrict__ rho) {
  __shared__ SplineMeta smeta[kMaxSplines];
  load_meta(smeta, meta, 2 * ntypes);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  const int off = offsets[i], cnt = counts[i];
  double rho_i = 0.0;
  for (int k = 0; k < cnt; ++k) {
    const int j = nbrs[off + k];
    double dx, dy, dz;
    const double r2 = pair_geom(pos, xi, yi, zi, j, box, &dx, &dy, &dz);
    if (r2 >= rc_sq) continue;
    const PairR<R> pr = pair_r<R>(r2);
    rho_i += static_cast<double>(
        spline_value<R>(smeta[ntypes + clamp_type(types[j], ntypes)], coeff, pr.r));
  }
  rho[i] = rho_i;
}

// Pass 2: fp_i = F'_{t_i}(rho_i); energy F(rho_i) (device_eam.cu:137-162).
template <class R, bool kEnergy>
__global__ void __launch_bounds__(kBlock)
eam_embed_csr_kernel(const int32_t* __restrict__ types, int natoms, int ntypes,
                     const SplineMeta* __restrict__ meta,
                     const typename Quad<R>::type* __restrict__ coeff,
                     const double* __restrict__ rho, R* __restrict__ fp,
                     double* __restrict__ partials, double* __restrict__ d_energy) {
  __shared__ SplineMeta smeta[kMaxSplines];
  load_meta(smeta, meta, ntypes);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (i < natoms) {
    const SplineVal<R> F = spline_both<R>(smeta[clamp_type(types[i], ntypes)], coeff, rho[i]);
    fp[i] = F.df;
    e = static_cast<double>(F.f);
  }
  if (kEnergy) emit_energy(e, partials, d_energy);
}

// Pass 3: F_i = sum_j -(phi' + fp_i rho'_{t_j} + fp_j rho'_{t_i})/r * d
// (device_eam.cu:176-278), plus 0.5 phi per pair of the full list.
template <class R, bool kEnergy>
__global__ void __launch_bounds__(kBlock)
eam_force_csr_kernel(const double* __restrict__ pos, R* __restrict__ forces,
                     const int32_t* __restrict__ types, const int32_t* __restrict__ nbrs,
                     const int32_t* __restrict__ offsets, const int32_t* __restrict__ counts,
                     int natoms, int ntypes, DevBox box, double rc_sq,
                     const SplineMeta* __restrict__ meta,
                     const typename Quad<R>::type* __restrict__ coeff,
                     const R* __restrict__ fp, double* __restrict__ partials,
                     double* __restrict__ d_energy) {
  __shared__ SplineMeta smeta[kMaxSplines];
  const int np = ntypes * (ntypes + 1) / 2;
  load_meta(smeta, meta, 2 * ntypes + np);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0;
  if (i < natoms) {
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const int ti = clamp_type(types[i], ntypes);
    const R fpi = fp[i];
    const int off = offsets[i], cnt = counts[i];
    for (int k = 0; k < cnt; ++k) {
      const int j = nbrs[off + k];
      double dx, dy, dz;
      const double r2 = pair_geom(pos, xi, yi, zi, j, box, &dx, &dy, &dz);
      if (r2 >= rc_sq) continue;
      const PairR<R> pr = pair_r<R>(r2);
      const int tj = clamp_type(types[j], ntypes);
      const int pidx = phi_index(ti, tj, ntypes);
      R dphi;
      if (kEnergy) {
        const SplineVal<R> ph = spline_both<R>(smeta[2 * ntypes + pidx], coeff, pr.r);
        pe += 0.5 * static_cast<double>(ph.f);
        dphi = ph.df;
      } else {
        dphi = spline_deriv<R>(smeta[2 * ntypes + pidx], coeff, pr.r);
      }
      const R drho_j = spline_deriv<R>(smeta[ntypes + tj], coeff, pr.r);
      const R drho_i = spline_deriv<R>(smeta[ntypes + ti], coeff, pr.r);
      const double fpair =
          static_cast<double>(-(dphi + fpi * drho_j + fp[j] * drho_i) * pr.rinv);
      fx = fma(fpair, dx, fx);
      fy = fma(fpair, dy, fy);
      fz = fma(fpair, dz, fz);
    }
    forces[3 * i] = static_cast<R>(fx);
    forces[3 * i + 1] = static_cast<R>(fy);
    forces[3 * i + 2] = static_cast<R>(fz);
  }
  if (kEnergy) emit_energy(pe, partials, d_energy);
}

template <class R, bool kAccum>
int morse_launch(const double* pos, void* forces, const int32_t* nbrs, const int32_t* offsets,
                 const int32_t* counts, int first, int count, const DevBox& box,
                 const tdmd_morse_params& p, double* d_energy, void* work, hipStream_t stream) {
  const int grid = grid_for(count);
  double* partials = static_cast<double*>(work);
  const MorseCoef c{p.d, p.alpha, p.r0, p.rc_sq};
  if (d_energy) {
    morse_csr_kernel<R, true, kAccum><<<grid, kBlock, 0, stream>>>(
        pos, static_cast<R*>(forces), nbrs, offsets, counts, first, count, box, c, partials,
        d_energy);
    TDMD_LAUNCH_CHECK();
    if (partials) return launch_sum_partials(partials, grid, d_energy, 1, stream);
  } else {
    morse_csr_kernel<R, false, kAccum><<<grid, kBlock, 0, stream>>>(
        pos, static_cast<R*>(forces), nbrs, offsets, counts, first, count, box, c, nullptr,
        nullptr);
    TDMD_LAUNCH_CHECK();
  }
  return TDMD_OK;
}

template <class R>
int eam_launch(const tdmd_eam_tables* t, const double* pos, void* forces, const int32_t* types,
               const int32_t* nbrs, const int32_t* offsets, const int32_t* counts, int natoms,
               const DevBox& box, double* rho, void* fpv, double* d_energy, void* work,
               hipStream_t stream) {
  using Q = typename Quad<R>::type;
  const int grid = grid_for(natoms);
  const double rc_sq = t->cutoff * t->cutoff;
  const auto* meta = static_cast<const SplineMeta*>(t->d_meta);
  const auto* coeff = static_cast<const Q*>(t->d_coeff);
  R* fp = static_cast<R*>(fpv);
  double* partials = static_cast<double*>(work);
  eam_density_csr_kernel<R><<<grid, kBlock, 0, stream>>>(pos, types, nbrs, offsets, counts,
                                                         natoms, t->ntypes, box, rc_sq, meta,
                                                         coeff, rho);
  TDMD_LAUNCH_CHECK();
  if (d_energy) {
    eam_embed_csr_kernel<R, true><<<grid, kBlock, 0, stream>>>(types, natoms, t->ntypes, meta,
                                                               coeff, rho, fp, partials, d_energy);
    TDMD_LAUNCH_CHECK();
    if (partials) {
      int rc = launch_sum_partials(partials, grid, d_energy, 1, stream);
      if (rc) return rc;
    }
    eam_force_csr_kernel<R, true><<<grid, kBlock, 0, stream>>>(
        pos, static_cast<R*>(forces), types, nbrs, offsets, counts, natoms, t->ntypes, box,
        rc_sq, meta, coeff, fp, partials, d_energy);
    TDMD_LAUNCH_CHECK();
    if (partials) return launch_sum_partials(partials, grid, d_energy, 1, stream);
  } else {
    eam_embed_csr_kernel<R, false><<<grid, kBlock, 0, stream>>>(types, natoms, t->ntypes, meta,
                                                                coeff, rho, fp, nullptr, nullptr);
    TDMD_LAUNCH_CHECK();
    eam_force_csr_kernel<R, false><<<grid, kBlock, 0, stream>>>(
        pos, static_cast<R*>(forces), types, nbrs, offsets, counts, natoms, t->ntypes, box,
        rc_sq, meta, coeff, fp, nullptr, nullptr);
    TDMD_LAUNCH_CHECK();
  }
  return TDMD_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free EAM compute: the per-atom densities, F'(rho) and the
// energy partials live in one file-scope device array instead of caller
// buffers. Its device address is looked up once per device
// (hipGetSymbolAddress) and kept; kernels get that address, never the
// array's host-side name. Every call on a device shares the array: calls
// that overlap on one device (several streams, several host threads) race.
// The kernels store every element they later read (rho, F'(rho), one
// partial per block), so nothing needs clearing between calls.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxAtoms = 1 << 20;
constexpr size_t kScratchRho = 0;                                             // double[N]
constexpr size_t kScratchFp = static_cast<size_t>(kNoAllocMaxAtoms) * 8;      // real[N]
constexpr size_t kScratchPartials = static_cast<size_t>(kNoAllocMaxAtoms) * 16;  // double[grid]
constexpr size_t kScratchBytes =
    kScratchPartials + static_cast<size_t>((kNoAllocMaxAtoms + kBlock - 1) / kBlock + 64) * 8;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};

int scratch_for_current_device(unsigned char** out) {
  int dev = 0;
  TDMD_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(TDMD_E_INVALID, "eam_compute_noalloc: device index");
  if (!g_scratch[dev]) {
    void* p = nullptr;
    TDMD_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(TDMD_E_HIP, "eam_compute_noalloc: no device address for the scratch array");
    g_scratch[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch[dev]);
  return TDMD_OK;
}

}  // namespace

}  // namespace tdmd_hip

using namespace tdmd_hip;

extern "C" {

int tdmd_morse_force(int32_t precision, const double* d_pos, void* d_forces,
                     const int32_t* d_neighbors, const int32_t* d_offsets,
                     const int32_t* d_counts, int32_t natoms, const tdmd_box* box,
                     const tdmd_morse_params* params, double* d_energy, void* d_work,
                     void* stream) {
  if (!box || !params || natoms < 0) return fail(TDMD_E_INVALID, "morse_force: bad arguments");
  if (natoms == 0) return TDMD_OK;  // device_morse.cu:141
  if (!d_pos || !d_forces || !d_offsets || !d_counts)
    return fail(TDMD_E_INVALID, "morse_force: null pointer");
  const DevBox db = make_devbox(*box);
  if (precision == TDMD_MIXED)
    return morse_launch<float, false>(d_pos, d_forces, d_neighbors, d_offsets, d_counts, 0,
                                      natoms, db, *params, d_energy, d_work, as_stream(stream));
  if (precision == TDMD_FP64)
    return morse_launch<double, false>(d_pos, d_forces, d_neighbors, d_offsets, d_counts, 0,
                                       natoms, db, *params, d_energy, d_work,
                                       as_stream(stream));
  return fail(TDMD_E_INVALID, "morse_force: bad precision");
}

int tdmd_morse_force_zone(int32_t precision, const double* d_pos, void* d_forces,
                          const int32_t* d_neighbors, const int32_t* d_offsets,
                          const int32_t* d_counts, int32_t first_atom, int32_t atom_count,
                          const tdmd_box* box, const tdmd_morse_params* params,
                          double* d_energy, void* d_work, void* stream) {
  if (!box || !params || first_atom < 0 || atom_count < 0)
    return fail(TDMD_E_INVALID, "morse_force_zone: bad arguments");
  if (atom_count == 0) return TDMD_OK;  // device_morse_zone.cu:126
  if (!d_pos || !d_forces || !d_offsets || !d_counts)
    return fail(TDMD_E_INVALID, "morse_force_zone: null pointer");
  const DevBox db = make_devbox(*box);
  if (precision == TDMD_MIXED)
    return morse_launch<float, true>(d_pos, d_forces, d_neighbors, d_offsets, d_counts,
                                     first_atom, atom_count, db, *params, d_energy, d_work,
                                     as_stream(stream));
  if (precision == TDMD_FP64)
    return morse_launch<double, true>(d_pos, d_forces, d_neighbors, d_offsets, d_counts,
                                      first_atom, atom_count, db, *params, d_energy, d_work,
                                      as_stream(stream));
  return fail(TDMD_E_INVALID, "morse_force_zone: bad precision");
}

int tdmd_eam_compute(const tdmd_eam_tables* t, const double* d_pos, void* d_forces,
                     const int32_t* d_types, const int32_t* d_neighbors,
                     const int32_t* d_offsets, const int32_t* d_counts, int32_t natoms,
                     const tdmd_box* box, double* d_rho, void* d_fp, double* d_energy,
                     void* d_work, void* stream) {
  if (!t || !box || natoms < 0) return fail(TDMD_E_INVALID, "eam_compute: bad arguments");
  if (natoms == 0) return TDMD_OK;  // device_eam.cu:345
  if (!d_pos || !d_forces || !d_types || !d_offsets || !d_counts || !d_rho || !d_fp)
    return fail(TDMD_E_INVALID, "eam_compute: null pointer");
  if (t->ntypes > 8) return fail(TDMD_E_INVALID, "eam_compute: at most 8 element types");
  const DevBox db = make_devbox(*box);
  if (t->precision == TDMD_MIXED)
    return eam_launch<float>(t, d_pos, d_forces, d_types, d_neighbors, d_offsets, d_counts,
                             natoms, db, d_rho, d_fp, d_energy, d_work, as_stream(stream));
  return eam_launch<double>(t, d_pos, d_forces, d_types, d_neighbors, d_offsets, d_counts,
                            natoms, db, d_rho, d_fp, d_energy, d_work, as_stream(stream));
}

int tdmd_eam_compute_noalloc(const tdmd_eam_tables* t, const double* d_pos, void* d_forces,
                             const int32_t* d_types, const int32_t* d_neighbors,
                             const int32_t* d_offsets, const int32_t* d_counts, int32_t natoms,
                             const tdmd_box* box, double* d_energy, void* stream) {
  if (!t || !box || natoms < 0) return fail(TDMD_E_INVALID, "eam_compute_noalloc: bad arguments");
  if (natoms > kNoAllocMaxAtoms)
    return fail(TDMD_E_INVALID, "eam_compute_noalloc: " + std::to_string(natoms) +
                                    " atoms exceed the scratch array's " +
                                    std::to_string(kNoAllocMaxAtoms));
  if (natoms == 0) return TDMD_OK;
  if (!d_pos || !d_forces || !d_types || !d_offsets || !d_counts)
    return fail(TDMD_E_INVALID, "eam_compute_noalloc: null pointer");
  if (t->ntypes > 8) return fail(TDMD_E_INVALID, "eam_compute_noalloc: at most 8 element types");
  unsigned char* base = nullptr;
  const int rc = scratch_for_current_device(&base);
  if (rc) return rc;
  double* rho = reinterpret_cast<double*>(base + kScratchRho);
  void* fp = base + kScratchFp;
  void* work = base + kScratchPartials;
  const DevBox db = make_devbox(*box);
  if (t->precision == TDMD_MIXED)
    return eam_launch<float>(t, d_pos, d_forces, d_types, d_neighbors, d_offsets, d_counts,
                             natoms, db, rho, fp, d_energy, work, as_stream(stream));
  return eam_launch<double>(t, d_pos, d_forces, d_types, d_neighbors, d_offsets, d_counts,
                            natoms, db, rho, fp, d_energy, work, as_stream(stream));
}

}  // extern "C"
