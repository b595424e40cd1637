// sched_lds.inl -- single-element EAM passes with the spline tables staged
// in LDS and a software-pipelined neighbour loop.
//
// One 1024-thread workgroup per tile of 1024 (spatially sorted) atoms. At
// entry the workgroup copies a window of spline intervals -- the last `nwin`
// intervals of the r grid, i.e. [r_lo, rc) -- from the coefficient table into
// LDS; a pair closer than r_lo reads the interval from global memory instead
// (same coefficients, same arithmetic). The window is sized to the LDS
// budget: for the 5000-point tables of the benchmark the whole table fits
// in mixed precision and the FP64 window starts near 1.8 A, below any
// distance a solid at a few hundred kelvin reaches. kWin = false is the
// specialisation for a window that holds every interval (no global path).
//
// The neighbour loop is for_neighbours (sched_kernels.inl), software
// pipelined in groups of kU slots.
//
// kStage: the workgroup also copies its block's halo (k_halo) into LDS,
// behind the table window, and gathers neighbour records from there through
// the 16-bit rows; a block whose halo did not fit reads the 32-bit rows and
// global records as without kStage. The window then gets what the halo
// (stage_cap records) leaves of the LDS budget.
//
// Force window entry: {phi b, phi c, phi d, rho b} + {rho c, rho d}
// (+ phi a when the energy is wanted); density window entry: rho {a,b,c,d}.
// For one element phi_00, rho_0 share the r grid (setfl: one Nr, dr).
#pragma once

namespace tdmd_hip {
namespace sched {

constexpr int kBigBlock = 1024;

struct LdsWindow {
  int k_lo;  // first interval held in LDS
  int nwin;  // intervals held (k_lo + nwin = n - 1)
};

template <class R>
struct Vec2;
template <>
struct Vec2<float> {
  using type = float2;
};
template <>
struct Vec2<double> {
  using type = double2;
};

// Same expressions as spline_value / spline_deriv (pair_math.hpp).
template <class R>
__device__ __forceinline__ R cubic_value(R a, R b, R c, R d, R dx) {
  return a + dx * (b + dx * (c + dx * d));
}
template <class R>
__device__ __forceinline__ R cubic_deriv(R b, R c, R d, R dx) {
  return b + dx * (static_cast<R>(2) * c + dx * static_cast<R>(3) * d);
}

__device__ __forceinline__ int big_block_index(int nblocks) {
  const int b = blockIdx.x;
  const int per = nblocks >> 3;
  return (b & 7) * per + (b >> 3);
}

// Copy the halo records of block blk into LDS; returns the block's mode
// (0: staged). The caller synchronises the workgroup before use.
template <class T>
__device__ __forceinline__ int stage_halo(const StepArgs& a, int blk, const T* __restrict__ rec,
                                          T* hrec) {
  const int mode = a.hmode[blk];
  if (mode == 0) {
    const int H = a.hn[blk];
    const int32_t* __restrict__ hj = a.hj + static_cast<size_t>(blk) * kHaloMax;
    for (int t = threadIdx.x; t < H; t += kBigBlock) hrec[t] = rec[hj[t]];
  }
  return mode;
}

// LDS bytes of a table window, rounded up so the halo behind it is 16-byte aligned.
__host__ __device__ constexpr size_t window_bytes(int nwin, size_t per) {
  return (static_cast<size_t>(nwin) * per + 15) & ~static_cast<size_t>(15);
}

template <class R, class Mir, bool kEnergy, bool kWin, int kU, bool kStage>
__global__ void __launch_bounds__(kBigBlock)
k_eam_density_lds(StepArgs a, const SplineMeta* __restrict__ meta,
                  const typename Quad<R>::type* __restrict__ coeff, double rc_sq, LdsWindow w,
                  int nblocks) {
  using T = typename Mir::T;
  using Q = typename Quad<R>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  Q* tab = reinterpret_cast<Q*>(smem);
  const SplineMeta mF = meta[0], mrho = meta[1];
  for (int q = threadIdx.x; q < w.nwin; q += kBigBlock) tab[q] = coeff[mrho.offset + w.k_lo + q];
  const int blk = big_block_index(nblocks);
  const T* __restrict__ M = static_cast<const T*>(a.M);
  T* hrec = reinterpret_cast<T*>(smem + window_bytes(w.nwin, sizeof(Q)));
  const int mode = kStage ? stage_halo<T>(a, blk, M, hrec) : 1;
  __syncthreads();
  const int i = blk * kBigBlock + threadIdx.x;
  double e = 0.0;
  if (i < a.natoms) {
    const T mi = M[i];
    const MirrorParams& mpi = mp_of<Mir>(a, i);
    const int n = a.cnt[i];
    double rho = 0.0;
    auto pair = [&](int, const T& mj, bool in_row) {
      double dx, dy, dz;
      const double r2 = Mir::delta(mi, mj, mpi, &dx, &dy, &dz);
      if (in_row && r2 < rc_sq) {
        const PairR<R> pr = pair_r<R>(r2);
        R t;
        const int idx = spline_locate<R>(mrho, pr.r, &t);
        const int q = idx - w.k_lo;
        const Q c = (!kWin || q >= 0) ? tab[q] : coeff[mrho.offset + idx];
        rho += static_cast<double>(cubic_value<R>(c.x, c.y, c.z, c.w, t));
      }
    };
    if (kStage && mode == 0)
      for_neighbours_staged<T, kU>(a, i, n, hrec, pair);
    else
      for_neighbours<T, kU>(a, i, n, M, pair);
    const SplineVal<R> F = spline_both<R>(mF, coeff, rho);
    static_cast<T*>(a.M2)[i] = Mir::with_w(mi, Mir::template pack_real<R>(F.df));
    e = static_cast<double>(F.f);
  }
  if (kEnergy) emit_partial<kBigBlock>(e, a.partials, blk);
}

template <class R, class Mir, int kEpi, bool kEnergy, bool kWin, int kU, bool kStage>
__global__ void __launch_bounds__(kBigBlock)
k_eam_force_lds(StepArgs a, const SplineMeta* __restrict__ meta,
                const typename Quad<R>::type* __restrict__ coeff, double rc_sq, LdsWindow w,
                int nblocks) {
  using T = typename Mir::T;
  using Q = typename Quad<R>::type;
  using V2 = typename Vec2<R>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  Q* tA = reinterpret_cast<Q*>(smem);              // {phi b, phi c, phi d, rho b}
  V2* tB = reinterpret_cast<V2*>(tA + w.nwin);     // {rho c, rho d}
  R* tE = reinterpret_cast<R*>(tB + w.nwin);       // phi a (energy only)
  const SplineMeta mrho = meta[1], mphi = meta[2];
  for (int q = threadIdx.x; q < w.nwin; q += kBigBlock) {
    const Q cp = coeff[mphi.offset + w.k_lo + q];
    const Q cr = coeff[mrho.offset + w.k_lo + q];
    Q x;
    x.x = cp.y;
    x.y = cp.z;
    x.z = cp.w;
    x.w = cr.y;
    V2 y;
    y.x = cr.z;
    y.y = cr.w;
    tA[q] = x;
    tB[q] = y;
    if (kEnergy) tE[q] = cp.x;
  }
  const int blk = big_block_index(nblocks);
  const T* __restrict__ M2 = static_cast<const T*>(a.M2);
  T* hrec = reinterpret_cast<T*>(
      smem + window_bytes(w.nwin, sizeof(Q) + sizeof(V2) + (kEnergy ? sizeof(R) : 0)));
  const int mode = kStage ? stage_halo<T>(a, blk, M2, hrec) : 1;
  __syncthreads();
  const int i = blk * kBigBlock + threadIdx.x;
  double red = 0.0;
  if (i < a.natoms) {
    const T mi = M2[i];
    const MirrorParams& mpi = mp_of<Mir>(a, i);
    const R fpi = Mir::template w_real<R>(mi);
    const int n = a.cnt[i];
    double fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0;
    auto pair = [&](int, const T& mj, bool in_row) {
      double dx, dy, dz;
      const double r2 = Mir::delta(mi, mj, mpi, &dx, &dy, &dz);
      if (in_row && r2 < rc_sq) {
        const PairR<R> pr = pair_r<R>(r2);
        R t;
        const int idx = spline_locate<R>(mphi, pr.r, &t);
        const int q = idx - w.k_lo;
        // every lane reads the window (below it: its first interval); the
        // rare pairs below the window then reload from the global table
        // under a wave-uniform branch, which keeps the next group's loads
        // in flight when no lane needs it
        const int qs = (!kWin || q >= 0) ? q : 0;
        const Q x = tA[qs];
        const V2 y = tB[qs];
        R dphi = cubic_deriv<R>(x.x, x.y, x.z, t);
        R drho = cubic_deriv<R>(x.w, y.x, y.y, t);
        R ephi = kEnergy ? cubic_value<R>(tE[qs], x.x, x.y, x.z, t) : static_cast<R>(0);
        if (kWin && __any(q < 0)) {
          if (q < 0) {
            const Q cp = coeff[mphi.offset + idx];
            const Q cr = coeff[mrho.offset + idx];
            dphi = cubic_deriv<R>(cp.y, cp.z, cp.w, t);
            drho = cubic_deriv<R>(cr.y, cr.z, cr.w, t);
            if (kEnergy) ephi = cubic_value<R>(cp.x, cp.y, cp.z, cp.w, t);
          }
        }
        if (kEnergy) pe += 0.5 * static_cast<double>(ephi);
        const R fpj = Mir::template w_real<R>(mj);
        const double fpair = static_cast<double>(-(dphi + fpi * drho + fpj * drho) * pr.rinv);
        fx = fma(fpair, dx, fx);
        fy = fma(fpair, dy, fy);
        fz = fma(fpair, dz, fz);
      }
    };
    if (kStage && mode == 0)
      for_neighbours_staged<T, kU>(a, i, n, hrec, pair);
    else
      for_neighbours<T, kU>(a, i, n, M2, pair);
    const double ke = epilogue<R, Mir, kEpi>(a, i, fx, fy, fz);
    red = kEnergy ? pe : ke;
  }
  if (kEnergy || kEpi == kKick2Ke) emit_partial<kBigBlock>(red, a.partials, blk);
}

}  // namespace sched
}  // namespace tdmd_hip
