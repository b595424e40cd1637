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
// kIso (fixed-point mirrors whose three axes share one unit u, e.g. cubic
// periodic boxes): pair differences stay in units, r^2 / u^2 is tested
// against rc^2 / u^2 and forces are summed in units and scaled once, which
// saves three FP64 products per neighbour slot.
//
// kStage: the workgroup also copies its block's halo (k_halo) into LDS,
// behind the table window, and gathers neighbour records from there through
// the 16-bit rows; a block whose halo did not fit reads the 32-bit rows and
// global records as without kStage. The window then gets what the halo
// (stage_cap records) leaves of the LDS budget.
//
// Force window entry: {phi b, phi c, phi d, rho b} + {rho c, rho d}
// (+ phi a when the energy is wanted; in the mixed force-only window 2c and
// 3d in place of c and d); density window entry: rho {a,b,c,d}.
// For one element phi_00, rho_0 share the r grid (setfl: one Nr, dr).
#pragma once

namespace tdmd_hip {
namespace sched {

constexpr int kBigBlock = 1024;
constexpr int kStageCap = 4608;  // halo records a staged pass holds in LDS (72 KiB of 16 B)

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

// Persistent block loop of the LDS passes. The grid holds a multiple of 8
// workgroups, at most nblocks (itself a multiple of 8). Workgroup g serves
// XCD x = g % 8, whose blocks are the contiguous run [x per, (x+1) per),
// per = nblocks / 8, and takes every (grid / 8)-th block of that run from
// the (g / 8)-th: neighbouring blocks, and so their halos, stay in one L2.
// The spline window is copied into LDS once per workgroup.
//
// With halo staging the copy of a block's halo is pipelined across the
// workgroup's blocks: while block b is evaluated, the records of its next
// block are being gathered into registers and the halo indices of the block
// after that are loaded, so a block starts with LDS stores only.
template <class T>
struct HaloPipe {
  static constexpr int kPer = (kStageCap + kBigBlock - 1) / kBigBlock;
  int j[kPer];  // halo indices of the block after next (-1: none)
  T v[kPer];    // halo records of the next block
  int jn = 0, jmode = 1;  // halo size and mode of the indices' block
  int vn = 0, vmode = 1;  // and of the records' block

  __device__ __forceinline__ void load_indices(const StepArgs& a, int blk) {
    jmode = a.hmode[blk];
    jn = jmode == 0 ? a.hn[blk] : 0;  // <= stage_cap <= kStageCap when staged
    const int32_t* __restrict__ hj = a.hj + static_cast<size_t>(blk) * kHaloMax;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int t = static_cast<int>(threadIdx.x) + u * kBigBlock;
      j[u] = t < jn ? hj[t] : -1;
    }
  }
  __device__ __forceinline__ void none() {
    jn = 0;
    jmode = 1;
#pragma unroll
    for (int u = 0; u < kPer; ++u) j[u] = -1;
  }
  __device__ __forceinline__ void gather_records(const T* __restrict__ rec) {
#pragma unroll
    for (int u = 0; u < kPer; ++u)
      if (j[u] >= 0) v[u] = gather(rec, j[u]);
    vn = jn;
    vmode = jmode;
  }
  __device__ __forceinline__ void store(T* hrec) const {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int t = static_cast<int>(threadIdx.x) + u * kBigBlock;
      if (t < vn) hrec[t] = v[u];
    }
  }
};

// Block sequence of this workgroup (see above).
struct BlockSeq {
  int base, first, per, step;
  __device__ __forceinline__ explicit BlockSeq(int nblocks) {
    per = nblocks >> 3;
    base = (static_cast<int>(blockIdx.x) & 7) * per;
    first = static_cast<int>(blockIdx.x) >> 3;
    step = static_cast<int>(gridDim.x) >> 3;
  }
};

// Prologue of the halo pipeline: records of the first block, indices of
// the second.
template <class T>
__device__ __forceinline__ void halo_prologue(const StepArgs& a, const BlockSeq& q,
                                              const T* __restrict__ rec, HaloPipe<T>& pipe) {
  pipe.load_indices(a, q.base + q.first);
  pipe.gather_records(rec);
  if (q.first + q.step < q.per)
    pipe.load_indices(a, q.base + q.first + q.step);
  else
    pipe.none();
}

// Top of block mm: store its halo, start the next block's gathers and the
// following block's index loads; returns the block's mode (0: staged).
template <class T>
__device__ __forceinline__ int halo_advance(const StepArgs& a, const BlockSeq& q, int mm,
                                            const T* __restrict__ rec, T* hrec, HaloPipe<T>& pipe) {
  const int mode = pipe.vmode;
  pipe.store(hrec);
  if (mm + q.step < q.per) {
    pipe.gather_records(rec);
    if (mm + 2 * q.step < q.per)
      pipe.load_indices(a, q.base + mm + 2 * q.step);
    else
      pipe.none();
  }
  return mode;
}

// LDS bytes of a table window, rounded up so the halo behind it is 16-byte aligned.
__host__ __device__ constexpr size_t window_bytes(int nwin, size_t per) {
  return (static_cast<size_t>(nwin) * per + 15) & ~static_cast<size_t>(15);
}

// Difference of two records (i minus j) and its squared length: in A, or in
// fixed-point units when kIso.
template <class Mir, bool kIso, class T>
__device__ __forceinline__ double pair_delta(const T& a, const T& b, const MirrorParams& mp,
                                             double* dx, double* dy, double* dz) {
  if constexpr (kIso) {
    *dx = static_cast<double>(Mir::wdiff(a.x, b.x));
    *dy = static_cast<double>(Mir::wdiff(a.y, b.y));
    *dz = static_cast<double>(Mir::wdiff(a.z, b.z));
    return *dx * *dx + *dy * *dy + *dz * *dz;
  } else {
    return Mir::delta(a, b, mp, dx, dy, dz);
  }
}

template <class R, class Mir, bool kEnergy, bool kWin, int kU, bool kStage, bool kIso>
__global__ void __launch_bounds__(kBigBlock)
k_eam_density_lds(StepArgs a, const SplineMeta* __restrict__ meta,
                  const typename Quad<R>::type* __restrict__ coeff, double rc_sq, LdsWindow w,
                  int nblocks) {
  using T = typename Mir::T;
  using Q = typename Quad<R>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  Q* tab = reinterpret_cast<Q*>(smem);
  pass_clock_begin(a.pclock);
  const SplineMeta mF = meta[0], mrho = meta[1];
  const SplineMetaF mrho_f = meta_f32(mrho);
  for (int q = threadIdx.x; q < w.nwin; q += kBigBlock) tab[q] = coeff[mrho.offset + w.k_lo + q];
  const T* __restrict__ M = static_cast<const T*>(a.M);
  T* hrec = reinterpret_cast<T*>(smem + window_bytes(w.nwin, sizeof(Q)));
  const BlockSeq seq(nblocks);
  HaloPipe<T> pipe;
  if (kStage) halo_prologue<T>(a, seq, M, pipe);
  for (int mm = seq.first; mm < seq.per; mm += seq.step) {
  const int blk = seq.base + mm;
  const int mode = kStage ? halo_advance<T>(a, seq, mm, M, hrec, pipe) : 1;
  __syncthreads();  // window (first block) and halo in LDS
  const int i = blk * kBigBlock + threadIdx.x;
  double e = 0.0;
  if (i < a.natoms) {
    const T mi = M[i];
    const MirrorParams& mpi = mp_of<Mir>(a, i);
    const int n = a.cnt[i];
    double rho = 0.0;
    const double u2 = kIso ? mpi.unit[0] * mpi.unit[0] : 1.0;
    const double rcut2 = kIso ? rc_sq / u2 : rc_sq;
    // staged blocks: skin-shell slots inside the cutoff are listed for the
    // force pass (ELL16 layout: the next slot is one element on, or
    // 64 x 8 - 7 on at the end of a group of eight)
    const bool cut = kStage && mode == 0;
    const int nn = cut ? a.cntn[i] : 0;
    uint16_t* cp = a.nbrc16 + (cut ? ell16(i, 0, a.cap) : 0);
    int c = 0;
    auto pair = [&](int l, const T& mj, bool in_row, int slot) {
      double dx, dy, dz;
      const double s2 = pair_delta<Mir, kIso>(mi, mj, mpi, &dx, &dy, &dz);
      if (in_row && s2 < rcut2) {
        R t;
        int idx;
        if constexpr (std::is_same_v<R, float>) {
          idx = locate_f32(mrho_f, kIso ? s2 * u2 : s2, &t);
        } else {
          idx = spline_locate<R>(mrho, pair_r<R>(kIso ? s2 * u2 : s2).r, &t);
        }
        const int q = idx - w.k_lo;
        const Q cf = (!kWin || q >= 0) ? tab[q] : coeff[mrho.offset + idx];
        rho += static_cast<double>(cubic_value<R>(cf.x, cf.y, cf.z, cf.w, t));
        if (cut && slot >= nn) {
          *cp = static_cast<uint16_t>(l);
          cp += (c & 7) == 7 ? 505 : 1;
          ++c;
        }
      }
    };
    if (kStage && mode == 0) {
      for_neighbours_staged<T, kU>(a, a.nbr16, i, n, hrec, pair);
      a.cntc[i] = c;
    } else {
      for_neighbours<T, kU>(a, i, n, M, [&](int j, const T& mj, bool r) { pair(j, mj, r, 0); });
    }
    const SplineVal<R> F = spline_both<R>(mF, coeff, rho);
    static_cast<T*>(a.M2)[i] = Mir::with_w(mi, Mir::template pack_real<R>(F.df));
    e = static_cast<double>(F.f);
  }
  if (kEnergy) emit_partial<kBigBlock>(e, a.partials, blk);
  __syncthreads();  // the next block's halo overwrites this one
  }
  pass_clock_end(a.pclock);
}

template <class R, class Mir, int kEpi, bool kEnergy, bool kWin, int kU, bool kStage,
          bool kIso>
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
  // mixed: pairs located by pair_t_f32, derivatives from b, 2c, 3d; the
  // force-only window holds 2c and 3d (the energy window c and d, scaled
  // at use: the same values)
  constexpr bool kF32 = std::is_same_v<R, float>;
  constexpr bool kPre = kF32 && !kEnergy;
  unsigned long long* const pclk = a.pclock ? a.pclock + 4 : nullptr;
  pass_clock_begin(pclk);
  const SplineMeta mrho = meta[1], mphi = meta[2];
  const ForceLocF floc = force_loc_f32(mphi);
  for (int q = threadIdx.x; q < w.nwin; q += kBigBlock) {
    const Q cp = coeff[mphi.offset + w.k_lo + q];
    const Q cr = coeff[mrho.offset + w.k_lo + q];
    Q x;
    x.x = cp.y;
    x.y = kPre ? static_cast<R>(2) * cp.z : cp.z;
    x.z = kPre ? static_cast<R>(3) * cp.w : cp.w;
    x.w = cr.y;
    V2 y;
    y.x = kPre ? static_cast<R>(2) * cr.z : cr.z;
    y.y = kPre ? static_cast<R>(3) * cr.w : cr.w;
    tA[q] = x;
    tB[q] = y;
    if (kEnergy) tE[q] = cp.x;
  }
  auto deriv = [](R b, R c, R d, R t) -> R {
    if constexpr (kPre)
      return cubic_deriv_s(b, c, d, t);
    else if constexpr (kF32)
      return cubic_deriv_s(b, 2.0f * c, 3.0f * d, t);
    else
      return cubic_deriv<R>(b, c, d, t);
  };
  const T* __restrict__ M2 = static_cast<const T*>(a.M2);
  T* hrec = reinterpret_cast<T*>(
      smem + window_bytes(w.nwin, sizeof(Q) + sizeof(V2) + (kEnergy ? sizeof(R) : 0)));
  const BlockSeq seq(nblocks);
  HaloPipe<T> pipe;
  if (kStage) halo_prologue<T>(a, seq, M2, pipe);
  for (int mm = seq.first; mm < seq.per; mm += seq.step) {
  const int blk = seq.base + mm;
  const int mode = kStage ? halo_advance<T>(a, seq, mm, M2, hrec, pipe) : 1;
  __syncthreads();  // window (first block) and halo in LDS
  const int i = blk * kBigBlock + threadIdx.x;
  double red = 0.0, wsum = 0.0;
  if (i < a.natoms) {
    const T mi = M2[i];
    const MirrorParams& mpi = mp_of<Mir>(a, i);
    const R fpi = Mir::template w_real<R>(mi);
    const bool cut = kStage && mode == 0;
    const int n = cut ? a.cntn[i] : a.cnt[i];
    double fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0, vir = 0.0;
    const double u = kIso ? mpi.unit[0] : 1.0;
    const double u2 = u * u;
    const double rcut2 = kIso ? rc_sq / u2 : rc_sq;
    auto pair = [&](int, const T& mj, bool in_row, int) {
      double dx, dy, dz;
      const double s2 = pair_delta<Mir, kIso>(mi, mj, mpi, &dx, &dy, &dz);
      if (in_row && s2 < rcut2) {
        const double r2 = kIso ? s2 * u2 : s2;
        R t, rinv;
        int idx, qs;
        bool below;
        // every lane reads the window (below it: its first interval); in
        // the rare wave with a pair below the window every lane then
        // re-reads its interval from the global table (the same values),
        // under a wave-uniform branch, which keeps the next group's loads
        // in flight when no lane needs it
        if constexpr (kF32) {
          const PairT p = pair_t_f32(floc, r2, kWin ? static_cast<float>(w.k_lo) : 0.0f);
          t = p.t;
          rinv = p.rinv;
          idx = p.idx;
          below = kWin && p.below;
          qs = idx - w.k_lo;  // idx >= k_lo
        } else {
          const PairR<R> pr = pair_r<R>(r2);
          idx = spline_locate<R>(mphi, pr.r, &t);
          rinv = pr.rinv;
          const int q = idx - w.k_lo;
          below = kWin && q < 0;
          qs = below ? 0 : q;
        }
        const Q x = tA[qs];
        const V2 y = tB[qs];
        R dphi = deriv(x.x, x.y, x.z, t);
        R drho = deriv(x.w, y.x, y.y, t);
        R ephi = kEnergy ? cubic_value<R>(tE[qs], x.x, x.y, x.z, t) : static_cast<R>(0);
        if (kWin && __builtin_amdgcn_ballot_w64(below) != 0) {
          if constexpr (kF32) {  // the interval unclamped by the window
            const PairT p = pair_t_f32(floc, r2);
            t = p.t;
            idx = p.idx;
          }
          const Q cp = coeff[mphi.offset + idx];  // c and d
          const Q cr = coeff[mrho.offset + idx];
          if constexpr (kF32) {
            dphi = cubic_deriv_f(cp, t);
            drho = cubic_deriv_f(cr, t);
          } else {
            dphi = cubic_deriv<R>(cp.y, cp.z, cp.w, t);
            drho = cubic_deriv<R>(cr.y, cr.z, cr.w, t);
          }
          if (kEnergy) ephi = cubic_value<R>(cp.x, cp.y, cp.z, cp.w, t);
        }
        if (kEnergy) pe += 0.5 * static_cast<double>(ephi);
        const R fpj = Mir::template w_real<R>(mj);
        const double fpair = static_cast<double>(-(dphi + fpi * drho + fpj * drho) * rinv);
        if (kEnergy) vir += 0.5 * fpair * r2;
        fx = fma(fpair, dx, fx);
        fy = fma(fpair, dy, fy);
        fz = fma(fpair, dz, fz);
      }
    };
    if (cut) {
      // the near slots, then the skin-shell slots the density pass found
      // inside the cutoff: the row's pairs in row order
      for_neighbours_staged<T, kU>(a, a.nbr16, i, n, hrec, pair);
      for_neighbours_staged<T, kU, true>(a, a.nbrc16, i, a.cntc[i], hrec, pair);
    } else {
      for_neighbours<T, kU>(a, i, n, M2, [&](int j, const T& mj, bool r) { pair(j, mj, r, 0); });
    }
    if constexpr (kIso) {  // sums in units
      fx *= u;
      fy *= u;
      fz *= u;
    }
    const double ke = epilogue<R, Mir, kEpi>(a, i, fx, fy, fz);
    red = kEnergy ? pe : ke;
    wsum = vir;
  }
  if (kEnergy || emits_ke(kEpi)) emit_partial<kBigBlock>(red, a.partials, blk);
  if (kEnergy) emit_partial<kBigBlock>(wsum, a.partials_w, blk);
  __syncthreads();  // the next block's halo overwrites this one
  }
  pass_clock_end(pclk);
}

}  // namespace sched
}  // namespace tdmd_hip
