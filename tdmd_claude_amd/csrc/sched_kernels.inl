// sched_kernels.inl -- kernels of the FastPipelineScheduler fast path
// (src/scheduler/fast_pipeline_scheduler.cu:87-232 re-designed for gfx950).
//
// Per NVE step (the reference's kick-drift / force pair,
// fast_pipeline_scheduler.cu:101-166):
//   [rebuild: bin, scan, scatter, order, permute, copy-back, list]
//   EAM density pass        M  -> M2 (rho, F'(rho) fused: reference passes
//                                 1+2, device_eam.cu:72-162)
//   force pass + epilogue   M2 -> F, then the cross-step kick and drift of
//                                 the next step in registers (the
//                                 reference's fused_full_kick_drift,
//                                 device_velocity_verlet.cu:88-130), which
//                                 writes the next M and tests the skin.
// The density pass reads M and writes M2; the force pass reads M2 and writes
// M, so no pass reads a record another thread of the same pass writes.
#pragma once

#include "nhc.hpp"
#include "sched_kernels.hpp"

namespace tdmd_hip {
namespace sched {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wrap1(double x, double lo, double L, int periodic) {
  // device_velocity_verlet.cu:49-61
  if (periodic) {
    if (x < lo)
      x += L;
    else if (x >= lo + L)
      x -= L;
  }
  return x;
}

template <class Mir>
__device__ __forceinline__ typename Mir::T encode_type(const MirrorParams& mp, double x,
                                                       double y, double z, int ti) {
  return Mir::encode(mp, x, y, z, ti);
}

template <class R>
__device__ __forceinline__ R* rp(void* p) {
  return static_cast<R*>(p);
}

template <int BLOCK = kBlock>
__device__ __forceinline__ void emit_partial(double v, double* partials, int slot) {
  const double t = block_sum<BLOCK>(v);
  if (threadIdx.x == 0) partials[slot] = t;
}

// Kick-and-drift for atom i with a given kick factor and an optional pre
// scale of v (NHC). Writes x, v and the gather record (into `rec`: M, or the
// Morse pass's other record buffer); sets the rebuild flag when the atom has
// moved more than skin/2 since the last build (minimum image).
// Streaming accesses of the integrator (positions, velocities, forces, build
// positions, types): each element is touched once per step, so they are
// loaded and stored non-temporally and leave the caches (L2, and MALL in
// front of HBM) to the neighbour rows and gather records that both passes
// re-read every step (1M-atom NVT step 286.6 -> 280.0 us).
template <class T>
__device__ __forceinline__ T ntl(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ void nts(T v, T* p) {
  __builtin_nontemporal_store(v, p);
}

// kKick2: the stored v lacks the previous step's closing half kick (its
// force pass only summed the KE of the kicked velocity, epilogue kKe); it is
// applied here first, as v + (dt/2) F / (m mvv2e), the epilogue's expression.
template <class R, class Mir, bool kKick2 = false>
__device__ __forceinline__ void kick_drift(const StepArgs& a, int i, double vscale,
                                           double kick_dt, double f0, double f1, double f2,
                                           void* rec) {
  const int ti = a.one_type ? 0 : ntl(&a.types[i]) - 1;
  const double mass = a.masses[ti + 1];
  double ux = ntl(&a.vx[i]), uy = ntl(&a.vy[i]), uz = ntl(&a.vz[i]);
  if constexpr (kKick2) {
    const double half = (0.5 * a.dt) / (mass * kMvv2e);
    ux = __fma_rn(half, f0, ux);
    uy = __fma_rn(half, f1, uy);
    uz = __fma_rn(half, f2, uz);
  }
  // explicit fused operations: the same roundings whether the half kick
  // above was applied here or by the force pass (kick 2 stored)
  const double factor = kick_dt / (mass * kMvv2e);
  const double vx = __fma_rn(factor, f0, vscale * ux);
  const double vy = __fma_rn(factor, f1, vscale * uy);
  const double vz = __fma_rn(factor, f2, vscale * uz);
  const MirrorParams& mp = mp_of<Mir>(a, i);
  const double x = wrap1(ntl(&a.px[i]) + a.dt * vx, mp.lo[0], mp.L[0], mp.periodic[0]);
  const double y = wrap1(ntl(&a.py[i]) + a.dt * vy, mp.lo[1], mp.L[1], mp.periodic[1]);
  const double z = wrap1(ntl(&a.pz[i]) + a.dt * vz, mp.lo[2], mp.L[2], mp.periodic[2]);
  nts(vx, &a.vx[i]);
  nts(vy, &a.vy[i]);
  nts(vz, &a.vz[i]);
  nts(x, &a.px[i]);
  nts(y, &a.py[i]);
  nts(z, &a.pz[i]);
  const typename Mir::T m = encode_type<Mir>(mp, x, y, z, ti);
  static_cast<typename Mir::T*>(rec)[i] = m;
  double dx, dy, dz;
  if constexpr (Mir::kFixed) {
    // displacement since the build from the 32-bit coordinates: the
    // wrapping difference is the minimum image on a periodic axis
    dx = static_cast<double>(Mir::wdiff(m.x, ntl(&reinterpret_cast<const int32_t*>(a.bx)[i]))) * mp.unit[0];
    dy = static_cast<double>(Mir::wdiff(m.y, ntl(&reinterpret_cast<const int32_t*>(a.by)[i]))) * mp.unit[1];
    dz = static_cast<double>(Mir::wdiff(m.z, ntl(&reinterpret_cast<const int32_t*>(a.bz)[i]))) * mp.unit[2];
  } else {
    dx = min_image(x - a.bx[i], mp.L[0], mp.hL[0], mp.periodic[0]);
    dy = min_image(y - a.by[i], mp.L[1], mp.hL[1], mp.periodic[1]);
    dz = min_image(z - a.bz[i], mp.L[2], mp.hL[2], mp.periodic[2]);
  }
  if (dx * dx + dy * dy + dz * dz > a.skin_half_sq) a.flags[0] = 1;
  if (!(isfinite(x) && isfinite(y) && isfinite(z) && Mir::representable(mp, x, y, z)))
    a.flags[2] = 1;
}

// Force-pass epilogue. The full kick-drift writes the next step's records
// into `rec` (M for the EAM passes, which gather M2; the other buffer of the
// pair for the Morse pass, which gathers M).
template <class R, class Mir, int kEpi>
__device__ __forceinline__ double epilogue(const StepArgs& a, int i, double fx, double fy,
                                           double fz, void* rec = nullptr) {
  const R f0 = static_cast<R>(fx), f1 = static_cast<R>(fy), f2 = static_cast<R>(fz);
  if (kEpi == kFullKickDrift) {
    kick_drift<R, Mir>(a, i, 1.0, a.dt, static_cast<double>(f0), static_cast<double>(f1),
                       static_cast<double>(f2), rec ? rec : a.M);
    return 0.0;
  }
  nts(f0, &rp<R>(a.fx)[i]);
  nts(f1, &rp<R>(a.fy)[i]);
  nts(f2, &rp<R>(a.fz)[i]);
  if (kEpi == kForceOnly) return 0.0;
  // kHalfKick / kKick2Ke: v += (dt/2) F / (m mvv2e)  (half_kick_kernel,
  // device_velocity_verlet.cu:14-32)
  const int ti = a.one_type ? 0 : ntl(&a.types[i]) - 1;
  const double mass = a.masses[ti + 1];
  const double factor = (0.5 * a.dt) / (mass * kMvv2e);
  const double vx = __fma_rn(factor, static_cast<double>(f0), ntl(&a.vx[i]));
  const double vy = __fma_rn(factor, static_cast<double>(f1), ntl(&a.vy[i]));
  const double vz = __fma_rn(factor, static_cast<double>(f2), ntl(&a.vz[i]));
  if (kEpi != kKe) {
    nts(vx, &a.vx[i]);
    nts(vy, &a.vy[i]);
    nts(vz, &a.vz[i]);
  }
  if (emits_ke(kEpi)) return 0.5 * mass * kMvv2e * (vx * vx + vy * vy + vz * vz);
  return 0.0;
}

// ---------------------------------------------------------------------------
// integrate-only kernels
// ---------------------------------------------------------------------------

// Window bootstrap (fused_half_kick_drift) and the NVT first kick: kick by
// dt/2 with the stored forces, drift, write M. NVT scales v by the chain's
// pending factor first.
template <class R, class Mir, bool kNvt, bool kFull, bool kKick2 = false>
__global__ void __launch_bounds__(kBlock) k_kick_drift(StepArgs a) {
  const int i = xcd_block(a.nblocks) * kBlock + threadIdx.x;
  if (i >= a.natoms) return;
  const double s = kNvt ? a.nhc[26] : 1.0;
  kick_drift<R, Mir, kKick2>(a, i, s, kFull ? a.dt : 0.5 * a.dt,
                             static_cast<double>(ntl(&rp<R>(a.fx)[i])),
                             static_cast<double>(ntl(&rp<R>(a.fy)[i])),
                             static_cast<double>(ntl(&rp<R>(a.fz)[i])), a.M);
}

// Mirror refresh (after upload or restore): M[i] = {x, type}.
template <class Mir>
__global__ void __launch_bounds__(kBlock) k_write_mirror(StepArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.natoms) return;
  static_cast<typename Mir::T*>(a.M)[i] =
      encode_type<Mir>(mp_of<Mir>(a, i), a.px[i], a.py[i], a.pz[i], a.types[i] - 1);
}

template <class R>
__global__ void __launch_bounds__(kBlock) k_ke_partials(StepArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double e = 0.0;
  if (i < a.natoms) {
    const double mass = a.masses[a.types[i]];
    e = 0.5 * mass * kMvv2e * (a.vx[i] * a.vx[i] + a.vy[i] * a.vy[i] + a.vz[i] * a.vz[i]);
  }
  emit_partial(e, a.partials, blockIdx.x);
}

__global__ void k_scale_v(StepArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.natoms) return;
  const double s = a.nhc[24];
  a.vx[i] *= s;
  a.vy[i] *= s;
  a.vz[i] *= s;
}

// ---------------------------------------------------------------------------
// NHC chain (one thread; device_nose_hoover.cu:206-289 on device state)
//   nhc[0..7] eta, [8..15] eta_v, [16..23] eta_m, [24] vscale (pending
//   factor not yet applied to v), [25] KE of the stored v, [26] kick scale,
//   [27] reduced sum scratch
// ---------------------------------------------------------------------------

// Before the VV step: half-step on the KE of the actual velocities
// (vscale^2 * stored KE), kick scale = s1 * vscale.
__global__ void k_nhc_pre(double* nhc, NhcParams p) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double ke = nhc[24] * nhc[24] * nhc[25];
  const double s1 = nhc_half_step(p.M, nhc, nhc + 8, nhc + 16, p.t_target, p.t_period, p.dt,
                                  p.n_dof, ke);
  nhc[26] = s1 * nhc[24];
  nhc[24] = 1.0;
}

// After the VV step: KE from the per-block partials of the kick-2 epilogue,
// half-step, pending scale = s2. kPre: the next step's k_nhc_pre follows in
// the same thread (same operations, one launch less per step).
template <bool kPre>
__global__ void k_nhc_post(double* nhc, const double* partials, int n, NhcParams p) {
  __shared__ double s[256];
  double t = 0.0;
  for (int k = threadIdx.x; k < n; k += 256) t += partials[k];
  s[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double ke = s[0];
    const double s2 = nhc_half_step(p.M, nhc, nhc + 8, nhc + 16, p.t_target, p.t_period, p.dt,
                                    p.n_dof, ke);
    nhc[25] = ke;
    nhc[24] = s2;
    if (kPre) {
      const double ke1 = nhc[24] * nhc[24] * nhc[25];
      const double s1 = nhc_half_step(p.M, nhc, nhc + 8, nhc + 16, p.t_target, p.t_period, p.dt,
                                      p.n_dof, ke1);
      nhc[26] = s1 * nhc[24];
      nhc[24] = 1.0;
    }
  }
}

// Stored-KE refresh (after upload, or after the pending scale was applied).
__global__ void k_nhc_set_ke(double* nhc, const double* partials, int n, int commit_scale) {
  __shared__ double s[256];
  double t = 0.0;
  for (int k = threadIdx.x; k < n; k += 256) t += partials[k];
  s[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    nhc[25] = s[0];
    if (commit_scale) nhc[24] = 1.0;
  }
}

// Window end: the pending factor has been applied to v by k_scale_v.
__global__ void k_nhc_commit(double* nhc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  nhc[25] *= nhc[24] * nhc[24];
  nhc[24] = 1.0;
}

// ---------------------------------------------------------------------------
// Neighbour loop shared by the force passes.
//
// Slots are visited in groups of kU. The indices of the next group are
// loaded while the current group's records are gathered and evaluated, so an
// index load (a coalesced 256-byte line per slot and wave, from HBM) overlaps
// a whole group of pair work instead of heading a dependent load chain. Rows
// are padded to a multiple of kRowPad slots with the atom itself, so group
// loads need no predicate; padded slots arrive with in_row = false.
// ---------------------------------------------------------------------------

constexpr int kRowPad = 8;  // rows padded to a multiple of this (cap too)
constexpr int kPairUnroll = 8;

// Slots k..k+kU-1 of atom i (k a multiple of 4): one 16-byte load per quad.
template <int kU>
__device__ __forceinline__ void load_slots(const StepArgs& a, int i, int k, int4 (&q)[kU / 4]) {
  const int4* __restrict__ q4 = reinterpret_cast<const int4*>(a.nbr);
#pragma unroll
  for (int h = 0; h < kU / 4; ++h) q[h] = q4[ell(i, k + 4 * h, a.cap) >> 2];
}

// Record j of a gather array, addressed as the array base plus an unsigned
// 32-bit byte offset (the host keeps natoms x record size below 4 GiB), so
// the index stays one 32-bit register instead of a sign-extended pair.
template <class T>
__device__ __forceinline__ T gather(const T* __restrict__ rec, int j) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(rec) +
                                     static_cast<uint32_t>(j) * static_cast<uint32_t>(sizeof(T)));
}

// The loop carries the next group's slots as the loaded quads and unpacks
// them at the top of the next iteration: registers the loads write are not
// copied before this group's pair work, so that work overlaps their latency
// (and the records' gathers) instead of waiting for them.
template <class T, int kU, class Pair>
__device__ __forceinline__ void for_neighbours(const StepArgs& a, int i, int n,
                                               const T* __restrict__ rec, Pair&& pair) {
  static_assert(kRowPad % kU == 0 && kU % 4 == 0, "groups must tile the padded row in quads");
  int4 q[kU / 4];
  if (n > 0) {
    load_slots<kU>(a, i, 0, q);
  } else {
#pragma unroll
    for (int h = 0; h < kU / 4; ++h) q[h] = make_int4(i, i, i, i);
  }
  for (int k = 0; k < n; k += kU) {
    int j[kU];
#pragma unroll
    for (int h = 0; h < kU / 4; ++h) {
      j[4 * h] = q[h].x;
      j[4 * h + 1] = q[h].y;
      j[4 * h + 2] = q[h].z;
      j[4 * h + 3] = q[h].w;
    }
    T mj[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) mj[u] = gather(rec, j[u]);
    if (k + kU < n) load_slots<kU>(a, i, k + kU, q);
#pragma unroll
    for (int u = 0; u < kU; ++u) pair(j[u], mj[u], k + u < n);
  }
}

// Packed 16-bit slots of one group: eight slots in a uint4, four in a uint2.
template <int kU>
struct Slots16;
template <>
struct Slots16<8> {
  using type = uint4;
};
template <>
struct Slots16<4> {
  using type = uint2;
};

// The packed slots k..k+kU-1 of atom i in 16-bit rows (k a multiple of kU).
template <int kU>
__device__ __forceinline__ typename Slots16<kU>::type load_local(const uint16_t* rows, int cap,
                                                                 int i, int k) {
  static_assert(kU == 4 || kU == 8, "four or eight slots per group");
  using P = typename Slots16<kU>::type;
  return reinterpret_cast<const P*>(rows)[ell16(i, k, cap) / kU];
}

__device__ __forceinline__ void unpack_local(const uint4& v, int (&ln)[8]) {
  const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    ln[2 * q] = static_cast<int>(w[q] & 0xffffu);
    ln[2 * q + 1] = static_cast<int>(w[q] >> 16);
  }
}
__device__ __forceinline__ void unpack_local(const uint2& v, int (&ln)[4]) {
  ln[0] = static_cast<int>(v.x & 0xffffu);
  ln[1] = static_cast<int>(v.x >> 16);
  ln[2] = static_cast<int>(v.y & 0xffffu);
  ln[3] = static_cast<int>(v.y >> 16);
}

// for_neighbours over a block whose halo records sit in LDS (`hrec`): the
// 16-bit rows index the halo, and the records come from LDS. The next
// group's slots stay packed until the top of the next iteration, so their
// load is waited for only after this group's pair work. pair(l, record,
// in_row, slot). kGuard: slots past n hold no valid index (read slot 0's
// record instead).
template <class T, int kU, bool kGuard = false, class Pair>
__device__ __forceinline__ void for_neighbours_staged(const StepArgs& a, const uint16_t* rows,
                                                      int i, int n, const T* hrec, Pair&& pair) {
  typename Slots16<kU>::type raw{};
  if (n > 0) raw = load_local<kU>(rows, a.cap, i, 0);
  for (int k = 0; k < n; k += kU) {
    int ln[kU];
    unpack_local(raw, ln);
    if (kGuard) {
#pragma unroll
      for (int u = 0; u < kU; ++u) ln[u] = k + u < n ? ln[u] : 0;
    }
    T mj[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) mj[u] = hrec[ln[u]];
    if (k + kU < n) raw = load_local<kU>(rows, a.cap, i, k + kU);
#pragma unroll
    for (int u = 0; u < kU; ++u) pair(ln[u], mj[u], k + u < n, k + u);
  }
}

// ---------------------------------------------------------------------------
// EAM passes (global tables; any number of elements)
// ---------------------------------------------------------------------------

__device__ __forceinline__ void load_meta(SplineMeta* sm, const SplineMeta* meta, int ns) {
  for (int k = threadIdx.x; k < ns; k += blockDim.x) sm[k] = meta[k];
  __syncthreads();
}

// A row split over kLanes neighbouring lanes (small systems, where one lane
// per atom leaves most of the chip idle): lane `part` takes quads part,
// part + kLanes, ... of the padded row; the four records of a quad are
// gathered before its pairs are evaluated.
template <class T, int kLanes, class Pair>
__device__ __forceinline__ void for_neighbours_split(const StepArgs& a, int i, int n, int part,
                                                     const T* __restrict__ rec, Pair&& pair) {
  const int4* __restrict__ q4 = reinterpret_cast<const int4*>(a.nbr);
  const int nq = (n + 3) >> 2;  // rows are padded to 8 slots: whole quads
  for (int q = part; q < nq; q += kLanes) {
    const int4 s = q4[ell(i, 4 * q, a.cap) >> 2];
    const T r0 = gather(rec, s.x), r1 = gather(rec, s.y), r2 = gather(rec, s.z),
            r3 = gather(rec, s.w);
    const int k = 4 * q;
    pair(s.x, r0, k < n);
    pair(s.y, r1, k + 1 < n);
    pair(s.z, r2, k + 2 < n);
    pair(s.w, r3, k + 3 < n);
  }
}

// Sum over the kLanes lanes of one atom (butterfly; every lane ends with
// the same bits, so the order is fixed).
template <int kLanes>
__device__ __forceinline__ double lane_sum(double v) {
#pragma unroll
  for (int w = kLanes / 2; w > 0; w >>= 1) v += __shfl_xor(v, w, 64);
  return v;
}

// Global-table EAM passes (several elements, or systems small enough that
// one workgroup per 1024 atoms would leave most CUs idle). kLanes > 1: the
// row of atom i is split over kLanes neighbouring lanes (quads p, p + kLanes,
// ... to lane p) and the lanes' partial sums are combined by a butterfly in
// a fixed order, so the result is run-to-run identical; the grid is kLanes
// times the one-lane grid and emits one energy / KE partial per 256 threads.
template <class R, class Mir, bool kEnergy, int kLanes>
__global__ void __launch_bounds__(kBlock)
k_eam_density(StepArgs a, const SplineMeta* __restrict__ meta,
              const typename Quad<R>::type* __restrict__ coeff, double rc_sq) {
  static_assert(kLanes >= 1 && 64 % kLanes == 0, "an atom's lanes share a wavefront");
  using T = typename Mir::T;
  pass_clock_begin(a.pclock);
  __shared__ SplineMeta sm[kMaxSplines];
  load_meta(sm, meta, 2 * a.ntypes);
  const int blk = xcd_block(gridDim.x);
  const int g = blk * kBlock + threadIdx.x;
  const int i = g / kLanes;
  const int part = g % kLanes;
  const T* __restrict__ M = static_cast<const T*>(a.M);
  double e = 0.0, rho = 0.0;
  T mi{};
  if (i < a.natoms) {
    mi = M[i];
    const MirrorParams& mpi = mp_of<Mir>(a, i);
    const int n = a.cnt[i];
    auto pair = [&](int, const T& mj, bool in_row) {
      double dx, dy, dz;
      const double r2 = Mir::delta(mi, mj, mpi, &dx, &dy, &dz);
      if (in_row && r2 < rc_sq) {
        const SplineMeta& m = sm[a.ntypes + Mir::type_of(mj)];
        if constexpr (std::is_same_v<R, float>) {
          float t;
          const auto c = coeff[m.offset + locate_f32(meta_f32(m), r2, &t)];
          rho += static_cast<double>(c.x + t * (c.y + t * (c.z + t * c.w)));
        } else {
          rho += spline_value<R>(m, coeff, pair_r<R>(r2).r);
        }
      }
    };
    if constexpr (kLanes == 1)
      for_neighbours<T, kPairUnroll>(a, i, n, M, pair);
    else
      for_neighbours_split<T, kLanes>(a, i, n, part, M, pair);
  }
  if constexpr (kLanes > 1) rho = lane_sum<kLanes>(rho);  // every lane takes part
  if (i < a.natoms && part == 0) {
    const int ti = Mir::type_of(mi);
    const SplineVal<R> F = spline_both<R>(sm[ti], coeff, rho);
    // two to kPackedTypes elements: the force pass reads a neighbour's
    // element from F'(rho)'s two lowest mantissa bits
    const bool typed = a.ntypes > 1 && a.ntypes <= kPackedTypes;
    static_cast<T*>(a.M2)[i] = Mir::with_w(mi, typed ? Mir::template pack_typed<R>(F.df, ti)
                                                     : Mir::template pack_real<R>(F.df));
    e = static_cast<double>(F.f);
  }
  if (kEnergy) emit_partial(e, a.partials, blk);
  pass_clock_end(a.pclock);
}

template <class R, class Mir, int kEpi, bool kEnergy, bool kMulti, int kLanes>
__global__ void __launch_bounds__(kBlock)
k_eam_force(StepArgs a, const SplineMeta* __restrict__ meta,
            const typename Quad<R>::type* __restrict__ coeff, double rc_sq) {
  static_assert(kLanes >= 1 && 64 % kLanes == 0, "an atom's lanes share a wavefront");
  using T = typename Mir::T;
  unsigned long long* const pclk = a.pclock ? a.pclock + 4 : nullptr;
  pass_clock_begin(pclk);
  __shared__ SplineMeta sm[kMaxSplines];
  const int nt = a.ntypes;
  load_meta(sm, meta, 2 * nt + nt * (nt + 1) / 2);
  const int blk = xcd_block(gridDim.x);
  const int g = blk * kBlock + threadIdx.x;
  const int i = g / kLanes;
  const int part = g % kLanes;
  const T* __restrict__ M2 = static_cast<const T*>(a.M2);
  using Q = typename Quad<R>::type;
  constexpr bool kF32 = std::is_same_v<R, float>;
  double red = 0.0, vir = 0.0, fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0;
  if (i < a.natoms) {
    const T mi = M2[i];
    const MirrorParams& mpi = mp_of<Mir>(a, i);
    const R fpi = Mir::template w_real<R>(mi);
    const int ti = kMulti ? a.types[i] - 1 : 0;
    const int n = a.cnt[i];
    const ForceLocF floc = force_loc_f32(sm[nt]);  // every r spline's grid
    auto pair = [&](int j, const T& mj, bool in_row) {
      double dx, dy, dz;
      const double r2 = Mir::delta(mi, mj, mpi, &dx, &dy, &dz);
      if (in_row && r2 < rc_sq) {
        const int tj = kMulti ? (a.ntypes <= kPackedTypes ? Mir::packed_type(mj) : a.types[j] - 1)
                              : 0;
        const int pidx = kMulti ? phi_index(ti, tj, nt) : 0;
        R dphi, drho_j, drho_i, rinv;
        if constexpr (kF32) {
          // mixed: pair_t_f32's interval and offset, derivatives from b,
          // 2c, 3d (the pair table holds 2c and 3d)
          const PairT p = pair_t_f32(floc, r2);
          rinv = p.rinv;
          if (kMulti && !kEnergy && a.pairtab) {
            const Q* e = static_cast<const Q*>(a.pairtab) +
                         (static_cast<size_t>(ti * nt + tj) * a.pair_nint + p.idx) * 3;
            const Q q0 = e[0], q1 = e[1], q2 = e[2];
            dphi = cubic_deriv_s(q0.x, q0.y, q0.z, p.t);
            drho_j = cubic_deriv_s(q0.w, q1.x, q1.y, p.t);
            drho_i = cubic_deriv_s(q1.z, q1.w, q2.x, p.t);
          } else {
            const Q cphi = coeff[sm[2 * nt + pidx].offset + p.idx];
            if (kEnergy)
              pe += 0.5 * static_cast<double>(
                              cphi.x + p.t * (cphi.y + p.t * (cphi.z + p.t * cphi.w)));
            dphi = cubic_deriv_f(cphi, p.t);
            drho_j = cubic_deriv_f(coeff[sm[nt + tj].offset + p.idx], p.t);
            drho_i = kMulti ? cubic_deriv_f(coeff[sm[nt + ti].offset + p.idx], p.t) : drho_j;
          }
        } else {
          const PairR<R> pr = pair_r<R>(r2);
          rinv = pr.rinv;
          if (kMulti && !kEnergy && a.pairtab) {
            // one interval (the three splines share the r grid), three
            // adjacent quads; the same expressions as spline_deriv
            R dx;
            const int idx = spline_locate<R>(sm[2 * nt + pidx], pr.r, &dx);
            const Q* e = static_cast<const Q*>(a.pairtab) +
                         (static_cast<size_t>(ti * nt + tj) * a.pair_nint + idx) * 3;
            const Q q0 = e[0], q1 = e[1], q2 = e[2];
            dphi = q0.x + dx * (static_cast<R>(2) * q0.y + dx * static_cast<R>(3) * q0.z);
            drho_j = q0.w + dx * (static_cast<R>(2) * q1.x + dx * static_cast<R>(3) * q1.y);
            drho_i = q1.z + dx * (static_cast<R>(2) * q1.w + dx * static_cast<R>(3) * q2.x);
          } else {
            if (kEnergy) {
              const SplineVal<R> ph = spline_both<R>(sm[2 * nt + pidx], coeff, pr.r);
              pe += 0.5 * static_cast<double>(ph.f);
              dphi = ph.df;
            } else {
              dphi = spline_deriv<R>(sm[2 * nt + pidx], coeff, pr.r);
            }
            drho_j = spline_deriv<R>(sm[nt + tj], coeff, pr.r);
            drho_i = kMulti ? spline_deriv<R>(sm[nt + ti], coeff, pr.r) : drho_j;
          }
        }
        const R fpj = Mir::template w_real<R>(mj);
        const double fpair = static_cast<double>(-(dphi + fpi * drho_j + fpj * drho_i) * rinv);
        if (kEnergy) vir += 0.5 * fpair * r2;
        fx = fma(fpair, dx, fx);
        fy = fma(fpair, dy, fy);
        fz = fma(fpair, dz, fz);
      }
    };
    if constexpr (kLanes == 1)
      for_neighbours<T, kPairUnroll>(a, i, n, M2, pair);
    else
      for_neighbours_split<T, kLanes>(a, i, n, part, M2, pair);
  }
  if constexpr (kLanes > 1) {  // every lane takes part in the shuffles
    fx = lane_sum<kLanes>(fx);
    fy = lane_sum<kLanes>(fy);
    fz = lane_sum<kLanes>(fz);
    if (kEnergy) {
      pe = lane_sum<kLanes>(pe);
      vir = lane_sum<kLanes>(vir);
    }
  }
  if (i < a.natoms && part == 0) {
    const double ke = epilogue<R, Mir, kEpi>(a, i, fx, fy, fz);
    red = kEnergy ? pe : ke;
  } else {
    vir = 0.0;
  }
  if (kEnergy || emits_ke(kEpi)) emit_partial(red, a.partials, blk);
  if (kEnergy) emit_partial(vir, a.partials_w, blk);
  pass_clock_end(pclk);
}

// ---------------------------------------------------------------------------
// Morse pass (morse.hpp:37-55; pair_math.hpp precision contract)
// ---------------------------------------------------------------------------

template <class R, class Mir, int kEpi, bool kEnergy, int kLanes>
__global__ void __launch_bounds__(kBlock) k_morse_force(StepArgs a, MorseCoef p) {
  // The pass gathers from M; with the fused kick-drift the next step's
  // records go to M2 and the host swaps the two buffers after the launch.
  // kLanes > 1: kLanes lanes per atom, the grid kLanes times larger.
  static_assert(kLanes >= 1 && 64 % kLanes == 0, "an atom's lanes share a wavefront");
  using T = typename Mir::T;
  unsigned long long* const pclk = a.pclock ? a.pclock + 4 : nullptr;
  pass_clock_begin(pclk);
  const int blk = xcd_block(gridDim.x);
  const int g = blk * kBlock + threadIdx.x;
  const int i = g / kLanes;
  const int part = g % kLanes;
  const T* __restrict__ M = static_cast<const T*>(a.M);
  double red = 0.0, fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0, vir = 0.0;
  if (i < a.natoms) {
    const T mi = M[i];
    const MirrorParams& mpi = mp_of<Mir>(a, i);
    const int n = a.cnt[i];
    auto pair = [&](int, const T& mj, bool in_row) {
      double dx, dy, dz;
      const double r2 = Mir::delta(mi, mj, mpi, &dx, &dy, &dz);
      if (in_row && r2 < p.rc_sq) {
        const PairR<R> pr = pair_r<R>(r2);
        R e;
        const double fpair = static_cast<double>(morse_fpair<R>(p, pr, &e));
        if (kEnergy) {
          pe += 0.5 * static_cast<double>(e);
          vir += 0.5 * fpair * r2;
        }
        fx = fma(fpair, dx, fx);
        fy = fma(fpair, dy, fy);
        fz = fma(fpair, dz, fz);
      }
    };
    if constexpr (kLanes == 1)
      for_neighbours<T, kPairUnroll>(a, i, n, M, pair);
    else
      for_neighbours_split<T, kLanes>(a, i, n, part, M, pair);
  }
  if constexpr (kLanes > 1) {  // every lane takes part in the shuffles
    fx = lane_sum<kLanes>(fx);
    fy = lane_sum<kLanes>(fy);
    fz = lane_sum<kLanes>(fz);
    if (kEnergy) {
      pe = lane_sum<kLanes>(pe);
      vir = lane_sum<kLanes>(vir);
    }
  }
  if (i < a.natoms && part == 0) {
    const double ke = epilogue<R, Mir, kEpi>(a, i, fx, fy, fz, a.M2);
    red = kEnergy ? pe : ke;
  } else {
    vir = 0.0;
  }
  if (kEnergy || emits_ke(kEpi)) emit_partial(red, a.partials, blk);
  if (kEnergy) emit_partial(vir, a.partials_w, blk);
  pass_clock_end(pclk);
}

// ---------------------------------------------------------------------------
// Rebuild: spatial sort and the ELL neighbour list. Every kernel first reads
// the rebuild flag and returns when it is clear, so the whole sequence can
// sit in a captured graph and only does work on steps that need it.
// ---------------------------------------------------------------------------

struct GridArgs {
  int n[3];
  double inv_cs[3];
  int ncells;
  const int32_t* rank_of;  // linear cell id -> Morton rank
  const int32_t* around;   // [ncells][27] by rank: ranks of the 27 cells
                           // around the cell (periodic wrap), increasing
  const uint8_t* around_img;  // the same entries' periodic images: 2 bits per
                              // axis, 1 = +box length, 2 = -box length
  int images;  // 1: every periodic axis has >= 3 cells no smaller than r_list,
               // so a stencil entry's image is the minimum image of any pair
               // within r_list that it holds (batches: per box, BoxGeo)
  int nl_split;  // 1: the fused rebuild searches each atom's cells with 32 lanes
  int32_t* ccount;         // [ncells] by rank
  int32_t* cstart;         // [ncells] by rank
  int32_t* cursor;         // [ncells] by rank
  int32_t* key;            // [N] rank of each atom's cell (old order)
  int32_t* sidx;           // [N] new slot -> old index
  int32_t* orig;           // [N] upload index
  int32_t* done;           // completion counter
  int32_t* wsum;           // [kMaxScanGroups] per-workgroup cell sums (fused rebuild)
};

constexpr int kMaxScanGroups = 4096;  // workgroups of the fused rebuild, at most
// Up to this many cells the fused rebuild scans them with one workgroup
// (~1.6 ns a cell) instead of the whole grid around one more barrier.
constexpr int kScanOneGroup = 4096;


struct PermArgs {
  // scratch arrays the permutation gathers into
  double *px, *py, *pz, *vx, *vy, *vz;
  void *fx, *fy, *fz;
  int32_t *types, *orig;
};

__global__ void k_set_flag(int32_t* flags) {
  if (threadIdx.x == 0 && blockIdx.x == 0) flags[0] = 1;
}

__global__ void __launch_bounds__(kBlock) k_zero_cells(const int32_t* flags, GridArgs g) {
  if (!flags[0]) return;
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= g.ncells) return;
  g.ccount[r] = 0;
  g.cursor[r] = 0;
}

// Binning grid of atom i's box; global cell id = off + (c2 n1 + c1) n0 + c0.
struct CellGrid {
  int n[3];
  double inv_cs[3];
  double lo[3];
  int off;
};

template <class Mir>
__device__ __forceinline__ CellGrid grid_of(const StepArgs& a, const GridArgs& g, int i) {
  CellGrid cg;
  if constexpr (Mir::kMB) {
    const BoxGeo& b = a.boxes[a.rep[i]];
    for (int d = 0; d < 3; ++d) {
      cg.n[d] = b.n[d];
      cg.inv_cs[d] = b.inv_cs[d];
      cg.lo[d] = b.mp.lo[d];
    }
    cg.off = b.cell_off;
  } else {
    for (int d = 0; d < 3; ++d) {
      cg.n[d] = g.n[d];
      cg.inv_cs[d] = g.inv_cs[d];
      cg.lo[d] = a.mp.lo[d];
    }
    cg.off = 0;
  }
  return cg;
}

__device__ __forceinline__ void cell_coords(const CellGrid& cg, double x, double y, double z,
                                            int c[3]) {
  const double p[3] = {x, y, z};
  for (int d = 0; d < 3; ++d) {
    int v = static_cast<int>((p[d] - cg.lo[d]) * cg.inv_cs[d]);
    v = v >= cg.n[d] ? cg.n[d] - 1 : v;
    c[d] = v < 0 ? 0 : v;
  }
}

// ---- per-item bodies, shared by the separate kernels (first build) and
// the fused rebuild kernel ----

// Adds one to ctr[key] for each active lane of the wavefront with one atomic
// per distinct key (atoms are in cell order from the previous sort, so a
// wave's lanes share a few cells and per-lane atomics on one address
// serialise); returns the counter's value before this lane's increment, in
// lane order within each key. Every active lane of the wave must call it.
__device__ __forceinline__ int wave_add_one(int32_t* ctr, int key) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  unsigned long long todo = __ballot(1);
  int old = 0;
  while (todo) {
    const int leader = __ffsll(todo) - 1;
    const int k = __shfl(key, leader, 64);
    const unsigned long long same = __ballot(key == k) & todo;
    int base = 0;
    if (lane == leader) base = atomicAdd(&ctr[k], __popcll(same));
    base = __shfl(base, leader, 64);
    if ((same >> lane) & 1ull) old = base + __popcll(same & ((1ull << lane) - 1ull));
    todo &= ~same;
  }
  return old;
}

template <class Mir>
__device__ __forceinline__ void bin_atom(const StepArgs& a, const GridArgs& g, int i) {
  int c[3];
  const CellGrid cg = grid_of<Mir>(a, g, i);
  cell_coords(cg, a.px[i], a.py[i], a.pz[i], c);
  const int rank = g.rank_of[cg.off + (c[2] * cg.n[1] + c[1]) * cg.n[0] + c[0]];
  g.key[i] = rank;
  (void)wave_add_one(g.ccount, rank);
}

// Exclusive scan of the cell counts by one workgroup of BLOCK threads
// (atomics-free): thread t owns a contiguous chunk of cells, sums it, the
// workgroup scans the chunk sums, and each thread writes its chunk's starts.
template <int BLOCK>
__device__ __forceinline__ void scan_cells(const GridArgs& g) {
  __shared__ int wsum[BLOCK / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int chunk = (g.ncells + BLOCK - 1) / BLOCK;
  const int c0 = min(static_cast<int>(threadIdx.x) * chunk, g.ncells);
  const int c1 = min(c0 + chunk, g.ncells);
  int v = 0;
  for (int c = c0; c < c1; ++c) v += g.ccount[c];
  int incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int o = __shfl_up(incl, off, 64);
    if (lane >= off) incl += o;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int base = incl - v;
  for (int k = 0; k < w; ++k) base += wsum[k];
  for (int c = c0; c < c1; ++c) {
    g.cstart[c] = base;
    base += g.ccount[c];
  }
}

// (The order of a cell's atoms after the scatter is the atomics' order;
// the ordering pass then sorts each cell by upload index.)
__device__ __forceinline__ void scatter_atom(const GridArgs& g, int i) {
  const int r = g.key[i];
  g.sidx[g.cstart[r] + wave_add_one(g.cursor, r)] = i;
}

// Inclusive scan of v over the workgroup (BLOCK threads); *total gets the
// workgroup's sum. Uses `ws` (BLOCK / 64 ints of LDS); ends synchronised.
template <int BLOCK>
__device__ __forceinline__ int block_scan_incl(int v, int* total, int* ws) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int o = __shfl_up(incl, off, 64);
    if (lane >= off) incl += o;
  }
  if (lane == 63) ws[w] = incl;
  __syncthreads();
  int before = 0, sum = 0;
  for (int k = 0; k < BLOCK / 64; ++k) {
    const int x = ws[k];
    before += k < w ? x : 0;
    sum += x;
  }
  __syncthreads();  // ws is reused by the next call
  *total = sum;
  return before + incl;
}

// The exclusive cell scan by the whole grid of the fused rebuild, in two
// phases with a grid barrier between them: workgroup w owns a contiguous
// chunk of cells; phase 1 sums its chunk into wsum[w]; phase 2 adds the sums
// of the chunks before it and scans its chunk. The same starts as
// scan_cells, which one workgroup computes alone.
template <int BLOCK>
__device__ __forceinline__ void scan_cells_grid(const GridArgs& g, int phase) {
  __shared__ int ws[BLOCK / 64];
  const int W = gridDim.x;
  const int chunk = (g.ncells + W - 1) / W;
  const int c0 = min(static_cast<int>(blockIdx.x) * chunk, g.ncells);
  const int c1 = min(c0 + chunk, g.ncells);
  int total = 0;
  if (phase == 1) {
    int v = 0;
    for (int c = c0 + threadIdx.x; c < c1; c += BLOCK) v += g.ccount[c];
    (void)block_scan_incl<BLOCK>(v, &total, ws);
    if (threadIdx.x == 0) g.wsum[blockIdx.x] = total;
    return;
  }
  int v = 0;
  for (int k = threadIdx.x; k < static_cast<int>(blockIdx.x); k += BLOCK) v += g.wsum[k];
  (void)block_scan_incl<BLOCK>(v, &total, ws);
  int run = total;  // cells before this chunk
  for (int c = c0; c < c1; c += BLOCK) {
    const int cc = c + threadIdx.x;
    const int n = cc < c1 ? g.ccount[cc] : 0;
    const int incl = block_scan_incl<BLOCK>(n, &total, ws);
    if (cc < c1) g.cstart[cc] = run + incl - n;
    run += total;
  }
}

// Deterministic order inside a cell: by upload index.
__device__ __forceinline__ void order_cell(const GridArgs& g, int r) {
  int32_t* seg = g.sidx + g.cstart[r];
  const int n = g.ccount[r];
  for (int x = 1; x < n; ++x) {
    const int32_t v = seg[x];
    const int32_t kv = g.orig[v];
    int y = x - 1;
    while (y >= 0 && g.orig[seg[y]] > kv) {
      seg[y + 1] = seg[y];
      --y;
    }
    seg[y + 1] = v;
  }
}

// The same order by one wavefront: lane l takes the cell's l-th atom and
// counts the atoms with a smaller upload index (indices are unique, so the
// counts are the sorted positions); all lanes read before any writes.
// Cells of more than 64 atoms fall back to one lane's insertion sort.
__device__ __forceinline__ void order_cell_wave(const GridArgs& g, int r) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  int32_t* seg = g.sidx + g.cstart[r];
  const int n = g.ccount[r];
  if (n > 64) {
    if (lane == 0) order_cell(g, r);
    return;
  }
  const int32_t v = lane < n ? seg[lane] : 0;
  const int32_t kv = lane < n ? g.orig[v] : 0x7fffffff;
  int pos = 0;
  for (int m = 0; m < n; ++m) pos += __shfl(kv, m, 64) < kv ? 1 : 0;
  if (lane < n) seg[pos] = v;
}

template <class R>
__device__ __forceinline__ void permute_atom(const StepArgs& a, const GridArgs& g,
                                             const PermArgs& t, int s) {
  const int o = g.sidx[s];
  t.px[s] = a.px[o];
  t.py[s] = a.py[o];
  t.pz[s] = a.pz[o];
  t.vx[s] = a.vx[o];
  t.vy[s] = a.vy[o];
  t.vz[s] = a.vz[o];
  rp<R>(t.fx)[s] = rp<R>(a.fx)[o];
  rp<R>(t.fy)[s] = rp<R>(a.fy)[o];
  rp<R>(t.fz)[s] = rp<R>(a.fz)[o];
  t.types[s] = a.types[o];
  t.orig[s] = g.orig[o];
}

template <class R, class Mir>
__device__ __forceinline__ void copyback_atom(const StepArgs& a, const GridArgs& g,
                                              const PermArgs& t, int s) {
  const double x = t.px[s], y = t.py[s], z = t.pz[s];
  a.px[s] = x;
  a.py[s] = y;
  a.pz[s] = z;
  if constexpr (Mir::kFixed) {
    const MirrorParams& mp = mp_of<Mir>(a, s);
    reinterpret_cast<int32_t*>(a.bx)[s] = Mir::enc1(x, mp.lo[0], mp.q[0]);
    reinterpret_cast<int32_t*>(a.by)[s] = Mir::enc1(y, mp.lo[1], mp.q[1]);
    reinterpret_cast<int32_t*>(a.bz)[s] = Mir::enc1(z, mp.lo[2], mp.q[2]);
  } else {
    a.bx[s] = x;
    a.by[s] = y;
    a.bz[s] = z;
  }
  a.vx[s] = t.vx[s];
  a.vy[s] = t.vy[s];
  a.vz[s] = t.vz[s];
  rp<R>(a.fx)[s] = rp<R>(t.fx)[s];
  rp<R>(a.fy)[s] = rp<R>(t.fy)[s];
  rp<R>(a.fz)[s] = rp<R>(t.fz)[s];
  const int ty = t.types[s];
  const_cast<int32_t*>(a.types)[s] = ty;
  g.orig[s] = t.orig[s];
  static_cast<typename Mir::T*>(a.M)[s] = Mir::encode(mp_of<Mir>(a, s), x, y, z, ty - 1);
}

// Candidate prefilter of the list search with the 32-bit fixed-point
// records (MirrorFixed): r^2 in FP32 from the records' wrapping differences.
// The records round each coordinate to one unit (at most L / 2^32, or
// W / 2^30 on an open axis; cells are at most 1024 per axis, so a unit is
// below 1e-6 of r_list) and FP32 adds ~1e-6 relative, so a candidate whose
// FP32 r^2 lies outside [1 - kPreMargin, 1 + kPreMargin] times a bound is on
// the same side of it as its exact FP64 r^2. Inside that band (a few
// candidates per million atoms) the exact test decides.
constexpr double kPreMargin = 1e-5;

// When the three axes share one unit u (cubic boxes), r^2 is formed in
// units^2 and the bounds are divided by u^2 instead: three products fewer
// per candidate, the same relative error.
struct PreBounds {
  float lo_list, hi_list, lo_near, hi_near;
  float ux, uy, uz;
  bool iso;
};

__device__ __forceinline__ PreBounds pre_bounds(const MirrorParams& mp, double r_list_sq,
                                                double near_sq) {
  PreBounds b;
  b.iso = mp.unit[0] == mp.unit[1] && mp.unit[1] == mp.unit[2];
  const double s = b.iso ? 1.0 / (mp.unit[0] * mp.unit[0]) : 1.0;
  b.lo_list = static_cast<float>(r_list_sq * s * (1.0 - kPreMargin));
  b.hi_list = static_cast<float>(r_list_sq * s * (1.0 + kPreMargin));
  b.lo_near = static_cast<float>(near_sq * s * (1.0 - kPreMargin));
  b.hi_near = static_cast<float>(near_sq * s * (1.0 + kPreMargin));
  b.ux = static_cast<float>(mp.unit[0]);
  b.uy = static_cast<float>(mp.unit[1]);
  b.uz = static_cast<float>(mp.unit[2]);
  return b;
}

template <bool kIsoU>
__device__ __forceinline__ float pre_r2(const int4& mj, const int4& mi, const PreBounds& b) {
  float dx = static_cast<float>(MirrorFixed::wdiff(mj.x, mi.x));
  float dy = static_cast<float>(MirrorFixed::wdiff(mj.y, mi.y));
  float dz = static_cast<float>(MirrorFixed::wdiff(mj.z, mi.z));
  if constexpr (!kIsoU) {
    dx *= b.ux;
    dy *= b.uy;
    dz *= b.uz;
  }
  return dx * dx + dy * dy + dz * dz;
}

// ELL row of atom i over the sorted atoms; returns the true count. Inclusion
// exactly as the CPU list (neighbor_list.cpp:58-62): d = x_j - x_i in FP64,
// single minimum image, unfused r2 < r_list^2. kCount: only count.
//
// Each row is partitioned: neighbours inside the interaction cutoff at build
// time first, the skin shell (rc <= r < r_list) after them. Atoms move less
// than half the skin between builds, so in the force passes the in-cutoff
// test is true for the leading slots and false for the trailing ones in
// nearly every lane of a wave, and whole waves skip the pair work of the
// shell. Near entries are written from the front and shell entries from the
// back of the row; the shell block is then moved down behind the near block.
template <bool kCount, class Mir>
__device__ __forceinline__ int nlist_atom(const StepArgs& a, const GridArgs& g, int i,
                                          double r_list_sq, double near_sq, int32_t* nbr,
                                          int32_t* cnt) {
  int count = 0, nn = 0, nf = 0;
  const int cap = a.cap;
  const double xi = a.px[i], yi = a.py[i], zi = a.pz[i];
  const MirrorParams& mp = mp_of<Mir>(a, i);
  const CellGrid cg = grid_of<Mir>(a, g, i);
  int c[3];
  cell_coords(cg, xi, yi, zi, c);
  // The 27 cells around c, visited in increasing rank. Atoms are sorted by
  // cell rank, so each row comes out in increasing atom index, and the
  // lanes of a wave (neighbouring atoms with overlapping neighbourhoods)
  // reach the same stretch of the record array at the same slot.
  const int32_t* __restrict__ around =
      g.around + 27 * static_cast<size_t>(
                          g.rank_of[cg.off + (c[2] * cg.n[1] + c[1]) * cg.n[0] + c[0]]);
  const size_t row = static_cast<size_t>(around - g.around);
  int4 mi_rec = make_int4(0, 0, 0, 0);
  PreBounds pb{};
  if constexpr (Mir::kFixed) {
    mi_rec = static_cast<const int4*>(a.M)[i];
    pb = pre_bounds(mp, r_list_sq, near_sq);
  }
  // One copy of the search per image mode, so the loop body has no branch
  // between a group's coordinate loads and they issue together.
  // Row slots are addressed from the row's first slot: slot k is
  // (k / 4) * 256 + k % 4 elements on (the ELL layout of ell()).
  int32_t* const rowp = kCount ? nullptr : nbr + ell(i, 0, cap);
  auto slot = [&](int k) -> int32_t& { return rowp[(k >> 2) * 256 + (k & 3)]; };
  auto search = [&](auto images, auto iso_units) {
    constexpr bool kImg = decltype(images)::value;
    constexpr bool kIsoU = decltype(iso_units)::value;
    for (int t = 0; t < 27; ++t) {
      const int r = around[t];
      const int start = g.cstart[r], end = start + g.ccount[r];
      // kImg: the cell's image offset replaces the per-candidate minimum
      // image test: x_j - x_i then +-L, the same two roundings, so r2 is
      // bit for bit that of the test for every pair within r_list
      double sx = 0.0, sy = 0.0, sz = 0.0;
      bool ix = false, iy = false, iz = false;
      if constexpr (kImg) {
        const int img = g.around_img[row + t];
        ix = (img & 3) != 0;
        iy = ((img >> 2) & 3) != 0;
        iz = ((img >> 4) & 3) != 0;
        sx = (img & 3) == 1 ? mp.L[0] : -mp.L[0];
        sy = ((img >> 2) & 3) == 1 ? mp.L[1] : -mp.L[1];
        sz = ((img >> 4) & 3) == 1 ? mp.L[2] : -mp.L[2];
      }
      auto exact_r2 = [&](int j) {
        double ddx = a.px[j] - xi, ddy = a.py[j] - yi, ddz = a.pz[j] - zi;
        if constexpr (kImg) {
          ddx = ix ? ddx + sx : ddx;
          ddy = iy ? ddy + sy : ddy;
          ddz = iz ? ddz + sz : ddz;
        } else {
          ddx = min_image(ddx, mp.L[0], mp.hL[0], mp.periodic[0]);
          ddy = min_image(ddy, mp.L[1], mp.hL[1], mp.periodic[1]);
          ddz = min_image(ddz, mp.L[2], mp.hL[2], mp.periodic[2]);
        }
        return r2_unfused(ddx, ddy, ddz);
      };
      if constexpr (Mir::kFixed) {
        // one 16-byte record per candidate, prefiltered in FP32; the exact
        // test where the prefilter is too close to call
        const char* __restrict__ Mb = static_cast<const char*>(a.M);
        for (int j0 = start; j0 < end; j0 += 8) {
          // A group of eight candidates: their prefilter decisions in
          // straight-line code, so the eight record loads are issued together
          // and consumed there (with the distance computed under each
          // candidate's branch the compiler re-issued the first load inside
          // its branch and waited on it); the exact test of the rare in-band
          // candidates and the appends follow. (Groups of four: 1M-atom
          // rebuild 2.09 ms; eight 2.00; sixteen 2.06.) The eight records
          // are read from the group's first at fixed offsets, also past the
          // cell's end (the buffer has kRecPad spare records at its end);
          // those slots are masked.
          const int4* __restrict__ g8 =
              reinterpret_cast<const int4*>(Mb + (static_cast<uint32_t>(j0) << 4));
          bool in[8], near[8], band[8];
          bool any_band = false;
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int j = j0 + q;
            const float r2f = pre_r2<kIsoU>(g8[q], mi_rec, pb);
            // bitwise, not short-circuit, so no branch splits the group;
            // a band is [lo, hi] around a bound: below hi and not below lo
            const bool valid = (j < end) & (j != i);
            const bool list_lo = r2f < pb.lo_list, list_hi = r2f <= pb.hi_list;
            const bool near_lo = r2f < pb.lo_near, near_hi = r2f <= pb.hi_near;
            band[q] = valid & ((list_hi & !list_lo) | (near_hi & !near_lo));
            in[q] = valid & list_lo;
            near[q] = valid & near_lo;
            any_band |= band[q];
          }
          if (any_band) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
              if (band[q]) {
                const double r2 = exact_r2(j0 + q);
                in[q] = r2 < r_list_sq;
                near[q] = r2 < near_sq;
              }
            }
          }
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            if (in[q]) {
              if (!kCount && nn + nf < cap) {
                if (near[q])
                  slot(nn++) = j0 + q;
                else
                  slot(cap - 1 - nf++) = j0 + q;
              }
              ++count;
            }
          }
        }
        continue;
      }
      // candidates in groups of four: the group's coordinates are loaded
      // together, then tested and appended in index order
      for (int j0 = start; j0 < end; j0 += 4) {
        double px[4], py[4], pz[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int j = min(j0 + q, end - 1);
          px[q] = a.px[j];
          py[q] = a.py[j];
          pz[q] = a.pz[j];
        }
        double r2q[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double ddx = px[q] - xi, ddy = py[q] - yi, ddz = pz[q] - zi;
          if constexpr (kImg) {
            ddx = ix ? ddx + sx : ddx;
            ddy = iy ? ddy + sy : ddy;
            ddz = iz ? ddz + sz : ddz;
          } else {
            ddx = min_image(ddx, mp.L[0], mp.hL[0], mp.periodic[0]);
            ddy = min_image(ddy, mp.L[1], mp.hL[1], mp.periodic[1]);
            ddz = min_image(ddz, mp.L[2], mp.hL[2], mp.periodic[2]);
          }
          r2q[q] = r2_unfused(ddx, ddy, ddz);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int j = j0 + q;
          if (j < end && j != i && r2q[q] < r_list_sq) {
            if (!kCount && nn + nf < cap) {
              if (r2q[q] < near_sq)
                slot(nn++) = j;
              else
                slot(cap - 1 - nf++) = j;
            }
            ++count;
          }
        }
      }
    }
  };
  bool images;
  if constexpr (Mir::kMB)
    images = a.boxes[a.rep[i]].images != 0;
  else
    images = g.images != 0;
  // (pb.iso is false unless the records are fixed-point)
  if (images) {
    if (pb.iso)
      search(std::true_type{}, std::true_type{});
    else
      search(std::true_type{}, std::false_type{});
  } else {
    if (pb.iso)
      search(std::false_type{}, std::true_type{});
    else
      search(std::false_type{}, std::false_type{});
  }
  if (!kCount) {
    // Shell entries sit at [cap - nf, cap); the gap [nn, cap - nf) has
    // room for the m that must move (the rest already lie in [nn, nn+nf)).
    const int gap = cap - nn - nf;
    const int m = nf < gap ? nf : gap;
    for (int q = 0; q < m; ++q) slot(nn + q) = slot(cap - 1 - q);
    const int n = nn + nf;  // == min(count, cap)
    cnt[i] = n;
    a.cntn[i] = nn;
    // Pad the row to a multiple of 8 with the atom itself, so unrolled
    // readers can load whole groups of slots (cap is a multiple of 8).
    for (int k = n; k < ((n + 7) & ~7); ++k) slot(k) = i;
  }
  return count;
}

// nlist_atom with 32 lanes per atom (two atoms per wavefront), for systems
// too small to fill the chip one lane per atom. Lane t < 27 of a group loads
// stencil entry t's cell range and image, shared by shuffles; each cell's
// candidates are then tested 32 at a time, and a hit's place in the row is
// its rank among the hits in candidate order (ballot, then a prefix count):
// hit k is kept when k < cap, near entries go to the front and shell entry
// q to cap - 1 - q, as in nlist_atom. After a workgroup-scope fence (this
// wave's row stores complete before it reads them back) the shell block is
// moved behind the near block and the row padded, as nlist_atom does.
// Distance arithmetic is nlist_atom's, so the row is bit for bit the same.
// Every lane of the wave runs the loop (the ballots and shuffles need
// them); `act` is false for a half with no atom, whose i is any valid atom.
template <class Mir>
__device__ __forceinline__ int nlist_atom_group(const StepArgs& a, const GridArgs& g, int i,
                                                bool act, double r_list_sq, double near_sq,
                                                int32_t* nbr, int32_t* cnt) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int l = lane & 31, sh = lane & 32;
  const uint32_t lt = (1u << l) - 1u;
  const int cap = a.cap;
  const double xi = a.px[i], yi = a.py[i], zi = a.pz[i];
  const MirrorParams& mp = mp_of<Mir>(a, i);
  const CellGrid cg = grid_of<Mir>(a, g, i);
  int c[3];
  cell_coords(cg, xi, yi, zi, c);
  const size_t row =
      27 * static_cast<size_t>(g.rank_of[cg.off + (c[2] * cg.n[1] + c[1]) * cg.n[0] + c[0]]);
  bool images;
  if constexpr (Mir::kMB)
    images = a.boxes[a.rep[i]].images != 0;
  else
    images = g.images != 0;
  int my_start = 0, my_end = 0, my_img = 0;
  if (act && l < 27) {
    const int r = g.around[row + l];
    my_start = g.cstart[r];
    my_end = my_start + g.ccount[r];
    my_img = g.around_img[row + l];
  }
  int4 mi_rec = make_int4(0, 0, 0, 0);
  PreBounds pb{};
  if constexpr (Mir::kFixed) {
    mi_rec = static_cast<const int4*>(a.M)[i];
    pb = pre_bounds(mp, r_list_sq, near_sq);
  }
  int count = 0, nn = 0, nf = 0;
  for (int t = 0; t < 27; ++t) {
    const int start = __shfl(my_start, sh + t);
    const int end = __shfl(my_end, sh + t);
    const int img = __shfl(my_img, sh + t);
    const bool ix = images && (img & 3) != 0;
    const bool iy = images && ((img >> 2) & 3) != 0;
    const bool iz = images && ((img >> 4) & 3) != 0;
    const double sx = (img & 3) == 1 ? mp.L[0] : -mp.L[0];
    const double sy = ((img >> 2) & 3) == 1 ? mp.L[1] : -mp.L[1];
    const double sz = ((img >> 4) & 3) == 1 ? mp.L[2] : -mp.L[2];
    for (int j0 = start; __any(j0 < end); j0 += 32) {
      const int j = j0 + l;
      const bool valid = j < end;
      const int jj = valid ? j : i;
      auto exact_r2 = [&]() {
        double ddx = a.px[jj] - xi, ddy = a.py[jj] - yi, ddz = a.pz[jj] - zi;
        if (images) {
          ddx = ix ? ddx + sx : ddx;
          ddy = iy ? ddy + sy : ddy;
          ddz = iz ? ddz + sz : ddz;
        } else {
          ddx = min_image(ddx, mp.L[0], mp.hL[0], mp.periodic[0]);
          ddy = min_image(ddy, mp.L[1], mp.hL[1], mp.periodic[1]);
          ddz = min_image(ddz, mp.L[2], mp.hL[2], mp.periodic[2]);
        }
        return r2_unfused(ddx, ddy, ddz);
      };
      bool in_list, in_near;
      if constexpr (Mir::kFixed) {
        // the record prefilter of nlist_atom, the exact test in its band
        const float r2f = pb.iso ? pre_r2<true>(static_cast<const int4*>(a.M)[jj], mi_rec, pb)
                                 : pre_r2<false>(static_cast<const int4*>(a.M)[jj], mi_rec, pb);
        if ((r2f >= pb.lo_list && r2f <= pb.hi_list) ||
            (r2f >= pb.lo_near && r2f <= pb.hi_near)) {
          const double r2 = exact_r2();
          in_list = r2 < r_list_sq;
          in_near = r2 < near_sq;
        } else {
          in_list = r2f < pb.lo_list;
          in_near = r2f < pb.lo_near;
        }
      } else {
        const double r2 = exact_r2();
        in_list = r2 < r_list_sq;
        in_near = r2 < near_sq;
      }
      const bool hit = valid && j != i && in_list;
      const uint32_t hg = static_cast<uint32_t>(__ballot(hit) >> sh);
      const bool kept = hit && count + __popc(hg & lt) < cap;
      const bool near = kept && in_near;
      const uint32_t kg = static_cast<uint32_t>(__ballot(kept) >> sh);
      const uint32_t ng = static_cast<uint32_t>(__ballot(near) >> sh);
      if (kept) {
        if (near)
          nbr[ell(i, nn + __popc(ng & lt), cap)] = j;
        else
          nbr[ell(i, cap - 1 - (nf + __popc(kg & ~ng & lt)), cap)] = j;
      }
      count += __popc(hg);
      nn += __popc(ng);
      nf += __popc(kg & ~ng);
    }
  }
  // the row stores above are complete before the shell entries are read back
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  if (act) {
    // Shell entries sit at [cap - nf, cap); the gap [nn, cap - nf) has room
    // for the m that move (source and destination ranges are disjoint).
    const int gap = cap - nn - nf;
    const int m = nf < gap ? nf : gap;
    for (int q = l; q < m; q += 32) nbr[ell(i, nn + q, cap)] = nbr[ell(i, cap - 1 - q, cap)];
    const int n = nn + nf;  // == min(count, cap)
    if (l == 0) {
      cnt[i] = n;
      a.cntn[i] = nn;
    }
    const int k = n + l;
    if (k < ((n + 7) & ~7)) nbr[ell(i, k, cap)] = i;
  }
  return act ? count : 0;
}

// Largest row count of the wave: recorded for sizing (kCount) or flagged as
// an overflow of the capacity.
template <bool kCount>
__device__ __forceinline__ void report_count(const StepArgs& a, int count, int32_t* maxcnt) {
  const int m = wave_max_i(count);
  if ((threadIdx.x & 63) == 0 && m > 0) {
    if (kCount) {
      atomicMax(maxcnt, m);
    } else if (m > a.cap) {
      atomicMax(&a.flags[1], m);
    }
  }
}

// ---- the separate kernels (first build; every one returns at once when
// the rebuild flag is clear) ----

template <class Mir>
__global__ void __launch_bounds__(kBlock) k_bin(StepArgs a, GridArgs g) {
  if (!a.flags[0]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.natoms) return;
  bin_atom<Mir>(a, g, i);
}

__global__ void __launch_bounds__(1024) k_cell_scan(const int32_t* flags, GridArgs g) {
  if (!flags[0]) return;
  scan_cells<1024>(g);
}

__global__ void __launch_bounds__(kBlock) k_scatter(StepArgs a, GridArgs g) {
  if (!a.flags[0]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.natoms) return;
  scatter_atom(g, i);
}

__global__ void __launch_bounds__(kBlock) k_order(StepArgs a, GridArgs g) {
  if (!a.flags[0]) return;
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= g.ncells) return;
  order_cell(g, r);
}

template <class R>
__global__ void __launch_bounds__(kBlock) k_permute(StepArgs a, GridArgs g, PermArgs t) {
  if (!a.flags[0]) return;
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= a.natoms) return;
  permute_atom<R>(a, g, t, s);
}

template <class R, class Mir>
__global__ void __launch_bounds__(kBlock) k_copyback(StepArgs a, GridArgs g, PermArgs t) {
  if (!a.flags[0]) return;
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= a.natoms) return;
  copyback_atom<R, Mir>(a, g, t, s);
}

template <bool kCount, class Mir>
__global__ void __launch_bounds__(kBlock)
k_nlist_ell(StepArgs a, GridArgs g, double r_list_sq, double near_sq, int32_t* nbr, int32_t* cnt,
            int32_t* maxcnt) {
  if (!kCount && !a.flags[0]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  int count = 0;
  if (i < a.natoms) count = nlist_atom<kCount, Mir>(a, g, i, r_list_sq, near_sq, nbr, cnt);
  report_count<kCount>(a, count, maxcnt);
}

// End of a rebuild: the last block to finish clears the flag (every block has
// read it by then), so the next step's kernels see a clean state.
__global__ void __launch_bounds__(kBlock) k_rebuild_done(int32_t* flags, int32_t* done,
                                                         int nblocks, int64_t* rebuilds) {
  if (!flags[0]) return;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const int prev = atomicAdd(done, 1);
    if (prev == nblocks - 1) {
      *done = 0;
      flags[4] = 1;  // the rows changed: halos are rebuilt next
      flags[0] = 0;
      *rebuilds += 1;
    }
  }
}

// ---------------------------------------------------------------------------
// Halos of the LDS-staged passes. Block b of those passes covers atoms
// [1024 b, 1024 b + 1024); its halo is the set of atoms its rows name (the
// block's own atoms included), in increasing index. A pass copies the halo's
// records into LDS once and gathers neighbour records from there, so a
// block reads ~3.5 records per atom from L2 instead of one per row slot.
//
// Built after every rebuild by one workgroup per block at a time -- as the
// last phase of the fused rebuild, or by k_halo after a rebuild by separate
// kernels (flags[4]). The set is kept as 64-atom chunks: an LDS hash maps a
// chunk id (j >> 6) to a 64-bit mask of its atoms present. Atoms are in
// cell order, so a row names its neighbours in runs of a few chunks (a
// stencil cell's atoms are adjacent indices): a thread ORs one mask per run
// into the set instead of inserting every slot. The chunk ids are sorted
// (a few hundred at most), a scan of their popcounts gives each chunk's
// first list position, the list goes out, and a row slot j becomes
// base(j >> 6) + popcount(mask(j >> 6) below bit j & 63). A halo larger
// than stage_cap (or than kHaloMax), or spread over more than kChunkMax
// chunks, leaves the block on its 32-bit rows. k_halo's last workgroup to
// finish clears flags[4].
// ---------------------------------------------------------------------------
constexpr int kHaloMax = 6144;                 // distinct atoms a staged halo may hold
constexpr int kChunkBits = 10;
constexpr int kChunkSlots = 1 << kChunkBits;   // hash slots of the chunk set
constexpr int kChunkMax = 768;                 // distinct chunks a halo may touch
static_assert(kHaloMax <= 65536 && kChunkMax <= kChunkSlots, "halo sizing");

constexpr int kHaloBlock = 1024;  // atoms per staged block (= the passes' workgroup)
static_assert(kChunkSlots == kHaloBlock, "one chunk slot per thread in the compaction");

// One slot of the chunk set: chunk id (empty = -1), first list position of
// its atoms, mask of its atoms present. 16 bytes: a probe is one LDS read.
struct alignas(16) HaloChunk {
  int32_t key;
  int32_t base;
  uint64_t mask;
};

struct HaloLds {
  HaloChunk ent[kChunkSlots];
  int32_t srt[kChunkSlots];   // chunk ids, sorted
  int32_t slot[kChunkSlots];  // hash slot of srt[p]
  int32_t wsum[kHaloBlock / 64];
  int nc, fill, ovf;
};

// Multiplicative (Fibonacci) hash of a chunk id onto the set's slots.
__device__ __forceinline__ unsigned chunk_hash(int c) {
  return (static_cast<unsigned>(c) * 2654435761u) >> (32 - kChunkBits);
}

// OR mask m into chunk c of the set (linear probing; a slot's key goes
// from -1 to its chunk once, by CAS). Counts new chunks in *nc; returns
// false once more than kChunkMax are held (the caller stops inserting) or
// the probe finds no slot.
__device__ __forceinline__ bool chunk_or(HaloChunk* ent, int c, uint64_t m, int* nc) {
  unsigned h = chunk_hash(c);
  for (int p = 0; p < kChunkSlots; ++p) {
    int k = __hip_atomic_load(&ent[h].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == -1) {
      k = atomicCAS(&ent[h].key, -1, c);
      if (k == -1) {
        if (atomicAdd(nc, 1) >= kChunkMax) return false;
        k = c;
      }
    }
    if (k == c) {
      atomicOr(reinterpret_cast<unsigned long long*>(&ent[h].mask),
               static_cast<unsigned long long>(m));
      return true;
    }
    h = (h + 1) & (kChunkSlots - 1);
  }
  return false;
}

// Slot of chunk c, which the set holds, read whole: {key, base, mask lo,
// mask hi}. The probe is bounded by the table size all the same.
__device__ __forceinline__ int4 chunk_find(const HaloChunk* ent, int c) {
  unsigned h = chunk_hash(c);
  int4 e = *reinterpret_cast<const int4*>(&ent[h]);
  for (int p = 1; e.x != c && p < kChunkSlots; ++p) {
    h = (h + 1) & (kChunkSlots - 1);
    e = *reinterpret_cast<const int4*>(&ent[h]);
  }
  return e;
}

// List position of atom j, whose chunk the set holds.
__device__ __forceinline__ uint32_t halo_pos(const HaloChunk* ent, int j) {
  const int4 e = chunk_find(ent, j >> 6);
  const uint64_t m = static_cast<uint32_t>(e.z) | (static_cast<uint64_t>(static_cast<uint32_t>(e.w)) << 32);
  const uint64_t below = (uint64_t{1} << (j & 63)) - 1u;
  return static_cast<uint32_t>(e.y + __popcll(m & below));
}

// One row slot j into the thread's current run (chunk cc, mask m); a slot
// in another chunk ORs the run into the set first (ok turns false when the
// set is full).
__device__ __forceinline__ void halo_merge(int j, int& cc, uint64_t& m, bool& ok, HaloLds& L) {
  if ((j >> 6) != cc) {
    ok = ok && chunk_or(L.ent, cc, m, &L.nc);
    cc = j >> 6;
    m = 0;
  }
  m |= uint64_t{1} << (j & 63);
}

// Halos of blocks blockIdx.x, blockIdx.x + gridDim.x, ... (< nblocks), by a
// workgroup of kHaloBlock threads; its LDS is `L`.
__device__ __forceinline__ void build_halos(const StepArgs& a, int nblocks, HaloLds& L) {
  constexpr int nt = kHaloBlock;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int cap = a.cap;
  int32_t* hj = const_cast<int32_t*>(a.hj);
  int32_t* hn = const_cast<int32_t*>(a.hn);
  int32_t* hmode = const_cast<int32_t*>(a.hmode);
  uint16_t* nbr16 = const_cast<uint16_t*>(a.nbr16);
  const int4* __restrict__ q4 = reinterpret_cast<const int4*>(a.nbr);
  for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
    L.ent[tid].key = -1;
    L.ent[tid].mask = 0;
    if (tid == 0) {
      L.nc = 0;
      L.fill = 0;
      L.ovf = 0;
    }
    __syncthreads();
    const int i = b * kHaloBlock + tid;
    if (i < a.natoms) {
      // runs of one chunk are merged in (cc, m) and ORed in when the chunk
      // changes; the row in quads, the next quad loaded while this one is
      // merged; slots past the count in the last quad are padding (i, whose
      // chunk starts the first run, since row padding names i itself)
      int cc = i >> 6;
      uint64_t m = uint64_t{1} << (i & 63);
      bool ok = true;
      // (the next quad is loaded by index, the last one reloading itself,
      // rather than selected between a load and a copy: a select of loaded
      // vectors kept them in scratch memory)
      const int nq = (a.cnt[i] + 3) >> 2;
      int4 cur = make_int4(i, i, i, i);
      if (nq > 0) cur = q4[ell(i, 0, cap) >> 2];
      for (int q = 0; q < nq && ok; ++q) {
        const int qn = q + 1 < nq ? q + 1 : q;
        const int4 nxt = q4[ell(i, 4 * qn, cap) >> 2];
        halo_merge(cur.x, cc, m, ok, L);
        halo_merge(cur.y, cc, m, ok, L);
        halo_merge(cur.z, cc, m, ok, L);
        halo_merge(cur.w, cc, m, ok, L);
        cur = nxt;
      }
      if (ok) ok = chunk_or(L.ent, cc, m, &L.nc);
      if (!ok) L.ovf = 1;
    }
    __syncthreads();
    if (L.ovf) {  // the same in every thread
      if (tid == 0) {
        hn[b] = kHaloMax + 1;
        hmode[b] = 1;
      }
      __syncthreads();  // LDS is reused by the next block
      continue;
    }
    // compaction: one slot per thread
    {
      const int k = L.ent[tid].key;
      if (k != -1) L.srt[atomicAdd(&L.fill, 1)] = k;
    }
    __syncthreads();
    const int C = L.fill;
    int P = 1;
    while (P < C) P <<= 1;
    if (tid >= C && tid < P) L.srt[tid] = 0x7fffffff;
    __syncthreads();
    // bitonic sort of srt[0, P), ascending
    for (int size = 2; size <= P; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        if (tid < (P >> 1)) {
          const int lo = 2 * tid - (tid & (stride - 1));
          const int hi = lo + stride;
          const bool up = (lo & size) == 0;
          const int x = L.srt[lo], y = L.srt[hi];
          if ((x > y) == up) {
            L.srt[lo] = y;
            L.srt[hi] = x;
          }
        }
        __syncthreads();
      }
    }
    // exclusive scan of the sorted chunks' popcounts: each chunk's base
    int h = 0, n = 0;
    if (tid < C) {
      h = static_cast<int>(chunk_hash(L.srt[tid]));
      while (L.ent[h].key != L.srt[tid]) h = (h + 1) & (kChunkSlots - 1);
      n = __popcll(L.ent[h].mask);
    }
    int incl = n;
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(incl, off, 64);
      if (lane >= off) incl += o;
    }
    if (lane == 63) L.wsum[w] = incl;
    __syncthreads();
    int base = incl - n, H = 0;
    for (int k = 0; k < nt / 64; ++k) {
      const int s = L.wsum[k];
      base += k < w ? s : 0;
      H += s;
    }
    if (H > a.stage_cap || H > kHaloMax) {  // the same in every thread
      if (tid == 0) {
        hn[b] = H;
        hmode[b] = 1;
      }
      __syncthreads();
      continue;
    }
    if (tid < C) {
      L.ent[h].base = base;
      L.slot[tid] = h;
    }
    __syncthreads();
    // the list: one wavefront per chunk, lane l writes atom 64 c + l if present
    for (int p = w; p < C; p += nt / 64) {
      const HaloChunk e = L.ent[L.slot[p]];
      if ((e.mask >> lane) & 1u) {
        const uint64_t below = (uint64_t{1} << lane) - 1u;
        hj[static_cast<size_t>(b) * kHaloMax + e.base + __popcll(e.mask & below)] =
            e.key * 64 + lane;
      }
    }
    if (i < a.natoms) {
      // the padded row in quads; a quad's four 16-bit positions are
      // adjacent in the 16-bit layout (eight slots per 16 bytes)
      const int mq = ((a.cnt[i] + 7) & ~7) >> 2;
      int4 cur = make_int4(i, i, i, i);
      if (mq > 0) cur = q4[ell(i, 0, cap) >> 2];
      for (int q = 0; q < mq; ++q) {
        const int qn = q + 1 < mq ? q + 1 : q;
        const int4 nxt = q4[ell(i, 4 * qn, cap) >> 2];
        const uint32_t p0 = halo_pos(L.ent, cur.x);
        const uint32_t p1 = halo_pos(L.ent, cur.y);
        const uint32_t p2 = halo_pos(L.ent, cur.z);
        const uint32_t p3 = halo_pos(L.ent, cur.w);
        *reinterpret_cast<uint2*>(nbr16 + ell16(i, 4 * q, cap)) =
            make_uint2(p0 | (p1 << 16), p2 | (p3 << 16));
        cur = nxt;
      }
    }
    if (tid == 0) {
      hn[b] = H;
      hmode[b] = 0;
    }
    __syncthreads();
  }
}

// Standalone halo build after a rebuild by separate kernels (flags[4]).
__global__ void __launch_bounds__(kHaloBlock) k_halo(StepArgs a, int nblocks, int32_t* done) {
  if (!a.flags[4]) return;
  __shared__ HaloLds L;
  build_halos(a, nblocks, L);
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(done, 1) == static_cast<int>(gridDim.x) - 1) {
      *done = 0;
      a.flags[4] = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// Fused rebuild: the whole sort + list sequence in one launch, its phases
// separated by a grid barrier, so a step that needs no rebuild costs one
// early-returning launch instead of nine.
//
// The grid is sized by the host to what the device holds resident at once
// (occupancy x CUs), so every workgroup of the launch runs concurrently and
// the barrier can complete. The barrier is a count + generation pair in
// device memory; its wait is bounded: a wait longer than kBarrierSpins
// polls sets flags[3] and the kernel returns (the host reports the error)
// instead of hanging.
//
// Only thread 0 of each workgroup touches the barrier. Its arrival is a
// release and its wait ends with an acquire fence at agent scope, and the
// workgroup barriers on either side extend both to the other threads: one
// L2 write-back and one invalidate per workgroup and phase, not one per
// wave. The polls themselves are relaxed loads: an acquiring poll would
// invalidate the XCD's L2 at every poll, under the workgroups still working
// (phases such as the cell scan run on one workgroup), and polls back off
// with a growing sleep.
// ---------------------------------------------------------------------------

// Arrival is two-level: the workgroups of XCD group x (blockIdx % 8 == x)
// count at sub[32 x], and the last of each group counts at `count`; the
// last of those bumps `gen`. Arrivals on one address are serialised, so 512
// workgroups on a single counter took ~30 us a barrier; eight counters of
// at most 64 arrivals each, on separate 128-byte lines, take a fraction.
struct GridBarrier {
  unsigned int count;  // groups arrived
  unsigned int gen;    // generation: bumped when every group has arrived
  unsigned int pad[30];
  unsigned int sub[8 * 32];  // workgroups arrived in group x: sub[32 x]
};

constexpr long kBarrierSpins = 1L << 22;

__device__ __forceinline__ bool grid_barrier(GridBarrier* b, int32_t* flags) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    ok = 1;
    const unsigned int gen =
        __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned int x = blockIdx.x & 7u;
    const unsigned int nx = (gridDim.x - x + 7u) >> 3;  // workgroups in group x
    const unsigned int ngroups = gridDim.x < 8u ? gridDim.x : 8u;
    unsigned int* sub = &b->sub[32u * x];
    const unsigned int arrived =
        __hip_atomic_fetch_add(sub, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    bool last = false;
    if (arrived == nx - 1u) {  // last of its group: arrive at the top
      __hip_atomic_store(sub, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned int top =
          __hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      last = top == ngroups - 1u;
    }
    if (last) {
      __hip_atomic_store(&b->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&b->gen, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      long spins = 0;
      while (__hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        if (spins < 16)
          __builtin_amdgcn_s_sleep(2);
        else if (spins < 256)
          __builtin_amdgcn_s_sleep(16);
        else
          __builtin_amdgcn_s_sleep(64);
        if (++spins > kBarrierSpins) {
          ok = 0;
          atomicExch(&flags[3], 1);
          break;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
  }
  __syncthreads();
  return ok != 0;
}

constexpr int kRebuildBlock = 1024;  // workgroup of the fused rebuild

// Registers are not capped: the list search's candidate groups need more
// than the 64 VGPRs that eight waves per SIMD (two workgroups per CU) would
// leave, and spilled under that cap (1M-atom rebuild 2.14 ms capped, 2.10
// uncapped; 0.67 / 0.58 ms at 256,000 atoms). The grid is as many
// workgroups as the device holds at once (one or two per CU).
template <class R, class Mir>
__global__ void __launch_bounds__(kRebuildBlock)
k_rebuild_fused(StepArgs a, GridArgs g, PermArgs t, double r_list_sq, double near_sq,
                int32_t* nbr, int32_t* cnt, GridBarrier* bar, int64_t* rebuilds, int forced) {
  // uniform: nothing writes the flag before the last phase
  if (!forced && !a.flags[0]) return;
  const int tid = blockIdx.x * kRebuildBlock + threadIdx.x;
  const int nthr = gridDim.x * kRebuildBlock;
  const int N = a.natoms;
  for (int r = tid; r < g.ncells; r += nthr) {
    g.ccount[r] = 0;
    g.cursor[r] = 0;
  }
  if (!grid_barrier(bar, a.flags)) return;
  for (int i = tid; i < N; i += nthr) bin_atom<Mir>(a, g, i);
  if (!grid_barrier(bar, a.flags)) return;
  if (g.ncells <= kScanOneGroup) {
    // few cells: one workgroup scans them faster than a second barrier
    if (blockIdx.x == 0) scan_cells<kRebuildBlock>(g);
  } else {
    scan_cells_grid<kRebuildBlock>(g, 1);
    if (!grid_barrier(bar, a.flags)) return;
    scan_cells_grid<kRebuildBlock>(g, 2);
  }
  if (!grid_barrier(bar, a.flags)) return;
  for (int i = tid; i < N; i += nthr) scatter_atom(g, i);
  if (!grid_barrier(bar, a.flags)) return;
  {
    // one cell per wavefront, cells dealt to the workgroups wave by wave
    const int nw = gridDim.x * (kRebuildBlock / 64);
    const int wv = (threadIdx.x >> 6) * gridDim.x + blockIdx.x;
    for (int r = wv; r < g.ncells; r += nw) order_cell_wave(g, r);
  }
  if (!grid_barrier(bar, a.flags)) return;
  for (int s = tid; s < N; s += nthr) permute_atom<R>(a, g, t, s);
  if (!grid_barrier(bar, a.flags)) return;
  for (int s = tid; s < N; s += nthr) copyback_atom<R, Mir>(a, g, t, s);
  if (!grid_barrier(bar, a.flags)) return;
  int count = 0;
  if (g.nl_split) {
    // two atoms per wavefront, 32 lanes each; the loop bound is uniform per wave
    const int gw = tid >> 6, nw = nthr >> 6;
    for (int base = 2 * gw; base < N; base += 2 * nw) {
      const int i = base + static_cast<int>((threadIdx.x >> 5) & 1);
      const bool act = i < N;
      const int c = nlist_atom_group<Mir>(a, g, act ? i : base, act, r_list_sq, near_sq, nbr,
                                          cnt);
      count = c > count ? c : count;
    }
  } else {
    for (int i = tid; i < N; i += nthr) {
      const int c = nlist_atom<false, Mir>(a, g, i, r_list_sq, near_sq, nbr, cnt);
      count = c > count ? c : count;
    }
  }
  report_count<false>(a, count, nullptr);
  if constexpr (Mir::kFixed) {
    if (a.stage_cap > 0) {  // halos of the staged passes, from the new rows
      __shared__ HaloLds L;
      if (!grid_barrier(bar, a.flags)) return;
      static_assert(kRebuildBlock == kHaloBlock, "halos are built by whole workgroups");
      build_halos(a, (N + kHaloBlock - 1) / kHaloBlock, L);
    }
  }
  // every thread read the flag at entry, before the first barrier
  if (tid == 0) {
    a.flags[0] = 0;
    *rebuilds += 1;
  }
}

}  // namespace sched
}  // namespace tdmd_hip
