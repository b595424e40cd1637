// sched_kernels.hpp -- device side of the FastPipelineScheduler fast path.
//
// Layout (all per-atom arrays are in the scheduler's spatial order, which is
// re-derived at every neighbour-list rebuild):
//   px,py,pz / vx,vy,vz   FP64 master state (SoA, coalesced)
//   bx,by,bz              positions at the last rebuild (skin test)
//   fx,fy,fz              forces in the force type R (kept for the window
//                         boundary kick and for download)
//   M  (mirror)           gather record {x, y, z, type} written after every
//                         drift; read by the density pass (and by Morse)
//   M2 (mirror)           {x, y, z, fp} written by the density pass; read
//                         by the EAM force pass -- one load per neighbour
//                         brings the neighbour's position and F'(rho_j)
//   nbr                   neighbour indices, wave-interleaved ("ELL"): the
//                         k-th neighbour of the 64 atoms of one wavefront
//                         is one contiguous 256-byte line, so every list
//                         read of a wave is a single coalesced load
//
// Mirror encodings:
//   MirrorF64    double4; pair difference and minimum image in FP64
//   MirrorFixed  int4 of 32-bit fixed-point coordinates. On a periodic
//                axis one box length is exactly 2^32 units, so the
//                wrapping 32-bit difference of two coordinates IS the
//                minimum-image difference (no compare, no select); on an
//                open axis the unit is W/2^30 with W = max(L, r_list) and
//                coordinates must stay inside [lo - 2W, lo + 2W).
//                16 bytes per gather instead of 32; positions are rounded
//                to the grid (L/2^32 = 5.4e-8 A for a 231 A box).
// Either way the pair geometry downstream is FP64 (pair_math.hpp).
#pragma once

#include "pair_math.hpp"

namespace tdmd_hip {
namespace sched {

constexpr int kMaxTypes = 8;
constexpr int kMaxSplines = 64;
// Several-element EAM with at most this many elements: the density pass
// stores F'(rho) in the force pass's records with its two lowest mantissa
// bits replaced by the atom's element index (pack_typed), and the force pass
// takes a neighbour's element from its record instead of gathering
// types[j] (one scattered load per neighbour slot). F' is then used with a
// relative perturbation below 4e-7 (FP32) / 7e-16 (FP64), the same value
// for the atom's own term and its neighbours'.
constexpr int kPackedTypes = 4;

// Spare records allocated past the last atom in each gather-record buffer:
// the list search loads a cell's candidates in whole groups of eight from
// the cell's first atom, masking the slots past the cell's end.
constexpr int kRecPad = 8;

struct MirrorParams {
  double lo[3];
  double q[3];     // fixed-point units per A (MirrorFixed)
  double unit[3];  // A per unit, exact: L 2^-32 (periodic) or W 2^-30 (open)
  double L[3];
  double hL[3];
  int periodic[3];
};

// Minimum image as selects (no divergent branches); same values as
// min_image() in common.hpp.
__device__ __forceinline__ double min_image_sel(double d, double L, double hL, int periodic) {
  const double up = d - L, dn = d + L;
  const double m = d > hL ? up : (d < -hL ? dn : d);
  return periodic ? m : d;
}

struct MirrorF64 {
  using T = double4;
  static constexpr bool kMB = false;
  static constexpr bool kFixed = false;
  __device__ __forceinline__ static T encode(const MirrorParams&, double x, double y, double z,
                                             double w) {
    return make_double4(x, y, z, w);
  }
  __device__ __forceinline__ static T with_w(T m, double w) {
    m.w = w;
    return m;
  }
  // Pair difference a - b with the minimum image, in FP64; returns r^2.
  __device__ __forceinline__ static double delta(const T& a, const T& b, const MirrorParams& mp,
                                                 double* dx, double* dy, double* dz) {
    *dx = min_image_sel(a.x - b.x, mp.L[0], mp.hL[0], mp.periodic[0]);
    *dy = min_image_sel(a.y - b.y, mp.L[1], mp.hL[1], mp.periodic[1]);
    *dz = min_image_sel(a.z - b.z, mp.L[2], mp.hL[2], mp.periodic[2]);
    return *dx * *dx + *dy * *dy + *dz * *dz;
  }
  __device__ __forceinline__ static bool representable(const MirrorParams&, double, double,
                                                       double) {
    return true;
  }
  __device__ __forceinline__ static int type_of(const T& m) { return static_cast<int>(m.w); }
  template <class R>
  __device__ __forceinline__ static R w_real(const T& m) {
    return static_cast<R>(m.w);
  }
  template <class R>
  __device__ __forceinline__ static double pack_real(R v) {
    return static_cast<double>(v);
  }
  // v with the element index t (0..3) in its two lowest mantissa bits
  template <class R>
  __device__ __forceinline__ static double pack_typed(R v, int t) {
    return __longlong_as_double((__double_as_longlong(static_cast<double>(v)) & ~3ll) | t);
  }
  __device__ __forceinline__ static int packed_type(const T& m) {
    return static_cast<int>(__double_as_longlong(m.w) & 3);
  }
};

struct MirrorFixed {
  using T = int4;
  static constexpr bool kMB = false;
  // build positions (skin test) kept as the 32-bit coordinates, in the
  // first 4 bytes per atom of bx/by/bz
  static constexpr bool kFixed = true;
  // Coordinate in units, reduced mod 2^32 (a periodic image on a periodic
  // axis; an open axis stays in range, see representable()).
  __device__ __forceinline__ static int enc1(double x, double lo, double q) {
    const long long v = __double2ll_rn((x - lo) * q);
    return static_cast<int>(static_cast<unsigned int>(static_cast<unsigned long long>(v)));
  }
  __device__ __forceinline__ static T encode(const MirrorParams& mp, double x, double y,
                                             double z, int w) {
    T m;
    m.x = enc1(x, mp.lo[0], mp.q[0]);
    m.y = enc1(y, mp.lo[1], mp.q[1]);
    m.z = enc1(z, mp.lo[2], mp.q[2]);
    m.w = w;
    return m;
  }
  __device__ __forceinline__ static bool rep1(double x, double lo, double q, int periodic) {
    const double u = (x - lo) * q;
    return periodic || (u >= -2147483648.0 && u < 2147483647.0);
  }
  __device__ __forceinline__ static bool representable(const MirrorParams& mp, double x,
                                                       double y, double z) {
    return rep1(x, mp.lo[0], mp.q[0], mp.periodic[0]) &&
           rep1(y, mp.lo[1], mp.q[1], mp.periodic[1]) &&
           rep1(z, mp.lo[2], mp.q[2], mp.periodic[2]);
  }
  __device__ __forceinline__ static T with_w(T m, int w) {
    m.w = w;
    return m;
  }
  __device__ __forceinline__ static int wdiff(int a, int b) {
    return static_cast<int>(static_cast<unsigned int>(a) - static_cast<unsigned int>(b));
  }
  // Wrapping 32-bit difference (the minimum image on a periodic axis),
  // scaled to A in FP64; returns r^2.
  __device__ __forceinline__ static double delta(const T& a, const T& b, const MirrorParams& mp,
                                                 double* dx, double* dy, double* dz) {
    *dx = static_cast<double>(wdiff(a.x, b.x)) * mp.unit[0];
    *dy = static_cast<double>(wdiff(a.y, b.y)) * mp.unit[1];
    *dz = static_cast<double>(wdiff(a.z, b.z)) * mp.unit[2];
    return *dx * *dx + *dy * *dy + *dz * *dz;
  }
  __device__ __forceinline__ static int type_of(const T& m) { return m.w; }
  template <class R>
  __device__ __forceinline__ static R w_real(const T& m) {
    return __int_as_float(m.w);
  }
  template <class R>
  __device__ __forceinline__ static int pack_real(R v) {
    return __float_as_int(static_cast<float>(v));
  }
  // v with the element index t (0..3) in its two lowest mantissa bits
  template <class R>
  __device__ __forceinline__ static int pack_typed(R v, int t) {
    return (__float_as_int(static_cast<float>(v)) & ~3) | t;
  }
  __device__ __forceinline__ static int packed_type(const T& m) { return m.w & 3; }
};

// A batch of independent boxes (replicas) advanced together: the same
// mirror encoding, with each atom's box looked up from its replica id.
template <class Base>
struct ManyBoxes : Base {
  static constexpr bool kMB = true;
};

// One box of a batch: its mirror parameters and binning grid; its cells are
// global cells [cell_off, cell_off + n0 n1 n2).
struct BoxGeo {
  MirrorParams mp;
  int n[3];
  double inv_cs[3];
  int cell_off;
  int images;  // 1: the list search may take stencil images (>= 3 cells of
               // >= r_list on every periodic axis)
};

// ELL index of the k-th neighbour of atom i. A wavefront's 64 rows are
// stored together, in quads: slots 4q..4q+3 of one row are adjacent (16
// bytes), and quad q of the 64 rows follows lane order, so a wave reads
// four slots of each of its rows with one coalesced 1 KiB load. cap is a
// multiple of 4.
__device__ __forceinline__ size_t ell(int i, int k, int cap) {
  return ((static_cast<size_t>(i >> 6) * (cap >> 2) + (k >> 2)) * 64 + (i & 63)) * 4 + (k & 3);
}

// Slot k of atom i in the 16-bit rows: eight slots of a row per 16 bytes,
// the 64 rows of a wavefront side by side (cap is a multiple of 8).
__device__ __forceinline__ size_t ell16(int i, int k, int cap) {
  return ((static_cast<size_t>(i >> 6) * (cap >> 3) + (k >> 3)) * 64 + (i & 63)) * 8 + (k & 7);
}

// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// run of atoms (spatially sorted) so neighbour gathers stay in its L2.
__device__ __forceinline__ int xcd_block(int nblocks) {
  const int b = blockIdx.x;
  const int per = nblocks >> 3;  // nblocks is a multiple of 8
  return (b & 7) * per + (b >> 3);
}

// Epilogue of the force pass.
enum Epilogue : int {
  kForceOnly = 0,    // store F (initial forces, thermo)
  kFullKickDrift = 1,// v += dt F/(m mvv2e); x += dt v (cross-step fused kick)
  kHalfKick = 2,     // v += dt/2 F/(m mvv2e); store F (window end)
  kKick2Ke = 3,      // NVT: v += dt/2 F/m; store F, v; KE partials
  kKe = 4            // NVT inside a window: KE partials of v + dt/2 F/m, store F
                     // only; the next kick-drift applies that half kick itself
};

// Epilogues that emit per-block KE partials.
constexpr bool emits_ke(int epi) { return epi == kKick2Ke || epi == kKe; }

// Nose-Hoover chain constants (device_nose_hoover.cu:206-289).
struct NhcParams {
  int M;
  int n_dof;
  double t_target, t_period, dt;
};

struct StepArgs {
  int natoms;
  int nblocks;  // grid size (multiple of 8)
  int cap;
  int ntypes;
  int one_type;  // 1: every atom is of type 1 (the integrator skips types[])
  double dt;
  double skin_half_sq;
  MirrorParams mp;
  // state
  double *px, *py, *pz, *vx, *vy, *vz, *bx, *by, *bz;
  void *fx, *fy, *fz;
  const int32_t* types;
  const double* masses;  // by 1-based type
  const int32_t* nbr;
  const int32_t* cnt;
  // Halos of the LDS-staged passes (blocks of 1024 atoms): halo b lists the
  // distinct atoms the rows of block b name, in increasing index; nbr16
  // holds those rows as positions in the list.
  const uint16_t* nbr16;
  const int32_t* hj;     // [blocks][kHaloMax]
  const int32_t* hn;     // [blocks] halo size
  const int32_t* hmode;  // [blocks] 0: staged in LDS, 1: reads the 32-bit rows
  int stage_cap;         // halo records a pass holds in LDS (0: no staging)
  // Staged blocks: each density pass lists, in row order, the skin-shell
  // slots (slot >= cntn) that are inside the cutoff (layout of nbr16, count
  // cntc); the force pass of the same positions reads the near slots of the
  // row and then this list.
  int32_t* cntn;  // [N] near slots of each row (inside the cutoff at build)
  uint16_t* nbrc16;
  int32_t* cntc;
  void* M;
  void* M2;
  int32_t* flags;        // [0] rebuild, [1] overflow max, [2] nonfinite,
                         // [3] barrier timeout, [4] halos stale
  double* partials;      // per-block FP64 partial sums
  double* partials_w;    // per-block virial partial sums (energy passes)
  double* nhc;           // device NHC state block
  // batches of boxes (ManyBoxes mirrors only)
  const BoxGeo* boxes;   // [nrep]
  const int32_t* rep;    // [N] replica of each slot (static: the sort keeps
                         // every replica in its own slot range)
  // Pass clock (measurement aid, null when off): 4 words per pass kind
  // (0 density, 1 force): earliest workgroup start of the running launch,
  // workgroups finished, sum of launch spans, launches (device wall clock)
  unsigned long long* pclock;
  // Several elements: per ordered element pair (ti, tj) and r interval, the
  // derivative coefficients the force pass needs, interleaved -- {phi_titj
  // b c d, rho_tj b | rho_tj c d, rho_ti b c | rho_ti d, 0, 0, 0} -- three
  // quads per interval instead of one quad from each of three tables; in
  // mixed precision 2c and 3d in place of c and d (null: the tables do not
  // share one r grid)
  const void* pairtab;
  int pair_nint;  // r intervals per element pair
};

// Pass clock: every workgroup stamps the device wall clock at entry; the
// last workgroup to finish a launch adds the span from the earliest stamp
// to its own finish to the running total and re-arms the words for the
// next launch. Device-scope atomics only (no fences: a release fence here
// would write back the XCD's L2 once per workgroup); every thread of the
// workgroup must reach pass_clock_end (it holds a workgroup barrier).
__device__ __forceinline__ void pass_clock_begin(unsigned long long* pc) {
  if (pc && threadIdx.x == 0)
    __hip_atomic_fetch_min(pc, static_cast<unsigned long long>(wall_clock64()), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void pass_clock_end(unsigned long long* pc) {
  if (!pc) return;
  __syncthreads();
  if (threadIdx.x == 0 &&
      __hip_atomic_fetch_add(pc + 1, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
          static_cast<unsigned long long>(gridDim.x) - 1) {
    const unsigned long long now = static_cast<unsigned long long>(wall_clock64());
    const unsigned long long t0 = __hip_atomic_load(pc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(pc + 2, now - t0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(pc + 3, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(pc, ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(pc + 1, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Mirror parameters of the box atom i belongs to.
template <class Mir>
__device__ __forceinline__ const MirrorParams& mp_of(const StepArgs& a, int i) {
  if constexpr (Mir::kMB)
    return a.boxes[a.rep[i]].mp;
  else
    return a.mp;
}

}  // namespace sched
}  // namespace tdmd_hip
