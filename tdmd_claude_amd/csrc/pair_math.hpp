// pair_math.hpp -- per-pair arithmetic shared by every force kernel.
//
// Precision contract (R = float for the mixed path, double for FP64):
//   * pair difference, minimum image, r^2, the cutoff test and r itself are
//     carried in FP64. On CDNA4 an FP64 FMA issues at the rate of an
//     unpacked FP32 FMA, so this costs registers, not throughput; it removes
//     the dominant mixed-precision error, rounding r to float (the
//     reference casts the FP64 difference to float first,
//     device_morse.cu:69-74).
//   * spline coefficients, the cubic in the interval offset dx, the
//     potential functions and fpair are computed in R; dx itself keeps the
//     precision of FP64 interval arithmetic: formed in FP64 (r - knot) and
//     then rounded, or, in the mixed EAM force passes, from a float seed and
//     its FP64 residual (pair_t_f32, below).
//   * force and density sums are FP64.
// For R = float, r comes from a float square root refined by one FP64
// Newton step (error ~1e-16 relative), avoiding the slow FP64 sqrt.
#pragma once

#include "tables.hpp"

namespace tdmd_hip {

template <class R>
struct PairR {
  double r;  // FP64 distance
  R rinv;    // 1 / r in R
};

template <class R>
__device__ __forceinline__ PairR<R> pair_r(double r2);

template <>
__device__ __forceinline__ PairR<double> pair_r<double>(double r2) {
  PairR<double> p;
  p.r = sqrt(r2);
  p.rinv = 1.0 / p.r;
  return p;
}

template <>
__device__ __forceinline__ PairR<float> pair_r<float>(double r2) {
  // Hardware square root and reciprocal (v_sqrt_f32 / v_rcp_f32, about
  // 1 ulp; pair distances are far from the denormal range), then one FP64
  // Newton step r = s + (r2 - s^2) / (2 s): the step squares the seed's
  // relative error, so r is accurate to ~1e-14 whichever seed is used.
  PairR<float> p;
  const float rf = __builtin_amdgcn_sqrtf(static_cast<float>(r2));
  const double rd = static_cast<double>(rf);
  p.rinv = __builtin_amdgcn_rcpf(rf);
  p.r = fma(fma(-rd, rd, r2), static_cast<double>(0.5f * p.rinv), rd);
  return p;
}

// Density pairs under mixed precision: r, the interval and the offset in
// FP32 from the squared distance rounded to float. The density pass's pair
// values reach the forces only through rho_i, a sum over ~60 pairs: an
// offset error of ~4e-7 A (float r, float interval coordinate) moves rho_i
// by ~1e-7 relative, far inside the force tolerance, and saves the FP64
// Newton step and FP64 interval arithmetic on every pair. The force pass
// keeps the offset's full precision (pair_t_f32), since its pair terms
// scale the FP64 difference vector.
struct SplineMetaF {
  float inv_dr;
  float off;  // rmin / dr
  float dr;
  int last;   // n - 2
};

__device__ __forceinline__ SplineMetaF meta_f32(const SplineMeta& m) {
  return SplineMetaF{static_cast<float>(m.inv_dr), static_cast<float>(m.rmin * m.inv_dr),
                     static_cast<float>(m.dr), m.n - 2};
}

// Interval of the pair at squared distance r2 and the offset into it.
__device__ __forceinline__ int locate_f32(const SplineMetaF& m, double r2, float* dx) {
  const float r = __builtin_amdgcn_sqrtf(static_cast<float>(r2));
  const float u = fmaf(r, m.inv_dr, -m.off);
  int idx = static_cast<int>(u);
  idx = idx < 0 ? 0 : idx;
  idx = idx > m.last ? m.last : idx;
  *dx = (u - static_cast<float>(idx)) * m.dr;
  return idx;
}

// Force pairs under mixed precision: 1/r, the interval and the offset t in
// FP32, with t as accurate as FP64 interval arithmetic would give it.
// One hardware reciprocal square root y of float(r^2) gives 1/r and the
// seed r0 = r^2 y (~2 ulp). The seed's error delta = (r^2 - r0^2) / (2 r0)
// comes from the FP64 residual (one Newton step, carried as a float beside
// r0 rather than added in FP64). The interval is floor(r0 / dr); the offset
// t = r0 - k dr + delta is formed in FP32 with dr = dr_hi + dr_lo
// (dr_hi = float(dr)): fma(-k, dr_hi, r0) rounds once, on a number below
// ~dr, so t keeps ~1e-10 A where float r0 - k dr alone would lose ~3e-7 A.
// A pair within ~1e-6 A of a knot may take the neighbouring interval, with
// t slightly outside [0, dr): the C1 FD-Hermite cubics agree there to
// rounding. Every r spline of the tables starts at r = 0 and shares one
// grid (setfl), which this relies on.
struct ForceLocF {
  float inv_dr;
  float dr_hi, dr_lo;
  float last;  // n - 2
};

__device__ __forceinline__ ForceLocF force_loc_f32(const SplineMeta& m) {
  const float hi = static_cast<float>(m.dr);
  return ForceLocF{static_cast<float>(m.inv_dr), hi,
                   static_cast<float>(m.dr - static_cast<double>(hi)),
                   static_cast<float>(m.n - 2)};
}

struct PairT {
  float rinv;
  float t;
  int idx;
  bool below;  // the interval lies below `lo` (idx and t are then not set)
};

// lo: the first interval the caller can read (a table window); the
// interval is clamped to [lo, n - 2] in float (one v_med3_f32).
__device__ __forceinline__ PairT pair_t_f32(const ForceLocF& m, double r2, float lo = 0.0f) {
  const float f = static_cast<float>(r2);
  const float y = __builtin_amdgcn_rsqf(f);
  const float r0 = f * y;
  const double e = fma(-static_cast<double>(r0), static_cast<double>(r0), r2);
  const float delta = static_cast<float>(e) * (0.5f * y);
  const float u = r0 * m.inv_dr;
  const float kf = floorf(__builtin_amdgcn_fmed3f(u, lo, m.last));
  PairT p;
  p.t = fmaf(-kf, m.dr_lo, fmaf(-kf, m.dr_hi, r0)) + delta;
  p.rinv = y;
  p.idx = static_cast<int>(kf);
  p.below = u < lo;
  return p;
}

// Derivative of a + t (b + t (c + t d)) from b, 2c and 3d: the mixed force
// pass's form (the tables it reads hold 2c and 3d, or it scales c and d by
// 2 and 3 in float, which gives the same values).
__device__ __forceinline__ float cubic_deriv_s(float b, float c2, float d3, float t) {
  return fmaf(t, fmaf(t, d3, c2), b);
}
__device__ __forceinline__ float cubic_deriv_f(const float4& q, float t) {
  return cubic_deriv_s(q.y, 2.0f * q.z, 3.0f * q.w, t);
}

template <class R>
struct SplineVal {
  R f, df;
};

// Interval and offset: idx = floor((x - rmin) / dr) clamped to [0, n-2]
// (spline_eval, device_eam.cu:36-50), with a stored reciprocal spacing; at a
// knot the two neighbouring cubics agree in value and slope (C1 FD-Hermite),
// so the choice there changes results only by rounding.
template <class R>
__device__ __forceinline__ int spline_locate(const SplineMeta& m, double x, R* dx) {
  int idx = static_cast<int>((x - m.rmin) * m.inv_dr);
  idx = idx < 0 ? 0 : idx;
  idx = idx > m.n - 2 ? m.n - 2 : idx;
  *dx = static_cast<R>(x - (m.rmin + static_cast<double>(idx) * m.dr));
  return idx;
}

template <class R>
__device__ __forceinline__ R spline_value(const SplineMeta& m,
                                          const typename Quad<R>::type* __restrict__ q,
                                          double x) {
  R dx;
  const int idx = spline_locate<R>(m, x, &dx);
  const auto c = q[m.offset + idx];
  return c.x + dx * (c.y + dx * (c.z + dx * c.w));
}

template <class R>
__device__ __forceinline__ R spline_deriv(const SplineMeta& m,
                                          const typename Quad<R>::type* __restrict__ q,
                                          double x) {
  R dx;
  const int idx = spline_locate<R>(m, x, &dx);
  const auto c = q[m.offset + idx];
  return c.y + dx * (static_cast<R>(2) * c.z + dx * static_cast<R>(3) * c.w);
}

template <class R>
__device__ __forceinline__ SplineVal<R> spline_both(const SplineMeta& m,
                                                    const typename Quad<R>::type* __restrict__ q,
                                                    double x) {
  R dx;
  const int idx = spline_locate<R>(m, x, &dx);
  const auto c = q[m.offset + idx];
  SplineVal<R> v;
  v.f = c.x + dx * (c.y + dx * (c.z + dx * c.w));
  v.df = c.y + dx * (static_cast<R>(2) * c.z + dx * static_cast<R>(3) * c.w);
  return v;
}

// Morse pair (morse.hpp:37-55): energy and fpair = -dU/dr / r.
struct MorseCoef {
  double D, alpha, r0, rc_sq;
};

template <class R>
__device__ __forceinline__ R morse_fpair(const MorseCoef& p, const PairR<R>& pr, R* energy);

template <>
__device__ __forceinline__ double morse_fpair<double>(const MorseCoef& p, const PairR<double>& pr,
                                                      double* energy) {
  // the reference expression order, in FP64
  const double e = exp(-p.alpha * (pr.r - p.r0));
  const double ome = 1.0 - e;
  *energy = p.D * (ome * ome - 1.0);
  return -(2.0 * p.D * p.alpha * ome * e) * pr.rinv;
}

template <>
__device__ __forceinline__ float morse_fpair<float>(const MorseCoef& p, const PairR<float>& pr,
                                                    float* energy) {
  // 1 - e from expm1 keeps the relative precision near r0, where e -> 1
  const float D = static_cast<float>(p.D), a = static_cast<float>(p.alpha);
  const float em1 = expm1f(-a * static_cast<float>(pr.r - p.r0));
  const float e = 1.0f + em1;
  const float ome = -em1;
  *energy = D * (ome * ome - 1.0f);
  return -(2.0f * D * a * ome * e) * pr.rinv;
}

}  // namespace tdmd_hip
