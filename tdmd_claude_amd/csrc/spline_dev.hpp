// spline_dev.hpp -- device evaluation of FD-Hermite splines
// (spline_eval / spline_eval_deriv, src/potentials/device_eam.cu:36-68).
//
// One interval is one {a,b,c,d} quad, so a lookup is a single 16-byte
// (float) or 32-byte (double) load. The interval index multiplies by the
// stored reciprocal spacing instead of dividing; at a knot the two choices
// can pick neighbouring intervals, whose cubics agree there in value and
// slope (the FD-Hermite interpolant is C1), so results differ only by
// rounding.
#pragma once

#include "tables.hpp"

namespace tdmd_hip {

template <class R>
struct SplineVal {
  R f, df;
};

template <class R>
__device__ __forceinline__ int spline_interval(const SplineMeta<R>& m, R r, R* dx) {
  int idx = static_cast<int>((r - m.rmin) * m.inv_dr);
  idx = idx < 0 ? 0 : idx;
  idx = idx > m.n - 2 ? m.n - 2 : idx;
  *dx = r - (m.rmin + static_cast<R>(idx) * m.dr);
  return idx;
}

template <class R>
__device__ __forceinline__ R spline_value(const SplineMeta<R>& m,
                                          const typename Quad<R>::type* __restrict__ q, R r) {
  R dx;
  const int idx = spline_interval(m, r, &dx);
  const auto c = q[m.offset + idx];
  return c.x + dx * (c.y + dx * (c.z + dx * c.w));
}

template <class R>
__device__ __forceinline__ R spline_deriv(const SplineMeta<R>& m,
                                          const typename Quad<R>::type* __restrict__ q, R r) {
  R dx;
  const int idx = spline_interval(m, r, &dx);
  const auto c = q[m.offset + idx];
  return c.y + dx * (static_cast<R>(2) * c.z + dx * static_cast<R>(3) * c.w);
}

template <class R>
__device__ __forceinline__ SplineVal<R> spline_both(const SplineMeta<R>& m,
                                                    const typename Quad<R>::type* __restrict__ q,
                                                    R r) {
  R dx;
  const int idx = spline_interval(m, r, &dx);
  const auto c = q[m.offset + idx];
  SplineVal<R> v;
  v.f = c.x + dx * (c.y + dx * (c.z + dx * c.w));
  v.df = c.y + dx * (static_cast<R>(2) * c.z + dx * static_cast<R>(3) * c.w);
  return v;
}

}  // namespace tdmd_hip
