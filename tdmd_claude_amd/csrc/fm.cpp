// fm.cpp -- force matching on the GPU: the loss L = sum_i |F_i - T_i|^2
// over a fixed set of configurations (snapshots) and its adjoint gradient
// with respect to the Morse parameters (D, alpha, r0) or the knot values of a
// single-element EAM (F(rho), rho(r), phi(r) as EamAlloy holds them). No
// reference counterpart: this is the fitting path built on the same pair
// arithmetic as the force kernels (pair_math.hpp).
//
// Layout: the snapshots are one concatenated system of nsnap * natoms atoms
// (AoS FP64 positions) sharing one box; each snapshot's full CSR list is
// built once at creation with global indices, so every pass is a single
// launch over all snapshots.
//
// Adjoint (ordered pairs i -> j, d = x_i - x_j, u = d / r, g = F - T):
//   Morse:  dL/dtheta = sum 2 (g_i . d) dfpair/dtheta        (one pass)
//   EAM:    w_ij = -2 (g_i - g_j) . u_ij
//           phi'  coefficients  (b,c,d)[k(r)] += w/2 * (1, 2t, 3t^2)     per ordered pair,
//           rho'  coefficients  (b,c,d)[k(r)] += w/2 (Fp_i+Fp_j) (1, 2t, 3t^2),
//           lam_i = sum_j w_ij rho'(r_ij),
//           F     coefficients  (b,c,d)[k(rho_i)] += lam_i (1, 2t, 3t^2),
//           mu_i = lam_i F''(rho_i),
//           rho   coefficients  (a,b,c,d)[k(r)] += mu_i (1, t, t^2, t^3).
//   (w/2 per ordered pair: the two orderings of a pair give w in total.)
// Coefficient gradients go back to knot values through the transpose of the
// FD-Hermite construction (eam_alloy.cpp:19-50).
//
// Mixed precision: pair geometry and r in FP64, spline and Morse arithmetic
// in FP32 (as the forward kernels), per-atom sums FP64; coefficient
// histograms are accumulated in FP32 in LDS per workgroup and flushed with
// FP64 atomics, so the gradient is reproducible to FP32 rounding of the
// per-workgroup partial sums, not bit for bit. FP64 mode accumulates straight
// into FP64 (global atomics).
#include <cstring>
#include <string>
#include <vector>

#include "pair_math.hpp"

struct tdmd_fm {
  int precision = TDMD_MIXED;
  int natoms = 0, nsnap = 0, ntot = 0;
  tdmd_box box{};
  double r_list = 0;
  double* d_pos = nullptr;     // [ntot][3]
  int32_t* d_cnt = nullptr;    // [ntot]
  int32_t* d_off = nullptr;    // [ntot]
  int32_t* d_nbr = nullptr;    // [total]
  int32_t* d_types = nullptr;  // [ntot], all 1
  double* d_target = nullptr;  // [ntot][3]
  double* d_g = nullptr;       // [ntot][3] residual F - T
  void* d_force = nullptr;     // [ntot][3] in R
  double* d_rho = nullptr;     // [ntot]
  void* d_fp = nullptr;        // [ntot] in R
  double* d_partials = nullptr;
  int nparts = 0;
  double* d_scalars = nullptr;  // [8]
  bool has_target = false;
  // EAM scratch
  tdmd_eam_tables* tables = nullptr;  // geometry + quads rewritten per call
  double* d_coef = nullptr;           // [4 niF + 8 niR] FP64 coefficients
  double* d_gcoef = nullptr;          // [4 niF + 8 niR]
  double* d_gS = nullptr;             // [niF + 2 niR]
  int nrho = 0, nr = 0;
  double drho = 0, dr = 0, cutoff = 0;
};

namespace tdmd_hip {

namespace {

constexpr int kFmBlock = 1024;  // adjoint workgroup (LDS histograms)
constexpr int kHist = 7;        // phi b,c,d; rho a,b,c,d

__device__ __forceinline__ double fm_geom(const double* __restrict__ pos, double xi, double yi,
                                          double zi, int j, const DevBox& box, double* dx,
                                          double* dy, double* dz) {
  *dx = min_image(xi - pos[3 * j], box.size[0], 0.5 * box.size[0], box.periodic[0]);
  *dy = min_image(yi - pos[3 * j + 1], box.size[1], 0.5 * box.size[1], box.periodic[1]);
  *dz = min_image(zi - pos[3 * j + 2], box.size[2], 0.5 * box.size[2], box.periodic[2]);
  return *dx * *dx + *dy * *dy + *dz * *dz;
}

template <class R>
__device__ __forceinline__ R fm_exp(R x);
template <>
__device__ __forceinline__ float fm_exp<float>(float x) {
  return expf(x);
}
template <>
__device__ __forceinline__ double fm_exp<double>(double x) {
  return exp(x);
}

// ---------------------------------------------------------------------------
// Morse: forces, residual, loss and dL/d(D, alpha, r0) in one pass.
// partials[blk * 4 + {0: L, 1: dD, 2: dalpha, 3: dr0}]
// ---------------------------------------------------------------------------
template <class R>
__global__ void __launch_bounds__(kBlock)
k_fm_morse(const double* __restrict__ pos, const int32_t* __restrict__ nbr,
           const int32_t* __restrict__ off, const int32_t* __restrict__ cnt, int ntot, DevBox box,
           MorseCoef p, const double* __restrict__ target, int has_target,
           double* __restrict__ forces_out, double* __restrict__ partials) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double l = 0.0, gD = 0.0, ga = 0.0, gr = 0.0;
  if (i < ntot) {
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const int o = off[i], c = cnt[i];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int k = 0; k < c; ++k) {
      double dx, dy, dz;
      const double r2 = fm_geom(pos, xi, yi, zi, nbr[o + k], box, &dx, &dy, &dz);
      if (r2 < p.rc_sq) {
        const PairR<R> pr = pair_r<R>(r2);
        R e;
        const double fpair = static_cast<double>(morse_fpair<R>(p, pr, &e));
        fx = fma(fpair, dx, fx);
        fy = fma(fpair, dy, fy);
        fz = fma(fpair, dz, fz);
      }
    }
    // forces as the product path stores them (R)
    const double Fx = static_cast<double>(static_cast<R>(fx));
    const double Fy = static_cast<double>(static_cast<R>(fy));
    const double Fz = static_cast<double>(static_cast<R>(fz));
    if (forces_out) {
      forces_out[3 * i] = Fx;
      forces_out[3 * i + 1] = Fy;
      forces_out[3 * i + 2] = Fz;
    }
    if (has_target) {
      const double gx = Fx - target[3 * i], gy = Fy - target[3 * i + 1],
                   gz = Fz - target[3 * i + 2];
      l = gx * gx + gy * gy + gz * gz;
      const R D = static_cast<R>(p.D), al = static_cast<R>(p.alpha);
      for (int k = 0; k < c; ++k) {
        double dx, dy, dz;
        const double r2 = fm_geom(pos, xi, yi, zi, nbr[o + k], box, &dx, &dy, &dz);
        if (r2 < p.rc_sq) {
          const PairR<R> pr = pair_r<R>(r2);
          const R x = static_cast<R>(pr.r - p.r0);
          const R e = fm_exp<R>(-al * x);
          const R one = static_cast<R>(1), two = static_cast<R>(2);
          const R ome = one - e, o2e = one - two * e;
          const R dfdD = -two * al * ome * e * pr.rinv;
          const R dfda = -two * D * e * (ome - al * o2e * x) * pr.rinv;
          const R dfdr = -two * D * al * al * o2e * e * pr.rinv;
          const double w = 2.0 * (gx * dx + gy * dy + gz * dz);
          gD = fma(w, static_cast<double>(dfdD), gD);
          ga = fma(w, static_cast<double>(dfda), ga);
          gr = fma(w, static_cast<double>(dfdr), gr);
        }
      }
    }
  }
  const double s0 = block_sum<kBlock>(l);
  const double s1 = block_sum<kBlock>(gD);
  const double s2 = block_sum<kBlock>(ga);
  const double s3 = block_sum<kBlock>(gr);
  if (threadIdx.x == 0) {
    partials[4 * blockIdx.x] = s0;
    partials[4 * blockIdx.x + 1] = s1;
    partials[4 * blockIdx.x + 2] = s2;
    partials[4 * blockIdx.x + 3] = s3;
  }
}

// Fixed-order sum of `ncomp` interleaved partial streams.
__global__ void k_fm_reduce(const double* __restrict__ partials, int nblk, int ncomp,
                            double* __restrict__ out) {
  __shared__ double s[256];
  for (int c = 0; c < ncomp; ++c) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) t += partials[b * ncomp + c];
    s[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = s[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// EAM
// ---------------------------------------------------------------------------

// FD-Hermite coefficients from knot values (build_fd_hermite on the device,
// same operation order, no contraction): spline s = 0 F, 1 rho, 2 phi.
// coef64: blocks [spline][4][ni]; quads: F, rho, phi intervals in a row.
template <class R>
__global__ void k_fm_build_coeffs(const double* __restrict__ vals, int nrho, double drho, int nr,
                                  double dr, double* __restrict__ coef64,
                                  typename Quad<R>::type* __restrict__ quads) {
#pragma clang fp contract(off)
  const int s = blockIdx.y;
  const int n = s == 0 ? nrho : nr;
  const double h = s == 0 ? drho : dr;
  const double* v = vals + (s == 0 ? 0 : (s == 1 ? nrho : nrho + nr));
  const int ni = n - 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  const double slope_left = (i > 0) ? (v[i] - v[i - 1]) / h : 0.0;
  const double slope_right = (v[i + 1] - v[i]) / h;
  const double slope_next = (i + 2 < n) ? (v[i + 2] - v[i + 1]) / h : slope_right;
  double m0 = (slope_left + slope_right) * 0.5;
  double m1 = (slope_right + slope_next) * 0.5;
  if (i == 0) m0 = slope_right;
  if (i == n - 2) m1 = slope_right;
  const double a = v[i], b = m0, c = (3.0 * slope_right - 2.0 * m0 - m1) / h,
               d = (m0 + m1 - 2.0 * slope_right) / (h * h);
  const int niF = nrho - 1, niR = nr - 1;
  const size_t base = s == 0 ? 0 : (s == 1 ? 4 * static_cast<size_t>(niF)
                                           : 4 * static_cast<size_t>(niF) + 4 * niR);
  coef64[base + i] = a;
  coef64[base + ni + i] = b;
  coef64[base + 2 * ni + i] = c;
  coef64[base + 3 * ni + i] = d;
  const int qo = s == 0 ? 0 : (s == 1 ? niF : niF + niR);
  typename Quad<R>::type q;
  q.x = static_cast<R>(a);
  q.y = static_cast<R>(b);
  q.z = static_cast<R>(c);
  q.w = static_cast<R>(d);
  quads[qo + i] = q;
}

// Residual g = F - T (F as stored, in R), loss partials (one per block).
template <class R>
__global__ void __launch_bounds__(kBlock)
k_fm_residual(const R* __restrict__ f, const double* __restrict__ target, int ntot,
              double* __restrict__ g, double* __restrict__ forces_out,
              double* __restrict__ partials, int has_target) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double l = 0.0;
  if (i < ntot) {
    for (int c = 0; c < 3; ++c) {
      const double F = static_cast<double>(f[3 * i + c]);
      if (forces_out) forces_out[3 * i + c] = F;
      if (has_target) {
        const double gc = F - target[3 * i + c];
        g[3 * i + c] = gc;
        l += gc * gc;
      }
    }
  }
  const double s = block_sum<kBlock>(l);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

struct FmSplines {
  SplineMeta F, rho, phi;
  int niF, niR;
};

template <bool kLds>
__device__ __forceinline__ void hist_add(float* hist, double* gglob, int which, int k, int ni,
                                         double v) {
  // which: 0..2 phi b,c,d ; 3..6 rho a,b,c,d
  if (kLds) {
    atomicAdd(&hist[which * ni + k], static_cast<float>(v));
  } else {
    // global layout after the F block: rho{a,b,c,d}, phi{a,b,c,d}
    const size_t blk = which < 3 ? static_cast<size_t>(4 + 1 + which) : static_cast<size_t>(which - 3);
    atomicAdd(&gglob[blk * ni + k], v);
  }
}

template <class R, bool kLds>
__global__ void __launch_bounds__(kFmBlock)
k_fm_eam_adjoint(const double* __restrict__ pos, const int32_t* __restrict__ nbr,
                 const int32_t* __restrict__ off, const int32_t* __restrict__ cnt, int ntot,
                 DevBox box, double rc_sq, const typename Quad<R>::type* __restrict__ quads,
                 FmSplines sp, const double* __restrict__ g, const double* __restrict__ rho,
                 const R* __restrict__ fp, double* __restrict__ gcoef) {
  extern __shared__ float hist[];
  const int niF = sp.niF, niR = sp.niR;
  double* gF = gcoef;                                  // [4][niF]
  double* gRP = gcoef + 4 * static_cast<size_t>(niF);  // rho[4][niR], phi[4][niR]
  if (kLds) {
    for (int q = threadIdx.x; q < kHist * niR; q += kFmBlock) hist[q] = 0.0f;
    __syncthreads();
  }
  for (int base = blockIdx.x * kFmBlock; base < ntot; base += gridDim.x * kFmBlock) {
    const int i = base + threadIdx.x;
    if (i >= ntot) continue;
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const double gix = g[3 * i], giy = g[3 * i + 1], giz = g[3 * i + 2];
    const R fpi = fp[i];
    const int o = off[i], c = cnt[i];
    double lam = 0.0;
    for (int k = 0; k < c; ++k) {
      const int j = nbr[o + k];
      double dx, dy, dz;
      const double r2 = fm_geom(pos, xi, yi, zi, j, box, &dx, &dy, &dz);
      if (r2 < rc_sq) {
        const PairR<R> pr = pair_r<R>(r2);
        R t;
        const int q = spline_locate<R>(sp.rho, pr.r, &t);
        const auto cr = quads[sp.rho.offset + q];
        const R drho = cr.y + t * (static_cast<R>(2) * cr.z + t * static_cast<R>(3) * cr.w);
        const double w = -2.0 *
                         ((gix - g[3 * j]) * dx + (giy - g[3 * j + 1]) * dy +
                          (giz - g[3 * j + 2]) * dz) /
                         pr.r;
        const double hw = 0.5 * w;  // the two orderings of the pair give w
        const double P = static_cast<double>(fpi + fp[j]);
        const double td = static_cast<double>(t);
        hist_add<kLds>(hist, gRP, 0, q, niR, hw);
        hist_add<kLds>(hist, gRP, 1, q, niR, 2.0 * td * hw);
        hist_add<kLds>(hist, gRP, 2, q, niR, 3.0 * td * td * hw);
        hist_add<kLds>(hist, gRP, 4, q, niR, hw * P);
        hist_add<kLds>(hist, gRP, 5, q, niR, 2.0 * td * hw * P);
        hist_add<kLds>(hist, gRP, 6, q, niR, 3.0 * td * td * hw * P);
        lam = fma(w, static_cast<double>(drho), lam);
      }
    }
    // F'(rho_i) coefficients and F''(rho_i)
    R tF;
    const int qF = spline_locate<R>(sp.F, rho[i], &tF);
    const double tFd = static_cast<double>(tF);
    atomicAdd(&gF[niF + qF], lam);
    atomicAdd(&gF[2 * niF + qF], 2.0 * tFd * lam);
    atomicAdd(&gF[3 * niF + qF], 3.0 * tFd * tFd * lam);
    const auto cF = quads[sp.F.offset + qF];
    const double mu =
        lam * static_cast<double>(static_cast<R>(2) * cF.z + static_cast<R>(6) * cF.w * tF);
    for (int k = 0; k < c; ++k) {
      double dx, dy, dz;
      const double r2 = fm_geom(pos, xi, yi, zi, nbr[o + k], box, &dx, &dy, &dz);
      if (r2 < rc_sq) {
        const PairR<R> pr = pair_r<R>(r2);
        R t;
        const int q = spline_locate<R>(sp.rho, pr.r, &t);
        const double td = static_cast<double>(t);
        hist_add<kLds>(hist, gRP, 3, q, niR, mu);
        hist_add<kLds>(hist, gRP, 4, q, niR, mu * td);
        hist_add<kLds>(hist, gRP, 5, q, niR, mu * td * td);
        hist_add<kLds>(hist, gRP, 6, q, niR, mu * td * td * td);
      }
    }
  }
  if (kLds) {
    __syncthreads();
    for (int q = threadIdx.x; q < kHist * niR; q += kFmBlock) {
      const float v = hist[q];
      if (v != 0.0f) {
        const int which = q / niR, k = q - which * niR;
        const size_t blk = which < 3 ? static_cast<size_t>(4 + 1 + which)
                                     : static_cast<size_t>(which - 3);
        atomicAdd(&gRP[blk * niR + k], static_cast<double>(v));
      }
    }
  }
}

// Transpose of the FD-Hermite map, step 1: gS[k] per interval.
__global__ void k_fm_coef_grad_s(const double* __restrict__ gcoef, int nrho, double drho, int nr,
                                 double dr, double* __restrict__ gS) {
  const int s = blockIdx.y;
  const int niF = nrho - 1, niR = nr - 1;
  const int ni = s == 0 ? niF : niR;
  const double h = s == 0 ? drho : dr;
  const double* gc4 = gcoef + (s == 0 ? 0 : (s == 1 ? 4 * static_cast<size_t>(niF)
                                                    : 4 * static_cast<size_t>(niF) + 4 * niR));
  double* out = gS + (s == 0 ? 0 : (s == 1 ? niF : niF + niR));
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ni) return;
  const double* gb = gc4 + ni;
  const double* gc = gc4 + 2 * ni;
  const double* gd = gc4 + 3 * ni;
  auto gm0 = [&](int q) { return gb[q] - 2.0 * gc[q] / h + gd[q] / (h * h); };
  auto gm1 = [&](int q) { return -gc[q] / h + gd[q] / (h * h); };
  double v = 3.0 * gc[k] / h - 2.0 * gd[k] / (h * h);
  v += (k == 0) ? gm0(0) : 0.5 * gm0(k);
  if (k + 1 <= ni - 1) v += 0.5 * gm0(k + 1);
  v += (k == ni - 1) ? gm1(k) : 0.5 * gm1(k);
  if (k >= 1) v += 0.5 * gm1(k - 1);
  out[k] = v;
}

// Step 2: knot-value gradient gv[m].
__global__ void k_fm_coef_grad_v(const double* __restrict__ gcoef, const double* __restrict__ gS,
                                 int nrho, double drho, int nr, double dr,
                                 double* __restrict__ gv) {
  const int s = blockIdx.y;
  const int niF = nrho - 1, niR = nr - 1;
  const int n = s == 0 ? nrho : nr, ni = n - 1;
  const double h = s == 0 ? drho : dr;
  const double* ga = gcoef + (s == 0 ? 0 : (s == 1 ? 4 * static_cast<size_t>(niF)
                                                   : 4 * static_cast<size_t>(niF) + 4 * niR));
  const double* S = gS + (s == 0 ? 0 : (s == 1 ? niF : niF + niR));
  double* out = gv + (s == 0 ? 0 : (s == 1 ? nrho : nrho + nr));
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double v = 0.0;
  if (m <= ni - 1) v += ga[m] - S[m] / h;
  if (m >= 1) v += S[m - 1] / h;
  out[m] = v;
}

__global__ void k_fm_offset(int32_t* __restrict__ nbr, const int32_t* __restrict__ off,
                            const int32_t* __restrict__ cnt, int n, int32_t shift) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int o = off[i], c = cnt[i];
  for (int k = 0; k < c; ++k) nbr[o + k] += shift;
}

int fm_free(tdmd_fm* F) {
  void* ptrs[] = {F->d_pos, F->d_cnt,  F->d_off,      F->d_nbr,     F->d_types,
                  F->d_target, F->d_g, F->d_force,    F->d_rho,     F->d_fp,
                  F->d_partials, F->d_scalars, F->d_coef, F->d_gcoef, F->d_gS};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (F->tables) eam_tables_destroy(F->tables);
  delete F;
  return TDMD_OK;
}

template <class T>
int fm_alloc(T** p, size_t n) {
  void* q = nullptr;
  const int rc = tdmd_malloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  *p = static_cast<T*>(q);
  return rc;
}

}  // namespace
}  // namespace tdmd_hip

using namespace tdmd_hip;

extern "C" {

int tdmd_fm_create(tdmd_fm** out, int32_t precision, const tdmd_box* box, int32_t natoms,
                   int32_t nsnap, const double* pos, double r_cut) {
  if (!out || !box || !pos || natoms <= 0 || nsnap <= 0 || !(r_cut > 0))
    return fail(TDMD_E_INVALID, "fm_create: bad arguments");
  if (precision != TDMD_MIXED && precision != TDMD_FP64)
    return fail(TDMD_E_INVALID, "fm_create: bad precision");
  const int64_t ntot64 = static_cast<int64_t>(natoms) * nsnap;
  if (ntot64 > (1LL << 30)) return fail(TDMD_E_INVALID, "fm_create: too many atoms");
  auto* F = new tdmd_fm();
  F->precision = precision;
  F->natoms = natoms;
  F->nsnap = nsnap;
  F->ntot = static_cast<int>(ntot64);
  F->box = *box;
  F->r_list = r_cut;
  const int N = F->ntot;
  const size_t rsz = precision == TDMD_MIXED ? 4 : 8;
  int rc = 0;
  if ((rc = fm_alloc(&F->d_pos, 3 * static_cast<size_t>(N))) ||
      (rc = fm_alloc(&F->d_cnt, N)) || (rc = fm_alloc(&F->d_off, N)) ||
      (rc = fm_alloc(&F->d_types, N)) || (rc = fm_alloc(&F->d_target, 3 * static_cast<size_t>(N))) ||
      (rc = fm_alloc(&F->d_g, 3 * static_cast<size_t>(N))) ||
      (rc = fm_alloc(reinterpret_cast<char**>(&F->d_force), 3 * rsz * N)) ||
      (rc = fm_alloc(&F->d_rho, N)) ||
      (rc = fm_alloc(reinterpret_cast<char**>(&F->d_fp), rsz * N)) ||
      (rc = fm_alloc(&F->d_scalars, 8))) {
    fm_free(F);
    return rc;
  }
  F->nparts = (N + kBlock - 1) / kBlock;
  if ((rc = fm_alloc(&F->d_partials, 4 * static_cast<size_t>(F->nparts) + 64))) {
    fm_free(F);
    return rc;
  }
  hipError_t e = hipMemcpy(F->d_pos, pos, 3 * sizeof(double) * N, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    fm_free(F);
    return hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
  }
  std::vector<int32_t> ones(static_cast<size_t>(N), 1);
  e = hipMemcpy(F->d_types, ones.data(), 4 * static_cast<size_t>(N), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    fm_free(F);
    return hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
  }
  // Per-snapshot full lists (cell build + count), one global scan, fill,
  // then shift each snapshot's indices to global numbering.
  int32_t ncells[3];
  double cs[3];
  if ((rc = tdmd_cell_grid(box, r_cut, ncells, cs))) {
    fm_free(F);
    return rc;
  }
  const int nc = ncells[0] * ncells[1] * ncells[2];
  size_t wb = 0;
  tdmd_cell_workspace_bytes(natoms, nc, &wb);
  std::vector<int32_t*> cc(static_cast<size_t>(nsnap)), co(static_cast<size_t>(nsnap)),
      ca(static_cast<size_t>(nsnap)), ac(static_cast<size_t>(nsnap));
  void* work = nullptr;
  auto cleanup_cells = [&]() {
    for (int s = 0; s < nsnap; ++s) {
      for (int32_t* p : {cc[static_cast<size_t>(s)], co[static_cast<size_t>(s)],
                         ca[static_cast<size_t>(s)], ac[static_cast<size_t>(s)]})
        if (p) (void)hipFree(p);
    }
    if (work) (void)hipFree(work);
  };
  if ((rc = tdmd_malloc(&work, wb))) {
    fm_free(F);
    return rc;
  }
  for (int s = 0; s < nsnap && !rc; ++s) {
    const size_t si = static_cast<size_t>(s);
    const double* ps = F->d_pos + 3 * static_cast<size_t>(s) * natoms;
    if ((rc = fm_alloc(&cc[si], nc)) || (rc = fm_alloc(&co[si], nc)) ||
        (rc = fm_alloc(&ca[si], natoms)) || (rc = fm_alloc(&ac[si], natoms)))
      break;
    if ((rc = tdmd_cell_build(ps, natoms, box, r_cut, cc[si], co[si], ca[si], ac[si], work,
                              nullptr)))
      break;
    rc = tdmd_nlist_count(ps, natoms, box, r_cut, cc[si], co[si], ca[si], ac[si], ncells,
                          F->d_cnt + static_cast<size_t>(s) * natoms, nullptr);
  }
  size_t sb = 0;
  void* swork = nullptr;
  int32_t* meta = nullptr;
  if (!rc) {
    tdmd_scan_workspace_bytes(N, &sb);
    if (!(rc = tdmd_malloc(&swork, sb)) && !(rc = fm_alloc(&meta, 2)))
      rc = tdmd_scan_exclusive_i32(F->d_cnt, F->d_off, N, meta, swork, nullptr);
  }
  int32_t hm[2] = {0, 0};
  if (!rc) {
    e = hipMemcpy(hm, meta, 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
  }
  int32_t* ovf = nullptr;
  if (!rc && !(rc = fm_alloc(&F->d_nbr, std::max(hm[0], 1))) && !(rc = fm_alloc(&ovf, 1))) {
    (void)hipMemset(ovf, 0, 4);
    for (int s = 0; s < nsnap && !rc; ++s) {
      const size_t si = static_cast<size_t>(s);
      const size_t a0 = static_cast<size_t>(s) * natoms;
      rc = tdmd_nlist_fill(F->d_pos + 3 * a0, natoms, box, r_cut, cc[si], co[si], ca[si], ac[si],
                           ncells, F->d_off + a0, F->d_nbr, F->d_cnt + a0, hm[1], ovf, nullptr);
      if (!rc && s > 0) {
        k_fm_offset<<<(natoms + kBlock - 1) / kBlock, kBlock>>>(F->d_nbr, F->d_off + a0,
                                                                F->d_cnt + a0, natoms,
                                                                static_cast<int32_t>(a0));
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) rc = hip_fail(le, "k_fm_offset", __FILE__, __LINE__);
      }
    }
  }
  if (!rc) {
    e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = hip_fail(e, "hipDeviceSynchronize", __FILE__, __LINE__);
  }
  cleanup_cells();
  if (swork) (void)hipFree(swork);
  if (meta) (void)hipFree(meta);
  if (ovf) (void)hipFree(ovf);
  if (rc) {
    fm_free(F);
    return rc;
  }
  *out = F;
  return TDMD_OK;
}

int tdmd_fm_destroy(tdmd_fm* F) {
  if (!F) return TDMD_OK;
  (void)hipDeviceSynchronize();
  return fm_free(F);
}

int tdmd_fm_info(const tdmd_fm* F, int32_t* ntot, int64_t* npairs) {
  if (!F) return fail(TDMD_E_INVALID, "fm_info: null");
  if (ntot) *ntot = F->ntot;
  if (npairs) {
    int32_t last[2];
    TDMD_HIP_TRY(hipMemcpy(&last[0], F->d_off + F->ntot - 1, 4, hipMemcpyDeviceToHost));
    TDMD_HIP_TRY(hipMemcpy(&last[1], F->d_cnt + F->ntot - 1, 4, hipMemcpyDeviceToHost));
    *npairs = static_cast<int64_t>(last[0]) + last[1];
  }
  return TDMD_OK;
}

int tdmd_fm_set_targets(tdmd_fm* F, const double* forces) {
  if (!F || !forces) return fail(TDMD_E_INVALID, "fm_set_targets: null");
  TDMD_HIP_TRY(hipMemcpy(F->d_target, forces, 3 * sizeof(double) * F->ntot,
                         hipMemcpyHostToDevice));
  F->has_target = true;
  return TDMD_OK;
}

int tdmd_fm_morse(tdmd_fm* F, const double params[3], double rc, double* loss, double grad[3],
                  double* forces_out) {
  if (!F || !params || !(rc > 0)) return fail(TDMD_E_INVALID, "fm_morse: bad arguments");
  if (rc > F->r_list) return fail(TDMD_E_INVALID, "fm_morse: cutoff beyond the list cutoff");
  if ((loss || grad) && !F->has_target) return fail(TDMD_E_STATE, "fm_morse: no targets set");
  const int N = F->ntot;
  const DevBox box = make_devbox(F->box);
  const MorseCoef p{params[0], params[1], params[2], rc * rc};
  double* d_fout = nullptr;
  if (forces_out) {
    const int rc2 = fm_alloc(&d_fout, 3 * static_cast<size_t>(N));
    if (rc2) return rc2;
  }
  const int nb = F->nparts;
  if (F->precision == TDMD_MIXED)
    k_fm_morse<float><<<nb, kBlock>>>(F->d_pos, F->d_nbr, F->d_off, F->d_cnt, N, box, p,
                                      F->d_target, F->has_target ? 1 : 0, d_fout, F->d_partials);
  else
    k_fm_morse<double><<<nb, kBlock>>>(F->d_pos, F->d_nbr, F->d_off, F->d_cnt, N, box, p,
                                       F->d_target, F->has_target ? 1 : 0, d_fout,
                                       F->d_partials);
  TDMD_LAUNCH_CHECK();
  k_fm_reduce<<<1, 256>>>(F->d_partials, nb, 4, F->d_scalars);
  TDMD_LAUNCH_CHECK();
  double h[4];
  TDMD_HIP_TRY(hipMemcpy(h, F->d_scalars, sizeof(h), hipMemcpyDeviceToHost));
  if (loss) *loss = h[0];
  if (grad) {
    grad[0] = h[1];
    grad[1] = h[2];
    grad[2] = h[3];
  }
  if (forces_out) {
    TDMD_HIP_TRY(hipMemcpy(forces_out, d_fout, 3 * sizeof(double) * N, hipMemcpyDeviceToHost));
    TDMD_HIP_TRY(hipFree(d_fout));
  }
  return TDMD_OK;
}

int tdmd_fm_eam(tdmd_fm* F, int32_t nrho, double drho, int32_t nr, double dr, double cutoff,
                const double* d_values, double* d_loss, double* d_grad, double* d_forces,
                void* stream) {
  if (!F || !d_values || nrho < 3 || nr < 3 || !(drho > 0) || !(dr > 0) || !(cutoff > 0))
    return fail(TDMD_E_INVALID, "fm_eam: bad arguments");
  if (cutoff > F->r_list) return fail(TDMD_E_INVALID, "fm_eam: cutoff beyond the list cutoff");
  if ((d_loss || d_grad) && !F->has_target) return fail(TDMD_E_STATE, "fm_eam: no targets set");
  hipStream_t s = as_stream(stream);
  const int N = F->ntot;
  const int niF = nrho - 1, niR = nr - 1;
  if (!F->tables || F->nrho != nrho || F->nr != nr || F->drho != drho || F->dr != dr ||
      F->cutoff != cutoff) {
    if (F->tables) eam_tables_destroy(F->tables);
    F->tables = nullptr;
    for (double** p : {&F->d_coef, &F->d_gcoef, &F->d_gS}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
    SetflData sd;
    sd.ntypes = 1;
    sd.nrho = nrho;
    sd.nr = nr;
    sd.drho = drho;
    sd.dr = dr;
    sd.cutoff = cutoff;
    sd.phi_values = true;
    sd.F.assign(static_cast<size_t>(nrho), 0.0);
    sd.rho.assign(static_cast<size_t>(nr), 0.0);
    sd.rphi.assign(static_cast<size_t>(nr), 0.0);
    int rc = eam_tables_create(&F->tables, F->precision, sd);
    if (rc) return rc;
    if ((rc = fm_alloc(&F->d_coef, 4 * static_cast<size_t>(niF) + 8 * static_cast<size_t>(niR))) ||
        (rc = fm_alloc(&F->d_gcoef, 4 * static_cast<size_t>(niF) + 8 * static_cast<size_t>(niR))) ||
        (rc = fm_alloc(&F->d_gS, static_cast<size_t>(niF) + 2 * static_cast<size_t>(niR))))
      return rc;
    F->nrho = nrho;
    F->nr = nr;
    F->drho = drho;
    F->dr = dr;
    F->cutoff = cutoff;
  }
  const bool mixed = F->precision == TDMD_MIXED;
  // FD-Hermite coefficients from the knot values, and the kernels' quads
  double* d_coef = F->d_coef;
  const dim3 gc((std::max(niF, niR) + 255) / 256, 3);
  if (mixed)
    k_fm_build_coeffs<float><<<gc, 256, 0, s>>>(d_values, nrho, drho, nr, dr, d_coef,
                                                static_cast<float4*>(F->tables->d_coeff));
  else
    k_fm_build_coeffs<double><<<gc, 256, 0, s>>>(d_values, nrho, drho, nr, dr, d_coef,
                                                 static_cast<double4*>(F->tables->d_coeff));
  TDMD_LAUNCH_CHECK();
  // forward: density, embedding, forces (the drop-in EAM path)
  int rc = tdmd_eam_compute(F->tables, F->d_pos, F->d_force, F->d_types, F->d_nbr, F->d_off,
                            F->d_cnt, N, &F->box, F->d_rho, F->d_fp, nullptr, nullptr, stream);
  if (rc) return rc;
  const int nb = F->nparts;
  if (mixed)
    k_fm_residual<float><<<nb, kBlock, 0, s>>>(static_cast<const float*>(F->d_force), F->d_target,
                                              N, F->d_g, d_forces, F->d_partials,
                                              F->has_target ? 1 : 0);
  else
    k_fm_residual<double><<<nb, kBlock, 0, s>>>(static_cast<const double*>(F->d_force),
                                               F->d_target, N, F->d_g, d_forces, F->d_partials,
                                               F->has_target ? 1 : 0);
  TDMD_LAUNCH_CHECK();
  if (d_loss) {
    k_fm_reduce<<<1, 256, 0, s>>>(F->d_partials, nb, 1, d_loss);
    TDMD_LAUNCH_CHECK();
  }
  if (d_grad) {
    rc = tdmd_fill_u32(F->d_gcoef, 0u,
                       2 * (4 * static_cast<size_t>(niF) + 8 * static_cast<size_t>(niR)), stream);
    if (rc) return rc;
    FmSplines sp;
    SplineMeta m3[3];
    TDMD_HIP_TRY(hipMemcpy(m3, F->tables->d_meta, sizeof(m3), hipMemcpyDeviceToHost));
    sp.F = m3[0];
    sp.rho = m3[1];
    sp.phi = m3[2];
    sp.niF = niF;
    sp.niR = niR;
    const DevBox box = make_devbox(F->box);
    int dev = 0, ncu = 0, lmax = 0;
    TDMD_HIP_TRY(hipGetDevice(&dev));
    TDMD_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    TDMD_HIP_TRY(hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const int tiles = (N + kFmBlock - 1) / kFmBlock;
    const int grid = std::max(1, std::min(tiles, ncu));
    const size_t lds = static_cast<size_t>(kHist) * niR * sizeof(float);
    const bool use_lds = mixed && lds + 1024 <= static_cast<size_t>(lmax);
    const double rc_sq = cutoff * cutoff;
    if (use_lds) {
      TDMD_HIP_TRY(hipFuncSetAttribute(
          reinterpret_cast<const void*>(&k_fm_eam_adjoint<float, true>),
          hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
      k_fm_eam_adjoint<float, true><<<grid, kFmBlock, lds, s>>>(
          F->d_pos, F->d_nbr, F->d_off, F->d_cnt, N, box, rc_sq,
          static_cast<const float4*>(F->tables->d_coeff), sp, F->d_g, F->d_rho,
          static_cast<const float*>(F->d_fp), F->d_gcoef);
    } else if (mixed) {
      k_fm_eam_adjoint<float, false><<<grid, kFmBlock, 0, s>>>(
          F->d_pos, F->d_nbr, F->d_off, F->d_cnt, N, box, rc_sq,
          static_cast<const float4*>(F->tables->d_coeff), sp, F->d_g, F->d_rho,
          static_cast<const float*>(F->d_fp), F->d_gcoef);
    } else {
      k_fm_eam_adjoint<double, false><<<grid, kFmBlock, 0, s>>>(
          F->d_pos, F->d_nbr, F->d_off, F->d_cnt, N, box, rc_sq,
          static_cast<const double4*>(F->tables->d_coeff), sp, F->d_g, F->d_rho,
          static_cast<const double*>(F->d_fp), F->d_gcoef);
    }
    TDMD_LAUNCH_CHECK();
    const dim3 gs((std::max(niF, niR) + 255) / 256, 3);
    k_fm_coef_grad_s<<<gs, 256, 0, s>>>(F->d_gcoef, nrho, drho, nr, dr, F->d_gS);
    TDMD_LAUNCH_CHECK();
    const dim3 gv((std::max(nrho, nr) + 255) / 256, 3);
    k_fm_coef_grad_v<<<gv, 256, 0, s>>>(F->d_gcoef, F->d_gS, nrho, drho, nr, dr, d_grad);
    TDMD_LAUNCH_CHECK();
  }
  TDMD_HIP_TRY(hipStreamSynchronize(s));
  return TDMD_OK;
}

int tdmd_fm_eam_coef_grad(const tdmd_fm* F, double* gcoef) {
  if (!F || !gcoef || !F->d_gcoef) return fail(TDMD_E_STATE, "fm_eam_coef_grad: no gradient yet");
  const size_t n = 4 * static_cast<size_t>(F->nrho - 1) + 8 * static_cast<size_t>(F->nr - 1);
  TDMD_HIP_TRY(hipMemcpy(gcoef, F->d_gcoef, n * sizeof(double), hipMemcpyDeviceToHost));
  return TDMD_OK;
}

}  // extern "C"
