// tables.cpp -- setfl parsing, FD-Hermite spline construction and the
// device layout of EAM tables (EamAlloy::read_setfl / SplineTable /
// DeviceEam::upload_tables, src/potentials/eam_alloy.cpp:19-192,
// src/potentials/device_eam.cu:282-340).
#include "tables.hpp"

#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace tdmd_hip {

void build_fd_hermite(const double* v, int n, double dr, double* a, double* b, double* c,
                      double* d) {
  // eam_alloy.cpp:19-50. Slopes: mean of the adjacent secants, the one-sided
  // secant at both ends. Coefficients of a + b dx + c dx^2 + d dx^3.
#pragma clang fp contract(off)
  if (n < 2) return;
  for (int i = 0; i < n - 1; ++i) {
    a[i] = v[i];
    const double slope_left = (i > 0) ? (v[i] - v[i - 1]) / dr : 0.0;
    const double slope_right = (v[i + 1] - v[i]) / dr;
    const double slope_next = (i + 2 < n) ? (v[i + 2] - v[i + 1]) / dr : slope_right;
    double m0 = (slope_left + slope_right) * 0.5;
    double m1 = (slope_right + slope_next) * 0.5;
    if (i == 0) m0 = slope_right;
    if (i == n - 2) m1 = slope_right;
    b[i] = m0;
    c[i] = (3.0 * slope_right - 2.0 * m0 - m1) / dr;
    d[i] = (m0 + m1 - 2.0 * slope_right) / (dr * dr);
  }
}

namespace {

bool read_values(std::istream& in, int count, double* out) {
  // eam_alloy.cpp:78-86: whitespace-separated, any number per line.
  for (int i = 0; i < count; ++i)
    if (!(in >> out[i])) return false;
  return true;
}

}  // namespace

int read_setfl(const std::string& path, SetflData* s) {
  std::ifstream file(path);
  if (!file.is_open()) return fail(TDMD_E_IO, "cannot open EAM file: " + path);
  std::string line;
  for (int i = 0; i < 3; ++i)  // eam_alloy.cpp:97-101
    if (!std::getline(file, line)) return fail(TDMD_E_IO, "premature end of EAM file");
  if (!std::getline(file, line)) return fail(TDMD_E_IO, "premature end of EAM file");
  {
    std::istringstream iss(line);  // eam_alloy.cpp:103-112
    if (!(iss >> s->ntypes) || s->ntypes <= 0)
      return fail(TDMD_E_IO, "bad element count line in EAM file");
  }
  if (!std::getline(file, line)) return fail(TDMD_E_IO, "premature end of EAM file");
  {
    std::istringstream iss(line);  // eam_alloy.cpp:114-121
    if (!(iss >> s->nrho >> s->drho >> s->nr >> s->dr >> s->cutoff) || s->nrho < 2 ||
        s->nr < 2)
      return fail(TDMD_E_IO, "bad grid line in EAM file");
  }
  const int nt = s->ntypes, np = nt * (nt + 1) / 2;
  s->F.assign(static_cast<size_t>(nt) * s->nrho, 0.0);
  s->rho.assign(static_cast<size_t>(nt) * s->nr, 0.0);
  s->rphi.assign(static_cast<size_t>(np) * s->nr, 0.0);
  for (int t = 0; t < nt; ++t) {
    // eam_alloy.cpp:129-148: the element header is the next non-blank line.
    bool got = false;
    while (std::getline(file, line)) {
      for (char ch : line)
        if (ch != ' ' && ch != '\t' && ch != '\r') {
          got = true;
          break;
        }
      if (got) break;
    }
    if (!got) return fail(TDMD_E_IO, "premature end of EAM file");
    if (!read_values(file, s->nrho, s->F.data() + static_cast<size_t>(t) * s->nrho) ||
        !read_values(file, s->nr, s->rho.data() + static_cast<size_t>(t) * s->nr))
      return fail(TDMD_E_IO, "unexpected end of EAM file while reading values");
  }
  for (int p = 0; p < np; ++p)  // eam_alloy.cpp:168-191 (i <= j row order)
    if (!read_values(file, s->nr, s->rphi.data() + static_cast<size_t>(p) * s->nr))
      return fail(TDMD_E_IO, "unexpected end of EAM file while reading values");
  return TDMD_OK;
}

template <class R>
static int upload_quads(tdmd_eam_tables* t) {
  using Q = typename Quad<R>::type;
  const size_t nq = t->a.size();
  std::vector<Q> q(nq);
  for (size_t k = 0; k < nq; ++k) {
    q[k].x = static_cast<R>(t->a[k]);
    q[k].y = static_cast<R>(t->b[k]);
    q[k].z = static_cast<R>(t->c[k]);
    q[k].w = static_cast<R>(t->d[k]);
  }
  const int nt = t->ntypes, np = nt * (nt + 1) / 2, ns = 2 * nt + np;
  std::vector<SplineMeta> meta(static_cast<size_t>(ns));
  for (int sidx = 0; sidx < ns; ++sidx) {
    const bool is_f = sidx < nt;
    SplineMeta& m = meta[static_cast<size_t>(sidx)];
    const double dr = is_f ? t->drho : t->dr;
    m.dr = dr;
    m.inv_dr = 1.0 / dr;
    m.rmin = 0.0;
    m.n = is_f ? t->nrho : t->nr;
    m.offset = t->offsets[static_cast<size_t>(sidx)];
  }
  t->coeff_bytes = nq * sizeof(Q);
  TDMD_HIP_TRY(hipMalloc(&t->d_coeff, t->coeff_bytes));
  TDMD_HIP_TRY(hipMalloc(&t->d_meta, meta.size() * sizeof(SplineMeta)));
  TDMD_HIP_TRY(hipMemcpy(t->d_coeff, q.data(), t->coeff_bytes, hipMemcpyHostToDevice));
  TDMD_HIP_TRY(hipMemcpy(t->d_meta, meta.data(), meta.size() * sizeof(SplineMeta),
                         hipMemcpyHostToDevice));
  return TDMD_OK;
}

int eam_tables_create(tdmd_eam_tables** out, int precision, const SetflData& s) {
  if (s.ntypes <= 0 || s.nrho < 2 || s.nr < 2 || !(s.dr > 0) || !(s.drho > 0) ||
      !(s.cutoff > 0))
    return fail(TDMD_E_INVALID, "eam tables: bad grid");
  if (precision != TDMD_FP64 && precision != TDMD_MIXED)
    return fail(TDMD_E_INVALID, "eam tables: bad precision");
  auto* t = new tdmd_eam_tables();
  t->precision = precision;
  t->ntypes = s.ntypes;
  t->nrho = s.nrho;
  t->nr = s.nr;
  t->drho = s.drho;
  t->dr = s.dr;
  t->cutoff = s.cutoff;
  const int nt = s.ntypes, np = nt * (nt + 1) / 2, ns = 2 * nt + np;
  size_t total = 0;
  t->offsets.resize(static_cast<size_t>(ns));
  for (int k = 0; k < ns; ++k) {
    t->offsets[static_cast<size_t>(k)] = static_cast<int>(total);
    total += static_cast<size_t>((k < nt ? s.nrho : s.nr) - 1);
  }
  t->a.resize(total);
  t->b.resize(total);
  t->c.resize(total);
  t->d.resize(total);
  std::vector<double> vals;
  for (int k = 0; k < ns; ++k) {
    const size_t o = static_cast<size_t>(t->offsets[static_cast<size_t>(k)]);
    if (k < nt) {
      build_fd_hermite(s.F.data() + static_cast<size_t>(k) * s.nrho, s.nrho, s.drho, &t->a[o],
                       &t->b[o], &t->c[o], &t->d[o]);
    } else if (k < 2 * nt) {
      build_fd_hermite(s.rho.data() + static_cast<size_t>(k - nt) * s.nr, s.nr, s.dr, &t->a[o],
                       &t->b[o], &t->c[o], &t->d[o]);
    } else {
      // eam_alloy.cpp:168-191: blocks arrive in i<=j row order and land at
      // phi_index(i, j); phi = (r phi) / r with phi(0) = 0.
      const int p_read = k - 2 * nt;
      int ii = 0, jj = 0, cnt = 0;
      for (int i = 0; i < nt; ++i)
        for (int j = i; j < nt; ++j, ++cnt)
          if (cnt == p_read) {
            ii = i;
            jj = j;
          }
      const int pidx = phi_index(ii, jj, nt);
      const size_t o2 = static_cast<size_t>(t->offsets[static_cast<size_t>(2 * nt + pidx)]);
      const double* rp = s.rphi.data() + static_cast<size_t>(p_read) * s.nr;
      if (s.phi_values) {
        vals.assign(rp, rp + s.nr);
      } else {
        vals.assign(static_cast<size_t>(s.nr), 0.0);
        for (int q = 1; q < s.nr; ++q) {
          const double r = static_cast<double>(q) * s.dr;
          vals[static_cast<size_t>(q)] = rp[q] / r;
        }
      }
      build_fd_hermite(vals.data(), s.nr, s.dr, &t->a[o2], &t->b[o2], &t->c[o2], &t->d[o2]);
    }
  }
  const int rc = precision == TDMD_MIXED ? upload_quads<float>(t) : upload_quads<double>(t);
  if (rc) {
    eam_tables_destroy(t);
    return rc;
  }
  *out = t;
  return TDMD_OK;
}

int eam_tables_clone(const tdmd_eam_tables* src, int precision, tdmd_eam_tables** out) {
  if (precision != TDMD_FP64 && precision != TDMD_MIXED)
    return fail(TDMD_E_INVALID, "eam tables: bad precision");
  auto* t = new tdmd_eam_tables();
  t->precision = precision;
  t->ntypes = src->ntypes;
  t->nrho = src->nrho;
  t->nr = src->nr;
  t->drho = src->drho;
  t->dr = src->dr;
  t->cutoff = src->cutoff;
  t->a = src->a;
  t->b = src->b;
  t->c = src->c;
  t->d = src->d;
  t->offsets = src->offsets;
  const int rc = precision == TDMD_MIXED ? upload_quads<float>(t) : upload_quads<double>(t);
  if (rc) {
    eam_tables_destroy(t);
    return rc;
  }
  *out = t;
  return TDMD_OK;
}

void eam_tables_destroy(tdmd_eam_tables* t) {
  if (!t) return;
  if (t->d_coeff) (void)hipFree(t->d_coeff);
  if (t->d_meta) (void)hipFree(t->d_meta);
  delete t;
}

}  // namespace tdmd_hip

using namespace tdmd_hip;

extern "C" {

int tdmd_spline_build(const double* values, int32_t n, double dr, double* a, double* b,
                      double* c, double* d) {
  if (!values || n < 2 || !(dr > 0) || !a || !b || !c || !d)
    return fail(TDMD_E_INVALID, "spline_build: bad arguments");
  build_fd_hermite(values, n, dr, a, b, c, d);
  return TDMD_OK;
}

int tdmd_setfl_read(const char* path, int32_t* ntypes, int32_t* nrho, double* drho, int32_t* nr,
                    double* dr, double* cutoff, double* F, double* rho, double* rphi) {
  if (!path || !ntypes || !nrho || !drho || !nr || !dr || !cutoff)
    return fail(TDMD_E_INVALID, "setfl_read: null pointer");
  SetflData s;
  const int rc = read_setfl(path, &s);
  if (rc) return rc;
  *ntypes = s.ntypes;
  *nrho = s.nrho;
  *drho = s.drho;
  *nr = s.nr;
  *dr = s.dr;
  *cutoff = s.cutoff;
  if (F) std::memcpy(F, s.F.data(), s.F.size() * sizeof(double));
  if (rho) std::memcpy(rho, s.rho.data(), s.rho.size() * sizeof(double));
  if (rphi) std::memcpy(rphi, s.rphi.data(), s.rphi.size() * sizeof(double));
  return TDMD_OK;
}

int tdmd_eam_tables_create(tdmd_eam_tables** out, int32_t precision, int32_t ntypes,
                           int32_t nrho, double drho, int32_t nr, double dr, double cutoff,
                           const double* F, const double* rho, const double* rphi) {
  if (!out || !F || !rho || !rphi || ntypes <= 0 || nrho < 2 || nr < 2)
    return fail(TDMD_E_INVALID, "eam_tables_create: bad arguments");
  SetflData s;
  s.ntypes = ntypes;
  s.nrho = nrho;
  s.nr = nr;
  s.drho = drho;
  s.dr = dr;
  s.cutoff = cutoff;
  const int np = ntypes * (ntypes + 1) / 2;
  s.F.assign(F, F + static_cast<size_t>(ntypes) * nrho);
  s.rho.assign(rho, rho + static_cast<size_t>(ntypes) * nr);
  s.rphi.assign(rphi, rphi + static_cast<size_t>(np) * nr);
  return eam_tables_create(out, precision, s);
}

int tdmd_eam_tables_create_values(tdmd_eam_tables** out, int32_t precision, int32_t ntypes,
                                  int32_t nrho, double drho, int32_t nr, double dr,
                                  double cutoff, const double* F, const double* rho,
                                  const double* phi) {
  if (!out || !F || !rho || !phi || ntypes <= 0 || nrho < 2 || nr < 2)
    return fail(TDMD_E_INVALID, "eam_tables_create_values: bad arguments");
  SetflData s;
  s.ntypes = ntypes;
  s.nrho = nrho;
  s.nr = nr;
  s.drho = drho;
  s.dr = dr;
  s.cutoff = cutoff;
  s.phi_values = true;
  const int np = ntypes * (ntypes + 1) / 2;
  s.F.assign(F, F + static_cast<size_t>(ntypes) * nrho);
  s.rho.assign(rho, rho + static_cast<size_t>(ntypes) * nr);
  s.rphi.assign(phi, phi + static_cast<size_t>(np) * nr);
  return eam_tables_create(out, precision, s);
}

int tdmd_eam_tables_destroy(tdmd_eam_tables* t) {
  eam_tables_destroy(t);
  return TDMD_OK;
}

int tdmd_eam_tables_info(const tdmd_eam_tables* t, int32_t* ntypes, double* cutoff) {
  if (!t) return fail(TDMD_E_INVALID, "eam_tables_info: null tables");
  if (ntypes) *ntypes = t->ntypes;
  if (cutoff) *cutoff = t->cutoff;
  return TDMD_OK;
}

}  // extern "C"
