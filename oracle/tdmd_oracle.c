/*
 * tdmd_oracle.c -- FP64 CPU restatement of TDMD's single-GPU step path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product). See
 * tdmd_oracle.h for the pinning of each function. Build with
 * -ffp-contract=off so that no multiply-add is fused: the reference CPU
 * build (g++ -O3, x86-64 baseline, no FMA) rounds every product and sum.
 */
#include "tdmd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* PBC helpers: src/core/math.hpp:78-131                                    */
/* ------------------------------------------------------------------------ */

void orc_minimum_image(double* d, const double* bs, const int* periodic) {
  /* math.hpp:81-104: one conditional +/- box per axis. */
  for (int a = 0; a < 3; ++a) {
    if (!periodic[a]) continue;
    if (d[a] > 0.5 * bs[a])
      d[a] -= bs[a];
    else if (d[a] < -0.5 * bs[a])
      d[a] += bs[a];
  }
}

void orc_wrap_position(double* p, const double* lo, const double* bs,
                       const int* periodic) {
  /* math.hpp:107-129: floor-based wrap into [lo, hi). */
  for (int a = 0; a < 3; ++a) {
    if (!periodic[a]) continue;
    double rel = p[a] - lo[a];
    rel -= floor(rel / bs[a]) * bs[a];
    p[a] = lo[a] + rel;
  }
}

static void box_size(const orc_box* b, double* bs) {
  /* box.hpp:23-25 */
  bs[0] = b->hi[0] - b->lo[0];
  bs[1] = b->hi[1] - b->lo[1];
  bs[2] = b->hi[2] - b->lo[2];
}

static double length_sq(const double* d) {
  /* math.hpp:56-63: dot(v, v) = x*x + y*y + z*z, left to right. */
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

/* ------------------------------------------------------------------------ */
/* Cell list: src/domain/cell_list.cpp:13-58, cell_list.hpp:47-72           */
/* ------------------------------------------------------------------------ */

static int imax(int a, int b) { return a > b ? a : b; }

int orc_cell_of(const orc_cells* cl, const double* p, const double* lo) {
  /* cell_list.hpp:60-72: truncating cast of (p - lo) / cell_size, clamp. */
  int ix = (int)((p[0] - lo[0]) / cl->cs[0]);
  int iy = (int)((p[1] - lo[1]) / cl->cs[1]);
  int iz = (int)((p[2] - lo[2]) / cl->cs[2]);
  if (ix >= cl->ncx) ix = cl->ncx - 1;
  if (iy >= cl->ncy) iy = cl->ncy - 1;
  if (iz >= cl->ncz) iz = cl->ncz - 1;
  if (ix < 0) ix = 0;
  if (iy < 0) iy = 0;
  if (iz < 0) iz = 0;
  return iz * cl->ncx * cl->ncy + iy * cl->ncx + ix;
}

static int cell_index(const orc_cells* cl, int ix, int iy, int iz) {
  /* cell_list.hpp:47-56: single wrap per axis. */
  if (ix < 0) ix += cl->ncx;
  else if (ix >= cl->ncx) ix -= cl->ncx;
  if (iy < 0) iy += cl->ncy;
  else if (iy >= cl->ncy) iy -= cl->ncy;
  if (iz < 0) iz += cl->ncz;
  else if (iz >= cl->ncz) iz -= cl->ncz;
  return iz * cl->ncx * cl->ncy + iy * cl->ncx + ix;
}

void orc_cells_free(orc_cells* cl) {
  free(cl->counts);
  free(cl->offsets);
  free(cl->atoms);
  memset(cl, 0, sizeof(*cl));
}

int orc_cells_build(orc_cells* cl, const double* pos, int64_t natoms,
                    const orc_box* box, double r_list) {
  if (!(r_list > 0.0) || natoms < 0) return -1;
  double bs[3];
  box_size(box, bs);
  /* cell_list.cpp:22-27: at least 3 cells per axis. */
  cl->ncx = imax(3, (int)floor(bs[0] / r_list));
  cl->ncy = imax(3, (int)floor(bs[1] / r_list));
  cl->ncz = imax(3, (int)floor(bs[2] / r_list));
  cl->cs[0] = bs[0] / cl->ncx;
  cl->cs[1] = bs[1] / cl->ncy;
  cl->cs[2] = bs[2] / cl->ncz;
  cl->ncells = cl->ncx * cl->ncy * cl->ncz;
  free(cl->counts);
  free(cl->offsets);
  free(cl->atoms);
  cl->counts = (int*)calloc((size_t)cl->ncells, sizeof(int));
  cl->offsets = (int*)malloc((size_t)cl->ncells * sizeof(int));
  cl->atoms = (int*)malloc((size_t)(natoms > 0 ? natoms : 1) * sizeof(int));
  int* placed = (int*)calloc((size_t)cl->ncells, sizeof(int));
  if (!cl->counts || !cl->offsets || !cl->atoms || !placed) {
    free(placed);
    return -2;
  }
  /* cell_list.cpp:31-45: counting sort. */
  for (int64_t i = 0; i < natoms; ++i) ++cl->counts[orc_cell_of(cl, pos + 3 * i, box->lo)];
  cl->offsets[0] = 0;
  for (int c = 1; c < cl->ncells; ++c)
    cl->offsets[c] = cl->offsets[c - 1] + cl->counts[c - 1];
  /* cell_list.cpp:48-57: placement in index order (stable). */
  for (int64_t i = 0; i < natoms; ++i) {
    int c = orc_cell_of(cl, pos + 3 * i, box->lo);
    cl->atoms[cl->offsets[c] + placed[c]] = (int)i;
    ++placed[c];
  }
  free(placed);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Half neighbour list: src/neighbors/neighbor_list.cpp:14-100              */
/* ------------------------------------------------------------------------ */

void orc_nlist_free(orc_nlist* nl) {
  free(nl->counts);
  free(nl->offsets);
  free(nl->nbrs);
  free(nl->build_pos);
  memset(nl, 0, sizeof(*nl));
}

int orc_nlist_build(orc_nlist* nl, const double* pos, int64_t natoms,
                    const orc_box* box, double r_cut, double r_skin) {
  /* neighbor_list.cpp:16-19 */
  const double r_list = r_cut + r_skin;
  const double r_list_sq = r_list * r_list;
  double bs[3];
  box_size(box, bs);
  orc_cells cl;
  memset(&cl, 0, sizeof(cl));
  if (orc_cells_build(&cl, pos, natoms, box, r_list) != 0) return -1;

  free(nl->counts);
  free(nl->offsets);
  free(nl->nbrs);
  free(nl->build_pos);
  nl->natoms = natoms;
  nl->r_cut = r_cut;
  nl->r_skin = r_skin;
  size_t n = (size_t)(natoms > 0 ? natoms : 1);
  nl->counts = (int*)calloc(n, sizeof(int));
  nl->offsets = (int*)calloc(n, sizeof(int));
  nl->build_pos = (double*)malloc(3 * n * sizeof(double));
  size_t cap = 64 * n, used = 0;
  nl->nbrs = (int*)malloc(cap * sizeof(int));
  if (!nl->counts || !nl->offsets || !nl->build_pos || !nl->nbrs) {
    orc_cells_free(&cl);
    return -2;
  }
  const int ncx = cl.ncx, ncy = cl.ncy;
  /* neighbor_list.cpp:35-69: rows are filled in i order; within a row the
   * 27-cell sweep dz, dy, dx (outer to inner) then the cell's atoms. */
  for (int64_t i = 0; i < natoms; ++i) {
    const double* pi = pos + 3 * i;
    const int ci = orc_cell_of(&cl, pi, box->lo);
    const int iz = ci / (ncx * ncy);
    const int iy = (ci - iz * ncx * ncy) / ncx;
    const int ix = ci - iz * ncx * ncy - iy * ncx;
    nl->offsets[i] = (int)used;
    int cnt = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int nci = cell_index(&cl, ix + dx, iy + dy, iz + dz);
          const int c = cl.counts[nci];
          const int* atoms = cl.atoms + cl.offsets[nci];
          for (int k = 0; k < c; ++k) {
            const int j = atoms[k];
            if (j <= (int)i) continue; /* half list: i < j */
            double d[3] = {pos[3 * j] - pi[0], pos[3 * j + 1] - pi[1],
                           pos[3 * j + 2] - pi[2]};
            orc_minimum_image(d, bs, box->periodic);
            const double r2 = length_sq(d);
            if (r2 < r_list_sq) {
              if (used == cap) {
                cap *= 2;
                int* nb = (int*)realloc(nl->nbrs, cap * sizeof(int));
                if (!nb) {
                  orc_cells_free(&cl);
                  return -2;
                }
                nl->nbrs = nb;
              }
              nl->nbrs[used++] = j;
              ++cnt;
            }
          }
        }
    nl->counts[i] = cnt;
  }
  nl->total = (int64_t)used;
  /* neighbor_list.cpp:86: snapshot for the skin check. */
  memcpy(nl->build_pos, pos, 3 * (size_t)natoms * sizeof(double));
  orc_cells_free(&cl);
  return 0;
}

int orc_nlist_needs_rebuild(const orc_nlist* nl, const double* pos,
                            int64_t natoms) {
  /* neighbor_list.cpp:89-99: plain difference (no minimum image), so an
   * atom that wraps through a periodic face also triggers a rebuild. */
  const double half_skin_sq = 0.25 * nl->r_skin * nl->r_skin;
  for (int64_t i = 0; i < natoms; ++i) {
    double d[3] = {pos[3 * i] - nl->build_pos[3 * i],
                   pos[3 * i + 1] - nl->build_pos[3 * i + 1],
                   pos[3 * i + 2] - nl->build_pos[3 * i + 2]};
    if (length_sq(d) > half_skin_sq) return 1;
  }
  return 0;
}

static int cmp_int(const void* a, const void* b) {
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

int orc_nlist_full(const orc_nlist* h, int* counts, int* offsets, int* nbrs) {
  const int64_t n = h->natoms;
  for (int64_t i = 0; i < n; ++i) counts[i] = h->counts[i];
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < h->counts[i]; ++k) ++counts[h->nbrs[h->offsets[i] + k]];
  int64_t acc = 0;
  for (int64_t i = 0; i < n; ++i) {
    offsets[i] = (int)acc;
    acc += counts[i];
  }
  int* fill = (int*)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
  if (!fill) return -2;
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < h->counts[i]; ++k) {
      const int j = h->nbrs[h->offsets[i] + k];
      nbrs[offsets[i] + fill[i]++] = j;
      nbrs[offsets[j] + fill[j]++] = (int)i;
    }
  for (int64_t i = 0; i < n; ++i) qsort(nbrs + offsets[i], (size_t)counts[i], sizeof(int), cmp_int);
  free(fill);
  return 0;
}

int64_t orc_bruteforce_full(const double* pos, int64_t natoms,
                            const orc_box* box, double r_list, int* counts,
                            int* offsets, int* nbrs, int64_t capacity) {
  const double r_list_sq = r_list * r_list;
  double bs[3];
  box_size(box, bs);
  int64_t used = 0;
  for (int64_t i = 0; i < natoms; ++i) {
    offsets[i] = (int)used;
    int cnt = 0;
    for (int64_t j = 0; j < natoms; ++j) {
      if (j == i) continue;
      double d[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1],
                     pos[3 * j + 2] - pos[3 * i + 2]};
      orc_minimum_image(d, bs, box->periodic);
      if (length_sq(d) < r_list_sq) {
        if (used < capacity) nbrs[used] = (int)j;
        ++used;
        ++cnt;
      }
    }
    counts[i] = cnt;
  }
  return used;
}

/* ------------------------------------------------------------------------ */
/* Morse: src/potentials/morse.hpp:37-55, force_compute.cpp:10-56          */
/* ------------------------------------------------------------------------ */

void orc_morse_pair(double D, double alpha, double r0, double rc_sq, double r2,
                    double* energy, double* fpair) {
  /* morse.hpp:37-55 */
  if (r2 >= rc_sq) {
    *energy = 0.0;
    *fpair = 0.0;
    return;
  }
  const double r = sqrt(r2);
  const double dr = r - r0;
  const double exp_val = exp(-alpha * dr);
  const double one_minus_exp = 1.0 - exp_val;
  *energy = D * (one_minus_exp * one_minus_exp - 1.0);
  const double dudr = 2.0 * D * alpha * one_minus_exp * exp_val;
  *fpair = -dudr / r;
}

double orc_morse_forces(const double* pos, double* f, int64_t natoms,
                        const orc_box* box, const orc_nlist* nl, double D,
                        double alpha, double r0, double rc) {
  /* morse.hpp:28-29: rc_sq = rc * rc. force_compute.cpp:10-56: zero, then
   * i-major sweep of the half list with Newton-3 scatter. */
  const double rc_sq = rc * rc;
  double bs[3];
  box_size(box, bs);
  memset(f, 0, 3 * (size_t)natoms * sizeof(double));
  double total = 0.0;
  for (int64_t i = 0; i < natoms; ++i) {
    const double* pi = pos + 3 * i;
    const int cnt = nl->counts[i];
    const int* nb = nl->nbrs + nl->offsets[i];
    for (int k = 0; k < cnt; ++k) {
      const int j = nb[k];
      double d[3] = {pi[0] - pos[3 * j], pi[1] - pos[3 * j + 1],
                     pi[2] - pos[3 * j + 2]};
      orc_minimum_image(d, bs, box->periodic);
      const double r2 = length_sq(d);
      double e, fp;
      orc_morse_pair(D, alpha, r0, rc_sq, r2, &e, &fp);
      total += e;
      const double fx = fp * d[0], fy = fp * d[1], fz = fp * d[2];
      f[3 * i] += fx;
      f[3 * i + 1] += fy;
      f[3 * i + 2] += fz;
      f[3 * j] -= fx;
      f[3 * j + 1] -= fy;
      f[3 * j + 2] -= fz;
    }
  }
  return total;
}

void orc_morse_forces_rows(const double* pos, double* f, int first, int count,
                           const int* counts, const int* offsets, const int* nbrs,
                           const orc_box* box, double D, double alpha, double r0,
                           double rc) {
  /* morse_force_zone_kernel (device_morse_zone.cu:29-112) in FP64: each row
   * i of [first, first + count) of a full list, d = x_i - x_j with one
   * minimum image, F_i += fpair d summed in row order, then ADDED to f. */
  const double rc_sq = rc * rc;
  double bs[3];
  box_size(box, bs);
  for (int i = first; i < first + count; ++i) {
    const double* pi = pos + 3 * (int64_t)i;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int* nb = nbrs + offsets[i];
    for (int k = 0; k < counts[i]; ++k) {
      const int j = nb[k];
      double d[3] = {pi[0] - pos[3 * (int64_t)j], pi[1] - pos[3 * (int64_t)j + 1],
                     pi[2] - pos[3 * (int64_t)j + 2]};
      orc_minimum_image(d, bs, box->periodic);
      const double r2 = length_sq(d);
      double e, fp;
      orc_morse_pair(D, alpha, r0, rc_sq, r2, &e, &fp);
      fx += fp * d[0];
      fy += fp * d[1];
      fz += fp * d[2];
    }
    f[3 * (int64_t)i] += fx;
    f[3 * (int64_t)i + 1] += fy;
    f[3 * (int64_t)i + 2] += fz;
  }
}

/* ------------------------------------------------------------------------ */
/* FD-Hermite spline: src/potentials/eam_alloy.cpp:19-72                    */
/* ------------------------------------------------------------------------ */

void orc_spline_free(orc_spline* s) {
  free(s->values);
  free(s->a);
  free(s->b);
  free(s->c);
  free(s->d);
  memset(s, 0, sizeof(*s));
}

void orc_spline_build(orc_spline* s) {
  /* eam_alloy.cpp:19-50: slopes are the mean of adjacent secants, one-sided
   * at both ends; coefficients of a + b dx + c dx^2 + d dx^3. */
  if (s->n < 2) return;
  const int ns = s->n;
  const double dr = s->dr;
  const double* v = s->values;
  free(s->a);
  free(s->b);
  free(s->c);
  free(s->d);
  s->a = (double*)malloc((size_t)(ns - 1) * sizeof(double));
  s->b = (double*)malloc((size_t)(ns - 1) * sizeof(double));
  s->c = (double*)malloc((size_t)(ns - 1) * sizeof(double));
  s->d = (double*)malloc((size_t)(ns - 1) * sizeof(double));
  for (int i = 0; i < ns - 1; ++i) {
    s->a[i] = v[i];
    const double slope_left = (i > 0) ? (v[i] - v[i - 1]) / dr : 0.0;
    const double slope_right = (v[i + 1] - v[i]) / dr;
    const double slope_next = (i + 2 < ns) ? (v[i + 2] - v[i + 1]) / dr : slope_right;
    double m0 = (slope_left + slope_right) * 0.5;
    double m1 = (slope_right + slope_next) * 0.5;
    if (i == 0) m0 = slope_right;
    if (i == ns - 2) m1 = slope_right;
    s->b[i] = m0;
    s->c[i] = (3.0 * slope_right - 2.0 * m0 - m1) / dr;
    s->d[i] = (m0 + m1 - 2.0 * slope_right) / (dr * dr);
  }
}

double orc_spline_eval(const orc_spline* s, double r) {
  /* eam_alloy.cpp:52-61 */
  if (s->n < 2) return 0.0;
  const double x = (r - s->rmin) / s->dr;
  int idx = (int)x;
  if (idx < 0) idx = 0;
  if (idx >= s->n - 1) idx = s->n - 2;
  const double dx = r - (s->rmin + (double)idx * s->dr);
  return s->a[idx] + dx * (s->b[idx] + dx * (s->c[idx] + dx * s->d[idx]));
}

double orc_spline_eval_deriv(const orc_spline* s, double r) {
  /* eam_alloy.cpp:63-72 */
  if (s->n < 2) return 0.0;
  const double x = (r - s->rmin) / s->dr;
  int idx = (int)x;
  if (idx < 0) idx = 0;
  if (idx >= s->n - 1) idx = s->n - 2;
  const double dx = r - (s->rmin + (double)idx * s->dr);
  return s->b[idx] + dx * (2.0 * s->c[idx] + dx * 3.0 * s->d[idx]);
}

/* ------------------------------------------------------------------------ */
/* EAM/alloy: src/potentials/eam_alloy.cpp:90-295                           */
/* ------------------------------------------------------------------------ */

int orc_eam_phi_index(int ntypes, int ti, int tj) {
  /* eam_alloy.hpp:88-92: flat upper triangle, i <= j. */
  if (ti > tj) {
    int t = ti;
    ti = tj;
    tj = t;
  }
  return ti * ntypes - ti * (ti + 1) / 2 + tj;
}

void orc_eam_free(orc_eam* e) {
  const int np = e->ntypes * (e->ntypes + 1) / 2;
  for (int t = 0; t < e->ntypes; ++t) {
    if (e->embedding) orc_spline_free(&e->embedding[t]);
    if (e->density) orc_spline_free(&e->density[t]);
  }
  for (int p = 0; p < np; ++p)
    if (e->phi) orc_spline_free(&e->phi[p]);
  free(e->embedding);
  free(e->density);
  free(e->phi);
  memset(e, 0, sizeof(*e));
}

static int eam_alloc(orc_eam* e, int ntypes) {
  e->ntypes = ntypes;
  const int np = ntypes * (ntypes + 1) / 2;
  e->embedding = (orc_spline*)calloc((size_t)ntypes, sizeof(orc_spline));
  e->density = (orc_spline*)calloc((size_t)ntypes, sizeof(orc_spline));
  e->phi = (orc_spline*)calloc((size_t)np, sizeof(orc_spline));
  return (e->embedding && e->density && e->phi) ? 0 : -2;
}

static void set_values(orc_spline* s, int n, double dr, const double* vals) {
  s->n = n;
  s->dr = dr;
  s->rmin = 0.0;
  s->values = (double*)malloc((size_t)n * sizeof(double));
  memcpy(s->values, vals, (size_t)n * sizeof(double));
}

static void set_phi_from_rphi(orc_spline* s, int nr, double dr,
                              const double* rphi) {
  /* eam_alloy.cpp:179-189: phi = (r phi)/r with r = k*dr, phi(0) = 0. */
  s->n = nr;
  s->dr = dr;
  s->rmin = 0.0;
  s->values = (double*)malloc((size_t)nr * sizeof(double));
  s->values[0] = 0.0;
  for (int k = 1; k < nr; ++k) {
    const double r = (double)k * dr;
    s->values[k] = rphi[k] / r;
  }
}

int orc_eam_from_arrays(orc_eam* e, int ntypes, int nrho, double drho, int nr,
                        double dr, double cutoff, const double* F,
                        const double* rho, const double* rphi) {
  memset(e, 0, sizeof(*e));
  if (ntypes <= 0 || nrho < 2 || nr < 2) return -1;
  if (eam_alloc(e, ntypes) != 0) return -2;
  e->cutoff = cutoff;
  e->nrho = nrho;
  e->nr = nr;
  e->drho = drho;
  e->dr = dr;
  for (int t = 0; t < ntypes; ++t) {
    set_values(&e->embedding[t], nrho, drho, F + (size_t)t * nrho);
    orc_spline_build(&e->embedding[t]);
    set_values(&e->density[t], nr, dr, rho + (size_t)t * nr);
    orc_spline_build(&e->density[t]);
  }
  int p = 0;
  for (int i = 0; i < ntypes; ++i)
    for (int j = i; j < ntypes; ++j, ++p) {
      const int pidx = orc_eam_phi_index(ntypes, i, j);
      set_phi_from_rphi(&e->phi[pidx], nr, dr, rphi + (size_t)p * nr);
      orc_spline_build(&e->phi[pidx]);
    }
  return 0;
}

/* getline into a growable buffer; returns length or -1 at EOF. */
static long read_line(FILE* f, char** buf, size_t* cap) {
  size_t len = 0;
  int ch;
  if (!*buf) {
    *cap = 256;
    *buf = (char*)malloc(*cap);
  }
  while ((ch = fgetc(f)) != EOF) {
    if (len + 1 >= *cap) {
      *cap *= 2;
      *buf = (char*)realloc(*buf, *cap);
    }
    if (ch == '\n') break;
    (*buf)[len++] = (char)ch;
  }
  if (ch == EOF && len == 0) return -1;
  (*buf)[len] = '\0';
  return (long)len;
}

static int read_values(FILE* f, int count, double* out) {
  /* eam_alloy.cpp:78-86: whitespace-separated tokens, any count per line. */
  for (int i = 0; i < count; ++i)
    if (fscanf(f, "%lf", &out[i]) != 1) return -1;
  return 0;
}

int orc_eam_read_setfl(orc_eam* e, const char* path, char* err, int errlen) {
  memset(e, 0, sizeof(*e));
  FILE* f = fopen(path, "r");
  if (!f) {
    snprintf(err, (size_t)errlen, "cannot open EAM file: %s", path);
    return -1;
  }
  char* line = NULL;
  size_t cap = 0;
  int rc = -1;
  double *Fv = NULL, *rhov = NULL, *rphiv = NULL;
  /* eam_alloy.cpp:97-101: three comment lines. */
  for (int i = 0; i < 3; ++i)
    if (read_line(f, &line, &cap) < 0) {
      snprintf(err, (size_t)errlen, "premature end of EAM file");
      goto done;
    }
  /* eam_alloy.cpp:103-112: ntypes and element names. */
  if (read_line(f, &line, &cap) < 0) {
    snprintf(err, (size_t)errlen, "premature end of EAM file");
    goto done;
  }
  int ntypes = 0;
  if (sscanf(line, "%d", &ntypes) != 1 || ntypes <= 0) {
    snprintf(err, (size_t)errlen, "bad element line");
    goto done;
  }
  /* eam_alloy.cpp:114-121: Nrho drho Nr dr cutoff. */
  if (read_line(f, &line, &cap) < 0) {
    snprintf(err, (size_t)errlen, "premature end of EAM file");
    goto done;
  }
  int nrho = 0, nr = 0;
  double drho = 0, dr = 0, cutoff = 0;
  if (sscanf(line, "%d %lf %d %lf %lf", &nrho, &drho, &nr, &dr, &cutoff) != 5) {
    snprintf(err, (size_t)errlen, "bad grid line");
    goto done;
  }
  const int np = ntypes * (ntypes + 1) / 2;
  Fv = (double*)malloc((size_t)ntypes * nrho * sizeof(double));
  rhov = (double*)malloc((size_t)ntypes * nr * sizeof(double));
  rphiv = (double*)malloc((size_t)np * nr * sizeof(double));
  for (int t = 0; t < ntypes; ++t) {
    /* eam_alloy.cpp:129-148: skip to the next non-blank line (the element
     * header), whose content is discarded. */
    int got = 0;
    long len;
    while ((len = read_line(f, &line, &cap)) >= 0) {
      for (long k = 0; k < len; ++k)
        if (line[k] != ' ' && line[k] != '\t' && line[k] != '\r') {
          got = 1;
          break;
        }
      if (got) break;
    }
    if (!got) {
      snprintf(err, (size_t)errlen, "premature end of EAM file");
      goto done;
    }
    if (read_values(f, nrho, Fv + (size_t)t * nrho) != 0 ||
        read_values(f, nr, rhov + (size_t)t * nr) != 0) {
      snprintf(err, (size_t)errlen, "unexpected end of EAM file while reading values");
      goto done;
    }
  }
  /* eam_alloy.cpp:168-191: r*phi blocks for i <= j in row order. */
  for (int p = 0; p < np; ++p)
    if (read_values(f, nr, rphiv + (size_t)p * nr) != 0) {
      snprintf(err, (size_t)errlen, "unexpected end of EAM file while reading values");
      goto done;
    }
  rc = orc_eam_from_arrays(e, ntypes, nrho, drho, nr, dr, cutoff, Fv, rhov, rphiv);
  if (rc != 0) snprintf(err, (size_t)errlen, "bad EAM tables");
done:
  free(line);
  free(Fv);
  free(rhov);
  free(rphiv);
  fclose(f);
  return rc;
}

double orc_eam_forces(const orc_eam* e, const double* pos, const int* types,
                      double* f, int64_t natoms, const orc_box* box,
                      const orc_nlist* nl, double* rho_out, double* fp_out,
                      double* embed_out) {
  /* eam_alloy.cpp:196-295 */
  const double rc_sq = e->cutoff * e->cutoff;
  double bs[3];
  box_size(box, bs);
  memset(f, 0, 3 * (size_t)natoms * sizeof(double));
  const size_t n = (size_t)(natoms > 0 ? natoms : 1);
  double* rho = (double*)calloc(n, sizeof(double));
  double* fp = (double*)calloc(n, sizeof(double));
  /* Pass 1 (eam_alloy.cpp:213-236): density gather, both ends of a pair. */
  for (int64_t i = 0; i < natoms; ++i) {
    const double* pi = pos + 3 * i;
    const int ti = types[i] - 1;
    const int* nb = nl->nbrs + nl->offsets[i];
    for (int k = 0; k < nl->counts[i]; ++k) {
      const int j = nb[k];
      double d[3] = {pi[0] - pos[3 * j], pi[1] - pos[3 * j + 1], pi[2] - pos[3 * j + 2]};
      orc_minimum_image(d, bs, box->periodic);
      const double r2 = length_sq(d);
      if (r2 >= rc_sq) continue;
      const double r = sqrt(r2);
      const int tj = types[j] - 1;
      rho[i] += orc_spline_eval(&e->density[tj], r);
      rho[j] += orc_spline_eval(&e->density[ti], r);
    }
  }
  /* Pass 2 (eam_alloy.cpp:238-246): F(rho) and F'(rho). */
  double energy_embed = 0.0;
  for (int64_t i = 0; i < natoms; ++i) {
    const int ti = types[i] - 1;
    energy_embed += orc_spline_eval(&e->embedding[ti], rho[i]);
    fp[i] = orc_spline_eval_deriv(&e->embedding[ti], rho[i]);
  }
  /* Pass 3 (eam_alloy.cpp:248-292): pair + embedding forces. */
  double energy_pair = 0.0;
  for (int64_t i = 0; i < natoms; ++i) {
    const double* pi = pos + 3 * i;
    const int ti = types[i] - 1;
    const int* nb = nl->nbrs + nl->offsets[i];
    for (int k = 0; k < nl->counts[i]; ++k) {
      const int j = nb[k];
      double d[3] = {pi[0] - pos[3 * j], pi[1] - pos[3 * j + 1], pi[2] - pos[3 * j + 2]};
      orc_minimum_image(d, bs, box->periodic);
      const double r2 = length_sq(d);
      if (r2 >= rc_sq) continue;
      const double r = sqrt(r2);
      const int tj = types[j] - 1;
      const int pidx = orc_eam_phi_index(e->ntypes, ti, tj);
      const double phi_val = orc_spline_eval(&e->phi[pidx], r);
      energy_pair += phi_val;
      const double dphi_dr = orc_spline_eval_deriv(&e->phi[pidx], r);
      const double drho_j_dr = orc_spline_eval_deriv(&e->density[tj], r);
      const double drho_i_dr = orc_spline_eval_deriv(&e->density[ti], r);
      const double fpair = -(dphi_dr + fp[i] * drho_j_dr + fp[j] * drho_i_dr) / r;
      const double fx = fpair * d[0], fy = fpair * d[1], fz = fpair * d[2];
      f[3 * i] += fx;
      f[3 * i + 1] += fy;
      f[3 * i + 2] += fz;
      f[3 * j] -= fx;
      f[3 * j + 1] -= fy;
      f[3 * j + 2] -= fz;
    }
  }
  if (rho_out) memcpy(rho_out, rho, (size_t)natoms * sizeof(double));
  if (fp_out) memcpy(fp_out, fp, (size_t)natoms * sizeof(double));
  if (embed_out) *embed_out = energy_embed;
  free(rho);
  free(fp);
  return energy_embed + energy_pair;
}

/* ------------------------------------------------------------------------ */
/* Force matching (no reference counterpart)                                */
/* ------------------------------------------------------------------------ */

double orc_fm_morse(const double* pos, int64_t natoms, const orc_box* box,
                    const orc_nlist* nl, double D, double alpha, double r0,
                    double rc, const double* target, double grad[3],
                    double* forces_out) {
  /* F_i = sum_j fpair(r_ij) d_ij, d_ij = x_i - x_j; for a half pair the
   * force enters i with +d and j with -d, so dL = 2 (g_i - g_j).d dfpair,
   * g = F - T. fpair = -2 D alpha (1-e) e / r, e = exp(-alpha (r - r0)). */
  const size_t n3 = 3 * (size_t)(natoms > 0 ? natoms : 1);
  double* f = (double*)malloc(n3 * sizeof(double));
  orc_morse_forces(pos, f, natoms, box, nl, D, alpha, r0, rc);
  if (forces_out) memcpy(forces_out, f, 3 * (size_t)natoms * sizeof(double));
  double L = 0.0;
  for (size_t k = 0; k < 3 * (size_t)natoms; ++k) {
    f[k] -= target[k];
    L += f[k] * f[k];
  }
  const double rc_sq = rc * rc;
  double bs[3];
  box_size(box, bs);
  grad[0] = grad[1] = grad[2] = 0.0;
  for (int64_t i = 0; i < natoms; ++i) {
    const double* pi = pos + 3 * i;
    const int* nb = nl->nbrs + nl->offsets[i];
    for (int k = 0; k < nl->counts[i]; ++k) {
      const int j = nb[k];
      double d[3] = {pi[0] - pos[3 * j], pi[1] - pos[3 * j + 1], pi[2] - pos[3 * j + 2]};
      orc_minimum_image(d, bs, box->periodic);
      const double r2 = length_sq(d);
      if (r2 >= rc_sq) continue;
      const double r = sqrt(r2);
      const double e = exp(-alpha * (r - r0));
      const double w = 2.0 * ((f[3 * i] - f[3 * j]) * d[0] + (f[3 * i + 1] - f[3 * j + 1]) * d[1] +
                              (f[3 * i + 2] - f[3 * j + 2]) * d[2]);
      grad[0] += w * (-2.0 * alpha * (1.0 - e) * e / r);
      grad[1] += w * (-2.0 * D * e * ((1.0 - e) - alpha * (1.0 - 2.0 * e) * (r - r0)) / r);
      grad[2] += w * (-2.0 * D * alpha * alpha * (1.0 - 2.0 * e) * e / r);
    }
  }
  free(f);
  return L;
}

int orc_eam_from_values(orc_eam* e, int nrho, double drho, int nr, double dr,
                        double cutoff, const double* F, const double* rho,
                        const double* phi) {
  memset(e, 0, sizeof(*e));
  if (nrho < 2 || nr < 2) return -1;
  if (eam_alloc(e, 1) != 0) return -2;
  e->cutoff = cutoff;
  e->nrho = nrho;
  e->nr = nr;
  e->drho = drho;
  e->dr = dr;
  set_values(&e->embedding[0], nrho, drho, F);
  orc_spline_build(&e->embedding[0]);
  set_values(&e->density[0], nr, dr, rho);
  orc_spline_build(&e->density[0]);
  set_values(&e->phi[0], nr, dr, phi);
  orc_spline_build(&e->phi[0]);
  return 0;
}

/* Interval and offset of x, as orc_spline_eval. */
static int spline_locate(const orc_spline* s, double x, double* t) {
  int idx = (int)((x - s->rmin) / s->dr);
  if (idx < 0) idx = 0;
  if (idx >= s->n - 1) idx = s->n - 2;
  *t = x - (s->rmin + (double)idx * s->dr);
  return idx;
}

void orc_spline_coef_grad_to_values(int n, double dr, const double* ga,
                                    const double* gb, const double* gc,
                                    const double* gd, double* gv) {
  /* Reverse of orc_spline_build. With S_k = (v_{k+1} - v_k)/dr:
   *   a_k = v_k, b_k = m0_k, c_k = (3 S_k - 2 m0_k - m1_k)/dr,
   *   d_k = (m0_k + m1_k - 2 S_k)/dr^2,
   *   m0_0 = S_0, m0_k = (S_{k-1} + S_k)/2; m1_{n-2} = S_{n-2},
   *   m1_k = (S_k + S_{k+1})/2. */
  const int ni = n - 1;
  double* gS = (double*)calloc((size_t)(ni > 0 ? ni : 1), sizeof(double));
  for (int k = 0; k < n; ++k) gv[k] = 0.0;
  for (int k = 0; k < ni; ++k) {
    const double gm0 = gb[k] - 2.0 * gc[k] / dr + gd[k] / (dr * dr);
    const double gm1 = -gc[k] / dr + gd[k] / (dr * dr);
    gS[k] += 3.0 * gc[k] / dr - 2.0 * gd[k] / (dr * dr);
    if (k == 0) {
      gS[0] += gm0;
    } else {
      gS[k - 1] += 0.5 * gm0;
      gS[k] += 0.5 * gm0;
    }
    if (k == ni - 1) {
      gS[k] += gm1;
    } else {
      gS[k] += 0.5 * gm1;
      gS[k + 1] += 0.5 * gm1;
    }
    gv[k] += ga[k];
  }
  for (int k = 0; k < ni; ++k) {
    gv[k + 1] += gS[k] / dr;
    gv[k] -= gS[k] / dr;
  }
  free(gS);
}

double orc_fm_eam(const orc_eam* e, const double* pos, int64_t natoms,
                  const orc_box* box, const orc_nlist* nl, const double* target,
                  double* gF, double* grho, double* gphi, double* gcoef,
                  double* forces_out) {
  /* Single element. With g = F - T, u = d/r, w = -2 (g_i - g_j).u per half
   * pair and Fp = F'(rho):
   *   dL = sum_pairs w [dphi' + (Fp_i + Fp_j) drho'] + sum_i lam_i dFp_i,
   *   lam_i = sum_{j} w_ij rho'(r_ij),
   *   dFp_i = dF'(rho_i)|coef + F''(rho_i) drho_i, drho_i = sum_j drho(r_ij). */
  const size_t N = (size_t)(natoms > 0 ? natoms : 1);
  double* f = (double*)malloc(3 * N * sizeof(double));
  double* rho = (double*)malloc(N * sizeof(double));
  double* fp = (double*)malloc(N * sizeof(double));
  double* lam = (double*)calloc(N, sizeof(double));
  int* types = (int*)malloc(N * sizeof(int));
  for (size_t i = 0; i < N; ++i) types[i] = 1;
  orc_eam_forces(e, pos, types, f, natoms, box, nl, rho, fp, NULL);
  if (forces_out) memcpy(forces_out, f, 3 * (size_t)natoms * sizeof(double));
  double L = 0.0;
  for (size_t k = 0; k < 3 * (size_t)natoms; ++k) {
    f[k] -= target[k];
    L += f[k] * f[k];
  }
  const orc_spline* sF = &e->embedding[0];
  const orc_spline* sR = &e->density[0];
  const orc_spline* sP = &e->phi[0];
  const int niF = sF->n - 1, niR = sR->n - 1;
  /* coefficient gradients, blocks [4][n-1] */
  double* cF = (double*)calloc(4 * (size_t)niF, sizeof(double));
  double* cR = (double*)calloc(4 * (size_t)niR, sizeof(double));
  double* cP = (double*)calloc(4 * (size_t)niR, sizeof(double));
  const double rc_sq = e->cutoff * e->cutoff;
  double bs[3];
  box_size(box, bs);
  for (int pass = 0; pass < 2; ++pass) {
    for (int64_t i = 0; i < natoms; ++i) {
      const double* pi = pos + 3 * i;
      const int* nb = nl->nbrs + nl->offsets[i];
      for (int k = 0; k < nl->counts[i]; ++k) {
        const int j = nb[k];
        double d[3] = {pi[0] - pos[3 * j], pi[1] - pos[3 * j + 1], pi[2] - pos[3 * j + 2]};
        orc_minimum_image(d, bs, box->periodic);
        const double r2 = length_sq(d);
        if (r2 >= rc_sq) continue;
        const double r = sqrt(r2);
        double t;
        const int q = spline_locate(sR, r, &t);
        if (pass == 0) {
          const double w = -2.0 * ((f[3 * i] - f[3 * j]) * d[0] + (f[3 * i + 1] - f[3 * j + 1]) * d[1] +
                                   (f[3 * i + 2] - f[3 * j + 2]) * d[2]) / r;
          const double P = fp[i] + fp[j];
          cP[1 * niR + q] += w;
          cP[2 * niR + q] += 2.0 * t * w;
          cP[3 * niR + q] += 3.0 * t * t * w;
          cR[1 * niR + q] += w * P;
          cR[2 * niR + q] += 2.0 * t * w * P;
          cR[3 * niR + q] += 3.0 * t * t * w * P;
          const double drho = orc_spline_eval_deriv(sR, r);
          lam[i] += w * drho;
          lam[j] += w * drho;
        } else {
          /* density values: mu_i rho(r_ij) enters for both ends */
          double ti;
          const int qi = spline_locate(sF, rho[i], &ti);
          double tj;
          const int qj = spline_locate(sF, rho[j], &tj);
          const double mu = lam[i] * (2.0 * sF->c[qi] + 6.0 * sF->d[qi] * ti) +
                            lam[j] * (2.0 * sF->c[qj] + 6.0 * sF->d[qj] * tj);
          cR[0 * niR + q] += mu;
          cR[1 * niR + q] += mu * t;
          cR[2 * niR + q] += mu * t * t;
          cR[3 * niR + q] += mu * t * t * t;
        }
      }
    }
  }
  for (int64_t i = 0; i < natoms; ++i) {
    double t;
    const int q = spline_locate(sF, rho[i], &t);
    cF[1 * niF + q] += lam[i];
    cF[2 * niF + q] += 2.0 * t * lam[i];
    cF[3 * niF + q] += 3.0 * t * t * lam[i];
  }
  orc_spline_coef_grad_to_values(sF->n, sF->dr, cF, cF + niF, cF + 2 * niF, cF + 3 * niF, gF);
  orc_spline_coef_grad_to_values(sR->n, sR->dr, cR, cR + niR, cR + 2 * niR, cR + 3 * niR, grho);
  orc_spline_coef_grad_to_values(sP->n, sP->dr, cP, cP + niR, cP + 2 * niR, cP + 3 * niR, gphi);
  if (gcoef) {
    memcpy(gcoef, cF, 4 * (size_t)niF * sizeof(double));
    memcpy(gcoef + 4 * (size_t)niF, cR, 4 * (size_t)niR * sizeof(double));
    memcpy(gcoef + 4 * (size_t)niF + 4 * (size_t)niR, cP, 4 * (size_t)niR * sizeof(double));
  }
  free(f);
  free(rho);
  free(fp);
  free(lam);
  free(types);
  free(cF);
  free(cR);
  free(cP);
  return L;
}

/* ------------------------------------------------------------------------ */
/* Velocity Verlet: src/integrator/velocity_verlet.cpp:11-48                */
/* ------------------------------------------------------------------------ */

void orc_half_kick(double* v, const double* f, const int* types,
                   const double* masses, int64_t natoms, double dt) {
  /* velocity_verlet.cpp:11-27 */
  const double half_dt = dt * 0.5;
  for (int64_t i = 0; i < natoms; ++i) {
    const double mass = masses[types[i]];
    const double factor = half_dt / (mass * ORC_MVV2E);
    v[3 * i] += factor * f[3 * i];
    v[3 * i + 1] += factor * f[3 * i + 1];
    v[3 * i + 2] += factor * f[3 * i + 2];
  }
}

void orc_drift(double* p, const double* v, int64_t natoms, double dt,
               const orc_box* box) {
  /* velocity_verlet.cpp:29-48: x += dt v, then floor-based wrap. */
  double bs[3];
  box_size(box, bs);
  for (int64_t i = 0; i < natoms; ++i) {
    p[3 * i] += dt * v[3 * i];
    p[3 * i + 1] += dt * v[3 * i + 1];
    p[3 * i + 2] += dt * v[3 * i + 2];
    orc_wrap_position(p + 3 * i, box->lo, bs, box->periodic);
  }
}

void orc_drift_single_image(double* p, const double* v, int64_t natoms,
                            double dt, const orc_box* box) {
  /* device_velocity_verlet.cu:34-66: one conditional +/- box per axis. */
  double bs[3];
  box_size(box, bs);
  for (int64_t i = 0; i < natoms; ++i)
    for (int a = 0; a < 3; ++a) {
      double x = p[3 * i + a] + dt * v[3 * i + a];
      if (box->periodic[a]) {
        if (x < box->lo[a])
          x += bs[a];
        else if (x >= box->lo[a] + bs[a])
          x -= bs[a];
      }
      p[3 * i + a] = x;
    }
}

double orc_kinetic_energy(const double* v, const int* types,
                          const double* masses, int64_t natoms) {
  /* tdmd_main.cpp:97-108 */
  double ke = 0.0;
  for (int64_t i = 0; i < natoms; ++i) {
    const double mass = masses[types[i]];
    ke += 0.5 * mass * ORC_MVV2E * length_sq(v + 3 * i);
  }
  return ke;
}

/* ------------------------------------------------------------------------ */
/* Nose-Hoover chain: src/integrator/device_nose_hoover.cu:187-289          */
/* ------------------------------------------------------------------------ */

void orc_nhc_init(orc_nhc* c, double t_target, double t_period, int M,
                  double dt, int n_dof) {
  /* device_nose_hoover.cu:187-204 */
  memset(c, 0, sizeof(*c));
  c->M = M > ORC_NHC_MAX ? ORC_NHC_MAX : M;
  c->t_target = t_target;
  c->t_period = t_period;
  c->dt = dt;
  c->n_dof = n_dof;
  const double kt = ORC_KB * t_target;
  const double tau2 = t_period * t_period;
  if (c->M > 0) c->eta_m[0] = (double)n_dof * kt * tau2;
  for (int i = 1; i < c->M; ++i) c->eta_m[i] = kt * tau2;
}

double orc_nhc_half_step(orc_nhc* c, double ke) {
  /* device_nose_hoover.cu:206-289 */
  const int M = c->M;
  if (M == 0) return 1.0;
  const double half_dt = 0.5 * c->dt;
  const double ke_target = 0.5 * (double)c->n_dof * ORC_KB * c->t_target;
  const double kt = ORC_KB * c->t_target;
  const double tau2 = c->t_period * c->t_period;
  c->eta_m[0] = (double)c->n_dof * kt * tau2;
  for (int i = 1; i < M; ++i) c->eta_m[i] = kt * tau2;
  const int last = M - 1;
  if (M > 1) {
    const double g_last = (c->eta_m[last - 1] * c->eta_v[last - 1] * c->eta_v[last - 1] -
                           ORC_KB * c->t_target) / c->eta_m[last];
    c->eta_v[last] += g_last * half_dt;
  }
  for (int k = M - 2; k >= 1; --k) {
    const double ef = exp(-c->eta_v[k + 1] * half_dt / 4.0);
    c->eta_v[k] *= ef;
    const double g_k = (c->eta_m[k - 1] * c->eta_v[k - 1] * c->eta_v[k - 1] -
                        ORC_KB * c->t_target) / c->eta_m[k];
    c->eta_v[k] += g_k * half_dt;
    c->eta_v[k] *= ef;
  }
  {
    const double ef = (M > 1) ? exp(-c->eta_v[1] * half_dt / 4.0) : 1.0;
    c->eta_v[0] *= ef;
    const double g0 = (2.0 * ke - 2.0 * ke_target) / c->eta_m[0];
    c->eta_v[0] += g0 * half_dt;
    c->eta_v[0] *= ef;
  }
  const double scale = exp(-c->eta_v[0] * half_dt);
  for (int k = 0; k < M; ++k) c->eta[k] += c->eta_v[k] * half_dt;
  return scale;
}

double orc_nhc_energy(const orc_nhc* c) {
  const double kt = ORC_KB * c->t_target;
  double e = 0.0;
  for (int k = 0; k < c->M; ++k) {
    e += 0.5 * c->eta_m[k] * c->eta_v[k] * c->eta_v[k];
    e += (k == 0 ? (double)c->n_dof * kt : kt) * c->eta[k];
  }
  return e;
}

void orc_scale_velocities(double* v, int64_t natoms, double s) {
  /* device_nose_hoover.cu:58-65 */
  for (int64_t i = 0; i < 3 * natoms; ++i) v[i] *= s;
}

/* ------------------------------------------------------------------------ */
/* Driver loop: src/drivers/tdmd_main.cpp:211-254; NVT order:              */
/* src/scheduler/pipeline_scheduler.cu:246-276                              */
/* ------------------------------------------------------------------------ */

static double compute_forces(const orc_run_cfg* cfg, const orc_box* box,
                             const double* pos, double* f, const int* types,
                             int64_t natoms, const orc_nlist* nl) {
  if (cfg->use_eam)
    return orc_eam_forces(cfg->eam, pos, types, f, natoms, box, nl, NULL, NULL, NULL);
  return orc_morse_forces(pos, f, natoms, box, nl, cfg->morse[0], cfg->morse[1],
                          cfg->morse[2], cfg->morse[3]);
}

static void record(orc_run_out* out, int step, double pe, double ke,
                   int64_t natoms, double nhc_e) {
  if (!out->thermo || out->nthermo >= out->thermo_cap) {
    ++out->nthermo;
    return;
  }
  double* row = out->thermo + 6 * out->nthermo;
  row[0] = step;
  row[1] = pe;
  row[2] = ke;
  row[3] = pe + ke;
  /* tdmd_main.cpp:112-117: T with 3N degrees of freedom. */
  row[4] = natoms > 0 ? 2.0 * ke / (3.0 * (double)natoms * ORC_KB) : 0.0;
  row[5] = nhc_e;
  ++out->nthermo;
}

int orc_run(const orc_run_cfg* cfg, const orc_box* box, double* pos,
            double* vel, double* f, const int* types, const double* masses,
            int64_t natoms, orc_run_out* out) {
  const double cutoff = cfg->use_eam ? cfg->eam->cutoff : cfg->morse[3];
  orc_nlist nl;
  memset(&nl, 0, sizeof(nl));
  out->rebuilds = 0;
  out->nthermo = 0;
  orc_nhc nhc;
  /* pipeline_scheduler.cu:33-41: n_dof = 3N - 3. */
  orc_nhc_init(&nhc, cfg->t_target, cfg->t_period, cfg->nvt ? cfg->nhc_length : 0,
               cfg->dt, (int)(3 * natoms - 3));
  if (orc_nlist_build(&nl, pos, natoms, box, cutoff, cfg->skin) != 0) return -1;
  double pe = compute_forces(cfg, box, pos, f, types, natoms, &nl);
  double ke = orc_kinetic_energy(vel, types, masses, natoms);
  record(out, 0, pe, ke, natoms, cfg->nvt ? orc_nhc_energy(&nhc) : 0.0);
  int since = 0;
  for (int step = 0; step < cfg->nsteps; ++step) {
    if (cfg->nvt) {
      const double k1 = orc_kinetic_energy(vel, types, masses, natoms);
      orc_scale_velocities(vel, natoms, orc_nhc_half_step(&nhc, k1));
    }
    orc_half_kick(vel, f, types, masses, natoms, cfg->dt);
    orc_drift(pos, vel, natoms, cfg->dt, box);
    ++since;
    int rebuild;
    if (cfg->rebuild_mode == 1)
      rebuild = since >= cfg->rebuild_every;
    else
      rebuild = orc_nlist_needs_rebuild(&nl, pos, natoms);
    if (rebuild) {
      if (orc_nlist_build(&nl, pos, natoms, box, cutoff, cfg->skin) != 0) {
        orc_nlist_free(&nl);
        return -1;
      }
      ++out->rebuilds;
      since = 0;
    }
    pe = compute_forces(cfg, box, pos, f, types, natoms, &nl);
    orc_half_kick(vel, f, types, masses, natoms, cfg->dt);
    if (cfg->nvt) {
      const double k2 = orc_kinetic_energy(vel, types, masses, natoms);
      orc_scale_velocities(vel, natoms, orc_nhc_half_step(&nhc, k2));
    }
    if (cfg->thermo_every > 0 && (step + 1) % cfg->thermo_every == 0) {
      ke = orc_kinetic_energy(vel, types, masses, natoms);
      record(out, step + 1, pe, ke, natoms, cfg->nvt ? orc_nhc_energy(&nhc) : 0.0);
    }
  }
  out->last_pe = pe;
  orc_nlist_free(&nl);
  return 0;
}
