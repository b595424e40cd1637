/*
 * tdmd_oracle.h -- FP64 CPU restatement of TDMD's single-GPU step path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it. The HIP library never links or calls it.
 *
 * Every function restates one reference routine (slava8519/tdmd_Claude); the
 * citation on each declaration names the reference file:line it follows.
 * Everything is double precision: the reference's FP64 build
 * (TDMD_PRECISION_FP64, src/core/types.hpp:11-17) makes real = double.
 *
 * Pinning (what checks each function beyond re-deriving it):
 *   cells/nlist  -- brute-force O(N^2) pair sets; reference tdmd_standalone
 *                   rebuild counts (oracle/_ref).
 *   morse        -- two-atom analytic JSON (verifylab two-atoms-morse),
 *                   U(r0) = -D, Newton-3 antisymmetry, finite differences,
 *                   reference tdmd_standalone PE/force dumps.
 *   spline/setfl -- knot reproduction a_i = v_i, hand-computed FD-Hermite
 *                   slopes, reference tdmd_standalone EAM dumps (Mishin Cu).
 *   eam forces   -- finite differences of the energy, brute-force density,
 *                   reference tdmd_standalone EAM PE/forces.
 *   vv / driver  -- reference tdmd_standalone thermo trajectories.
 *   nhc          -- PARITY UNPINNED against the reference binary (the
 *                   reference NHC lives in a CUDA translation unit that does
 *                   not compile here); pinned by hand-computed half-steps,
 *                   the closed form for M=1, and <T> statistics.
 */
#ifndef TDMD_ORACLE_H
#define TDMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Orthorhombic box, src/core/box.hpp:17-26. */
typedef struct {
  double lo[3];
  double hi[3];
  int periodic[3];
} orc_box;

/* Metal-unit constants, src/core/constants.hpp:15,28. */
#define ORC_KB 8.617333262e-5
#define ORC_MVV2E 1.0364269e-4

/* ---- cell list: src/domain/cell_list.cpp:13-58, cell_list.hpp:47-72 ---- */
typedef struct {
  int ncx, ncy, ncz;
  double cs[3];
  int ncells;
  int* counts;  /* [ncells] */
  int* offsets; /* [ncells] exclusive */
  int* atoms;   /* [natoms] atom indices grouped by cell, index order */
} orc_cells;

int orc_cells_build(orc_cells* cl, const double* pos, int64_t natoms,
                    const orc_box* box, double r_list);
int orc_cell_of(const orc_cells* cl, const double* p, const double* lo);
void orc_cells_free(orc_cells* cl);

/* ---- half neighbour list: src/neighbors/neighbor_list.cpp:14-100 ---- */
typedef struct {
  int64_t natoms;
  double r_cut, r_skin;
  int* counts;   /* [natoms] */
  int* offsets;  /* [natoms] */
  int* nbrs;     /* [total] j > i only */
  int64_t total;
  double* build_pos; /* [3*natoms] */
} orc_nlist;

int orc_nlist_build(orc_nlist* nl, const double* pos, int64_t natoms,
                    const orc_box* box, double r_cut, double r_skin);
int orc_nlist_needs_rebuild(const orc_nlist* nl, const double* pos,
                            int64_t natoms);
void orc_nlist_free(orc_nlist* nl);

/* Full list = symmetric closure of the half list, rows sorted ascending.
 * The reference GPU list (src/neighbors/device_neighbor_list.cu:93-206)
 * stores both (i,j) and (j,i); this is the set it must equal. */
int orc_nlist_full(const orc_nlist* half, int* counts, int* offsets,
                   int* nbrs);

/* Brute force O(N^2) full list (minimum image, r2 < r_list^2). */
int64_t orc_bruteforce_full(const double* pos, int64_t natoms,
                            const orc_box* box, double r_list, int* counts,
                            int* offsets, int* nbrs, int64_t capacity);

/* ---- PBC: src/core/math.hpp:78-131 ---- */
void orc_minimum_image(double* d, const double* box_size, const int* periodic);
void orc_wrap_position(double* p, const double* lo, const double* box_size,
                       const int* periodic);

/* ---- Morse: src/potentials/morse.hpp:37-55,
 *             src/potentials/force_compute.cpp:10-56 ---- */
void orc_morse_pair(double D, double alpha, double r0, double rc_sq,
                    double r2, double* energy, double* fpair);
double orc_morse_forces(const double* pos, double* forces, int64_t natoms,
                        const orc_box* box, const orc_nlist* nl, double D,
                        double alpha, double r0, double rc);
/* compute_morse_gpu_zone (device_morse_zone.cu:29-112), FP64: rows
 * [first, first + count) of a full list; forces ADDED to f. */
void orc_morse_forces_rows(const double* pos, double* f, int first, int count,
                           const int* counts, const int* offsets, const int* nbrs,
                           const orc_box* box, double D, double alpha, double r0,
                           double rc);

/* ---- spline: src/potentials/eam_alloy.cpp:19-72 ---- */
typedef struct {
  int n;
  double dr, rmin;
  double* values; /* [n] */
  double *a, *b, *c, *d; /* [n-1] */
} orc_spline;

void orc_spline_build(orc_spline* s);
double orc_spline_eval(const orc_spline* s, double r);
double orc_spline_eval_deriv(const orc_spline* s, double r);
void orc_spline_free(orc_spline* s);

/* ---- EAM/alloy: src/potentials/eam_alloy.cpp:90-295,
 *                  src/potentials/eam_alloy.hpp:84-92 ---- */
typedef struct {
  int ntypes;
  double cutoff;
  int nrho, nr;
  double drho, dr;
  orc_spline* embedding; /* [ntypes] F(rho) */
  orc_spline* density;   /* [ntypes] rho(r) */
  orc_spline* phi;       /* [ntypes*(ntypes+1)/2] phi(r), upper triangle */
} orc_eam;

int orc_eam_read_setfl(orc_eam* e, const char* path, char* err, int errlen);
/* Same post-processing as read_setfl, from in-memory arrays:
 * F[t*nrho+k], rho[t*nr+k], rphi[p*nr+k] with p in i<=j row order. */
int orc_eam_from_arrays(orc_eam* e, int ntypes, int nrho, double drho, int nr,
                        double dr, double cutoff, const double* F,
                        const double* rho, const double* rphi);
int orc_eam_phi_index(int ntypes, int ti, int tj);
/* 3-pass EAM over the half list. rho_out / fp_out may be NULL. */
double orc_eam_forces(const orc_eam* e, const double* pos, const int* types,
                      double* forces, int64_t natoms, const orc_box* box,
                      const orc_nlist* nl, double* rho_out, double* fp_out,
                      double* embed_energy_out);
void orc_eam_free(orc_eam* e);

/* ---- force matching (no reference counterpart: the gradient the fitting
 * harness needs, stated here in FP64 as the checker of the GPU adjoint).
 * L = sum_i |F_i - T_i|^2 with F from orc_morse_forces / orc_eam_forces over
 * the half list `nl`. Pinned by central finite differences of L. ---- */

/* dL/d(D, alpha, r0) -> grad[3]; returns L. forces_out (3N) may be NULL. */
double orc_fm_morse(const double* pos, int64_t natoms, const orc_box* box,
                    const orc_nlist* nl, double D, double alpha, double r0,
                    double rc, const double* target, double grad[3],
                    double* forces_out);

/* Single-element EAM whose parameters are the spline knot values as EamAlloy
 * holds them (F[nrho], rho[nr], phi[nr] with phi = (r phi)/r). */
int orc_eam_from_values(orc_eam* e, int nrho, double drho, int nr, double dr,
                        double cutoff, const double* F, const double* rho,
                        const double* phi);
/* dL/d(F knots, rho knots, phi knots) -> gF[nrho], grho[nr], gphi[nr], through
 * the FD-Hermite coefficients (eam_alloy.cpp:19-50); returns L.
 * gcoef (optional, 4*(nrho-1) + 8*(nr-1) doubles) receives the gradient with
 * respect to the cubic coefficients: F {a,b,c,d}, rho {a,b,c,d},
 * phi {a,b,c,d}, each block [4][n-1]. */
double orc_fm_eam(const orc_eam* e, const double* pos, int64_t natoms,
                  const orc_box* box, const orc_nlist* nl, const double* target,
                  double* gF, double* grho, double* gphi, double* gcoef,
                  double* forces_out);

/* Transpose of the FD-Hermite map for one spline: knot-value gradient gv[n]
 * from coefficient gradients ga, gb, gc, gd [n-1]. */
void orc_spline_coef_grad_to_values(int n, double dr, const double* ga,
                                    const double* gb, const double* gc,
                                    const double* gd, double* gv);

/* ---- velocity Verlet: src/integrator/velocity_verlet.cpp:11-48 ---- */
void orc_half_kick(double* vel, const double* forces, const int* types,
                   const double* masses, int64_t natoms, double dt);
void orc_drift(double* pos, const double* vel, int64_t natoms, double dt,
               const orc_box* box);
/* Single-image wrap of the GPU drift, device_velocity_verlet.cu:34-66. */
void orc_drift_single_image(double* pos, const double* vel, int64_t natoms,
                            double dt, const orc_box* box);

/* KE, tdmd_main.cpp:97-108. */
double orc_kinetic_energy(const double* vel, const int* types,
                          const double* masses, int64_t natoms);

/* ---- NHC: src/integrator/device_nose_hoover.cu:187-289 ---- */
#define ORC_NHC_MAX 8
typedef struct {
  int M;
  double t_target, t_period, dt;
  int n_dof;
  double eta[ORC_NHC_MAX], eta_v[ORC_NHC_MAX], eta_m[ORC_NHC_MAX];
} orc_nhc;

void orc_nhc_init(orc_nhc* c, double t_target, double t_period, int M,
                  double dt, int n_dof);
double orc_nhc_half_step(orc_nhc* c, double ke);
/* Conserved-energy bookkeeping term of the chain (not in the reference,
 * which never computes it): sum_k 0.5 Q_k v_k^2 + n_dof kT eta_0
 * + kT sum_{k>0} eta_k. */
double orc_nhc_energy(const orc_nhc* c);
void orc_scale_velocities(double* vel, int64_t natoms, double s);

/* ---- driver loop: src/drivers/tdmd_main.cpp:211-254 (NVE, displacement
 * rebuild) and src/scheduler/pipeline_scheduler.cu:246-276 (NVT order).
 * thermo[k] rows = {step, pe, ke, te, T(3N dof), nhc_energy}. ---- */
typedef struct {
  int use_eam;
  double morse[4]; /* D, alpha, r0, rc */
  const orc_eam* eam;
  double dt, skin;
  int nsteps, thermo_every;
  int rebuild_mode;   /* 0 = displacement (driver), 1 = fixed interval */
  int rebuild_every;  /* fixed-interval mode */
  int nvt;            /* 1 = NHC around each step */
  double t_target, t_period;
  int nhc_length;
} orc_run_cfg;

typedef struct {
  int rebuilds;
  int nthermo;
  double* thermo; /* [nthermo*6] caller-allocated, capacity thermo_cap */
  int thermo_cap;
  double last_pe;
} orc_run_out;

int orc_run(const orc_run_cfg* cfg, const orc_box* box, double* pos,
            double* vel, double* forces, const int* types,
            const double* masses, int64_t natoms, orc_run_out* out);

#ifdef __cplusplus
}
#endif

#endif /* TDMD_ORACLE_H */
