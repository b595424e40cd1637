/*
 * tdmd_hip.h -- C-ABI of the MI355X (gfx950) implementation of TDMD's
 * single-GPU step path: cell binning, Verlet neighbour lists, Morse and EAM
 * forces, velocity Verlet and the Nose-Hoover chain, and the
 * FastPipelineScheduler that drives them.
 *
 * Plain pointers and sizes only. Every entry point returns 0 (TDMD_OK) or a
 * tdmd_status code; tdmd_last_error() returns the message of the last
 * failure on the calling thread. Device pointers come from tdmd_malloc (or
 * any hipMalloc'd memory). `stream` is a hipStream_t passed as void*; NULL
 * means the null stream. Entry points that launch kernels never allocate
 * device memory unless the comment says so.
 *
 * Each declaration cites the reference (slava8519/tdmd_Claude) interface it
 * replaces, file:line.
 */
#ifndef TDMD_HIP_H
#define TDMD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDMD_ABI_VERSION 1

typedef enum {
  TDMD_OK = 0,
  TDMD_E_INVALID = 1,   /* bad argument (reference: TDMD_ASSERT abort)      */
  TDMD_E_HIP = 2,       /* HIP runtime failure (reference: TDMD_CUDA_CHECK) */
  TDMD_E_OVERFLOW = 3,  /* neighbour capacity exceeded                      */
  TDMD_E_NONFINITE = 4, /* NaN/Inf found in state                           */
  TDMD_E_NOMEM = 5,
  TDMD_E_STATE = 6,     /* call out of order (e.g. run before upload)       */
  TDMD_E_IO = 7         /* malformed input file (reference: tdmd::Error)    */
} tdmd_status;

/* Precision of `real` and ForceVec at the boundary
 * (src/core/types.hpp:11-17,57-73): FP64 -> real = double, forces double;
 * MIXED -> real = float, forces float. Positions and velocities are double
 * in both. */
typedef enum { TDMD_FP64 = 0, TDMD_MIXED = 1 } tdmd_precision;

/* Orthorhombic box, src/core/box.hpp:17-26. */
typedef struct {
  double lo[3];
  double hi[3];
  int32_t periodic[3];
} tdmd_box;

/* ------------------------------------------------------------------ */
/* Runtime (DeviceBuffer<T>, src/core/device_buffer.cuh:29-116)        */
/* ------------------------------------------------------------------ */
int tdmd_abi_version(void);
const char* tdmd_last_error(void);
int tdmd_device_count(int32_t* count);
int tdmd_set_device(int32_t device);
int tdmd_device_name(char* buf, int32_t buflen);
int tdmd_malloc(void** dptr, size_t bytes);
int tdmd_free(void* dptr);
/* stream == NULL: synchronous copy (DeviceBuffer::copy_from_host :78-82). */
int tdmd_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
int tdmd_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
int tdmd_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
/* Kernel fill (graph-capture safe); DeviceBuffer::zero :103-107. */
int tdmd_fill_u32(void* dst, uint32_t value, size_t count, void* stream);
/* Kernel copy of `count` 32-bit words (DeviceBuffer device-to-device copy;
 * also the known-size kernel that calibrates HBM byte counters). */
int tdmd_copy_u32(void* dst, const void* src, size_t count, void* stream);
int tdmd_stream_create(void** stream);
int tdmd_stream_destroy(void* stream);
int tdmd_stream_synchronize(void* stream);
int tdmd_device_synchronize(void);
int tdmd_event_create(void** ev);
int tdmd_event_destroy(void* ev);
int tdmd_event_record(void* ev, void* stream);
int tdmd_event_elapsed_ms(void* start, void* stop, float* ms);

/* ------------------------------------------------------------------ */
/* Prefix scan (replaces cub::DeviceScan::ExclusiveSum + DeviceReduce  */
/* ::Max + pack_meta_kernel, src/neighbors/device_neighbor_list.cu:    */
/* 147-154,211-243). Atomics-free. meta_out[0] = total, [1] = max.     */
/* ------------------------------------------------------------------ */
int tdmd_scan_workspace_bytes(int64_t n, size_t* bytes);
int tdmd_scan_exclusive_i32(const int32_t* d_in, int32_t* d_out, int64_t n,
                            int32_t* d_meta_out, void* d_work, void* stream);

/* ------------------------------------------------------------------ */
/* Cell list: DeviceCellList (src/domain/device_cell_list.cuh:16-54,   */
/* device_cell_list.cu:59-150). Grid: n = max(3, floor(L / r_list)).   */
/* Within a cell atoms are stored in increasing index order (the       */
/* reference's deterministic order, device_cell_list.cu:114-136).      */
/* ------------------------------------------------------------------ */
int tdmd_cell_grid(const tdmd_box* box, double r_list, int32_t ncells[3],
                   double cell_size[3]);
int tdmd_cell_workspace_bytes(int64_t natoms, int32_t ncells_total,
                              size_t* bytes);
int tdmd_cell_build(const double* d_pos, int64_t natoms, const tdmd_box* box,
                    double r_list, int32_t* d_cell_counts,
                    int32_t* d_cell_offsets, int32_t* d_cell_atoms,
                    int32_t* d_atom_cells, void* d_work, void* stream);

/* ------------------------------------------------------------------ */
/* Full CSR neighbour list: DeviceNeighborList::build                  */
/* (src/neighbors/device_neighbor_list.cuh:20-65,                      */
/*  device_neighbor_list.cu:158-278). Inclusion is decided in FP64:    */
/* d = x_j - x_i, single minimum image, r2 = dx*dx + dy*dy + dz*dz     */
/* (unfused), r2 < r_list^2 -- the CPU list's test                     */
/* (src/neighbors/neighbor_list.cpp:58-62), so the per-atom sets equal */
/* the symmetric closure of the CPU half list bit for bit.             */
/* ------------------------------------------------------------------ */
int tdmd_nlist_count(const double* d_pos, int64_t natoms, const tdmd_box* box,
                     double r_list, const int32_t* d_cell_counts,
                     const int32_t* d_cell_offsets,
                     const int32_t* d_cell_atoms, const int32_t* d_atom_cells,
                     const int32_t ncells[3], int32_t* d_counts, void* stream);
/* Writes neighbours at d_offsets[i]; d_counts gets the true counts; atoms
 * beyond max_neighbors are dropped and *d_overflow gets max count seen. */
int tdmd_nlist_fill(const double* d_pos, int64_t natoms, const tdmd_box* box,
                    double r_list, const int32_t* d_cell_counts,
                    const int32_t* d_cell_offsets, const int32_t* d_cell_atoms,
                    const int32_t* d_atom_cells, const int32_t ncells[3],
                    const int32_t* d_offsets, int32_t* d_neighbors,
                    int32_t* d_counts, int32_t max_neighbors,
                    int32_t* d_overflow, void* stream);

/* ------------------------------------------------------------------ */
/* Morse: MorseParams + compute_morse_gpu                              */
/* (src/potentials/device_morse.cuh:14-41, device_morse.cu:136-170).   */
/* Forces are STORED (the reference code stores, :110-117); *d_energy  */
/* is ADDED TO (caller zeroes). d_work == NULL: energy partials are    */
/* added with one FP64 atomic per block; otherwise d_work (8 bytes per */
/* block of 256 atoms) gives a fixed-order, run-to-run identical sum.  */
/* d_forces is float[3N] for TDMD_MIXED, double[3N] for TDMD_FP64.     */
/* ------------------------------------------------------------------ */
typedef struct {
  double d, alpha, r0, rc, rc_sq;
} tdmd_morse_params;

int tdmd_morse_force(int32_t precision, const double* d_pos, void* d_forces,
                     const int32_t* d_neighbors, const int32_t* d_offsets,
                     const int32_t* d_counts, int32_t natoms,
                     const tdmd_box* box, const tdmd_morse_params* params,
                     double* d_energy, void* d_work, void* stream);
/* Zone pass (compute_morse_gpu_zone, src/potentials/device_morse_zone.cuh:
 * 14-22): atoms [first_atom, first_atom + atom_count) over the whole CSR
 * list and position array; forces are ADDED to d_forces (the caller zeroes
 * the zone, device_morse_zone.cu:110-112); *d_energy is added to. d_work: as
 * tdmd_morse_force, one double per 256 zone atoms. */
int tdmd_morse_force_zone(int32_t precision, const double* d_pos, void* d_forces,
                          const int32_t* d_neighbors, const int32_t* d_offsets,
                          const int32_t* d_counts, int32_t first_atom,
                          int32_t atom_count, const tdmd_box* box,
                          const tdmd_morse_params* params, double* d_energy,
                          void* d_work, void* stream);

/* ------------------------------------------------------------------ */
/* FD-Hermite spline (SplineTable::build_splines,                      */
/* src/potentials/eam_alloy.cpp:19-50). Host routine; a,b,c,d get n-1. */
/* ------------------------------------------------------------------ */
int tdmd_spline_build(const double* values, int32_t n, double dr,
                      double* a, double* b, double* c, double* d);

/* setfl reader (EamAlloy::read_setfl, src/potentials/eam_alloy.cpp:    */
/* 90-192). Two calls: with NULL arrays it fills the sizes; then with   */
/* arrays F[ntypes*nrho], rho[ntypes*nr], rphi[np*nr] (np = n(n+1)/2,   */
/* pair blocks in i<=j row order as the reference reads them).          */
int tdmd_setfl_read(const char* path, int32_t* ntypes, int32_t* nrho,
                    double* drho, int32_t* nr, double* dr, double* cutoff,
                    double* F, double* rho, double* rphi);

/* ------------------------------------------------------------------ */
/* EAM tables on the device: DeviceEam::upload_tables                  */
/* (src/potentials/device_eam.cuh:26-75, device_eam.cu:282-340).       */
/* Tables are uploaded once; one spline interval is one {a,b,c,d}      */
/* quad (float4 in MIXED, double4 in FP64) so a lookup is one load.    */
/* ------------------------------------------------------------------ */
typedef struct tdmd_eam_tables tdmd_eam_tables;

int tdmd_eam_tables_create(tdmd_eam_tables** out, int32_t precision,
                           int32_t ntypes, int32_t nrho, double drho,
                           int32_t nr, double dr, double cutoff,
                           const double* F, const double* rho,
                           const double* rphi);
/* Same, from knot values as EamAlloy holds them after read_setfl
 * (eam_alloy.cpp:152-191): F and rho as read, phi = (r phi)/r with
 * phi[0] = 0, pair blocks in phi_index order (eam_alloy.hpp:88-92). This is
 * what DeviceEam::upload_tables(const EamAlloy&) receives. */
int tdmd_eam_tables_create_values(tdmd_eam_tables** out, int32_t precision,
                                  int32_t ntypes, int32_t nrho, double drho,
                                  int32_t nr, double dr, double cutoff,
                                  const double* F, const double* rho,
                                  const double* phi);
int tdmd_eam_tables_destroy(tdmd_eam_tables* t);
int tdmd_eam_tables_info(const tdmd_eam_tables* t, int32_t* ntypes,
                         double* cutoff);

/* DeviceEam::compute (device_eam.cu:342-421): density, embedding, force
 * over a full CSR list; forces stored, *d_energy added to. d_rho (double
 * [N]) and d_fp (real [N]) are caller-provided per-atom scratch (the
 * reference allocates them inside compute, :353-357). d_work as for
 * tdmd_morse_force. */
int tdmd_eam_compute(const tdmd_eam_tables* t, const double* d_pos,
                     void* d_forces, const int32_t* d_types,
                     const int32_t* d_neighbors, const int32_t* d_offsets,
                     const int32_t* d_counts, int32_t natoms,
                     const tdmd_box* box, double* d_rho, void* d_fp,
                     double* d_energy, void* d_work, void* stream);
/* No counterpart in the reference: tdmd_eam_compute for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The same kernels; rho, F'(rho) and the
 * fixed-order energy partials live in a library-owned device array of
 * ~16.8 MB per device (held until the process ends), so at most
 * 1,048,576 atoms (larger natoms: TDMD_E_INVALID, nothing launched). It
 * calls no hipMalloc/hipFree. Calls on one device share that array: calls
 * that overlap on one device (other streams, other threads) race. */
int tdmd_eam_compute_noalloc(const tdmd_eam_tables* t, const double* d_pos,
                             void* d_forces, const int32_t* d_types,
                             const int32_t* d_neighbors, const int32_t* d_offsets,
                             const int32_t* d_counts, int32_t natoms,
                             const tdmd_box* box, double* d_energy, void* stream);

/* ------------------------------------------------------------------ */
/* Velocity Verlet (src/integrator/device_velocity_verlet.cuh:13-59).  */
/* FP64 arithmetic; d_forces and d_masses are `real`/ForceVec of       */
/* `precision`; masses are indexed by the 1-based type.                */
/* ------------------------------------------------------------------ */
int tdmd_half_kick(int32_t precision, double* d_vel, const void* d_forces,
                   const int32_t* d_types, const void* d_masses,
                   int32_t natoms, double dt, void* stream);
int tdmd_drift(double* d_pos, const double* d_vel, int32_t natoms, double dt,
               const tdmd_box* box, void* stream);
/* full_kick = 0: kick by dt/2 then drift (device_fused_half_kick_drift);
 * full_kick = 1: kick by dt then drift (device_fused_full_kick_drift). */
int tdmd_fused_kick_drift(int32_t precision, int32_t full_kick, double* d_vel,
                          double* d_pos, const void* d_forces,
                          const int32_t* d_types, const void* d_masses,
                          int32_t natoms, double dt, const tdmd_box* box,
                          void* stream);
int tdmd_zero_forces(int32_t precision, void* d_forces, int32_t natoms,
                     void* stream);
/* Zone variants (src/integrator/device_velocity_verlet_zone.cuh:12-26):
 * the kernels above over atoms [first_atom, first_atom + atom_count). */
int tdmd_half_kick_zone(int32_t precision, double* d_vel, const void* d_forces,
                        const int32_t* d_types, const void* d_masses,
                        int32_t first_atom, int32_t atom_count, double dt,
                        void* stream);
int tdmd_drift_zone(double* d_pos, const double* d_vel, int32_t first_atom,
                    int32_t atom_count, double dt, const tdmd_box* box,
                    void* stream);
int tdmd_zero_forces_zone(int32_t precision, void* d_forces, int32_t first_atom,
                          int32_t atom_count, void* stream);

/* ------------------------------------------------------------------ */
/* Nose-Hoover support (src/integrator/device_nose_hoover.cuh:26-88).  */
/* The reductions write their FP64 result to device memory (no host    */
/* round trip); d_work holds one double per block of 256 atoms.        */
/* ------------------------------------------------------------------ */
int tdmd_reduce_workspace_bytes(int64_t n, size_t* bytes);
int tdmd_compute_ke(int32_t precision, const double* d_vel,
                    const int32_t* d_types, const void* d_masses,
                    int32_t natoms, double* d_ke_out, void* d_work,
                    void* stream);
int tdmd_scale_velocities(double* d_vel, int32_t natoms, double factor,
                          void* stream);
/* max |v| (device_compute_vmax :159-183), result sqrt'ed, on device. */
int tdmd_compute_vmax(const double* d_vel, int32_t natoms, double* d_vmax_out,
                      void* d_work, void* stream);
/* NoseHooverChain::half_step (device_nose_hoover.cu:206-289), host. The
 * chain state is eta[M], eta_v[M], eta_m[M]; returns the scale factor. */
double tdmd_nhc_half_step(int32_t M, double* eta, double* eta_v,
                          double* eta_m, double t_target, double t_period,
                          double dt, int32_t n_dof, double ke);

/* ------------------------------------------------------------------ */
/* FastPipelineScheduler (src/scheduler/fast_pipeline_scheduler.cuh:   */
/* 36-192, fast_pipeline_scheduler.cu:15-232), one GPU, one stream.    */
/* Host arrays are in upload order; download returns upload order.     */
/* ------------------------------------------------------------------ */
typedef struct tdmd_sched tdmd_sched;

typedef enum {
  TDMD_REBUILD_FIXED = 0,        /* every rebuild_every steps (reference) */
  TDMD_REBUILD_DISPLACEMENT = 1  /* max |dx| > skin/2, decided on device  */
} tdmd_rebuild_policy;

typedef struct {
  /* FastPipelineConfig (fast_pipeline_scheduler.cuh:42-46) */
  double dt;
  double r_skin;
  int32_t rebuild_every;
  /* additions */
  int32_t precision;      /* tdmd_precision of the force evaluation       */
  int32_t rebuild_policy; /* tdmd_rebuild_policy                          */
  double t_target;        /* > 0: NVT with a Nose-Hoover chain            */
  double t_period;        /* NHC coupling period (ps)                     */
  int32_t nhc_length;     /* chain length                                 */
  int32_t use_graph;      /* capture step sequences as hipGraphs          */
  int32_t block_size;     /* 0 = default/autotuned                        */
  int32_t mirror_fp64;    /* MIXED only: gather FP64 positions (1) or the */
                          /* 32-bit fixed-point mirror (0)                */
  int32_t max_neighbors;  /* 0 = sized from the first build               */
  int32_t pair_unroll;    /* neighbour slots per pipelined group: 0 (=4   */
                          /* mixed, 8 FP64), 4 or 8 (single-element EAM  */
                          /* passes)                                      */
  int32_t lds_budget;     /* LDS bytes per workgroup for spline windows:  */
                          /* 0 = device limit                             */
  int32_t split_rebuild;  /* 1: neighbour rebuild as separate kernels;    */
                          /* 0 (default): one fused launch per step       */
  int32_t disable_halo;   /* 1: no LDS halo staging in the single-element */
                          /* EAM passes (0: staged where records are 16 B)*/
  int32_t pair_lanes;     /* lanes per neighbour row in the force passes:  */
                          /* 0 = by size (8 up to 131072 atoms for Morse, */
                          /* 65536 for EAM, else 1),                      */
                          /* 1, 4 or 8; EAM with more than one lane runs  */
                          /* the global-table passes (no LDS tables/halo) */
  int32_t cell_pct;       /* binning cells at least this percentage of     */
                          /* r_list per axis: 0 (= 100) or 100..400       */
  int32_t reserved[2];
} tdmd_sched_config;

void tdmd_sched_config_default(tdmd_sched_config* cfg);

int tdmd_sched_create_morse(tdmd_sched** out, const tdmd_box* box,
                            int32_t natoms, const tdmd_morse_params* params,
                            const tdmd_sched_config* cfg);
/* EAM tables given as setfl content (see tdmd_setfl_read). */
int tdmd_sched_create_eam(tdmd_sched** out, const tdmd_box* box,
                          int32_t natoms, int32_t ntypes, int32_t nrho,
                          double drho, int32_t nr, double dr, double cutoff,
                          const double* F, const double* rho,
                          const double* rphi, const tdmd_sched_config* cfg);
/* EAM from tables already on the device (the FastPipelineScheduler(box,
 * natoms, const EamAlloy&, cfg) constructor after DeviceEam::upload_tables,
 * fast_pipeline_scheduler.cu:27-40). The scheduler keeps its own copy in
 * cfg->precision; `tables` stays owned by the caller. */
int tdmd_sched_create_eam_tables(tdmd_sched** out, const tdmd_box* box,
                                 int32_t natoms, const tdmd_eam_tables* tables,
                                 const tdmd_sched_config* cfg);
/* A batch of nrep independent systems (replicas) advanced together: box
 * boxes[r] holds natoms[r] atoms; upload/download take the replicas'
 * arrays concatenated in replica order. Every step pass is one launch over
 * all replicas (no reference counterpart: the reference runs one system per
 * scheduler). NVE only; the rebuild test is shared by all replicas. */
int tdmd_sched_create_eam_batch(tdmd_sched** out, int32_t nrep,
                                const tdmd_box* boxes, const int32_t* natoms,
                                int32_t ntypes, int32_t nrho, double drho,
                                int32_t nr, double dr, double cutoff,
                                const double* F, const double* rho,
                                const double* rphi,
                                const tdmd_sched_config* cfg);
int tdmd_sched_create_morse_batch(tdmd_sched** out, int32_t nrep,
                                  const tdmd_box* boxes, const int32_t* natoms,
                                  const tdmd_morse_params* params,
                                  const tdmd_sched_config* cfg);
int tdmd_sched_destroy(tdmd_sched* s);
/* upload (fast_pipeline_scheduler.cu:49-73). pos, vel, forces: double[3N];
 * masses: double[n_masses] indexed by 1-based type. forces may be NULL. */
int tdmd_sched_upload(tdmd_sched* s, const double* pos, const double* vel,
                      const double* forces, const int32_t* types,
                      const int32_t* ids, const double* masses,
                      int32_t n_masses);
/* run_until (fast_pipeline_scheduler.cu:195-223): one stream sync. */
int tdmd_sched_run_until(tdmd_sched* s, int32_t target_step);
/* download (fast_pipeline_scheduler.cu:75-85); any pointer may be NULL. */
int tdmd_sched_download(const tdmd_sched* s, double* pos, double* vel,
                        double* forces, int32_t* types, int32_t* ids,
                        int32_t natoms);
/* stats (fast_pipeline_scheduler.cuh:49-53): out[0] ticks,
 * [1] kernel_launches, [2] rebuilds, [3] graph_launches. */
int tdmd_sched_stats(const tdmd_sched* s, int64_t out[4]);
int tdmd_sched_current_step(const tdmd_sched* s, int32_t* step);
int tdmd_sched_interaction_cutoff(const tdmd_sched* s, double* rc);
/* Thermo at the current step (velocities at integer time):
 * out[0] PE, [1] KE, [2] NHC chain energy, [3] T (3N dof), [4] max
 * neighbour count, [5] mean neighbour count. Runs one energy pass. */
int tdmd_sched_thermo(tdmd_sched* s, double out[6]);
/* Observables at the current step (no reference counterpart; the thermo
 * and pressure a driver reports, tdmd_main.cpp:97-122, computed on the
 * device in FP64). One energy pass. virial W = sum over pairs of r_ij .
 * f_ij (eV); pressure P = (2 KE + W) / (3 V) in bar; T from 3N degrees of
 * freedom as the reference's thermo line; conserved = PE + KE + chain
 * energy (NVT) or PE + KE (NVE). Batches: sums over the replicas, V the
 * summed box volumes. */
typedef struct tdmd_observables {
  int64_t step;
  int64_t natoms;
  double pe, ke, te, temp;
  double nhc_energy, conserved;
  double virial, volume, pressure;
  int32_t max_neighbors;
  int32_t pad;
  double mean_neighbors;      /* mean row length (pairs within r_list)     */
  double mean_near_neighbors; /* mean pairs within the cutoff at the build */
} tdmd_observables;
int tdmd_sched_observables(tdmd_sched* s, tdmd_observables* out);
/* Timing window (no reference counterpart; measurement aid): runs `steps`
 * steps without rebuild tests, with HIP events around each pass of the
 * steady-state step, and returns the mean durations: ms_out[0] = EAM
 * density pass (Morse: kick-drift), ms_out[1] = force pass (EAM NVE: with
 * the fused kick-drift epilogue; EAM NVT: with the closing half kick and the
 * KE partials, the pass of the thermostatted step). Advances the state like
 * run_until. NVT is accepted for EAM schedulers only. */
int tdmd_sched_time_passes(tdmd_sched* s, int32_t steps, float* ms_out);
/* In-window pass timing (no reference counterpart; measurement aid): with
 * `on`, the density and force passes of every step (inside the captured
 * hipGraphs too) time themselves on the device wall clock: each workgroup
 * stamps its start, the last one to finish adds the launch's span (first
 * workgroup start to last workgroup finish) to a running sum. Each call
 * with `on` resets the sums. tdmd_sched_pass_times returns the mean span
 * per launch since then: ms_out[0] = EAM density pass (Morse: 0),
 * ms_out[1] = force pass. Graphs captured with timing on are kept apart
 * from those without. (HIP events recorded inside a captured graph are not
 * timing events on this runtime, hence the device clock.) */
int tdmd_sched_pass_timing(tdmd_sched* s, int32_t on);
int tdmd_sched_pass_times(tdmd_sched* s, float* ms_out);
/* Autotuner (no reference counterpart): times `iters` force-path passes
 * (EAM density + force, Morse force; force-only epilogue) for each
 * candidate kernel shape on the uploaded state -- lanes per neighbour row
 * {1, 4, 8}, and with one lane and single-element EAM tables the LDS
 * passes with {8, 4} neighbour slots per pipelined group (x LDS spline
 * window {device limit, 79 KiB} where no halo is staged) -- and keeps the
 * fastest (the starting shape when within 1 % of it). The passes are then
 * run once more with the chosen shape, so the state holds that shape's
 * forces: a choice equal to the starting shape leaves the trajectory
 * bitwise unchanged, another lane count changes the order of the row sums
 * only. Returns out[0] = pair_unroll, out[1] = lds_budget bytes (0 off the
 * LDS passes); tdmd_sched_kernel_shape reports the lanes. Runs the initial
 * force pass if needed. */
int tdmd_sched_autotune(tdmd_sched* s, int32_t iters, int32_t out[2]);
/* The current kernel shape (no reference counterpart): out[0] = lanes per
 * neighbour row, out[1] = pair_unroll, out[2] = lds_budget. */
int tdmd_sched_kernel_shape(const tdmd_sched* s, int32_t out[3]);
/* Halo staging (no reference counterpart; diagnostics): out[0] = blocks of
 * 1024 atoms, out[1] = blocks whose halo is staged in LDS, out[2] = largest
 * halo (atoms), out[3] = halo atoms summed over all blocks. All zero when
 * the passes do not stage halos. Synchronises the scheduler's stream. */
int tdmd_sched_halo_info(tdmd_sched* s, int64_t out[4]);
/* Stream the scheduler runs on (for event timing). */
int tdmd_sched_stream(const tdmd_sched* s, void** stream);

/* ------------------------------------------------------------------ */
/* Force matching (no reference counterpart: the differentiable path  */
/* of the fitting harness). A fixed set of nsnap configurations of    */
/* natoms atoms in one box; L = sum |F - T|^2 over all atoms and its  */
/* gradient by a hand-written adjoint on the same pair arithmetic as  */
/* the force kernels. One-element potentials.                          */
/* ------------------------------------------------------------------ */
typedef struct tdmd_fm tdmd_fm;

/* pos: host double[nsnap][natoms][3]. Builds each snapshot's full list
 * with cutoff r_cut once (allocates). */
int tdmd_fm_create(tdmd_fm** out, int32_t precision, const tdmd_box* box,
                   int32_t natoms, int32_t nsnap, const double* pos,
                   double r_cut);
int tdmd_fm_destroy(tdmd_fm* fm);
/* Total atoms and listed (ordered) pairs. */
int tdmd_fm_info(const tdmd_fm* fm, int32_t* ntot, int64_t* npairs);
/* Target forces, host double[nsnap][natoms][3]. */
int tdmd_fm_set_targets(tdmd_fm* fm, const double* forces);
/* Morse (D, alpha, r0) with cutoff rc <= r_cut; host outputs, any may be
 * NULL: loss, grad[3] = dL/d(D, alpha, r0), forces_out[3 ntot]. */
int tdmd_fm_morse(tdmd_fm* fm, const double params[3], double rc,
                  double* loss, double grad[3], double* forces_out);
/* Single-element EAM from knot values on the device, d_values =
 * [F(nrho) | rho(nr) | phi(nr)] (phi = (r phi)/r, EamAlloy's form).
 * Device outputs, any may be NULL: d_loss[1], d_grad (same layout as
 * d_values), d_forces[3 ntot]. Synchronises `stream` before returning. */
int tdmd_fm_eam(tdmd_fm* fm, int32_t nrho, double drho, int32_t nr,
                double dr, double cutoff, const double* d_values,
                double* d_loss, double* d_grad, double* d_forces,
                void* stream);
/* Coefficient gradient of the last tdmd_fm_eam call (host copy):
 * F{a,b,c,d}[nrho-1], rho{a,b,c,d}[nr-1], phi{a,b,c,d}[nr-1]. */
int tdmd_fm_eam_coef_grad(const tdmd_fm* fm, double* gcoef);

/* ------------------------------------------------------------------ */
/* Zone pipeline: PipelineScheduler (src/scheduler/                    */
/* pipeline_scheduler.cuh:26-133, pipeline_scheduler.cu:15-309) with   */
/* ZonePartition (src/domain/zone_partition.hpp, .cpp:15-143) and the  */
/* zone state machine (src/scheduler/zone.hpp:17-64). Morse only, as   */
/* the reference. Slabs along x; atoms sorted by slab at upload; zone   */
/* steps issued on a pool of n_streams streams (one if deterministic). */
/* Host arrays: positions / velocities double[3N], forces and masses   */
/* in `real` of `precision` (float for TDMD_MIXED). download returns   */
/* the zone order with the ids, as the reference does. Allocates.      */
/* ------------------------------------------------------------------ */
typedef struct tdmd_pipe tdmd_pipe;

/* PipelineConfig (pipeline_scheduler.cuh:26-45). */
typedef struct {
  double dt;
  double r_skin;
  int32_t n_zones;        /* 0: max(3, floor(Lx / rc)) */
  int32_t n_streams;
  int32_t rebuild_every;
  int32_t deterministic;  /* one stream, each zone step waited for */
  double t_target;        /* > 0: NVT (Nose-Hoover chain) */
  double t_period;
  int32_t nhc_length;
  int32_t adaptive_dt;    /* dt = clamp(c2 rc / vmax, dt_min, dt_max) */
  double dt_max;
  double dt_min;
  double c2;
  int32_t precision;      /* tdmd_precision */
  int32_t reserved[7];
} tdmd_pipe_config;

/* Zone (zone.hpp:43-64): state is the ZoneState code (0 Free .. 6
 * Sending); lo/hi the slab's bounding box. */
typedef struct {
  int32_t id;
  int32_t atom_offset;
  int32_t natoms;
  int32_t time_step;
  int32_t state;
  int32_t lattice_index[3];
  double lo[3];
  double hi[3];
} tdmd_zone_info;

/* PipelineScheduler::PipelineScheduler (pipeline_scheduler.cu:15-38). */
int tdmd_pipe_create(tdmd_pipe** out, const tdmd_box* box, int32_t natoms,
                     const tdmd_morse_params* params,
                     const tdmd_pipe_config* cfg);
int tdmd_pipe_destroy(tdmd_pipe* p);
/* upload (:40-82): types are 1-based indices into masses[n_masses]. */
int tdmd_pipe_upload(tdmd_pipe* p, const double* pos, const double* vel,
                     const void* forces, const int32_t* types,
                     const int32_t* ids, const void* masses,
                     int32_t n_masses);
/* run_until (:196-268): every zone at time_step >= target_step. */
int tdmd_pipe_run_until(tdmd_pipe* p, int32_t target_step);
/* download (:279-287), zone order. */
int tdmd_pipe_download(const tdmd_pipe* p, double* pos, double* vel,
                       void* forces, int32_t* types, int32_t* ids,
                       int32_t natoms);
/* partition() (:74-76): zone count, the zones, and a zone's neighbour
 * zones (ZonePartition::zone_neighbors; *n gets the full count, at most
 * cap are written). */
int tdmd_pipe_zone_count(const tdmd_pipe* p, int32_t* n);
int tdmd_pipe_zones(const tdmd_pipe* p, tdmd_zone_info* out, int32_t n);
int tdmd_pipe_zone_neighbors(const tdmd_pipe* p, int32_t zone, int32_t* out,
                             int32_t cap, int32_t* n);
/* min_time_step (:270-277); stats (PipelineStats, pipeline_scheduler.cuh:
 * 48-53): out[0] kernel_launches (zone steps issued), [1] dep_check_calls,
 * [2] dep_check_failures, [3] ticks; the current (adaptive) dt. */
int tdmd_pipe_min_time_step(const tdmd_pipe* p, int32_t* out);
int tdmd_pipe_stats(const tdmd_pipe* p, int64_t out[4]);
int tdmd_pipe_current_dt(const tdmd_pipe* p, double* dt);

#ifdef __cplusplus
}
#endif

#endif /* TDMD_HIP_H */
