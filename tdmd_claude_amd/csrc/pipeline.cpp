// pipeline.cpp -- the zone pipeline (PipelineScheduler,
// src/scheduler/pipeline_scheduler.cuh:26-133, pipeline_scheduler.cu:15-309)
// on one GPU.
//
// The box is cut into slabs along x (ZonePartition,
// src/domain/zone_partition.cpp:15-140); atoms are sorted by slab once, at
// upload, and each zone owns a contiguous index range from then on. A zone
// step is the velocity-Verlet step restricted to the zone's atoms (half
// kick, drift, zero forces, the zone's forces over the whole position array,
// half kick), issued on one stream of a small pool. A zone may start a step
// when every zone within r_cut + skin has finished at least as many steps
// and none of them is running (pipeline_scheduler.cu:98-126). With
// `deterministic` the pool has one stream and every zone step is waited
// for, so zones advance in index order, one sweep per step, each zone's
// forces seeing its lower neighbours already moved and its upper ones not
// yet: the trajectory is a fixed function of the input. Otherwise the
// order follows the device's completion order.
//
// Kept as the reference has it, because results depend on it:
//   * forces of all atoms are recomputed from the current positions at the
//     start of every run_until (pipeline_scheduler.cu:200-206);
//   * the list is built once after upload; the periodic rebuild test only
//     drains the pipeline, since the rebuild flag is raised by upload alone
//     (pipeline_scheduler.cu:85-90,210-216);
//   * NVT applies the chain half-steps around every tick (a tick is a sweep
//     of the zones that may start), not around every step
//     (pipeline_scheduler.cu:230-261);
//   * host arithmetic in `real` (float for TDMD_MIXED): zone width, list
//     radius, time step, chain scale factors.
//
// The kernels are the library's own: cell list + CSR list
// (tdmd_cell_build, tdmd_nlist_count/fill), tdmd_morse_force(_zone), the
// zone velocity-Verlet functions, tdmd_compute_ke / tdmd_compute_vmax /
// tdmd_scale_velocities and the host chain step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "common.hpp"

using namespace tdmd_hip;

namespace {

// Zone states of the zone state machine (src/scheduler/zone.hpp:17-25).
enum ZoneStateCode : int32_t { kFree = 0, kReady = 3, kComputing = 4, kDone = 5 };

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t n) {
    if (n <= bytes && p) return TDMD_OK;
    if (p) {
      const int rc = tdmd_free(p);
      p = nullptr;
      bytes = 0;
      if (rc) return rc;
    }
    const int rc = tdmd_malloc(&p, std::max<size_t>(n, 16));
    if (rc) return rc;
    bytes = std::max<size_t>(n, 16);
    return TDMD_OK;
  }
  void release() {
    if (p) (void)tdmd_free(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

}  // namespace

struct tdmd_pipe {
  tdmd_box box{};
  int natoms = 0;
  tdmd_morse_params mp{};
  tdmd_pipe_config cfg{};
  bool mixed = false;
  // zones
  int nz = 0;
  double zone_width = 0, lo_x = 0;
  std::vector<tdmd_zone_info> zones;
  std::vector<std::vector<int32_t>> zone_nbrs;
  // stream pool (src/scheduler/stream_pool.cu:7-63)
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<char> busy;
  std::vector<int> zone_stream;
  hipStream_t main = nullptr;
  // state in zone order: positions / velocities Vec3D, forces ForceVec
  DevBuf pos, vel, frc, types, ids, masses;
  int n_masses = 0;
  // neighbour list (DeviceNeighborList)
  DevBuf ccount, coff, catoms, acells, cwork, counts, offsets, nbrs, meta, ovf, swork;
  DevBuf scalar, rwork;
  int32_t ncells[3] = {0, 0, 0};
  bool uploaded = false;
  bool needs_rebuild = true;
  int steps_since = 0;
  double dt = 0;
  // chain (NoseHooverChain, host state)
  int nhc_m = 0, n_dof = 0;
  std::vector<double> eta, eta_v, eta_m;
  struct {
    int64_t kernel_launches = 0, dep_check_calls = 0, dep_check_failures = 0, ticks = 0;
  } stats;

  ~tdmd_pipe() {
    for (auto* b : {&pos, &vel, &frc, &types, &ids, &masses, &ccount, &coff, &catoms, &acells,
                    &cwork, &counts, &offsets, &nbrs, &meta, &ovf, &swork, &scalar, &rwork})
      b->release();
    for (auto e : events)
      if (e) (void)hipEventDestroy(e);
    for (auto s : streams)
      if (s) (void)hipStreamDestroy(s);
    if (main) (void)hipStreamDestroy(main);
  }

  // A value the reference holds in `real`: float under TDMD_MIXED. One
  // operation in double rounded to float equals the float operation.
  double rl(double x) const { return mixed ? static_cast<double>(static_cast<float>(x)) : x; }
  size_t real_bytes() const { return mixed ? 4 : 8; }
  int32_t prec() const { return mixed ? TDMD_MIXED : TDMD_FP64; }

  // ZonePartition::zone_of (zone_partition.cpp:44-49); x arrives as `real`
  int zone_of(double x) const {
    int z = static_cast<int>(rl(rl(rl(x) - lo_x) / zone_width));
    if (z < 0) z = 0;
    if (z >= nz) z = nz - 1;
    return z;
  }

  int min_time_step() const {
    int m = std::numeric_limits<int>::max();
    for (const auto& z : zones) m = std::min(m, static_cast<int>(z.time_step));
    return m;
  }

  // DeviceNeighborList::build over all atoms (device_neighbor_list.cu:158-278)
  int rebuild_nlist() {
    if (!needs_rebuild) return TDMD_OK;
    const double r_list = rl(mp.rc + cfg.r_skin);
    double cs[3];
    int rc;
    if ((rc = tdmd_cell_grid(&box, r_list, ncells, cs))) return rc;
    const int nc = ncells[0] * ncells[1] * ncells[2];
    const size_t na = static_cast<size_t>(std::max(natoms, 1));
    size_t wb = 0;
    if ((rc = tdmd_cell_workspace_bytes(natoms, nc, &wb))) return rc;
    if ((rc = ccount.ensure(4 * static_cast<size_t>(nc))) || (rc = coff.ensure(4 * static_cast<size_t>(nc))) ||
        (rc = catoms.ensure(4 * na)) || (rc = acells.ensure(4 * na)) || (rc = cwork.ensure(wb)) ||
        (rc = counts.ensure(4 * na)) || (rc = offsets.ensure(4 * na)) || (rc = meta.ensure(8)) ||
        (rc = ovf.ensure(4)))
      return rc;
    if ((rc = tdmd_cell_build(pos.as<double>(), natoms, &box, r_list, ccount.as<int32_t>(),
                              coff.as<int32_t>(), catoms.as<int32_t>(), acells.as<int32_t>(),
                              cwork.p, main)))
      return rc;
    if ((rc = tdmd_nlist_count(pos.as<double>(), natoms, &box, r_list, ccount.as<int32_t>(),
                               coff.as<int32_t>(), catoms.as<int32_t>(), acells.as<int32_t>(),
                               ncells, counts.as<int32_t>(), main)))
      return rc;
    if ((rc = tdmd_scan_workspace_bytes(natoms, &wb)) || (rc = swork.ensure(wb))) return rc;
    if ((rc = tdmd_scan_exclusive_i32(counts.as<int32_t>(), offsets.as<int32_t>(), natoms,
                                      meta.as<int32_t>(), swork.p, main)))
      return rc;
    int32_t m[2] = {0, 0};
    if ((rc = tdmd_memcpy_dtoh(m, meta.p, sizeof(m), main))) return rc;
    TDMD_HIP_TRY(hipStreamSynchronize(main));
    if ((rc = nbrs.ensure(4 * static_cast<size_t>(std::max(m[0], 1))))) return rc;
    if ((rc = tdmd_fill_u32(ovf.p, 0u, 1, main))) return rc;
    if ((rc = tdmd_nlist_fill(pos.as<double>(), natoms, &box, r_list, ccount.as<int32_t>(),
                              coff.as<int32_t>(), catoms.as<int32_t>(), acells.as<int32_t>(),
                              ncells, offsets.as<int32_t>(), nbrs.as<int32_t>(),
                              counts.as<int32_t>(), m[1], ovf.as<int32_t>(), main)))
      return rc;
    TDMD_HIP_TRY(hipStreamSynchronize(main));
    needs_rebuild = false;
    steps_since = 0;
    return TDMD_OK;
  }

  // PipelineScheduler::check_deps (pipeline_scheduler.cu:98-126)
  bool check_deps(int z) {
    ++stats.dep_check_calls;
    const int target = zones[static_cast<size_t>(z)].time_step;
    for (int32_t n : zone_nbrs[static_cast<size_t>(z)]) {
      if (n == z) continue;
      const auto& o = zones[static_cast<size_t>(n)];
      if (o.time_step < target || o.state == kComputing) {
        ++stats.dep_check_failures;
        return false;
      }
    }
    return true;
  }

  // StreamPool::try_acquire (stream_pool.cu:26-42)
  int try_acquire() {
    for (size_t k = 0; k < streams.size(); ++k)
      if (busy[k] && hipEventQuery(events[k]) == hipSuccess) busy[k] = 0;
    for (size_t k = 0; k < streams.size(); ++k)
      if (!busy[k]) {
        busy[k] = 1;
        return static_cast<int>(k);
      }
    return -1;
  }

  // PipelineScheduler::launch_zone_step (pipeline_scheduler.cu:128-168)
  int launch_zone_step(int z, int sid) {
    auto& zn = zones[static_cast<size_t>(z)];
    hipStream_t s = streams[static_cast<size_t>(sid)];
    const int32_t first = zn.atom_offset, count = zn.natoms;
    const int32_t p = prec();
    int rc;
    if ((rc = tdmd_half_kick_zone(p, vel.as<double>(), frc.p, types.as<int32_t>(), masses.p, first,
                                  count, dt, s)) ||
        (rc = tdmd_drift_zone(pos.as<double>(), vel.as<double>(), first, count, dt, &box, s)) ||
        (rc = tdmd_zero_forces_zone(p, frc.p, first, count, s)) ||
        (rc = tdmd_morse_force_zone(p, pos.as<double>(), frc.p, nbrs.as<int32_t>(),
                                    offsets.as<int32_t>(), counts.as<int32_t>(), first, count, &box,
                                    &mp, nullptr, nullptr, s)) ||
        (rc = tdmd_half_kick_zone(p, vel.as<double>(), frc.p, types.as<int32_t>(), masses.p, first,
                                  count, dt, s)))
      return rc;
    TDMD_HIP_TRY(hipEventRecord(events[static_cast<size_t>(sid)], s));
    busy[static_cast<size_t>(sid)] = 1;
    zone_stream[static_cast<size_t>(z)] = sid;
    zn.state = kComputing;  // Ready -> Computing
    ++stats.kernel_launches;
    return TDMD_OK;
  }

  // PipelineScheduler::poll_completions (pipeline_scheduler.cu:170-189)
  void poll_completions() {
    for (int z = 0; z < nz; ++z) {
      auto& zn = zones[static_cast<size_t>(z)];
      if (zn.state != kComputing) continue;
      const int sid = zone_stream[static_cast<size_t>(z)];
      if (sid >= 0 && hipEventQuery(events[static_cast<size_t>(sid)]) == hipSuccess) {
        ++zn.time_step;  // Computing -> Done
        busy[static_cast<size_t>(sid)] = 0;
        zone_stream[static_cast<size_t>(z)] = -1;
        zn.state = kReady;  // Done -> Ready on one rank
      }
    }
  }

  // PipelineScheduler::tick (pipeline_scheduler.cu:191-224)
  int tick(bool* progress) {
    ++stats.ticks;
    poll_completions();
    *progress = false;
    for (int z = 0; z < nz; ++z) {
      if (zones[static_cast<size_t>(z)].state != kReady) continue;
      if (!check_deps(z)) continue;
      const int sid = try_acquire();
      if (sid < 0) break;
      const int rc = launch_zone_step(z, sid);
      if (rc) return rc;
      *progress = true;
      if (cfg.deterministic) {
        TDMD_HIP_TRY(hipStreamSynchronize(streams[static_cast<size_t>(sid)]));
        poll_completions();
      }
    }
    return TDMD_OK;
  }

  int drain() {
    TDMD_HIP_TRY(hipDeviceSynchronize());
    poll_completions();
    return TDMD_OK;
  }

  int ke_host(double* ke) {
    int rc;
    size_t wb = 0;
    if ((rc = tdmd_reduce_workspace_bytes(natoms, &wb)) || (rc = rwork.ensure(wb)) ||
        (rc = scalar.ensure(8)))
      return rc;
    if ((rc = tdmd_compute_ke(prec(), vel.as<double>(), types.as<int32_t>(), masses.p, natoms,
                              scalar.as<double>(), rwork.p, main)))
      return rc;
    if ((rc = tdmd_memcpy_dtoh(ke, scalar.p, 8, main))) return rc;
    TDMD_HIP_TRY(hipStreamSynchronize(main));
    return TDMD_OK;
  }

  int nhc_scale() {
    double ke = 0;
    int rc;
    if ((rc = ke_host(&ke))) return rc;
    const double s = rl(tdmd_nhc_half_step(nhc_m, eta.data(), eta_v.data(), eta_m.data(),
                                           cfg.t_target, cfg.t_period, cfg.dt, n_dof, ke));
    if ((rc = tdmd_scale_velocities(vel.as<double>(), natoms, s, main))) return rc;
    TDMD_HIP_TRY(hipDeviceSynchronize());
    return TDMD_OK;
  }

  // PipelineScheduler::run_until (pipeline_scheduler.cu:196-268)
  int run_until(int target) {
    if (!uploaded) return fail(TDMD_E_STATE, "pipeline: run_until before upload");
    int rc;
    if ((rc = rebuild_nlist())) return rc;
    // step-0 forces of all atoms from the current positions
    if ((rc = tdmd_zero_forces(prec(), frc.p, natoms, main)) ||
        (rc = tdmd_morse_force(prec(), pos.as<double>(), frc.p, nbrs.as<int32_t>(),
                               offsets.as<int32_t>(), counts.as<int32_t>(), natoms, &box, &mp,
                               nullptr, nullptr, main)))
      return rc;
    TDMD_HIP_TRY(hipDeviceSynchronize());
    while (min_time_step() < target) {
      ++steps_since;
      if (steps_since >= cfg.rebuild_every) {
        if ((rc = drain()) || (rc = rebuild_nlist())) return rc;
      }
      if (cfg.adaptive_dt) {
        if ((rc = drain())) return rc;
        size_t wb = 0;
        if ((rc = tdmd_reduce_workspace_bytes(natoms, &wb)) || (rc = rwork.ensure(wb)) ||
            (rc = scalar.ensure(8)))
          return rc;
        if ((rc = tdmd_compute_vmax(vel.as<double>(), natoms, scalar.as<double>(), rwork.p, main)))
          return rc;
        double vmax = 0;
        if ((rc = tdmd_memcpy_dtoh(&vmax, scalar.p, 8, main))) return rc;
        TDMD_HIP_TRY(hipStreamSynchronize(main));
        vmax = rl(vmax);
        if (vmax > 0) {
          dt = std::min(rl(cfg.dt_max), rl(rl(rl(cfg.c2) * rl(mp.rc)) / vmax));
          dt = std::max(dt, rl(cfg.dt_min));
        }
      }
      bool progress = false;
      if (nhc_m > 0) {
        // chain half-step, one tick, chain half-step (pipelined NVT is not
        // attempted: pipeline_scheduler.cu:230-261)
        if ((rc = drain()) || (rc = nhc_scale()) || (rc = tick(&progress))) return rc;
        if (!progress && (rc = drain())) return rc;
        if ((rc = drain()) || (rc = nhc_scale())) return rc;
      } else {
        if ((rc = tick(&progress))) return rc;
        if (!progress && (rc = drain())) return rc;
      }
    }
    return drain();
  }
};

extern "C" {

int tdmd_pipe_create(tdmd_pipe** out, const tdmd_box* box, int32_t natoms,
                     const tdmd_morse_params* params, const tdmd_pipe_config* cfg) {
  if (!out || !box || !params || !cfg) return fail(TDMD_E_INVALID, "pipe_create: null argument");
  *out = nullptr;
  if (natoms <= 0) return fail(TDMD_E_INVALID, "pipe_create: natoms must be positive");
  if (cfg->precision != TDMD_FP64 && cfg->precision != TDMD_MIXED)
    return fail(TDMD_E_INVALID, "pipe_create: bad precision");
  if (!(cfg->dt > 0) || !(cfg->r_skin >= 0) || cfg->rebuild_every < 1 || cfg->n_streams < 1 ||
      cfg->n_streams > 64)
    return fail(TDMD_E_INVALID, "pipe_create: bad dt / r_skin / rebuild_every / n_streams");
  if (!(params->rc > 0)) return fail(TDMD_E_INVALID, "pipe_create: rc must be positive");
  if (cfg->t_target > 0 && (cfg->nhc_length < 1 || cfg->nhc_length > 8 || !(cfg->t_period > 0)))
    return fail(TDMD_E_INVALID, "pipe_create: NHC chain length must be 1..8, t_period > 0");
  if (cfg->adaptive_dt && !(cfg->dt_min > 0 && cfg->dt_max >= cfg->dt_min && cfg->c2 > 0))
    return fail(TDMD_E_INVALID, "pipe_create: adaptive dt needs 0 < dt_min <= dt_max, c2 > 0");
  auto* P = new (std::nothrow) tdmd_pipe();
  if (!P) return fail(TDMD_E_NOMEM, "pipe_create: out of host memory");
  P->box = *box;
  P->natoms = natoms;
  P->mp = *params;
  P->cfg = *cfg;
  P->mixed = cfg->precision == TDMD_MIXED;
  P->dt = P->rl(cfg->dt);
  // ZonePartition::build (zone_partition.cpp:15-42): slabs along x
  const double lx = box->hi[0] - box->lo[0];
  int nzr = cfg->n_zones;
  if (nzr <= 0) nzr = std::max(3, static_cast<int>(std::floor(lx / P->rl(params->rc))));
  if (nzr < 3) {
    delete P;
    return fail(TDMD_E_INVALID, "pipe_create: need at least 3 zones for PBC");
  }
  P->nz = nzr;
  P->lo_x = P->rl(box->lo[0]);
  P->zone_width = P->rl(lx / nzr);
  P->zones.resize(static_cast<size_t>(nzr));
  for (int i = 0; i < nzr; ++i) {
    tdmd_zone_info& z = P->zones[static_cast<size_t>(i)];
    z = tdmd_zone_info{};
    z.id = i;
    z.lattice_index[0] = i;
    z.lo[0] = box->lo[0] + static_cast<double>(i) * (lx / nzr);
    z.hi[0] = box->lo[0] + static_cast<double>(i + 1) * (lx / nzr);
    for (int d = 1; d < 3; ++d) {
      z.lo[d] = box->lo[d];
      z.hi[d] = box->hi[d];
    }
    z.state = kFree;
  }
  // build_zone_neighbors (zone_partition.cpp:123-143): zones within
  // ceil(r_list / width) slabs, periodic, sorted, unique
  const double r_list = P->rl(params->rc + cfg->r_skin);
  const int span = static_cast<int>(std::ceil(P->rl(r_list / P->zone_width)));
  P->zone_nbrs.resize(static_cast<size_t>(nzr));
  for (int z = 0; z < nzr; ++z) {
    auto& v = P->zone_nbrs[static_cast<size_t>(z)];
    for (int d = -span; d <= span; ++d) {
      int n = z + d;
      if (n < 0)
        n += nzr;
      else if (n >= nzr)
        n -= nzr;
      v.push_back(n);
    }
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  const int ns = cfg->deterministic ? 1 : cfg->n_streams;
  P->streams.assign(static_cast<size_t>(ns), nullptr);
  P->events.assign(static_cast<size_t>(ns), nullptr);
  P->busy.assign(static_cast<size_t>(ns), 0);
  P->zone_stream.assign(static_cast<size_t>(nzr), -1);
  for (int k = 0; k < ns; ++k) {
    hipError_t e = hipStreamCreate(&P->streams[static_cast<size_t>(k)]);
    if (e == hipSuccess)
      e = hipEventCreateWithFlags(&P->events[static_cast<size_t>(k)], hipEventDisableTiming);
    if (e != hipSuccess) {
      delete P;
      return hip_fail(e, "pipe_create: stream pool", __FILE__, __LINE__);
    }
  }
  const hipError_t e = hipStreamCreate(&P->main);
  if (e != hipSuccess) {
    delete P;
    return hip_fail(e, "pipe_create: stream", __FILE__, __LINE__);
  }
  if (cfg->t_target > 0) {
    P->nhc_m = cfg->nhc_length;
    P->n_dof = 3 * natoms - 3;
    P->eta.assign(8, 0.0);
    P->eta_v.assign(8, 0.0);
    P->eta_m.assign(8, 0.0);
  }
  *out = P;
  return TDMD_OK;
}

int tdmd_pipe_destroy(tdmd_pipe* P) {
  if (!P) return TDMD_OK;
  (void)hipDeviceSynchronize();
  delete P;
  return TDMD_OK;
}

int tdmd_pipe_upload(tdmd_pipe* P, const double* pos, const double* vel, const void* forces,
                     const int32_t* types, const int32_t* ids, const void* masses,
                     int32_t n_masses) {
  if (!P || !pos || !vel || !forces || !types || !ids || !masses || n_masses < 1)
    return fail(TDMD_E_INVALID, "pipe_upload: null argument or no masses");
  const size_t n = static_cast<size_t>(P->natoms);
  for (size_t i = 0; i < n; ++i)
    if (types[i] < 1 || types[i] >= n_masses)
      return fail(TDMD_E_INVALID, "pipe_upload: type outside the mass table");
  // ZonePartition::assign_atoms (zone_partition.cpp:51-121): stable
  // counting sort by zone
  std::vector<int32_t> zone(n), count(static_cast<size_t>(P->nz), 0);
  for (size_t i = 0; i < n; ++i) {
    zone[i] = P->zone_of(pos[3 * i]);
    ++count[static_cast<size_t>(zone[i])];
  }
  std::vector<int32_t> offset(static_cast<size_t>(P->nz), 0), placed(static_cast<size_t>(P->nz), 0);
  for (int z = 1; z < P->nz; ++z)
    offset[static_cast<size_t>(z)] = offset[static_cast<size_t>(z - 1)] + count[static_cast<size_t>(z - 1)];
  std::vector<int32_t> perm(n);
  for (size_t i = 0; i < n; ++i) {
    const auto z = static_cast<size_t>(zone[i]);
    perm[static_cast<size_t>(offset[z] + placed[z]++)] = static_cast<int32_t>(i);
  }
  const size_t rb = P->real_bytes();
  std::vector<double> hp(3 * n), hv(3 * n);
  std::vector<unsigned char> hf(3 * n * rb);
  std::vector<int32_t> ht(n), hid(n);
  const auto* fb = static_cast<const unsigned char*>(forces);
  for (size_t k = 0; k < n; ++k) {
    const auto o = static_cast<size_t>(perm[k]);
    for (int d = 0; d < 3; ++d) {
      hp[3 * k + d] = pos[3 * o + d];
      hv[3 * k + d] = vel[3 * o + d];
    }
    std::copy(fb + 3 * o * rb, fb + 3 * (o + 1) * rb, hf.begin() + static_cast<std::ptrdiff_t>(3 * k * rb));
    ht[k] = types[o];
    hid[k] = ids[o];
  }
  int rc;
  if ((rc = P->pos.ensure(24 * n)) || (rc = P->vel.ensure(24 * n)) ||
      (rc = P->frc.ensure(3 * rb * n)) || (rc = P->types.ensure(4 * n)) ||
      (rc = P->ids.ensure(4 * n)) || (rc = P->masses.ensure(rb * static_cast<size_t>(n_masses))))
    return rc;
  if ((rc = tdmd_memcpy_htod(P->pos.p, hp.data(), 24 * n, nullptr)) ||
      (rc = tdmd_memcpy_htod(P->vel.p, hv.data(), 24 * n, nullptr)) ||
      (rc = tdmd_memcpy_htod(P->frc.p, hf.data(), 3 * rb * n, nullptr)) ||
      (rc = tdmd_memcpy_htod(P->types.p, ht.data(), 4 * n, nullptr)) ||
      (rc = tdmd_memcpy_htod(P->ids.p, hid.data(), 4 * n, nullptr)) ||
      (rc = tdmd_memcpy_htod(P->masses.p, masses, rb * static_cast<size_t>(n_masses), nullptr)))
    return rc;
  P->n_masses = n_masses;
  for (int z = 0; z < P->nz; ++z) {
    tdmd_zone_info& zi = P->zones[static_cast<size_t>(z)];
    zi.atom_offset = offset[static_cast<size_t>(z)];
    zi.natoms = count[static_cast<size_t>(z)];
    zi.state = kReady;
    zi.time_step = 0;
  }
  P->uploaded = true;
  P->needs_rebuild = true;
  return TDMD_OK;
}

int tdmd_pipe_run_until(tdmd_pipe* P, int32_t target_step) {
  if (!P) return fail(TDMD_E_INVALID, "pipe_run_until: null scheduler");
  return P->run_until(target_step);
}

int tdmd_pipe_download(const tdmd_pipe* P, double* pos, double* vel, void* forces,
                       int32_t* types, int32_t* ids, int32_t natoms) {
  if (!P || !pos || !vel || !forces || !types || !ids)
    return fail(TDMD_E_INVALID, "pipe_download: null argument");
  if (!P->uploaded) return fail(TDMD_E_STATE, "pipe_download: nothing uploaded");
  if (natoms != P->natoms) return fail(TDMD_E_INVALID, "pipe_download: natoms mismatch");
  TDMD_HIP_TRY(hipDeviceSynchronize());
  const size_t n = static_cast<size_t>(natoms);
  int rc;
  if ((rc = tdmd_memcpy_dtoh(pos, P->pos.p, 24 * n, nullptr)) ||
      (rc = tdmd_memcpy_dtoh(vel, P->vel.p, 24 * n, nullptr)) ||
      (rc = tdmd_memcpy_dtoh(forces, P->frc.p, 3 * P->real_bytes() * n, nullptr)) ||
      (rc = tdmd_memcpy_dtoh(types, P->types.p, 4 * n, nullptr)) ||
      (rc = tdmd_memcpy_dtoh(ids, P->ids.p, 4 * n, nullptr)))
    return rc;
  return TDMD_OK;
}

int tdmd_pipe_zone_count(const tdmd_pipe* P, int32_t* n) {
  if (!P || !n) return fail(TDMD_E_INVALID, "pipe_zone_count: null argument");
  *n = P->nz;
  return TDMD_OK;
}

int tdmd_pipe_zones(const tdmd_pipe* P, tdmd_zone_info* out, int32_t n) {
  if (!P || !out || n < P->nz) return fail(TDMD_E_INVALID, "pipe_zones: null or short output");
  std::copy(P->zones.begin(), P->zones.end(), out);
  return TDMD_OK;
}

int tdmd_pipe_zone_neighbors(const tdmd_pipe* P, int32_t zone, int32_t* out, int32_t cap,
                             int32_t* n) {
  if (!P || !n || zone < 0 || zone >= P->nz)
    return fail(TDMD_E_INVALID, "pipe_zone_neighbors: bad zone");
  const auto& v = P->zone_nbrs[static_cast<size_t>(zone)];
  *n = static_cast<int32_t>(v.size());
  if (out) std::copy(v.begin(), v.begin() + std::min<size_t>(v.size(), static_cast<size_t>(std::max(cap, 0))), out);
  return TDMD_OK;
}

int tdmd_pipe_min_time_step(const tdmd_pipe* P, int32_t* out) {
  if (!P || !out) return fail(TDMD_E_INVALID, "pipe_min_time_step: null argument");
  *out = P->min_time_step();
  return TDMD_OK;
}

int tdmd_pipe_stats(const tdmd_pipe* P, int64_t out[4]) {
  if (!P || !out) return fail(TDMD_E_INVALID, "pipe_stats: null argument");
  out[0] = P->stats.kernel_launches;
  out[1] = P->stats.dep_check_calls;
  out[2] = P->stats.dep_check_failures;
  out[3] = P->stats.ticks;
  return TDMD_OK;
}

int tdmd_pipe_current_dt(const tdmd_pipe* P, double* dt) {
  if (!P || !dt) return fail(TDMD_E_INVALID, "pipe_current_dt: null argument");
  *dt = P->dt;
  return TDMD_OK;
}

}  // extern "C"
