// pipeline_scheduler.hpp -- PipelineScheduler (src/scheduler/
// pipeline_scheduler.cuh:26-133): the zone pipeline on one GPU, over
// tdmd_pipe_* (the scheduling, streams and kernels live in the library).
//
// partition() mirrors the library's zones: their ranges after upload, their
// time steps and states after each run_until. download() returns the zone
// order with the ids, as the reference does.
#pragma once

#include <vector>

#include "../core/hip_abi.hpp"
#include "../domain/zone_partition.hpp"
#include "../potentials/device_morse.hpp"

namespace tdmd::scheduler {

struct PipelineConfig {
  real dt{real{0.001}};
  real r_skin{real{1.0}};
  i32 n_zones{0};  // auto
  i32 n_streams{4};
  i32 rebuild_every{10};
  bool deterministic{false};  // one stream, zones in index order
  real t_target{0};           // > 0: NVT (Nose-Hoover chain)
  real t_period{real{0.1}};
  i32 nhc_length{3};
  bool adaptive_dt{false};  // dt = clamp(c2 rc / v_max, dt_min, dt_max)
  real dt_max{real{0.002}};
  real dt_min{real{0.0001}};
  real c2{real{0.05}};
};

struct PipelineStats {
  i64 kernel_launches{0};
  i64 dep_check_calls{0};
  i64 dep_check_failures{0};
  i64 ticks{0};
};

class PipelineScheduler {
 public:
  PipelineScheduler(const Box& box, i32 natoms, const potentials::MorseParams& params,
                    const PipelineConfig& cfg)
      : natoms_(natoms) {
    const tdmd_box b = hip::to_c(box);
    const tdmd_morse_params p = potentials::to_c(params);
    tdmd_pipe_config c{};
    c.dt = static_cast<double>(cfg.dt);
    c.r_skin = static_cast<double>(cfg.r_skin);
    c.n_zones = cfg.n_zones;
    c.n_streams = cfg.n_streams;
    c.rebuild_every = cfg.rebuild_every;
    c.deterministic = cfg.deterministic ? 1 : 0;
    c.t_target = static_cast<double>(cfg.t_target);
    c.t_period = static_cast<double>(cfg.t_period);
    c.nhc_length = cfg.nhc_length;
    c.adaptive_dt = cfg.adaptive_dt ? 1 : 0;
    c.dt_max = static_cast<double>(cfg.dt_max);
    c.dt_min = static_cast<double>(cfg.dt_min);
    c.c2 = static_cast<double>(cfg.c2);
    c.precision = hip::kPrecision;
    TDMD_HIP_CALL(tdmd_pipe_create(&p_, &b, natoms, &p, &c));
    partition_.build(box, params.rc, cfg.n_zones);
    partition_.build_zone_neighbors(params.rc + cfg.r_skin);
  }
  ~PipelineScheduler() {
    if (p_) (void)tdmd_pipe_destroy(p_);
  }
  PipelineScheduler(const PipelineScheduler&) = delete;
  PipelineScheduler& operator=(const PipelineScheduler&) = delete;

  void upload(const PositionVec* positions, const VelocityVec* velocities,
              const ForceVec* forces, const i32* types, const i32* ids, const real* masses,
              i32 n_masses) {
    TDMD_HIP_CALL(tdmd_pipe_upload(p_, reinterpret_cast<const double*>(positions),
                                   reinterpret_cast<const double*>(velocities), forces, types,
                                   ids, masses, n_masses));
    refresh();
  }

  void run_until(i32 target_step) {
    TDMD_HIP_CALL(tdmd_pipe_run_until(p_, target_step));
    refresh();
  }

  void download(PositionVec* positions, VelocityVec* velocities, ForceVec* forces, i32* types,
                i32* ids, i32 natoms) const {
    TDMD_HIP_CALL(tdmd_pipe_download(p_, reinterpret_cast<double*>(positions),
                                     reinterpret_cast<double*>(velocities), forces, types, ids,
                                     natoms));
  }

  const domain::ZonePartition& partition() const noexcept { return partition_; }
  const PipelineStats& stats() const noexcept { return stats_; }
  i32 min_time_step() const noexcept {
    i32 m = 0;
    bool first = true;
    for (const auto& z : partition_.zones()) {
      m = first ? z.time_step : (z.time_step < m ? z.time_step : m);
      first = false;
    }
    return m;
  }
  // The time step in use (changes with adaptive_dt).
  double current_dt() const {
    double dt = 0;
    TDMD_HIP_CALL(tdmd_pipe_current_dt(p_, &dt));
    return dt;
  }

 private:
  void refresh() {
    std::vector<tdmd_zone_info> info(static_cast<std::size_t>(partition_.n_zones()));
    TDMD_HIP_CALL(tdmd_pipe_zones(p_, info.data(), partition_.n_zones()));
    auto& zones = partition_.zones();
    for (std::size_t k = 0; k < info.size(); ++k) {
      zones[k].atom_offset = info[k].atom_offset;
      zones[k].natoms_in_zone = info[k].natoms;
      zones[k].time_step = info[k].time_step;
      zones[k].state = static_cast<ZoneState>(info[k].state);
    }
    int64_t s[4];
    TDMD_HIP_CALL(tdmd_pipe_stats(p_, s));
    stats_ = PipelineStats{s[0], s[1], s[2], s[3]};
  }

  tdmd_pipe* p_{nullptr};
  i32 natoms_;
  domain::ZonePartition partition_;
  PipelineStats stats_;
};

}  // namespace tdmd::scheduler
