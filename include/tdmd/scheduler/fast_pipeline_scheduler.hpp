// fast_pipeline_scheduler.hpp -- FastPipelineScheduler
// (src/scheduler/fast_pipeline_scheduler.cuh:36-192) over tdmd_sched_*.
//
// Same construction, upload / run_until / download (upload order) and
// telemetry as the reference. FastPipelineConfig keeps the reference's three
// fields first; the fields after them select precision-independent options
// of this implementation (rebuild policy, NVT, graphs) and default to the
// reference's behaviour: fixed-interval rebuild, NVE.
#pragma once

#include <vector>

#include "../core/hip_abi.hpp"
#include "../potentials/device_eam.hpp"
#include "../potentials/device_morse.hpp"

namespace tdmd::scheduler {

enum class PotentialKind {
  Morse,
  Eam,
};

enum class RebuildPolicy {
  FixedInterval,  // every rebuild_every steps (reference)
  Displacement,   // max displacement > skin / 2, decided on the device
};

struct FastPipelineConfig {
  real dt{real{0.001}};
  real r_skin{real{1.0}};
  i32 rebuild_every{10};
  // additions
  RebuildPolicy rebuild_policy{RebuildPolicy::FixedInterval};
  double t_target{0.0};  // > 0: NVT with a Nose-Hoover chain
  double t_period{0.1};
  i32 nhc_length{3};
  bool use_graph{true};
  i32 max_neighbors{0};
  i32 pair_unroll{0};
  i32 lds_budget{0};
  bool split_rebuild{false};
  bool disable_halo{false};  // no LDS halo staging in the EAM passes
};

struct FastPipelineStats {
  i64 ticks{0};
  i64 kernel_launches{0};
  i64 rebuilds{0};
  i64 graph_launches{0};
};

class FastPipelineScheduler {
 public:
  FastPipelineScheduler(const Box& box, i32 natoms, const potentials::MorseParams& params,
                        const FastPipelineConfig& cfg)
      : natoms_(natoms), kind_(PotentialKind::Morse), cutoff_(params.rc) {
    const tdmd_box b = hip::to_c(box);
    const tdmd_morse_params p = potentials::to_c(params);
    const tdmd_sched_config c = to_c(cfg);
    TDMD_HIP_CALL(tdmd_sched_create_morse(&s_, &b, natoms, &p, &c));
  }

  // The EamAlloy is read only here: its tables go to the device, and the
  // scheduler keeps no reference to it (fast_pipeline_scheduler.cu:27-40).
  template <class EamAlloyLike>
  FastPipelineScheduler(const Box& box, i32 natoms, const EamAlloyLike& eam,
                        const FastPipelineConfig& cfg)
      : natoms_(natoms), kind_(PotentialKind::Eam), cutoff_(eam.cutoff()) {
    potentials::DeviceEam tables;
    tables.upload_tables(eam);
    const tdmd_box b = hip::to_c(box);
    const tdmd_sched_config c = to_c(cfg);
    TDMD_HIP_CALL(tdmd_sched_create_eam_tables(&s_, &b, natoms, tables.handle(), &c));
  }

  ~FastPipelineScheduler() {
    if (s_) (void)tdmd_sched_destroy(s_);
  }
  FastPipelineScheduler(const FastPipelineScheduler&) = delete;
  FastPipelineScheduler& operator=(const FastPipelineScheduler&) = delete;

  void upload(const PositionVec* positions, const VelocityVec* velocities,
              const ForceVec* forces, const i32* types, const i32* ids, const real* masses,
              i32 n_types) {
    const auto n = static_cast<std::size_t>(natoms_);
    std::vector<double> f(3 * n, 0.0);
    if (forces)
      for (std::size_t i = 0; i < n; ++i) {
        f[3 * i] = static_cast<double>(forces[i].x);
        f[3 * i + 1] = static_cast<double>(forces[i].y);
        f[3 * i + 2] = static_cast<double>(forces[i].z);
      }
    std::vector<double> m(masses, masses + n_types);
    TDMD_HIP_CALL(tdmd_sched_upload(s_, reinterpret_cast<const double*>(positions),
                                    reinterpret_cast<const double*>(velocities), f.data(), types,
                                    ids, m.data(), n_types));
  }

  void run_until(i32 target_step) { TDMD_HIP_CALL(tdmd_sched_run_until(s_, target_step)); }

  void download(PositionVec* positions, VelocityVec* velocities, ForceVec* forces, i32* types,
                i32* ids, i32 natoms) const {
    std::vector<double> f(forces ? 3 * static_cast<std::size_t>(natoms) : 0);
    TDMD_HIP_CALL(tdmd_sched_download(s_, reinterpret_cast<double*>(positions),
                                      reinterpret_cast<double*>(velocities),
                                      forces ? f.data() : nullptr, types, ids, natoms));
    if (forces)
      for (std::size_t i = 0; i < static_cast<std::size_t>(natoms); ++i)
        forces[i] = ForceVec{static_cast<force_t>(f[3 * i]), static_cast<force_t>(f[3 * i + 1]),
                             static_cast<force_t>(f[3 * i + 2])};
  }

  const FastPipelineStats& stats() const {
    int64_t out[4];
    TDMD_HIP_CALL(tdmd_sched_stats(s_, out));
    stats_ = FastPipelineStats{out[0], out[1], out[2], out[3]};
    return stats_;
  }
  i32 current_step() const {
    i32 step = 0;
    TDMD_HIP_CALL(tdmd_sched_current_step(s_, &step));
    return step;
  }
  PotentialKind potential_kind() const noexcept { return kind_; }
  real interaction_cutoff() const noexcept { return cutoff_; }

  // PE, KE, NHC energy, T (3N dof), max and mean neighbour count.
  std::vector<double> thermo() {
    std::vector<double> out(6, 0.0);
    TDMD_HIP_CALL(tdmd_sched_thermo(s_, out.data()));
    return out;
  }
  // Device FP64 observables: energies, T, chain and conserved energy,
  // virial, volume, pressure (bar), neighbour counts.
  tdmd_observables observables() {
    tdmd_observables o{};
    TDMD_HIP_CALL(tdmd_sched_observables(s_, &o));
    return o;
  }
  hipStream_t stream() const {
    void* s = nullptr;
    TDMD_HIP_CALL(tdmd_sched_stream(s_, &s));
    return static_cast<hipStream_t>(s);
  }

 private:
  static tdmd_sched_config to_c(const FastPipelineConfig& cfg) {
    tdmd_sched_config c;
    tdmd_sched_config_default(&c);
    c.dt = static_cast<double>(cfg.dt);
    c.r_skin = static_cast<double>(cfg.r_skin);
    c.rebuild_every = cfg.rebuild_every;
    c.precision = hip::kPrecision;
    c.rebuild_policy = cfg.rebuild_policy == RebuildPolicy::Displacement
                           ? TDMD_REBUILD_DISPLACEMENT
                           : TDMD_REBUILD_FIXED;
    c.t_target = cfg.t_target;
    c.t_period = cfg.t_period;
    c.nhc_length = cfg.nhc_length;
    c.use_graph = cfg.use_graph ? 1 : 0;
    c.max_neighbors = cfg.max_neighbors;
    c.pair_unroll = cfg.pair_unroll;
    c.lds_budget = cfg.lds_budget;
    c.split_rebuild = cfg.split_rebuild ? 1 : 0;
    c.disable_halo = cfg.disable_halo ? 1 : 0;
    return c;
  }

  tdmd_sched* s_ = nullptr;
  i32 natoms_;
  PotentialKind kind_;
  real cutoff_;
  mutable FastPipelineStats stats_{};
};

}  // namespace tdmd::scheduler
