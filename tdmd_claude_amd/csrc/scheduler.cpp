// scheduler.cpp -- FastPipelineScheduler on gfx950
// (src/scheduler/fast_pipeline_scheduler.cuh:36-192,
//  src/scheduler/fast_pipeline_scheduler.cu:15-232).
//
// Same contract as the reference: construct with a box, an atom count and a
// potential; upload; run_until(step) (velocity Verlet with the cross-step
// fused kick, one stream, one synchronisation per call); download in upload
// order. Re-designed underneath:
//   * atoms are kept in a Morton-ordered cell sort, re-derived at each
//     neighbour rebuild on the device (no host round trip, no
//     cudaDeviceSynchronize -- compare device_cell_list.cu:103-111,148);
//   * neighbour lists are wave-interleaved (ELL) for coalesced reads;
//   * EAM runs as two passes per step: density+embedding, then force with
//     the next step's kick and drift fused into its epilogue;
//   * the rebuild trigger is decided on the device (fixed interval as the
//     reference, or max displacement > skin/2), so whole step sequences are
//     captured once as hipGraphs and replayed;
//   * NVT (Nose-Hoover chain, device_nose_hoover.cu:187-289) runs on the
//     device with the KE reduced from the kick epilogue's block partials.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include <hip/hip_ext.h>

#include "sched_kernels.inl"
#include "sched_lds.inl"

namespace tdmd_hip {

int read_setfl(const std::string& path, SetflData* out);

namespace {

using namespace sched;

// Rows of small systems are split over kSplitLanes lanes in the force
// passes (Morse, and EAM through the global-table passes): one lane per atom
// would give 32,000 atoms 125 workgroups on 256 CUs, each lane walking ~80
// slots in turn; the LDS EAM passes would run 32 workgroups. Thresholds from
// the measured pass times (mixed, 8 lanes against one): Morse 32,000 atoms
// 10.6 against 19.9 us, 108,000 26.3 against 28.9, 256,000 57.9 against
// 51.1; EAM density + force 32,000 22.7 against 53.7 us (LDS passes),
// 108,000 74.6 against 59.3.
constexpr int kSplitLanes = 8;
constexpr int kSplitMax = 131072;     // Morse
constexpr int kEamSplitMax = 65536;   // EAM
// The fused rebuild's list search takes 32 lanes per atom up to this size.
constexpr int kNlistSplitMax = 131072;

// Operation codes of a step sequence (also the graph-cache key alphabet).
enum Op : char {
  OP_BOOT = 'b',        // half kick + drift (window start, NVE)
  OP_KD_FULL = 'k',     // full kick + drift as its own launch (Morse pass timing)
  OP_NVT_PRE = 'p',     // chain half-step before the step
  OP_KD_NVT = 'q',      // scaled half kick + drift
  OP_KD_NVT2 = 'Q',     // the previous step's closing half kick, then as 'q'
  OP_REB_FORCED = 'R',  // set flag + rebuild
  OP_REB_COND = 'r',    // rebuild if the device flag is set
  OP_DENS = 'd',        // EAM density pass
  OP_F_FULL = 'F',      // force + next step's kick-drift (Morse: into the other record buffer)
  OP_F_ONLY = 'o',      // force, store F
  OP_F_HALF = 'h',      // force + closing half kick
  OP_F_KICK2 = '2',     // force + half kick + KE partials (NVT)
  OP_F_KE = 'e',        // force + KE partials of the half-kicked v, v not stored
                        // (NVT inside a window: the next 'Q' applies the kick)
  OP_NVT_POST = 'P',    // chain half-step after the step
  OP_NVT_POST_PRE = 'N',// chain half-step after the step + the next step's
  OP_NVT_FINAL = 'Z'    // apply pending chain scale, commit
};

int round_up(int x, int m) { return (x + m - 1) / m * m; }

uint32_t spread3(uint32_t v) {
  uint32_t r = 0;
  for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
  return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// Engine: precision- and mirror-specific launches.
// ---------------------------------------------------------------------------
struct EngineBase {
  virtual ~EngineBase() = default;
  virtual int init() = 0;
  virtual int launch(char op, hipStream_t s) = 0;
  virtual int write_mirror(hipStream_t s) = 0;
  virtual int energy_pass(hipStream_t s, double* d_out) = 0;  // d_out[0..2]
  virtual int count_pass(hipStream_t s, int32_t* d_max) = 0;
  virtual int fill_pass(hipStream_t s) = 0;
  virtual int sort_pass(hipStream_t s) = 0;
  // Time the candidate kernel shapes on the current state and keep the
  // fastest; out = {pair_unroll, lds_budget bytes (0: not applicable)}.
  virtual int tune(int iters, int32_t* out) = 0;
  // {lanes per row, neighbour slots per pipelined group, LDS window budget}
  virtual void shape(int32_t* out) const = 0;
  // Bracket the next force-path launch (EAM density / force, Morse
  // kick-drift / force) with events stamped by the dispatch itself.
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  void time_next(hipEvent_t a, hipEvent_t b) {
    ev_start = a;
    ev_stop = b;
  }
};

}  // namespace tdmd_hip

struct tdmd_sched {
  int kind = 0;  // 0 Morse, 1 EAM
  tdmd_box box{};
  int natoms = 0;
  int nblocks = 0;   // 256-thread grid (multiple of 8)
  int nb_force = 0;  // grid of the EAM/Morse passes = block partials they emit
  int lanes = 1;  // lanes per atom of the force passes (Morse, global-table EAM)
  bool stage_ok = false;  // halo buffers exist: the staged LDS passes can be chosen
  int nb_rebuild = 0;  // fused rebuild grid (all resident at once); 0: separate kernels
  tdmd_hip::sched::GridBarrier* d_bar = nullptr;
  tdmd_sched_config cfg{};
  tdmd_morse_params morse{};
  tdmd_eam_tables* tables = nullptr;
  hipStream_t stream = nullptr;
  double cutoff = 0, r_list = 0;
  tdmd_hip::sched::StepArgs a{};
  tdmd_hip::sched::GridArgs g{};
  tdmd_hip::sched::PermArgs t{};
  tdmd_hip::sched::NhcParams nhcp{};
  std::unique_ptr<tdmd_hip::EngineBase> eng;
  std::vector<void*> allocs;
  int32_t* d_maxcnt = nullptr;
  int64_t* d_rebuilds = nullptr;
  double* d_scalars = nullptr;
  std::vector<int32_t> h_types, h_ids;
  std::vector<double> h_masses;
  bool uploaded = false, initial_forces = false;
  // Morse: the gather records alternate between two buffers (the fused
  // kick-drift writes the next step's into the one not being gathered);
  // true when a.M / a.M2 are exchanged relative to their allocation
  bool rec_swapped = false;
  int current_step = 0, steps_since = 0;
  int64_t ticks = 0, launches = 0, host_rebuilds = 0, graph_launches = 0;
  // In-window pass timing: the passes' own device clock (StepArgs::pclock)
  bool pass_timing = false;
  unsigned long long* d_pclock = nullptr;  // 8 words: density, force
  int clock_khz = 0;                       // device wall clock rate
  std::map<std::string, hipGraphExec_t> graphs;
  std::map<std::string, hipGraph_t> graph_defs;
  std::map<std::string, int64_t> graph_kernels;
  std::map<std::string, bool> graph_swaps;  // the graph exchanges the record buffers
  // batches of independent boxes (replicas), atoms of replica r in slots
  // [rep_off[r], rep_off[r+1])
  bool multibox = false;
  std::vector<tdmd_hip::sched::BoxGeo> h_boxes;
  std::vector<int> rep_off;
};

namespace tdmd_hip {
namespace {

// Destroys every captured graph (the caller has synchronised the stream).
void drop_graphs(tdmd_sched* S) {
  for (auto& kv : S->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto& kv : S->graph_defs) (void)hipGraphDestroy(kv.second);
  S->graphs.clear();
  S->graph_defs.clear();
  S->graph_kernels.clear();
  S->graph_swaps.clear();
}

template <class R, class Mir>
struct Engine final : EngineBase {
  tdmd_sched* S;
  explicit Engine(tdmd_sched* s) : S(s) {}

  using Q = typename Quad<R>::type;

  template <class K, class... Args>
  void go(K kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s, Args... args) {
    if (ev_start) {
      hipExtLaunchKernelGGL(kernel, grid, block, static_cast<uint32_t>(lds), s, ev_start, ev_stop,
                            0u, args...);
      ev_start = ev_stop = nullptr;
    } else {
      kernel<<<grid, block, lds, s>>>(args...);
    }
  }

  const SplineMeta* meta() const { return static_cast<const SplineMeta*>(S->tables->d_meta); }
  const Q* coeff() const { return static_cast<const Q*>(S->tables->d_coeff); }
  double rc_sq() const { return S->tables->cutoff * S->tables->cutoff; }
  MorseCoef mp() const {
    return MorseCoef{S->morse.d, S->morse.alpha, S->morse.r0, S->morse.rc_sq};
  }

  // Single-element EAM runs the LDS-table kernels (sched_lds.inl); the
  // windows are sized once from the LDS budget (the device's per-workgroup
  // limit unless the configuration asks for less, e.g. to fit two
  // workgroups per CU).
  bool lds = false;
  // Halo staging (k_halo + the kStage passes) for 16-byte gather records.
  static constexpr bool kCanStage = sizeof(typename Mir::T) == 16;
  bool stage = false;
  bool lds_ok = false;  // single-element tables on one r grid: the LDS passes can run
  int nb_lds = 0;       // 1024-atom blocks of the LDS passes (multiple of 8)
  bool iso = false;  // fixed-point mirror with one unit on all axes (every box)
  int ncu = 0;
  size_t lds_cu = 160 * 1024;  // LDS of one CU
  int32_t* d_halo_done = nullptr;
  size_t stage_bytes() const { return stage ? static_cast<size_t>(kStageCap) * 16 : 0; }
  int unroll = 4;
  int lds_grid = 0;   // persistent workgroups of the LDS passes
  int lds_limit = 0;  // per-workgroup dynamic LDS the kernels are allowed
  int nint = 0;       // r-grid intervals
  int budget = 0;     // current window budget
  LdsWindow w_dens[2]{}, w_force[2]{};  // [kEnergy]
  size_t b_dens[2]{}, b_force[2]{};

  static LdsWindow window(int nint, size_t per, size_t budget, size_t* bytes) {
    LdsWindow w;
    w.nwin = static_cast<int>(std::min<size_t>(static_cast<size_t>(nint), budget / per));
    w.k_lo = nint - w.nwin;
    *bytes = window_bytes(w.nwin, per);
    return w;
  }
  template <class K>
  static int allow_lds(K* fn, size_t bytes) {
    TDMD_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     static_cast<int>(bytes)));
    return TDMD_OK;
  }

  template <bool kWin, int kU, bool kSt, bool kIs>
  int allow_all() {
    int rc = 0;
    const size_t L = static_cast<size_t>(lds_limit);
    if ((rc = allow_lds(&k_eam_density_lds<R, Mir, false, kWin, kU, kSt, kIs>, L)) ||
        (rc = allow_lds(&k_eam_density_lds<R, Mir, true, kWin, kU, kSt, kIs>, L)) ||
        (rc = allow_lds(&k_eam_force_lds<R, Mir, kForceOnly, true, kWin, kU, kSt, kIs>, L)) ||
        (rc = allow_lds(&k_eam_force_lds<R, Mir, kForceOnly, false, kWin, kU, kSt, kIs>, L)) ||
        (rc = allow_lds(&k_eam_force_lds<R, Mir, kFullKickDrift, false, kWin, kU, kSt, kIs>,
                        L)) ||
        (rc = allow_lds(&k_eam_force_lds<R, Mir, kHalfKick, false, kWin, kU, kSt, kIs>, L)) ||
        (rc = allow_lds(&k_eam_force_lds<R, Mir, kKick2Ke, false, kWin, kU, kSt, kIs>, L)) ||
        (rc = allow_lds(&k_eam_force_lds<R, Mir, kKe, false, kWin, kU, kSt, kIs>, L)))
      return rc;
    return TDMD_OK;
  }
  template <bool kSt, bool kIs>
  int allow_shapes() {
    int rc = 0;
    if ((rc = allow_all<false, 4, kSt, kIs>()) || (rc = allow_all<true, 4, kSt, kIs>()) ||
        (rc = allow_all<false, 8, kSt, kIs>()) || (rc = allow_all<true, 8, kSt, kIs>()))
      return rc;
    return TDMD_OK;
  }

  // Windows for a budget (bytes, clamped to the allowed limit) and unroll.
  // With halo staging the windows get what the halo leaves.
  void configure(int u, int bytes) {
    budget = std::min(bytes > 0 ? bytes : lds_limit, lds_limit);
    const size_t wb = static_cast<size_t>(budget) - std::min(static_cast<size_t>(budget), stage_bytes());
    using V2 = typename Vec2<R>::type;
    for (int e = 0; e < 2; ++e) {
      w_dens[e] = window(nint, sizeof(Q), wb, &b_dens[e]);
      w_force[e] = window(nint, sizeof(Q) + sizeof(V2) + (e ? sizeof(R) : 0), wb, &b_force[e]);
    }
    unroll = u;
    // one workgroup per LDS slot of the device: as many as the CU's LDS
    // holds of the largest pass (at most two of 1024 threads), per CU
    const size_t per_wg = std::max(b_dens[0], b_force[0]) + stage_bytes() + 1024;
    const int per_cu = std::max(1, std::min(2, static_cast<int>(lds_cu / per_wg)));
    lds_grid = std::max(8, std::min(nb_lds, (std::max(ncu, 8) / 8) * 8 * per_cu));
  }

  // Interleaved derivative table of the several-element force pass (see
  // StepArgs::pairtab), from the tables' FP64 coefficients in the kernels'
  // precision: the same values the per-spline quads hold; in mixed
  // precision c and d are stored as 2c and 3d, scaled in float (the values
  // the per-spline path forms at use).
  int build_pair_table() {
    const tdmd_eam_tables* t = S->tables;
    const int nt = t->ntypes;
    SplineMeta m[kMaxSplines];
    const int ns = 2 * nt + nt * (nt + 1) / 2;
    TDMD_HIP_TRY(hipMemcpy(m, t->d_meta, ns * sizeof(SplineMeta), hipMemcpyDeviceToHost));
    for (int k = nt + 1; k < ns; ++k)  // rho and phi splines on one r grid
      if (m[k].n != m[nt].n || m[k].dr != m[nt].dr || m[k].rmin != m[nt].rmin) return TDMD_OK;
    const int nint = m[nt].n - 1;
    std::vector<Q> h(static_cast<size_t>(nt) * nt * nint * 3);
    constexpr bool kF32 = std::is_same_v<R, float>;
    auto co = [&](const std::vector<double>& v, int spline, int k) {
      return static_cast<R>(v[static_cast<size_t>(t->offsets[static_cast<size_t>(spline)] + k)]);
    };
    auto c2 = [&](int spline, int k) {
      const R c = co(t->c, spline, k);
      return kF32 ? static_cast<R>(2) * c : c;
    };
    auto d3 = [&](int spline, int k) {
      const R d = co(t->d, spline, k);
      return kF32 ? static_cast<R>(3) * d : d;
    };
    for (int ti = 0; ti < nt; ++ti)
      for (int tj = 0; tj < nt; ++tj) {
        const int sp = 2 * nt + phi_index(ti, tj, nt), sj = nt + tj, si = nt + ti;
        for (int k = 0; k < nint; ++k) {
          Q* e = &h[(static_cast<size_t>(ti * nt + tj) * nint + k) * 3];
          e[0].x = co(t->b, sp, k);
          e[0].y = c2(sp, k);
          e[0].z = d3(sp, k);
          e[0].w = co(t->b, sj, k);
          e[1].x = c2(sj, k);
          e[1].y = d3(sj, k);
          e[1].z = co(t->b, si, k);
          e[1].w = c2(si, k);
          e[2].x = d3(si, k);
          e[2].y = e[2].z = e[2].w = static_cast<R>(0);
        }
      }
    void* p = nullptr;
    if (int rc = tdmd_malloc(&p, h.size() * sizeof(Q))) return rc;
    S->allocs.push_back(p);
    TDMD_HIP_TRY(hipMemcpy(p, h.data(), h.size() * sizeof(Q), hipMemcpyHostToDevice));
    S->a.pairtab = p;
    S->a.pair_nint = nint;
    return TDMD_OK;
  }

  // Kernel path for `lanes` lanes per row: one lane with single-element
  // tables runs the LDS passes (halo-staged where the halo buffers exist);
  // otherwise the global-table passes with rows split over `lanes` lanes.
  // The fused rebuild builds halos only while the staged passes are chosen.
  void set_path(int lanes) {
    S->lanes = lanes;
    lds = lanes == 1 && lds_ok;
    stage = lds && S->stage_ok;
    S->a.stage_cap = stage ? kStageCap : 0;
    S->nb_force = lds ? nb_lds : S->nblocks * lanes;
  }

  int init() override {
    // rows split over kSplitLanes lanes when one lane per atom would leave
    // the chip mostly idle (partials: one per block of the larger grid)
    S->lanes = S->cfg.pair_lanes > 0
                   ? S->cfg.pair_lanes
                   : (S->natoms <= (S->kind == 1 ? kEamSplitMax : kSplitMax) ? kSplitLanes : 1);
    S->nb_force = S->nblocks * S->lanes;
    S->nb_rebuild = 0;
    if (!S->cfg.split_rebuild) {
      // every workgroup of the fused rebuild must be resident at once
      int dev = 0, ncu = 0, per_cu = 0;
      TDMD_HIP_TRY(hipGetDevice(&dev));
      TDMD_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
      TDMD_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void*>(&k_rebuild_fused<R, Mir>), kRebuildBlock, 0));
      // at most two workgroups per CU: a margin under the occupancy limit.
      // Small systems search each atom's cells with 32 lanes (one lane per
      // atom leaves the search a chain of dependent loads on a few CUs) --
      // under the fixed-interval policy. Under the displacement policy the
      // launch is a no-op check on nearly every step (a 32,000-atom Morse
      // system rebuilds once in 10,000 steps), and a grid sized for 32
      // lanes per atom makes every check dispatch 16 times more workgroups:
      // there the search takes one lane per atom.
      const bool split = S->natoms <= kNlistSplitMax &&
                         S->cfg.rebuild_policy != TDMD_REBUILD_DISPLACEMENT;
      const int64_t lanes = static_cast<int64_t>(S->natoms) * (split ? 32 : 1);
      const int need = static_cast<int>(
          std::max<int64_t>((lanes + kRebuildBlock - 1) / kRebuildBlock, 1));
      S->nb_rebuild = std::max(1, std::min(std::min(per_cu, 2) * ncu, need));
      S->nb_rebuild = std::min(S->nb_rebuild, sched::kMaxScanGroups);
      if (per_cu < 1) S->nb_rebuild = 0;  // cannot guarantee residency: separate kernels
      S->g.nl_split = split && S->nb_rebuild > 0 ? 1 : 0;
    }
    if (S->kind == 1 && S->tables->ntypes > 1) return build_pair_table();
    if (S->kind != 1 || S->tables->ntypes != 1) return TDMD_OK;
    // rho_0 and phi_00 must share the r grid (setfl: one Nr, dr).
    SplineMeta m[3];
    TDMD_HIP_TRY(hipMemcpy(m, S->tables->d_meta, sizeof(m), hipMemcpyDeviceToHost));
    if (m[1].n != m[2].n || m[1].dr != m[2].dr || m[1].rmin != m[2].rmin || m[1].n < 2)
      return TDMD_OK;
    int dev = 0, lmax = 0;
    TDMD_HIP_TRY(hipGetDevice(&dev));
    TDMD_HIP_TRY(hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    // The dynamic window plus the kernels' static LDS (block reduction)
    // must fit the per-workgroup limit; keep 1 KiB for the static part.
    lds_limit = std::min(lmax, 160 * 1024) - 1024;
    if (S->cfg.lds_budget > 0 && S->cfg.lds_budget < 4096)
      return fail(TDMD_E_INVALID, "scheduler: lds_budget below 4096 bytes");
    nint = m[1].n - 1;
    int rc = 0;
    if ((rc = allow_shapes<false, false>())) return rc;
    TDMD_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int lcu = 0;
    if (hipDeviceGetAttribute(&lcu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) ==
            hipSuccess &&
        lcu > 0)
      lds_cu = static_cast<size_t>(lcu);
    lds_ok = true;
    nb_lds = round_up((S->natoms + kBigBlock - 1) / kBigBlock, 8);
    if constexpr (kCanStage) {
      iso = true;
      auto same_unit = [](const MirrorParams& p) {
        return p.unit[0] == p.unit[1] && p.unit[0] == p.unit[2];
      };
      if (S->multibox) {
        for (const BoxGeo& b : S->h_boxes) iso = iso && same_unit(b.mp);
      } else {
        iso = same_unit(S->a.mp);
      }
      if (iso && (rc = allow_shapes<false, true>())) return rc;
      if (!S->cfg.disable_halo && lds_limit >= static_cast<int>(kStageCap) * 16 + 16384) {
        if ((rc = (iso ? allow_shapes<true, true>() : allow_shapes<true, false>()))) return rc;
        const size_t nbf = static_cast<size_t>(nb_lds);
        void* p = nullptr;
        const size_t sizes[5] = {nbf * kHaloMax * 4, nbf * 4, nbf * 4, 4,
                                 static_cast<size_t>(round_up(S->natoms, 64)) * 4};
        void* ptrs[5];
        for (int k = 0; k < 5; ++k) {
          if ((rc = tdmd_malloc(&p, sizes[k]))) return rc;
          S->allocs.push_back(p);
          ptrs[k] = p;
        }
        S->a.hj = static_cast<const int32_t*>(ptrs[0]);
        S->a.hn = static_cast<const int32_t*>(ptrs[1]);
        S->a.hmode = static_cast<const int32_t*>(ptrs[2]);
        d_halo_done = static_cast<int32_t*>(ptrs[3]);
        S->a.cntc = static_cast<int32_t*>(ptrs[4]);
        TDMD_HIP_TRY(hipMemset(ptrs[2], 0xff, sizes[2]));  // no halo yet: not staged
        TDMD_HIP_TRY(hipMemset(d_halo_done, 0, 4));
        S->stage_ok = true;
      }
    }
    set_path(S->lanes);
    // default group: four slots in mixed precision, eight in FP64, the
    // faster shape of each on the 1M-atom EAM state (mixed force pass 108 us
    // with four against 111 with eight; FP64 305 us with eight against 317)
    configure(S->cfg.pair_unroll != 0 ? S->cfg.pair_unroll : (std::is_same_v<R, float> ? 4 : 8),
              S->cfg.lds_budget);
    return TDMD_OK;
  }

  // Halos of the staged passes, after a rebuild (returns at once otherwise).
  void halo(hipStream_t s) {
    if (!stage) return;
    const int grid = std::max(1, std::min(S->nb_force, 2 * std::max(ncu, 1)));
    k_halo<<<grid, kHaloBlock, 0, s>>>(S->a, (S->natoms + kHaloBlock - 1) / kHaloBlock,
                                       d_halo_done);
    ++S->launches;
  }

  void shape(int32_t* out) const override {
    out[0] = S->lanes;
    out[1] = lds ? unroll : kPairUnroll;
    out[2] = lds ? budget : 0;
  }

  // The force-path passes of one step with the force-only epilogue: they
  // rewrite M2 and F from the current positions and change nothing else.
  void probe_passes(hipStream_t s) {
    if (S->kind == 1) density<false>(s);
    force<kForceOnly, false>(s);
  }

  int tune(int iters, int32_t* out) override {
    // Candidates: lanes per row {1, 4, 8}; with one lane and single-element
    // tables the LDS passes, in two pipelined group sizes (and, without
    // halo staging, two spline-window budgets).
    struct Cand {
      int lanes, u, b;
    };
    std::vector<Cand> cands;
    if (lds_ok) {
      cands.push_back({1, 8, lds_limit});
      cands.push_back({1, 4, lds_limit});
      if (!S->stage_ok) {
        cands.push_back({1, 8, 80 * 1024 - 1024});
        cands.push_back({1, 4, 80 * 1024 - 1024});
      }
    } else {
      cands.push_back({1, 0, 0});
    }
    cands.push_back({4, 0, 0});
    cands.push_back({8, 0, 0});
    hipStream_t s = S->stream;
    hipEvent_t e0, e1;
    TDMD_HIP_TRY(hipEventCreate(&e0));
    TDMD_HIP_TRY(hipEventCreate(&e1));
    const Cand start{S->lanes, lds ? unroll : 0, lds ? budget : 0};
    std::vector<float> times;
    for (const Cand& c : cands) {
      if (int rc = apply_shape(c.lanes, c.u, c.b, s)) return rc;
      for (int k = 0; k < 2; ++k) probe_passes(s);
      TDMD_HIP_TRY(hipEventRecord(e0, s));
      for (int k = 0; k < iters; ++k) probe_passes(s);
      TDMD_HIP_TRY(hipEventRecord(e1, s));
      TDMD_HIP_TRY(hipEventSynchronize(e1));
      TDMD_LAUNCH_CHECK();
      float ms = 0.0f;
      TDMD_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      times.push_back(ms);
    }
    // the fastest shape; the starting shape when it is within 1 % of it
    size_t k_best = 0;
    for (size_t k = 1; k < cands.size(); ++k)
      if (times[k] < times[k_best]) k_best = k;
    Cand pick = cands[k_best];
    for (size_t k = 0; k < cands.size(); ++k) {
      const Cand& c = cands[k];
      const bool is_start =
          c.lanes == start.lanes && (c.lanes != 1 || !lds_ok || (c.u == start.u && c.b == start.b));
      if (is_start && times[k] <= 1.01f * times[k_best]) pick = c;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (int rc = apply_shape(pick.lanes, pick.u, pick.b, s)) return rc;
    // M2 and F as the chosen shape computes them, so the trajectory goes on
    // exactly as it would have with that shape (bitwise unchanged when the
    // choice is the starting shape)
    probe_passes(s);
    TDMD_LAUNCH_CHECK();
    out[0] = lds ? unroll : kPairUnroll;
    out[1] = lds ? budget : 0;
    return TDMD_OK;
  }

  // Switch the kernel shape (tuner): path, group size, window budget. A
  // switch to the staged passes rebuilds the halos of the current rows.
  int apply_shape(int lanes, int u, int b, hipStream_t s) {
    const bool was_staged = stage;
    set_path(lanes);
    if (lds) configure(u, b);
    if (stage && !was_staged) {
      TDMD_HIP_TRY(hipMemsetAsync(S->a.flags + 4, 0xff, 4, s));  // halos due
      halo(s);
      TDMD_LAUNCH_CHECK();
    }
    return TDMD_OK;
  }

  template <bool kEnergy, bool kWin, int kU, bool kSt, bool kIs>
  void density_go(hipStream_t s) {
    go(k_eam_density_lds<R, Mir, kEnergy, kWin, kU, kSt, kIs>, lds_grid, kBigBlock,
       b_dens[kEnergy] + (kSt ? stage_bytes() : 0), s, S->a, meta(), coeff(), rc_sq(),
       w_dens[kEnergy], S->nb_force);
  }
  template <bool kEnergy, bool kWin, int kU>
  void density_lds(hipStream_t s) {
    if constexpr (kCanStage) {
      if (stage) {
        iso ? density_go<kEnergy, kWin, kU, true, true>(s)
            : density_go<kEnergy, kWin, kU, true, false>(s);
        return;
      }
      if (iso) {
        density_go<kEnergy, kWin, kU, false, true>(s);
        return;
      }
    }
    density_go<kEnergy, kWin, kU, false, false>(s);
  }
  template <bool kEnergy>
  void density(hipStream_t s) {
    if (!lds) {
      const int nb = S->nblocks;
      if (S->lanes == 8)
        go(k_eam_density<R, Mir, kEnergy, 8>, nb * 8, kBlock, 0, s, S->a, meta(), coeff(),
           rc_sq());
      else if (S->lanes == 4)
        go(k_eam_density<R, Mir, kEnergy, 4>, nb * 4, kBlock, 0, s, S->a, meta(), coeff(),
           rc_sq());
      else
        go(k_eam_density<R, Mir, kEnergy, 1>, nb, kBlock, 0, s, S->a, meta(), coeff(),
           rc_sq());
      return;
    }
    const bool win = w_dens[kEnergy].k_lo > 0;
    if (unroll == 8)
      win ? density_lds<kEnergy, true, 8>(s) : density_lds<kEnergy, false, 8>(s);
    else
      win ? density_lds<kEnergy, true, 4>(s) : density_lds<kEnergy, false, 4>(s);
  }

  template <int kEpi, bool kEnergy, bool kWin, int kU, bool kSt, bool kIs>
  void force_go(hipStream_t s) {
    go(k_eam_force_lds<R, Mir, kEpi, kEnergy, kWin, kU, kSt, kIs>, lds_grid, kBigBlock,
       b_force[kEnergy] + (kSt ? stage_bytes() : 0), s, S->a, meta(), coeff(), rc_sq(),
       w_force[kEnergy], S->nb_force);
  }
  template <int kEpi, bool kEnergy, bool kWin, int kU>
  void force_lds(hipStream_t s) {
    if constexpr (kCanStage) {
      if (stage) {
        iso ? force_go<kEpi, kEnergy, kWin, kU, true, true>(s)
            : force_go<kEpi, kEnergy, kWin, kU, true, false>(s);
        return;
      }
      if (iso) {
        force_go<kEpi, kEnergy, kWin, kU, false, true>(s);
        return;
      }
    }
    force_go<kEpi, kEnergy, kWin, kU, false, false>(s);
  }

  template <int kEpi, bool kEnergy, bool kMulti>
  void eam_force_global(hipStream_t s) {
    const int nb = S->nblocks;
    if (S->lanes == 8)
      go(k_eam_force<R, Mir, kEpi, kEnergy, kMulti, 8>, nb * 8, kBlock, 0, s, S->a, meta(),
         coeff(), rc_sq());
    else if (S->lanes == 4)
      go(k_eam_force<R, Mir, kEpi, kEnergy, kMulti, 4>, nb * 4, kBlock, 0, s, S->a, meta(),
         coeff(), rc_sq());
    else
      go(k_eam_force<R, Mir, kEpi, kEnergy, kMulti, 1>, nb, kBlock, 0, s, S->a, meta(),
         coeff(), rc_sq());
  }

  template <int kEpi, bool kEnergy>
  void force(hipStream_t s) {
    const int nb = S->nblocks;
    if (S->kind == 1) {
      const bool win = w_force[kEnergy].k_lo > 0;
      if (lds && unroll == 8)
        win ? force_lds<kEpi, kEnergy, true, 8>(s) : force_lds<kEpi, kEnergy, false, 8>(s);
      else if (lds)
        win ? force_lds<kEpi, kEnergy, true, 4>(s) : force_lds<kEpi, kEnergy, false, 4>(s);
      else if (S->tables->ntypes > 1)
        eam_force_global<kEpi, kEnergy, true>(s);
      else
        eam_force_global<kEpi, kEnergy, false>(s);
    } else {
      if (S->lanes == 8)
        go(k_morse_force<R, Mir, kEpi, kEnergy, 8>, nb * 8, kBlock, 0, s, S->a, mp());
      else if (S->lanes == 4)
        go(k_morse_force<R, Mir, kEpi, kEnergy, 4>, nb * 4, kBlock, 0, s, S->a, mp());
      else
        go(k_morse_force<R, Mir, kEpi, kEnergy, 1>, nb, kBlock, 0, s, S->a, mp());
      if constexpr (kEpi == kFullKickDrift) {
        // the epilogue wrote the next step's records into M2: it is now M
        std::swap(S->a.M, S->a.M2);
        S->rec_swapped = !S->rec_swapped;
      }
    }
  }

  // Conditional rebuild: one fused launch, or the separate kernels.
  // forced: rebuild whatever the flag says (the fused kernel takes it as an
  // argument; the separate kernels need the flag raised first)
  int rebuild(hipStream_t s, bool forced) {
    if (S->nb_rebuild > 0) {
      k_rebuild_fused<R, Mir><<<S->nb_rebuild, kRebuildBlock, 0, s>>>(
          S->a, S->g, S->t, S->r_list * S->r_list, S->cutoff * S->cutoff,
          const_cast<int32_t*>(S->a.nbr), const_cast<int32_t*>(S->a.cnt), S->d_bar,
          S->d_rebuilds, forced ? 1 : 0);
      ++S->launches;  // the fused kernel also builds the halos
      TDMD_LAUNCH_CHECK();
      return TDMD_OK;
    }
    const int rc = sort_pass(s);
    return rc ? rc : fill_pass(s);
  }

  int sort_pass(hipStream_t s) override {
    const int gb = grid_for(S->natoms), gc = grid_for(S->g.ncells);
    k_zero_cells<<<gc, kBlock, 0, s>>>(S->a.flags, S->g);
    k_bin<Mir><<<gb, kBlock, 0, s>>>(S->a, S->g);
    k_cell_scan<<<1, 1024, 0, s>>>(S->a.flags, S->g);
    k_scatter<<<gb, kBlock, 0, s>>>(S->a, S->g);
    k_order<<<gc, kBlock, 0, s>>>(S->a, S->g);
    k_permute<R><<<gb, kBlock, 0, s>>>(S->a, S->g, S->t);
    k_copyback<R, Mir><<<gb, kBlock, 0, s>>>(S->a, S->g, S->t);
    S->launches += 7;
    TDMD_LAUNCH_CHECK();
    return TDMD_OK;
  }
  int fill_pass(hipStream_t s) override {
    const int gb = grid_for(S->natoms);
    k_nlist_ell<false, Mir><<<gb, kBlock, 0, s>>>(S->a, S->g, S->r_list * S->r_list,
                                             S->cutoff * S->cutoff,
                                             const_cast<int32_t*>(S->a.nbr),
                                             const_cast<int32_t*>(S->a.cnt), nullptr);
    k_rebuild_done<<<gb, kBlock, 0, s>>>(S->a.flags, S->g.done, gb, S->d_rebuilds);
    S->launches += 2;
    halo(s);
    TDMD_LAUNCH_CHECK();
    return TDMD_OK;
  }
  int count_pass(hipStream_t s, int32_t* d_max) override {
    k_nlist_ell<true, Mir><<<grid_for(S->natoms), kBlock, 0, s>>>(S->a, S->g,
                                                                  S->r_list * S->r_list,
                                                             S->cutoff * S->cutoff, nullptr,
                                                             nullptr, d_max);
    TDMD_LAUNCH_CHECK();
    return TDMD_OK;
  }
  int write_mirror(hipStream_t s) override {
    k_write_mirror<Mir><<<grid_for(S->natoms), kBlock, 0, s>>>(S->a);
    TDMD_LAUNCH_CHECK();
    return TDMD_OK;
  }
  int energy_pass(hipStream_t s, double* d_out) override {
    const int nb = S->nb_force;
    if (S->kind == 1) {
      density<true>(s);
      launch_sum_partials(S->a.partials, nb, d_out + 0, 0, s);
    } else {
      tdmd_fill_u32(d_out, 0u, 2, s);
    }
    force<kForceOnly, true>(s);
    launch_sum_partials(S->a.partials, nb, d_out + 1, 0, s);
    launch_sum_partials(S->a.partials_w, nb, d_out + 3, 0, s);
    k_ke_partials<R><<<grid_for(S->natoms), kBlock, 0, s>>>(S->a);
    launch_sum_partials(S->a.partials, grid_for(S->natoms), d_out + 2, 0, s);
    TDMD_LAUNCH_CHECK();
    return TDMD_OK;
  }

  int launch(char op, hipStream_t s) override {
    const int nb = S->nblocks;
    const int gb = grid_for(S->natoms);
    switch (op) {
      case OP_BOOT:
        k_kick_drift<R, Mir, false, false><<<nb, kBlock, 0, s>>>(S->a);
        break;
      case OP_KD_FULL:
        go(k_kick_drift<R, Mir, false, true>, nb, kBlock, 0, s, S->a);
        break;
      case OP_KD_NVT:
        k_kick_drift<R, Mir, true, false><<<nb, kBlock, 0, s>>>(S->a);
        break;
      case OP_KD_NVT2:
        k_kick_drift<R, Mir, true, false, true><<<nb, kBlock, 0, s>>>(S->a);
        break;
      case OP_NVT_PRE:
        k_nhc_pre<<<1, 64, 0, s>>>(S->a.nhc, S->nhcp);
        break;
      case OP_NVT_POST:
        k_nhc_post<false><<<1, 256, 0, s>>>(S->a.nhc, S->a.partials, S->nb_force, S->nhcp);
        break;
      case OP_NVT_POST_PRE:
        k_nhc_post<true><<<1, 256, 0, s>>>(S->a.nhc, S->a.partials, S->nb_force, S->nhcp);
        break;
      case OP_NVT_FINAL:
        k_scale_v<<<gb, kBlock, 0, s>>>(S->a);
        k_nhc_commit<<<1, 64, 0, s>>>(S->a.nhc);
        ++S->launches;
        break;
      case OP_REB_FORCED: {
        if (S->nb_rebuild == 0) {
          k_set_flag<<<1, 64, 0, s>>>(S->a.flags);
          ++S->launches;
        }
        return rebuild(s, true);
      }
      case OP_REB_COND:
        return rebuild(s, false);
      case OP_DENS:
        density<false>(s);
        break;
      case OP_F_FULL:
        force<kFullKickDrift, false>(s);
        break;
      case OP_F_ONLY:
        force<kForceOnly, false>(s);
        break;
      case OP_F_HALF:
        force<kHalfKick, false>(s);
        break;
      case OP_F_KICK2:
        force<kKick2Ke, false>(s);
        break;
      case OP_F_KE:
        force<kKe, false>(s);
        break;
      default:
        return fail(TDMD_E_STATE, std::string("scheduler: unknown op ") + op);
    }
    ++S->launches;
    TDMD_LAUNCH_CHECK();
    return TDMD_OK;
  }
};

template <class T>
int dalloc(tdmd_sched* S, T** p, size_t n) {
  void* q = nullptr;
  const int rc = tdmd_malloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (rc) return rc;
  S->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return TDMD_OK;
}

int sched_alloc(tdmd_sched* S) {
  const int N = S->natoms;
  const size_t fsz = S->cfg.precision == TDMD_MIXED ? sizeof(float) : sizeof(double);
  StepArgs& a = S->a;
  int rc = 0;
  double** d9[9] = {&a.px, &a.py, &a.pz, &a.vx, &a.vy, &a.vz, &a.bx, &a.by, &a.bz};
  for (auto* p : d9)
    if ((rc = dalloc(S, p, N))) return rc;
  void** f3[3] = {&a.fx, &a.fy, &a.fz};
  for (auto* p : f3) {
    char* q;
    if ((rc = dalloc(S, &q, N * fsz))) return rc;
    *p = q;
  }
  int32_t* types;
  if ((rc = dalloc(S, &types, N))) return rc;
  a.types = types;
  double* masses;
  if ((rc = dalloc(S, &masses, 64))) return rc;
  a.masses = masses;
  const bool fixed = S->cfg.precision == TDMD_MIXED && !S->cfg.mirror_fp64;
  const size_t msz = fixed ? 16 : 32;
  // gathers address records by a 32-bit byte offset
  if (static_cast<uint64_t>(N) * msz >= (uint64_t{1} << 32))
    return fail(TDMD_E_INVALID, "scheduler: at most " + std::to_string((uint64_t{1} << 32) / msz - 1) +
                                    " atoms in one scheduler at this precision");
  // kRecPad spare records past the last atom: the list search reads its
  // candidates in whole groups of eight from the first of a cell's atoms
  char *M, *M2;
  if ((rc = dalloc(S, &M, (static_cast<size_t>(N) + sched::kRecPad) * msz))) return rc;
  if ((rc = dalloc(S, &M2, (static_cast<size_t>(N) + sched::kRecPad) * msz))) return rc;
  a.M = M;
  a.M2 = M2;
  if ((rc = dalloc(S, &a.flags, 8))) return rc;
  // room for the per-block partials of the widest pass (Morse rows split
  // over kSplitLanes lanes)
  if ((rc = dalloc(S, &a.partials, static_cast<size_t>(S->nblocks) * kSplitLanes + 64))) return rc;
  if ((rc = dalloc(S, &a.partials_w, static_cast<size_t>(S->nblocks) * kSplitLanes + 64))) return rc;
  if ((rc = dalloc(S, &a.nhc, 64))) return rc;
  int32_t* cnt;
  if ((rc = dalloc(S, &cnt, round_up(N, 64)))) return rc;
  a.cnt = cnt;
  if ((rc = dalloc(S, &a.cntn, round_up(N, 64)))) return rc;
  // permutation scratch
  PermArgs& t = S->t;
  double** t6[6] = {&t.px, &t.py, &t.pz, &t.vx, &t.vy, &t.vz};
  for (auto* p : t6)
    if ((rc = dalloc(S, p, N))) return rc;
  void** tf[3] = {&t.fx, &t.fy, &t.fz};
  for (auto* p : tf) {
    char* q;
    if ((rc = dalloc(S, &q, N * fsz))) return rc;
    *p = q;
  }
  if ((rc = dalloc(S, &t.types, N))) return rc;
  if ((rc = dalloc(S, &t.orig, N))) return rc;
  GridArgs& g = S->g;
  int32_t* rank_of;
  if ((rc = dalloc(S, &rank_of, g.ncells))) return rc;
  g.rank_of = rank_of;
  if ((rc = dalloc(S, &g.ccount, g.ncells))) return rc;
  if ((rc = dalloc(S, &g.cstart, g.ncells))) return rc;
  if ((rc = dalloc(S, &g.cursor, g.ncells))) return rc;
  if ((rc = dalloc(S, &g.key, N))) return rc;
  if ((rc = dalloc(S, &g.sidx, N))) return rc;
  if ((rc = dalloc(S, &g.orig, N))) return rc;
  if ((rc = dalloc(S, &g.done, 4))) return rc;
  if ((rc = dalloc(S, &g.wsum, sched::kMaxScanGroups))) return rc;
  if ((rc = dalloc(S, &S->d_maxcnt, 4))) return rc;
  if ((rc = dalloc(S, &S->d_rebuilds, 1))) return rc;
  if ((rc = dalloc(S, &S->d_scalars, 16))) return rc;
  if ((rc = dalloc(S, &S->d_bar, 1))) return rc;
  TDMD_HIP_TRY(hipMemset(S->d_bar, 0, sizeof(GridBarrier)));
  // Morton rank of every cell of each box's grid; a box's cells take the
  // global ranks [cell_off, cell_off + ncells_box), so the sort keeps every
  // box in its own slot range.
  std::vector<int32_t> h_rank(static_cast<size_t>(g.ncells));
  const int nbox = S->multibox ? static_cast<int>(S->h_boxes.size()) : 1;
  for (int b = 0; b < nbox; ++b) {
    const int* n = S->multibox ? S->h_boxes[static_cast<size_t>(b)].n : g.n;
    const int off = S->multibox ? S->h_boxes[static_cast<size_t>(b)].cell_off : 0;
    const int nc = n[0] * n[1] * n[2];
    std::vector<std::pair<uint32_t, int>> order(static_cast<size_t>(nc));
    for (int z = 0, c = 0; z < n[2]; ++z)
      for (int y = 0; y < n[1]; ++y)
        for (int x = 0; x < n[0]; ++x, ++c)
          order[static_cast<size_t>(c)] = {spread3(x) | (spread3(y) << 1) | (spread3(z) << 2), c};
    std::sort(order.begin(), order.end());
    for (int r = 0; r < nc; ++r)
      h_rank[static_cast<size_t>(off + order[static_cast<size_t>(r)].second)] = off + r;
  }
  TDMD_HIP_TRY(hipMemcpy(rank_of, h_rank.data(), h_rank.size() * 4, hipMemcpyHostToDevice));
  // The list search's stencil: for each cell (row = its rank), the ranks of
  // the 27 cells around it with periodic wrap, in increasing rank, the order
  // in which the search visits them (atoms are sorted by cell rank, so rows
  // come out in increasing atom index), and each entry's periodic image:
  // +L where the stencil wrapped past the upper face, -L past the lower.
  // When every periodic axis has at least 3 cells no smaller than r_list,
  // a pair within r_list lies in exactly one stencil entry, whose image is
  // its minimum image; the search then skips the per-candidate test.
  std::vector<int32_t> h_around(27 * static_cast<size_t>(g.ncells));
  std::vector<uint8_t> h_img(27 * static_cast<size_t>(g.ncells));
  bool images = true;
  for (int b = 0; b < nbox; ++b) {
    const int* n = S->multibox ? S->h_boxes[static_cast<size_t>(b)].n : g.n;
    const int off = S->multibox ? S->h_boxes[static_cast<size_t>(b)].cell_off : 0;
    const MirrorParams& mp = S->multibox ? S->h_boxes[static_cast<size_t>(b)].mp : S->a.mp;
    bool box_images = true;
    for (int d = 0; d < 3; ++d)
      box_images = box_images && (!mp.periodic[d] || (n[d] >= 3 && mp.L[d] / n[d] >= S->r_list));
    if (S->multibox) S->h_boxes[static_cast<size_t>(b)].images = box_images ? 1 : 0;
    images = images && box_images;
    for (int z = 0; z < n[2]; ++z)
      for (int y = 0; y < n[1]; ++y)
        for (int x = 0; x < n[0]; ++x) {
          std::pair<int32_t, uint8_t> e[27];
          for (int t = 0; t < 27; ++t) {
            const int c[3] = {x + t % 3 - 1, y + (t / 3) % 3 - 1, z + t / 9 - 1};
            int w[3];
            unsigned code = 0;
            for (int d = 0; d < 3; ++d) {
              w[d] = (c[d] + n[d]) % n[d];
              if (mp.periodic[d] && c[d] >= n[d]) code |= 1u << (2 * d);  // +L
              if (mp.periodic[d] && c[d] < 0) code |= 2u << (2 * d);      // -L
            }
            e[t] = {h_rank[static_cast<size_t>(off + (w[2] * n[1] + w[1]) * n[0] + w[0])],
                    static_cast<uint8_t>(code)};
          }
          std::sort(e, e + 27);
          const int own = h_rank[static_cast<size_t>(off + (z * n[1] + y) * n[0] + x)];
          for (int t = 0; t < 27; ++t) {
            h_around[27 * static_cast<size_t>(own) + t] = e[t].first;
            h_img[27 * static_cast<size_t>(own) + t] = e[t].second;
          }
        }
  }
  int32_t* around;
  if ((rc = dalloc(S, &around, h_around.size()))) return rc;
  TDMD_HIP_TRY(hipMemcpy(around, h_around.data(), h_around.size() * 4, hipMemcpyHostToDevice));
  g.around = around;
  uint8_t* img;
  if ((rc = dalloc(S, &img, h_img.size()))) return rc;
  TDMD_HIP_TRY(hipMemcpy(img, h_img.data(), h_img.size(), hipMemcpyHostToDevice));
  g.around_img = img;
  g.images = images ? 1 : 0;  // batches: decided per box (BoxGeo::images)
  if (S->multibox) {
    BoxGeo* boxes;
    int32_t* rep;
    if ((rc = dalloc(S, &boxes, S->h_boxes.size()))) return rc;
    if ((rc = dalloc(S, &rep, N))) return rc;
    std::vector<int32_t> h_rep(static_cast<size_t>(N));
    for (size_t b = 0; b < S->h_boxes.size(); ++b)
      for (int i = S->rep_off[b]; i < S->rep_off[b + 1]; ++i) h_rep[static_cast<size_t>(i)] = static_cast<int32_t>(b);
    TDMD_HIP_TRY(hipMemcpy(boxes, S->h_boxes.data(), S->h_boxes.size() * sizeof(BoxGeo),
                           hipMemcpyHostToDevice));
    TDMD_HIP_TRY(hipMemcpy(rep, h_rep.data(), h_rep.size() * 4, hipMemcpyHostToDevice));
    a.boxes = boxes;
    a.rep = rep;
  }
  TDMD_HIP_TRY(hipMemset(g.ccount, 0, static_cast<size_t>(g.ncells) * 4));
  TDMD_HIP_TRY(hipMemset(g.cursor, 0, static_cast<size_t>(g.ncells) * 4));
  TDMD_HIP_TRY(hipMemset(g.done, 0, 16));
  TDMD_HIP_TRY(hipMemset(a.flags, 0, 32));
  TDMD_HIP_TRY(hipMemset(S->d_rebuilds, 0, 8));
  TDMD_HIP_TRY(hipMemset(a.nhc, 0, 64 * 8));
  const double one = 1.0;  // no pending chain scale
  TDMD_HIP_TRY(hipMemcpy(a.nhc + 24, &one, 8, hipMemcpyHostToDevice));
  TDMD_HIP_TRY(hipDeviceSynchronize());
  return TDMD_OK;
}

// Mirror parameters and binning grid of one box.
int setup_box(const tdmd_box& box, double r_list, double cell_min, MirrorParams* mp, int n[3],
              double inv_cs[3]) {
  for (int d = 0; d < 3; ++d) {
    mp->lo[d] = box.lo[d];
    mp->L[d] = box.hi[d] - box.lo[d];
    mp->hL[d] = 0.5 * mp->L[d];
    mp->periodic[d] = box.periodic[d] ? 1 : 0;
    if (!(mp->L[d] > 0)) return fail(TDMD_E_INVALID, "scheduler: box must have positive size");
  }
  // 32-bit fixed point (MirrorFixed): a periodic axis spans exactly 2^32
  // units; an open axis gets unit W 2^-30, W = max(L, r_list), so any listed
  // pair difference stays far below 2^31 units.
  for (int d = 0; d < 3; ++d) {
    const double W = mp->periodic[d] ? mp->L[d] : std::max(mp->L[d], r_list);
    const int e = mp->periodic[d] ? 32 : 30;
    mp->q[d] = std::ldexp(1.0, e) / W;
    mp->unit[d] = std::ldexp(W, -e);
  }
  // Binning grid: cells no smaller than cell_min >= r_list (with a relative
  // margin so rounding in the bin computation can never push a neighbour
  // out of the 27-cell stencil), at least 3 per axis (cell_list.cpp:22-27).
  for (int d = 0; d < 3; ++d) {
    n[d] = std::max(3, static_cast<int>(std::floor(mp->L[d] / (cell_min * (1.0 + 1e-9)))));
    inv_cs[d] = static_cast<double>(n[d]) / mp->L[d];
    if (n[d] > 1024) return fail(TDMD_E_INVALID, "scheduler: more than 1024 cells per axis");
  }
  return TDMD_OK;
}

// boxes.size() == 1 and !multibox: one system. multibox: a batch of
// independent boxes with counts[r] atoms each, concatenated in box order.
int sched_setup(tdmd_sched* S, const std::vector<tdmd_box>& boxes, const std::vector<int>& counts,
                const tdmd_sched_config& cfg, bool multibox) {
  int64_t total = 0;
  for (int c : counts) {
    if (c <= 0) return fail(TDMD_E_INVALID, "scheduler: natoms must be positive");
    total += c;
  }
  if (boxes.empty() || boxes.size() != counts.size() || total > (1LL << 30))
    return fail(TDMD_E_INVALID, "scheduler: bad box / atom counts");
  const int natoms = static_cast<int>(total);
  if (!(cfg.dt > 0) || !(cfg.r_skin >= 0) || cfg.rebuild_every < 1)
    return fail(TDMD_E_INVALID, "scheduler: bad dt / r_skin / rebuild_every");
  if (cfg.precision != TDMD_FP64 && cfg.precision != TDMD_MIXED)
    return fail(TDMD_E_INVALID, "scheduler: bad precision");
  if (cfg.t_target > 0 && (cfg.nhc_length < 1 || cfg.nhc_length > 8 || !(cfg.t_period > 0)))
    return fail(TDMD_E_INVALID, "scheduler: NHC chain length must be 1..8, t_period > 0");
  if (multibox && cfg.t_target > 0)
    return fail(TDMD_E_INVALID, "scheduler: a batch of boxes runs NVE (one thermostat per box "
                                "is not provided)");
  if (cfg.pair_unroll != 0 && cfg.pair_unroll != 4 && cfg.pair_unroll != 8)
    return fail(TDMD_E_INVALID, "scheduler: pair_unroll must be 0 (default), 4 or 8");
  if (cfg.lds_budget < 0) return fail(TDMD_E_INVALID, "scheduler: lds_budget must be >= 0");
  if (cfg.pair_lanes != 0 && cfg.pair_lanes != 1 && cfg.pair_lanes != 4 && cfg.pair_lanes != 8)
    return fail(TDMD_E_INVALID, "scheduler: pair_lanes must be 0 (by size), 1, 4 or 8");
  if (cfg.cell_pct != 0 && (cfg.cell_pct < 100 || cfg.cell_pct > 400))
    return fail(TDMD_E_INVALID, "scheduler: cell_pct must be 0 or 100..400");
  S->box = boxes[0];
  S->natoms = natoms;
  S->cfg = cfg;
  S->multibox = multibox;
  S->nblocks = round_up(grid_for(natoms), 8);
  S->r_list = S->cutoff + cfg.r_skin;
  GridArgs& g = S->g;
  const double cell_min = S->r_list * (cfg.cell_pct > 0 ? cfg.cell_pct : 100) / 100.0;
  int rc = setup_box(boxes[0], S->r_list, cell_min, &S->a.mp, g.n, g.inv_cs);
  if (rc) return rc;
  g.ncells = g.n[0] * g.n[1] * g.n[2];
  if (multibox) {
    S->h_boxes.resize(boxes.size());
    S->rep_off.assign(1, 0);
    int cells = 0;
    for (size_t b = 0; b < boxes.size(); ++b) {
      BoxGeo& bg = S->h_boxes[b];
      std::memset(&bg, 0, sizeof(bg));
      if ((rc = setup_box(boxes[b], S->r_list, cell_min, &bg.mp, bg.n, bg.inv_cs))) return rc;
      bg.cell_off = cells;
      cells += bg.n[0] * bg.n[1] * bg.n[2];
      S->rep_off.push_back(S->rep_off.back() + counts[b]);
    }
    g.ncells = cells;
  }
  S->a.natoms = natoms;
  S->a.ntypes = S->tables ? S->tables->ntypes : 1;
  if (S->a.ntypes < 1 || S->a.ntypes > kMaxTypes ||
      2 * S->a.ntypes + S->a.ntypes * (S->a.ntypes + 1) / 2 > kMaxSplines)
    return fail(TDMD_E_INVALID, "scheduler: 1..8 element types");
  S->a.nblocks = S->nblocks;
  S->a.dt = cfg.dt;
  S->a.skin_half_sq = 0.25 * cfg.r_skin * cfg.r_skin;
  S->nhcp.M = cfg.t_target > 0 ? cfg.nhc_length : 0;
  S->nhcp.n_dof = 3 * natoms - 3;  // pipeline_scheduler.cu:38
  S->nhcp.t_target = cfg.t_target;
  S->nhcp.t_period = cfg.t_period;
  S->nhcp.dt = cfg.dt;
  TDMD_HIP_TRY(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
  rc = sched_alloc(S);
  if (rc) return rc;
  const bool fixed = cfg.precision == TDMD_MIXED && !cfg.mirror_fp64;
  if (multibox) {
    if (cfg.precision == TDMD_FP64)
      S->eng.reset(new Engine<double, ManyBoxes<MirrorF64>>(S));
    else if (fixed)
      S->eng.reset(new Engine<float, ManyBoxes<MirrorFixed>>(S));
    else
      S->eng.reset(new Engine<float, ManyBoxes<MirrorF64>>(S));
  } else if (cfg.precision == TDMD_FP64) {
    S->eng.reset(new Engine<double, MirrorF64>(S));
  } else if (fixed) {
    S->eng.reset(new Engine<float, MirrorFixed>(S));
  } else {
    S->eng.reset(new Engine<float, MirrorF64>(S));
  }
  return S->eng->init();
}

int sched_setup(tdmd_sched* S, const tdmd_box& box, int natoms, const tdmd_sched_config& cfg) {
  return sched_setup(S, std::vector<tdmd_box>{box}, std::vector<int>{natoms}, cfg, false);
}

void sched_free(tdmd_sched* S) {
  if (S->stream) (void)hipStreamSynchronize(S->stream);
  drop_graphs(S);
  for (void* p : S->allocs) (void)hipFree(p);
  if (S->a.nbr) (void)hipFree(const_cast<int32_t*>(S->a.nbr));
  if (S->a.nbr16) (void)hipFree(const_cast<uint16_t*>(S->a.nbr16));
  if (S->a.nbrc16) (void)hipFree(S->a.nbrc16);
  if (S->tables) eam_tables_destroy(S->tables);
  if (S->stream) (void)hipStreamDestroy(S->stream);
}

int check_flags(tdmd_sched* S) {
  int32_t f[4];
  TDMD_HIP_TRY(hipMemcpy(f, S->a.flags, 16, hipMemcpyDeviceToHost));
  if (f[2])
    return fail(TDMD_E_NONFINITE,
                "scheduler: non-finite position, or an atom outside the representable range "
                "of an open axis (lo - 2W .. lo + 2W, W = max(box length, r_list))");
  if (f[3]) {
    // The rebuild stopped between phases: the atoms may be half permuted and
    // the barrier's arrival counters hold the arrivals of workgroups that
    // returned. The kernel has ended (the stream was synchronised), so the
    // counters are cleared here; the state is unusable until the next upload.
    TDMD_HIP_TRY(hipMemset(S->d_bar, 0, sizeof(GridBarrier)));
    TDMD_HIP_TRY(hipMemset(S->a.flags, 0, 32));
    S->uploaded = false;
    S->initial_forces = false;
    return fail(TDMD_E_HIP, "scheduler: fused rebuild grid barrier timed out (workgroups not "
                            "co-resident); upload again, or use split_rebuild");
  }
  if (f[1] > S->a.cap)
    return fail(TDMD_E_OVERFLOW, "scheduler: neighbour capacity " + std::to_string(S->a.cap) +
                                     " exceeded (max count " + std::to_string(f[1]) + ")");
  return TDMD_OK;
}

// First neighbour build: sort, count exactly, size the list, fill.
int first_build(tdmd_sched* S) {
  hipStream_t s = S->stream;
  k_set_flag<<<1, 64, 0, s>>>(S->a.flags);
  TDMD_LAUNCH_CHECK();
  int rc = S->eng->sort_pass(s);
  if (rc) return rc;
  TDMD_HIP_TRY(hipMemsetAsync(S->d_maxcnt, 0, 4, s));
  if ((rc = S->eng->count_pass(s, S->d_maxcnt))) return rc;
  int32_t maxc = 0;
  TDMD_HIP_TRY(hipMemcpyAsync(&maxc, S->d_maxcnt, 4, hipMemcpyDeviceToHost, s));
  TDMD_HIP_TRY(hipStreamSynchronize(s));
  int cap = S->cfg.max_neighbors > 0 ? S->cfg.max_neighbors
                                     : round_up(static_cast<int>(std::ceil(maxc * 1.25)) + 8, 4);
  cap = round_up(std::max(cap, kRowPad), kRowPad);  // rows are padded to kRowPad slots
  if (S->a.nbr == nullptr || cap > S->a.cap) {
    if (S->a.nbr) TDMD_HIP_TRY(hipFree(const_cast<int32_t*>(S->a.nbr)));
    if (S->a.nbr16) TDMD_HIP_TRY(hipFree(const_cast<uint16_t*>(S->a.nbr16)));
    if (S->a.nbrc16) TDMD_HIP_TRY(hipFree(S->a.nbrc16));
    S->a.nbr = nullptr;
    S->a.nbr16 = nullptr;
    S->a.nbrc16 = nullptr;
    const size_t slots = static_cast<size_t>(round_up(S->natoms, 64)) * cap;
    void* p = nullptr;
    if ((rc = tdmd_malloc(&p, slots * 4))) return rc;
    S->a.nbr = static_cast<int32_t*>(p);
    if (S->stage_ok) {  // 16-bit rows of the staged passes, full and cut
      if ((rc = tdmd_malloc(&p, slots * 2))) return rc;
      S->a.nbr16 = static_cast<uint16_t*>(p);
      if ((rc = tdmd_malloc(&p, slots * 2))) return rc;
      S->a.nbrc16 = static_cast<uint16_t*>(p);
    }
    S->a.cap = cap;
    drop_graphs(S);
  }
  rc = S->eng->fill_pass(s);
  if (rc) return rc;
  ++S->host_rebuilds;
  return TDMD_OK;
}

int launch_ops(tdmd_sched* S, const std::string& ops, hipStream_t s) {
  for (char op : ops) {
    const int rc = S->eng->launch(op, s);
    if (rc) return rc;
  }
  return TDMD_OK;
}

// Enqueue a sequence of ops, as one cached graph when enabled.
int run_ops(tdmd_sched* S, const std::string& ops) {
  if (ops.empty()) return TDMD_OK;
  hipStream_t s = S->stream;
  if (!S->cfg.use_graph) return launch_ops(S, ops, s);
  // a graph holds the record buffers of the state it was captured in (and
  // the pass clock, when pass timing is on)
  const std::string key = ops + (S->rec_swapped ? "/1" : "/0") + (S->pass_timing ? "t" : "");
  auto it = S->graphs.find(key);
  if (it == S->graphs.end()) {
    const int64_t saved = S->launches;
    const bool swapped0 = S->rec_swapped;
    TDMD_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    {
      const int rc = launch_ops(S, ops, s);
      if (rc) {
        hipGraph_t dummy;
        (void)hipStreamEndCapture(s, &dummy);
        return rc;
      }
    }
    hipGraph_t gdef;
    TDMD_HIP_TRY(hipStreamEndCapture(s, &gdef));
    hipGraphExec_t exec;
    TDMD_HIP_TRY(hipGraphInstantiate(&exec, gdef, nullptr, nullptr, 0));
    S->graph_defs[key] = gdef;
    S->graph_kernels[key] = S->launches - saved;
    S->graph_swaps[key] = S->rec_swapped != swapped0;
    // capture enqueued nothing: back to the state before it, then replay
    if (S->graph_swaps[key]) {
      std::swap(S->a.M, S->a.M2);
      S->rec_swapped = swapped0;
    }
    it = S->graphs.emplace(key, exec).first;
    S->launches = saved;
  }
  S->launches += S->graph_kernels[key];  // kernels the graph replays
  if (S->graph_swaps[key]) {  // the replay leaves the records in the other buffer
    std::swap(S->a.M, S->a.M2);
    S->rec_swapped = !S->rec_swapped;
  }
  TDMD_HIP_TRY(hipGraphLaunch(it->second, s));
  ++S->graph_launches;
  return TDMD_OK;
}

int compute_initial_forces(tdmd_sched* S) {
  // fast_pipeline_scheduler.cu:168-193: build the list, compute F(t=0).
  int rc = first_build(S);
  if (rc) return rc;
  S->steps_since = 0;
  hipStream_t s = S->stream;
  if (S->kind == 1) {
    if ((rc = S->eng->launch(OP_DENS, s))) return rc;
  }
  if ((rc = S->eng->launch(OP_F_ONLY, s))) return rc;
  if (S->nhcp.M > 0) {
    // Stored KE of the uploaded velocities (the chain state was reset by
    // upload, with no pending scale).
    k_ke_partials<double><<<grid_for(S->natoms), kBlock, 0, s>>>(S->a);
    k_nhc_set_ke<<<1, 256, 0, s>>>(S->a.nhc, S->a.partials, grid_for(S->natoms), 0);
    TDMD_LAUNCH_CHECK();
  }
  S->initial_forces = true;
  return TDMD_OK;
}

constexpr int kChunk = 16;  // steps per captured graph

int run_window(tdmd_sched* S, int n) {
  // Step ops; the rebuild test happens after the kick-drift of each step
  // (fast_pipeline_scheduler.cu:101-166 + maybe_rebuild_nlist :225-232).
  const bool nvt = S->nhcp.M > 0;
  const bool eam = S->kind == 1;
  const bool disp = S->cfg.rebuild_policy == TDMD_REBUILD_DISPLACEMENT;
  std::vector<std::string> steps;
  steps.reserve(static_cast<size_t>(n));
  for (int k = 1; k <= n; ++k) {
    std::string st;
    if (nvt) {
      // later steps: the chain's pre half-step is fused into 'N', and the
      // previous step's closing half kick into this kick-drift
      if (k == 1) st += static_cast<char>(OP_NVT_PRE);
      st += static_cast<char>(k == 1 ? OP_KD_NVT : OP_KD_NVT2);
    } else if (k == 1) {
      st += static_cast<char>(OP_BOOT);
    }  // NVE: later steps' kick-drift is the previous force pass's epilogue
    if (disp) {
      st += static_cast<char>(OP_REB_COND);
    } else if (S->steps_since >= S->cfg.rebuild_every) {
      st += static_cast<char>(OP_REB_FORCED);
      S->steps_since = 0;
      ++S->host_rebuilds;
    }
    if (eam) st += static_cast<char>(OP_DENS);
    if (nvt) {
      // the window's last force pass stores the kicked velocities (the
      // window end scales them; the next window starts from them)
      st += static_cast<char>(k < n ? OP_F_KE : OP_F_KICK2);
      st += static_cast<char>(k < n ? OP_NVT_POST_PRE : OP_NVT_POST);
    } else if (k == n) {
      st += static_cast<char>(OP_F_HALF);
    } else {
      st += static_cast<char>(OP_F_FULL);
    }
    if (nvt && k == n) st += static_cast<char>(OP_NVT_FINAL);
    steps.push_back(st);
    ++S->steps_since;
    ++S->current_step;
    ++S->ticks;
  }
  for (size_t k = 0; k < steps.size(); k += kChunk) {
    std::string ops;
    for (size_t q = k; q < std::min(steps.size(), k + kChunk); ++q) ops += steps[q];
    const int rc = run_ops(S, ops);
    if (rc) return rc;
  }
  return TDMD_OK;
}

}  // namespace
}  // namespace tdmd_hip

using namespace tdmd_hip;

extern "C" {

void tdmd_sched_config_default(tdmd_sched_config* cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->dt = 0.001;  // FastPipelineConfig defaults, fast_pipeline_scheduler.cuh:42-46
  cfg->r_skin = 1.0;
  cfg->rebuild_every = 10;
  cfg->precision = TDMD_MIXED;
  cfg->rebuild_policy = TDMD_REBUILD_FIXED;
  cfg->t_target = 0.0;
  cfg->t_period = 0.1;
  cfg->nhc_length = 3;
  cfg->use_graph = 1;
  cfg->mirror_fp64 = 0;
}

int tdmd_sched_create_morse(tdmd_sched** out, const tdmd_box* box, int32_t natoms,
                            const tdmd_morse_params* params, const tdmd_sched_config* cfg) {
  if (!out || !box || !params || !cfg) return fail(TDMD_E_INVALID, "sched_create_morse: null");
  if (!(params->rc > 0)) return fail(TDMD_E_INVALID, "sched_create_morse: rc must be positive");
  auto* S = new tdmd_sched();
  S->kind = 0;
  S->morse = *params;
  S->cutoff = params->rc;  // fast_pipeline_scheduler.cu:23 (cutoff_ = params.rc)
  const int rc = sched_setup(S, *box, natoms, *cfg);
  if (rc) {
    sched_free(S);
    delete S;
    return rc;
  }
  *out = S;
  return TDMD_OK;
}

int tdmd_sched_create_eam(tdmd_sched** out, const tdmd_box* box, int32_t natoms, int32_t ntypes,
                          int32_t nrho, double drho, int32_t nr, double dr, double cutoff,
                          const double* F, const double* rho, const double* rphi,
                          const tdmd_sched_config* cfg) {
  if (!out || !box || !cfg || !F || !rho || !rphi)
    return fail(TDMD_E_INVALID, "sched_create_eam: null");
  if (ntypes < 1 || ntypes > sched::kMaxTypes)
    return fail(TDMD_E_INVALID, "sched_create_eam: 1..8 element types");
  auto* S = new tdmd_sched();
  S->kind = 1;
  int rc = tdmd_eam_tables_create(&S->tables, cfg->precision, ntypes, nrho, drho, nr, dr, cutoff,
                                  F, rho, rphi);
  if (rc) {
    delete S;
    return rc;
  }
  S->cutoff = cutoff;  // fast_pipeline_scheduler.cu:35 (cutoff_ = eam.cutoff())
  rc = sched_setup(S, *box, natoms, *cfg);
  if (rc) {
    sched_free(S);
    delete S;
    return rc;
  }
  *out = S;
  return TDMD_OK;
}

int tdmd_sched_create_eam_tables(tdmd_sched** out, const tdmd_box* box, int32_t natoms,
                                 const tdmd_eam_tables* tables, const tdmd_sched_config* cfg) {
  if (!out || !box || !cfg || !tables)
    return fail(TDMD_E_INVALID, "sched_create_eam_tables: null");
  if (tables->ntypes < 1 || tables->ntypes > sched::kMaxTypes)
    return fail(TDMD_E_INVALID, "sched_create_eam_tables: 1..8 element types");
  auto* S = new tdmd_sched();
  S->kind = 1;
  int rc = eam_tables_clone(tables, cfg->precision, &S->tables);
  if (rc) {
    delete S;
    return rc;
  }
  S->cutoff = tables->cutoff;
  rc = sched_setup(S, *box, natoms, *cfg);
  if (rc) {
    sched_free(S);
    delete S;
    return rc;
  }
  *out = S;
  return TDMD_OK;
}

int tdmd_sched_create_eam_batch(tdmd_sched** out, int32_t nrep, const tdmd_box* boxes,
                                const int32_t* natoms, int32_t ntypes, int32_t nrho, double drho,
                                int32_t nr, double dr, double cutoff, const double* F,
                                const double* rho, const double* rphi,
                                const tdmd_sched_config* cfg) {
  if (!out || !boxes || !natoms || !cfg || !F || !rho || !rphi || nrep < 1)
    return fail(TDMD_E_INVALID, "sched_create_eam_batch: bad arguments");
  if (ntypes < 1 || ntypes > sched::kMaxTypes)
    return fail(TDMD_E_INVALID, "sched_create_eam_batch: 1..8 element types");
  auto* S = new tdmd_sched();
  S->kind = 1;
  int rc = tdmd_eam_tables_create(&S->tables, cfg->precision, ntypes, nrho, drho, nr, dr, cutoff,
                                  F, rho, rphi);
  if (rc) {
    delete S;
    return rc;
  }
  S->cutoff = cutoff;
  rc = sched_setup(S, std::vector<tdmd_box>(boxes, boxes + nrep),
                   std::vector<int>(natoms, natoms + nrep), *cfg, true);
  if (rc) {
    sched_free(S);
    delete S;
    return rc;
  }
  *out = S;
  return TDMD_OK;
}

int tdmd_sched_create_morse_batch(tdmd_sched** out, int32_t nrep, const tdmd_box* boxes,
                                  const int32_t* natoms, const tdmd_morse_params* params,
                                  const tdmd_sched_config* cfg) {
  if (!out || !boxes || !natoms || !params || !cfg || nrep < 1)
    return fail(TDMD_E_INVALID, "sched_create_morse_batch: bad arguments");
  if (!(params->rc > 0)) return fail(TDMD_E_INVALID, "sched_create_morse_batch: rc must be > 0");
  auto* S = new tdmd_sched();
  S->kind = 0;
  S->morse = *params;
  S->cutoff = params->rc;
  const int rc = sched_setup(S, std::vector<tdmd_box>(boxes, boxes + nrep),
                             std::vector<int>(natoms, natoms + nrep), *cfg, true);
  if (rc) {
    sched_free(S);
    delete S;
    return rc;
  }
  *out = S;
  return TDMD_OK;
}

int tdmd_sched_autotune(tdmd_sched* S, int32_t iters, int32_t out[2]) {
  if (!S || !out || iters < 1) return fail(TDMD_E_INVALID, "sched_autotune: bad arguments");
  if (!S->uploaded) return fail(TDMD_E_STATE, "sched_autotune: upload first");
  if (!S->initial_forces) {
    const int rc = compute_initial_forces(S);
    if (rc) return rc;
  }
  const int rc = S->eng->tune(iters, out);
  if (rc) return rc;
  // captured graphs hold the previous kernel choice
  TDMD_HIP_TRY(hipStreamSynchronize(S->stream));
  drop_graphs(S);
  S->cfg.pair_unroll = out[0];
  S->cfg.lds_budget = out[1];
  S->cfg.pair_lanes = S->lanes;
  return check_flags(S);
}

int tdmd_sched_kernel_shape(const tdmd_sched* S, int32_t out[3]) {
  if (!S || !out || !S->eng) return fail(TDMD_E_INVALID, "sched_kernel_shape: null");
  S->eng->shape(out);
  return TDMD_OK;
}

int tdmd_sched_destroy(tdmd_sched* s) {
  if (!s) return TDMD_OK;
  sched_free(s);
  delete s;
  return TDMD_OK;
}

int tdmd_sched_upload(tdmd_sched* S, const double* pos, const double* vel, const double* forces,
                      const int32_t* types, const int32_t* ids, const double* masses,
                      int32_t n_masses) {
  if (!S || !pos || !vel || !types || !masses || n_masses < 2 || n_masses > 64)
    return fail(TDMD_E_INVALID, "sched_upload: bad arguments (masses indexed by 1-based type)");
  const int N = S->natoms;
  const int ntypes_max = S->kind == 1 ? S->tables->ntypes : n_masses - 1;
  for (int i = 0; i < N; ++i)
    if (types[i] < 1 || types[i] > ntypes_max || types[i] >= n_masses)
      return fail(TDMD_E_INVALID, "sched_upload: atom type out of range at index " +
                                      std::to_string(i));
  for (int k = 1; k < n_masses; ++k)
    if (!(masses[k] > 0)) return fail(TDMD_E_INVALID, "sched_upload: masses must be positive");
  S->h_types.assign(types, types + N);
  S->h_ids.assign(N, 0);
  if (ids) S->h_ids.assign(ids, ids + N);
  S->h_masses.assign(masses, masses + n_masses);
  const bool mixed = S->cfg.precision == TDMD_MIXED;
  std::vector<double> buf(static_cast<size_t>(N));
  std::vector<float> fbuf(static_cast<size_t>(N));
  double* pd[3] = {S->a.px, S->a.py, S->a.pz};
  double* vd[3] = {S->a.vx, S->a.vy, S->a.vz};
  void* fd[3] = {S->a.fx, S->a.fy, S->a.fz};
  for (int d = 0; d < 3; ++d) {
    for (int i = 0; i < N; ++i) buf[static_cast<size_t>(i)] = pos[3 * i + d];
    TDMD_HIP_TRY(hipMemcpy(pd[d], buf.data(), N * 8, hipMemcpyHostToDevice));
    for (int i = 0; i < N; ++i) buf[static_cast<size_t>(i)] = vel[3 * i + d];
    TDMD_HIP_TRY(hipMemcpy(vd[d], buf.data(), N * 8, hipMemcpyHostToDevice));
    for (int i = 0; i < N; ++i) buf[static_cast<size_t>(i)] = forces ? forces[3 * i + d] : 0.0;
    if (mixed) {
      for (int i = 0; i < N; ++i) fbuf[static_cast<size_t>(i)] = static_cast<float>(buf[static_cast<size_t>(i)]);
      TDMD_HIP_TRY(hipMemcpy(fd[d], fbuf.data(), N * 4, hipMemcpyHostToDevice));
    } else {
      TDMD_HIP_TRY(hipMemcpy(fd[d], buf.data(), N * 8, hipMemcpyHostToDevice));
    }
  }
  TDMD_HIP_TRY(hipMemcpy(const_cast<int32_t*>(S->a.types), types, N * 4, hipMemcpyHostToDevice));
  // captured graphs hold the kernel arguments, this flag among them
  const int one_type = std::all_of(types, types + N, [](int32_t t) { return t == 1; }) ? 1 : 0;
  if (one_type != S->a.one_type) {
    TDMD_HIP_TRY(hipStreamSynchronize(S->stream));
    drop_graphs(S);
    S->a.one_type = one_type;
  }
  std::vector<double> m64(64, 1.0);
  std::copy(masses, masses + n_masses, m64.begin());
  TDMD_HIP_TRY(hipMemcpy(const_cast<double*>(S->a.masses), m64.data(), 64 * 8, hipMemcpyHostToDevice));
  std::vector<int32_t> iota(static_cast<size_t>(N));
  std::iota(iota.begin(), iota.end(), 0);
  TDMD_HIP_TRY(hipMemcpy(S->g.orig, iota.data(), N * 4, hipMemcpyHostToDevice));
  TDMD_HIP_TRY(hipMemset(S->a.flags, 0, 32));
  if (S->d_bar) TDMD_HIP_TRY(hipMemset(S->d_bar, 0, sizeof(GridBarrier)));
  std::vector<double> nhc0(64, 0.0);  // fresh chain, no pending scale
  nhc0[24] = 1.0;
  TDMD_HIP_TRY(hipMemcpy(S->a.nhc, nhc0.data(), 64 * 8, hipMemcpyHostToDevice));
  TDMD_HIP_TRY(hipDeviceSynchronize());
  const int rc = S->eng->write_mirror(S->stream);
  if (rc) return rc;
  TDMD_HIP_TRY(hipStreamSynchronize(S->stream));
  // fast_pipeline_scheduler.cu:70-72
  S->uploaded = true;
  S->initial_forces = false;
  S->current_step = 0;
  S->steps_since = 0;
  return TDMD_OK;
}

int tdmd_sched_run_until(tdmd_sched* S, int32_t target_step) {
  if (!S) return fail(TDMD_E_INVALID, "sched_run_until: null");
  if (!S->uploaded) return fail(TDMD_E_STATE, "sched_run_until: upload() first");
  int rc;
  if (!S->initial_forces && (rc = compute_initial_forces(S))) return rc;
  if (S->current_step < target_step) {
    rc = run_window(S, target_step - S->current_step);
    if (rc) return rc;
  }
  TDMD_HIP_TRY(hipStreamSynchronize(S->stream));
  return check_flags(S);
}

int tdmd_sched_download(const tdmd_sched* S, double* pos, double* vel, double* forces,
                        int32_t* types, int32_t* ids, int32_t natoms) {
  if (!S || natoms != S->natoms) return fail(TDMD_E_INVALID, "sched_download: natoms mismatch");
  if (!S->uploaded) return fail(TDMD_E_STATE, "sched_download: nothing uploaded");
  const int N = S->natoms;
  TDMD_HIP_TRY(hipStreamSynchronize(S->stream));
  std::vector<int32_t> orig(static_cast<size_t>(N));
  TDMD_HIP_TRY(hipMemcpy(orig.data(), S->g.orig, N * 4, hipMemcpyDeviceToHost));
  std::vector<double> buf(static_cast<size_t>(N));
  std::vector<float> fbuf(static_cast<size_t>(N));
  const double* pd[3] = {S->a.px, S->a.py, S->a.pz};
  const double* vd[3] = {S->a.vx, S->a.vy, S->a.vz};
  const void* fd[3] = {S->a.fx, S->a.fy, S->a.fz};
  const bool mixed = S->cfg.precision == TDMD_MIXED;
  for (int d = 0; d < 3; ++d) {
    if (pos) {
      TDMD_HIP_TRY(hipMemcpy(buf.data(), pd[d], N * 8, hipMemcpyDeviceToHost));
      for (int s = 0; s < N; ++s) pos[3 * orig[static_cast<size_t>(s)] + d] = buf[static_cast<size_t>(s)];
    }
    if (vel) {
      TDMD_HIP_TRY(hipMemcpy(buf.data(), vd[d], N * 8, hipMemcpyDeviceToHost));
      for (int s = 0; s < N; ++s) vel[3 * orig[static_cast<size_t>(s)] + d] = buf[static_cast<size_t>(s)];
    }
    if (forces) {
      if (mixed) {
        TDMD_HIP_TRY(hipMemcpy(fbuf.data(), fd[d], N * 4, hipMemcpyDeviceToHost));
        for (int s = 0; s < N; ++s)
          forces[3 * orig[static_cast<size_t>(s)] + d] = fbuf[static_cast<size_t>(s)];
      } else {
        TDMD_HIP_TRY(hipMemcpy(buf.data(), fd[d], N * 8, hipMemcpyDeviceToHost));
        for (int s = 0; s < N; ++s)
          forces[3 * orig[static_cast<size_t>(s)] + d] = buf[static_cast<size_t>(s)];
      }
    }
  }
  if (types) std::copy(S->h_types.begin(), S->h_types.end(), types);
  if (ids) std::copy(S->h_ids.begin(), S->h_ids.end(), ids);
  return TDMD_OK;
}

int tdmd_sched_stats(const tdmd_sched* S, int64_t out[4]) {
  if (!S || !out) return fail(TDMD_E_INVALID, "sched_stats: null");
  int64_t dev_rebuilds = 0;
  TDMD_HIP_TRY(hipMemcpy(&dev_rebuilds, S->d_rebuilds, 8, hipMemcpyDeviceToHost));
  out[0] = S->ticks;
  out[1] = S->launches;
  out[2] = dev_rebuilds;
  out[3] = S->graph_launches;
  return TDMD_OK;
}

int tdmd_sched_current_step(const tdmd_sched* S, int32_t* step) {
  if (!S || !step) return fail(TDMD_E_INVALID, "sched_current_step: null");
  *step = S->current_step;
  return TDMD_OK;
}

int tdmd_sched_interaction_cutoff(const tdmd_sched* S, double* rc) {
  if (!S || !rc) return fail(TDMD_E_INVALID, "sched_interaction_cutoff: null");
  *rc = S->cutoff;
  return TDMD_OK;
}

int tdmd_sched_time_passes(tdmd_sched* S, int32_t steps, float* ms_out) {
  // A window of `steps` steps with HIP events around each pass of the
  // steady-state step: EAM NVE -> [density, force + kick-drift], EAM NVT ->
  // [density, force + KE partials] (the chain half-steps and the scaled
  // kick-drift run between them, untimed), Morse -> [kick-drift,
  // force]. No rebuild test inside the window.
  if (!S || !ms_out || steps < 2) return fail(TDMD_E_INVALID, "sched_time_passes: bad args");
  if (!S->uploaded) return fail(TDMD_E_STATE, "sched_time_passes: upload() first");
  const bool nvt = S->nhcp.M > 0;
  const bool eam = S->kind == 1;
  if (nvt && !eam) return fail(TDMD_E_STATE, "sched_time_passes: NVT timing is for EAM");
  int rc;
  if (!S->initial_forces && (rc = compute_initial_forces(S))) return rc;
  hipStream_t s = S->stream;
  const char a_op = eam ? OP_DENS : OP_KD_FULL;
  const char b_op = nvt ? OP_F_KE : eam ? OP_F_FULL : OP_F_ONLY;
  // Events are stamped by the kernel dispatches themselves
  // (hipExtLaunchKernel), so each interval is one kernel's execution.
  std::vector<hipEvent_t> ev(static_cast<size_t>(4 * steps));
  for (auto& e : ev) TDMD_HIP_TRY(hipEventCreate(&e));
  if (nvt) {
    if ((rc = S->eng->launch(OP_NVT_PRE, s))) return rc;
  } else {
    if ((rc = S->eng->launch(OP_BOOT, s))) return rc;
    if (!eam && (rc = S->eng->launch(OP_F_ONLY, s))) return rc;
  }
  for (int k = 0; k < steps - 1; ++k) {
    if (nvt && (rc = S->eng->launch(k == 0 ? OP_KD_NVT : OP_KD_NVT2, s))) return rc;
    S->eng->time_next(ev[4 * k], ev[4 * k + 1]);
    if ((rc = S->eng->launch(a_op, s))) return rc;
    S->eng->time_next(ev[4 * k + 2], ev[4 * k + 3]);
    if ((rc = S->eng->launch(b_op, s))) return rc;
    S->eng->time_next(nullptr, nullptr);
    if (nvt && (rc = S->eng->launch(OP_NVT_POST_PRE, s))) return rc;
  }
  if (nvt) {
    for (char op : {OP_KD_NVT2, OP_DENS, OP_F_KICK2, OP_NVT_POST, OP_NVT_FINAL})
      if ((rc = S->eng->launch(op, s))) return rc;
  } else {
    if (eam && (rc = S->eng->launch(OP_DENS, s))) return rc;
    if (!eam && (rc = S->eng->launch(OP_KD_FULL, s))) return rc;
    if ((rc = S->eng->launch(OP_F_HALF, s))) return rc;
  }
  TDMD_HIP_TRY(hipStreamSynchronize(s));
  double ta = 0, tb = 0;
  for (int k = 0; k < steps - 1; ++k) {
    float x, y;
    TDMD_HIP_TRY(hipEventElapsedTime(&x, ev[4 * k], ev[4 * k + 1]));
    TDMD_HIP_TRY(hipEventElapsedTime(&y, ev[4 * k + 2], ev[4 * k + 3]));
    ta += x;
    tb += y;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  ms_out[0] = static_cast<float>(ta / (steps - 1));
  ms_out[1] = static_cast<float>(tb / (steps - 1));
  S->current_step += steps;
  S->steps_since += steps;
  S->ticks += steps;
  return check_flags(S);
}

int tdmd_sched_pass_timing(tdmd_sched* S, int32_t on) {
  if (!S) return fail(TDMD_E_INVALID, "sched_pass_timing: null");
  if (on && !S->d_pclock) {
    int dev = 0;
    TDMD_HIP_TRY(hipGetDevice(&dev));
    TDMD_HIP_TRY(hipDeviceGetAttribute(&S->clock_khz, hipDeviceAttributeWallClockRate, dev));
    if (S->clock_khz <= 0) return fail(TDMD_E_HIP, "sched_pass_timing: no device wall clock");
    int rc = dalloc(S, &S->d_pclock, 8);
    if (rc) return rc;
  }
  TDMD_HIP_TRY(hipStreamSynchronize(S->stream));
  if (on) {
    const unsigned long long init[8] = {~0ull, 0, 0, 0, ~0ull, 0, 0, 0};
    TDMD_HIP_TRY(hipMemcpy(S->d_pclock, init, sizeof(init), hipMemcpyHostToDevice));
  }
  S->pass_timing = on != 0;
  S->a.pclock = on ? S->d_pclock : nullptr;
  return TDMD_OK;
}

int tdmd_sched_pass_times(tdmd_sched* S, float* ms_out) {
  if (!S || !ms_out) return fail(TDMD_E_INVALID, "sched_pass_times: null");
  if (!S->pass_timing) return fail(TDMD_E_STATE, "sched_pass_times: pass timing is off");
  TDMD_HIP_TRY(hipStreamSynchronize(S->stream));
  unsigned long long w[8];
  TDMD_HIP_TRY(hipMemcpy(w, S->d_pclock, sizeof(w), hipMemcpyDeviceToHost));
  if (w[7] == 0) return fail(TDMD_E_STATE, "sched_pass_times: no force pass ran since timing on");
  for (int p = 0; p < 2; ++p)
    ms_out[p] = w[4 * p + 3] ? static_cast<float>(static_cast<double>(w[4 * p + 2]) /
                                                  static_cast<double>(w[4 * p + 3]) / S->clock_khz)
                             : 0.0f;
  return TDMD_OK;
}

int tdmd_sched_halo_info(tdmd_sched* S, int64_t out[4]) {
  if (!S || !out) return fail(TDMD_E_INVALID, "sched_halo_info: null");
  for (int k = 0; k < 4; ++k) out[k] = 0;
  if (S->a.stage_cap == 0) return TDMD_OK;
  const int nb = (S->natoms + 1023) / 1024;
  std::vector<int32_t> hn(static_cast<size_t>(nb)), mode(static_cast<size_t>(nb));
  TDMD_HIP_TRY(hipStreamSynchronize(S->stream));
  TDMD_HIP_TRY(hipMemcpy(hn.data(), S->a.hn, hn.size() * 4, hipMemcpyDeviceToHost));
  TDMD_HIP_TRY(hipMemcpy(mode.data(), S->a.hmode, mode.size() * 4, hipMemcpyDeviceToHost));
  out[0] = nb;
  for (int b = 0; b < nb; ++b) {
    out[1] += mode[static_cast<size_t>(b)] == 0;
    out[2] = std::max<int64_t>(out[2], hn[static_cast<size_t>(b)]);
    out[3] += hn[static_cast<size_t>(b)];
  }
  return TDMD_OK;
}

int tdmd_sched_stream(const tdmd_sched* S, void** stream) {
  if (!S || !stream) return fail(TDMD_E_INVALID, "sched_stream: null");
  *stream = S->stream;
  return TDMD_OK;
}

int tdmd_sched_observables(tdmd_sched* S, tdmd_observables* o) {
  if (!S || !o) return fail(TDMD_E_INVALID, "sched_observables: null");
  if (!S->uploaded) return fail(TDMD_E_STATE, "sched_observables: upload() first");
  double th[6];
  int rc = tdmd_sched_thermo(S, th);
  if (rc) return rc;
  double sc[4];
  TDMD_HIP_TRY(hipMemcpy(sc, S->d_scalars, 32, hipMemcpyDeviceToHost));
  double vol = 0.0;
  if (S->multibox) {
    for (const auto& b : S->h_boxes) vol += b.mp.L[0] * b.mp.L[1] * b.mp.L[2];
  } else {
    vol = S->a.mp.L[0] * S->a.mp.L[1] * S->a.mp.L[2];
  }
  o->step = S->current_step;
  o->natoms = S->natoms;
  o->pe = th[0];
  o->ke = th[1];
  o->te = th[0] + th[1];
  o->temp = th[3];
  o->nhc_energy = th[2];
  o->conserved = o->te + th[2];
  o->virial = sc[3];
  o->volume = vol;
  // P = (2 KE + W) / (3 V), eV/A^3 -> bar (1 eV/A^3 = 1.602176634e6 bar)
  o->pressure = (2.0 * th[1] + sc[3]) / (3.0 * vol) * 1.602176634e6;
  o->max_neighbors = static_cast<int32_t>(th[4]);
  o->mean_neighbors = th[5];
  std::vector<int32_t> nn(static_cast<size_t>(S->natoms));
  TDMD_HIP_TRY(hipMemcpy(nn.data(), S->a.cntn, nn.size() * 4, hipMemcpyDeviceToHost));
  int64_t sn = 0;
  for (int32_t v : nn) sn += v;
  o->mean_near_neighbors = static_cast<double>(sn) / S->natoms;
  return TDMD_OK;
}

int tdmd_sched_thermo(tdmd_sched* S, double out[6]) {
  if (!S || !out) return fail(TDMD_E_INVALID, "sched_thermo: null");
  if (!S->uploaded) return fail(TDMD_E_STATE, "sched_thermo: upload() first");
  int rc;
  if (!S->initial_forces && (rc = compute_initial_forces(S))) return rc;
  if ((rc = S->eng->energy_pass(S->stream, S->d_scalars))) return rc;
  double sc[3];
  TDMD_HIP_TRY(hipMemcpyAsync(sc, S->d_scalars, 24, hipMemcpyDeviceToHost, S->stream));
  std::vector<int32_t> cnt(static_cast<size_t>(S->natoms));
  TDMD_HIP_TRY(hipMemcpyAsync(cnt.data(), S->a.cnt, S->natoms * 4, hipMemcpyDeviceToHost, S->stream));
  double nhc[32];
  TDMD_HIP_TRY(hipMemcpyAsync(nhc, S->a.nhc, 32 * 8, hipMemcpyDeviceToHost, S->stream));
  TDMD_HIP_TRY(hipStreamSynchronize(S->stream));
  out[0] = sc[0] + sc[1];
  out[1] = sc[2];
  double enhc = 0.0;
  const int M = S->nhcp.M;
  const double kt = kBoltzmann * S->cfg.t_target;
  for (int k = 0; k < M; ++k)
    enhc += 0.5 * nhc[16 + k] * nhc[8 + k] * nhc[8 + k] +
            (k == 0 ? static_cast<double>(S->nhcp.n_dof) * kt : kt) * nhc[k];
  out[2] = enhc;
  out[3] = 2.0 * sc[2] / (3.0 * static_cast<double>(S->natoms) * kBoltzmann);
  int64_t sum = 0;
  int mx = 0;
  for (int v : cnt) {
    sum += v;
    mx = std::max(mx, v);
  }
  out[4] = mx;
  out[5] = static_cast<double>(sum) / S->natoms;
  return check_flags(S);
}

}  // extern "C"
