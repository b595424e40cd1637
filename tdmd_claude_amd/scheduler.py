"""Python mirror of FastPipelineScheduler over the C-ABI.

Same names and meaning as the reference C++ class
(src/scheduler/fast_pipeline_scheduler.cuh:36-192): construct with a box, an
atom count and a potential (Morse parameters or EAM/alloy tables), upload,
run_until(step), download in upload order, stats(). Errors raise
``capi.TdmdError`` (the reference aborts or throws).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import enum

import numpy as np

from . import capi
from .inputs import SetflTables, System


class PotentialKind(enum.IntEnum):
    Morse = 0
    Eam = 1


@dataclasses.dataclass
class FastPipelineConfig:
    """FastPipelineConfig (fast_pipeline_scheduler.cuh:42-46) plus precision,
    rebuild policy, NVT and launch options."""

    dt: float = 0.001
    r_skin: float = 1.0
    rebuild_every: int = 10
    precision: str = "mixed"  # "mixed" | "fp64"
    rebuild_policy: str = "fixed"  # "fixed" | "displacement"
    t_target: float = 0.0  # > 0: NVT (Nose-Hoover chain)
    t_period: float = 0.1
    nhc_length: int = 3
    use_graph: bool = True
    mirror_fp64: bool = False
    max_neighbors: int = 0
    pair_unroll: int = 0  # neighbour slots per pipelined group (0 = default)
    lds_budget: int = 0  # LDS bytes per workgroup for spline windows (0 = device limit)
    split_rebuild: bool = False  # rebuild as separate kernels instead of one fused launch
    disable_halo: bool = False  # no LDS halo staging in the single-element EAM passes
    pair_lanes: int = 0  # lanes per neighbour row: 0 = by size, 1, 4 or 8
    cell_pct: int = 0  # binning cells >= this % of r_list per axis (0 = 100)

    def to_c(self) -> capi.SchedConfig:
        c = capi.SchedConfig()
        capi.lib().tdmd_sched_config_default(C.byref(c))
        c.dt = self.dt
        c.r_skin = self.r_skin
        c.rebuild_every = self.rebuild_every
        c.precision = capi.TDMD_MIXED if self.precision == "mixed" else capi.TDMD_FP64
        c.rebuild_policy = (capi.REBUILD_DISPLACEMENT if self.rebuild_policy == "displacement"
                            else capi.REBUILD_FIXED)
        c.t_target = self.t_target
        c.t_period = self.t_period
        c.nhc_length = self.nhc_length
        c.use_graph = int(self.use_graph)
        c.mirror_fp64 = int(self.mirror_fp64)
        c.max_neighbors = self.max_neighbors
        c.pair_unroll = self.pair_unroll
        c.lds_budget = self.lds_budget
        c.split_rebuild = int(self.split_rebuild)
        c.disable_halo = int(self.disable_halo)
        c.pair_lanes = self.pair_lanes
        c.cell_pct = self.cell_pct
        return c


@dataclasses.dataclass
class FastPipelineStats:
    """FastPipelineStats (fast_pipeline_scheduler.cuh:49-53) + graph launches."""

    ticks: int
    kernel_launches: int
    rebuilds: int
    graph_launches: int


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class FastPipelineScheduler:
    """One-GPU step scheduler; potential chosen by the constructor argument."""

    def __init__(self, lo, hi, natoms: int, potential, cfg: FastPipelineConfig | None = None,
                 periodic=(1, 1, 1)):
        self.cfg = cfg or FastPipelineConfig()
        self.natoms = int(natoms)
        self._box = capi.Box.make(lo, hi, periodic)
        h = C.c_void_p()
        c = self.cfg.to_c()
        if isinstance(potential, SetflTables):
            t = potential
            self._F = np.ascontiguousarray(t.F, dtype=np.float64)
            self._rho = np.ascontiguousarray(t.rho, dtype=np.float64)
            self._rphi = np.ascontiguousarray(t.rphi, dtype=np.float64)
            capi.call("tdmd_sched_create_eam", C.byref(h), C.byref(self._box), self.natoms,
                      t.ntypes, t.nrho, t.drho, t.nr, t.dr, t.cutoff, _pd(self._F),
                      _pd(self._rho), _pd(self._rphi), C.byref(c))
            self.kind = PotentialKind.Eam
        else:
            d, alpha, r0, rc = potential
            p = capi.MorseParams.make(d, alpha, r0, rc)
            capi.call("tdmd_sched_create_morse", C.byref(h), C.byref(self._box), self.natoms,
                      C.byref(p), C.byref(c))
            self.kind = PotentialKind.Morse
        self._h = h

    @classmethod
    def for_system(cls, sys: System, potential, cfg: FastPipelineConfig | None = None):
        s = cls(sys.lo, sys.hi, sys.natoms, potential, cfg, sys.periodic)
        s.upload_system(sys)
        return s

    def upload(self, pos, vel, forces, types, ids, masses) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        types = np.ascontiguousarray(types, dtype=np.int32)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        masses = np.ascontiguousarray(masses, dtype=np.float64)
        f = None if forces is None else np.ascontiguousarray(forces, dtype=np.float64)
        capi.call("tdmd_sched_upload", self._h, _pd(pos), _pd(vel),
                  None if f is None else _pd(f), _pi(types), _pi(ids), _pd(masses), len(masses))

    def upload_system(self, sys: System) -> None:
        self.upload(sys.pos, sys.vel, None, sys.types, sys.ids, sys.masses)

    def run_until(self, target_step: int) -> None:
        capi.call("tdmd_sched_run_until", self._h, int(target_step))

    def download(self):
        n = self.natoms
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        f = np.zeros((n, 3))
        types = np.zeros(n, dtype=np.int32)
        ids = np.zeros(n, dtype=np.int32)
        capi.call("tdmd_sched_download", self._h, _pd(pos), _pd(vel), _pd(f), _pi(types),
                  _pi(ids), n)
        return pos, vel, f, types, ids

    def stats(self) -> FastPipelineStats:
        out = (C.c_int64 * 4)()
        capi.call("tdmd_sched_stats", self._h, out)
        return FastPipelineStats(*[int(x) for x in out])

    def current_step(self) -> int:
        s = C.c_int32()
        capi.call("tdmd_sched_current_step", self._h, C.byref(s))
        return s.value

    def potential_kind(self) -> PotentialKind:
        return self.kind

    def interaction_cutoff(self) -> float:
        rc = C.c_double()
        capi.call("tdmd_sched_interaction_cutoff", self._h, C.byref(rc))
        return rc.value

    def thermo(self) -> dict:
        out = (C.c_double * 6)()
        capi.call("tdmd_sched_thermo", self._h, out)
        return dict(pe=out[0], ke=out[1], nhc=out[2], temp=out[3], max_nbr=out[4],
                    mean_nbr=out[5])

    def observables(self) -> dict:
        """PE, KE, total, T, chain and conserved energy, virial (eV), volume
        (A^3), pressure (bar), neighbour counts at the current step."""
        o = capi.Observables()
        capi.call("tdmd_sched_observables", self._h, C.byref(o))
        return {name: getattr(o, name) for name, _ in capi.Observables._fields_ if name != "pad"}

    def time_passes(self, steps: int = 20) -> tuple[float, float]:
        """Mean ms of the two passes of a steady-state step (NVE, or NVT for EAM;
        advances the state)."""
        out = (C.c_float * 2)()
        capi.call("tdmd_sched_time_passes", self._h, int(steps), out)
        return float(out[0]), float(out[1])

    def pass_timing(self, on: bool = True) -> None:
        """Time the density and force passes of every step on the device
        clock (inside the captured graphs too); each call with on=True
        resets the sums."""
        capi.call("tdmd_sched_pass_timing", self._h, int(on))

    def pass_times(self) -> tuple[float, float]:
        """Mean ms per launch of the density and force passes since pass
        timing was (re)started."""
        out = (C.c_float * 2)()
        capi.call("tdmd_sched_pass_times", self._h, out)
        return float(out[0]), float(out[1])

    def autotune(self, iters: int = 10) -> dict:
        """Pick the fastest kernel shape on the uploaded state (lanes per row,
        and for the LDS passes the group size and window); returns
        {"pair_lanes", "pair_unroll", "lds_budget"}."""
        out = (C.c_int32 * 2)()
        capi.call("tdmd_sched_autotune", self._h, int(iters), out)
        return self.kernel_shape()

    def kernel_shape(self) -> dict:
        """The kernel shape in use: {"pair_lanes", "pair_unroll", "lds_budget"}."""
        out = (C.c_int32 * 3)()
        capi.call("tdmd_sched_kernel_shape", self._h, out)
        return {"pair_lanes": int(out[0]), "pair_unroll": int(out[1]),
                "lds_budget": int(out[2])}

    def halo_info(self) -> dict:
        """LDS halo staging: 1024-atom blocks, blocks staged, largest and total halo."""
        out = (C.c_int64 * 4)()
        capi.call("tdmd_sched_halo_info", self._h, out)
        return {"blocks": int(out[0]), "staged": int(out[1]), "max_halo": int(out[2]),
                "halo_atoms": int(out[3])}

    def stream(self) -> C.c_void_p:
        s = C.c_void_p()
        capi.call("tdmd_sched_stream", self._h, C.byref(s))
        return s

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            capi.lib().tdmd_sched_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchScheduler(FastPipelineScheduler):
    """A batch of independent systems (replicas, each in its own box) advanced
    together: every pass of a step is one launch over all replicas
    (tdmd_sched_create_{eam,morse}_batch). NVE. Arrays are uploaded and
    downloaded per replica."""

    def __init__(self, systems: list, potential, cfg: FastPipelineConfig | None = None):
        self.cfg = cfg or FastPipelineConfig()
        self.counts = np.array([s.natoms for s in systems], dtype=np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        self.natoms = int(self.offsets[-1])
        n = len(systems)
        self._boxes = (capi.Box * n)(*[capi.Box.make(s.lo, s.hi, s.periodic) for s in systems])
        h = C.c_void_p()
        c = self.cfg.to_c()
        if isinstance(potential, SetflTables):
            t = potential
            self._F = np.ascontiguousarray(t.F, dtype=np.float64)
            self._rho = np.ascontiguousarray(t.rho, dtype=np.float64)
            self._rphi = np.ascontiguousarray(t.rphi, dtype=np.float64)
            capi.call("tdmd_sched_create_eam_batch", C.byref(h), n, self._boxes, _pi(self.counts),
                      t.ntypes, t.nrho, t.drho, t.nr, t.dr, t.cutoff, _pd(self._F),
                      _pd(self._rho), _pd(self._rphi), C.byref(c))
            self.kind = PotentialKind.Eam
        else:
            d, alpha, r0, rc = potential
            p = capi.MorseParams.make(d, alpha, r0, rc)
            capi.call("tdmd_sched_create_morse_batch", C.byref(h), n, self._boxes,
                      _pi(self.counts), C.byref(p), C.byref(c))
            self.kind = PotentialKind.Morse
        self._h = h

    @classmethod
    def for_systems(cls, systems: list, potential, cfg: FastPipelineConfig | None = None):
        b = cls(systems, potential, cfg)
        b.upload_systems(systems)
        return b

    def upload_systems(self, systems: list) -> None:
        masses = systems[0].masses
        for s in systems[1:]:
            if len(s.masses) != len(masses) or not np.array_equal(s.masses, masses):
                raise ValueError("all replicas must share the per-type masses")
        pos = np.concatenate([s.pos for s in systems])
        vel = np.concatenate([s.vel for s in systems])
        types = np.concatenate([s.types for s in systems])
        ids = np.concatenate([s.ids for s in systems])
        self.upload(pos, vel, None, types, ids, masses)

    def download_replicas(self):
        """[(pos, vel, forces, types, ids) of each replica]."""
        pos, vel, f, types, ids = self.download()
        out = []
        for r in range(len(self.counts)):
            a, b = self.offsets[r], self.offsets[r + 1]
            out.append((pos[a:b], vel[a:b], f[a:b], types[a:b], ids[a:b]))
        return out


def tune_cell_size(systems, potential, cfg: FastPipelineConfig | None = None,
                   candidates=(100, 125, 150), steps: int = 30, rebuild_every: int = 10) -> dict:
    """Cell-size tuner: times a trial scheduler on the given state for each
    candidate binning cell size (percent of r_list) -- `steps` steps with a
    list rebuild every `rebuild_every` steps, after a warm-up window of the
    same shape -- and returns the fastest as {"cell_pct", "ms_per_step":
    {pct: ms}}. The cell size changes the spatial sort (atoms are ordered by
    cell) and so the order of the row sums, never which pairs interact: runs
    with different cell sizes agree to rounding. `systems` is one System or
    a list of replicas (BatchScheduler)."""
    import time

    base = cfg or FastPipelineConfig()
    times = {}
    for pct in candidates:
        trial = dataclasses.replace(base, cell_pct=int(pct), rebuild_policy="fixed",
                                    rebuild_every=rebuild_every)
        if isinstance(systems, list):
            sched = BatchScheduler.for_systems(systems, potential, trial)
        else:
            sched = FastPipelineScheduler.for_system(systems, potential, trial)
        sched.run_until(steps)  # initial build, graph capture
        sched.run_until(2 * steps)
        capi.call("tdmd_device_synchronize")
        t0 = time.perf_counter()
        sched.run_until(3 * steps)
        times[int(pct)] = 1000.0 * (time.perf_counter() - t0) / steps
        sched.close()
    best = min(times, key=times.get)
    return {"cell_pct": best, "ms_per_step": times}
