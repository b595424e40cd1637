"""The zone pipeline: PipelineScheduler over the gfx950 library.

Mirrors ``tdmd::scheduler::PipelineScheduler``
(src/scheduler/pipeline_scheduler.cuh:26-133): Morse forces, slabs along x
(ZonePartition, src/domain/zone_partition.cpp:15-143), zone steps on a pool
of streams, a deterministic mode with one stream and a fixed zone order,
optional Nose-Hoover chain and adaptive time step. The work happens in
``tdmd_pipe_*`` (csrc/pipeline.cpp); this class only moves host arrays.

``download`` returns the zone order with the atom ids, as the reference
does; ``download_sorted`` returns it sorted by id.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import capi
from .inputs import System

ZONE_STATES = {0: "Free", 1: "Receiving", 2: "Received", 3: "Ready", 4: "Computing",
               5: "Done", 6: "Sending"}


@dataclasses.dataclass
class PipelineConfig:
    """PipelineConfig (pipeline_scheduler.cuh:26-45) + the boundary's precision."""

    dt: float = 0.001
    r_skin: float = 1.0
    n_zones: int = 0  # auto: max(3, floor(Lx / rc))
    n_streams: int = 4
    rebuild_every: int = 10
    deterministic: bool = False
    t_target: float = 0.0
    t_period: float = 0.1
    nhc_length: int = 3
    adaptive_dt: bool = False
    dt_max: float = 0.002
    dt_min: float = 0.0001
    c2: float = 0.05
    precision: str = "fp64"

    def to_c(self) -> capi.PipeConfig:
        c = capi.PipeConfig()
        c.dt, c.r_skin = self.dt, self.r_skin
        c.n_zones, c.n_streams, c.rebuild_every = self.n_zones, self.n_streams, self.rebuild_every
        c.deterministic = int(self.deterministic)
        c.t_target, c.t_period, c.nhc_length = self.t_target, self.t_period, self.nhc_length
        c.adaptive_dt = int(self.adaptive_dt)
        c.dt_max, c.dt_min, c.c2 = self.dt_max, self.dt_min, self.c2
        if self.precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        c.precision = capi.TDMD_MIXED if self.precision == "mixed" else capi.TDMD_FP64
        return c


@dataclasses.dataclass
class PipelineStats:
    kernel_launches: int
    dep_check_calls: int
    dep_check_failures: int
    ticks: int


@dataclasses.dataclass
class Zone:
    id: int
    atom_offset: int
    natoms_in_zone: int
    time_step: int
    state: str
    lo: tuple
    hi: tuple


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class PipelineScheduler:
    """Zone pipeline on one GPU (Morse)."""

    def __init__(self, lo, hi, natoms: int, morse, cfg: PipelineConfig | None = None,
                 periodic=(1, 1, 1)):
        self.cfg = cfg or PipelineConfig()
        self.natoms = int(natoms)
        self._real = np.float32 if self.cfg.precision == "mixed" else np.float64
        self._box = capi.Box.make(lo, hi, periodic)
        d, alpha, r0, rc = morse
        self._params = capi.MorseParams.make(d, alpha, r0, rc)
        h = C.c_void_p()
        c = self.cfg.to_c()
        capi.call("tdmd_pipe_create", C.byref(h), C.byref(self._box), self.natoms,
                  C.byref(self._params), C.byref(c))
        self._h = h

    @classmethod
    def for_system(cls, sys: System, morse, cfg: PipelineConfig | None = None):
        s = cls(sys.lo, sys.hi, sys.natoms, morse, cfg, sys.periodic)
        s.upload(sys.pos, sys.vel, None, sys.types, sys.ids, sys.masses)
        return s

    def upload(self, pos, vel, forces, types, ids, masses) -> None:
        n = self.natoms
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 3)
        f = (np.zeros((n, 3), dtype=self._real) if forces is None
             else np.ascontiguousarray(forces, dtype=self._real).reshape(n, 3))
        types = np.ascontiguousarray(types, dtype=np.int32)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        masses = np.ascontiguousarray(masses, dtype=self._real)
        capi.call("tdmd_pipe_upload", self._h, _pd(pos), _pd(vel),
                  f.ctypes.data_as(C.c_void_p), _pi(types), _pi(ids),
                  masses.ctypes.data_as(C.c_void_p), len(masses))

    def run_until(self, target_step: int) -> None:
        capi.call("tdmd_pipe_run_until", self._h, int(target_step))

    def download(self):
        """(pos, vel, forces, types, ids) in zone order; forces in `real`."""
        n = self.natoms
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        f = np.zeros((n, 3), dtype=self._real)
        types = np.zeros(n, dtype=np.int32)
        ids = np.zeros(n, dtype=np.int32)
        capi.call("tdmd_pipe_download", self._h, _pd(pos), _pd(vel),
                  f.ctypes.data_as(C.c_void_p), _pi(types), _pi(ids), n)
        return pos, vel, f, types, ids

    def download_sorted(self):
        """The same, rows sorted by atom id."""
        pos, vel, f, types, ids = self.download()
        o = np.argsort(ids, kind="stable")
        return pos[o], vel[o], f[o].astype(np.float64), types[o], ids[o]

    def zones(self) -> list[Zone]:
        """partition().zones() (pipeline_scheduler.cuh:74-76)."""
        n = C.c_int32()
        capi.call("tdmd_pipe_zone_count", self._h, C.byref(n))
        buf = (capi.ZoneInfo * n.value)()
        capi.call("tdmd_pipe_zones", self._h, buf, n.value)
        return [Zone(z.id, z.atom_offset, z.natoms, z.time_step, ZONE_STATES[z.state],
                     tuple(z.lo), tuple(z.hi)) for z in buf]

    def zone_neighbors(self, zone: int) -> list[int]:
        n = C.c_int32()
        capi.call("tdmd_pipe_zone_neighbors", self._h, int(zone), None, 0, C.byref(n))
        out = np.zeros(max(n.value, 1), dtype=np.int32)
        capi.call("tdmd_pipe_zone_neighbors", self._h, int(zone), _pi(out), n.value, C.byref(n))
        return [int(x) for x in out[:n.value]]

    def min_time_step(self) -> int:
        out = C.c_int32()
        capi.call("tdmd_pipe_min_time_step", self._h, C.byref(out))
        return out.value

    def stats(self) -> PipelineStats:
        out = (C.c_int64 * 4)()
        capi.call("tdmd_pipe_stats", self._h, out)
        return PipelineStats(*[int(x) for x in out])

    def current_dt(self) -> float:
        out = C.c_double()
        capi.call("tdmd_pipe_current_dt", self._h, C.byref(out))
        return out.value

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            capi.lib().tdmd_pipe_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
