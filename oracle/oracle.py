"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. Each function documents the
reference routine it restates (see tdmd_oracle.h).
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_BIN = HERE / "_ref" / "tdmd_standalone"

_D = C.c_double
_I = C.c_int
_I64 = C.c_int64
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)


class OrcBox(C.Structure):
    _fields_ = [("lo", _D * 3), ("hi", _D * 3), ("periodic", _I * 3)]


class OrcCells(C.Structure):
    _fields_ = [("ncx", _I), ("ncy", _I), ("ncz", _I), ("cs", _D * 3), ("ncells", _I),
                ("counts", _PI), ("offsets", _PI), ("atoms", _PI)]


class OrcNlist(C.Structure):
    _fields_ = [("natoms", _I64), ("r_cut", _D), ("r_skin", _D), ("counts", _PI),
                ("offsets", _PI), ("nbrs", _PI), ("total", _I64), ("build_pos", _PD)]


class OrcSpline(C.Structure):
    _fields_ = [("n", _I), ("dr", _D), ("rmin", _D), ("values", _PD), ("a", _PD), ("b", _PD),
                ("c", _PD), ("d", _PD)]


class OrcEam(C.Structure):
    _fields_ = [("ntypes", _I), ("cutoff", _D), ("nrho", _I), ("nr", _I), ("drho", _D),
                ("dr", _D), ("embedding", C.POINTER(OrcSpline)),
                ("density", C.POINTER(OrcSpline)), ("phi", C.POINTER(OrcSpline))]


class OrcNhc(C.Structure):
    _fields_ = [("M", _I), ("t_target", _D), ("t_period", _D), ("dt", _D), ("n_dof", _I),
                ("eta", _D * 8), ("eta_v", _D * 8), ("eta_m", _D * 8)]


class OrcRunCfg(C.Structure):
    _fields_ = [("use_eam", _I), ("morse", _D * 4), ("eam", C.POINTER(OrcEam)), ("dt", _D),
                ("skin", _D), ("nsteps", _I), ("thermo_every", _I), ("rebuild_mode", _I),
                ("rebuild_every", _I), ("nvt", _I), ("t_target", _D), ("t_period", _D),
                ("nhc_length", _I)]


class OrcRunOut(C.Structure):
    _fields_ = [("rebuilds", _I), ("nthermo", _I), ("thermo", _PD), ("thermo_cap", _I),
                ("last_pe", _D)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-C", str(HERE), "liboracle.so"], check=True,
                           capture_output=True)
        L = C.CDLL(str(LIB_PATH))
        L.orc_cells_build.argtypes = [C.POINTER(OrcCells), _PD, _I64, C.POINTER(OrcBox), _D]
        L.orc_cells_free.argtypes = [C.POINTER(OrcCells)]
        L.orc_nlist_build.argtypes = [C.POINTER(OrcNlist), _PD, _I64, C.POINTER(OrcBox), _D, _D]
        L.orc_nlist_needs_rebuild.argtypes = [C.POINTER(OrcNlist), _PD, _I64]
        L.orc_nlist_free.argtypes = [C.POINTER(OrcNlist)]
        L.orc_nlist_full.argtypes = [C.POINTER(OrcNlist), _PI, _PI, _PI]
        L.orc_bruteforce_full.argtypes = [_PD, _I64, C.POINTER(OrcBox), _D, _PI, _PI, _PI, _I64]
        L.orc_bruteforce_full.restype = _I64
        L.orc_morse_pair.argtypes = [_D, _D, _D, _D, _D, _PD, _PD]
        L.orc_morse_forces.argtypes = [_PD, _PD, _I64, C.POINTER(OrcBox), C.POINTER(OrcNlist),
                                       _D, _D, _D, _D]
        L.orc_morse_forces.restype = _D
        L.orc_morse_forces_rows.argtypes = [_PD, _PD, _I, _I, _PI, _PI, _PI, C.POINTER(OrcBox),
                                            _D, _D, _D, _D]
        L.orc_spline_build.argtypes = [C.POINTER(OrcSpline)]
        L.orc_spline_eval.argtypes = [C.POINTER(OrcSpline), _D]
        L.orc_spline_eval.restype = _D
        L.orc_spline_eval_deriv.argtypes = [C.POINTER(OrcSpline), _D]
        L.orc_spline_eval_deriv.restype = _D
        L.orc_eam_read_setfl.argtypes = [C.POINTER(OrcEam), C.c_char_p, C.c_char_p, _I]
        L.orc_eam_from_arrays.argtypes = [C.POINTER(OrcEam), _I, _I, _D, _I, _D, _D, _PD, _PD,
                                          _PD]
        L.orc_eam_forces.argtypes = [C.POINTER(OrcEam), _PD, _PI, _PD, _I64, C.POINTER(OrcBox),
                                     C.POINTER(OrcNlist), _PD, _PD, _PD]
        L.orc_eam_forces.restype = _D
        L.orc_eam_free.argtypes = [C.POINTER(OrcEam)]
        L.orc_half_kick.argtypes = [_PD, _PD, _PI, _PD, _I64, _D]
        L.orc_drift.argtypes = [_PD, _PD, _I64, _D, C.POINTER(OrcBox)]
        L.orc_drift_single_image.argtypes = [_PD, _PD, _I64, _D, C.POINTER(OrcBox)]
        L.orc_kinetic_energy.argtypes = [_PD, _PI, _PD, _I64]
        L.orc_kinetic_energy.restype = _D
        L.orc_nhc_init.argtypes = [C.POINTER(OrcNhc), _D, _D, _I, _D, _I]
        L.orc_nhc_half_step.argtypes = [C.POINTER(OrcNhc), _D]
        L.orc_nhc_half_step.restype = _D
        L.orc_nhc_energy.argtypes = [C.POINTER(OrcNhc)]
        L.orc_nhc_energy.restype = _D
        L.orc_fm_morse.argtypes = [_PD, _I64, C.POINTER(OrcBox), C.POINTER(OrcNlist), _D, _D,
                                   _D, _D, _PD, _PD, _PD]
        L.orc_fm_morse.restype = _D
        L.orc_eam_from_values.argtypes = [C.POINTER(OrcEam), _I, _D, _I, _D, _D, _PD, _PD, _PD]
        L.orc_fm_eam.argtypes = [C.POINTER(OrcEam), _PD, _I64, C.POINTER(OrcBox),
                                 C.POINTER(OrcNlist), _PD, _PD, _PD, _PD, _PD, _PD]
        L.orc_fm_eam.restype = _D
        L.orc_spline_coef_grad_to_values.argtypes = [_I, _D, _PD, _PD, _PD, _PD, _PD]
        L.orc_run.argtypes = [C.POINTER(OrcRunCfg), C.POINTER(OrcBox), _PD, _PD, _PD, _PI, _PD,
                              _I64, C.POINTER(OrcRunOut)]
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(_PD)


def _pi(a):
    return a.ctypes.data_as(_PI)


def box(lo, hi, periodic=(1, 1, 1)) -> OrcBox:
    b = OrcBox()
    for a in range(3):
        b.lo[a], b.hi[a], b.periodic[a] = float(lo[a]), float(hi[a]), int(periodic[a])
    return b


def cells(pos, lo, hi, r_list, periodic=(1, 1, 1)):
    """CellList::build (src/domain/cell_list.cpp:13-58)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    cl = OrcCells()
    b = box(lo, hi, periodic)
    if lib().orc_cells_build(C.byref(cl), _pd(pos), len(pos), C.byref(b), r_list) != 0:
        raise RuntimeError("orc_cells_build failed")
    nc = cl.ncells
    out = dict(n=(cl.ncx, cl.ncy, cl.ncz), cs=tuple(cl.cs),
               counts=np.ctypeslib.as_array(cl.counts, (nc,)).copy(),
               offsets=np.ctypeslib.as_array(cl.offsets, (nc,)).copy(),
               atoms=np.ctypeslib.as_array(cl.atoms, (max(len(pos), 1),))[:len(pos)].copy())
    lib().orc_cells_free(C.byref(cl))
    return out


class HalfList:
    """NeighborList (src/neighbors/neighbor_list.cpp:14-100), owned C state."""

    def __init__(self, pos, lo, hi, r_cut, r_skin, periodic=(1, 1, 1)):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.box = box(lo, hi, periodic)
        self.nl = OrcNlist()
        if lib().orc_nlist_build(C.byref(self.nl), _pd(self.pos), len(self.pos),
                                 C.byref(self.box), r_cut, r_skin) != 0:
            raise RuntimeError("orc_nlist_build failed")
        n = len(self.pos)
        self.counts = np.ctypeslib.as_array(self.nl.counts, (n,)).copy()
        self.offsets = np.ctypeslib.as_array(self.nl.offsets, (n,)).copy()
        self.nbrs = np.ctypeslib.as_array(self.nl.nbrs, (max(self.nl.total, 1),))[
            :self.nl.total].copy()

    def needs_rebuild(self, pos) -> bool:
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        return bool(lib().orc_nlist_needs_rebuild(C.byref(self.nl), _pd(pos), len(pos)))

    def full(self):
        """Symmetric closure with sorted rows: (counts, offsets, nbrs)."""
        n = len(self.pos)
        counts = np.zeros(n, dtype=np.int32)
        offsets = np.zeros(n, dtype=np.int32)
        nbrs = np.zeros(max(2 * self.nl.total, 1), dtype=np.int32)
        lib().orc_nlist_full(C.byref(self.nl), _pi(counts), _pi(offsets), _pi(nbrs))
        return counts, offsets, nbrs[:2 * self.nl.total]

    def __del__(self):
        try:
            lib().orc_nlist_free(C.byref(self.nl))
        except Exception:
            pass


def bruteforce_full(pos, lo, hi, r_list, periodic=(1, 1, 1)):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = len(pos)
    counts = np.zeros(n, dtype=np.int32)
    offsets = np.zeros(n, dtype=np.int32)
    cap = 400 * n
    nbrs = np.zeros(cap, dtype=np.int32)
    b = box(lo, hi, periodic)
    tot = lib().orc_bruteforce_full(_pd(pos), n, C.byref(b), r_list, _pi(counts), _pi(offsets),
                                    _pi(nbrs), cap)
    assert tot <= cap
    return counts, offsets, nbrs[:tot]


def morse_pair(D, alpha, r0, rc, r2):
    e, f = C.c_double(), C.c_double()
    lib().orc_morse_pair(D, alpha, r0, rc * rc, r2, C.byref(e), C.byref(f))
    return e.value, f.value


def morse_forces(half: HalfList, D, alpha, r0, rc):
    """compute_pair_forces with MorsePair (force_compute.cpp:10-56)."""
    n = len(half.pos)
    f = np.zeros((n, 3))
    pe = lib().orc_morse_forces(_pd(half.pos), _pd(f), n, C.byref(half.box), C.byref(half.nl),
                                D, alpha, r0, rc)
    return pe, f


def spline_coeffs(values, dr):
    """SplineTable::build_splines (eam_alloy.cpp:19-50) -> (a, b, c, d)."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    s = OrcSpline()
    s.n = len(v)
    s.dr = dr
    s.rmin = 0.0
    s.values = _pd(v)
    lib().orc_spline_build(C.byref(s))
    m = len(v) - 1
    out = tuple(np.ctypeslib.as_array(getattr(s, k), (m,)).copy() for k in "abcd")
    # a/b/c/d were malloc'd by the oracle; values belong to numpy: free only coeffs.
    libc = C.CDLL(None)
    for k in "abcd":
        libc.free(C.cast(getattr(s, k), C.c_void_p))
    return out


class Eam:
    """EamAlloy (eam_alloy.cpp:90-295) in the oracle."""

    def __init__(self, tables=None, path=None):
        self.e = OrcEam()
        if path is not None:
            err = C.create_string_buffer(256)
            if lib().orc_eam_read_setfl(C.byref(self.e), str(path).encode(), err, 256) != 0:
                raise RuntimeError(err.value.decode())
        else:
            t = tables
            F = np.ascontiguousarray(t.F, dtype=np.float64)
            rho = np.ascontiguousarray(t.rho, dtype=np.float64)
            rphi = np.ascontiguousarray(t.rphi, dtype=np.float64)
            if lib().orc_eam_from_arrays(C.byref(self.e), t.ntypes, t.nrho, t.drho, t.nr, t.dr,
                                         t.cutoff, _pd(F), _pd(rho), _pd(rphi)) != 0:
                raise RuntimeError("orc_eam_from_arrays failed")

    @property
    def cutoff(self):
        return self.e.cutoff

    def spline(self, kind, idx):
        s = getattr(self.e, kind)[idx]
        m = s.n - 1
        return dict(n=s.n, dr=s.dr, **{k: np.ctypeslib.as_array(getattr(s, k), (m,)).copy()
                                        for k in "abcd"})

    def eval(self, kind, idx, r, deriv=False):
        s = C.byref(getattr(self.e, kind)[idx])
        fn = lib().orc_spline_eval_deriv if deriv else lib().orc_spline_eval
        return fn(s, float(r))

    def forces(self, half: HalfList, types):
        n = len(half.pos)
        types = np.ascontiguousarray(types, dtype=np.int32)
        f = np.zeros((n, 3))
        rho = np.zeros(n)
        fp = np.zeros(n)
        emb = C.c_double()
        pe = lib().orc_eam_forces(C.byref(self.e), _pd(half.pos), _pi(types), _pd(f), n,
                                  C.byref(half.box), C.byref(half.nl), _pd(rho), _pd(fp),
                                  C.byref(emb))
        return dict(pe=pe, forces=f, rho=rho, fp=fp, embed=emb.value)

    def __del__(self):
        try:
            lib().orc_eam_free(C.byref(self.e))
        except Exception:
            pass


class EamValues(Eam):
    """Single-element EAM from knot values as EamAlloy holds them:
    F[nrho], rho[nr], phi[nr] (phi = (r phi)/r)."""

    def __init__(self, F, rho, phi, drho, dr, cutoff):
        self.e = OrcEam()
        self._v = [np.ascontiguousarray(x, dtype=np.float64) for x in (F, rho, phi)]
        F, rho, phi = self._v
        if lib().orc_eam_from_values(C.byref(self.e), len(F), drho, len(rho), dr, cutoff,
                                     _pd(F), _pd(rho), _pd(phi)) != 0:
            raise RuntimeError("orc_eam_from_values failed")


def fm_morse(half: HalfList, D, alpha, r0, rc, target):
    """Force-matching loss sum |F - T|^2 and its gradient wrt (D, alpha, r0)."""
    n = len(half.pos)
    target = np.ascontiguousarray(target, dtype=np.float64)
    grad = np.zeros(3)
    f = np.zeros((n, 3))
    L = lib().orc_fm_morse(_pd(half.pos), n, C.byref(half.box), C.byref(half.nl), D, alpha, r0,
                           rc, _pd(target), _pd(grad), _pd(f))
    return L, grad, f


def fm_eam(eam: EamValues, half: HalfList, target):
    """Loss and gradient wrt the F, rho, phi knot values (and the cubic
    coefficients, blocks F{a,b,c,d}, rho{a,b,c,d}, phi{a,b,c,d})."""
    n = len(half.pos)
    target = np.ascontiguousarray(target, dtype=np.float64)
    nrho, nr = eam.e.nrho, eam.e.nr
    gF, grho, gphi = np.zeros(nrho), np.zeros(nr), np.zeros(nr)
    gcoef = np.zeros(4 * (nrho - 1) + 8 * (nr - 1))
    f = np.zeros((n, 3))
    L = lib().orc_fm_eam(C.byref(eam.e), _pd(half.pos), n, C.byref(half.box), C.byref(half.nl),
                         _pd(target), _pd(gF), _pd(grho), _pd(gphi), _pd(gcoef), _pd(f))
    return L, dict(F=gF, rho=grho, phi=gphi, coef=gcoef), f


def spline_coef_grad_to_values(n, dr, ga, gb, gc, gd):
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (ga, gb, gc, gd)]
    gv = np.zeros(n)
    lib().orc_spline_coef_grad_to_values(n, dr, *[_pd(x) for x in arrs], _pd(gv))
    return gv


def kinetic_energy(vel, types, masses):
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    types = np.ascontiguousarray(types, dtype=np.int32)
    masses = np.ascontiguousarray(masses, dtype=np.float64)
    return lib().orc_kinetic_energy(_pd(vel), _pi(types), _pd(masses), len(vel))


def half_kick(vel, forces, types, masses, dt):
    lib().orc_half_kick(_pd(vel), _pd(np.ascontiguousarray(forces, dtype=np.float64)),
                        _pi(np.ascontiguousarray(types, dtype=np.int32)),
                        _pd(np.ascontiguousarray(masses, dtype=np.float64)), len(vel), dt)


def drift(pos, vel, lo, hi, dt, single_image=False, periodic=(1, 1, 1)):
    b = box(lo, hi, periodic)
    fn = lib().orc_drift_single_image if single_image else lib().orc_drift
    fn(_pd(pos), _pd(np.ascontiguousarray(vel, dtype=np.float64)), len(pos), dt, C.byref(b))


class Nhc:
    """NoseHooverChain (device_nose_hoover.cu:187-289)."""

    def __init__(self, t_target, t_period, M, dt, n_dof):
        self.c = OrcNhc()
        lib().orc_nhc_init(C.byref(self.c), t_target, t_period, M, dt, n_dof)

    def half_step(self, ke):
        return lib().orc_nhc_half_step(C.byref(self.c), ke)

    def energy(self):
        return lib().orc_nhc_energy(C.byref(self.c))

    def state(self):
        M = self.c.M
        return (np.array(self.c.eta[:M]), np.array(self.c.eta_v[:M]), np.array(self.c.eta_m[:M]))


def run(sys, *, morse=None, eam: Eam | None = None, dt=0.001, skin=1.0, nsteps=0,
        thermo_every=100, rebuild_mode=0, rebuild_every=10, nvt=False, t_target=0.0,
        t_period=0.1, nhc_length=3):
    """The driver loop (tdmd_main.cpp:211-254), optionally wrapped in NHC.

    Returns (pos, vel, forces, thermo[k, 6], rebuilds); `sys` is not modified.
    """
    pos = np.ascontiguousarray(sys.pos, dtype=np.float64).copy()
    vel = np.ascontiguousarray(sys.vel, dtype=np.float64).copy()
    f = np.zeros_like(pos)
    types = np.ascontiguousarray(sys.types, dtype=np.int32)
    masses = np.ascontiguousarray(sys.masses, dtype=np.float64)
    cfg = OrcRunCfg()
    cfg.use_eam = 1 if eam is not None else 0
    if morse is not None:
        for k in range(4):
            cfg.morse[k] = float(morse[k])
    if eam is not None:
        cfg.eam = C.pointer(eam.e)
    cfg.dt, cfg.skin, cfg.nsteps, cfg.thermo_every = dt, skin, nsteps, thermo_every
    cfg.rebuild_mode, cfg.rebuild_every = rebuild_mode, rebuild_every
    cfg.nvt, cfg.t_target, cfg.t_period, cfg.nhc_length = int(nvt), t_target, t_period, nhc_length
    cap = nsteps // max(thermo_every, 1) + 2
    th = np.zeros((cap, 6))
    out = OrcRunOut()
    out.thermo = _pd(th)
    out.thermo_cap = cap
    b = box(sys.lo, sys.hi, sys.periodic)
    rc = lib().orc_run(C.byref(cfg), C.byref(b), _pd(pos), _pd(vel), _pd(f), _pi(types),
                       _pd(masses), len(pos), C.byref(out))
    if rc != 0:
        raise RuntimeError("orc_run failed")
    return pos, vel, f, th[:out.nthermo].copy(), out.rebuilds


# ---------------------------------------------------------------------------
# Zone pipeline, deterministic mode (PipelineScheduler with
# cfg.deterministic, src/scheduler/pipeline_scheduler.cu:15-309)
# ---------------------------------------------------------------------------

def morse_forces_rows(pos, f, first, count, full, lo, hi, morse, periodic=(1, 1, 1)):
    """compute_morse_gpu_zone (device_morse_zone.cu:29-112), FP64: adds the
    row sums of atoms [first, first + count) of the full list to f."""
    counts, offsets, nbrs = full
    b = box(lo, hi, periodic)
    D, alpha, r0, rc = morse
    lib().orc_morse_forces_rows(_pd(pos), _pd(f), int(first), int(count), _pi(counts),
                                _pi(offsets), _pi(nbrs), C.byref(b), D, alpha, r0, rc)


def zone_partition(lo_x, lx, r_cut, n_zones=0):
    """ZonePartition::build (zone_partition.cpp:15-42): slabs along x,
    n = max(3, floor(Lx / r_cut)) when not given; returns (n, width)."""
    if n_zones <= 0:
        n_zones = max(3, int(np.floor(lx / r_cut)))
    if n_zones < 3:
        raise ValueError("need at least 3 zones for PBC")
    return n_zones, lx / n_zones


def zone_neighbors(n_zones, width, r_list):
    """build_zone_neighbors (zone_partition.cpp:123-143): zones within
    ceil(r_list / width) slabs, periodic, sorted, unique."""
    span = int(np.ceil(r_list / width))
    out = []
    for z in range(n_zones):
        v = sorted({(z + d) % n_zones for d in range(-span, span + 1)})
        out.append(v)
    return out


def assign_atoms(x, lo_x, width, n_zones):
    """assign_atoms (zone_partition.cpp:51-121): zone_of (:44-49) and a
    stable counting sort; returns (perm, offsets, counts)."""
    z = ((np.asarray(x) - lo_x) / width).astype(np.int64)  # truncation, as a C cast
    z = np.clip(z, 0, n_zones - 1)
    counts = np.bincount(z, minlength=n_zones)
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]])
    perm = np.argsort(z, kind="stable")
    return perm, offsets, counts


def pipeline_run(sys, morse, *, targets, dt=0.001, skin=1.0, n_zones=0, nvt=None,
                 adaptive=None):
    """Deterministic zone pipeline. Each tick sweeps the zones in index
    order (one stream, each zone step waited for, so every dependency is
    met): half kick, drift, zero, row forces of the zone over the whole
    position array, half kick. Forces of all atoms are recomputed at the
    start of every run_until (`targets` lists the run_until calls); the
    list is built from the uploaded positions and never again (the
    reference's rebuild flag is set by upload only). nvt = (t_target,
    t_period, M): chain half-steps on the global KE around every tick.
    adaptive = (dt_min, dt_max, c2): dt = clamp(c2 rc / vmax) before every
    tick. Returns (pos, vel, forces, types, ids, zone offsets, counts), zone
    order."""
    lo = np.asarray(sys.lo, dtype=np.float64)
    hi = np.asarray(sys.hi, dtype=np.float64)
    D, alpha, r0, rc = morse
    nz, width = zone_partition(lo[0], hi[0] - lo[0], rc, n_zones)
    perm, offsets, counts = assign_atoms(sys.pos[:, 0], lo[0], width, nz)
    pos = np.ascontiguousarray(sys.pos[perm], dtype=np.float64)
    vel = np.ascontiguousarray(sys.vel[perm], dtype=np.float64)
    types = np.ascontiguousarray(sys.types[perm], dtype=np.int32)
    ids = np.ascontiguousarray(sys.ids[perm], dtype=np.int32)
    masses = np.ascontiguousarray(sys.masses, dtype=np.float64)
    n = len(pos)
    half = HalfList(pos, lo, hi, rc, skin, sys.periodic)
    full = half.full()
    f = np.zeros_like(pos)
    chain = Nhc(nvt[0], nvt[1], nvt[2], dt, 3 * n - 3) if nvt else None
    step = 0
    for target in targets:
        f[:] = 0.0
        morse_forces_rows(pos, f, 0, n, full, lo, hi, morse, sys.periodic)
        while step < target:
            if adaptive:
                vmax = float(np.sqrt(np.max(np.sum(vel * vel, axis=1))))
                if vmax > 0:
                    dt = max(min(adaptive[1], adaptive[2] * rc / vmax), adaptive[0])
            if chain:
                vel *= chain.half_step(kinetic_energy(vel, types, masses))
            for z in range(nz):
                a, c = int(offsets[z]), int(counts[z])
                if c == 0:
                    continue
                sl = slice(a, a + c)
                half_kick(vel[sl], f[sl], types[sl], masses, dt)
                drift(pos[sl], vel[sl], lo, hi, dt, single_image=True, periodic=sys.periodic)
                f[sl] = 0.0
                morse_forces_rows(pos, f, a, c, full, lo, hi, morse, sys.periodic)
                half_kick(vel[sl], f[sl], types[sl], masses, dt)
            if chain:
                vel *= chain.half_step(kinetic_energy(vel, types, masses))
            step += 1
    return pos, vel, f, types, ids, offsets, counts
