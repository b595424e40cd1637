"""Synthetic inputs for the step path: FCC lattices, seeded velocities,
LAMMPS data files and EAM/alloy (setfl) tables.

* FCC generator -- the basis and id order of tools/gen_fcc_data.py:20-70
  (z, y, x, then basis; ids 1-based).
* Velocities -- the recipe of init_velocities,
  benchmarks/phase1_baseline/bench_pipeline_scheduler.cu:137-190: Gaussian
  per component with sigma = sqrt(kB T / (m mvv2e)), mass-weighted centre of
  mass removed, rescaled exactly to T with n_dof = 3N - 3. The Gaussian
  stream is numpy's PCG64 (portable), not libstdc++'s normal_distribution.
* LAMMPS data reader/writer -- atomic style, orthogonal box, Masses / Atoms /
  Velocities (src/io/lammps_data_reader.cpp:33-205).
* setfl writer -- the layout EamAlloy::read_setfl reads
  (src/potentials/eam_alloy.cpp:90-192), pair blocks in i <= j row order.
* Johnson-form Cu EAM -- rho(r) = f_e exp(-beta (r/r_e - 1)), a Morse-like
  pair phi(r), F(rho) = -E0 sqrt(rho/rho_e) + F2 (rho/rho_e - 1)^2, both r
  functions tapered to zero over [r_s, r_c] by a C2 smoothstep so energy is
  conserved through the cutoff; tabulated on 5000-point grids.
"""
from __future__ import annotations

import dataclasses
import pathlib

import numpy as np

KB = 8.617333262e-5  # eV/K, src/core/constants.hpp:15
MVV2E = 1.0364269e-4  # src/core/constants.hpp:28

FCC_BASIS = np.array([(0.0, 0.0, 0.0), (0.5, 0.5, 0.0), (0.5, 0.0, 0.5), (0.0, 0.5, 0.5)])
MASSES = {"Cu": 63.546, "Ni": 58.6934, "Co": 58.9332, "Cr": 51.9961}


@dataclasses.dataclass
class System:
    """Host-side system state in upload order (SystemState, src/core/system_state.hpp:20-50)."""

    lo: np.ndarray  # (3,)
    hi: np.ndarray  # (3,)
    pos: np.ndarray  # (N, 3) float64
    vel: np.ndarray  # (N, 3) float64
    types: np.ndarray  # (N,) int32, 1-based
    ids: np.ndarray  # (N,) int32
    masses: np.ndarray  # (ntypes + 1,) float64, index 0 unused
    periodic: tuple = (1, 1, 1)

    @property
    def natoms(self) -> int:
        return int(self.pos.shape[0])

    @property
    def box_size(self) -> np.ndarray:
        return self.hi - self.lo

    def copy(self) -> "System":
        return System(self.lo.copy(), self.hi.copy(), self.pos.copy(), self.vel.copy(),
                      self.types.copy(), self.ids.copy(), self.masses.copy(), self.periodic)


def fcc_lattice(nx: int, ny: int, nz: int, a: float, mass: float = MASSES["Cu"]) -> System:
    """Perfect FCC crystal, ids in z-y-x-basis order (tools/gen_fcc_data.py:49-57)."""
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    pos = ((cells[:, None, :] + FCC_BASIS[None, :, :]) * a).reshape(-1, 3)
    n = pos.shape[0]
    return System(
        lo=np.zeros(3),
        hi=np.array([nx * a, ny * a, nz * a], dtype=np.float64),
        pos=pos,
        vel=np.zeros((n, 3)),
        types=np.ones(n, dtype=np.int32),
        ids=np.arange(1, n + 1, dtype=np.int32),
        masses=np.array([0.0, mass]),
    )


def wrap(sys: System) -> None:
    """Floor-based wrap into [lo, hi) (src/core/math.hpp:107-129)."""
    L = sys.box_size
    rel = sys.pos - sys.lo
    rel -= np.floor(rel / L) * L
    sys.pos = sys.lo + rel


def jitter_uniform(sys: System, amp: float, seed: int) -> None:
    """Add U(-amp, amp) per component, then wrap."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sys.pos = sys.pos + rng.uniform(-amp, amp, size=sys.pos.shape)
    wrap(sys)


def displace_gaussian(sys: System, sigma: float, seed: int) -> None:
    """Add N(0, sigma^2) per component, then wrap."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sys.pos = sys.pos + rng.normal(0.0, sigma, size=sys.pos.shape)
    wrap(sys)


def init_velocities(sys: System, t_target: float, seed: int) -> None:
    """Maxwell-Boltzmann velocities at exactly t_target (3N-3 dof), zero momentum.

    Follows init_velocities, bench_pipeline_scheduler.cu:137-190.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    m = sys.masses[sys.types]
    sigma = np.sqrt(KB * t_target / (m * MVV2E))
    v = sigma[:, None] * rng.standard_normal(size=(sys.natoms, 3))
    com = (m[:, None] * v).sum(axis=0) / m.sum()
    v -= com
    ke = float((0.5 * m * MVV2E * (v * v).sum(axis=1)).sum())
    n_dof = 3 * sys.natoms - 3
    t_cur = 2.0 * ke / (n_dof * KB)
    if t_cur > 0:
        v *= np.sqrt(t_target / t_cur)
    sys.vel = v


def kinetic_energy(vel: np.ndarray, types: np.ndarray, masses: np.ndarray) -> float:
    """0.5 m mvv2e |v|^2 summed in atom order (tdmd_main.cpp:97-108)."""
    m = masses[types]
    return float((0.5 * m * MVV2E * (vel * vel).sum(axis=1)).sum())


# ---------------------------------------------------------------------------
# LAMMPS data files
# ---------------------------------------------------------------------------

def write_lammps_data(path, sys: System, title: str = "tdmd synthetic system") -> None:
    """Atomic-style data file the reference reader accepts (lammps_data_reader.cpp)."""
    ntypes = len(sys.masses) - 1
    lines = [title, "", f"{sys.natoms} atoms", f"{ntypes} atom types", ""]
    for a, name in enumerate(("x", "y", "z")):
        lines.append(f"{sys.lo[a]:.17g} {sys.hi[a]:.17g} {name}lo {name}hi")
    lines += ["", "Masses", ""]
    lines += [f"{t} {sys.masses[t]:.17g}" for t in range(1, ntypes + 1)]
    lines += ["", "Atoms # atomic", ""]
    p = sys.pos
    lines += [f"{sys.ids[i]} {sys.types[i]} {p[i, 0]:.17g} {p[i, 1]:.17g} {p[i, 2]:.17g}"
              for i in range(sys.natoms)]
    if np.any(sys.vel != 0):
        v = sys.vel
        lines += ["", "Velocities", ""]
        lines += [f"{sys.ids[i]} {v[i, 0]:.17g} {v[i, 1]:.17g} {v[i, 2]:.17g}"
                  for i in range(sys.natoms)]
    pathlib.Path(path).write_text("\n".join(lines) + "\n")


def read_lammps_data(path) -> System:
    """Atomic-style reader (lammps_data_reader.cpp:33-205): atoms stored at id-1."""
    text = pathlib.Path(path).read_text().splitlines()
    natoms = ntypes = 0
    lo = np.zeros(3)
    hi = np.zeros(3)
    k = 1
    section_words = ("Atoms", "Velocities", "Masses", "Pair Coeffs", "Bond Coeffs")
    while k < len(text):
        t = text[k].strip()
        if t and (t in section_words or t.startswith("Atoms ") or t.startswith("Pair Coeffs")
                  or t.startswith("Bond Coeffs")):
            break
        if "xlo xhi" in t:
            lo[0], hi[0] = map(float, t.split()[:2])
        elif "ylo yhi" in t:
            lo[1], hi[1] = map(float, t.split()[:2])
        elif "zlo zhi" in t:
            lo[2], hi[2] = map(float, t.split()[:2])
        elif "atom types" in t:
            ntypes = int(t.split()[0])
        elif "atoms" in t:
            natoms = int(t.split()[0])
        k += 1
    if natoms <= 0:
        raise ValueError("no atoms in data file")
    pos = np.zeros((natoms, 3))
    vel = np.zeros((natoms, 3))
    types = np.zeros(natoms, dtype=np.int32)
    ids = np.zeros(natoms, dtype=np.int32)
    masses = np.ones(ntypes + 1)
    masses[0] = 0.0
    while k < len(text):
        sec = text[k].split("#")[0].strip()
        k += 1
        if sec not in ("Masses", "Atoms", "Velocities"):
            continue
        count = ntypes if sec == "Masses" else natoms
        got = 0
        while got < count and k < len(text):
            t = text[k].strip()
            k += 1
            if not t:
                continue
            f = t.split()
            if sec == "Masses":
                masses[int(f[0])] = float(f[1])
            elif sec == "Atoms":
                i = int(f[0]) - 1
                ids[i] = int(f[0])
                types[i] = int(f[1])
                pos[i] = [float(x) for x in f[2:5]]
            else:
                vel[int(f[0]) - 1] = [float(x) for x in f[1:4]]
            got += 1
    return System(lo, hi, pos, vel, types, ids, masses)


# ---------------------------------------------------------------------------
# EAM/alloy tables
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class SetflTables:
    """setfl content: F[t, nrho], rho[t, nr], rphi[p, nr] (p in i <= j row order)."""

    elements: list
    masses: list
    nrho: int
    drho: float
    nr: int
    dr: float
    cutoff: float
    F: np.ndarray
    rho: np.ndarray
    rphi: np.ndarray

    @property
    def ntypes(self) -> int:
        return len(self.elements)


def smoothstep_taper(r: np.ndarray, r_s: float, r_c: float) -> np.ndarray:
    x = np.clip((r - r_s) / (r_c - r_s), 0.0, 1.0)
    return 1.0 - x ** 3 * (10.0 - 15.0 * x + 6.0 * x * x)


@dataclasses.dataclass
class JohnsonParams:
    """Analytic Johnson-form EAM parameters.

    Defaults are fitted (Nelder-Mead on a 256-atom crystal) to Cu at
    a = 3.615 A: cohesive energy -3.540 eV/atom, pressure 0.04 GPa, bulk
    modulus 141 GPa, Einstein spring constant 7.0 eV/A^2.
    """

    element: str = "Cu"
    mass: float = MASSES["Cu"]
    a: float = 3.615
    f_e: float = 1.0
    beta: float = 4.5397
    D_p: float = 0.0592
    alpha_p: float = 1.4911
    r_p: float = 3.2012  # pair minimum; 0 -> r_e
    E0: float = 3.4842
    F2: float = 0.2274
    r_s: float = 4.8
    r_c: float = 5.5

    @property
    def r_e(self) -> float:
        return self.a / np.sqrt(2.0)

    @property
    def r_min(self) -> float:
        return self.r_p if self.r_p > 0 else self.r_e

    @property
    def rho_e(self) -> float:
        return 12.0 * self.f_e

    def rho(self, r):
        return self.f_e * np.exp(-self.beta * (r / self.r_e - 1.0)) * smoothstep_taper(
            r, self.r_s, self.r_c)

    def phi(self, r):
        x = r - self.r_min
        return self.D_p * (np.exp(-2.0 * self.alpha_p * x) - 2.0 * np.exp(-self.alpha_p * x)) * \
            smoothstep_taper(r, self.r_s, self.r_c)

    def F(self, rho):
        q = rho / self.rho_e
        return -self.E0 * np.sqrt(q) + self.F2 * (q - 1.0) ** 2


def johnson_setfl(params: JohnsonParams | None = None, nr: int = 5000, nrho: int = 5000,
                  rho_max_factor: float = 3.0) -> SetflTables:
    """Tabulate the analytic functions on setfl grids: r_k = k dr, dr = r_c/(nr-1)."""
    p = params or JohnsonParams()
    dr = p.r_c / (nr - 1)
    drho = rho_max_factor * p.rho_e / (nrho - 1)
    r = np.arange(nr) * dr
    rho_grid = np.arange(nrho) * drho
    return SetflTables(
        elements=[p.element], masses=[p.mass], nrho=nrho, drho=drho, nr=nr, dr=dr,
        cutoff=p.r_c, F=p.F(rho_grid)[None, :], rho=p.rho(r)[None, :],
        rphi=(r * p.phi(r))[None, :])


# Ni/Co/Cr parameters of the three-element Johnson-form alloy: the Cu fit
# rescaled to a = 3.57 A with element-specific density, embedding and pair
# strengths (synthetic, for the batched-replica workload; not a fitted alloy).
ALLOY_NICOCR = {
    "Ni": dict(f_e=1.00, E0=4.10, D_p=0.0640, beta=4.60, alpha_p=1.52, F2=0.26),
    "Co": dict(f_e=1.05, E0=4.20, D_p=0.0660, beta=4.50, alpha_p=1.48, F2=0.25),
    "Cr": dict(f_e=0.92, E0=3.90, D_p=0.0700, beta=4.70, alpha_p=1.45, F2=0.24),
}


def johnson_alloy_setfl(elements=("Ni", "Co", "Cr"), a: float = 3.57, r_c: float = 5.0,
                        nr: int = 5000, nrho: int = 5000,
                        rho_max_factor: float = 3.0) -> SetflTables:
    """Multi-element Johnson-form EAM/alloy tables: per-element F and rho, and
    per-pair phi with Johnson's mixing rule
    phi_ab = (f_b / f_a phi_aa + f_a / f_b phi_bb) / 2 (equilibrium densities
    f at the nearest-neighbour distance), pair blocks in i <= j row order."""
    scale = a / 3.615
    ps = []
    for el in elements:
        kw = dict(ALLOY_NICOCR[el])
        ps.append(JohnsonParams(element=el, mass=MASSES[el], a=a, r_p=3.2012 * scale,
                                r_s=r_c - 0.6, r_c=r_c, **kw))
    dr = r_c / (nr - 1)
    drho = rho_max_factor * max(p.rho_e for p in ps) / (nrho - 1)
    r = np.arange(nr) * dr
    rho_grid = np.arange(nrho) * drho
    F = np.stack([p.F(rho_grid) for p in ps])
    rho = np.stack([p.rho(r) for p in ps])
    fe = [p.rho(np.array([p.r_e]))[0] for p in ps]
    rphi = []
    for i in range(len(ps)):
        for j in range(i, len(ps)):
            if i == j:
                phi = ps[i].phi(r)
            else:
                phi = 0.5 * (fe[j] / fe[i] * ps[i].phi(r) + fe[i] / fe[j] * ps[j].phi(r))
            rphi.append(r * phi)
    return SetflTables(elements=list(elements), masses=[MASSES[e] for e in elements], nrho=nrho,
                       drho=drho, nr=nr, dr=dr, cutoff=r_c, F=F, rho=rho, rphi=np.stack(rphi))


def alloy_replicas(nrep: int, seed: int, n_min: int = 500, n_max: int = 32000,
                   a: float = 3.57, t_range=(300.0, 1200.0), elements=("Ni", "Co", "Cr")):
    """Replicas of random equiatomic FCC alloys: for each, a size drawn
    log-uniform in [n_min, n_max] and rounded to the nearest 4 n^3 (n >= 3
    cells, so the 27-cell stencil and the minimum image hold), species
    uniform at random, Maxwell-Boltzmann velocities at T ~ U(t_range).
    Returns (list of System, list of temperatures)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    masses = np.array([1.0] + [MASSES[e] for e in elements])
    systems, temps = [], []
    for r in range(nrep):
        target = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        n = max(3, int(round((target / 4.0) ** (1.0 / 3.0))))
        s = fcc_lattice(n, n, n, a)
        s.types = rng.integers(1, len(elements) + 1, size=s.natoms).astype(np.int32)
        s.masses = masses.copy()
        t = float(rng.uniform(*t_range))
        init_velocities(s, t, int(rng.integers(0, 2**31 - 1)))
        systems.append(s)
        temps.append(t)
    return systems, temps


def write_setfl(path, t: SetflTables, comment: str = "tdmd synthetic EAM/alloy") -> None:
    """setfl text in the layout EamAlloy::read_setfl parses (eam_alloy.cpp:90-192)."""
    def block(vals):
        v = np.asarray(vals, dtype=np.float64)
        rows = [" ".join(f"{x:.16e}" for x in v[k:k + 5]) for k in range(0, len(v), 5)]
        return "\n".join(rows)

    out = [comment, "generated by tdmd_claude_amd.inputs", "units metal",
           f"{t.ntypes} " + " ".join(t.elements),
           f"{t.nrho} {t.drho:.16e} {t.nr} {t.dr:.16e} {t.cutoff:.16e}"]
    for e in range(t.ntypes):
        out.append(f"0 {t.masses[e]:.10g} 0.0 fcc")
        out.append(block(t.F[e]))
        out.append(block(t.rho[e]))
    for p in range(t.rphi.shape[0]):
        out.append(block(t.rphi[p]))
    pathlib.Path(path).write_text("\n".join(out) + "\n")


def read_setfl(path) -> SetflTables:
    """Token-based setfl reader with the reference's header handling."""
    lines = pathlib.Path(path).read_text().splitlines()
    head = lines[3].split()
    ntypes = int(head[0])
    elements = head[1:1 + ntypes]
    g = lines[4].split()
    nrho, drho, nr, dr, cutoff = int(g[0]), float(g[1]), int(g[2]), float(g[3]), float(g[4])
    k = 5
    F = np.zeros((ntypes, nrho))
    rho = np.zeros((ntypes, nr))
    masses = []
    rest = []
    for e in range(ntypes):
        while not lines[k].strip():
            k += 1
        masses.append(float(lines[k].split()[1]))
        k += 1
        toks = []
        while len(toks) < nrho + nr:
            toks += lines[k].split()
            k += 1
        F[e] = np.array(toks[:nrho], dtype=np.float64)
        rho[e] = np.array(toks[nrho:nrho + nr], dtype=np.float64)
        rest = toks[nrho + nr:]
    npair = ntypes * (ntypes + 1) // 2
    toks = list(rest)
    while len(toks) < npair * nr and k < len(lines):
        toks += lines[k].split()
        k += 1
    rphi = np.array(toks[:npair * nr], dtype=np.float64).reshape(npair, nr)
    return SetflTables(elements, masses, nrho, drho, nr, dr, cutoff, F, rho, rphi)
