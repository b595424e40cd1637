"""Driver outputs of the step path: the thermo line and LAMMPS custom force
dumps.

* ``thermo_line`` is the driver's per-step report (print_thermo,
  src/drivers/tdmd_main.cpp:110-122): T from 3N degrees of freedom,
  T = 2 KE / (3 N kB), fixed-width ``%20.14f`` energies.
* ``write_force_dump`` writes ``dump custom id type x y z fx fy fz``
  (write_lammps_force_dump, tdmd_main.cpp:65-95): reals with 17 significant
  digits, box bounds always tagged ``pp pp pp`` as the reference does.
* ``read_dump`` reads such a dump back, LAMMPS's own included
  (read_lammps_dump, src/io/dump_reader.cpp:15-54): atoms sorted by id.

The C++ drop-in has the same functions in ``include/tdmd/io/``; the two
write byte-identical files.
"""
from __future__ import annotations

import dataclasses

import numpy as np

KB = 8.617333262e-5  # eV/K (src/core/constants.hpp:15)


def thermo_line(step: int, pe: float, ke: float, natoms: int) -> str:
    """The driver's thermo line for one step (energies in eV)."""
    temp = 2.0 * ke / (3.0 * natoms * KB) if natoms > 0 else 0.0
    return "Step %8d  PE %20.14f  KE %20.14f  TE %20.14f  T %10.4f" % (
        step, pe, ke, pe + ke, temp)


def _g17(x: float) -> str:
    return format(float(x), ".17g")


def write_force_dump(path, step: int, lo, hi, ids, types, pos, forces) -> None:
    """LAMMPS custom dump of positions and forces, in the given atom order."""
    ids = np.asarray(ids)
    types = np.asarray(types)
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    forces = np.asarray(forces, dtype=np.float64).reshape(-1, 3)
    n = len(ids)
    if not (len(types) == n and len(pos) == n and len(forces) == n):
        raise ValueError("ids, types, positions and forces must have one entry per atom")
    lines = ["ITEM: TIMESTEP", str(int(step)), "ITEM: NUMBER OF ATOMS", str(n),
             "ITEM: BOX BOUNDS pp pp pp"]
    lines += [f"{_g17(lo[d])} {_g17(hi[d])}" for d in range(3)]
    lines.append("ITEM: ATOMS id type x y z fx fy fz")
    for k in range(n):
        p, f = pos[k], forces[k]
        lines.append(f"{int(ids[k])} {int(types[k])} {_g17(p[0])} {_g17(p[1])} {_g17(p[2])} "
                     f"{_g17(f[0])} {_g17(f[1])} {_g17(f[2])}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


@dataclasses.dataclass
class Dump:
    step: int
    ids: np.ndarray     # (N,) int32, ascending
    types: np.ndarray   # (N,) int32
    pos: np.ndarray     # (N, 3) float64
    forces: np.ndarray  # (N, 3) float64


def read_dump(path) -> Dump:
    """Atoms of a ``id type x y z fx fy fz`` dump, sorted by id. Raises
    ValueError on a missing atom count, a short file or a malformed line."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    natoms, step, k = 0, 0, 0
    while k < len(lines):
        line = lines[k]
        if "ITEM: TIMESTEP" in line and k + 1 < len(lines):
            step = int(lines[k + 1])
        if "ITEM: NUMBER OF ATOMS" in line and k + 1 < len(lines):
            natoms = int(lines[k + 1])
        k += 1
        if "ITEM: ATOMS" in line:
            break
    if natoms == 0:
        raise ValueError("no atoms in dump file")
    if k + natoms > len(lines):
        raise ValueError("premature end of dump")
    rows = []
    for line in lines[k:k + natoms]:
        f = line.split()
        if len(f) < 8:
            raise ValueError("bad dump line: " + line)
        try:
            rows.append((int(f[0]), int(f[1]), *map(float, f[2:8])))
        except ValueError as e:
            raise ValueError("bad dump line: " + line) from e
    rows.sort(key=lambda r: r[0])
    a = np.array([r[2:] for r in rows], dtype=np.float64)
    return Dump(step, np.array([r[0] for r in rows], dtype=np.int32),
                np.array([r[1] for r in rows], dtype=np.int32), a[:, 0:3].copy(),
                a[:, 3:6].copy())
