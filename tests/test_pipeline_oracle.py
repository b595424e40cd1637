"""Pins of the zone-pipeline oracle (oracle.pipeline_run and its parts).

* The zone partition against the reference's own partition tests
  (tests/unit/test_zone_partition.cpp: every atom in one zone, atoms inside
  their slab, neighbour counts 4 and 3 for the 20 A box).
* Row-wise Morse forces (compute_morse_gpu_zone restated) against the
  half-list Newton-3 sweep, a different algorithm pinned by golden vectors.
* The sweep order: two atoms in different zones, one step, worked out here
  from the closed-form Morse force, independent of the oracle's C code.
  Plain velocity Verlet (every zone moved before any force) gives the
  lower zone a different force, which this test tells apart.
* One occupied zone: the sweep is velocity Verlet, so the oracle equals the
  pinned driver loop.
* A lattice at rest stays at rest (the reference's DeterministicMatchesM3
  case, tests/unit/test_pipeline_scheduler.cu:44-123).
"""
import numpy as np
import pytest

from conftest import MORSE_CU
from oracle import oracle as O
from tdmd_claude_amd import inputs as I

MVV2E = 1.0364269e-4


def fcc(n, T=300.0, seed=42, jitter=0.02):
    s = I.fcc_lattice(n, n, n, 3.615)
    if jitter:
        I.jitter_uniform(s, jitter, 4242)
    if T:
        I.init_velocities(s, T, seed)
    return s


def test_partition_matches_reference_cases():
    s = fcc(4, T=0.0, jitter=0.0)  # cu_fcc_256: 14.46 A box
    nz, w = O.zone_partition(s.lo[0], s.hi[0] - s.lo[0], 6.0)
    assert nz == 3 and w == pytest.approx(14.46 / 3)
    perm, off, cnt = O.assign_atoms(s.pos[:, 0], s.lo[0], w, nz)
    assert sorted(perm.tolist()) == list(range(s.natoms))  # each atom once
    assert cnt.sum() == s.natoms
    x = s.pos[perm, 0]
    for z in range(nz):
        seg = x[off[z]:off[z] + cnt[z]]
        assert np.all(seg >= s.lo[0] + z * w) and np.all(seg < s.lo[0] + (z + 1) * w + 1e-10)
    # ZoneNeighbors: box 20, 4 zones of 5 A
    nz, w = O.zone_partition(0.0, 20.0, 5.0, 4)
    assert [len(v) for v in O.zone_neighbors(nz, w, 6.5)] == [4, 4, 4, 4]
    assert [len(v) for v in O.zone_neighbors(nz, w, 4.0)] == [3, 3, 3, 3]
    assert O.zone_neighbors(nz, w, 4.0)[0] == [0, 1, 3]
    with pytest.raises(ValueError):
        O.zone_partition(0.0, 20.0, 5.0, 2)


def test_row_forces_equal_half_list_sweep():
    s = fcc(4)
    D, alpha, r0, rc = MORSE_CU
    half = O.HalfList(s.pos, s.lo, s.hi, rc, 1.0)
    _, f_half = O.morse_forces(half, D, alpha, r0, rc)
    f = np.zeros_like(f_half)
    full = half.full()
    O.morse_forces_rows(half.pos, f, 0, 100, full, s.lo, s.hi, MORSE_CU)
    O.morse_forces_rows(half.pos, f, 100, s.natoms - 100, full, s.lo, s.hi, MORSE_CU)
    scale = np.sqrt(np.mean(f_half ** 2))
    assert np.abs(f - f_half).max() < 1e-12 * scale
    # rows add: a second pass doubles them
    O.morse_forces_rows(half.pos, f, 0, s.natoms, full, s.lo, s.hi, MORSE_CU)
    assert np.abs(f - 2 * f_half).max() < 2e-12 * scale


def _morse_f(xa, xb, L):
    """Closed-form Morse force on the atom at xa from the one at xb (x axis)."""
    D, alpha, r0, rc = MORSE_CU
    d = xa - xb
    d -= L * np.round(d / L)
    r = abs(d)
    if r >= rc:
        return 0.0
    e = np.exp(-alpha * (r - r0))
    dudr = 2.0 * D * alpha * (1.0 - e) * e
    return -dudr / r * d


def test_sweep_order_two_atoms():
    L = 30.0
    s = I.System(lo=np.zeros(3), hi=np.full(3, L),
                 pos=np.array([[8.0, 5.0, 5.0], [12.5, 5.0, 5.0]]),
                 vel=np.array([[0.3, 0.0, 0.0], [-0.1, 0.0, 0.0]]),
                 types=np.array([1, 1], dtype=np.int32), ids=np.array([1, 2], dtype=np.int32),
                 masses=np.array([0.0, 63.546]))
    dt = 0.002
    pos, vel, f, types, ids, off, cnt = O.pipeline_run(s, MORSE_CU, targets=[1], dt=dt)
    # five 6 A slabs: atom 1 in zone 1, atom 2 in zone 2
    assert list(ids) == [1, 2] and list(cnt) == [0, 1, 1, 0, 0]
    m = 63.546 * MVV2E
    xa, xb, va, vb = 8.0, 12.5, 0.3, -0.1
    fa0, fb0 = _morse_f(xa, xb, L), _morse_f(xb, xa, L)
    # the lower zone first: B has not moved when A's new force is taken
    va += 0.5 * dt * fa0 / m
    xa += dt * va
    fa1 = _morse_f(xa, xb, L)
    va += 0.5 * dt * fa1 / m
    # the upper zone sees A already moved
    vb += 0.5 * dt * fb0 / m
    xb += dt * vb
    fb1 = _morse_f(xb, xa, L)
    vb += 0.5 * dt * fb1 / m
    assert pos[0, 0] == pytest.approx(xa, abs=1e-14) and pos[1, 0] == pytest.approx(xb, abs=1e-14)
    assert vel[0, 0] == pytest.approx(va, abs=1e-13) and vel[1, 0] == pytest.approx(vb, abs=1e-13)
    assert f[0, 0] == pytest.approx(fa1, rel=1e-12) and f[1, 0] == pytest.approx(fb1, rel=1e-12)
    # plain velocity Verlet would give A the force of B's new position
    assert abs(_morse_f(xa, xb, L) - fa1) > 1e-6


def test_one_occupied_zone_is_velocity_verlet():
    c = fcc(2, T=300.0, jitter=0.0)  # a 32-atom FCC cluster, x in [0, 5.42] ...
    L = 40.0
    s = I.System(lo=np.zeros(3), hi=np.full(3, L), pos=c.pos + 1.0, vel=c.vel.copy(),
                 types=c.types.copy(), ids=c.ids.copy(), masses=c.masses.copy())
    # ... inside the first of six 6.67 A slabs
    nz, w = O.zone_partition(0.0, L, MORSE_CU[3])
    assert nz == 6 and s.pos[:, 0].max() < w
    pos, vel, f, types, ids, off, cnt = O.pipeline_run(s, MORSE_CU, targets=[10], dt=0.001)
    assert cnt[0] == s.natoms
    rpos, rvel, rf, _, _ = O.run(s, morse=MORSE_CU, nsteps=10, thermo_every=10, rebuild_mode=1,
                                 rebuild_every=10)
    o = np.argsort(ids)
    assert np.abs(pos[o] - rpos).max() < 1e-12
    assert np.abs(vel[o] - rvel).max() < 1e-11
    assert np.abs(f[o] - rf).max() < 1e-10 * np.sqrt(np.mean(rf ** 2))


def test_lattice_at_rest_stays_at_rest():
    s = fcc(4, T=0.0, jitter=0.0)
    pos, vel, f, types, ids, off, cnt = O.pipeline_run(s, MORSE_CU, targets=[20], dt=0.001)
    o = np.argsort(ids)
    assert np.abs(pos[o] - s.pos).max() < 1e-12
    assert np.abs(vel).max() < 1e-10
