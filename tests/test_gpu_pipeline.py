"""The zone pipeline (PipelineScheduler) on the GPU.

Deterministic mode is a fixed function of the input: it is compared with the
oracle's zone sweep (oracle.pipeline_run), in FP64 to rounding and in mixed
precision to the force tolerance. The pipelined mode (several streams, order
set by completion) is checked as the reference checks it
(tests/unit/test_pipeline_scheduler.cu:125-216): NVE energy conservation
over 1000 steps and every zone reaching the target step.
"""
import numpy as np
import pytest

from conftest import MORSE_CU
from gpu_helpers import say
from oracle import oracle as O
from tdmd_claude_amd import inputs as I
from tdmd_claude_amd.pipeline import PipelineConfig, PipelineScheduler

pytestmark = pytest.mark.gpu


def fcc(n, T=300.0, seed=42, jitter=0.02):
    s = I.fcc_lattice(n, n, n, 3.615)
    if jitter:
        I.jitter_uniform(s, jitter, 4242)
    if T:
        I.init_velocities(s, T, seed)
    return s


def by_id(arrs, ids):
    o = np.argsort(ids, kind="stable")
    return [a[o] for a in arrs]


def wrapped_diff(a, b, L):
    d = a - b
    return d - np.round(d / L) * L


def energy(pos, vel, s):
    half = O.HalfList(pos, s.lo, s.hi, MORSE_CU[3], 1.0)
    pe, _ = O.morse_forces(half, *MORSE_CU)
    return pe + O.kinetic_energy(vel, s.types, s.masses)


def test_partition_and_upload_order():
    s = fcc(5)
    say("pipeline partition")
    p = PipelineScheduler.for_system(s, MORSE_CU, PipelineConfig(deterministic=True))
    zones = p.zones()
    nz, w = O.zone_partition(s.lo[0], s.hi[0] - s.lo[0], MORSE_CU[3])
    perm, off, cnt = O.assign_atoms(s.pos[:, 0], s.lo[0], w, nz)
    assert len(zones) == nz
    assert [z.atom_offset for z in zones] == off.tolist()
    assert [z.natoms_in_zone for z in zones] == cnt.tolist()
    assert all(z.state == "Ready" and z.time_step == 0 for z in zones)
    for z in range(nz):
        assert p.zone_neighbors(z) == O.zone_neighbors(nz, w, MORSE_CU[3] + 1.0)[z]
    pos, vel, f, types, ids = p.download()
    assert np.array_equal(ids, s.ids[perm])
    assert np.array_equal(pos, s.pos[perm])


@pytest.mark.parametrize("targets", [[60], [17, 60]])
def test_deterministic_fp64_matches_oracle(targets):
    s = fcc(5)
    say("pipeline deterministic fp64", targets)
    p = PipelineScheduler.for_system(s, MORSE_CU, PipelineConfig(deterministic=True))
    for t in targets:
        p.run_until(t)
    pos, vel, f, _, ids = p.download()
    rpos, rvel, rf, _, rids, _, _ = O.pipeline_run(s, MORSE_CU, targets=targets)
    assert np.array_equal(ids, rids)
    L = s.hi - s.lo
    dp = np.abs(wrapped_diff(pos, rpos, L)).max()
    dv = np.abs(vel - rvel).max()
    df = np.abs(f - rf).max() / np.sqrt(np.mean(rf ** 2))
    say("  dpos", dp, "dvel", dv, "dF", df)
    assert dp < 1e-10 and dv < 1e-9 and df < 1e-9
    assert p.min_time_step() == targets[-1]
    st = p.stats()
    nz = len(p.zones())
    assert st.kernel_launches == nz * targets[-1] and st.ticks == targets[-1]


def test_deterministic_mixed_within_force_tolerance():
    s = fcc(5)
    say("pipeline deterministic mixed")
    p = PipelineScheduler.for_system(s, MORSE_CU,
                                     PipelineConfig(deterministic=True, precision="mixed"))
    p.run_until(20)
    pos, vel, f, _, ids = p.download()
    rpos, rvel, rf, _, rids, _, _ = O.pipeline_run(s, MORSE_CU, targets=[20],
                                                   dt=float(np.float32(0.001)))
    assert np.array_equal(ids, rids)
    df = np.abs(f.astype(np.float64) - rf).max() / np.sqrt(np.mean(rf ** 2))
    dp = np.abs(wrapped_diff(pos, rpos, s.hi - s.lo)).max()
    say("  dF", df, "dpos", dp)
    assert df < 1e-5 and dp < 1e-7


def test_deterministic_nvt_matches_oracle():
    s = fcc(5)
    say("pipeline deterministic nvt")
    cfg = PipelineConfig(deterministic=True, t_target=500.0, t_period=0.05, nhc_length=3)
    p = PipelineScheduler.for_system(s, MORSE_CU, cfg)
    p.run_until(40)
    pos, vel, f, _, ids = p.download()
    rpos, rvel, rf, _, rids, _, _ = O.pipeline_run(s, MORSE_CU, targets=[40],
                                                   nvt=(500.0, 0.05, 3))
    dv = np.abs(vel - rvel).max()
    dp = np.abs(wrapped_diff(pos, rpos, s.hi - s.lo)).max()
    say("  dpos", dp, "dvel", dv)
    assert dp < 1e-10 and dv < 1e-9


def test_deterministic_adaptive_dt_matches_oracle():
    s = fcc(5, T=600.0)
    say("pipeline adaptive dt")
    cfg = PipelineConfig(deterministic=True, adaptive_dt=True, dt_max=0.002, dt_min=0.0001,
                         c2=0.001)
    p = PipelineScheduler.for_system(s, MORSE_CU, cfg)
    p.run_until(30)
    pos, vel, f, _, ids = p.download()
    rpos, rvel, rf, _, _, _, _ = O.pipeline_run(s, MORSE_CU, targets=[30],
                                                adaptive=(0.0001, 0.002, 0.001))
    vmax = np.sqrt(np.max(np.sum(rvel ** 2, axis=1)))
    say("  dt", p.current_dt(), "vmax", vmax)
    assert 0.0001 <= p.current_dt() <= 0.002 and p.current_dt() < 0.001
    assert np.abs(wrapped_diff(pos, rpos, s.hi - s.lo)).max() < 1e-10
    assert np.abs(vel - rvel).max() < 1e-9


def test_lattice_at_rest_stays_at_rest():
    s = fcc(4, T=0.0, jitter=0.0)
    say("pipeline lattice at rest")
    p = PipelineScheduler.for_system(s, MORSE_CU, PipelineConfig(deterministic=True))
    p.run_until(100)
    pos, vel, f, _, ids = p.download_sorted()
    assert np.abs(pos - s.pos).max() < 1e-10 and np.abs(vel).max() < 1e-9


def test_pipelined_nve_conserves_energy():
    # The zone sweep is not velocity Verlet: a zone's forces mix its
    # neighbours' old and new positions. Its energy error on thermal Cu is
    # |dE/E| = 3.2e-4 over 1000 steps of 1 fs (the oracle's sweep gives the
    # same; velocity Verlet gives 3e-6) and 4.9e-5 at 0.5 fs: the
    # reference's 1e-4 bound is checked at 0.5 fs.
    s = fcc(5)
    say("pipelined nve, 4 streams")
    p = PipelineScheduler.for_system(s, MORSE_CU, PipelineConfig(n_streams=4, dt=0.0005))
    p.run_until(1)
    pos, vel, _, _, _ = p.download_sorted()
    e0 = energy(pos, vel, s)
    p.run_until(1000)
    pos, vel, _, _, _ = p.download_sorted()
    e1 = energy(pos, vel, s)
    drift = abs((e1 - e0) / e0)
    st = p.stats()
    say("  |dE/E|", drift, "ticks", st.ticks, "dep failures", st.dep_check_failures)
    assert drift < 1e-4
    assert all(z.time_step >= 1000 for z in p.zones())


def test_pipelined_three_zones_follow_the_sweep():
    # Three slabs are all neighbours of each other: one zone runs at a time,
    # in index order, whatever the stream count, so the pipelined run is the
    # deterministic sweep. As in the reference, a tick that finds a zone
    # ready after the last zone finished still starts it: zone 0 ends one
    # step ahead (run_until tests the minimum before each tick only), and
    # zones 1 and 2 hold exactly the sweep's state.
    s = fcc(5)
    say("pipelined 3 zones vs sweep")
    p = PipelineScheduler.for_system(s, MORSE_CU, PipelineConfig(n_streams=4))
    zones = p.zones()
    assert len(zones) == 3
    p.run_until(200)
    assert [z.time_step for z in p.zones()] == [201, 200, 200]
    pos, vel, _, _, ids = p.download()
    rpos, rvel, _, _, rids, _, _ = O.pipeline_run(s, MORSE_CU, targets=[200])
    assert np.array_equal(ids, rids)
    rest = slice(zones[1].atom_offset, s.natoms)
    dp = np.abs(wrapped_diff(pos[rest], rpos[rest], s.hi - s.lo)).max()
    dv = np.abs(vel[rest] - rvel[rest]).max()
    say("  zones 1-2: dpos", dp, "dvel", dv)
    assert dp < 1e-12 and dv < 1e-10


def test_zone_time_steps_advance():
    s = fcc(4)
    say("pipeline zones advance, 2 streams")
    p = PipelineScheduler.for_system(s, MORSE_CU, PipelineConfig(n_streams=2))
    p.run_until(10)
    assert p.min_time_step() >= 10
    assert all(z.time_step >= 10 for z in p.zones())
    st = p.stats()
    assert st.kernel_launches >= 10 * len(p.zones()) and st.ticks > 0


def test_errors():
    from tdmd_claude_amd import capi

    s = fcc(3)
    with pytest.raises(capi.TdmdError):
        PipelineScheduler(s.lo, s.hi, s.natoms, MORSE_CU, PipelineConfig(n_zones=2))
    p = PipelineScheduler(s.lo, s.hi, s.natoms, MORSE_CU, PipelineConfig())
    with pytest.raises(capi.TdmdError):
        p.run_until(1)  # nothing uploaded
    bad = s.types.copy()
    bad[0] = 5
    with pytest.raises(capi.TdmdError):
        p.upload(s.pos, s.vel, None, bad, s.ids, s.masses)
