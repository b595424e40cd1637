"""The step path at the benchmark shapes, against the FP64 oracle.

* 1,048,576-atom Johnson Cu EAM (64^3 FCC cells, 300 K, seed 7): above
  131,072 atoms the fused rebuild searches each atom's cells with one lane,
  its grid is the whole chip (two 1024-thread workgroups per CU), and the
  LDS passes serve four 1024-atom blocks per persistent workgroup with the
  next block's halo gathered while the current one is evaluated. Forced
  rebuilds (fixed interval) make every run here go through the fused
  rebuild at least twice.
* 143,748-atom Morse Cu (33^3 cells): above the sizes where the Morse rows
  are split over 8 lanes and the list search takes 32 lanes per atom.
* 64 Ni/Co/Cr replicas (571,384 atoms, 5000-point three-element tables):
  the batched multi-element passes.

The metric harness of the reference runs the scheduler at benchmark size
(benchmarks/phase1_baseline/bench_pipeline_scheduler.cu:257-296); these
tests are its parity counterpart.
"""
import numpy as np
import pytest

from conftest import MORSE_CU, force_rel_err
from gpu_helpers import say
from oracle import oracle as O
from tdmd_claude_amd import inputs as I
from tdmd_claude_amd.scheduler import BatchScheduler, FastPipelineConfig, FastPipelineScheduler

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-5  # max |dF| / RMS |F| against the FP64 oracle
FP64_TOL = 1e-9


def c3_system():
    s = I.fcc_lattice(64, 64, 64, 3.615)
    I.init_velocities(s, 300.0, 7)
    return s


def wrapped_max(a, b, s):
    L = s.hi - s.lo
    d = a - b
    d -= np.round(d / L) * L
    return float(np.abs(d).max())


def at_positions(s, pos, vel):
    t = s.copy()
    t.pos, t.vel = pos, vel
    return t


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_c3_forces_after_fused_rebuilds(prec):
    s = c3_system()
    tables = I.johnson_setfl()
    say(f"c3 forces after fused rebuilds, {prec}, {s.natoms} atoms")
    cfg = FastPipelineConfig(precision=prec, rebuild_policy="fixed", rebuild_every=5)
    sched = FastPipelineScheduler.for_system(s, tables, cfg)
    sched.run_until(12)  # rebuilds at steps 6 and 11
    st = sched.stats()
    pos, vel, f, _, _ = sched.download()
    sched.close()
    assert st.rebuilds == 3
    _, _, rf, _, _ = O.run(at_positions(s, pos, vel), eam=O.Eam(tables=tables), nsteps=0)
    err = force_rel_err(f, rf)
    say(f"  step-12 force err {err:.3e}")
    assert err < (MIXED_TOL if prec == "mixed" else FP64_TOL)


def test_c3_fused_rebuild_and_halo_staging_are_bitwise_neutral():
    # The fused one-launch rebuild against the separate kernels, and the
    # LDS-staged passes against global gathers, at 1M atoms: same sort, same
    # rows, same records in the same order, so the same bits.
    s = c3_system()
    tables = I.johnson_setfl()
    out = []
    for split, nohalo in ((False, False), (True, False), (False, True)):
        say(f"c3 bitwise, split_rebuild={split} disable_halo={nohalo}")
        cfg = FastPipelineConfig(precision="mixed", rebuild_policy="fixed", rebuild_every=5,
                                 split_rebuild=split, disable_halo=nohalo)
        sched = FastPipelineScheduler.for_system(s, tables, cfg)
        sched.run_until(12)
        pos, vel, f, _, _ = sched.download()
        out.append((pos, vel, f, sched.stats(), sched.halo_info()))
        sched.close()
    fused, split, nohalo = out
    assert fused[3].rebuilds == split[3].rebuilds == 3
    assert fused[3].kernel_launches < split[3].kernel_launches
    say(f"  halos {fused[4]}")
    assert fused[4]["blocks"] == 1024 and fused[4]["staged"] == 1024
    assert nohalo[4]["blocks"] == 0
    for other in (split, nohalo):
        for k in range(3):
            assert np.array_equal(fused[k], other[k])


def test_c3_fp64_nvt_trajectory_matches_oracle():
    # The benchmark's thermostat (NHC of 3, tau 0.1 ps, 300 K) with two fused
    # rebuilds, FP64, against the oracle driver loop.
    s = c3_system()
    tables = I.johnson_setfl()
    say("c3 fp64 nvt trajectory")
    cfg = FastPipelineConfig(precision="fp64", rebuild_policy="fixed", rebuild_every=3,
                             t_target=300.0, t_period=0.1, nhc_length=3)
    sched = FastPipelineScheduler.for_system(s, tables, cfg)
    sched.run_until(8)  # rebuilds at steps 4 and 7
    assert sched.stats().rebuilds == 3
    pos, vel, f, _, _ = sched.download()
    chain = sched.thermo()["nhc"]
    sched.close()
    rpos, rvel, rf, th, _ = O.run(s, eam=O.Eam(tables=tables), nsteps=8, thermo_every=8,
                                  rebuild_mode=1, rebuild_every=3, nvt=True, t_target=300.0,
                                  t_period=0.1, nhc_length=3)
    dx = wrapped_max(pos, rpos, s)
    say(f"  max |dx| {dx:.2e}, max |dv| {np.abs(vel - rvel).max():.2e}")
    assert dx < 1e-10
    assert np.abs(vel - rvel).max() < 1e-8
    assert force_rel_err(f, rf) < FP64_TOL
    assert chain == pytest.approx(th[-1, 5], rel=1e-9)


def test_morse_above_split_sizes_matches_oracle():
    # 33^3 cells: the Morse pass takes one lane per row and the fused
    # rebuild one lane per atom; FP64 trajectory against the oracle, and the
    # mixed fused rebuild against the separate kernels bit for bit
    s = I.fcc_lattice(33, 33, 33, 3.615)
    I.init_velocities(s, 300.0, 42)
    I.jitter_uniform(s, 0.02, 4242)
    say(f"morse {s.natoms} atoms")
    cfg = FastPipelineConfig(precision="fp64", rebuild_policy="fixed", rebuild_every=5)
    sched = FastPipelineScheduler.for_system(s, MORSE_CU, cfg)
    sched.run_until(12)
    pos, vel, f, _, _ = sched.download()
    sched.close()
    rpos, rvel, rf, _, _ = O.run(s, morse=MORSE_CU, nsteps=12, thermo_every=12, rebuild_mode=1,
                                 rebuild_every=5)
    dx = wrapped_max(pos, rpos, s)
    say(f"  fp64 max |dx| {dx:.2e}")
    assert dx < 1e-10
    assert force_rel_err(f, rf) < FP64_TOL
    out = []
    for split in (False, True):
        cfg = FastPipelineConfig(precision="mixed", rebuild_policy="fixed", rebuild_every=5,
                                 split_rebuild=split)
        sched = FastPipelineScheduler.for_system(s, MORSE_CU, cfg)
        sched.run_until(12)
        out.append(sched.download()[:3])
        sched.close()
    for k in range(3):
        assert np.array_equal(out[0][k], out[1][k])
    assert force_rel_err(out[0][2], O.run(at_positions(s, out[0][0], out[0][1]), morse=MORSE_CU,
                                          nsteps=0)[2]) < MIXED_TOL


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_c4_batch_forces_per_replica(prec):
    # 64 replicas of the batched workload (master seed 2024), forced
    # rebuilds, every replica's forces at its step-12 positions against the
    # oracle's three-element EAM
    systems, _ = I.alloy_replicas(64, 2024)
    tables = I.johnson_alloy_setfl()
    say(f"c4 batch {prec}: {len(systems)} replicas, {sum(s.natoms for s in systems)} atoms")
    cfg = FastPipelineConfig(precision=prec, rebuild_policy="fixed", rebuild_every=5)
    batch = BatchScheduler.for_systems(systems, tables, cfg)
    batch.run_until(12)
    assert batch.stats().rebuilds == 3
    got = batch.download_replicas()
    batch.close()
    eam = O.Eam(tables=tables)
    worst = 0.0
    for k, s in enumerate(systems):
        pos = got[k][0]
        half = O.HalfList(pos, s.lo, s.hi, tables.cutoff, 0.0)
        f_ref = eam.forces(half, s.types)["forces"]
        assert np.array_equal(got[k][3], s.types)
        worst = max(worst, force_rel_err(got[k][2], f_ref))
    say(f"  worst replica force err {worst:.3e}")
    assert worst < (MIXED_TOL if prec == "mixed" else FP64_TOL)
