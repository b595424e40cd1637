"""Batched replicas (one launch per pass over all replicas, each in its own
box): every replica's trajectory equals the same system run alone, for the
three-element Ni/Co/Cr alloy (FP64 and mixed) and for Morse; mixed forces
within the stated tolerance of the FP64 oracle per replica."""
import numpy as np
import pytest

from gpu_helpers import say
from oracle import oracle as O
from tdmd_claude_amd import inputs as I
from tdmd_claude_amd.scheduler import BatchScheduler, FastPipelineConfig, FastPipelineScheduler

MIXED_TOL = 1e-5


def replicas():
    systems, _ = I.alloy_replicas(3, 7, n_min=108, n_max=900)
    for k, s in enumerate(systems):
        I.displace_gaussian(s, 0.03, 100 + k)
    return systems


def max_dx(a, b, s):
    L = s.box_size
    d = a - b
    d -= np.round(d / L) * L
    return float(np.max(np.abs(d)))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_alloy_matches_individual_runs(prec):
    say(f"batch alloy {prec}")
    tables = I.johnson_alloy_setfl(nr=1000, nrho=1000)
    systems = replicas()
    cfg = FastPipelineConfig(precision=prec, rebuild_policy="displacement", r_skin=0.5)
    batch = BatchScheduler.for_systems(systems, tables, cfg)
    batch.run_until(60)
    got = batch.download_replicas()
    assert batch.stats().ticks == 60
    for k, s in enumerate(systems):
        single = FastPipelineScheduler.for_system(s, tables, cfg)
        single.run_until(60)
        pos, vel, f, types, ids = single.download()
        single.close()
        dx = max_dx(got[k][0], pos, s)
        say(f"  replica {k} ({s.natoms} atoms): max |dx| {dx:.2e}")
        assert np.array_equal(got[k][3], types) and np.array_equal(got[k][4], ids)
        assert dx < (1e-10 if prec == "fp64" else 1e-6)
    batch.close()


@pytest.mark.gpu
def test_batch_mixed_forces_vs_oracle():
    say("batch forces vs oracle")
    tables = I.johnson_alloy_setfl(nr=1000, nrho=1000)
    systems = replicas()
    batch = BatchScheduler.for_systems(systems, tables, FastPipelineConfig(precision="mixed"))
    batch.run_until(0)
    got = batch.download_replicas()
    eam = O.Eam(tables=tables)
    for k, s in enumerate(systems):
        half = O.HalfList(s.pos, s.lo, s.hi, tables.cutoff, 0.0)
        f_ref = eam.forces(half, s.types)["forces"]
        rms = np.sqrt(np.mean(np.sum(f_ref * f_ref, axis=1)))
        err = np.max(np.linalg.norm(got[k][2] - f_ref, axis=1)) / rms
        say(f"  replica {k}: force err {err:.2e}")
        assert err < MIXED_TOL
    batch.close()


@pytest.mark.gpu
def test_batch_morse_fp64_matches_oracle_trajectories():
    say("batch morse")
    morse = (0.3429, 1.3588, 2.866, 6.0)
    systems = []
    for k, n in enumerate((3, 4)):
        s = I.fcc_lattice(n, n, n, 3.615)
        I.init_velocities(s, 300.0 + 100 * k, 42 + k)
        systems.append(s)
    cfg = FastPipelineConfig(precision="fp64")
    batch = BatchScheduler.for_systems(systems, morse, cfg)
    batch.run_until(50)
    got = batch.download_replicas()
    for k, s in enumerate(systems):
        pos, vel, f, thermo, _ = O.run(s, morse=morse, nsteps=50, thermo_every=50, rebuild_mode=1,
                                       rebuild_every=10)
        dx = max_dx(got[k][0], pos, s)
        say(f"  replica {k}: max |dx| vs oracle {dx:.2e}")
        assert dx < 1e-9
    batch.close()


@pytest.mark.gpu
def test_batch_rejects_nvt():
    from tdmd_claude_amd import capi

    systems = replicas()[:2]
    with pytest.raises(capi.TdmdError):
        BatchScheduler(systems, I.johnson_alloy_setfl(nr=500, nrho=500),
                       FastPipelineConfig(t_target=300.0))
