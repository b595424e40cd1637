"""FastPipelineScheduler on the GPU against the FP64 oracle's driver loop.

The oracle runs the reference CPU driver's loop (tdmd_main.cpp:211-254) with
the fast path's fixed-interval rebuild (fast_pipeline_scheduler.cu:225-232);
the GPU runs the scheduler with its fused cross-step kick. In FP64 the two
agree to rounding over short runs; in mixed precision forces agree to 1e-5
of the RMS force.
"""
import numpy as np
import pytest

from conftest import GOLDEN, MORSE_CU, force_rel_err, load_golden, system_from_golden
from gpu_helpers import say
from oracle import oracle as O
from tdmd_claude_amd import capi
from tdmd_claude_amd import inputs as I
from tdmd_claude_amd.scheduler import FastPipelineConfig, FastPipelineScheduler, PotentialKind

pytestmark = pytest.mark.gpu


def c1():
    return system_from_golden(load_golden("ref_c1_morse"))


def fcc(n, T=300.0, seed=42, jitter=0.02):
    s = I.fcc_lattice(n, n, n, 3.615)
    if jitter:
        I.jitter_uniform(s, jitter, 4242)
    I.init_velocities(s, T, seed)
    return s


def potential(name):
    if name == "morse":
        return MORSE_CU
    if name == "mishin":
        return I.read_setfl(GOLDEN / "Cu_mishin1.eam.alloy")
    return I.johnson_setfl()


def oracle_run(s, pot, nsteps, **kw):
    if pot == "morse":
        return O.run(s, morse=MORSE_CU, nsteps=nsteps, thermo_every=max(nsteps, 1),
                     rebuild_mode=1, rebuild_every=10, **kw)
    return O.run(s, eam=O.Eam(tables=potential(pot)), nsteps=nsteps,
                 thermo_every=max(nsteps, 1), rebuild_mode=1, rebuild_every=10, **kw)


def gpu_run(s, pot, nsteps, chunks=None, **cfg):
    sched = FastPipelineScheduler.for_system(s, potential(pot), FastPipelineConfig(**cfg))
    for t in (chunks or [nsteps]):
        sched.run_until(t)
    pos, vel, f, types, ids = sched.download()
    return sched, pos, vel, f


@pytest.mark.parametrize("lanes", [0, 1])
@pytest.mark.parametrize("pot", ["morse", "johnson", "mishin"])
def test_fp64_trajectory_matches_oracle(pot, lanes):
    # lanes 0: rows split over 8 lanes (the default at this size; EAM then
    # runs the global-table passes); 1: one lane per row (EAM: LDS passes)
    s = c1() if pot == "morse" else fcc(5)
    say("sched fp64", pot, s.natoms, "lanes", lanes)
    sched, pos, vel, f = gpu_run(s, pot, 100, precision="fp64", pair_lanes=lanes)
    rpos, rvel, rf, th, reb = oracle_run(s, pot, 100)
    L = s.hi - s.lo
    dpos = pos - rpos
    dpos -= np.round(dpos / L) * L
    say("  max dpos", np.abs(dpos).max(), "dvel", np.abs(vel - rvel).max())
    assert np.abs(dpos).max() < 1e-10
    assert np.abs(vel - rvel).max() < 1e-8
    assert force_rel_err(f, rf) < 1e-9
    st = sched.stats()
    assert st.ticks == 100 and st.rebuilds == 1 + 9  # initial + steps 11, 21, ..., 91
    assert sched.current_step() == 100
    assert sched.potential_kind() == (PotentialKind.Morse if pot == "morse" else PotentialKind.Eam)


@pytest.mark.parametrize("lanes", [0, 1])
@pytest.mark.parametrize("pot", ["morse", "johnson"])
@pytest.mark.parametrize("mirror_fp64", [False, True])
def test_mixed_forces_and_trajectory_vs_fp64_oracle(pot, mirror_fp64, lanes):
    s = fcc(6)
    say("sched mixed", pot, "mirror_fp64", mirror_fp64, "lanes", lanes)
    sched = FastPipelineScheduler.for_system(
        s, potential(pot), FastPipelineConfig(precision="mixed", mirror_fp64=mirror_fp64,
                                              pair_lanes=lanes))
    sched.run_until(0)  # initial forces only
    _, _, f0, _, _ = sched.download()
    _, _, rf0, _, _ = oracle_run(s, pot, 0)
    err = force_rel_err(f0, rf0)
    say("  step-0 rel force err", err)
    assert err < 1e-5
    sched.run_until(50)
    pos, vel, f, _, _ = sched.download()
    rpos, rvel, rf, _, _ = oracle_run(s, pot, 50)
    L = s.hi - s.lo
    dpos = pos - rpos
    dpos -= np.round(dpos / L) * L
    say("  50-step max dpos", np.abs(dpos).max(), "force err", force_rel_err(f, rf))
    assert np.abs(dpos).max() < 1e-5
    assert force_rel_err(f, rf) < 1e-4


def test_thermo_matches_oracle():
    s = fcc(5)
    for prec in ("fp64", "mixed"):
        say("thermo", prec)
        sched = FastPipelineScheduler.for_system(s, potential("johnson"),
                                                 FastPipelineConfig(precision=prec))
        th = sched.thermo()
        _, _, _, rth, _ = oracle_run(s, "johnson", 0)
        tol = 1e-12 if prec == "fp64" else 1e-6
        assert th["pe"] == pytest.approx(rth[0, 1], rel=tol)
        assert th["ke"] == pytest.approx(rth[0, 2], rel=1e-13)
        assert th["temp"] == pytest.approx(rth[0, 4], rel=1e-12)


@pytest.mark.parametrize("pot", ["johnson", "morse"])
def test_graph_and_direct_launch_are_bitwise_identical(pot):
    # Morse alternates its record buffers every step: graphs are keyed by the
    # buffer state they were captured in (windows of odd length here)
    s = fcc(5)
    say("graph vs direct", pot)
    out = []
    for g in (True, False):
        _, pos, vel, f = gpu_run(s, pot, 41, chunks=[7, 24, 41], precision="mixed",
                                 use_graph=g, rebuild_policy="displacement")
        out.append((pos, vel, f))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_reupload_with_two_types_matches_a_fresh_scheduler():
    # When every atom is of type 1 the integrator does not read the types
    # (a flag that captured graphs hold). Re-uploading a two-type, two-mass
    # state into a scheduler whose graphs were captured for one type must
    # give what a fresh scheduler gives, bit for bit.
    s1 = fcc(5)
    s2 = s1.copy()
    s2.types = np.where(np.arange(s2.natoms) % 3 == 0, 2, 1).astype(np.int32)
    s2.masses = np.append(s1.masses, 2.0 * s1.masses[1])
    say("re-upload with two types")
    cfg = FastPipelineConfig(precision="mixed", rebuild_policy="displacement")
    sched = FastPipelineScheduler.for_system(s1, MORSE_CU, cfg)
    sched.run_until(20)
    sched.upload_system(s2)
    sched.run_until(20)
    got = sched.download()
    sched.close()
    fresh = FastPipelineScheduler.for_system(s2, MORSE_CU, cfg)
    fresh.run_until(20)
    want = fresh.download()
    fresh.close()
    for a, b in zip(got[:3], want[:3]):
        assert np.array_equal(a, b)


def test_download_returns_upload_order():
    s = fcc(5)
    say("upload order")
    perm = np.random.Generator(np.random.PCG64(5)).permutation(s.natoms)
    p = s.copy()
    p.pos, p.vel, p.types, p.ids = s.pos[perm], s.vel[perm], s.types[perm], s.ids[perm]
    _, pos_a, vel_a, f_a = gpu_run(s, "johnson", 30, precision="fp64")
    sched, pos_b, vel_b, f_b = gpu_run(p, "johnson", 30, precision="fp64")
    _, _, _, types, ids = sched.download()
    assert np.array_equal(ids, p.ids) and np.array_equal(types, p.types)
    assert np.abs(pos_b - pos_a[perm]).max() < 1e-11
    assert np.abs(vel_b - vel_a[perm]).max() < 1e-9


@pytest.mark.parametrize("pot", ["johnson", "morse"])
def test_chunked_run_until_reproduces_single_window(pot):
    s = fcc(5)
    say("chunked run_until", pot)
    _, pos_a, vel_a, _ = gpu_run(s, pot, 60, precision="fp64")
    _, pos_b, vel_b, _ = gpu_run(s, pot, 60, chunks=[10, 35, 60], precision="fp64")
    assert np.abs(pos_a - pos_b).max() < 1e-11
    assert np.abs(vel_a - vel_b).max() < 1e-9


def test_displacement_rebuild_policy():
    # A crystal never moves skin/2 = 0.5 A; a 0.2 A skin makes the device-side
    # trigger fire every few steps. With valid lists the forces do not depend
    # on when lists are rebuilt, so the run must agree with the oracle's
    # displacement-triggered driver loop (which rebuilds on its own schedule).
    s = fcc(5, T=600.0)
    say("displacement policy")
    sched, pos, vel, f = gpu_run(s, "johnson", 200, precision="fp64", r_skin=0.2,
                                 rebuild_policy="displacement")
    st = sched.stats()
    say("  rebuilds", st.rebuilds)
    assert 3 < st.rebuilds < 200
    rpos, rvel, rf, _, rreb = O.run(s, eam=O.Eam(tables=potential("johnson")), nsteps=200,
                                    thermo_every=200, skin=0.2, rebuild_mode=0)
    L = s.hi - s.lo
    dpos = pos - rpos
    dpos -= np.round(dpos / L) * L
    say("  max dpos", np.abs(dpos).max(), "oracle rebuilds", rreb)
    assert np.abs(dpos).max() < 1e-9
    assert force_rel_err(f, rf) < 1e-8


@pytest.mark.parametrize("pot", ["johnson", "morse"])
def test_large_box_trajectory_matches_oracle(pot):
    # 7^3 cells (25.3 A): at least three cells no smaller than r_list per
    # axis, so the list search takes each stencil cell's periodic image
    # instead of a per-candidate minimum-image test; the rows (and so the
    # FP64 trajectory) must be those of the oracle's list
    s = fcc(7)
    say("large box", pot, s.natoms)
    sched, pos, vel, f = gpu_run(s, pot, 40, precision="fp64", rebuild_every=5)
    rpos, rvel, rf, th, reb = O.run(s, eam=None if pot == "morse" else O.Eam(tables=potential(pot)),
                                    morse=MORSE_CU if pot == "morse" else None, nsteps=40,
                                    thermo_every=40, rebuild_mode=1, rebuild_every=5)
    L = s.hi - s.lo
    dpos = pos - rpos
    dpos -= np.round(dpos / L) * L
    say("  max dpos", np.abs(dpos).max())
    assert np.abs(dpos).max() < 1e-10
    assert force_rel_err(f, rf) < 1e-9
    assert sched.stats().rebuilds == 1 + 7


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("pot,chunks,lanes", [("johnson", None, 0), ("johnson", [7, 31, 60], 0),
                                              ("johnson", None, 1), ("morse", None, 0),
                                              ("mishin", [25, 60], 0), ("mishin", None, 1)])
def test_nvt_matches_oracle_chain(prec, pot, chunks, lanes):
    # windows end with the chain half-step after the step only
    s = fcc(5, T=300.0)
    say("nvt", prec, pot, chunks, "lanes", lanes)
    cfg = dict(precision=prec, t_target=400.0, t_period=0.05, nhc_length=3, pair_lanes=lanes)
    sched, pos, vel, f = gpu_run(s, pot, 60, chunks=chunks, **cfg)
    rpos, rvel, rf, th, _ = oracle_run(s, pot, 60, nvt=True, t_target=400.0,
                                       t_period=0.05, nhc_length=3)
    ke = I.kinetic_energy(vel, s.types, s.masses)
    say("  ke", ke, th[-1, 2])
    tol = 1e-10 if prec == "fp64" else 1e-5
    assert ke == pytest.approx(th[-1, 2], rel=tol)
    assert sched.thermo()["nhc"] == pytest.approx(th[-1, 5], rel=1e-6 if prec == "mixed" else 1e-9)


def test_nvt_reaches_target_temperature():
    s = fcc(6, T=150.0)
    say("nvt temperature")
    sched = FastPipelineScheduler.for_system(
        s, potential("johnson"),
        FastPipelineConfig(precision="mixed", t_target=300.0, t_period=0.1,
                           rebuild_policy="displacement"))
    temps = []
    for k in range(1, 31):
        sched.run_until(100 * k)
        temps.append(sched.thermo()["temp"])
    late = np.array(temps[10:]) * (3 * s.natoms) / (3 * s.natoms - 3)
    say("  <T>", late.mean())
    assert abs(late.mean() - 300.0) < 0.1 * 300.0


def test_overflow_is_reported_not_written():
    s = fcc(4)
    say("overflow")
    sched = FastPipelineScheduler.for_system(
        s, potential("johnson"), FastPipelineConfig(precision="fp64", max_neighbors=24))
    with pytest.raises(capi.TdmdError) as e:
        sched.run_until(5)
    assert e.value.status == 3


def test_launch_accounting():
    s = fcc(4)
    say("launch accounting")
    sched = FastPipelineScheduler.for_system(s, potential("johnson"),
                                             FastPipelineConfig(precision="mixed"))
    sched.run_until(1)
    a = sched.stats()
    sched.run_until(101)
    b = sched.stats()
    # EAM: density + force(+kick-drift) per step, one bootstrap per window,
    # and a 10-launch rebuild every 10 steps.
    per_step = (b.kernel_launches - a.kernel_launches) / 100.0
    assert 2.0 <= per_step <= 3.2
    assert b.ticks - a.ticks == 100
    # Morse: force(+kick-drift) and the rebuild check per step
    sched = FastPipelineScheduler.for_system(s, MORSE_CU, FastPipelineConfig(
        precision="mixed", rebuild_policy="displacement"))
    sched.run_until(1)
    a = sched.stats()
    sched.run_until(101)
    b = sched.stats()
    per_step = (b.kernel_launches - a.kernel_launches) / 100.0
    say("  morse launches per step", per_step)
    assert 2.0 <= per_step <= 2.2
    # EAM NVT: kick-drift, rebuild check, density, force(+kick), chain
    sched = FastPipelineScheduler.for_system(s, potential("johnson"), FastPipelineConfig(
        precision="mixed", rebuild_policy="displacement", t_target=300.0))
    sched.run_until(1)
    a = sched.stats()
    sched.run_until(101)
    b = sched.stats()
    per_step = (b.kernel_launches - a.kernel_launches) / 100.0
    say("  nvt launches per step", per_step)
    assert 5.0 <= per_step <= 5.2


@pytest.mark.parametrize("pot,lanes", [("johnson", 1), ("johnson", 0), ("morse", 0),
                                       ("alloy", 0)])
def test_autotune_picks_a_shape_and_leaves_the_trajectory_unchanged(pot, lanes):
    # the tuner times lanes per row {1, 4, 8} (and the LDS passes' shapes),
    # keeps the fastest and reruns the passes with it: the same shape as the
    # start leaves the trajectory bitwise unchanged, another lane count only
    # the order of the row sums (FP64: rounding)
    if pot == "alloy":
        from golden.make_golden import alloy_system, alloy_tables

        s, params = alloy_system("alloy3"), alloy_tables("alloy3")
    else:
        s, params = fcc(6), potential(pot)
    say("autotune", pot, lanes)
    cfg = FastPipelineConfig(precision="fp64", pair_lanes=lanes)
    ref = FastPipelineScheduler.for_system(s, params, cfg)
    start = ref.kernel_shape()
    ref.run_until(10)  # the same two windows as the tuned run
    ref.run_until(20)
    p0, v0, _, _, _ = ref.download()
    tuned = FastPipelineScheduler.for_system(s, params, cfg)
    tuned.run_until(10)
    choice = tuned.autotune(3)
    say("  start", start, "choice", choice)
    assert choice == tuned.kernel_shape()
    assert choice["pair_lanes"] in (1, 4, 8)
    if pot == "johnson" and choice["pair_lanes"] == 1:
        assert choice["pair_unroll"] in (4, 8) and choice["lds_budget"] > 0
    tuned.run_until(20)
    p1, v1, _, _, _ = tuned.download()
    L = s.hi - s.lo
    d = p1 - p0
    d -= np.round(d / L) * L
    if choice == start:
        assert np.array_equal(p1, p0) and np.array_equal(v1, v0)
    else:
        assert np.abs(d).max() < 1e-11


def test_kernel_shape_by_size():
    # 8 lanes per row up to 65,536 atoms for EAM (global-table passes) and
    # 131,072 for Morse; one lane (EAM: the LDS passes) above
    s = fcc(5)
    for params, want in ((potential("johnson"), 8), (MORSE_CU, 8)):
        sched = FastPipelineScheduler.for_system(s, params, FastPipelineConfig())
        assert sched.kernel_shape()["pair_lanes"] == want
        sched.close()
    sched = FastPipelineScheduler.for_system(s, potential("johnson"),
                                             FastPipelineConfig(pair_lanes=1))
    shape = sched.kernel_shape()
    assert shape["pair_lanes"] == 1 and shape["lds_budget"] > 0
    sched.close()


@pytest.mark.parametrize("pot", ["johnson", "morse13"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fused_rebuild_equals_separate_kernels(prec, pot):
    # the one-launch rebuild and the nine-kernel rebuild build the same sort
    # and the same rows: trajectories are bitwise identical. Systems this
    # small take the 32-lane list search in the fused rebuild and one lane
    # per atom in the separate kernels. morse13: a 13 A cutoff, ~290 atoms
    # per cell (several 32-candidate rounds per cell) and ~1000-slot rows.
    s = fcc(6) if pot == "johnson" else fcc(13)
    params = potential("johnson") if pot == "johnson" else MORSE_CU[:3] + (13.0,)
    say("fused rebuild", prec, pot)
    out = []
    for split in (False, True):
        cfg = FastPipelineConfig(precision=prec, rebuild_every=5, split_rebuild=split)
        sched = FastPipelineScheduler.for_system(s, params, cfg)
        sched.run_until(30)
        pos, vel, f, _, _ = sched.download()
        out.append((pos, vel, f, sched.stats()))
        sched.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][3].rebuilds == out[1][3].rebuilds == 1 + 5
    assert out[0][3].kernel_launches < out[1][3].kernel_launches


def test_long_cutoff_morse_matches_oracle():
    # a 13 A Morse cutoff: ~290 atoms per cell, ~1000 neighbours per row and
    # three cells per axis, so the 27 cells around any atom are the whole
    # box: rows, padding and capacity sizing far from the metal benchmarks
    s = fcc(13)
    morse = MORSE_CU[:3] + (13.0,)
    say("long-cutoff morse", s.natoms)
    sched = FastPipelineScheduler.for_system(s, morse, FastPipelineConfig(precision="fp64"))
    sched.run_until(12)
    pos, vel, f, _, _ = sched.download()
    rpos, rvel, rf, _, reb = O.run(s, morse=morse, nsteps=12, thermo_every=12, rebuild_mode=1,
                                   rebuild_every=10)
    L = s.hi - s.lo
    dpos = pos - rpos
    dpos -= np.round(dpos / L) * L
    say("  max dpos", np.abs(dpos).max(), "force err", force_rel_err(f, rf))
    assert np.abs(dpos).max() < 1e-10
    assert force_rel_err(f, rf) < 1e-9
    assert sched.stats().rebuilds == 2  # initial + step 11


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_open_axis_slab_matches_oracle(prec):
    # (001) slab: z open with vacuum above and below the crystal
    s = fcc(6)
    s.lo[2] -= 6.0
    s.hi[2] += 6.0
    s.periodic = (1, 1, 0)
    say("open-z slab", prec)
    sched, pos, vel, f = gpu_run(s, "johnson", 40, precision=prec)
    rpos, rvel, rf, _, _ = oracle_run(s, "johnson", 40)
    L = s.hi - s.lo
    dpos = pos - rpos
    dpos[:, :2] -= np.round(dpos[:, :2] / L[:2]) * L[:2]
    err = force_rel_err(f, rf)
    say("  max dpos", np.abs(dpos).max(), "force err", err)
    if prec == "fp64":
        assert np.abs(dpos).max() < 1e-10 and err < 1e-9
    else:
        assert np.abs(dpos).max() < 1e-5 and err < 1e-4


@pytest.mark.parametrize("r_c", [None, 8.5])
def test_halo_staging_is_bitwise_neutral(r_c):
    # staging changes where the passes read neighbour records from (LDS), not
    # which records or in what order: trajectories with and without it are
    # bitwise identical. At r_c 8.5 A a block's halo (~6000 atoms) exceeds
    # the LDS capacity and the blocks fall back to their 32-bit rows.
    s = fcc(12)  # 6912 atoms: six full blocks of 1024 and one of 768
    tables = (potential("johnson") if r_c is None else
              I.johnson_setfl(I.JohnsonParams(r_c=r_c), nr=2000, nrho=2000))
    say("halo staging", r_c)
    out = []
    for off in (False, True):
        cfg = FastPipelineConfig(precision="mixed", rebuild_every=5, disable_halo=off,
                                 pair_lanes=1)
        sched = FastPipelineScheduler.for_system(s, tables, cfg)
        sched.run_until(20)
        pos, vel, f, _, _ = sched.download()
        out.append((pos, vel, f, sched.halo_info()))
        sched.close()
    info = out[0][3]
    say("  halos", info)
    assert info["blocks"] == 7 and out[1][3]["blocks"] == 0
    if r_c is None:
        assert info["staged"] == 7 and info["max_halo"] <= 4608
    else:
        assert info["staged"] < 7
    for k in range(3):
        assert np.array_equal(out[0][k], out[1][k])


@pytest.mark.parametrize("name", ["alloy2", "alloy3"])
def test_alloy_matches_reference_driver_golden(name):
    # two- and three-element EAM on random-species crystals against the
    # reference CPU driver's own outputs (tests/golden/ref_alloy*.npz): the
    # FP64 scheduler's 100-step trajectory, and the mixed step-0 forces
    from golden.make_golden import alloy_tables

    g = load_golden(f"ref_{name}")
    s = system_from_golden(g)
    tables = alloy_tables(name)
    nsteps = int(g["thermo"][-1, 0])
    say("alloy vs reference golden", name, s.natoms)
    sched = FastPipelineScheduler.for_system(
        s, tables, FastPipelineConfig(precision="fp64", rebuild_policy="displacement"))
    sched.run_until(nsteps)
    pos, vel, f, _, _ = sched.download()
    sched.close()
    L = s.hi - s.lo
    dpos = pos - g["pos_final"]
    dpos -= np.round(dpos / L) * L
    say("  fp64 max dpos", np.abs(dpos).max(), "force err", force_rel_err(f, g["f_final"]))
    assert np.abs(dpos).max() < 1e-10
    assert force_rel_err(f, g["f_final"]) < 1e-9
    sched = FastPipelineScheduler.for_system(s, tables, FastPipelineConfig(precision="mixed"))
    sched.run_until(0)
    f0 = sched.download()[2]
    sched.close()
    err = force_rel_err(f0, g["f0"])
    say("  mixed step-0 force err", err)
    assert err < 1e-5


@pytest.mark.parametrize("pot", ["johnson", "morse", "alloy"])
def test_row_split_is_deterministic_and_agrees_with_one_lane(pot):
    # 1, 4 and 8 lanes per row sum a row's pairs in different (fixed)
    # orders: each repeats bit for bit, and the FP64 trajectories agree to
    # rounding; the EAM one-lane run is the LDS pass, the split runs the
    # global-table passes
    if pot == "alloy":
        from golden.make_golden import alloy_system, alloy_tables

        s, params = alloy_system("alloy3"), alloy_tables("alloy3")
    else:
        s, params = fcc(6), potential(pot)
    say("row split", pot)
    out = {}
    for lanes in (1, 4, 8, 8):
        cfg = FastPipelineConfig(precision="fp64", rebuild_every=5, pair_lanes=lanes)
        sched = FastPipelineScheduler.for_system(s, params, cfg)
        sched.run_until(40)
        pos, vel, f, _, _ = sched.download()
        sched.close()
        if lanes in out:
            assert np.array_equal(out[lanes][0], pos) and np.array_equal(out[lanes][2], f)
        out[lanes] = (pos, vel, f)
    L = s.hi - s.lo
    for lanes in (4, 8):
        d = out[lanes][0] - out[1][0]
        d -= np.round(d / L) * L
        say(f"  lanes {lanes}: max |dx| vs one lane {np.abs(d).max():.2e}")
        assert np.abs(d).max() < 1e-11
        assert force_rel_err(out[lanes][2], out[1][2]) < 1e-11


@pytest.mark.parametrize("pot", ["johnson", "morse"])
def test_pass_timing_leaves_the_trajectory_unchanged(pot):
    # events recorded around the passes inside the captured graphs time the
    # passes of the window itself and change nothing in it
    s = fcc(6)
    say("pass timing", pot)
    out = []
    for on in (False, True):
        cfg = FastPipelineConfig(precision="mixed", pair_lanes=1,
                                 t_target=300.0 if pot == "johnson" else 0.0)
        sched = FastPipelineScheduler.for_system(s, potential(pot), cfg)
        sched.pass_timing(on)
        sched.run_until(37)
        out.append(sched.download()[:3])
        if on:
            dens, force = sched.pass_times()
            say(f"  density {dens:.4f} ms, force {force:.4f} ms")
            assert 0.0 < force < 50.0
            assert (0.0 < dens < 50.0) if pot == "johnson" else dens == 0.0
        sched.close()
    for k in range(3):
        assert np.array_equal(out[0][k], out[1][k])


def test_cell_size_tuner_agrees_to_rounding():
    # the cell size changes the spatial sort and so the order of the row
    # sums, never the pairs: the tuner's pick and the default give the same
    # FP64 trajectory to rounding
    from tdmd_claude_amd.scheduler import tune_cell_size

    s = fcc(9)  # 9 cells of 3.6 A per axis at 100 %, 6 at 150 %
    say("cell-size tuner")
    res = tune_cell_size(s, potential("johnson"), FastPipelineConfig(precision="fp64"),
                         candidates=(100, 150), steps=10)
    say("  ", res)
    assert res["cell_pct"] in (100, 150) and set(res["ms_per_step"]) == {100, 150}
    out = []
    for pct in (100, 150):
        sched = FastPipelineScheduler.for_system(
            s, potential("johnson"), FastPipelineConfig(precision="fp64", cell_pct=pct))
        sched.run_until(40)
        out.append(sched.download()[:3])
        sched.close()
    L = s.hi - s.lo
    d = out[1][0] - out[0][0]
    d -= np.round(d / L) * L
    say(f"  max |dx| 150 % vs 100 %: {np.abs(d).max():.2e}")
    assert np.abs(d).max() < 1e-10
    assert force_rel_err(out[1][2], out[0][2]) < 1e-10


@pytest.mark.parametrize("mirror_fp64,lz", [(False, 26.0), (True, 26.0), (False, 29.5)])
def test_list_counts_match_oracle_at_the_cutoffs(mirror_fp64, lz):
    # The list search decides most candidates from the 32-bit fixed-point
    # records in FP32, with the exact FP64 test inside a 1e-5 relative band
    # around r_list^2 and rc^2. Pairs placed 1e-7 from r_list and from rc
    # (inside the band) and 3e-5 from them (outside it): the rows' total,
    # near total and longest row must be the oracle's exact lists, at the
    # first build (one lane per atom) and after a fused rebuild (32 lanes).
    # A cubic box takes the prefilter in units^2 (one unit on all axes); a
    # box longer in z has a larger unit there and takes the per-axis scaling.
    rng = np.random.Generator(np.random.PCG64(17))
    rc, skin = 6.0, 1.0
    L = np.array([26.0, 26.0, lz])
    r_list = rc + skin
    base = rng.uniform(0.0, L, size=(300, 3))
    extra = []
    for d in (r_list * (1 - 1e-7), r_list * (1 + 1e-7), rc * (1 - 1e-7), rc * (1 + 1e-7),
              r_list * (1 - 3e-5), r_list * (1 + 3e-5), rc * (1 - 3e-5), rc * (1 + 3e-5)):
        for _ in range(6):
            p = rng.uniform(0.0, L, size=3)
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            extra += [p, np.mod(p + d * u, L)]
    pos = np.concatenate([base, np.array(extra)])
    s = I.System(lo=np.zeros(3), hi=L.copy(), pos=pos, vel=np.zeros_like(pos),
                 types=np.ones(len(pos), dtype=np.int32),
                 ids=np.arange(1, len(pos) + 1, dtype=np.int32), masses=np.array([1.0, 63.546]))
    say("list counts at the cutoffs, mirror_fp64", mirror_fp64, s.natoms)
    full = O.HalfList(s.pos, s.lo, s.hi, rc, skin).counts
    near = O.HalfList(s.pos, s.lo, s.hi, rc, 0.0).counts
    sched = FastPipelineScheduler.for_system(
        s, MORSE_CU[:3] + (rc,), FastPipelineConfig(precision="mixed", mirror_fp64=mirror_fp64,
                                                    rebuild_every=1, r_skin=skin, dt=1e-9))
    for step in (0, 1):
        sched.run_until(step)
        if step == 1:  # the fused rebuild, after a step too short to move the pairs off
            pos1 = sched.download()[0]
            full = O.HalfList(pos1, s.lo, s.hi, rc, skin).counts
            near = O.HalfList(pos1, s.lo, s.hi, rc, 0.0).counts
        obs = sched.observables()
        n_full = 2 * int(np.sum(full))
        say(f"  step {step}: rows {obs['mean_neighbors'] * s.natoms:.0f} vs {n_full}, near "
            f"{obs['mean_near_neighbors'] * s.natoms:.0f} vs {2 * int(np.sum(near))}")
        assert round(obs["mean_neighbors"] * s.natoms) == n_full
        assert round(obs["mean_near_neighbors"] * s.natoms) == 2 * int(np.sum(near))
    sched.close()
