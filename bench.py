#!/usr/bin/env python3
"""Throughput of the FastPipelineScheduler step on one MI355X.

Workload (the 1M-atom EAM configuration): FCC Cu 64x64x64 cells
(1,048,576 atoms, a = 3.615 A, box 231.36 A), Johnson-form Cu EAM tabulated
on 5000-point FD-Hermite splines (cutoff 5.5 A), skin 1.0 A with the
skin/2 displacement rebuild, 300 K Maxwell-Boltzmann velocities (seed 7),
NVT Nose-Hoover chain of 3 (tau 0.1 ps, 300 K), dt 1 fs. Synthetic inputs,
state resident in HBM before the timed region. Each timed window is
preceded by the warm-up and one untimed window of the same shape, so the
timed window only replays captured hipGraphs.

One JSON line on stdout: the mixed-precision throughput (value), the FP64
path's throughput over the same windows, the force error of the mixed path
against the FP64 CPU oracle, the NVE energy drift (mixed and FP64, from a
fixed thermalised state), a roofline object for the dominant kernel (the
force pass of the timed NVT step) and a CPU baseline (the reference's own
FP64 CPU driver, compiled from its sources, on one core). Under "extras":
the 32,000-atom Morse NVE workload, the reference's own tiny / small /
medium systems beside its published timesteps per second, 64 batched alloy
replicas, and force matching with the adjoint gradient.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tdmd_claude_amd import capi  # noqa: E402
from tdmd_claude_amd import inputs as I  # noqa: E402
from tdmd_claude_amd.scheduler import FastPipelineConfig, FastPipelineScheduler  # noqa: E402

METRIC = "Matom-steps/s, 1M-atom EAM Cu, mixed vs FP64 at ≤1e-5 rel force error"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


THERMAL_STEPS = 1100  # NVT steps from the lattice to the state of the pass timing and drift


def make_system(cells: int) -> I.System:
    s = I.fcc_lattice(cells, cells, cells, 3.615)
    I.init_velocities(s, 300.0, 7)
    return s


def nvt_cfg(precision: str) -> FastPipelineConfig:
    return FastPipelineConfig(dt=0.001, r_skin=1.0, precision=precision,
                              rebuild_policy="displacement", t_target=300.0, t_period=0.1,
                              nhc_length=3, use_graph=True)


def timed_run(sched: FastPipelineScheduler, warmup: int, steps: int,
              pass_clock: bool = False) -> tuple[float, float]:
    """Warm up, run one untimed window of the timed window's shape (so every
    hipGraph the timed window replays is captured and instantiated before
    it), then time exactly `steps` steps: host clock and HIP events on the
    scheduler's stream, device synchronised on both sides. pass_clock: after
    the timed window, one more window of the same shape in which the passes
    time themselves on the device clock (its graphs captured in a
    pre-capture window of their own), so the instrumentation's atomics stay
    out of the timed window; read the result with sched.pass_times()."""
    sched.run_until(warmup)
    if pass_clock:  # capture the instrumented graphs
        sched.pass_timing(True)
        sched.run_until(sched.current_step() + steps)
        sched.pass_timing(False)
    sched.run_until(sched.current_step() + steps)  # graph pre-capture, untimed
    start = sched.current_step()
    stream = sched.stream()
    e0, e1 = capi.C.c_void_p(), capi.C.c_void_p()
    capi.call("tdmd_event_create", capi.C.byref(e0))
    capi.call("tdmd_event_create", capi.C.byref(e1))
    capi.call("tdmd_device_synchronize")
    t0 = time.perf_counter()
    capi.call("tdmd_event_record", e0, stream)
    sched.run_until(start + steps)  # ends with a stream synchronisation
    capi.call("tdmd_event_record", e1, stream)
    capi.call("tdmd_device_synchronize")
    wall = time.perf_counter() - t0
    ms = capi.C.c_float()
    capi.call("tdmd_event_elapsed_ms", e0, e1, capi.C.byref(ms))
    capi.call("tdmd_event_destroy", e0)
    capi.call("tdmd_event_destroy", e1)
    if pass_clock:  # the passes of the next window of the same shape
        sched.pass_timing(True)
        sched.run_until(start + 2 * steps)
    return wall, ms.value / 1000.0


def thermal_state(system: I.System, tables) -> I.System:
    """The workload's state after THERMAL_STEPS NVT steps (mixed) from the
    lattice: the fixed starting point of the pass timing, the drift legs and
    the profile runs, independent of --steps and --warmup."""
    sched = FastPipelineScheduler.for_system(system, tables, nvt_cfg("mixed"))
    sched.run_until(THERMAL_STEPS)
    pos, vel, _, _, _ = sched.download()
    sched.close()
    state = system.copy()
    state.pos, state.vel = pos, vel
    return state


def bytes_per_atom(precision: str, mirror_fp64: bool, nbar: float, staged: float = 0.0,
                   halo: float = 0.0, nnear: float | None = None) -> dict:
    """Bytes per atom of the two steady-state passes in this build's own
    layout (the compressed model, beside the reference-layout figure of
    alg_bytes_force): each per-atom array touched once per pass, the
    neighbour rows read once per pass, neighbour records served from cache
    (L2, or LDS for staged blocks). Rows are 4 bytes a slot; for the
    fraction `staged` of blocks whose halo sits in LDS they are 2 bytes a
    slot plus the 4-byte halo list (`halo` atoms per atom), and the force
    pass reads only the `nnear` slots inside the cutoff at the build (its
    list of skin-shell slots inside the cutoff, 0-3 entries, is not
    counted). Build positions for the skin test are 32-bit coordinates with
    the fixed-point mirror, FP64 otherwise."""
    fixed = precision == "mixed" and not mirror_fp64
    rec = 16 if fixed else 32  # M / M2 record
    xb = 12 if fixed else 24  # positions at the last build
    fb = 12 if precision == "mixed" else 24  # forces
    nn = nbar if nnear is None else nnear
    rows_d = staged * (2 * nbar + 4 * halo) + (1 - staged) * 4 * nbar
    rows_f = staged * (2 * nn + 4 * halo) + (1 - staged) * 4 * nbar
    density = rec + 4 + rows_d + rec  # read M_i, count, rows; write M2_i
    # (one element: the integrator reads no types)
    # NVE force pass: read M2_i, count, x, v, x_build, rows; write x, v, M_i
    force_nve = rec + 4 + 24 + 24 + xb + rows_f + 24 + 24 + rec
    # NVT force pass inside a window: read M2_i, count, v, rows; write F (the
    # kicked v only enters the KE; the next kick-drift applies that kick)
    force_nvt = rec + 4 + 24 + rows_f + fb
    return dict(density=density, force_nve=force_nve, force_nvt=force_nvt)


def alg_bytes_force(precision: str, nbar: float) -> float:
    """Algorithmic bytes per atom of the EAM force pass in the reference's
    data layout: read x (24), type (4), F'(rho) (4 / 8), the row offset and
    count (8) and the int32 full-list row (4 nbar); write F (12 / 24)."""
    return (52.0 if precision == "mixed" else 68.0) + 4.0 * nbar


def alg_bytes_step(precision: str, nbar: float, n_rb: float) -> float:
    """Algorithmic bytes per atom-step of the whole EAM step in the
    reference's layout: kick-drift, density, embedding and force passes, and
    the list rebuild spread over the n_rb steps between rebuilds
    (220 + 8 nbar + (104 + 4 nbar) / n_rb mixed, 252 + ... FP64)."""
    base = 220.0 if precision == "mixed" else 252.0
    return base + 8.0 * nbar + (104.0 + 4.0 * nbar) / max(n_rb, 1.0)


def energy_slope(t_ps, e, discard: float = 0.25) -> float:
    """Least-squares slope of total energy per atom against time after the
    first `discard` of the run (the transient), in eV/atom/ns: the method of
    the reference's NVE drift check (verifylab/cases/nve-drift/check.py)."""
    t = np.asarray(t_ps)
    e = np.asarray(e)
    keep = t >= t[0] + discard * (t[-1] - t[0])
    return float(np.polyfit(t[keep], e[keep], 1)[0] * 1000.0)


def nve_drift(system: I.System, potential, precision: str, steps: int, every: int) -> dict:
    """NVE from the given state (skin/2 displacement rebuild): total energy
    per atom sampled every `every` steps, slope after a 25 % transient."""
    cfg = FastPipelineConfig(dt=0.001, r_skin=1.0, precision=precision,
                             rebuild_policy="displacement")
    sched = FastPipelineScheduler.for_system(system, potential, cfg)
    sched.pass_timing(True)
    t, e = [], []
    for k in range(0, steps + 1, every):
        sched.run_until(k)
        th = sched.thermo()
        t.append(k * 0.001)
        e.append((th["pe"] + th["ke"]) / system.natoms)
    rebuilds = sched.stats().rebuilds
    force_ms = sched.pass_times()[1]
    sched.close()
    e = np.array(e)
    return dict(drift_ev_per_atom_per_ns=energy_slope(t, e),
                max_dev_ev_per_atom=float(np.abs(e - e[0]).max()), steps=steps,
                sample_every=every, discarded="first 25 %", precision=precision,
                rebuilds=rebuilds, force_pass_ms=force_ms)


def bench_morse_c2(steps: int) -> dict:
    """32,000-atom Morse Cu (20^3 cells, 300 K, +-0.02 A jitter), NVE, skin/2
    displacement rebuild: mixed and FP64 throughput, and the energy drift of
    each over a separate run of the same length (energy every 500 steps,
    slope after a 25 % transient)."""
    s = I.fcc_lattice(20, 20, 20, 3.615)
    I.init_velocities(s, 300.0, 42)
    I.jitter_uniform(s, 0.02, 4242)
    morse = (0.3429, 1.3588, 2.866, 6.0)
    out = {"workload": "Morse Cu 20x20x20 FCC cells (32,000 atoms), NVE, dt 1 fs, skin 1 A, "
                       "skin/2 displacement rebuild", "steps": steps}
    for prec in ("mixed", "fp64"):
        cfg = FastPipelineConfig(dt=0.001, r_skin=1.0, precision=prec,
                                 rebuild_policy="displacement")
        sched = FastPipelineScheduler.for_system(s, morse, cfg)
        wall, _ = timed_run(sched, 100, steps)
        out[prec] = {"value": round(s.natoms * steps / wall / 1e6, 2),
                     "ms_per_step": round(1000 * wall / steps, 4),
                     "rebuilds": sched.stats().rebuilds}
        sched.close()
        d = nve_drift(s, morse, prec, steps, 500)
        out[prec]["drift_ev_per_atom_per_ns"] = d["drift_ev_per_atom_per_ns"]
        out[prec]["drift_max_dev_ev_per_atom"] = d["max_dev_ev_per_atom"]
    out["mixed_over_fp64"] = round(out["mixed"]["value"] / out["fp64"]["value"], 3)
    return out


# The reference's published mixed-precision timesteps per second for its
# tiny / small / medium Cu systems (RTX 5080, CUDA 12.6), the best figure
# published for each: Morse with the relative-coordinate trick
# (docs/06-decisions/0007-precision-contract.md:369-371), EAM from the
# LAMMPS A/B session (docs/05-benchmarks/lammps-ab-results.md:38-52).
REF_SIZES = (("tiny", 4), ("small", 10), ("medium", 20))
REF_TS = {"morse": (14814, 9139, 6927), "eam": (4802, 3288, 2645)}


def bench_reference_sizes(steps: int) -> dict:
    """The reference's own benchmark systems: FCC Cu (a = 3.615 A) of 4^3,
    10^3 and 20^3 cells (256, 4000, 32000 atoms), Morse (D 0.3429 eV,
    alpha 1.3588 1/A, r0 2.866 A, rc 6.0 A) and Mishin Cu EAM (its setfl:
    10001-point tables, rc 5.50679 A), mixed precision, NVE, dt 1 fs, skin
    1.0 A, list rebuilt every 10 steps, 300 K velocities (seed 42). Timesteps
    per second beside the reference's published figures for the same
    systems."""
    eam = I.read_setfl(os.path.join(ROOT, "tests", "golden", "Cu_mishin1.eam.alloy"))
    morse = (0.3429, 1.3588, 2.866, 6.0)
    out = {"workload": "FCC Cu 256 / 4000 / 32000 atoms, Morse and Mishin Cu EAM, mixed, NVE, "
                       "dt 1 fs, skin 1 A, rebuild every 10 steps",
           "steps": steps, "reference_hardware": "RTX 5080, CUDA 12.6"}
    for pot, params in (("morse", morse), ("eam", eam)):
        rows = {}
        for (name, cells), ref_ts in zip(REF_SIZES, REF_TS[pot]):
            s = I.fcc_lattice(cells, cells, cells, 3.615)
            I.init_velocities(s, 300.0, 42)
            cfg = FastPipelineConfig(dt=0.001, r_skin=1.0, precision="mixed",
                                     rebuild_policy="fixed", rebuild_every=10)
            sched = FastPipelineScheduler.for_system(s, params, cfg)
            wall, _ = timed_run(sched, 100, steps)
            ts = steps / wall
            rows[name] = {"natoms": s.natoms, "ts_per_s": round(ts, 1),
                          "matom_steps_per_s": round(s.natoms * ts / 1e6, 2),
                          "reference_ts_per_s": ref_ts, "over_reference": round(ts / ref_ts, 2)}
            sched.close()
        out[pot] = rows
    return out


def bench_replicas(steps: int) -> dict:
    """64 replicas of random equiatomic Ni/Co/Cr FCC alloys (a = 3.57 A, sizes
    log-uniform in [500, 32000], T ~ U(300, 1200) K, master seed 2024),
    three-element Johnson-form EAM (5000-point tables, cutoff 5.0 A), NVE,
    dt 1 fs, all replicas advanced together (one launch per pass), shape
    chosen by the autotuner: mixed and FP64 throughput."""
    from tdmd_claude_amd.scheduler import BatchScheduler, tune_cell_size

    systems, temps = I.alloy_replicas(64, 2024)
    tables = I.johnson_alloy_setfl()
    natoms = sum(s.natoms for s in systems)
    out = {"workload": "64 Ni/Co/Cr FCC replicas, sizes log-uniform in [500, 32000] atoms, "
                       "3-element EAM (5000-point tables, rc 5.0 A), NVE, dt 1 fs, skin 1 A "
                       "displacement rebuild", "natoms": natoms, "replicas": len(systems),
           "smallest": min(s.natoms for s in systems), "largest": max(s.natoms for s in systems),
           "steps": steps}
    for prec in ("mixed", "fp64"):
        cfg = FastPipelineConfig(dt=0.001, r_skin=1.0, precision=prec,
                                 rebuild_policy="displacement")
        # tuner: the binning cell size on trial schedulers, then the kernel
        # shape (lanes per row) in place
        cells = tune_cell_size(systems, tables, cfg, candidates=(100, 125, 150))
        cfg.cell_pct = cells["cell_pct"]
        b = BatchScheduler.for_systems(systems, tables, cfg)
        b.run_until(0)
        tuned = b.autotune(5)
        tuned["cell_pct"] = cells["cell_pct"]
        tuned["cell_ms_per_step"] = {k: round(v, 4) for k, v in cells["ms_per_step"].items()}
        wall, _ = timed_run(b, 50, steps)
        out[prec] = {"value": round(natoms * steps / wall / 1e6, 2),
                     "ms_per_step": round(1000 * wall / steps, 4),
                     "rebuilds": b.stats().rebuilds, "autotune": tuned}
        b.close()
    out["mixed_over_fp64"] = round(out["mixed"]["value"] / out["fp64"]["value"], 3)
    return out


def bench_force_matching(iters: int) -> dict:
    """Force matching over 256 snapshots of 10^3-cell Cu (4,000 atoms, N(0, 0.05^2)
    displacements, seeds 5..260): loss+adjoint evaluation time for Morse and for the
    5000-point EAM knot values (mixed), and Adam fits of `iters` iterations: Morse
    (D, alpha, r0), and the EAM tables (relative hat basis, 64 nodes per function)
    from the unperturbed tables to targets of a perturbed Johnson parameter set."""
    import torch

    from tdmd_claude_amd import fitting

    base, snaps = fitting.fc_snapshots(10, 256, 0.05, 5)
    t0 = time.perf_counter()
    fm_m = fitting.ForceMatchSet(base.lo, base.hi, snaps, 6.0, "mixed")
    setup = time.perf_counter() - t0
    morse_t = (0.36, 1.30, 2.90)
    fm_m.set_targets(fm_m.morse_forces(morse_t, 6.0))
    theta = (0.3429, 1.3588, 2.866)
    fm_m.morse(theta, 6.0)
    n_eval = 20
    capi.call("tdmd_device_synchronize")
    t0 = time.perf_counter()
    for _ in range(n_eval):
        fm_m.morse(theta, 6.0)
    morse_ms = 1000 * (time.perf_counter() - t0) / n_eval
    x0 = torch.tensor(theta, dtype=torch.float64)
    _, mlosses = fitting.fit(lambda p: fitting.MorseForceMatch.apply(p, fm_m, 6.0), x0,
                             iters=iters, lr=5e-3)
    npairs_m = fm_m.npairs()
    fm_m.close()

    tab = I.johnson_setfl()
    values, grid = fitting.eam_values(tab)
    pt = I.JohnsonParams(beta=4.5397 * 1.02, E0=3.4842 * 0.98, D_p=0.0592 * 1.03)
    vt, gt = fitting.eam_values(I.johnson_setfl(pt))
    fm_e = fitting.ForceMatchSet(base.lo, base.hi, snaps, grid.cutoff, "mixed")
    dev = torch.device("cuda")
    _, _, ftgt = fm_e.eam(torch.tensor(vt, dtype=torch.float64, device=dev), gt,
                          want_loss=False, want_grad=False, want_forces=True)
    fm_e.set_targets(ftgt.cpu().numpy())
    v0 = torch.tensor(values, dtype=torch.float64, device=dev)
    fm_e.eam(v0, grid)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n_eval):
        fm_e.eam(v0, grid)
    torch.cuda.synchronize()
    eam_ms = 1000 * (time.perf_counter() - t0) / n_eval
    basis = fitting.RelativeHatBasis(v0, grid, nodes=64)
    t0 = time.perf_counter()
    _, elosses = fitting.fit(
        lambda c: fitting.EamForceMatch.apply(basis.expand(c), fm_e, grid), basis.zeros(),
        iters=iters, lr=1e-3)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    npairs_e = fm_e.npairs()
    fm_e.close()
    natoms = fm_e.ntot
    return {
        "workload": "256 snapshots x 4,000-atom FCC Cu (sigma 0.05 A), mixed, forces + "
                    "adjoint gradient per evaluation",
        "atoms": natoms, "list_setup_s": round(setup, 3),
        "morse": {"eval_ms": round(morse_ms, 3), "pairs": npairs_m,
                  "matom_evals_per_s": round(natoms / morse_ms / 1e3, 1),
                  "adam_iters": iters, "loss_start": mlosses[0], "loss_end": mlosses[-1]},
        "eam": {"eval_ms": round(eam_ms, 3), "pairs": npairs_e,
                "matom_evals_per_s": round(natoms / eam_ms / 1e3, 1),
                "parameters": grid.nvalues, "fit_basis": "relative hat, 64 nodes per function",
                "adam_iters": iters,
                "fit_s": round(fit_s, 3), "loss_start": elosses[0], "loss_end": elosses[-1]},
    }


def profile_doc(name: str, natoms: int):
    """A committed rocprofv3 summary under profiles/ (JSON), when it was taken
    on this workload; else None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        doc = json.load(fh)
    return doc if doc.get("natoms") == natoms else None


def measured_traffic(natoms: int, mirror_fp64: bool, kernel: str):
    """HBM bytes per launch of a pass from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (profiles/eam_passes_1M_mixed_traffic.json)."""
    doc = None if mirror_fp64 else profile_doc("eam_passes_1M_mixed_traffic.json", natoms)
    if doc is None or kernel not in doc["kernels"]:
        return None
    k = doc["kernels"][kernel]
    return {"bytes": round(k["hbm_bytes"]), "bytes_per_atom": round(k["bytes_per_atom"], 1),
            "source": "profiles/eam_passes_1M_mixed_traffic.json"}


def measured_flops(natoms: int, mirror_fp64: bool, kernel: str):
    """FP32 and FP64 operations per launch of a pass from the committed
    rocprofv3 SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F{32,64} runs
    (profiles/eam_passes_1M_mixed_valu.json): 64 lanes per wave-level
    instruction, two operations per FMA."""
    doc = None if mirror_fp64 else profile_doc("eam_passes_1M_mixed_valu.json", natoms)
    if doc is None or kernel not in doc["kernels"]:
        return None
    k = doc["kernels"][kernel]
    return {"fp32_ops": k["fp32_ops"], "fp64_ops": k["fp64_ops"],
            "source": "profiles/eam_passes_1M_mixed_valu.json"}


REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "tdmd_standalone")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_reference(cells: int, steps: int) -> dict:
    """The reference's own CPU driver (src/drivers/tdmd_main.cpp, FP64, one
    thread), compiled from the reference sources into oracle/_ref by
    oracle/Makefile, on the workload's input: the lattice with its 300 K
    velocities and the same EAM tables as a setfl file. Two runs, 0 and
    `steps` steps; their difference is the cost of the steps (the input
    parsing, the first list build and the first forces cancel)."""
    import shutil
    import subprocess
    import tempfile

    system = make_system(cells)
    N = system.natoms
    tmp = tempfile.mkdtemp(prefix="tdmd_cpu_")
    try:
        data = os.path.join(tmp, "c3.data")
        setfl = os.path.join(tmp, "c3.eam.alloy")
        I.write_lammps_data(data, system)
        I.write_setfl(setfl, I.johnson_setfl())
        cpus = sorted(os.sched_getaffinity(0))
        os.sched_setaffinity(0, {cpus[-1]})  # this process and the driver: one core
        times, rebuilds = [], []
        for n in (0, steps):
            cmd = [REF_DRIVER, "--data", data, "--eam", setfl, "--nsteps", str(n),
                   "--thermo", str(max(n, 1))]
            t0 = time.perf_counter()
            res = subprocess.run(cmd, capture_output=True, text=True, timeout=1800)
            times.append(time.perf_counter() - t0)
            if res.returncode != 0:
                raise RuntimeError(f"reference driver failed: {res.stderr[-500:]}")
            reb = [ln for ln in res.stdout.splitlines() if ln.startswith("Neighbor list")]
            rebuilds.append(int(reb[-1].split()[-1]) if reb else -1)
            log(f"cpu baseline: reference driver, {n} steps, {times[-1]:.1f} s")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    per_step = (times[1] - times[0]) / steps
    return dict(
        value=N / per_step / 1e6, unit="Matom-steps/s", cores=1, kind="reference",
        sample=(f"{N}-atom EAM Cu lattice with 300 K velocities (the workload's input), {steps} "
                "NVE steps of the reference driver tdmd_standalone (FP64, built from "
                "src/drivers/tdmd_main.cpp and its CPU sources), timed as the difference of a "
                f"{steps}-step and a 0-step run: {per_step:.2f} s/step; its skin test rebuilt "
                f"the list {rebuilds[1] - rebuilds[0]} times in {steps} steps (it measures "
                "displacement without the minimum image, so an atom crossing the periodic "
                "boundary triggers a rebuild); one core of "
                f"{len(cpus)}, {cpu_model()}, in its own process beside the GPU legs"),
        cpu_model=cpu_model(), host_cpus=len(cpus), s_per_step=per_step)


class CpuBaseline:
    """The CPU baseline leg. With the reference driver built, cpu_reference()
    runs in a child process on one core, started before the GPU legs and
    joined after them (its own interpreter: no lock shared with the timed
    GPU windows). Otherwise the oracle's C restatement of the same loop
    (kind "port") is timed in this process at the end."""

    def __init__(self, system: I.System, tables, cells: int, steps: int):
        import subprocess

        self.system, self.tables, self.steps = system, tables, steps
        self.proc = None
        if os.access(REF_DRIVER, os.X_OK):
            self.proc = subprocess.Popen(
                [sys.executable, os.path.abspath(__file__), "--cpu-only", "--cells", str(cells),
                 "--cpu-steps", str(steps)], stdout=subprocess.PIPE, text=True)

    def join(self):
        import subprocess

        if self.proc is None:
            return self._run_port()
        while True:
            try:
                out, _ = self.proc.communicate(timeout=30)
                break
            except subprocess.TimeoutExpired:
                log("cpu baseline: reference driver still running")
        if self.proc.returncode != 0:
            raise RuntimeError(f"cpu baseline failed (exit {self.proc.returncode})")
        return json.loads(out.strip().splitlines()[-1])

    def _run_port(self):
        from oracle import oracle as O

        eam = O.Eam(tables=self.tables)
        N = self.system.natoms
        t0 = time.perf_counter()
        O.run(self.system, eam=eam, nsteps=0, thermo_every=1)
        t1 = time.perf_counter()
        O.run(self.system, eam=eam, nsteps=self.steps, thermo_every=self.steps)
        t2 = time.perf_counter()
        per_step = ((t2 - t1) - (t1 - t0)) / self.steps
        return dict(value=N / per_step / 1e6, unit="Matom-steps/s", cores=1, kind="port",
                    sample=(f"{N}-atom EAM Cu lattice with 300 K velocities, {self.steps} steps "
                            "of the oracle's FP64 driver loop (the reference driver was not "
                            f"built here), {per_step:.2f} s/step; one core, {cpu_model()}"),
                    cpu_model=cpu_model(), host_cpus=len(os.sched_getaffinity(0)),
                    s_per_step=per_step)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--nve-steps", type=int, default=20000,
                    help="NVE steps of each 1M-atom drift leg (mixed and FP64)")
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="steps of the CPU baseline (reference driver)")
    ap.add_argument("--mirror-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--c2-steps", type=int, default=10000,
                    help="Morse 32k-atom NVE steps per precision (0: skip)")
    ap.add_argument("--ref-steps", type=int, default=5000,
                    help="steps per system of the reference-size legs (0: skip)")
    ap.add_argument("--replica-steps", type=int, default=1000,
                    help="64-replica alloy steps per precision (0: skip)")
    ap.add_argument("--fm-iters", type=int, default=200,
                    help="force-matching Adam iterations (0: skip the force-matching leg)")
    ap.add_argument("--fm-only", action="store_true",
                    help="run only the force-matching leg and print its JSON (child process)")
    ap.add_argument("--cpu-only", action="store_true",
                    help="run only the reference CPU driver leg and print its JSON (child process)")
    ap.add_argument("--ref-only", action="store_true",
                    help="run only the reference-size legs and print their JSON")
    ap.add_argument("--profile", action="store_true",
                    help="short mixed run only (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the step path runs on a single GPU here; "
              "run with --gpus 1", file=sys.stderr)
        return 2
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if world != 1:
        print("bench.py is single-process; WORLD_SIZE must be 1", file=sys.stderr)
        return 2

    if args.fm_only:
        # torch first: it carries its own HIP runtime, and the library then
        # binds to the same one in this process
        import torch  # noqa: F401
        print(json.dumps(bench_force_matching(args.fm_iters)), flush=True)
        return 0

    if args.cpu_only:
        print(json.dumps(cpu_reference(args.cells, args.cpu_steps)), flush=True)
        return 0

    if args.ref_only:
        capi.call("tdmd_set_device", 0)
        print(json.dumps(bench_reference_sizes(args.ref_steps)), flush=True)
        return 0

    tables = I.johnson_setfl()
    system = make_system(args.cells)
    N = system.natoms
    cpu_job = None
    if not args.no_cpu_baseline and not args.profile:
        cpu_job = CpuBaseline(system, tables, args.cells, args.cpu_steps)

    capi.call("tdmd_set_device", 0)
    name = capi.C.create_string_buffer(256)
    capi.call("tdmd_device_name", name, 256)
    log("device:", name.value.decode())
    log(f"system: {N} atoms, box {system.hi[0]:.2f} A")

    if args.profile:
        # a plain copy kernel of known size calibrates the HBM byte counters
        nwords = 1 << 28  # 1 GiB
        src, dst = capi.DeviceArray(nwords, np.uint32), capi.DeviceArray(nwords, np.uint32)
        capi.call("tdmd_copy_u32", dst.ptr, src.ptr, nwords, None)
        capi.call("tdmd_device_synchronize")
        src.free()
        dst.free()
        sched = FastPipelineScheduler.for_system(
            system, tables, FastPipelineConfig(**{**nvt_cfg("mixed").__dict__,
                                                  "mirror_fp64": args.mirror_fp64}))
        wall, dev = timed_run(sched, args.warmup, args.steps, pass_clock=True)
        ta, tb = sched.pass_times()
        log(f"profile run: {args.steps} NVT steps in {wall:.3f} s; passes of the timed window: "
            f"density {ta:.4f} ms, force {tb:.4f} ms")
        pos_e, vel_e, _, _, _ = sched.download()
        sched.close()
        state = system.copy()
        state.pos, state.vel = pos_e, vel_e
        # a list rebuild every 5 steps: the fused rebuild's own cost in a trace
        reb = FastPipelineScheduler.for_system(
            state, tables, FastPipelineConfig(precision="mixed", rebuild_policy="fixed",
                                              rebuild_every=5, mirror_fp64=args.mirror_fp64))
        wall, _ = timed_run(reb, 0, 20)
        log(f"rebuild every 5 steps: 20 steps in {wall:.4f} s")
        return 0

    # ---- mixed precision (the metric) ----
    cfg = nvt_cfg("mixed")
    cfg.mirror_fp64 = args.mirror_fp64
    sched = FastPipelineScheduler.for_system(system, tables, cfg)
    log("mixed: warmup, graph pre-capture window, timed window")
    # the passes' mean durations from their own device clock, over the
    # window of the same shape that follows the timed one
    wall, dev = timed_run(sched, args.warmup, args.steps, pass_clock=True)
    t_dens, t_force = sched.pass_times()
    th = sched.thermo()
    value = N * args.steps / wall / 1e6
    stats = sched.stats()
    log(f"mixed: {args.steps} steps, wall {wall:.4f} s (events {dev:.4f} s) -> {value:.1f} "
        f"Matom-steps/s; T {th['temp']:.1f} K, mean neighbours {th['mean_nbr']:.2f}, "
        f"rebuilds {stats.rebuilds}")
    obs = sched.observables()
    nbar, nnear = obs["mean_neighbors"], obs["mean_near_neighbors"]
    halo = sched.halo_info()
    pos_m, vel_m, f_m, _, _ = sched.download()
    sched.close()

    # ---- FP64 path, same workload, same windows ----
    s64 = FastPipelineScheduler.for_system(system, tables, nvt_cfg("fp64"))
    log("fp64: warmup, graph pre-capture window, timed window")
    w64, d64 = timed_run(s64, args.warmup, args.steps)
    v64 = N * args.steps / w64 / 1e6
    fp64 = dict(value=v64, steps=args.steps, ms_per_step=1000 * w64 / args.steps,
                rebuilds=s64.stats().rebuilds)
    s64.close()
    log(f"fp64: {v64:.1f} Matom-steps/s")

    # ---- the roofline of the timed window's force pass ----
    staged = halo["staged"] / halo["blocks"] if halo["blocks"] else 0.0
    halo_pa = halo["halo_atoms"] / N if halo["blocks"] else 0.0
    bpa = bytes_per_atom("mixed", args.mirror_fp64, nbar, staged, halo_pa, nnear)
    b_alg = alg_bytes_force("mixed", nbar)
    achieved = N * b_alg / (t_force * 1e-3) / 1e9
    traffic = measured_traffic(N, args.mirror_fp64, "force_nvt")
    flops = measured_flops(N, args.mirror_fp64, "force_nvt")
    log(f"timed-window passes: density {t_dens:.4f} ms, force {t_force:.4f} ms; force pass "
        f"{achieved:.0f} GB/s on {b_alg:.1f} B/atom")
    # steps between rebuilds over the whole mixed run (warmup, two pre-capture
    # windows, the timed and the pass-clock window; the first build is not a
    # rebuild)
    run_steps = args.warmup + 4 * args.steps
    n_rb = run_steps / (stats.rebuilds - 1) if stats.rebuilds > 1 else float(run_steps)

    # ---- NVE energy drift from the thermalised state, mixed and FP64 ----
    state = thermal_state(system, tables)
    drift = {}
    for prec in ("mixed", "fp64"):
        drift[prec] = nve_drift(state, tables, prec, args.nve_steps, 500)
        log(f"NVE drift {prec}: {drift[prec]['drift_ev_per_atom_per_ns']:.3e} eV/atom/ns "
            f"over {args.nve_steps} steps")
    t_force_nve = drift["mixed"]["force_pass_ms"]

    # ---- other workloads of the path (not the metric) ----
    extras = {}
    if args.c2_steps > 0:
        extras["morse_32k"] = bench_morse_c2(args.c2_steps)
        log(f"Morse 32k: mixed {extras['morse_32k']['mixed']['value']} / fp64 "
            f"{extras['morse_32k']['fp64']['value']} Matom-steps/s")
    if args.ref_steps > 0:
        extras["reference_sizes"] = bench_reference_sizes(args.ref_steps)
        for pot in ("morse", "eam"):
            log(f"reference sizes, {pot}: " + ", ".join(
                f"{k} {v['ts_per_s']} ts/s ({v['over_reference']}x)"
                for k, v in extras["reference_sizes"][pot].items()))
    if args.replica_steps > 0:
        extras["replicas_64"] = bench_replicas(args.replica_steps)
        log(f"64 replicas ({extras['replicas_64']['natoms']} atoms): mixed "
            f"{extras['replicas_64']['mixed']['value']} / fp64 "
            f"{extras['replicas_64']['fp64']['value']} Matom-steps/s")
    if args.fm_iters > 0:
        # own process: torch brings its own HIP runtime, this process already
        # runs on the system one
        import subprocess

        res = subprocess.run([sys.executable, os.path.abspath(__file__), "--fm-only",
                              "--fm-iters", str(args.fm_iters)],
                             capture_output=True, text=True, timeout=600)
        if res.returncode != 0:
            raise RuntimeError(f"force-matching leg failed:\n{res.stderr[-2000:]}")
        extras["force_matching"] = json.loads(res.stdout.strip().splitlines()[-1])
        fmx = extras["force_matching"]
        log(f"force matching: Morse eval {fmx['morse']['eval_ms']} ms, EAM eval "
            f"{fmx['eam']['eval_ms']} ms, EAM fit {fmx['eam']['fit_s']} s "
            f"(loss {fmx['eam']['loss_start']:.4e} -> {fmx['eam']['loss_end']:.4e})")

    # ---- force check against the FP64 oracle (the checker), then the CPU baseline ----
    from oracle import oracle as O

    ref_state = system.copy()
    ref_state.pos, ref_state.vel = pos_m, vel_m
    _, _, f_ref, _, _ = O.run(ref_state, eam=O.Eam(tables=tables), nsteps=0, thermo_every=1)
    rms = float(np.sqrt(np.mean(np.sum(f_ref * f_ref, axis=1))))
    force_err = float(np.max(np.linalg.norm(f_m - f_ref, axis=1)) / rms)
    log(f"force error vs FP64 oracle: {force_err:.3e} (max |dF| / RMS |F|, RMS {rms:.3f})")
    cpu = cpu_job.join() if cpu_job is not None else None
    if cpu is not None:
        log(f"cpu baseline ({cpu['kind']}): {cpu['value']:.4f} Matom-steps/s")

    step_bytes = alg_bytes_step("mixed", nbar, n_rb)
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Matom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * wall / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed: f32 tables/pair functions, f64 geometry and sums",
        "data": "synthetic FCC Cu lattice, seeded 300 K velocities, Johnson-form EAM tables",
        "config": {
            "workload": "EAM Cu 64x64x64 FCC cells (1,048,576 atoms), NVT NHC(3) 300 K, "
                        "dt 1 fs, skin 1.0 A displacement rebuild",
            "natoms": N,
            "potential": "Johnson-form Cu EAM, 5000-point FD-Hermite tables, rc 5.5 A",
            "mirror": "fp64" if args.mirror_fp64 else "fixed32",
            "mean_neighbours": round(nbar, 2),
            "rebuilds": stats.rebuilds,
            "kernel_launches_per_step": round(stats.kernel_launches / max(stats.ticks, 1), 2),
            "untimed_precapture_steps": 2 * args.steps,
            "pass_clock_window_steps": args.steps,
        },
        "events_s": round(dev, 5),
        "fp64": {k: round(v, 4) if isinstance(v, float) else v for k, v in fp64.items()},
        "mixed_over_fp64": round(value / fp64["value"], 3),
        "force_rel_err_vs_fp64_oracle": force_err,
        "nve_energy_drift": {
            "drift_ev_per_atom_per_ns": drift["mixed"]["drift_ev_per_atom_per_ns"],
            "start": f"the state after {THERMAL_STEPS} NVT steps from the lattice",
            "mixed": drift["mixed"], "fp64": drift["fp64"]},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_eam_force_lds<float, MirrorFixed, 4, ...> (force + KE partials of the "
                      "half-kicked velocities: the force pass of the timed NVT step)",
            "window": "the timed window's steps, repeated with the pass clock on",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else traffic["bytes"],
            "traffic_per_atom": None if traffic is None else traffic["bytes_per_atom"],
            "bytes_per_atom": round(b_alg, 1),
            "bytes_per_atom_rule": "52 + 4 nbar: x, type, F'(rho), row offset and count, "
                                   "int32 full-list row, F (the reference's layout)",
            "bytes_per_atom_own_layout": round(bpa["force_nvt"], 1),
            "avg_ms": round(t_force, 5),
            "avg_ms_source": "the pass's own device-clock spans (first workgroup start to "
                             "last workgroup finish), mean over the launches of the window of "
                             "the timed window's shape that follows it",
            "density_pass_ms": round(t_dens, 5),
            "nve_force_pass_ms": round(t_force_nve, 5),
            "fp32_tflops": None if flops is None else round(flops["fp32_ops"] / (t_force * 1e9), 2),
            "fp64_tflops": None if flops is None else round(flops["fp64_ops"] / (t_force * 1e9), 2),
            "fp32_peak_tflops": 157.3,
            "fp64_peak_tflops": 78.6,
            "step_bytes_per_atom": round(step_bytes, 1),
            "step_achieved": round(N * step_bytes * args.steps / wall / 1e9, 1),
            "halo_staged_blocks": f"{halo['staged']}/{halo['blocks']}",
            "halo_atoms_per_atom": round(halo_pa, 3),
            "near_slots_per_atom": round(nnear, 2),
        },
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
