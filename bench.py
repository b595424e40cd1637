#!/usr/bin/env python3
"""Throughput of the FastPipelineScheduler step on one MI355X.

Workload (the 1M-atom EAM configuration): FCC Cu 64x64x64 cells
(1,048,576 atoms, a = 3.615 A, box 231.36 A), Johnson-form Cu EAM tabulated
on 5000-point FD-Hermite splines (cutoff 5.5 A), skin 1.0 A with the
skin/2 displacement rebuild, 300 K Maxwell-Boltzmann velocities (seed 7),
NVT Nose-Hoover chain of 3 (tau 0.1 ps, 300 K), dt 1 fs. Synthetic inputs,
state resident in HBM before the timed region.

One JSON line on stdout: the mixed-precision throughput (value), the FP64
path's throughput on the same workload, the force error of the mixed path
against the FP64 CPU oracle, the NVE energy drift, a roofline object for the
dominant kernel (the EAM force pass) and a CPU baseline (the oracle's FP64
driver loop on one core, bounded sample). Under "extras": the 32,000-atom
Morse NVE workload, the reference's own tiny / small / medium systems beside
its published timesteps per second, 64 batched alloy replicas, and force
matching with the adjoint gradient.
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


THERMAL_STEPS = 1100  # NVT steps before the pass timing: the default warmup + timed steps


def make_system(cells: int) -> I.System:
    s = I.fcc_lattice(cells, cells, cells, 3.615)
    I.init_velocities(s, 300.0, 7)
    return s


def nvt_cfg(precision: str) -> FastPipelineConfig:
    return FastPipelineConfig(dt=0.001, r_skin=1.0, precision=precision,
                              rebuild_policy="displacement", t_target=300.0, t_period=0.1,
                              nhc_length=3, use_graph=True)


def timed_run(sched: FastPipelineScheduler, warmup: int, steps: int) -> tuple[float, float]:
    """Warm up, then time exactly `steps` steps: host clock and HIP events on
    the scheduler's stream, device synchronised on both sides."""
    sched.run_until(warmup)
    stream = sched.stream()
    e0, e1 = capi.C.c_void_p(), capi.C.c_void_p()
    capi.call("tdmd_event_create", capi.C.byref(e0))
    capi.call("tdmd_event_create", capi.C.byref(e1))
    capi.call("tdmd_device_synchronize")
    t0 = time.perf_counter()
    capi.call("tdmd_event_record", e0, stream)
    sched.run_until(warmup + steps)  # ends with a stream synchronisation
    capi.call("tdmd_event_record", e1, stream)
    capi.call("tdmd_device_synchronize")
    wall = time.perf_counter() - t0
    ms = capi.C.c_float()
    capi.call("tdmd_event_elapsed_ms", e0, e1, capi.C.byref(ms))
    capi.call("tdmd_event_destroy", e0)
    capi.call("tdmd_event_destroy", e1)
    return wall, ms.value / 1000.0


def bytes_per_atom(precision: str, mirror_fp64: bool, nbar: float, staged: float = 0.0,
                   halo: float = 0.0, nnear: float | None = None) -> dict:
    """Algorithmic HBM bytes per atom of the two steady-state passes: each
    per-atom array touched once per pass, the neighbour rows read once per
    pass, neighbour records served from cache (L2, or LDS for staged
    blocks). Rows are 4 bytes a slot; for the fraction `staged` of blocks
    whose halo sits in LDS they are 2 bytes a slot plus the 4-byte halo list
    (`halo` atoms per atom), and the force pass reads only the `nnear` slots
    inside the cutoff at the build (its list of skin-shell slots inside the
    cutoff, 0-3 entries, is not counted). Build positions for the skin test
    are 32-bit coordinates with the fixed-point mirror, FP64 otherwise."""
    fixed = precision == "mixed" and not mirror_fp64
    rec = 16 if fixed else 32  # M / M2 record
    xb = 12 if fixed else 24  # positions at the last build
    nn = nbar if nnear is None else nnear
    rows_d = staged * (2 * nbar + 4 * halo) + (1 - staged) * 4 * nbar
    rows_f = staged * (2 * nn + 4 * halo) + (1 - staged) * 4 * nbar
    density = rec + 4 + rows_d + rec  # read M_i, count, rows; write M2_i
    # force pass: read M2_i, count, type, x, v, x_build, rows; write x, v, M_i
    force = rec + 4 + 4 + 24 + 24 + xb + rows_f + 24 + 24 + rec
    return dict(density=density, force=force, step=density + force)


def bench_morse_c2(steps: int) -> dict:
    """32,000-atom Morse Cu (20^3 cells, 300 K, +-0.02 A jitter), NVE, skin/2
    displacement rebuild: mixed and FP64 throughput and the mixed energy drift."""
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
        sched.run_until(0)
        th0 = sched.thermo()
        wall, _ = timed_run(sched, 100, steps)
        th1 = sched.thermo()
        e0 = (th0["pe"] + th0["ke"]) / s.natoms
        e1 = (th1["pe"] + th1["ke"]) / s.natoms
        out[prec] = {"value": round(s.natoms * steps / wall / 1e6, 2),
                     "ms_per_step": round(1000 * wall / steps, 4),
                     "energy_change_ev_per_atom": e1 - e0,
                     "drift_ev_per_atom_per_ns": (e1 - e0) / ((steps + 100) * 1e-6),
                     "rebuilds": sched.stats().rebuilds}
        sched.close()
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
    from tdmd_claude_amd.scheduler import BatchScheduler

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
        b = BatchScheduler.for_systems(systems, tables, cfg)
        b.run_until(0)
        tuned = b.autotune(5)
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


def measured_traffic(natoms: int, mirror_fp64: bool):
    """HBM bytes per launch of the force pass from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (profiles/eam_passes_1M_mixed_traffic.json),
    when they were taken on this workload; else None."""
    path = os.path.join(ROOT, "profiles", "eam_passes_1M_mixed_traffic.json")
    if mirror_fp64 or not os.path.exists(path):
        return None
    with open(path) as fh:
        doc = json.load(fh)
    if doc.get("natoms") != natoms:
        return None
    k = doc["kernels"]["force_nve"]
    return {"bytes": round(k["hbm_bytes"]), "bytes_per_atom": round(k["bytes_per_atom"], 1),
            "source": "profiles/eam_passes_1M_mixed_traffic.json"}


def nve_drift(system: I.System, tables, precision: str, steps: int, every: int) -> dict:
    """Mixed NVE run from the given state: slope of total energy per atom."""
    cfg = FastPipelineConfig(dt=0.001, r_skin=1.0, precision=precision,
                             rebuild_policy="displacement")
    sched = FastPipelineScheduler.for_system(system, tables, cfg)
    t, e = [], []
    for k in range(0, steps + 1, every):
        sched.run_until(k)
        th = sched.thermo()
        t.append(k * 0.001)
        e.append((th["pe"] + th["ke"]) / system.natoms)
    t = np.array(t)
    e = np.array(e)
    slope = np.polyfit(t, e, 1)[0]  # eV/atom/ps
    sched.close()
    return dict(drift_ev_per_atom_per_ns=float(slope * 1000.0),
                max_dev_ev_per_atom=float(np.abs(e - e[0]).max()), steps=steps,
                precision=precision)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--fp64-steps", type=int, default=500)
    ap.add_argument("--nve-steps", type=int, default=2000)
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

    if args.ref_only:
        capi.call("tdmd_set_device", 0)
        print(json.dumps(bench_reference_sizes(args.ref_steps)), flush=True)
        return 0

    capi.call("tdmd_set_device", 0)
    name = capi.C.create_string_buffer(256)
    capi.call("tdmd_device_name", name, 256)
    log("device:", name.value.decode())
    tables = I.johnson_setfl()
    system = make_system(args.cells)
    N = system.natoms
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
        wall, dev = timed_run(sched, args.warmup, args.steps)
        log(f"profile run: {args.steps} steps in {wall:.3f} s")
        # on to the state the benchmark's pass timing sees (1100 NVT steps)
        sched.run_until(max(THERMAL_STEPS, args.warmup + args.steps))
        pos_t, vel_t, _, _, _ = sched.download()
        sched.close()
        state = system.copy()
        state.pos, state.vel = pos_t, vel_t
        nve = FastPipelineScheduler.for_system(
            state, tables, FastPipelineConfig(precision="mixed", mirror_fp64=args.mirror_fp64))
        ta, tb = nve.time_passes(30)
        log(f"passes (thermalised state): density {ta:.4f} ms, force {tb:.4f} ms")
        nve.close()
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
    log("mixed: warmup + timed run")
    wall, dev = timed_run(sched, args.warmup, args.steps)
    th = sched.thermo()
    value = N * args.steps / wall / 1e6
    log(f"mixed: {args.steps} steps, wall {wall:.4f} s (events {dev:.4f} s) -> {value:.1f} "
        f"Matom-steps/s; T {th['temp']:.1f} K, mean neighbours {th['mean_nbr']:.2f}")
    stats = sched.stats()
    pos_m, vel_m, f_m, _, _ = sched.download()
    sched.close()

    # ---- per-pass timing of the steady-state NVE step (roofline) ----
    state = system.copy()
    state.pos, state.vel = pos_m, vel_m
    nve = FastPipelineScheduler.for_system(
        state, tables, FastPipelineConfig(precision="mixed", mirror_fp64=args.mirror_fp64))
    nve.run_until(0)
    obs = nve.observables()
    nbar, nnear = obs["mean_neighbors"], obs["mean_near_neighbors"]
    t_dens, t_force = nve.time_passes(30)
    halo = nve.halo_info()
    nve.close()
    staged = halo["staged"] / halo["blocks"] if halo["blocks"] else 0.0
    halo_pa = halo["halo_atoms"] / N if halo["blocks"] else 0.0
    bpa = bytes_per_atom("mixed", args.mirror_fp64, nbar, staged, halo_pa, nnear)
    traffic = measured_traffic(N, args.mirror_fp64)
    achieved = N * bpa["force"] / (t_force * 1e-3) / 1e9
    log(f"passes: density {t_dens:.4f} ms, force {t_force:.4f} ms; force pass "
        f"{achieved:.0f} GB/s algorithmic ({bpa['force']:.0f} B/atom)")

    # ---- FP64 path, same workload ----
    fp64 = None
    if args.fp64_steps > 0:
        s64 = FastPipelineScheduler.for_system(system, tables, nvt_cfg("fp64"))
        log("fp64: warmup + timed run")
        w64, d64 = timed_run(s64, args.warmup, args.fp64_steps)
        v64 = N * args.fp64_steps / w64 / 1e6
        s64.close()
        fp64 = dict(value=v64, steps=args.fp64_steps, ms_per_step=1000 * w64 / args.fp64_steps)
        log(f"fp64: {v64:.1f} Matom-steps/s")

    # ---- NVE energy drift (mixed) ----
    drift = nve_drift(state, tables, "mixed", args.nve_steps, max(args.nve_steps // 10, 1))
    log(f"NVE drift: {drift['drift_ev_per_atom_per_ns']:.3e} eV/atom/ns")

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

    # ---- CPU baseline leg: FP64 oracle driver (also the force checker) ----
    cpu = None
    force_err = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        eam = O.Eam(tables=tables)
        ref_state = system.copy()
        ref_state.pos = pos_m
        ref_state.vel = vel_m
        t0 = time.perf_counter()
        _, _, f_ref, _, _ = O.run(ref_state, eam=eam, nsteps=0, thermo_every=1)
        t1 = time.perf_counter()
        rms = float(np.sqrt(np.mean(np.sum(f_ref * f_ref, axis=1))))
        force_err = float(np.max(np.linalg.norm(f_m - f_ref, axis=1)) / rms)
        log(f"force error vs FP64 oracle: {force_err:.3e} (max |dF| / RMS |F|, RMS {rms:.3f})")
        nsteps_cpu = 2
        t2 = time.perf_counter()
        O.run(ref_state, eam=eam, nsteps=nsteps_cpu, thermo_every=nsteps_cpu)
        t3 = time.perf_counter()
        per_step = ((t3 - t2) - (t1 - t0)) / nsteps_cpu
        cpu = dict(value=N / per_step / 1e6, unit="Matom-steps/s", cores=1, kind="port",
                   sample=(f"{N}-atom EAM Cu (the GPU run's state), {nsteps_cpu} steps of the "
                           "FP64 driver loop (tdmd_main.cpp:237-254) after its list build; "
                           f"{per_step:.2f} s/step on {os.cpu_count()} host CPUs, 1 used"))
        log(f"cpu baseline: {cpu['value']:.4f} Matom-steps/s")

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
        },
        "events_s": round(dev, 5),
        "fp64": None if fp64 is None else {k: round(v, 4) if isinstance(v, float) else v
                                           for k, v in fp64.items()},
        "mixed_over_fp64": None if fp64 is None else round(value / fp64["value"], 3),
        "force_rel_err_vs_fp64_oracle": force_err,
        "nve_energy_drift": drift,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_eam_force_lds<float, MirrorFixed> (force + fused NVE kick-drift)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else traffic["bytes"],
            "traffic_per_atom": None if traffic is None else traffic["bytes_per_atom"],
            "bytes_per_atom": round(bpa["force"], 1),
            "avg_ms": round(t_force, 5),
            "density_pass_ms": round(t_dens, 5),
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
