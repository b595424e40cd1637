"""Times the oracle's FP64 driver loop beside the reference's own CPU driver
(oracle/_ref/tdmd_standalone, built by `make -C oracle ref`) on the same
inputs, one core each, to show how the restatement's speed compares with
the reference CPU path it stands in for where the reference driver is not
built. Test infrastructure; prints one JSON line.

  python oracle/fairness.py [--c2-steps 200]

Inputs: C1, the 256-atom Morse Cu lattice with U(-0.02, 0.02) jitter (seed
4242) and 300 K velocities (seed 42), 1000 steps; C2, the 32,000-atom system
built the same way, --c2-steps steps; C3, the 1,048,576-atom Johnson EAM
lattice with 300 K velocities (seed 7), 3 steps. Both drivers rebuild their
lists by the reference's own skin test. Each figure is the difference of an
n-step and a 0-step run (input parsing and the first build cancel).
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import subprocess
import sys
import tempfile
import time

HERE = pathlib.Path(__file__).resolve().parent
sys.path[0] = str(HERE.parent)  # the repository root, not oracle/ (oracle.py would shadow the package)

from oracle import oracle as O  # noqa: E402
from tdmd_claude_amd import inputs as I  # noqa: E402

REF = HERE / "_ref" / "tdmd_standalone"
MORSE = (0.3429, 1.3588, 2.866, 6.0)


def ref_seconds(data, n, pot_args):
    t0 = time.perf_counter()
    subprocess.run([str(REF), "--data", str(data), "--nsteps", str(n), "--thermo", str(max(n, 1)),
                    *pot_args], capture_output=True, check=True)
    return time.perf_counter() - t0


def port_seconds(s, n, **kw):
    t0 = time.perf_counter()
    O.run(s, nsteps=n, thermo_every=max(n, 1), **kw)
    return time.perf_counter() - t0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--c2-steps", type=int, default=200)
    args = ap.parse_args()
    if not REF.exists():
        print(f"{REF} missing: run `make -C oracle ref` first", file=sys.stderr)
        return 2
    os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[-1]})
    out = {}
    with tempfile.TemporaryDirectory() as td:
        tmp = pathlib.Path(td)
        cases = []
        for name, cells, steps in (("C1", 4, 1000), ("C2", 20, args.c2_steps)):
            s = I.fcc_lattice(cells, cells, cells, 3.615)
            I.jitter_uniform(s, 0.02, 4242)
            I.init_velocities(s, 300.0, 42)
            cases.append((name, s, steps, ["--morse", ",".join(map(str, MORSE))],
                          dict(morse=MORSE)))
        s = I.fcc_lattice(64, 64, 64, 3.615)
        I.init_velocities(s, 300.0, 7)
        t = I.johnson_setfl()
        setfl = tmp / "johnson.eam.alloy"
        I.write_setfl(setfl, t)
        cases.append(("C3", s, 3, ["--eam", str(setfl)], dict(eam=O.Eam(tables=t))))
        for name, s, steps, pot_args, kw in cases:
            data = tmp / f"{name}.data"
            I.write_lammps_data(data, s)
            r = (ref_seconds(data, steps, pot_args) - ref_seconds(data, 0, pot_args)) / steps
            p = (port_seconds(s, steps, **kw) - port_seconds(s, 0, **kw)) / steps
            out[name] = {"natoms": s.natoms, "steps": steps, "reference_s_per_step": r,
                         "oracle_s_per_step": p, "oracle_over_reference": p / r}
            print(f"{name}: reference {r * 1e3:.2f} ms/step, oracle {p * 1e3:.2f} ms/step "
                  f"({p / r:.2f}x)", file=sys.stderr, flush=True)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
