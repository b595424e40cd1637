"""The C++ drop-in interface (include/tdmd/) driven by a compiled C++
program, as the reference's own code would call it: neighbour list vs brute
force, Morse vs host FP64, FastPipelineScheduler vs the reference's step
sequence assembled from the free functions, DeviceEam vs finite
differences, NVE energy, observables with the thermo line and a force dump
round trip, NHC reductions, error translation. One run per
precision macro; each check prints a PASS/FAIL line."""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent
BIN = ROOT / "cpp" / "_bin"
SETFL = ROOT / "golden" / "Cu_mishin1.eam.alloy"


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_cpp_dropin_program(prec, tmp_path):
    exe = BIN / f"dropin_test_{prec}"
    assert exe.exists(), f"{exe} missing: run __graft_entry__.build()"
    print(f"running {exe.name}", flush=True)
    res = subprocess.run([str(exe), str(SETFL), str(tmp_path / "dropin.dump")],
                         capture_output=True, text=True, timeout=110)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 failure(s)" in res.stdout
    assert res.stdout.count("PASS") == 11


def test_cpp_dropin_program_is_built():
    for prec in ("fp64", "mixed"):
        assert (BIN / f"dropin_test_{prec}").exists()
