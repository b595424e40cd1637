"""Shared fixtures. GPU tests are marked ``gpu`` and call the HIP library
through its C-ABI; everything else runs on the CPU."""
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

MORSE_CU = (0.3429, 1.3588, 2.866, 6.0)  # configs: D, alpha, r0, rc


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


def load_golden(name):
    return dict(np.load(GOLDEN / f"{name}.npz"))


def system_from_golden(g):
    from tdmd_claude_amd import inputs as I

    n = len(g["pos0"])
    return I.System(lo=g["lo"], hi=g["hi"], pos=g["pos0"].copy(), vel=g["vel0"].copy(),
                    types=g["types"].astype(np.int32), ids=np.arange(1, n + 1, dtype=np.int32),
                    masses=g["masses"])


def force_rel_err(f, f_ref):
    """Max over atoms of |F - F_ref| divided by the RMS force magnitude of F_ref."""
    rms = np.sqrt(np.mean(np.sum(f_ref * f_ref, axis=1)))
    return float(np.max(np.linalg.norm(f - f_ref, axis=1)) / rms)
