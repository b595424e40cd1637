"""The reference's NVE drift case on the GPU scheduler
(verifylab/cases/nve-drift/check.py, tolerance.toml): its own input -- 4000
Cu atoms, FCC, 100 K Maxwell-Boltzmann velocities (tests/golden/
cu_fcc_4000_T100K.data, a copy of the case's data file) -- with Mishin Cu
EAM, dt 1 fs, skin 1 A, 20,000 NVE steps, total energy every 500 steps; the
first 25 % of the samples are discarded as the lattice's transient, and the
relative drift |slope(TE vs t)| / |mean TE| must stay within the case's
thresholds: 5e-7 per ps in FP64, 5e-5 per ps in mixed precision."""
import numpy as np
import pytest

from conftest import GOLDEN
from gpu_helpers import say
from tdmd_claude_amd import inputs as I
from tdmd_claude_amd.scheduler import FastPipelineConfig, FastPipelineScheduler

pytestmark = pytest.mark.gpu

THRESHOLD = {"fp64": 5e-7, "mixed": 5e-5}  # tolerance.toml [drift]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_reference_nve_drift_case(prec):
    s = I.read_lammps_data(GOLDEN / "cu_fcc_4000_T100K.data")
    eam = I.read_setfl(GOLDEN / "Cu_mishin1.eam.alloy")
    say(f"nve drift case, {prec}, {s.natoms} atoms")
    sched = FastPipelineScheduler.for_system(
        s, eam, FastPipelineConfig(precision=prec, dt=0.001, r_skin=1.0,
                                   rebuild_policy="displacement"))
    t, te = [], []
    for k in range(0, 20001, 500):
        sched.run_until(k)
        th = sched.thermo()
        t.append(k * 0.001)
        te.append(th["pe"] + th["ke"])
    sched.close()
    t, te = np.array(t), np.array(te)
    keep = np.arange(len(t)) >= int(0.25 * len(t))
    slope = np.polyfit(t[keep], te[keep], 1)[0]
    rel = abs(slope) / abs(te[keep].mean())
    say(f"  relative drift {rel:.3e} /ps (threshold {THRESHOLD[prec]:.0e})")
    assert rel <= THRESHOLD[prec]
