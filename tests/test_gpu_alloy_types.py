"""Several-element EAM at the edge of the packed element index: with at most
four elements the force pass reads a neighbour's element from the two low
mantissa bits of its F'(rho) record, with more it gathers types[j]. Four
elements use the largest packed index (3); five take the gather. Both run
against the FP64 oracle: the FP64 scheduler's trajectory, and mixed
step-0 forces within the stated tolerance.

The tables are distinct per element: the Ni/Co/Cr Johnson-form functions,
the fourth and fifth element scaled copies of the first two (F, rho and
the pair functions of every pair that involves them), so a mix-up of any
two element indices changes the forces."""
import numpy as np
import pytest

from gpu_helpers import say
from oracle import oracle as O
from tdmd_claude_amd import inputs as I
from tdmd_claude_amd.scheduler import FastPipelineConfig, FastPipelineScheduler

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-5


def pair_index(a: int, b: int, n: int) -> int:
    """Row of pair (a, b), a <= b, in setfl's i <= j order."""
    return a * n - a * (a - 1) // 2 + (b - a)


def tables_of(nel: int) -> I.SetflTables:
    base = I.johnson_alloy_setfl(elements=("Ni", "Co", "Cr"), nr=1000, nrho=1000)
    src = [k % 3 for k in range(nel)]  # element k copies base element src[k]
    sF = [1.0 + 0.05 * (k // 3) for k in range(nel)]
    sR = [1.0 + 0.04 * (k // 3) for k in range(nel)]
    F = np.stack([base.F[src[k]] * sF[k] for k in range(nel)])
    rho = np.stack([base.rho[src[k]] * sR[k] for k in range(nel)])
    rphi = []
    for a in range(nel):
        for b in range(a, nel):
            lo, hi = sorted((src[a], src[b]))
            scale = 1.0 + 0.03 * (a // 3 + b // 3)
            rphi.append(base.rphi[pair_index(lo, hi, 3)] * scale)
    names = [f"E{k}" for k in range(nel)]
    masses = [base.masses[src[k]] for k in range(nel)]
    return I.SetflTables(elements=names, masses=masses, nrho=base.nrho, drho=base.drho,
                         nr=base.nr, dr=base.dr, cutoff=base.cutoff, F=F, rho=rho,
                         rphi=np.stack(rphi))


def system_of(nel: int, tables: I.SetflTables) -> I.System:
    s = I.fcc_lattice(5, 5, 5, 3.57)
    rng = np.random.Generator(np.random.PCG64(31 + nel))
    s.types = rng.integers(1, nel + 1, size=s.natoms).astype(np.int32)
    s.masses = np.array([1.0] + list(tables.masses))
    I.jitter_uniform(s, 0.02, 41 + nel)
    I.init_velocities(s, 600.0, 51 + nel)
    return s


@pytest.mark.parametrize("nel", [4, 5])
def test_element_index_packed_and_gathered(nel):
    tables = tables_of(nel)
    s = system_of(nel, tables)
    assert set(np.unique(s.types)) == set(range(1, nel + 1))
    say(f"alloy of {nel} elements, {s.natoms} atoms")
    eam = O.Eam(tables=tables)
    # FP64 trajectory against the oracle driver
    sched = FastPipelineScheduler.for_system(
        s, tables, FastPipelineConfig(precision="fp64", rebuild_policy="fixed",
                                      rebuild_every=10, pair_lanes=1))
    sched.run_until(40)
    pos, vel, f, _, _ = sched.download()
    sched.close()
    rpos, rvel, rf, _, _ = O.run(s, eam=eam, nsteps=40, rebuild_mode=1, rebuild_every=10)
    L = s.hi - s.lo
    d = pos - rpos
    d -= np.round(d / L) * L
    say(f"  fp64 40 steps: max |dx| {np.abs(d).max():.2e}")
    assert np.abs(d).max() < 1e-10
    # mixed step-0 forces against the oracle
    sched = FastPipelineScheduler.for_system(
        s, tables, FastPipelineConfig(precision="mixed", pair_lanes=1))
    sched.run_until(0)
    f0 = sched.download()[2]
    sched.close()
    _, _, rf0, _, _ = O.run(s, eam=eam, nsteps=0)
    rms = np.sqrt(np.mean(np.sum(rf0 * rf0, axis=1)))
    err = float(np.max(np.linalg.norm(f0 - rf0, axis=1)) / rms)
    say(f"  mixed step-0 force err {err:.2e}")
    assert err < MIXED_TOL
