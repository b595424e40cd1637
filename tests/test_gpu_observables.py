"""Device observables (tdmd_sched_observables). The virial is pinned by the
oracle's FP64 energy: under a uniform scaling of positions and box by
lambda, W = sum over pairs of r_ij . f_ij = -dU/dlambda at lambda = 1, taken
here by a fourth-order central difference. The chain energy is pinned by conservation
of the NVT extended energy, and the dump writer end to end by the LAMMPS
dump the reference's run-0 case holds."""
import numpy as np
import pytest

from conftest import GOLDEN, MORSE_CU, load_golden
from gpu_helpers import say
from oracle import oracle as O
from tdmd_claude_amd import inputs as I
from tdmd_claude_amd import output as OUT
from tdmd_claude_amd.scheduler import FastPipelineConfig, FastPipelineScheduler

pytestmark = pytest.mark.gpu

KB = 8.617333262e-5
EV_PER_A3_IN_BAR = 1.602176634e6


def scaled(s, lam):
    t = s.copy()
    t.pos = s.lo + (s.pos - s.lo) * lam
    t.hi = s.lo + (s.hi - s.lo) * lam
    return t


def oracle_pe(s, pot, tables):
    if pot == "morse":
        pe, _ = O.morse_forces(O.HalfList(s.pos, s.lo, s.hi, MORSE_CU[3], 1.0), *MORSE_CU)
        return pe
    e = O.Eam(tables=tables)
    return e.forces(O.HalfList(s.pos, s.lo, s.hi, e.cutoff, 1.0), s.types)["pe"]


@pytest.mark.parametrize("pot", ["morse", "johnson"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_virial_is_minus_dU_dlambda(pot, prec):
    s = I.fcc_lattice(5, 5, 5, 3.55)  # compressed: W of a few hundred eV
    I.jitter_uniform(s, 0.05, 11)
    I.init_velocities(s, 300.0, 5)
    tables = None if pot == "morse" else I.johnson_setfl()
    # fourth-order central difference (the second-order one is ~1e-6 off on
    # this strongly curved U(lambda) at h = 1e-4)
    h = 1e-4
    u = {k: oracle_pe(scaled(s, 1 + k * h), pot, tables) for k in (-2, -1, 1, 2)}
    w_fd = -(8 * (u[1] - u[-1]) - (u[2] - u[-2])) / (12 * h)
    sched = FastPipelineScheduler.for_system(s, MORSE_CU if pot == "morse" else tables,
                                             FastPipelineConfig(precision=prec))
    sched.run_until(0)
    o = sched.observables()
    sched.close()
    say(f"virial {pot} {prec}: device {o['virial']:.9f} fd {w_fd:.9f}")
    assert abs(o["virial"] - w_fd) < (1e-7 if prec == "fp64" else 2e-5) * abs(w_fd)
    V = float(np.prod(s.hi - s.lo))
    assert abs(o["volume"] - V) < 1e-9 * V
    # the reported pressure carries the same virial and the 3N-dof KE
    assert abs(o["pressure"] - (2 * o["ke"] + w_fd) / (3 * V) * EV_PER_A3_IN_BAR) < \
        (1e-6 if prec == "fp64" else 3e-5) * abs(o["pressure"])
    assert abs(o["temp"] - 2 * o["ke"] / (3 * s.natoms * KB)) < 1e-9 * o["temp"]
    assert o["natoms"] == s.natoms


def test_nvt_extended_energy_is_conserved():
    # One thermostat: the reference's half-step is a consistent splitting, so
    # PE + KE + chain energy is conserved to O(dt) while energy flows through
    # the chain. (Longer chains drift: the reference damps the thermostat
    # velocities with half the friction the chain equations need,
    # device_nose_hoover.cu:251-261; the oracle tests record that.)
    s = I.fcc_lattice(5, 5, 5, 3.615)
    I.init_velocities(s, 600.0, 9)
    cfg = FastPipelineConfig(precision="fp64", t_target=300.0, t_period=0.1, nhc_length=1)
    sched = FastPipelineScheduler.for_system(s, I.johnson_setfl(), cfg)
    sched.run_until(0)
    o0 = sched.observables()
    h, chain = [], []
    for t in range(100, 1601, 100):
        sched.run_until(t)
        o = sched.observables()
        h.append(o["conserved"])
        chain.append(o["nhc_energy"])
        assert o["step"] == t and abs(o["conserved"] - o["te"] - o["nhc_energy"]) < 1e-9
    sched.close()
    dev = np.abs(np.array(h) - o0["conserved"]).max()
    exchanged = np.abs(chain).max()
    say(f"NVT M=1: chain energy up to {exchanged:.4f} eV, extended energy deviation {dev:.2e} eV")
    assert exchanged > 0.5
    assert dev < 0.02 * exchanged


def test_force_dump_of_the_run0_case(tmp_path):
    # verifylab run0-force-match: our dump of the step-0 state parses to the
    # LAMMPS dump's atoms, and the forces agree (perfect lattice: ~1e-15)
    s = I.read_lammps_data(GOLDEN / "cu_fcc_256.data")
    sched = FastPipelineScheduler.for_system(s, I.read_setfl(GOLDEN / "Cu_mishin1.eam.alloy"),
                                             FastPipelineConfig(precision="fp64"))
    sched.run_until(0)
    pos, vel, f, types, ids = sched.download()
    o = sched.observables()
    sched.close()
    p = tmp_path / "run0.dump"
    OUT.write_force_dump(p, o["step"], s.lo, s.hi, ids, types, pos, f)
    ours, lmp = OUT.read_dump(p), OUT.read_dump(GOLDEN / "forces_eam.dump")
    assert np.array_equal(ours.ids, lmp.ids) and np.array_equal(ours.types, lmp.types)
    assert np.array_equal(ours.pos, lmp.pos)
    assert np.abs(ours.forces - lmp.forces).max() < 1e-12
    # PE of this case from the reference's compiled CPU EAM (golden fixture;
    # LAMMPS reports -906.29591 eV)
    pe_ref = float(load_golden("ref_lattice")["eam_pe"])
    say(f"run0 PE {o['pe']:.10f} eV, reference {pe_ref:.10f}")
    assert abs(o["pe"] - pe_ref) < 1e-10 * abs(pe_ref)
