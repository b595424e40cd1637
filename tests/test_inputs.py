"""Input generators and file formats (CPU)."""
import numpy as np
import pytest

from conftest import GOLDEN
from tdmd_claude_amd import inputs as I


def test_fcc_matches_reference_fixture():
    # tools/gen_fcc_data.py produced tests/data/cu_fcc_256.data; same basis,
    # order and box.
    ref = I.read_lammps_data(GOLDEN / "cu_fcc_256.data")
    s = I.fcc_lattice(4, 4, 4, 3.615)
    assert np.allclose(s.pos, ref.pos, atol=1e-10, rtol=0)
    assert np.allclose(s.hi, ref.hi, atol=1e-10)
    assert np.array_equal(s.ids, ref.ids) and np.array_equal(s.types, ref.types)
    assert s.masses[1] == ref.masses[1]


def test_velocities_zero_momentum_exact_temperature():
    s = I.fcc_lattice(5, 5, 5, 3.615)
    I.init_velocities(s, 300.0, 42)
    m = s.masses[s.types]
    assert np.abs((m[:, None] * s.vel).sum(axis=0)).max() < 1e-10
    ke = I.kinetic_energy(s.vel, s.types, s.masses)
    t = 2 * ke / ((3 * s.natoms - 3) * I.KB)
    assert t == pytest.approx(300.0, rel=1e-12)
    s2 = I.fcc_lattice(5, 5, 5, 3.615)
    I.init_velocities(s2, 300.0, 42)
    assert np.array_equal(s.vel, s2.vel)  # seeded


def test_data_file_roundtrip_with_velocities(tmp_path):
    s = I.fcc_lattice(3, 3, 3, 3.615)
    I.jitter_uniform(s, 0.02, 4242)
    I.init_velocities(s, 300.0, 42)
    p = tmp_path / "x.data"
    I.write_lammps_data(p, s)
    b = I.read_lammps_data(p)
    assert np.array_equal(b.pos, s.pos) and np.array_equal(b.vel, s.vel)
    assert np.array_equal(b.types, s.types) and np.array_equal(b.masses, s.masses)


def test_reads_reference_multitype_fixture():
    s = I.read_lammps_data(GOLDEN / "4atoms.data")
    assert s.natoms == 4 and list(s.types) == [1, 1, 2, 2]
    assert s.masses[1] == 26.9815 and s.masses[2] == 63.546
    assert np.allclose(s.vel[3], [1.0, 1.1, 1.2])


def test_jitter_stays_in_box():
    s = I.fcc_lattice(4, 4, 4, 3.615)
    I.jitter_uniform(s, 0.02, 4242)
    assert (s.pos >= s.lo).all() and (s.pos < s.hi).all()


def test_johnson_tables_are_smooth_at_cutoff():
    t = I.johnson_setfl()
    assert abs(t.rho[0, -1]) < 1e-15 and abs(t.rphi[0, -1]) < 1e-15


def test_johnson_model_is_cu_like():
    from oracle import oracle as O

    t = I.johnson_setfl()
    e = O.Eam(tables=t)
    s = I.fcc_lattice(4, 4, 4, 3.615)
    pe = e.forces(O.HalfList(s.pos, s.lo, s.hi, t.cutoff, 0.3), s.types)["pe"] / s.natoms
    assert pe == pytest.approx(-3.54, abs=0.01)
    s.pos[0, 0] += 0.01  # Einstein spring ~7 eV/A^2
    f = e.forces(O.HalfList(s.pos, s.lo, s.hi, t.cutoff, 0.3), s.types)["forces"]
    assert 6.0 < -f[0, 0] / 0.01 < 8.0
