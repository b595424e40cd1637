"""Thermo line and LAMMPS force dumps (tdmd_main.cpp:65-122,
src/io/dump_reader.cpp:15-54): the Python writers, the C++ headers of the
drop-in (tests/cpp/io_test.cpp), and the files the reference's own cases
hold. Host-only: no device."""
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, MORSE_CU
from oracle import oracle as O
from tdmd_claude_amd import inputs as I
from tdmd_claude_amd import output as OUT
from tdmd_claude_amd.build import CPP_BIN

# the consumer's pattern: verifylab/cases/nve-drift/check.py:64-66 parses
# every thermo line with it
THERMO_RE = re.compile(r"Step\s+(\d+)\s+PE\s+([-\d.eE+]+)\s+KE\s+([-\d.eE+]+)\s+"
                       r"TE\s+([-\d.eE+]+)\s+T\s+([-\d.eE+]+)")


def io_bin(prec):
    exe = CPP_BIN / f"io_test_{prec}"
    if not exe.exists():
        pytest.skip("C++ io test not built (python -m tdmd_claude_amd.build)")
    return str(exe)


def test_thermo_line_matches_the_oracle_driver():
    # the oracle restates the driver loop's thermo (tdmd_main.cpp:211-254);
    # its rows parse back out of the formatted lines
    s = I.fcc_lattice(3, 3, 3, 3.615)
    I.init_velocities(s, 300.0, 3)
    _, _, _, th, _ = O.run(s, morse=MORSE_CU, nsteps=20, thermo_every=10)
    for row in th:
        line = OUT.thermo_line(int(row[0]), row[1], row[2], s.natoms)
        m = THERMO_RE.fullmatch(line)
        assert m, line
        step, pe, ke, te, t = (float(g) for g in m.groups())
        assert step == row[0]
        assert abs(pe - row[1]) < 1e-13 and abs(ke - row[2]) < 1e-13
        assert abs(te - row[3]) < 1e-13 and abs(t - row[4]) < 1e-4


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_cpp_thermo_line_equals_python(prec):
    exe = io_bin(prec)
    for step, pe, ke, n in [(0, -906.2973631, 0.0, 256), (100, -3.5e6, 12345.678901234567, 10**6),
                            (12345678, 1.0e-3, 2.5, 1)]:
        out = subprocess.run([exe, "thermo", str(step), repr(pe), repr(ke), str(n)],
                             capture_output=True, text=True, check=True).stdout.strip()
        assert out == OUT.thermo_line(step, pe, ke, n)


def test_dump_roundtrip_is_exact(tmp_path):
    rng = np.random.default_rng(1)
    n = 50
    ids = rng.permutation(n) + 1
    types = rng.integers(1, 4, n)
    pos = rng.uniform(-10, 10, (n, 3))
    f = rng.normal(size=(n, 3)) * 1e-3
    p = tmp_path / "d.dump"
    OUT.write_force_dump(p, 7, (0, 0, 0), (20.0, 21.5, 22.25), ids, types, pos, f)
    d = OUT.read_dump(p)
    order = np.argsort(ids)
    assert d.step == 7
    assert np.array_equal(d.ids, ids[order]) and np.array_equal(d.types, types[order])
    assert np.array_equal(d.pos, pos[order]) and np.array_equal(d.forces, f[order])


@pytest.mark.parametrize("which", ["morse", "eam"])
def test_read_the_reference_lammps_dumps(which):
    # verifylab run0-force-match: LAMMPS dumps of the 256-atom lattice
    d = OUT.read_dump(GOLDEN / f"forces_{which}.dump")
    s = I.read_lammps_data(GOLDEN / "cu_fcc_256.data")
    assert np.array_equal(d.ids, np.arange(1, 257))
    assert np.array_equal(d.pos, s.pos) and np.array_equal(d.types, s.types)
    assert np.abs(d.forces).max() < 1e-12  # perfect lattice: rounding noise


def test_cpp_writer_equals_python_writer(tmp_path):
    # FP64 build: the same atoms give byte-identical files
    exe = io_bin("fp64")
    src = GOLDEN / "forces_eam.dump"
    d = OUT.read_dump(src)
    lo, hi = (0.0, 0.0, 0.0), (14.46, 14.46, 14.46)
    cpp = tmp_path / "cpp.dump"
    py = tmp_path / "py.dump"
    subprocess.run([exe, "copy", str(src), str(cpp), "3", "0", "14.46", "0", "14.46", "0",
                    "14.46"], check=True)
    OUT.write_force_dump(py, 3, lo, hi, d.ids, d.types, d.pos, d.forces)
    assert cpp.read_bytes() == py.read_bytes()


def test_cpp_mixed_writer_rounds_forces_to_float(tmp_path):
    exe = io_bin("mixed")
    src = GOLDEN / "forces_eam.dump"
    out = tmp_path / "m.dump"
    subprocess.run([exe, "copy", str(src), str(out), "0", "0", "1", "0", "1", "0", "1"],
                   check=True)
    a, b = OUT.read_dump(src), OUT.read_dump(out)
    assert np.array_equal(a.pos, b.pos)
    assert np.array_equal(b.forces, a.forces.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("body,msg", [("ITEM: TIMESTEP\n0\n", "no atoms"),
                                      ("ITEM: NUMBER OF ATOMS\n3\nITEM: ATOMS id type x y z "
                                       "fx fy fz\n1 1 0 0 0 0 0 0\n", "premature end"),
                                      ("ITEM: NUMBER OF ATOMS\n1\nITEM: ATOMS id type x y z "
                                       "fx fy fz\n1 1 0 0 zero 0 0 0\n", "bad dump line")])
def test_malformed_dumps_are_errors(tmp_path, body, msg):
    p = tmp_path / "bad.dump"
    p.write_text(body)
    with pytest.raises(ValueError, match=msg):
        OUT.read_dump(p)
    r = subprocess.run([io_bin("fp64"), "read", str(p)], capture_output=True, text=True)
    assert r.returncode == 2 and msg in r.stdout


# ---------------------------------------------------------------------------
# LAMMPS data files and setfl tables (lammps_data_reader.cpp:33-205,
# eam_alloy.cpp:90-192): the C++ headers against the Python reader/writers
# and the reference's own data files.
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["cu_fcc_256.data", "cu_fcc_32.data", "4atoms.data"])
def test_cpp_data_reader_writer_equal_python(tmp_path, name):
    src = GOLDEN / name
    ref = I.read_lammps_data(src)
    for prec in ("fp64", "mixed"):
        out = tmp_path / f"{prec}.data"
        res = subprocess.run([io_bin(prec), "data_copy", str(src), str(out)], capture_output=True,
                             text=True)
        assert res.returncode == 0, res.stdout + res.stderr
        assert int(res.stdout) == ref.natoms
        back = I.read_lammps_data(out)
        assert np.array_equal(back.ids, ref.ids) and np.array_equal(back.types, ref.types)
        assert np.array_equal(back.lo, ref.lo) and np.array_equal(back.hi, ref.hi)
        if prec == "fp64":
            py = tmp_path / "py.data"
            I.write_lammps_data(py, ref)
            assert out.read_bytes() == py.read_bytes()
        else:  # numbers parsed as float, as the reference's mixed build does
            assert np.abs(back.pos - ref.pos).max() <= 1e-6 * max(1.0, np.abs(ref.pos).max())
            assert np.allclose(back.masses, ref.masses, rtol=1e-7)


def test_cpp_data_velocities_round_trip(tmp_path):
    s = I.fcc_lattice(2, 2, 2, 3.615)
    I.init_velocities(s, 500.0, 9)
    src = tmp_path / "v.data"
    I.write_lammps_data(src, s)
    out = tmp_path / "c.data"
    res = subprocess.run([io_bin("fp64"), "data_copy", str(src), str(out)], capture_output=True,
                         text=True)
    assert res.returncode == 0, res.stdout
    assert out.read_bytes() == src.read_bytes()
    back = I.read_lammps_data(out)
    assert np.array_equal(back.vel, s.vel)


@pytest.mark.parametrize("broken,msg", [
    ("drop_box", "incomplete box definition"),
    ("drop_count", "no atoms in data file"),
    ("short_atoms", "premature end in Atoms"),
    ("bad_id", "atom id outside 1..N"),
    ("bad_row", "bad Atoms line"),
])
def test_cpp_data_reader_errors(tmp_path, broken, msg):
    lines = (GOLDEN / "4atoms.data").read_text().splitlines()
    if broken == "drop_box":
        lines = [ln for ln in lines if "ylo yhi" not in ln]
    elif broken == "drop_count":
        lines = [ln for ln in lines if not ln.strip().endswith(" atoms")]
    elif broken == "short_atoms":
        k = max(i for i, ln in enumerate(lines) if ln.strip() and ln.split()[0].isdigit()
                and len(ln.split()) == 5)
        lines = lines[:k]
        # no later section: the file ends inside Atoms
        lines = [ln for ln in lines]
    elif broken == "bad_id":
        k = next(i for i, ln in enumerate(lines) if ln.strip().startswith("Atoms")) + 2
        f = lines[k].split()
        f[0] = "99"
        lines[k] = " ".join(f)
    elif broken == "bad_row":
        k = next(i for i, ln in enumerate(lines) if ln.strip().startswith("Atoms")) + 2
        lines[k] = lines[k].split()[0] + " 1 x y z"
    p = tmp_path / "broken.data"
    p.write_text("\n".join(lines) + "\n")
    res = subprocess.run([io_bin("fp64"), "data_copy", str(p), str(tmp_path / "o.data")],
                         capture_output=True, text=True)
    assert res.returncode == 2 and msg in res.stdout, res.stdout


def test_cpp_setfl_writer_equals_python(tmp_path):
    t = I.johnson_alloy_setfl(nr=120, nrho=90)
    tab = tmp_path / "tables.txt"
    with open(tab, "w") as fh:
        fh.write(f"{t.ntypes} {t.nrho} {t.drho!r} {t.nr} {t.dr!r} {t.cutoff!r}\n")
        fh.write(" ".join(t.elements) + "\n")
        fh.write(" ".join(repr(float(m)) for m in t.masses) + "\n")
        for e in range(t.ntypes):
            fh.write(" ".join(repr(float(x)) for x in t.F[e]) + "\n")
            fh.write(" ".join(repr(float(x)) for x in t.rho[e]) + "\n")
        for p in range(t.rphi.shape[0]):
            fh.write(" ".join(repr(float(x)) for x in t.rphi[p]) + "\n")
    out = tmp_path / "c.setfl"
    res = subprocess.run([io_bin("fp64"), "setfl_write", str(tab), str(out)], capture_output=True,
                         text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    py = tmp_path / "py.setfl"
    I.write_setfl(py, t)
    assert out.read_bytes() == py.read_bytes()
    back = I.read_setfl(out)
    assert back.elements == list(t.elements) and np.allclose(back.rphi, t.rphi, rtol=1e-15)
