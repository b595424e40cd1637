"""Pins of the FP64 CPU oracle (oracle/) -- CPU only.

The oracle is checked against (a) golden vectors written by the reference's
own CPU driver (tests/golden/make_golden.py), bit for bit; (b) the
reference's fixtures (LAMMPS run-0 dumps, the two-atom analytic JSON);
(c) closed forms, brute force and finite differences.
"""
import json

import numpy as np
import pytest

from conftest import GOLDEN, MORSE_CU, load_golden, system_from_golden
from oracle import oracle as O
from tdmd_claude_amd import inputs as I


def test_lattice_pe_forces_bitexact_vs_reference():
    g = load_golden("ref_lattice")
    s = I.read_lammps_data(GOLDEN / "cu_fcc_256.data")
    h = O.HalfList(s.pos, s.lo, s.hi, MORSE_CU[3], 1.0)
    pe, f = O.morse_forces(h, *MORSE_CU)
    assert pe == g["morse_pe"]
    assert np.array_equal(f, g["morse_f"])
    e = O.Eam(path=GOLDEN / "Cu_mishin1.eam.alloy")
    h2 = O.HalfList(s.pos, s.lo, s.hi, e.cutoff, 1.0)
    r = e.forces(h2, s.types)
    assert r["pe"] == g["eam_pe"]
    assert np.array_equal(r["forces"], g["eam_f"])


@pytest.mark.parametrize("which", ["morse", "eam"])
def test_lattice_forces_vs_lammps_dump(which):
    # verifylab run0-force-match: LAMMPS forces on the perfect lattice are
    # rounding noise (~1e-15 eV/A); the oracle's must be too.
    d = np.loadtxt(GOLDEN / f"forces_{which}.dump", skiprows=9)
    d = d[np.argsort(d[:, 0])]
    s = I.read_lammps_data(GOLDEN / "cu_fcc_256.data")
    assert np.array_equal(s.pos, d[:, 2:5])
    if which == "morse":
        h = O.HalfList(s.pos, s.lo, s.hi, MORSE_CU[3], 1.0)
        _, f = O.morse_forces(h, *MORSE_CU)
    else:
        e = O.Eam(path=GOLDEN / "Cu_mishin1.eam.alloy")
        f = e.forces(O.HalfList(s.pos, s.lo, s.hi, e.cutoff, 1.0), s.types)["forces"]
    assert np.abs(f - d[:, 5:8]).max() < 1e-13


@pytest.mark.parametrize("name,pot", [("ref_c1_morse", "morse"), ("ref_c1_mishin", "mishin"),
                                      ("ref_c1_johnson", "johnson"), ("ref_alloy2", "alloy2"),
                                      ("ref_alloy3", "alloy3")])
def test_trajectory_bitexact_vs_reference_driver(name, pot):
    # alloy2 / alloy3: two- and three-element setfl tables on random-species
    # crystals, so the type -> element mapping (rho of the neighbour's
    # element, F of the atom's own) and the i <= j pair-block order
    # (eam_alloy.cpp:167-191) are pinned against the reference's reader and
    # force loop
    g = load_golden(name)
    s = system_from_golden(g)
    th = g["thermo"]
    nsteps = int(th[-1, 0])
    every = int(th[1, 0] - th[0, 0])
    if pot == "morse":
        kw = dict(morse=MORSE_CU)
    elif pot == "mishin":
        kw = dict(eam=O.Eam(path=GOLDEN / "Cu_mishin1.eam.alloy"))
    elif pot.startswith("alloy"):
        from golden.make_golden import alloy_tables

        kw = dict(eam=O.Eam(tables=alloy_tables(pot)))
    else:
        kw = dict(eam=O.Eam(tables=I.johnson_setfl()))
    pos, vel, f, thermo, rebuilds = O.run(s, nsteps=nsteps, thermo_every=every, **kw)
    # thermo is printed with 14 decimals by the reference driver
    assert np.allclose(thermo[:, 1:3], th[:, 1:3], rtol=0, atol=6e-15 * 1000)
    assert np.abs(thermo[:, 1] - th[:, 1]).max() <= 5e-14 * max(1.0, np.abs(th[:, 1]).max())
    assert np.array_equal(pos, g["pos_final"])
    assert np.array_equal(f, g["f_final"])
    assert rebuilds == int(g["rebuilds"])
    assert np.array_equal(O.run(s, nsteps=0, **kw)[2], g["f0"])


@pytest.mark.parametrize("name", ["ref_alloy2", "ref_alloy3"])
def test_alloy_golden_inputs_are_regenerated(name):
    # the fixtures' inputs are what the generator makes (so the GPU tests can
    # rebuild them), and every element is present
    from golden.make_golden import alloy_system

    g = load_golden(name)
    s = alloy_system(name.split("_")[1])
    assert np.array_equal(s.pos, g["pos0"]) and np.array_equal(s.vel, g["vel0"])
    assert np.array_equal(s.types, g["types"]) and np.array_equal(s.masses, g["masses"])
    assert len(np.unique(g["types"])) == len(g["masses"]) - 1


def test_two_atom_morse_analytic():
    ref = json.loads((GOLDEN / "two_atoms_morse_analytic.json").read_text())
    p = ref["potential"]
    r = ref["step0"]["r_A"]
    e, fpair = O.morse_pair(p["D0_eV"], p["alpha_inv_A"], p["r0_A"], p["cutoff_A"], r * r)
    assert e == pytest.approx(ref["step0"]["U_eV"], rel=1e-14)
    # force on atom 1 = fpair * (x1 - x2) = fpair * (-r)
    assert fpair * (-r) == pytest.approx(ref["step0"]["force_on_atom_1_eV_per_A"][0], rel=1e-14)


def test_morse_closed_forms():
    D, a, r0, rc = MORSE_CU
    e, f = O.morse_pair(D, a, r0, rc, r0 * r0)
    assert e == pytest.approx(-D, rel=1e-15) and abs(f) < 1e-15
    e, f = O.morse_pair(D, a, r0, rc, rc * rc)
    assert e == 0.0 and f == 0.0  # r2 >= rc^2 -> zero, no shift (morse.hpp:37-42)
    for r in (2.3, 2.9, 4.0, 5.5):
        h = 1e-6
        ep, _ = O.morse_pair(D, a, r0, rc, (r + h) ** 2)
        em, _ = O.morse_pair(D, a, r0, rc, (r - h) ** 2)
        _, f = O.morse_pair(D, a, r0, rc, r * r)
        assert f * r == pytest.approx(-(ep - em) / (2 * h), rel=1e-7)


def test_fd_hermite_spline_hand_values():
    v = np.array([0.0, 1.0, 4.0, 9.0, 16.0])  # x^2 on a unit grid
    a, b, c, d = O.spline_coeffs(v, 1.0)
    assert np.array_equal(a, v[:-1])  # knots reproduced
    # slopes: one-sided at the ends, mean of the adjacent secants inside
    assert b.tolist() == [1.0, 2.0, 4.0, 6.0]
    # interior intervals get exact derivatives of x^2 -> the cubic is x^2
    for x0 in (1, 2):
        for t in (0.25, 0.5, 0.75):
            val = a[x0] + t * (b[x0] + t * (c[x0] + t * d[x0]))
            assert val == pytest.approx((x0 + t) ** 2, rel=1e-15)


def test_spline_eval_clamps_index():
    t = I.johnson_setfl(nr=200, nrho=200)
    e = O.Eam(tables=t)
    s = e.spline("density", 0)
    # beyond the last knot the last interval's cubic is extrapolated
    r = t.dr * (t.nr - 1) + 0.3
    dx = r - (t.nr - 2) * t.dr
    expect = s["a"][-1] + dx * (s["b"][-1] + dx * (s["c"][-1] + dx * s["d"][-1]))
    assert e.eval("density", 0, r) == pytest.approx(expect, rel=1e-14)


def random_gas(n, L, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(0, L, size=(n, 3))


@pytest.mark.parametrize("case", ["c1", "gas", "thin"])
def test_half_list_equals_bruteforce(case):
    if case == "c1":
        s = system_from_golden(load_golden("ref_c1_morse"))
        pos, lo, hi, rc, skin = s.pos, s.lo, s.hi, 6.0, 1.0
    elif case == "gas":
        pos, lo, hi, rc, skin = random_gas(400, 21.0, 7), np.zeros(3), np.full(3, 21.0), 4.0, 0.7
    else:  # three cells along one axis, cell edge < r_list
        pos, lo, hi, rc, skin = random_gas(300, 1.0, 3) * [14.0, 30.0, 30.0], np.zeros(3), \
            np.array([14.0, 30.0, 30.0]), 5.0, 1.0
    h = O.HalfList(pos, lo, hi, rc, skin)
    counts, offsets, nbrs = h.full()
    bc, bo, bn = O.bruteforce_full(pos, lo, hi, rc + skin)
    assert np.array_equal(counts, bc)
    for i in range(len(pos)):
        assert np.array_equal(np.sort(nbrs[offsets[i]:offsets[i] + counts[i]]),
                              np.sort(bn[bo[i]:bo[i] + bc[i]]))
    # half list: only j > i
    for i in range(len(pos)):
        row = h.nbrs[h.offsets[i]:h.offsets[i] + h.counts[i]]
        assert np.all(row > i)


def test_needs_rebuild_half_skin():
    s = system_from_golden(load_golden("ref_c1_morse"))
    h = O.HalfList(s.pos, s.lo, s.hi, 6.0, 1.0)
    p = s.pos.copy()
    p[10, 0] += 0.4999
    assert not h.needs_rebuild(p)
    p[10, 0] += 0.0002
    assert h.needs_rebuild(p)


def test_eam_forces_finite_difference():
    s = I.fcc_lattice(3, 3, 3, 3.615)
    I.jitter_uniform(s, 0.08, 11)
    e = O.Eam(tables=I.johnson_setfl())

    def energy(pos):
        return e.forces(O.HalfList(pos, s.lo, s.hi, e.cutoff, 0.5), s.types)["pe"]

    f = e.forces(O.HalfList(s.pos, s.lo, s.hi, e.cutoff, 0.5), s.types)["forces"]
    h = 1e-5
    for i in (0, 17, 50):
        for a in range(3):
            p = s.pos.copy()
            p[i, a] += h
            ep = energy(p)
            p[i, a] -= 2 * h
            em = energy(p)
            assert f[i, a] == pytest.approx(-(ep - em) / (2 * h), abs=2e-7)


def test_eam_density_bruteforce():
    s = I.fcc_lattice(3, 3, 3, 3.615)
    I.jitter_uniform(s, 0.05, 5)
    t = I.johnson_setfl()
    e = O.Eam(tables=t)
    r = e.forces(O.HalfList(s.pos, s.lo, s.hi, e.cutoff, 0.5), s.types)
    L = s.hi - s.lo
    rho = np.zeros(s.natoms)
    for i in range(s.natoms):
        d = s.pos[i] - s.pos
        d -= np.where(d > 0.5 * L, L, 0) - np.where(d < -0.5 * L, L, 0)
        rr = np.sqrt((d * d).sum(axis=1))
        for j in np.nonzero((rr < e.cutoff) & (np.arange(s.natoms) != i))[0]:
            rho[i] += e.eval("density", 0, rr[j])
    assert np.allclose(r["rho"], rho, rtol=1e-12, atol=0)


def test_nhc_half_step_by_hand():
    kB = 8.617333262e-5
    T, tp, dt, ndof, M = 300.0, 0.1, 0.001, 765, 3
    c = O.Nhc(T, tp, M, dt, ndof)
    ke = 12.3
    s = c.half_step(ke)
    # published chain update: outermost first, inward sweep with exp(-v_{k+1} dt/8)
    kt = kB * T
    Q = [ndof * kt * tp * tp, kt * tp * tp, kt * tp * tp]
    v = [0.0, 0.0, 0.0]
    h = 0.5 * dt
    v[2] += (Q[1] * v[1] ** 2 - kt) / Q[2] * h
    ef = np.exp(-v[2] * h / 4)
    v[1] = (v[1] * ef + (Q[0] * v[0] ** 2 - kt) / Q[1] * h) * ef
    ef = np.exp(-v[1] * h / 4)
    v[0] = (v[0] * ef + (2 * ke - ndof * kt) / Q[0] * h) * ef
    assert s == pytest.approx(np.exp(-v[0] * h), rel=1e-14)
    eta, eta_v, eta_m = c.state()
    assert eta_v == pytest.approx(v, rel=1e-12)
    assert eta == pytest.approx([x * h for x in v], rel=1e-12)


def test_nhc_single_thermostat_closed_form():
    c = O.Nhc(500.0, 0.05, 1, 0.002, 30)
    kB = 8.617333262e-5
    ke = 1.0
    Q = 30 * kB * 500.0 * 0.05 ** 2
    v0 = (2 * ke - 30 * kB * 500.0) / Q * 0.001
    assert c.half_step(ke) == pytest.approx(np.exp(-v0 * 0.001), rel=1e-14)


def _chain_run(t_target, t_period, dt, kes, M=3, ndof=765):
    c = O.Nhc(t_target, t_period, M, dt, ndof)
    scales = [c.half_step(ke) for ke in kes]
    eta, eta_v, eta_m = c.state()
    return np.array(scales), eta, eta_v, eta_m


_KES = [12.3, 9.1, 15.7, 11.0, 8.4, 13.9]  # uneven, so every chain velocity moves


def test_nhc_invariant_under_energy_and_temperature_scaling():
    # (KE, T) -> (c KE, c T): every G_k is a difference of energies over a
    # mass that scales with kT, so the chain velocities, positions and the
    # velocity scale do not change. c = 4 is exact in binary floating point,
    # so the half-steps must agree bit for bit; a mass or force term missing
    # its kT (Q_1 = n_dof tau^2, G_k = (Q v^2 - 1)/Q, ...) breaks it.
    a = _chain_run(300.0, 0.1, 0.001, _KES)
    b = _chain_run(1200.0, 0.1, 0.001, [4.0 * ke for ke in _KES])
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert np.array_equal(4.0 * a[3], b[3])
    assert np.all(a[2] != 0.0)


def test_nhc_invariant_under_time_scaling():
    # (tau, dt, v_k) -> (c tau, c dt, v_k / c): Q ~ tau^2, each kick is G dt,
    # each friction factor exp(-v dt), so with chains started at rest the
    # velocities come out divided by c and the positions and the velocity
    # scale are unchanged, bit for bit for c = 2. A mass linear in tau, or
    # a dt factor that does not pair with a 1/tau^2 or a v (dt^2, a missing
    # half), breaks it.
    a = _chain_run(300.0, 0.1, 0.001, _KES)
    b = _chain_run(300.0, 0.2, 0.002, _KES)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], 2.0 * b[2])
    assert np.array_equal(4.0 * a[3], b[3])


def test_nvt_holds_temperature():
    s = I.fcc_lattice(4, 4, 4, 3.615)
    I.init_velocities(s, 600.0, 9)  # start hot, thermostat to 300 K
    _, _, _, th, _ = O.run(s, eam=O.Eam(tables=I.johnson_setfl()), nsteps=3000,
                           thermo_every=50, nvt=True, t_target=300.0, t_period=0.1)
    t_late = th[th[:, 0] >= 1500, 4] * (3 * s.natoms) / (3 * s.natoms - 3)
    assert abs(t_late.mean() - 300.0) < 0.15 * 300.0


def test_nhc_conserved_energy_single_thermostat():
    # With one thermostat the reference half-step is a consistent (first
    # order) splitting, so KE + PE + chain energy is conserved up to O(dt).
    # Longer chains are not: the reference damps v_k by exp(-v_{k+1} dt/8)
    # twice per dt/2 force update (device_nose_hoover.cu:251-261), half the
    # friction the chain equations need, so the extended energy drifts.
    s = I.fcc_lattice(4, 4, 4, 3.615)
    I.init_velocities(s, 600.0, 9)
    dev = {}
    for M in (1, 3):
        _, _, _, th, _ = O.run(s, eam=O.Eam(tables=I.johnson_setfl()), nsteps=3000,
                               thermo_every=100, nvt=True, t_target=300.0, t_period=0.1,
                               nhc_length=M)
        h = th[:, 3] + th[:, 5]
        dev[M] = np.abs(h - h[0]).max()
        if M == 1:
            exchanged = np.abs(th[:, 5]).max()
            assert exchanged > 0.5
            assert dev[1] < 0.02 * exchanged
    assert dev[3] > 5 * dev[1]


# ---------------------------------------------------------------------------
# Force matching: the oracle's analytic gradients against central finite
# differences of the oracle's own loss (FP64). These pin the checker of the
# GPU adjoint; a dropped term (e.g. the density chain through F'') or a
# wrong factor fails them.
# ---------------------------------------------------------------------------

def _fm_system():
    s = I.fcc_lattice(3, 3, 3, 3.615)
    I.displace_gaussian(s, 0.05, 5)
    return s


def test_fm_morse_gradient_finite_difference():
    s = _fm_system()
    half = O.HalfList(s.pos, s.lo, s.hi, 6.0, 0.0)
    _, target = O.morse_forces(half, 0.36, 1.30, 2.90, 6.0)
    th = np.array([0.3429, 1.3588, 2.866])
    L, g, f = O.fm_morse(half, *th, 6.0, target)
    _, f_ref = O.morse_forces(half, *th, 6.0)
    assert np.array_equal(f, f_ref)
    assert L == pytest.approx(float(np.sum((f_ref - target) ** 2)), rel=1e-14)
    for k in range(3):
        h = 1e-6
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        fd = (O.fm_morse(half, *tp, 6.0, target)[0] - O.fm_morse(half, *tm, 6.0, target)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-7)


def _eam_values(t):
    F = np.asarray(t.F).reshape(-1)[: t.nrho]
    rho = np.asarray(t.rho).reshape(-1)[: t.nr]
    rphi = np.asarray(t.rphi).reshape(-1)[: t.nr]
    phi = np.zeros(t.nr)
    phi[1:] = rphi[1:] / (np.arange(1, t.nr) * t.dr)
    return F, rho, phi


def test_fm_eam_gradient_finite_difference():
    s = _fm_system()
    t = I.johnson_setfl(nr=500, nrho=500)
    F, rho, phi = _eam_values(t)
    half = O.HalfList(s.pos, s.lo, s.hi, t.cutoff, 0.0)
    tgt_model = O.EamValues(F * 1.03, rho * 0.98, phi * 1.05, t.drho, t.dr, t.cutoff)
    target = tgt_model.forces(half, np.ones(len(s.pos), dtype=np.int32))["forces"]
    model = O.EamValues(F, rho, phi, t.drho, t.dr, t.cutoff)
    L, G, f = O.fm_eam(model, half, target)
    assert L == pytest.approx(float(np.sum((f - target) ** 2)), rel=1e-14)
    rng = np.random.default_rng(3)
    base = {"F": F, "rho": rho, "phi": phi}
    for name in ("F", "rho", "phi"):
        gv = G[name]
        live = np.nonzero(np.abs(gv) > 1e-3 * np.abs(gv).max())[0]
        for k in rng.choice(live, 3, replace=False):
            h = 1e-5 * max(1.0, abs(base[name][k]))

            def loss(delta):
                v = {kk: vv.copy() for kk, vv in base.items()}
                v[name][k] += delta
                m = O.EamValues(v["F"], v["rho"], v["phi"], t.drho, t.dr, t.cutoff)
                return O.fm_eam(m, half, target)[0]

            fd = (loss(h) - loss(-h)) / (2 * h)
            # F'' of the FD-Hermite spline jumps at knots, which limits the
            # finite difference for the density knots
            assert gv[k] == pytest.approx(fd, rel=2e-5 if name == "rho" else 1e-8), (name, k)


def test_spline_value_gradient_transpose():
    # <H^T gc, dv> == <gc, H dv> for the FD-Hermite map H (values -> coefficients)
    rng = np.random.default_rng(11)
    n, dr = 37, 0.1
    gc = rng.normal(size=(4, n - 1))
    dv = rng.normal(size=n)
    gv = O.spline_coef_grad_to_values(n, dr, *gc)
    coef = np.array(O.spline_coeffs(dv, dr))
    assert float(gv @ dv) == pytest.approx(float(np.sum(gc * coef)), rel=1e-12)
