"""Force matching on the GPU (tdmd_fm_*, tdmd_claude_amd.fitting): loss and
adjoint gradients against the FP64 oracle (whose gradients are pinned by
finite differences in test_oracle.py), the GPU FP64 gradient against its own
central differences, and short Adam fits through torch.autograd."""
import numpy as np
import pytest

from gpu_helpers import say
from oracle import oracle as O
from tdmd_claude_amd import fitting
from tdmd_claude_amd import inputs as I

MORSE = (0.3429, 1.3588, 2.866)
MORSE_TARGET = (0.36, 1.30, 2.90)
RC_MORSE = 6.0


def snapshots(n=3, cells=3):
    return fitting.fc_snapshots(cells, n, 0.05, 5)


def oracle_morse(base, snaps, params, targets=None):
    L, g, fs = 0.0, np.zeros(3), []
    for k, p in enumerate(snaps):
        half = O.HalfList(p, base.lo, base.hi, RC_MORSE, 0.0)
        if targets is None:
            fs.append(O.morse_forces(half, *params, RC_MORSE)[1])
        else:
            l, gk, f = O.fm_morse(half, *params, RC_MORSE, targets[k])
            L += l
            g += gk
            fs.append(f)
    return L, g, np.stack(fs)


def eam_case(nr=500):
    t = I.johnson_setfl(nr=nr, nrho=nr)
    values, grid = fitting.eam_values(t)
    F, rho, phi = np.split(values, [grid.nrho, grid.nrho + grid.nr])
    return t, values, grid, (F, rho, phi)


def oracle_eam(base, snaps, grid, F, rho, phi, targets=None):
    model = O.EamValues(F, rho, phi, grid.drho, grid.dr, grid.cutoff)
    L, G, fs = 0.0, None, []
    gco = None
    for k, p in enumerate(snaps):
        half = O.HalfList(p, base.lo, base.hi, grid.cutoff, 0.0)
        if targets is None:
            fs.append(model.forces(half, np.ones(len(p), dtype=np.int32))["forces"])
            continue
        l, g, f = O.fm_eam(model, half, targets[k])
        L += l
        gv = np.concatenate([g["F"], g["rho"], g["phi"]])
        G = gv if G is None else G + gv
        gco = g["coef"] if gco is None else gco + g["coef"]
        fs.append(f)
    return L, G, gco, np.stack(fs)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fm_morse_loss_and_gradient_vs_oracle(prec):
    say(f"fm morse {prec}")
    base, snaps = snapshots()
    _, _, targets = oracle_morse(base, snaps, MORSE_TARGET)
    L0, g0, f0 = oracle_morse(base, snaps, MORSE, targets)
    fm = fitting.ForceMatchSet(base.lo, base.hi, snaps, RC_MORSE, prec)
    fm.set_targets(targets)
    L, g, f = fm.morse(MORSE, RC_MORSE, want_forces=True)
    tol = 1e-10 if prec == "fp64" else 2e-4
    rms = np.sqrt(np.mean(np.sum(f0.reshape(-1, 3) ** 2, axis=1)))
    assert np.max(np.abs(f - f0.reshape(-1, 3))) / rms < (1e-12 if prec == "fp64" else 1e-5)
    assert L == pytest.approx(L0, rel=tol)
    assert np.linalg.norm(g - g0) / np.linalg.norm(g0) < tol
    fm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,nr", [("fp64", 500), ("mixed", 500), ("mixed", 5000)])
def test_fm_eam_loss_and_gradient_vs_oracle(prec, nr):
    import torch

    say(f"fm eam {prec} nr={nr}")
    base, snaps = snapshots(n=2)
    t, values, grid, (F, rho, phi) = eam_case(nr)
    _, _, _, targets = oracle_eam(base, snaps, grid, F * 1.03, rho * 0.98, phi * 1.05)
    L0, G0, C0, f0 = oracle_eam(base, snaps, grid, F, rho, phi, targets)
    fm = fitting.ForceMatchSet(base.lo, base.hi, snaps, grid.cutoff, prec)
    fm.set_targets(targets)
    v = torch.tensor(values, dtype=torch.float64, device="cuda")
    loss, grad, forces = fm.eam(v, grid, want_forces=True)
    torch.cuda.synchronize()
    f = forces.cpu().numpy()
    rms = np.sqrt(np.mean(np.sum(f0.reshape(-1, 3) ** 2, axis=1)))
    ferr = np.max(np.linalg.norm(f - f0.reshape(-1, 3), axis=1)) / rms
    L = float(loss.cpu()[0])
    G = grad.cpu().numpy()
    C = fm.eam_coef_grad(grid)
    gerr = np.linalg.norm(G - G0) / np.linalg.norm(G0)
    cerr = np.linalg.norm(C - C0) / np.linalg.norm(C0)
    say(f"  force err {ferr:.2e}, loss {L:.6e} vs {L0:.6e}, grad err {gerr:.2e}, coef err {cerr:.2e}")
    if prec == "fp64":
        assert ferr < 1e-12
        assert L == pytest.approx(L0, rel=1e-10)
        assert gerr < 1e-9 and cerr < 1e-9
    else:
        assert ferr < 1e-5
        assert L == pytest.approx(L0, rel=2e-4)
        assert gerr < 2e-3 and cerr < 2e-3
    fm.close()


@pytest.mark.gpu
def test_fm_eam_fp64_gradient_matches_its_finite_differences():
    import torch

    say("fm eam fd")
    base, snaps = snapshots(n=2)
    t, values, grid, (F, rho, phi) = eam_case(500)
    _, _, _, targets = oracle_eam(base, snaps, grid, F * 1.03, rho * 0.98, phi * 1.05)
    fm = fitting.ForceMatchSet(base.lo, base.hi, snaps, grid.cutoff, "fp64")
    fm.set_targets(targets)
    v = torch.tensor(values, dtype=torch.float64, device="cuda")
    _, grad, _ = fm.eam(v, grid)
    G = grad.cpu().numpy()
    rng = np.random.default_rng(7)
    checked = 0
    for lo, hi, tol in [(0, grid.nrho, 1e-6), (grid.nrho, grid.nrho + grid.nr, 3e-5),
                        (grid.nrho + grid.nr, grid.nvalues, 1e-6)]:
        live = lo + np.nonzero(np.abs(G[lo:hi]) > 1e-3 * np.abs(G[lo:hi]).max())[0]
        for k in rng.choice(live, 3, replace=False):
            h = 1e-5 * max(1.0, abs(values[k]))
            vp, vm = v.clone(), v.clone()
            vp[k] += h
            vm[k] -= h
            lp = float(fm.eam(vp, grid, want_grad=False)[0].cpu()[0])
            lm = float(fm.eam(vm, grid, want_grad=False)[0].cpu()[0])
            fd = (lp - lm) / (2 * h)
            assert G[k] == pytest.approx(fd, rel=tol), k
            checked += 1
    assert checked == 9
    fm.close()


@pytest.mark.gpu
def test_fm_adam_fits_reduce_the_loss():
    import torch

    say("fm adam")
    base, snaps = snapshots(n=2)
    # Morse: recover perturbed parameters
    _, _, targets = oracle_morse(base, snaps, MORSE_TARGET)
    fm = fitting.ForceMatchSet(base.lo, base.hi, snaps, RC_MORSE, "mixed")
    fm.set_targets(targets)
    x0 = torch.tensor(MORSE, dtype=torch.float64)
    x, losses = fitting.fit(lambda p: fitting.MorseForceMatch.apply(p, fm, RC_MORSE), x0,
                            iters=150, lr=5e-3)
    assert losses[-1] < 0.05 * losses[0]
    fm.close()
    # EAM knot values through autograd
    t, values, grid, (F, rho, phi) = eam_case(500)
    _, _, _, targets = oracle_eam(base, snaps, grid, F * 1.03, rho * 0.98, phi * 1.05)
    fm = fitting.ForceMatchSet(base.lo, base.hi, snaps, grid.cutoff, "mixed")
    fm.set_targets(targets)
    v0 = torch.tensor(values, dtype=torch.float64, device="cuda")
    basis = fitting.RelativeHatBasis(v0, grid, nodes=32)
    c, losses = fitting.fit(
        lambda c: fitting.EamForceMatch.apply(basis.expand(c), fm, grid), basis.zeros(),
        iters=40, lr=1e-3)
    say(f"  EAM fit loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert losses[-1] < 0.1 * losses[0]
    fm.close()
