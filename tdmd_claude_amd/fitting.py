"""Force-matching fit on the GPU: loss L = sum |F(theta) - F_target|^2 over a
fixed set of configurations and its adjoint gradient, behind
``torch.autograd.Function`` so a torch optimizer can drive it.

theta is either the Morse triple (D, alpha, r0) or the knot values of a
one-element EAM in EamAlloy's form: [F(rho) | rho(r) | phi(r)] with
phi = (r phi) / r (src/potentials/eam_alloy.cpp:152-191). The loss and the
gradient come from the HIP library (tdmd_fm_*); torch only holds the device
buffers and runs the optimizer.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import capi
from .inputs import SetflTables, System, displace_gaussian, fcc_lattice

try:  # torch is the plumbing of the fitting harness only
    import torch
except ImportError:  # pragma: no cover
    torch = None


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclasses.dataclass
class EamGrid:
    nrho: int
    drho: float
    nr: int
    dr: float
    cutoff: float

    @property
    def nvalues(self) -> int:
        return self.nrho + 2 * self.nr


def eam_values(t: SetflTables) -> tuple[np.ndarray, EamGrid]:
    """Knot values [F | rho | phi] of a one-element setfl table."""
    if t.ntypes != 1:
        raise ValueError("force matching fits one-element EAM tables")
    F = np.asarray(t.F, dtype=np.float64).reshape(-1)[: t.nrho]
    rho = np.asarray(t.rho, dtype=np.float64).reshape(-1)[: t.nr]
    rphi = np.asarray(t.rphi, dtype=np.float64).reshape(-1)[: t.nr]
    r = np.arange(t.nr) * t.dr
    phi = np.zeros(t.nr)
    phi[1:] = rphi[1:] / r[1:]
    return np.concatenate([F, rho, phi]), EamGrid(t.nrho, t.drho, t.nr, t.dr, t.cutoff)


class ForceMatchSet:
    """A fixed set of configurations on the device (tdmd_fm)."""

    def __init__(self, lo, hi, snapshots: np.ndarray, r_cut: float, precision: str = "mixed",
                 periodic=(1, 1, 1)):
        snaps = np.ascontiguousarray(snapshots, dtype=np.float64)
        if snaps.ndim != 3 or snaps.shape[2] != 3:
            raise ValueError("snapshots must have shape (nsnap, natoms, 3)")
        self.nsnap, self.natoms = snaps.shape[0], snaps.shape[1]
        self.precision = precision
        self._box = capi.Box.make(lo, hi, periodic)
        h = C.c_void_p()
        prec = capi.TDMD_MIXED if precision == "mixed" else capi.TDMD_FP64
        capi.call("tdmd_fm_create", C.byref(h), prec, C.byref(self._box), self.natoms,
                  self.nsnap, _pd(snaps), float(r_cut))
        self._h = h

    @property
    def ntot(self) -> int:
        return self.nsnap * self.natoms

    def npairs(self) -> int:
        n = C.c_int32()
        p = C.c_int64()
        capi.call("tdmd_fm_info", self._h, C.byref(n), C.byref(p))
        return p.value

    def set_targets(self, forces: np.ndarray) -> None:
        f = np.ascontiguousarray(forces, dtype=np.float64).reshape(self.ntot, 3)
        capi.call("tdmd_fm_set_targets", self._h, _pd(f))

    # -- Morse -------------------------------------------------------------
    def morse(self, params, rc: float, want_forces: bool = False):
        """(loss, grad[3], forces or None) for Morse (D, alpha, r0)."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        loss = C.c_double()
        grad = np.zeros(3)
        f = np.zeros((self.ntot, 3)) if want_forces else None
        capi.call("tdmd_fm_morse", self._h, _pd(p), float(rc), C.byref(loss), _pd(grad),
                  None if f is None else _pd(f))
        return loss.value, grad, f

    def morse_forces(self, params, rc: float) -> np.ndarray:
        p = np.ascontiguousarray(params, dtype=np.float64)
        f = np.zeros((self.ntot, 3))
        capi.call("tdmd_fm_morse", self._h, _pd(p), float(rc), None, None, _pd(f))
        return f

    # -- EAM (device buffers) ------------------------------------------------
    def eam(self, values, grid: EamGrid, want_loss=True, want_grad=True, want_forces=False):
        """values: torch float64 tensor on the GPU ([F | rho | phi]).
        Returns (loss tensor[1] or None, grad tensor or None, forces or None)."""
        if torch is None:
            raise RuntimeError("torch is required for device buffers")
        v = values.detach().contiguous()
        if v.dtype != torch.float64 or not v.is_cuda or v.numel() != grid.nvalues:
            raise ValueError("values must be a float64 CUDA tensor of nrho + 2 nr entries")
        loss = torch.zeros(1, dtype=torch.float64, device=v.device) if want_loss else None
        grad = torch.zeros_like(v) if want_grad else None
        forces = (torch.zeros((self.ntot, 3), dtype=torch.float64, device=v.device)
                  if want_forces else None)
        stream = torch.cuda.current_stream(v.device).cuda_stream
        capi.call("tdmd_fm_eam", self._h, grid.nrho, grid.drho, grid.nr, grid.dr, grid.cutoff,
                  C.c_void_p(v.data_ptr()),
                  None if loss is None else C.c_void_p(loss.data_ptr()),
                  None if grad is None else C.c_void_p(grad.data_ptr()),
                  None if forces is None else C.c_void_p(forces.data_ptr()),
                  C.c_void_p(stream))
        return loss, grad, forces

    def eam_coef_grad(self, grid: EamGrid) -> np.ndarray:
        g = np.zeros(4 * (grid.nrho - 1) + 8 * (grid.nr - 1))
        capi.call("tdmd_fm_eam_coef_grad", self._h, _pd(g))
        return g

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            capi.lib().tdmd_fm_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


if torch is not None:

    class MorseForceMatch(torch.autograd.Function):
        """L(D, alpha, r0); params is a float64 tensor of 3 (any device)."""

        @staticmethod
        def forward(ctx, params, fm: ForceMatchSet, rc: float):
            p = params.detach().cpu().numpy()
            loss, grad, _ = fm.morse(p, rc)
            ctx.save_for_backward(torch.as_tensor(grad, dtype=params.dtype, device=params.device))
            return torch.tensor(loss, dtype=params.dtype, device=params.device)

        @staticmethod
        def backward(ctx, gout):
            (grad,) = ctx.saved_tensors
            return gout * grad, None, None

    class EamForceMatch(torch.autograd.Function):
        """L(knot values); values is a float64 CUDA tensor [F | rho | phi]."""

        @staticmethod
        def forward(ctx, values, fm: ForceMatchSet, grid: EamGrid):
            loss, grad, _ = fm.eam(values, grid)
            ctx.save_for_backward(grad)
            return loss[0]

        @staticmethod
        def backward(ctx, gout):
            (grad,) = ctx.saved_tensors
            return gout * grad, None, None


class RelativeHatBasis:
    """Smooth relative parametrisation of EAM knot values for fitting:
    v = v0 * (1 + P c), with P piecewise-linear hat functions on `nodes`
    evenly spaced nodes per function (F, rho, phi). Adam's per-coordinate
    steps on raw knots put sign-like noise on neighbouring knots, which the
    spline derivative amplifies by 1/dr; relative hat coefficients move the
    tables smoothly and on every magnitude scale at once. The gradient wrt c
    is the chain rule through P on the knot-value gradient of the adjoint."""

    def __init__(self, values0, grid: EamGrid, nodes: int = 32):
        self.v0 = values0.detach().clone()
        dev, dt = values0.device, values0.dtype
        self.P = [self._hat(grid.nrho, nodes, dev, dt), self._hat(grid.nr, nodes, dev, dt),
                  self._hat(grid.nr, nodes, dev, dt)]
        self.nodes = nodes

    @staticmethod
    def _hat(n, k, dev, dt):
        x = torch.arange(n, dtype=dt, device=dev)[:, None]
        nodes = torch.linspace(0, n - 1, k, dtype=dt, device=dev)[None, :]
        h = (n - 1) / (k - 1)
        return torch.clamp(1.0 - torch.abs(x - nodes) / h, min=0.0)

    def zeros(self):
        return torch.zeros(3 * self.nodes, dtype=self.v0.dtype, device=self.v0.device)

    def expand(self, c):
        cs = torch.split(c, self.nodes)
        rel = torch.cat([P @ ci for P, ci in zip(self.P, cs)])
        return self.v0 * (1.0 + rel)


def fc_snapshots(cells: int, nsnap: int, sigma: float, seed: int, a: float = 3.615):
    """nsnap FCC configurations with N(0, sigma^2) displacements, seeds
    seed, seed+1, ... -> (System of the first, array (nsnap, N, 3))."""
    base = fcc_lattice(cells, cells, cells, a)
    snaps = []
    for k in range(nsnap):
        s = base.copy()
        displace_gaussian(s, sigma, seed + k)
        snaps.append(s.pos)
    return base, np.stack(snaps)


def fit(loss_fn, x0, iters: int = 200, lr: float = 1e-3, log_every: int = 0):
    """Adam on loss_fn(x) from x0 (a leaf tensor). Returns (x, losses)."""
    x = x0.detach().clone().requires_grad_(True)
    opt = torch.optim.Adam([x], lr=lr)
    losses = []
    for it in range(iters):
        opt.zero_grad()
        loss = loss_fn(x)
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
        if log_every and it % log_every == 0:
            print(f"iter {it}: loss {losses[-1]:.6e}", flush=True)
    return x.detach(), losses
