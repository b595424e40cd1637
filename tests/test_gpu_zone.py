"""Zone pipeline kernels (the PipelineScheduler's per-zone calls):
compute_morse_gpu_zone (src/potentials/device_morse_zone.cu:29-157) and the
zone velocity-Verlet functions (src/integrator/device_velocity_verlet_
zone.cu:14-118). A zone pass computes the same per-atom sums as the whole
pass, restricted to [first, first + count), and ADDS forces; the zone
integrators touch only their atoms."""
import ctypes as C

import numpy as np
import pytest

from conftest import MORSE_CU
from gpu_helpers import box_of, csr_build, morse_gpu, say
from oracle import oracle as O
from tdmd_claude_amd import capi
from tdmd_claude_amd import inputs as I
from tdmd_claude_amd.capi import DeviceArray as D

pytestmark = pytest.mark.gpu

FP64, MIXED = capi.TDMD_FP64, capi.TDMD_MIXED
E_INVALID = 1  # tdmd_status TDMD_E_INVALID
ZONES = [(0, 700), (700, 1), (701, 0), (701, 1347)]  # 2048 atoms, uneven, one empty


def system():
    s = I.fcc_lattice(8, 8, 8, 3.615)
    I.jitter_uniform(s, 0.03, 17)
    I.init_velocities(s, 400.0, 5)
    return s


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_morse_zones_compose_the_whole_pass(prec):
    s = system()
    N = s.natoms
    say("morse zones", prec)
    g = csr_build(s.pos, s.lo, s.hi, MORSE_CU[3], 1.0)
    f_all, e_all = morse_gpu(prec, s.pos, s.lo, s.hi, g, MORSE_CU)
    rdt = np.float32 if prec == MIXED else np.float64
    f = D((N, 3), rdt)
    f.upload(np.full((N, 3), 7.0, dtype=rdt))  # stale values the zeroing must clear
    e = D(1, np.float64)
    e.zero()
    work = D(N // 256 + 2, np.float64)
    p = capi.MorseParams.make(*MORSE_CU)
    box = box_of(s.lo, s.hi)
    for first, count in ZONES:
        capi.call("tdmd_zero_forces_zone", prec, f.ptr, first, count, None)
        capi.call("tdmd_morse_force_zone", prec, g["cells"]["d_pos"].ptr, f.ptr, g["nbrs"].ptr,
                  g["offsets"].ptr, g["counts"].ptr, first, count, C.byref(box), C.byref(p),
                  e.ptr, work.ptr, None)
    capi.call("tdmd_device_synchronize")
    fz = f.download().astype(np.float64)
    ez = float(e.download()[0])
    # per atom the same row, the same order: bitwise the whole pass
    assert np.array_equal(fz, f_all)
    assert ez == pytest.approx(e_all, rel=1e-12)
    if prec == FP64:
        h = O.HalfList(s.pos, s.lo, s.hi, MORSE_CU[3], 1.0)
        pe_ref, f_ref = O.morse_forces(h, *MORSE_CU)
        assert np.abs(fz - f_ref).max() < 1e-12 and ez == pytest.approx(pe_ref, rel=1e-12)
    # forces are added: a second pass over one zone doubles that zone only
    first, count = ZONES[0]
    capi.call("tdmd_morse_force_zone", prec, g["cells"]["d_pos"].ptr, f.ptr, g["nbrs"].ptr,
              g["offsets"].ptr, g["counts"].ptr, first, count, C.byref(box), C.byref(p), None,
              None, None)
    capi.call("tdmd_device_synchronize")
    f2 = f.download().astype(np.float64)
    assert np.array_equal(f2[first:first + count], 2 * fz[first:first + count])
    assert np.array_equal(f2[first + count:], fz[first + count:])


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_zone_integrators_touch_only_their_atoms(prec):
    s = system()
    N = s.natoms
    rdt = np.float32 if prec == MIXED else np.float64
    rng = np.random.Generator(np.random.PCG64(8))
    f = rng.normal(0, 1.0, (N, 3)).astype(rdt)
    masses = s.masses.astype(rdt)
    d_v, d_x, d_f = D.from_host(s.vel), D.from_host(s.pos), D.from_host(f)
    d_t, d_m = D.from_host(s.types), D.from_host(masses)
    box = box_of(s.lo, s.hi)
    first, count, dt = 300, 1000, 0.001
    capi.call("tdmd_half_kick_zone", prec, d_v.ptr, d_f.ptr, d_t.ptr, d_m.ptr, first, count, dt,
              None)
    capi.call("tdmd_drift_zone", d_x.ptr, d_v.ptr, first, count, dt, C.byref(box), None)
    capi.call("tdmd_zero_forces_zone", prec, d_f.ptr, 0, 200, None)
    capi.call("tdmd_device_synchronize")
    v, x = s.vel.copy(), s.pos.copy()
    z = slice(first, first + count)
    vz, xz = v[z].copy(), x[z].copy()
    O.half_kick(vz, f[z].astype(np.float64), s.types[z], masses.astype(np.float64), dt)
    O.drift(xz, vz, s.lo, s.hi, dt, single_image=True)
    v[z], x[z] = vz, xz
    gv, gx, gf = d_v.download(), d_x.download(), d_f.download()
    assert np.allclose(gv, v, rtol=1e-14, atol=1e-15) and np.allclose(gx, x, rtol=1e-14, atol=1e-13)
    outside = np.ones(N, bool)
    outside[z] = False
    assert np.array_equal(gv[outside], s.vel[outside]) and np.array_equal(gx[outside], s.pos[outside])
    assert not gf[:200].any() and np.array_equal(gf[200:], f[200:])


def test_zone_arguments_are_checked():
    box = box_of(np.zeros(3), np.full(3, 10.0))
    p = capi.MorseParams.make(*MORSE_CU)
    lib = capi.lib()
    assert lib.tdmd_morse_force_zone(FP64, None, None, None, None, None, -1, 5, C.byref(box),
                                     C.byref(p), None, None, None) == E_INVALID
    assert lib.tdmd_morse_force_zone(FP64, None, None, None, None, None, 3, 0, C.byref(box),
                                     C.byref(p), None, None, None) == 0  # empty zone: no-op
    assert lib.tdmd_half_kick_zone(FP64, None, None, None, None, 0, -2, 0.001, None) == \
        E_INVALID
    assert lib.tdmd_drift_zone(None, None, 0, 0, 0.001, C.byref(box), None) == 0
    assert lib.tdmd_zero_forces_zone(FP64, None, 0, 4, None) == E_INVALID
