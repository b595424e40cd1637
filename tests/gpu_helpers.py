"""Helpers that drive the drop-in C-ABI entry points on the GPU (tests only)."""
import ctypes as C
import sys

import numpy as np

from tdmd_claude_amd import capi
from tdmd_claude_amd.capi import DeviceArray as D


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def box_of(lo, hi, periodic=(1, 1, 1)):
    return capi.Box.make(lo, hi, periodic)


def cell_build(pos, lo, hi, r_list):
    box = box_of(lo, hi)
    n = (C.c_int32 * 3)()
    cs = (C.c_double * 3)()
    capi.call("tdmd_cell_grid", C.byref(box), r_list, n, cs)
    ncells = n[0] * n[1] * n[2]
    N = len(pos)
    d_pos = D.from_host(np.ascontiguousarray(pos, dtype=np.float64))
    cc, co = D(ncells, np.int32), D(ncells, np.int32)
    ca, ac = D(max(N, 1), np.int32), D(max(N, 1), np.int32)
    wb = C.c_size_t()
    capi.call("tdmd_cell_workspace_bytes", N, ncells, C.byref(wb))
    work = D(wb.value, np.uint8)
    capi.call("tdmd_cell_build", d_pos.ptr, N, C.byref(box), r_list, cc.ptr, co.ptr, ca.ptr,
              ac.ptr, work.ptr, None)
    capi.call("tdmd_device_synchronize")
    return dict(box=box, n=n, ncells=ncells, d_pos=d_pos, cc=cc, co=co, ca=ca, ac=ac, work=work)


def csr_build(pos, lo, hi, r_cut, r_skin):
    """DeviceNeighborList::build through the C-ABI: count, scan, fill."""
    r_list = r_cut + r_skin
    c = cell_build(pos, lo, hi, r_list)
    N = len(pos)
    counts = D(N, np.int32)
    offsets = D(N, np.int32)
    capi.call("tdmd_nlist_count", c["d_pos"].ptr, N, C.byref(c["box"]), r_list, c["cc"].ptr,
              c["co"].ptr, c["ca"].ptr, c["ac"].ptr, c["n"], counts.ptr, None)
    sb = C.c_size_t()
    capi.call("tdmd_scan_workspace_bytes", N, C.byref(sb))
    sw = D(sb.value, np.uint8)
    meta = D(2, np.int32)
    capi.call("tdmd_scan_exclusive_i32", counts.ptr, offsets.ptr, N, meta.ptr, sw.ptr, None)
    total, mx = meta.download()
    nbrs = D(max(int(total), 1), np.int32)
    counts2 = D(N, np.int32)
    ovf = D(1, np.int32)
    ovf.zero()
    capi.call("tdmd_nlist_fill", c["d_pos"].ptr, N, C.byref(c["box"]), r_list, c["cc"].ptr,
              c["co"].ptr, c["ca"].ptr, c["ac"].ptr, c["n"], offsets.ptr, nbrs.ptr, counts2.ptr,
              int(mx), ovf.ptr, None)
    capi.call("tdmd_device_synchronize")
    return dict(cells=c, counts=counts2, offsets=offsets, nbrs=nbrs, total=int(total),
                max=int(mx), overflow=int(ovf.download()[0]), counts_pass1=counts)


def morse_gpu(prec, pos, lo, hi, csr, params, energy=True, deterministic=True):
    N = len(pos)
    dt = np.float32 if prec == capi.TDMD_MIXED else np.float64
    f = D((N, 3), dt)
    e = D(1, np.float64)
    e.zero()
    work = D(N // 256 + 2, np.float64) if deterministic else None
    p = capi.MorseParams.make(*params)
    box = box_of(lo, hi)
    capi.call("tdmd_morse_force", prec, csr["cells"]["d_pos"].ptr, f.ptr, csr["nbrs"].ptr,
              csr["offsets"].ptr, csr["counts"].ptr, N, C.byref(box), C.byref(p),
              e.ptr if energy else None, work.ptr if work is not None else None, None)
    capi.call("tdmd_device_synchronize")
    return f.download().astype(np.float64), float(e.download()[0])


class EamTables:
    def __init__(self, prec, t):
        self.h = C.c_void_p()
        F = np.ascontiguousarray(t.F, dtype=np.float64)
        rho = np.ascontiguousarray(t.rho, dtype=np.float64)
        rphi = np.ascontiguousarray(t.rphi, dtype=np.float64)
        capi.call("tdmd_eam_tables_create", C.byref(self.h), prec, t.ntypes, t.nrho, t.drho,
                  t.nr, t.dr, t.cutoff, capi.ptr(F), capi.ptr(rho), capi.ptr(rphi))
        self.prec = prec
        self.cutoff = t.cutoff

    def __del__(self):
        try:
            capi.lib().tdmd_eam_tables_destroy(self.h)
        except Exception:
            pass


def eam_gpu(tabs: EamTables, pos, types, lo, hi, csr, energy=True):
    N = len(pos)
    dt = np.float32 if tabs.prec == capi.TDMD_MIXED else np.float64
    f = D((N, 3), dt)
    rho = D(N, np.float64)
    fp = D(N, dt)
    e = D(1, np.float64)
    e.zero()
    work = D(N // 256 + 2, np.float64)
    d_types = D.from_host(np.ascontiguousarray(types, dtype=np.int32))
    box = box_of(lo, hi)
    capi.call("tdmd_eam_compute", tabs.h, csr["cells"]["d_pos"].ptr, f.ptr, d_types.ptr,
              csr["nbrs"].ptr, csr["offsets"].ptr, csr["counts"].ptr, N, C.byref(box), rho.ptr,
              fp.ptr, e.ptr if energy else None, work.ptr, None)
    capi.call("tdmd_device_synchronize")
    return dict(forces=f.download().astype(np.float64), pe=float(e.download()[0]),
                rho=rho.download(), fp=fp.download().astype(np.float64))
