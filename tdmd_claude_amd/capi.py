"""ctypes binding of the C-ABI in ``include/tdmd_hip.h``.

Loads ``tdmd_claude_amd/lib/libtdmd_hip.so`` (built in-tree by
``tdmd_claude_amd.build``). There is no fallback: if the library is missing
or fails to load, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libtdmd_hip.so"

TDMD_OK = 0
TDMD_FP64 = 0
TDMD_MIXED = 1
REBUILD_FIXED = 0
REBUILD_DISPLACEMENT = 1

STATUS = {
    0: "TDMD_OK",
    1: "TDMD_E_INVALID",
    2: "TDMD_E_HIP",
    3: "TDMD_E_OVERFLOW",
    4: "TDMD_E_NONFINITE",
    5: "TDMD_E_NOMEM",
    6: "TDMD_E_STATE",
    7: "TDMD_E_IO",
}


class TdmdError(RuntimeError):
    """A C-ABI call returned a non-zero tdmd_status."""

    def __init__(self, status: int, func: str, msg: str):
        super().__init__(f"{func} -> {STATUS.get(status, status)}: {msg}")
        self.status = status


class Box(C.Structure):
    """tdmd_box (src/core/box.hpp:17-26)."""

    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3), ("periodic", C.c_int32 * 3)]

    @classmethod
    def make(cls, lo, hi, periodic=(1, 1, 1)) -> "Box":
        b = cls()
        for a in range(3):
            b.lo[a] = float(lo[a])
            b.hi[a] = float(hi[a])
            b.periodic[a] = int(bool(periodic[a]))
        return b


class MorseParams(C.Structure):
    """tdmd_morse_params (MorseParams, src/potentials/device_morse.cuh:14-20)."""

    _fields_ = [("d", C.c_double), ("alpha", C.c_double), ("r0", C.c_double),
                ("rc", C.c_double), ("rc_sq", C.c_double)]

    @classmethod
    def make(cls, d, alpha, r0, rc) -> "MorseParams":
        return cls(float(d), float(alpha), float(r0), float(rc), float(rc) * float(rc))


class SchedConfig(C.Structure):
    """tdmd_sched_config (FastPipelineConfig + additions)."""

    _fields_ = [
        ("dt", C.c_double),
        ("r_skin", C.c_double),
        ("rebuild_every", C.c_int32),
        ("precision", C.c_int32),
        ("rebuild_policy", C.c_int32),
        ("t_target", C.c_double),
        ("t_period", C.c_double),
        ("nhc_length", C.c_int32),
        ("use_graph", C.c_int32),
        ("block_size", C.c_int32),
        ("mirror_fp64", C.c_int32),
        ("max_neighbors", C.c_int32),
        ("pair_unroll", C.c_int32),
        ("lds_budget", C.c_int32),
        ("split_rebuild", C.c_int32),
        ("disable_halo", C.c_int32),
        ("pair_lanes", C.c_int32),
        ("cell_pct", C.c_int32),
        ("reserved", C.c_int32 * 2),
    ]


class Observables(C.Structure):
    """tdmd_observables."""

    _fields_ = [
        ("step", C.c_int64),
        ("natoms", C.c_int64),
        ("pe", C.c_double),
        ("ke", C.c_double),
        ("te", C.c_double),
        ("temp", C.c_double),
        ("nhc_energy", C.c_double),
        ("conserved", C.c_double),
        ("virial", C.c_double),
        ("volume", C.c_double),
        ("pressure", C.c_double),
        ("max_neighbors", C.c_int32),
        ("pad", C.c_int32),
        ("mean_neighbors", C.c_double),
        ("mean_near_neighbors", C.c_double),
    ]


class PipeConfig(C.Structure):
    """tdmd_pipe_config (PipelineConfig)."""

    _fields_ = [
        ("dt", C.c_double),
        ("r_skin", C.c_double),
        ("n_zones", C.c_int32),
        ("n_streams", C.c_int32),
        ("rebuild_every", C.c_int32),
        ("deterministic", C.c_int32),
        ("t_target", C.c_double),
        ("t_period", C.c_double),
        ("nhc_length", C.c_int32),
        ("adaptive_dt", C.c_int32),
        ("dt_max", C.c_double),
        ("dt_min", C.c_double),
        ("c2", C.c_double),
        ("precision", C.c_int32),
        ("reserved", C.c_int32 * 7),
    ]


class ZoneInfo(C.Structure):
    """tdmd_zone_info (Zone)."""

    _fields_ = [
        ("id", C.c_int32),
        ("atom_offset", C.c_int32),
        ("natoms", C.c_int32),
        ("time_step", C.c_int32),
        ("state", C.c_int32),
        ("lattice_index", C.c_int32 * 3),
        ("lo", C.c_double * 3),
        ("hi", C.c_double * 3),
    ]


_VP = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_D = C.c_double
_SZ = C.c_size_t
_PD = C.POINTER(C.c_double)
_PI32 = C.POINTER(C.c_int32)

_PROTOS = {
    "tdmd_abi_version": (C.c_int, []),
    "tdmd_last_error": (C.c_char_p, []),
    "tdmd_device_count": (C.c_int, [_PI32]),
    "tdmd_set_device": (C.c_int, [_I32]),
    "tdmd_device_name": (C.c_int, [C.c_char_p, _I32]),
    "tdmd_malloc": (C.c_int, [C.POINTER(_VP), _SZ]),
    "tdmd_free": (C.c_int, [_VP]),
    "tdmd_memcpy_htod": (C.c_int, [_VP, _VP, _SZ, _VP]),
    "tdmd_memcpy_dtoh": (C.c_int, [_VP, _VP, _SZ, _VP]),
    "tdmd_memcpy_dtod": (C.c_int, [_VP, _VP, _SZ, _VP]),
    "tdmd_fill_u32": (C.c_int, [_VP, C.c_uint32, _SZ, _VP]),
    "tdmd_copy_u32": (C.c_int, [_VP, _VP, _SZ, _VP]),
    "tdmd_stream_create": (C.c_int, [C.POINTER(_VP)]),
    "tdmd_stream_destroy": (C.c_int, [_VP]),
    "tdmd_stream_synchronize": (C.c_int, [_VP]),
    "tdmd_device_synchronize": (C.c_int, []),
    "tdmd_event_create": (C.c_int, [C.POINTER(_VP)]),
    "tdmd_event_destroy": (C.c_int, [_VP]),
    "tdmd_event_record": (C.c_int, [_VP, _VP]),
    "tdmd_event_elapsed_ms": (C.c_int, [_VP, _VP, C.POINTER(C.c_float)]),
    "tdmd_scan_workspace_bytes": (C.c_int, [_I64, C.POINTER(_SZ)]),
    "tdmd_scan_exclusive_i32": (C.c_int, [_VP, _VP, _I64, _VP, _VP, _VP]),
    "tdmd_cell_grid": (C.c_int, [C.POINTER(Box), _D, _PI32, _PD]),
    "tdmd_cell_workspace_bytes": (C.c_int, [_I64, _I32, C.POINTER(_SZ)]),
    "tdmd_cell_build": (C.c_int, [_VP, _I64, C.POINTER(Box), _D, _VP, _VP, _VP, _VP, _VP, _VP]),
    "tdmd_nlist_count": (C.c_int, [_VP, _I64, C.POINTER(Box), _D, _VP, _VP, _VP, _VP, _PI32,
                                   _VP, _VP]),
    "tdmd_nlist_fill": (C.c_int, [_VP, _I64, C.POINTER(Box), _D, _VP, _VP, _VP, _VP, _PI32,
                                  _VP, _VP, _VP, _I32, _VP, _VP]),
    "tdmd_morse_force": (C.c_int, [_I32, _VP, _VP, _VP, _VP, _VP, _I32, C.POINTER(Box),
                                   C.POINTER(MorseParams), _VP, _VP, _VP]),
    "tdmd_spline_build": (C.c_int, [_PD, _I32, _D, _PD, _PD, _PD, _PD]),
    "tdmd_setfl_read": (C.c_int, [C.c_char_p, _PI32, _PI32, _PD, _PI32, _PD, _PD, _PD, _PD,
                                  _PD]),
    "tdmd_eam_tables_create": (C.c_int, [C.POINTER(_VP), _I32, _I32, _I32, _D, _I32, _D, _D,
                                         _PD, _PD, _PD]),
    "tdmd_eam_tables_create_values": (C.c_int, [C.POINTER(_VP), _I32, _I32, _I32, _D, _I32,
                                                _D, _D, _PD, _PD, _PD]),
    "tdmd_eam_tables_destroy": (C.c_int, [_VP]),
    "tdmd_eam_tables_info": (C.c_int, [_VP, _PI32, _PD]),
    "tdmd_eam_compute": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _I32, C.POINTER(Box),
                                   _VP, _VP, _VP, _VP, _VP]),
    "tdmd_eam_compute_noalloc": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _I32,
                                           C.POINTER(Box), _VP, _VP]),
    "tdmd_half_kick": (C.c_int, [_I32, _VP, _VP, _VP, _VP, _I32, _D, _VP]),
    "tdmd_drift": (C.c_int, [_VP, _VP, _I32, _D, C.POINTER(Box), _VP]),
    "tdmd_fused_kick_drift": (C.c_int, [_I32, _I32, _VP, _VP, _VP, _VP, _VP, _I32, _D,
                                        C.POINTER(Box), _VP]),
    "tdmd_zero_forces": (C.c_int, [_I32, _VP, _I32, _VP]),
    "tdmd_morse_force_zone": (C.c_int, [_I32, _VP, _VP, _VP, _VP, _VP, _I32, _I32,
                                        C.POINTER(Box), C.POINTER(MorseParams), _VP, _VP, _VP]),
    "tdmd_half_kick_zone": (C.c_int, [_I32, _VP, _VP, _VP, _VP, _I32, _I32, _D, _VP]),
    "tdmd_drift_zone": (C.c_int, [_VP, _VP, _I32, _I32, _D, C.POINTER(Box), _VP]),
    "tdmd_zero_forces_zone": (C.c_int, [_I32, _VP, _I32, _I32, _VP]),
    "tdmd_reduce_workspace_bytes": (C.c_int, [_I64, C.POINTER(_SZ)]),
    "tdmd_compute_ke": (C.c_int, [_I32, _VP, _VP, _VP, _I32, _VP, _VP, _VP]),
    "tdmd_scale_velocities": (C.c_int, [_VP, _I32, _D, _VP]),
    "tdmd_compute_vmax": (C.c_int, [_VP, _I32, _VP, _VP, _VP]),
    "tdmd_nhc_half_step": (C.c_double, [_I32, _PD, _PD, _PD, _D, _D, _D, _I32, _D]),
    "tdmd_sched_config_default": (None, [C.POINTER(SchedConfig)]),
    "tdmd_sched_create_morse": (C.c_int, [C.POINTER(_VP), C.POINTER(Box), _I32,
                                          C.POINTER(MorseParams), C.POINTER(SchedConfig)]),
    "tdmd_sched_create_eam": (C.c_int, [C.POINTER(_VP), C.POINTER(Box), _I32, _I32, _I32, _D,
                                        _I32, _D, _D, _PD, _PD, _PD, C.POINTER(SchedConfig)]),
    "tdmd_sched_create_eam_tables": (C.c_int, [C.POINTER(_VP), C.POINTER(Box), _I32, _VP,
                                               C.POINTER(SchedConfig)]),
    "tdmd_sched_create_eam_batch": (C.c_int, [C.POINTER(_VP), _I32, C.POINTER(Box), _PI32,
                                              _I32, _I32, _D, _I32, _D, _D, _PD, _PD, _PD,
                                              C.POINTER(SchedConfig)]),
    "tdmd_sched_create_morse_batch": (C.c_int, [C.POINTER(_VP), _I32, C.POINTER(Box), _PI32,
                                                C.POINTER(MorseParams), C.POINTER(SchedConfig)]),
    "tdmd_sched_destroy": (C.c_int, [_VP]),
    "tdmd_sched_upload": (C.c_int, [_VP, _PD, _PD, _PD, _PI32, _PI32, _PD, _I32]),
    "tdmd_sched_run_until": (C.c_int, [_VP, _I32]),
    "tdmd_sched_download": (C.c_int, [_VP, _PD, _PD, _PD, _PI32, _PI32, _I32]),
    "tdmd_sched_stats": (C.c_int, [_VP, C.POINTER(_I64)]),
    "tdmd_sched_current_step": (C.c_int, [_VP, _PI32]),
    "tdmd_sched_interaction_cutoff": (C.c_int, [_VP, _PD]),
    "tdmd_sched_thermo": (C.c_int, [_VP, _PD]),
    "tdmd_sched_observables": (C.c_int, [_VP, C.POINTER(Observables)]),
    "tdmd_sched_stream": (C.c_int, [_VP, C.POINTER(_VP)]),
    "tdmd_sched_halo_info": (C.c_int, [_VP, C.POINTER(C.c_int64)]),
    "tdmd_sched_time_passes": (C.c_int, [_VP, _I32, C.POINTER(C.c_float)]),
    "tdmd_sched_pass_timing": (C.c_int, [_VP, _I32]),
    "tdmd_sched_pass_times": (C.c_int, [_VP, C.POINTER(C.c_float)]),
    "tdmd_sched_autotune": (C.c_int, [_VP, _I32, _PI32]),
    "tdmd_sched_kernel_shape": (C.c_int, [_VP, _PI32]),
    "tdmd_pipe_create": (C.c_int, [C.POINTER(_VP), C.POINTER(Box), _I32, C.POINTER(MorseParams),
                                   C.POINTER(PipeConfig)]),
    "tdmd_pipe_destroy": (C.c_int, [_VP]),
    "tdmd_pipe_upload": (C.c_int, [_VP, _PD, _PD, _VP, _PI32, _PI32, _VP, _I32]),
    "tdmd_pipe_run_until": (C.c_int, [_VP, _I32]),
    "tdmd_pipe_download": (C.c_int, [_VP, _PD, _PD, _VP, _PI32, _PI32, _I32]),
    "tdmd_pipe_zone_count": (C.c_int, [_VP, _PI32]),
    "tdmd_pipe_zones": (C.c_int, [_VP, C.POINTER(ZoneInfo), _I32]),
    "tdmd_pipe_zone_neighbors": (C.c_int, [_VP, _I32, _PI32, _I32, _PI32]),
    "tdmd_pipe_min_time_step": (C.c_int, [_VP, _PI32]),
    "tdmd_pipe_stats": (C.c_int, [_VP, C.POINTER(C.c_int64)]),
    "tdmd_pipe_current_dt": (C.c_int, [_VP, _PD]),
    "tdmd_fm_create": (C.c_int, [C.POINTER(_VP), _I32, C.POINTER(Box), _I32, _I32, _PD, _D]),
    "tdmd_fm_destroy": (C.c_int, [_VP]),
    "tdmd_fm_info": (C.c_int, [_VP, _PI32, C.POINTER(_I64)]),
    "tdmd_fm_set_targets": (C.c_int, [_VP, _PD]),
    "tdmd_fm_morse": (C.c_int, [_VP, _PD, _D, _PD, _PD, _PD]),
    "tdmd_fm_eam": (C.c_int, [_VP, _I32, _D, _I32, _D, _D, _VP, _VP, _VP, _VP, _VP]),
    "tdmd_fm_eam_coef_grad": (C.c_int, [_VP, _PD]),
}

_lib = None


def lib() -> C.CDLL:
    """Load the gfx950 library (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m tdmd_claude_amd.build`")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _PROTOS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def declared_symbols() -> list[str]:
    return list(_PROTOS)


def check(status: int, func: str = "") -> None:
    if status != TDMD_OK:
        msg = lib().tdmd_last_error()
        raise TdmdError(status, func, msg.decode() if msg else "")


def call(name: str, *args):
    """Call a status-returning entry point and raise on failure."""
    st = getattr(lib(), name)(*args)
    check(st, name)


def ptr(a: np.ndarray, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


class DeviceArray:
    """Owning device allocation holding a numpy-typed array (DeviceBuffer<T>)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        call("tdmd_malloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p

    @classmethod
    def from_host(cls, a: np.ndarray, dtype=None) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        call("tdmd_memcpy_htod", self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, None)

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        call("tdmd_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, None)
        return out

    def zero(self) -> None:
        call("tdmd_fill_u32", self.ptr, 0, self.nbytes // 4, None)
        call("tdmd_device_synchronize")

    def free(self) -> None:
        if self.ptr is not None and self.ptr.value:
            lib().tdmd_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
