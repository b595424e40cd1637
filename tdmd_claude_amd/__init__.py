"""MI355X (gfx950) implementation of TDMD's single-GPU MD step path.

The compute lives in ``lib/libtdmd_hip.so`` (hand-written HIP kernels behind
the C-ABI of ``include/tdmd_hip.h``); this package holds its ctypes binding
(:mod:`.capi`), the FastPipelineScheduler mirror (:mod:`.scheduler`) and the
input generators (:mod:`.inputs`).
"""
from . import capi, inputs  # noqa: F401

__all__ = ["capi", "inputs", "scheduler"]
