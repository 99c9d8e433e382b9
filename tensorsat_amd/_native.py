"""
ctypes binding of ``libtensorsat_amd.so`` (``include/tensorsat_amd.h``).

The HIP library is the only compute path: there is no CPU or PyTorch fallback.  If the
library is missing or no gfx950 device is usable, :func:`lib` raises.  PyTorch is used
only to allocate device memory and to name the current HIP stream.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libtensorsat_amd.so"

TS_OK = 0
TS_EINVAL = -1
TS_ENOMEM = -2
TS_EUNSUPPORTED = -3
TS_EOVERFLOW = -4
TS_EHIP = -5
BATCH_LDS_MAX = 65536
BATCH_SEGMENTS = 8  # queue counters per dispatch (TS_BATCH_SEGMENTS)
BATCH_COUNTER_STRIDE = 64  # int32s between queue counters (TS_BATCH_COUNTER_STRIDE)
BATCH_COUNTER_INTS = BATCH_SEGMENTS * BATCH_COUNTER_STRIDE

MODE_GT0 = 0
MODE_MOD256 = 1
MODES = {"GT0": MODE_GT0, "MOD256": MODE_MOD256}

# Every symbol declared in include/tensorsat_amd.h.
EXPORTS = (
    "ts_init", "ts_version", "ts_contract2", "ts_rewire", "ts_spider", "ts_union", "ts_stack",
    "ts_select", "ts_as_scalar", "ts_diagonal", "ts_count_contract2", "ts_bytes_to_u64",
    "ts_program_run", "ts_program_count", "ts_batch_run", "ts_batch_max_workgroups",
    "ts_batch_compact", "ts_batch_count", "ts_batch_max_workgroups_lds", "ts_gnn_gemm", "ts_gnn_segsum", "ts_gnn_relu_back",
    "ts_gnn_head", "ts_gnn_adam", "ts_gnn_cast_bf16", "ts_dense_config", "ts_dense_tiles", "ts_dense_tile_of",
    "ts_search_run", "ts_search_run_queue", "ts_search_decide_noalloc", "ts_program_run_fused", "ts_program_run_gang",
)

_lib = None
_ready = False

vp = ctypes.c_void_p
i32 = ctypes.c_int32
i64 = ctypes.c_int64
P = ctypes.POINTER


class Batch(ctypes.Structure):
    _fields_ = [
        ("jobs", vp), ("order", vp), ("num_jobs", i32), ("steps", vp), ("leaves", vp),
        ("arenas", vp), ("arena_stride", i64), ("num_workgroups", i32), ("mode", i32),
        ("verdicts", vp), ("counter", vp), ("early_exit", i32), ("job_done", vp),
        ("num_jobs_dev", vp), ("lds_arena", i32),
    ]


def load() -> ctypes.CDLL:
    """Load the shared library without touching the GPU (symbol checks on CPU)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m tensorsat_amd.build hip`"
            )
        L = ctypes.CDLL(str(LIB_PATH))
        L.ts_init.argtypes = [ctypes.c_int]
        L.ts_version.restype = ctypes.c_char_p
        L.ts_contract2.argtypes = [vp, i32, P(i32), vp, i32, P(i32), vp, i32, P(i32), i32, vp]
        L.ts_count_contract2.argtypes = [vp, i32, P(i32), vp, i32, P(i32), vp, i32, P(i32), vp, vp]
        L.ts_rewire.argtypes = [vp, i32, vp, i32, P(i32), i32, vp]
        L.ts_diagonal.argtypes = [vp, i32, P(i32), vp, i32, vp]
        L.ts_spider.argtypes = [vp, i32, i32, i32, vp]
        L.ts_union.argtypes = [vp, vp, vp, i64, vp]
        L.ts_stack.argtypes = [P(vp), i32, i32, i32, i32, vp, vp]
        L.ts_select.argtypes = [vp, i32, P(i32), vp, vp]
        L.ts_as_scalar.argtypes = [vp, i64, P(i32), vp]
        L.ts_bytes_to_u64.argtypes = [vp, vp, i64, vp]
        L.ts_program_run.argtypes = [vp, vp, vp, vp, i32, vp, vp, vp]
        L.ts_program_count.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.ts_program_run_fused.argtypes = [vp, vp, vp, vp, vp, i32, i32, vp]
        L.ts_program_run_gang.argtypes = [vp, i32, vp, vp, vp, vp, i64, vp, i32, i32, vp]
        L.ts_batch_run.argtypes = [P(Batch), vp]
        L.ts_batch_compact.argtypes = [vp, i32, vp, vp, vp, vp, vp]
        L.ts_batch_count.argtypes = [P(Batch), vp, vp, vp]
        i8 = ctypes.c_int64
        f32 = ctypes.c_float
        L.ts_gnn_gemm.argtypes = [vp, i8, i8, vp, i8, i8, vp, i8, i32, i32, i32, vp, i32, i32, i32, vp]
        L.ts_gnn_segsum.argtypes = [vp, vp, vp, i32, vp, i8, i32, i32, i32, vp, i8, i32, i32, i32, i32, vp]
        L.ts_gnn_relu_back.argtypes = [vp, i8, vp, i8, i32, i32, vp, i8, vp, vp]
        L.ts_gnn_head.argtypes = [vp, i8, i32, vp, vp, vp, vp, f32, vp, vp, vp, vp, vp, i32, i32, vp]
        L.ts_gnn_adam.argtypes = [vp, vp, vp, vp, vp, i8, f32, f32, f32, f32, i32, f32, vp]
        L.ts_gnn_cast_bf16.argtypes = [vp, vp, i8, vp]
        L.ts_batch_max_workgroups.argtypes = [P(i32)]
        L.ts_batch_max_workgroups_lds.argtypes = [i64, i32, P(i32)]
        L.ts_dense_config.argtypes = [i32, i32]
        L.ts_dense_tiles.argtypes = [i32, i32, i32, i32, i32]
        L.ts_dense_tile_of.argtypes = [i32, i32, i32, i32, P(i32)]
        L.ts_search_run.argtypes = [vp, vp]
        L.ts_search_run_queue.argtypes = [vp, vp, i32, vp, vp, i32, vp]
        L.ts_search_decide_noalloc.argtypes = [vp, P(i32), vp]
        _lib = L
    return _lib


def lib() -> ctypes.CDLL:
    """Load the library and initialise the current torch device (must be gfx950)."""
    global _ready
    L = load()
    if not _ready:
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("tensorsat_amd needs a ROCm GPU (gfx950); none is visible")
        check(L.ts_init(torch.cuda.current_device()), "ts_init")
        _ready = True
    return L


def check(rc: int, what: str) -> None:
    if rc == TS_OK:
        return
    if rc == TS_EINVAL:
        raise ValueError(f"{what}: invalid arguments")
    if rc == TS_EUNSUPPORTED:
        raise NotImplementedError(f"{what}: unsupported case")
    if rc == TS_EOVERFLOW:
        raise OverflowError(f"{what}: count exceeds 2^64")
    if rc == TS_ENOMEM:
        raise MemoryError(f"{what}: out of memory")
    raise RuntimeError(f"{what}: HIP error ({rc})")


def stream_ptr():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def i32arr(xs) -> ctypes.Array:
    xs = list(xs)
    return (i32 * max(1, len(xs)))(*xs)


def dptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def host_ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)
