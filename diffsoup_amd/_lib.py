"""ctypes binding of libdiffsoup_hip.so (the C-ABI declared in include/diffsoup_hip.h).

The library is built in-tree by ``python -m diffsoup_amd.build_lib``.  There is
no fallback: if the library is missing or fails to load, every operator raises.
"""
from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdiffsoup_hip.so")

DS_OK = 0
DS_STAT_COVERED, DS_STAT_PAIRS, DS_STAT_OVERFLOW, DS_STAT_BAD_INDEX = 0, 1, 2, 3
DS_STAT_COUNT = 8

_lib = None

_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_sz = ctypes.c_size_t
_f = ctypes.c_float
_ip = ctypes.POINTER(ctypes.c_int)

_SIGNATURES = {
    "ds_version": ([], ctypes.c_char_p),
    "ds_last_error": ([], ctypes.c_char_p),
    "ds_feats_at_level": ([_i], ctypes.c_uint32),
    "ds_raster_tile_count": ([_i, _ip, _ip, _i], _i64),
    "ds_raster_workspace_size": ([_i, _ip, _ip, _i, _i, _i64], _sz),
    "ds_raster_stats_offset": ([], _i),
    "ds_rasterize_fwd": ([_i, _ip, _ip, _i, _vp, _i, _vp, _i, _vp, _i, _u64, _i, _vp, _vp, _sz, _i64, _vp], _i),
    "ds_opacity_aux_bwd": ([_i, _ip, _ip, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _vp, _vp, _i, _vp, _sz, _i64, _vp], _i),
    "ds_pixel_l1": ([_i64, _i, _vp, _vp, _vp, _vp], _i),
    "ds_multires_color_fwd": ([_i64, _vp, _i, _i, _i, _i, _vp, _i, _vp, _vp], _i),
    "ds_multires_color_bwd": ([_i64, _vp, _i, _i, _i, _i, _vp, _i, _vp, _vp], _i),
    "ds_accumulate_fwd": ([_i, _i, _i, _i, _i, _vp, _vp, _vp], _i),
    "ds_accumulate_bwd": ([_i, _i, _i, _i, _i, _vp, _vp, _vp], _i),
    "ds_sh2": ([_i, _ip, _ip, _vp, _vp, _vp, _vp], _i),
    "ds_edge_grad_bwd": ([_i, _ip, _ip, _i, _vp, _vp, _vp, _i, _vp, _vp, _i, _vp, _vp], _i),
    "ds_fill_f32": ([_vp, _i64, _f, _vp], _i),
    "ds_count_triangle_ids": ([_i64, _vp, _i, _vp, _i, _vp], _i),
}

# Entry points added by later translation units bind when present.
_OPTIONAL = {}


class DiffSoupError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DiffSoupError(
                f"{LIB_PATH} is missing: build it with `python -m diffsoup_amd.build_lib` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        for name, (args, res) in _OPTIONAL.items():
            if hasattr(L, name):
                fn = getattr(L, name)
                fn.argtypes = args
                fn.restype = res
        _lib = L
    return _lib


def register(name, args, res=_i):
    """Declare an entry point's ctypes signature (used by sibling modules)."""
    _OPTIONAL[name] = (args, res)
    if _lib is not None and hasattr(_lib, name):
        fn = getattr(_lib, name)
        fn.argtypes = args
        fn.restype = res


def check(rc: int, what: str) -> None:
    if rc != DS_OK:
        msg = lib().ds_last_error().decode(errors="replace")
        raise DiffSoupError(f"{what} failed ({rc}): {msg}")


def int_array(values):
    arr = (ctypes.c_int * len(values))(*[int(v) for v in values])
    return arr


def ptr(t: torch.Tensor | None):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def stream_ptr(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_device(t: torch.Tensor, name: str) -> None:
    if not t.is_cuda:
        raise DiffSoupError(f"{name} must be a GPU tensor (the HIP library has no CPU path)")
