"""Loader for libstencils_hip.so (the C ABI in include/stencils_hip.h).

There is no CPU fallback: if the library is missing or no HIP device is
visible, every accelerated call raises ``BackendUnavailable``.
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

from . import _abi

PKG_ROOT = Path(__file__).resolve().parent.parent  # stencils.jl_amd/
LIB_PATH = Path(os.environ.get("STENCILS_HIP_LIB", PKG_ROOT / "lib" / "libstencils_hip.so"))

# Every symbol include/stencils_hip.h declares; tests check the library exports them.
EXPORTS = {
    "sth_version": (ctypes.c_int, []),
    "sth_last_error": (ctypes.c_char_p, []),
    "sth_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "sth_ctx_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sth_alloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "sth_free": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "sth_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]),
    "sth_synchronize": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "sth_update_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_abi.SthArray),
                                           ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                           ctypes.c_void_p]),
    "sth_map": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_abi.SthArray),
                               ctypes.POINTER(_abi.SthArray), ctypes.POINTER(_abi.SthStencil),
                               ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                               ctypes.POINTER(_abi.SthOp), ctypes.c_uint32, ctypes.c_void_p]),
    "sth_iterate": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_abi.SthArray),
                                   ctypes.POINTER(_abi.SthArray), ctypes.POINTER(_abi.SthStencil),
                                   ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                   ctypes.POINTER(_abi.SthOp), ctypes.c_int64, ctypes.c_uint32,
                                   ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p]),
    "sth_map_multi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(_abi.SthTerm),
                                     ctypes.POINTER(_abi.SthArray), ctypes.c_int32, ctypes.c_uint32,
                                     ctypes.c_void_p]),
    "sth_scatter": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_abi.SthTerm), ctypes.POINTER(_abi.SthTerm),
                                   ctypes.POINTER(_abi.SthArray), ctypes.POINTER(_abi.SthScatterRule), ctypes.c_uint32,
                                   ctypes.c_void_p]),
    "sth_map_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(_abi.SthJob),
                                     ctypes.c_uint32, ctypes.c_void_p]),
    "sth_map_batch_static": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(_abi.SthJob),
                                            ctypes.c_uint32, ctypes.c_void_p]),
    "sth_jit_selftest": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "sth_iterate_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(_abi.SthJob),
                                         ctypes.POINTER(ctypes.c_int64), ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p]),
    "sth_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_abi.SthArray),
                                    ctypes.POINTER(_abi.SthArray), ctypes.POINTER(_abi.SthStencil),
                                    ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                    ctypes.POINTER(_abi.SthOp), ctypes.c_void_p, ctypes.c_char_p,
                                    ctypes.c_size_t]),
    "sth_tuning_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.c_size_t)]),
    "sth_set_variant": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "sth_server_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_abi.SthServerConfig),
                                         ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "sth_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_abi.SthRequest)]),
    "sth_server_start": (ctypes.c_int, [ctypes.c_void_p]),
    "sth_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_abi.SthCompletion), ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
    "sth_server_drain": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "sth_server_stats_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_abi.SthServerStats)]),
    "sth_server_destroy": (ctypes.c_int, [ctypes.c_void_p]),
}


class BackendUnavailable(RuntimeError):
    """libstencils_hip.so is not built or no HIP device is visible."""


class StencilsHipError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


_lock = threading.Lock()
_lib = None
_ctx = {}


def load_library(path: Path | None = None):
    """dlopen the library and attach prototypes (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path else LIB_PATH
        if not p.exists():
            raise BackendUnavailable(
                f"{p} is not built: run `python __graft_entry__.py` (build()) first")
        lib = ctypes.CDLL(str(p))
        for name, (res, args) in EXPORTS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def lib():
    return load_library()


def check(status: int):
    if status == _abi.STH_OK:
        return
    msg = lib().sth_last_error().decode(errors="replace")
    if status == _abi.STH_ENOTSUP:
        from .ops import Unsupported
        raise Unsupported(msg)
    if status == _abi.STH_EINVAL:
        raise ValueError(msg)
    raise StencilsHipError(status, msg)


def tuning_table(device: int = 0) -> dict:
    """The autotuner's decisions: {family|bucket: variant}."""
    L = lib()
    need = ctypes.c_size_t(0)
    check(L.sth_tuning_table(context(device), None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(max(need.value, 1))
    check(L.sth_tuning_table(context(device), buf, need.value, None))
    return dict(line.split("\t", 1) for line in buf.value.decode().splitlines() if "\t" in line)


def set_variant(pattern: str = "", device: int = 0):
    """Force the first kernel variant whose name contains `pattern` ("" = tuner)."""
    check(lib().sth_set_variant(context(device), pattern.encode()))


def context(device: int = 0):
    """The library context of `device`, created on first use."""
    c = _ctx.get(device)
    if c is not None:
        return c
    import torch
    if not torch.cuda.is_available():
        raise BackendUnavailable("no HIP device is visible; the GPU path has no CPU fallback")
    L = lib()
    with _lock:
        c = _ctx.get(device)
        if c is None:
            h = ctypes.c_void_p()
            st = L.sth_ctx_create(device, ctypes.byref(h))
            if st != _abi.STH_OK:
                raise StencilsHipError(st, L.sth_last_error().decode())
            c = h
            _ctx[device] = c
        return c
