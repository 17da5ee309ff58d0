"""Serving loop over the C ABI (sth_server_*): continuous batching of
iterated stencil solves with admission by an HBM budget.

A request is `steps` calls of ``A = mapstencil!(f, A)`` on a
SwitchingStencilArray (src/gatherstencil.jl:77-83).  The server copies the
initial state into its arena when the request is admitted, advances every
admitted request together (one batched launch per step for all of them;
requests of 2^21 cells or more with a two-step kernel run their own
launches) and copies the final state to the request's result buffer.

    srv = Server(hbm_budget=1 << 30)
    tag = srv.submit(Laplacian(0.15), A, steps=500, arrival=0.01)
    srv.start(); srv.drain()
    for c in srv.completions(): ...   # latency = c.finished - c.arrival
"""
from __future__ import annotations

import ctypes
import itertools

import numpy as np

from . import _abi, _lib
from ._abi import SthCompletion, SthRequest, SthServerConfig, SthServerStats
from .array import (SwitchingStencilArray, colmajor_copy, colmajor_empty, data_ptr,
                    elem_strides, is_device)
from .gather import _Call, _op_flags, _padval_bytes, _stream, result_dtype, stencil_descriptor
from .ops import resolve

SERIAL = 1


class Completion:
    """One finished request: server-clock seconds and the result buffer."""

    def __init__(self, c: SthCompletion, result):
        self.tag = int(c.tag)
        self.status = int(c.status)
        self.arrival = float(c.arrival)
        self.admitted = float(c.admitted)
        self.finished = float(c.finished)
        self.steps = int(c.steps)
        self.cells = int(c.cells)
        self.result = result

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def _dense_parent(x, device):
    """A dense column-major device copy of parent `x` (or `x` itself)."""
    shape = tuple(x.shape)
    dense = tuple(elem_strides(x)) == tuple(int(np.prod(shape[:d])) for d in range(len(shape)))
    if is_device(x) and dense:
        return x
    return colmajor_copy(x, device)


class Server:
    """Continuous-batching server on one device (see the module docstring)."""

    def __init__(self, hbm_budget: int, max_active: int = 0, round_steps: int = 8,
                 serial: bool = False, device: int | None = None):
        import torch
        L = _lib.lib()
        self._L = L
        self.device = torch.cuda.current_device() if device is None else device
        cfg = SthServerConfig(int(hbm_budget), int(max_active), int(round_steps),
                              SERIAL if serial else 0, 0)
        h = ctypes.c_void_p()
        _lib.check(L.sth_server_create(_lib.context(self.device), ctypes.byref(cfg),
                                       _stream(self.device), ctypes.byref(h)))
        self._h = h
        self._keep = {}
        self._tags = itertools.count(1)
        self._done = []

    def submit(self, f, A: SwitchingStencilArray, steps: int, arrival: float = 0.0,
               result=None, tag: int | None = None) -> int:
        """Queue `steps` calls of mapstencil!(f, A); returns the request tag.
        The final state (parent layout, dense) lands in `result` (a device
        tensor allocated here unless given)."""
        red = resolve(f)
        if result_dtype(f, A) != A.sth_dtype():
            raise ValueError("the result type differs from the array's element type")
        src = _dense_parent(A.source, self.device)
        if result is None:
            result = colmajor_empty(tuple(src.shape), src.dtype, src.device)
        dt = A.sth_dtype()
        keep = _Call()
        req = SthRequest()
        h = A.halo()
        req.state = _abi.make_array(data_ptr(src), dt, A.size(), elem_strides(src), h)
        req.stencil = stencil_descriptor(A.stencil, dt, keep)
        req.op = red.op(_op_flags(dt))
        pv = _padval_bytes(A)
        if pv is not None:
            req.padval = ctypes.cast(keep.hold(ctypes.create_string_buffer(pv, len(pv))), ctypes.c_void_p)
        req.boundary = A.boundary.code
        req.padding = A.padding.code
        req.steps = int(steps)
        req.result = data_ptr(result)
        req.arrival = float(arrival)
        t = next(self._tags) if tag is None else int(tag)
        req.tag = t
        _lib.check(self._L.sth_server_submit(self._h, ctypes.byref(req)))
        self._keep[t] = (src, result, keep)
        return t

    def start(self):
        _lib.check(self._L.sth_server_start(self._h))

    def drain(self, timeout: float = 600.0):
        _lib.check(self._L.sth_server_drain(self._h, float(timeout)))
        self._collect(0)

    def _collect(self, wait_ms: int):
        buf = (SthCompletion * 256)()
        n = ctypes.c_int32(0)
        while True:
            _lib.check(self._L.sth_server_poll(self._h, buf, 256, ctypes.byref(n), wait_ms))
            for i in range(n.value):
                c = buf[i]
                src, result, _ = self._keep.pop(int(c.tag))
                self._done.append(Completion(c, result))
            if n.value < 256:
                return
            wait_ms = 0

    def completions(self, wait_ms: int = 0) -> list[Completion]:
        """Every completion collected so far (in completion order)."""
        self._collect(wait_ms)
        return list(self._done)

    def stats(self) -> dict:
        st = SthServerStats()
        _lib.check(self._L.sth_server_stats_get(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in SthServerStats._fields_}

    def close(self):
        if self._h:
            _lib.check(self._L.sth_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_latency_summary(done: list[Completion]) -> dict:
    """p50 / p99 / max latency (seconds) and throughput over completions."""
    lat = np.array([c.latency for c in done])
    span = max(c.finished for c in done) - min(c.arrival for c in done)
    cell_steps = sum(c.cells * c.steps for c in done)
    return {"requests": len(done), "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "max_s": float(lat.max()),
            "mean_s": float(lat.mean()), "span_s": float(span),
            "cell_steps_per_s": float(cell_steps / span) if span > 0 else 0.0}
