"""sth_map_batch_static (the batched entry point that allocates nothing)
against sth_map_batch, through the C ABI: two calls in a row, the second
with a smaller batch, each bit-identical to the allocating entry point on
the same jobs; oversize batches are rejected before any work."""
import ctypes

import numpy as np
import pytest

import stencils_hip as sh
from stencils_hip import _abi, _lib

pytestmark = pytest.mark.gpu


def _jobs(torch, n, max_side, seed):
    """A mixed batch on the device: batched catalogue shapes and ops, one job
    large enough to run as its own launch, a Halo job (halo refresh first),
    an array too narrow for a full wave, and a runtime-offset shape that
    runs alone."""
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda", torch.cuda.current_device())
    shapes = [sh.Window(1), sh.Circle(3), sh.VonNeumann(2), sh.Moore(1)]
    fs = [sh.sum, sh.maximum, sh.mean, sh.minimum]
    calls = []
    for i in range(n):
        s0 = int(rng.integers(3, max_side))
        s1 = int(rng.integers(8, max_side))
        st = shapes[i % len(shapes)]
        f = fs[(i // len(shapes)) % len(fs)]
        pad = sh.Halo() if i % 7 == 3 else sh.Conditional()
        bc = sh.Reflect() if i % 7 == 3 else sh.Remove(np.float32(0.25))
        if min(s0, s1) <= st.radius():
            s0, s1 = s0 + st.radius() + 1, s1 + st.radius() + 1
        t = sh.colmajor_empty((s0, s1), torch.float32, dev)
        t.copy_(torch.rand((s1, s0), device=dev).T)
        sa = sh.StencilArray(t, st, boundary=bc, padding=pad)
        calls.append((f, sh.colmajor_empty((s0, s1), torch.float32, dev), sa))
    big = sh.colmajor_empty((4096, 2048), torch.float32, dev)
    big.copy_(torch.rand((2048, 4096), device=dev).T)
    calls.append((sh.mean, sh.colmajor_empty((4096, 2048), torch.float32, dev),
                  sh.StencilArray(big, sh.Window(1), boundary=sh.Wrap())))
    odd = sh.colmajor_empty((37, 29), torch.float32, dev)
    odd.copy_(torch.rand((29, 37), device=dev).T)
    calls.append((sh.sum, sh.colmajor_empty((37, 29), torch.float32, dev),
                  sh.StencilArray(odd, sh.Positional((2, -1), (0, 0), (-1, 1)), boundary=sh.Wrap())))
    return calls


def _run_both(torch, calls):
    """Results of the allocating and of the allocation-free entry point on
    the same jobs (dests cleared in between)."""
    plan = sh.BatchPlan(calls)
    plan.run()
    ref = [d.clone() for _, d, _ in calls]
    for _, d, _ in calls:
        d.fill_(float("nan"))
    torch.cuda.synchronize()
    L = _lib.lib()
    strm = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(L.sth_map_batch_static(_lib.context(plan.dev), plan.n, plan.jobs, 0, strm))
    got = [d.clone() for _, d, _ in calls]
    return ref, got


def test_static_batch_matches_allocating_batch(gpu):
    import torch
    first = _jobs(torch, 120, 400, 1)
    ref, got = _run_both(torch, first)
    for i, (r, g) in enumerate(zip(ref, got)):
        assert torch.equal(r, g), f"job {i} of the first call differs"
    # a second call, smaller, right after: the fixed table keeps the first
    # call's larger contents beyond the new jobs, which must not matter
    second = _jobs(torch, 9, 90, 2)
    ref, got = _run_both(torch, second)
    for i, (r, g) in enumerate(zip(ref, got)):
        assert torch.equal(r, g), f"job {i} of the second call differs"


def test_static_batch_rejects_more_jobs_than_its_table(gpu):
    import torch
    L = _lib.lib()
    n = _abi.STATIC_BATCH_MAX_JOBS + 1
    jobs = (_abi.SthJob * n)()
    strm = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = L.sth_map_batch_static(_lib.context(torch.cuda.current_device()), n, jobs, 0, strm)
    assert rc == _abi.STH_EINVAL
    assert b"STH_STATIC_BATCH_MAX_JOBS" in L.sth_last_error()
