"""Batched launches (sth_map_batch): every result equals the one-call-per-job
path bit for bit, and the oracle."""
import numpy as np
import pytest

import stencils_hip as sh

pytestmark = pytest.mark.gpu

SHAPES = [sh.Window(1), sh.Circle(3), sh.VonNeumann(2), sh.Moore(2), sh.Window(3),
          sh.Positional((2, -2), (0, 0), (-2, 1))]  # the last runs outside the batch
OPS = [sh.sum, sh.maximum, sh.mean, sh.minimum]


def _jobs(seed, n, dtypes=(np.float32, np.float64), pads=("cond",)):
    import torch
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda")
    jobs = []
    for i in range(n):
        s0 = int(rng.integers(8, 300))
        s1 = int(rng.integers(8, 300))
        dt = dtypes[i % len(dtypes)]
        st = SHAPES[int(rng.integers(len(SHAPES)))]
        f = OPS[int(rng.integers(len(OPS)))]
        bc = [sh.Remove(np.dtype(dt).type(0)), sh.Wrap(), sh.Reflect(), sh.Remove(np.dtype(dt).type(2))][i % 4]
        pad = pads[i % len(pads)]
        A = (rng.random((s0, s1)) * 2 - 1).astype(dt)
        t = sh.colmajor_empty(A.shape, torch.float32 if dt == np.float32 else torch.float64, dev)
        t.copy_(torch.from_numpy(np.ascontiguousarray(A)).to(dev))
        padding = sh.Halo() if pad == "halo" else sh.Conditional()
        sa = sh.StencilArray(t, st, boundary=bc, padding=padding)
        dest = sh.colmajor_empty(A.shape, t.dtype, dev)
        jobs.append((f, dest, sa, A, st, bc, padding))
    return jobs


@pytest.mark.parametrize("serial", [False, True])
def test_batch_equals_single_calls(gpu, serial):
    jobs = _jobs(1, 40)
    sh.mapstencil_batch([(f, d, sa) for f, d, sa, *_ in jobs], serial=serial)
    for f, d, sa, *_ in jobs:
        ref = sh.mapstencil(f, sa)
        assert np.array_equal(d.cpu().numpy(), ref.cpu().numpy()), f"{sa!r} {f}"


def test_batch_matches_oracle(gpu, oracle):
    jobs = _jobs(2, 24, pads=("cond", "halo"))
    sh.mapstencil_batch([(f, d, sa) for f, d, sa, *_ in jobs])
    for f, d, sa, A, st, bc, padding in jobs:
        o = oracle.mapstencil(f, sh.StencilArray(np.asfortranarray(A), st, boundary=bc, padding=padding))
        assert np.array_equal(d.cpu().numpy(), o), f"{st} {bc} {padding} {f}"


def test_batch_2d_stencil_on_3d_array(gpu, oracle):
    """Trailing dims pass through (src/stencil.jl:93-95) inside a batch too."""
    import torch
    dev = torch.device("cuda")
    rng = np.random.default_rng(3)
    A3 = rng.random((20, 17, 9))
    t3 = sh.colmajor_empty(A3.shape, torch.float64, dev)
    t3.copy_(torch.from_numpy(np.ascontiguousarray(A3)).to(dev))
    sa3 = sh.StencilArray(t3, sh.Window(1), boundary=sh.Wrap())
    d3 = sh.colmajor_empty(A3.shape, torch.float64, dev)
    sh.mapstencil_batch([(sh.mean, d3, sa3)])
    o = oracle.mapstencil(sh.mean, sh.StencilArray(np.asfortranarray(A3), sh.Window(1), boundary=sh.Wrap()))
    assert np.array_equal(d3.cpu().numpy(), o)
