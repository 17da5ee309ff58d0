"""The five workloads as they actually run, compared with the oracle.

Each test drives the HIP path through the same entry points, layouts, step
schedules and element types as the corresponding benchmark workload and
compares every result with the oracle on the same inputs:

  - mean blur 1000 x 1000 Float64, Window(1), Remove(0.0), Conditional,
    and the same call at 16384 x 16384;
  - Game of Life 8192 x 8192 (full size) and 1024 x 1024 x 1000 generations,
    Wrap, Halo, through the hipGraph schedules the iteration engine uses for
    67 steps and more (replayed 16-unit two-generation graphs, and 32-step
    graphs without temporal blocking) and through plain launches;
  - the 7 x 7 Gaussian (sigma 1.5), Reflect, Halo, in Float32 and in BF16
    storage, at 16384 x 16384 and on a ragged size;
  - 3-D heat (VonNeumann(1,3) Laplacian(0.15), Wrap) in BF16 storage: one
    step, iterations (two steps per launch) and served requests;
  - the batched workload's job mix (the batch tests cover it).

Integer and float results must be bit-identical; BF16 storage results must
be bit-identical to the oracle's f32-accumulate restatement of the same
arithmetic.
"""
import numpy as np
import pytest

import stencils_hip as sh

pytestmark = pytest.mark.gpu


def _threads(oracle):
    return max(1, min(16, oracle.host_cores()))


def _dev(torch):
    return torch.device("cuda", torch.cuda.current_device())


def _to_device(sh, torch, A, tdt=None):
    t = sh.colmajor_empty(A.shape, tdt or torch.from_numpy(A[:1].copy()).dtype, _dev(torch))
    src = torch.from_numpy(np.ascontiguousarray(A)).to(_dev(torch))
    t.copy_(src if tdt is None else src.to(tdt))
    return t


def _raw16(torch, t):
    """Raw 16-bit patterns of a device tensor, as a host uint16 array in the
    same (Julia) index order, column-major."""
    return np.asfortranarray(t.cpu().view(torch.int16).numpy().view(np.uint16))


# --------------------------------------------------------------- mean blur

def test_mean_blur_1000_f64(gpu, oracle):
    """mapstencil(mean, StencilArray(A, Window(1); Remove(0.0), Conditional))
    on the README's 1000 x 1000 Float64 input (PCG64 seed 1), device
    resident.  The stated bound is 1e-14 relative; the result is bit-exact."""
    import torch
    A = np.random.Generator(np.random.PCG64(1)).random((1000, 1000))
    sa = sh.StencilArray(_to_device(sh, torch, A), sh.Window(1), boundary=sh.Remove(0.0))
    g = sh.mapstencil(sh.mean, sa).cpu().numpy()
    o = oracle.mapstencil(sh.mean, sh.StencilArray(np.asfortranarray(A), sh.Window(1), boundary=sh.Remove(0.0)),
                          nthreads=_threads(oracle))
    np.testing.assert_allclose(g, o, rtol=1e-14, atol=0)
    np.testing.assert_array_equal(g, o)


def test_mean_blur_16384_f64(gpu, oracle):
    """The same call on 16384 x 16384 Float64 (PCG64 seed 11; 2 GiB, past
    32-bit byte offsets), the size the single-array bandwidth figure is
    quoted at: bit-exact against the oracle."""
    import torch
    A = np.random.Generator(np.random.PCG64(11)).random((16384, 16384))
    sa = sh.StencilArray(_to_device(sh, torch, A), sh.Window(1), boundary=sh.Remove(0.0))
    g = sh.mapstencil(sh.mean, sa).cpu().numpy()
    del sa
    o = oracle.mapstencil(sh.mean, sh.StencilArray(np.asfortranarray(A), sh.Window(1), boundary=sh.Remove(0.0)),
                          nthreads=_threads(oracle))
    np.testing.assert_array_equal(g, o)


# ------------------------------------------------------------ Game of Life

_LIFE_ORACLE = {}


def _life_oracle(oracle, n, steps):
    key = (n, steps)
    if key not in _LIFE_ORACLE:
        _LIFE_ORACLE.clear()  # keep at most one full-size pair in host memory
        A = (np.random.Generator(np.random.PCG64(2)).random((n, n)) < 0.3).astype(np.uint8)
        O = sh.SwitchingStencilArray(np.asfortranarray(A), sh.Moore(1), boundary=sh.Wrap(), padding=sh.Halo())
        O = oracle.mapstencil_steps(sh.LifeRule(), O, steps, nthreads=_threads(oracle))
        _LIFE_ORACLE[key] = (A, O.source, O.dest)
    return _LIFE_ORACLE[key]


@pytest.mark.parametrize("n,steps,graph,fuse", [
    (1024, 1000, True, True),     # 15 replays of the 16-unit fused graph + remainder graph
    (1024, 1000, True, False),    # 31 replays of the 32-step graph + remainder graph
    (1024, 1000, False, True),    # the same fused schedule as plain launches
    (8192, 70, True, True),       # full size: one 16-unit replay + a tail of 5 single steps
    (8192, 70, True, False),      # full size: two 32-step replays + 5
], ids=lambda v: str(v))
def test_life_schedules(gpu, oracle, n, steps, graph, fuse):
    """Game of Life B3/S23, Wrap, Halo, device-resident SwitchingStencilArray
    through sth_iterate; both parents (halos included) must equal the
    step-by-step oracle bit for bit."""
    import torch
    A, osrc, odst = _life_oracle(oracle, n, steps)
    t = _to_device(sh, torch, A)
    G = sh.SwitchingStencilArray(t, sh.Moore(1), boundary=sh.Wrap(), padding=sh.Halo())
    flags = 0 if fuse else sh._abi.NO_FUSE
    G = sh.mapstencil_steps(sh.LifeRule(), G, steps, graph=graph, flags=flags)
    np.testing.assert_array_equal(G.source.cpu().numpy(), osrc)
    np.testing.assert_array_equal(G.dest.cpu().numpy(), odst)


def test_life_graph_replays_are_reused(gpu, oracle):
    """Two 1000-generation calls on the same buffers replay the cached graphs
    (the second call captures nothing) and keep matching the oracle."""
    import torch
    A, _, _ = _life_oracle(oracle, 1024, 1000)
    t = _to_device(sh, torch, A)
    G = sh.SwitchingStencilArray(t, sh.Moore(1), boundary=sh.Wrap(), padding=sh.Halo())
    G = sh.mapstencil_steps(sh.LifeRule(), G, 1000, graph=True)
    G = sh.mapstencil_steps(sh.LifeRule(), G, 1000, graph=True)
    O = sh.SwitchingStencilArray(np.asfortranarray(A), sh.Moore(1), boundary=sh.Wrap(), padding=sh.Halo())
    O = oracle.mapstencil_steps(sh.LifeRule(), O, 2000, nthreads=_threads(oracle))
    np.testing.assert_array_equal(G.source.cpu().numpy(), O.source)
    np.testing.assert_array_equal(G.dest.cpu().numpy(), O.dest)


# ------------------------------------------------------------ Gaussian 7x7

def _gaussian_weights(sigma=1.5, R=3):
    a = np.arange(-R, R + 1)
    W = np.exp(-(a[:, None] ** 2 + a[None, :] ** 2) / (2 * sigma ** 2))
    return (W / W.sum()).astype(np.float32)


@pytest.mark.parametrize("shape", [(16384, 16384), (1029, 771)], ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("tname", ["float32", "bfloat16"])
def test_gaussian_reflect_halo(gpu, oracle, tname, shape):
    """Kernel(Window(3), W) with the 7 x 7 Gaussian, Reflect, Halo{:out}
    (the halo refresh runs through the 16-bit halo kernel for BF16 and the
    streaming kernel reads it).  Float32 is bit-exact against the oracle;
    BF16 storage is bit-exact against the oracle's f32-accumulate
    restatement, whose error against FP64 is bounded below."""
    import torch
    tdt = getattr(torch, tname)
    A = np.random.Generator(np.random.PCG64(3)).random(shape, dtype=np.float32)
    W = _gaussian_weights()
    K = sh.Kernel(sh.Window(3), W)
    sa = sh.StencilArray(_to_device(sh, torch, A, tdt), K, boundary=sh.Reflect(), padding=sh.Halo())
    g = sh.mapstencil(sh.kernelproduct, sa)
    assert g.dtype == tdt
    nt = _threads(oracle)
    if tname == "float32":
        o = oracle.mapstencil(sh.kernelproduct, sh.StencilArray(np.asfortranarray(A), K, boundary=sh.Reflect(),
                                                                padding=sh.Halo()), nthreads=nt)
        np.testing.assert_array_equal(g.cpu().numpy(), o)
        return
    parent = _raw16(torch, sa.parent())  # the interior as uploaded; halo refreshed by the call
    o = oracle.map_raw(parent, sh._abi.BF16, shape, 3, sh.Window(3).offsets(), 2, sh._abi.REFLECT, None,
                       sh._abi.HALO_OUT, sh.kernelproduct.op(sh._abi.OPF_ACC_F32), W, sh._abi.F32, nthreads=nt)
    gr = _raw16(torch, g)
    bad = np.argwhere(gr != o)
    assert bad.size == 0, f"{len(bad)} cells differ, first {bad[:4].tolist()}"
    # the refreshed halo equals the oracle's update_boundary!
    np.testing.assert_array_equal(_raw16(torch, sa.parent()), parent)
    # error of the BF16 result against FP64, on a crop (values in [0, 1))
    cx, cy = slice(0, min(shape[0], 600)), slice(0, min(shape[1], 600))
    crop = np.asfortranarray(A[: cx.stop + 3, : cy.stop + 3].astype(np.float64))
    f64 = oracle.mapstencil(sh.kernelproduct, sh.StencilArray(crop, sh.Kernel(sh.Window(3), W.astype(np.float64)),
                                                              boundary=sh.Reflect()))
    got = g[cx, cy].float().cpu().numpy().astype(np.float64)
    inner = (slice(0, cx.stop - (3 if cx.stop < shape[0] else 0)), slice(0, cy.stop - (3 if cy.stop < shape[1] else 0)))
    err = np.abs(got[inner] - f64[: cx.stop, : cy.stop][inner])
    assert err.max() < 8e-3, err.max()   # input rounding to bf16 (2^-9) plus one output rounding


# ------------------------------------------------------- 3-D heat, BF16

def _bf16_cube(torch, shape, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    t = sh.colmajor_empty(shape, torch.bfloat16, _dev(torch))
    t.copy_(torch.rand(shape, generator=g, device="cuda").to(torch.bfloat16))
    return t


def test_heat3d_bf16_single_step(gpu, oracle):
    """One step of Laplacian(0.15) over VonNeumann(1,3), Wrap, BF16 storage:
    the streaming body, bit-exact against the oracle's restatement."""
    import torch
    lap = sh.Laplacian(0.15)
    for shape in ((128, 128, 128), (67, 45, 31)):
        t = _bf16_cube(torch, shape, sum(shape))
        g = sh.mapstencil(lap, sh.StencilArray(t, sh.VonNeumann(1, 3), boundary=sh.Wrap()))
        o = oracle.map_raw(_raw16(torch, t), sh._abi.BF16, shape, 0, sh.VonNeumann(1, 3).offsets(), 3,
                           sh._abi.WRAP, None, sh._abi.CONDITIONAL, lap.op(sh._abi.OPF_ACC_F32),
                           nthreads=_threads(oracle))
        np.testing.assert_array_equal(_raw16(torch, g), o, err_msg=str(shape))


@pytest.mark.parametrize("steps", [9, 70, 141])
@pytest.mark.parametrize("shape", [(128, 128, 128), (70, 41, 33)], ids=lambda s: "x".join(map(str, s)))
def test_heat3d_bf16_iterate(gpu, oracle, shape, steps):
    """mapstencil! iterations of the BF16 heat step through sth_iterate
    (two steps per launch in the steady state, hipGraph replays) against
    the oracle's step-by-step restatement, both buffers compared."""
    import torch
    lap = sh.Laplacian(0.15)
    t = _bf16_cube(torch, shape, 7 * steps)
    a0 = _raw16(torch, t)
    G = sh.SwitchingStencilArray(t, sh.VonNeumann(1, 3), boundary=sh.Wrap())
    G = sh.mapstencil_steps(lap, G, steps, graph=True)
    a, b = a0.copy(order="F"), a0.copy(order="F")
    cur = oracle.iterate_raw(a, b, sh._abi.BF16, shape, 0, sh.VonNeumann(1, 3).offsets(), 3, sh._abi.WRAP,
                             None, sh._abi.CONDITIONAL, lap.op(sh._abi.OPF_ACC_F32), steps,
                             nthreads=_threads(oracle))
    new, old = (b, a) if cur == 1 else (a, b)
    np.testing.assert_array_equal(_raw16(torch, G.source), new)
    np.testing.assert_array_equal(_raw16(torch, G.dest), old)


def test_heat3d_bf16_served(gpu, oracle):
    """BF16 heat requests through the continuous-batching server (a mix of
    small batched requests and one 128^3 request on its own two-step
    launches): every finished state equals the oracle's step-by-step
    restatement of the same request."""
    import torch
    lap = sh.Laplacian(0.15)
    rng = np.random.default_rng(55)
    reqs = []
    for i, (side, steps) in enumerate(((32, 10), (48, 37), (128, 64), (40, 101), (33, 12))):
        t = _bf16_cube(torch, (side,) * 3, 100 + i)
        reqs.append((sh.SwitchingStencilArray(t, sh.VonNeumann(1, 3), boundary=sh.Wrap()), steps,
                     float(rng.uniform(0, 0.01))))
    srv = sh.Server(hbm_budget=1 << 30, round_steps=8)
    tags = {}
    for S, steps, arr in reqs:
        tags[srv.submit(lap, S, steps, arrival=arr)] = (S, steps)
    torch.cuda.synchronize()
    srv.start()
    srv.drain(timeout=300)
    done = srv.completions()
    srv.close()
    assert len(done) == len(reqs) and all(c.status == 0 for c in done)
    for c in done:
        S, steps = tags[c.tag]
        a = _raw16(torch, S.source)
        b = a.copy(order="F")
        cur = oracle.iterate_raw(a, b, sh._abi.BF16, S.size(), 0, sh.VonNeumann(1, 3).offsets(), 3,
                                 sh._abi.WRAP, None, sh._abi.CONDITIONAL, lap.op(sh._abi.OPF_ACC_F32), steps,
                                 nthreads=_threads(oracle))
        exp = b if cur == 1 else a
        np.testing.assert_array_equal(_raw16(torch, c.result), exp, err_msg=f"side {S.size()[0]} steps {steps}")


# ------------------------------------------------------ hood (keyword) form

@pytest.mark.parametrize("pad", ["cond", "halo_out", "halo_in"])
@pytest.mark.parametrize("bname", ["remove", "wrap", "reflect"])
def test_hood_form(gpu, oracle, bname, pad):
    """gatherstencil(f, hood, A; boundary, padding) (src/gatherstencil.jl:
    22-39): the dest has A's size, shrunk by 2R per dim for Halo{:in}; host
    and device parents give the oracle's result."""
    import torch
    A = np.random.default_rng(60).random((57, 41))
    bc = {"remove": sh.Remove(0.5), "wrap": sh.Wrap(), "reflect": sh.Reflect()}[bname]
    padding = {"cond": sh.Conditional(), "halo_out": sh.Halo("out"), "halo_in": sh.Halo("in")}[pad]
    for hood, f in ((sh.Window(2), sh.mean), (sh.VonNeumann(1), sh.maximum),
                    (sh.Kernel(sh.Circle(2), np.arange(len(sh.Circle(2)), dtype=np.float64)), sh.kernelproduct)):
        o = oracle.mapstencil(f, sh.StencilArray(A.copy(order="F"), hood, boundary=bc, padding=padding))
        g_host = sh.mapstencil(f, hood, A.copy(order="F"), boundary=bc, padding=padding)
        assert isinstance(g_host, np.ndarray)
        exp_shape = tuple(s - 2 * hood.radius() for s in A.shape) if pad == "halo_in" else A.shape
        assert g_host.shape == exp_shape == o.shape
        np.testing.assert_array_equal(g_host, o)
        g_dev = sh.mapstencil(f, hood, _to_device(sh, torch, A), boundary=bc, padding=padding)
        assert g_dev.is_cuda
        np.testing.assert_array_equal(g_dev.cpu().numpy(), o)
