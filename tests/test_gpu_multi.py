"""Multi-array gather on the GPU (sth_map_multi) against the oracle.

mapstencil(f, A, B, ...) and mapstencil!(f, S::SwitchingStencilArray, B, ...)
(src/gatherstencil.jl:84-113) for `Terms` functions: results bit-identical
to the oracle, Halo arguments' parents refreshed as the oracle refreshes
them, and the reference's own cases (test/array.jl:312-383).
"""
import numpy as np
import pytest

import stencils_hip as sh

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    t = sh.colmajor_empty(a.shape, torch.from_numpy(np.zeros(1, a.dtype)).dtype, torch.device("cuda"))
    rev = tuple(reversed(range(a.ndim)))
    t.copy_(torch.from_numpy(np.ascontiguousarray(a.transpose(rev))).permute(*rev))
    return t


def _host(t):
    return np.asfortranarray(t.cpu().numpy())


def test_reference_cases(gpu, oracle):
    """test/array.jl:312-383 through the GPU path (host arrays: copied in,
    results copied back)."""
    A = np.asfortranarray([[1, 2, 3], [4, 5, 6], [7, 8, 9]], dtype=np.float64)
    B = np.asfortranarray(10 * A)
    r = sh.mapstencil(sh.Terms(sh.center, sh.center), sh.StencilArray(A.copy(order="F"), sh.Moore(1)),
                      sh.StencilArray(B.copy(order="F"), sh.Moore(1)))
    np.testing.assert_array_equal(r, A + B)
    r2 = sh.mapstencil(sh.Terms(sh.sum, sh.sum), sh.StencilArray(A.copy(order="F"), sh.Moore(1)),
                       sh.StencilArray(B.copy(order="F"), sh.Moore(1)))
    assert r2.shape == A.shape
    o2 = oracle.mapstencil_multi(sh.Terms(sh.sum, sh.sum), (sh.StencilArray(A.copy(order="F"), sh.Moore(1)),
                                                            sh.StencilArray(B.copy(order="F"), sh.Moore(1))))
    np.testing.assert_array_equal(r2, o2)
    # SwitchingStencilArray with a read-only StencilArray, three steps
    m = np.asfortranarray([[0, 0, 0], [0, 1, 0], [0, 0, 0]], dtype=np.float64)
    ro = np.asfortranarray([[1, 1, 1], [1, 0, 1], [1, 1, 1]], dtype=np.float64)
    f = sh.Terms(sh.center, sh.Scaled(0.1, sh.sum))
    G = sh.SwitchingStencilArray(m.copy(order="F"), sh.Moore(1))
    O = sh.SwitchingStencilArray(m.copy(order="F"), sh.Moore(1))
    gsa, osa = sh.StencilArray(ro.copy(order="F"), sh.Moore(1)), sh.StencilArray(ro.copy(order="F"), sh.Moore(1))
    for _ in range(3):
        G = sh.mapstencil_(f, G, gsa)
        O = oracle.mapstencil_multi_switching(f, O, (osa,))
    assert G.source[1, 1] > m[1, 1]
    np.testing.assert_array_equal(G.source, O.source)
    np.testing.assert_array_equal(G.dest, O.dest)
    # three StencilArrays
    arrs = [np.full((5, 5), v, dtype=np.float64, order="F") for v in (1.0, 2.0, 3.0)]
    r3 = sh.mapstencil(sh.Terms(sh.center, sh.center, sh.center), *[sh.StencilArray(a, sh.Moore(1)) for a in arrs])
    assert (r3[1:4, 1:4] == 6.0).all()
    # a StencilArray and a plain array (read at I)
    r4 = sh.mapstencil(sh.Terms(sh.center, sh.center), sh.StencilArray(A.copy(order="F"), sh.Moore(1)), B)
    np.testing.assert_array_equal(r4, A + B)


SHAPES = {1: [(37,), (300,)], 2: [(67, 23), (5, 130), (64, 16)], 3: [(13, 9, 7), (33, 8, 5)]}
STENCILS = {1: [sh.VonNeumann(1, 1), sh.Positional((-1,), (0,), (2,)), sh.Moore(2, 1)],
            2: [sh.Window(1), sh.VonNeumann(2), sh.Moore(1), sh.Circle(2), sh.Positional((2, -1), (0, 0), (-1, 2))],
            3: [sh.VonNeumann(1, 3), sh.Moore(1, 3), sh.Positional((1, 0, -1), (0, 0, 0), (0, 2, 0))]}
PADS = [sh.Conditional(), sh.Halo("out"), sh.Halo("in")]


def _random_case(rng, dt, nd, shapes=SHAPES):
    """(Terms, [numpy parents or plain arrays], [per-argument (stencil,
    boundary, padding) or None]) for one random multi-array call."""
    shape = shapes[nd][int(rng.integers(len(shapes[nd])))]
    floats = np.dtype(dt).kind == "f"
    k = int(rng.integers(2, 5))
    terms, specs, arrays = [], [], []
    for i in range(k):
        if floats:
            A = (rng.random(shape) * 2 - 1).astype(dt)
        elif dt == np.uint8:
            A = (rng.random(shape) < 0.4).astype(dt) * rng.integers(1, 4, shape).astype(dt)
        else:
            A = rng.integers(-50, 51, shape).astype(dt)
        A = np.asfortranarray(A)
        if i > 0 and rng.random() < 0.2:
            terms.append(sh.center)
            specs.append(None)  # a plain array, read at I
            arrays.append(A)
            continue
        st = STENCILS[nd][int(rng.integers(len(STENCILS[nd])))]
        if floats:
            pool = [sh.sum, sh.mean, sh.maximum, sh.minimum, sh.center, sh.kernelproduct]
            if st.offsets() == sh.VonNeumann(1, nd).offsets():
                pool.append(sh.Laplacian(0.2))
        elif dt == np.uint8:
            pool = [sh.maximum, sh.minimum, sh.center, sh.LifeRule(), sh.sum]
        else:
            pool = [sh.maximum, sh.minimum, sh.center, sh.sum]
        g = pool[int(rng.integers(len(pool)))]
        if g is sh.kernelproduct:
            st = sh.Kernel(st, rng.standard_normal(len(st)).astype(dt))
        pv = dt(0.25) if floats else dt(3)
        bc = [sh.Remove(pv), sh.Wrap(), sh.Reflect()][int(rng.integers(3))]
        pad = PADS[int(rng.integers(len(PADS)))]
        if floats and rng.random() < 0.35:
            g = sh.Scaled(float(dt(rng.choice([0.5, 0.1, -3.0, 1.5]))), g)
        terms.append(g)
        specs.append((st, bc, pad))
        arrays.append(A)
    combine = "*" if rng.random() < 0.3 else "+"
    return sh.Terms(*terms, combine=combine), arrays, specs


def _args(arrays, specs, device):
    """Device or host arguments.  A Halo("in") argument's parent is the array
    with R cells around it (their values are the halo, overwritten by the
    refresh), so every argument has the same logical size."""
    out = []
    for A, spec in zip(arrays, specs):
        if spec is not None and isinstance(spec[2], sh.Halo) and spec[2].mode == "in":
            R = spec[0].radius()
            A = np.asfortranarray(np.pad(A, R, mode="wrap"))
        data = _dev(A) if device else A.copy(order="F")
        out.append(data if spec is None else sh.StencilArray(data, spec[0], boundary=spec[1], padding=spec[2]))
    return out


@pytest.fixture
def int_wrap():
    old = sh.set_integer_fold_wraps(True)
    yield
    sh.set_integer_fold_wraps(old)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("dt", [np.float32, np.float64, np.uint8, np.int32], ids=lambda d: np.dtype(d).name)
def test_random_multi_against_oracle(gpu, oracle, int_wrap, dt, nd, seed):
    """Random Terms over 2-4 arguments (stencils, boundaries, paddings,
    plain arrays, scaled terms, `+` and `*`): the device result and every
    Halo parent equal the oracle's."""
    rng = np.random.default_rng(1000 * nd + 10 * seed + np.dtype(dt).itemsize)
    f, arrays, specs = _random_case(rng, dt, nd)
    gargs = _args(arrays, specs, device=True)
    oargs = _args(arrays, specs, device=False)
    g = sh.mapstencil(f, *gargs)
    o = oracle.mapstencil_multi(f, tuple(oargs))
    np.testing.assert_array_equal(_host(g), o, err_msg=f"{f} {[s[:2] if s else None for s in specs]}")
    for ga, oa in zip(gargs, oargs):
        if isinstance(ga, sh.StencilArray):
            np.testing.assert_array_equal(_host(ga.parent()), oa.parent())


LARGE = {2: [(1024, 1024), (1100, 960)], 3: [(128, 128, 64)]}


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("dt", [np.float32, np.float64, np.uint8, np.int32], ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("mode", ["fused", "passes"])
def test_random_multi_large_against_oracle(gpu, oracle, int_wrap, dt, nd, seed, mode):
    """Calls of 2^20 cells and more run runtime-compiled streaming kernels:
    one kernel for the whole call when two or more stencil arguments are
    unweighted ("fused": every fold and centre term combined in registers),
    otherwise, or with STH_NO_FUSE ("passes"), one pass per stencil argument
    (the running value read back from dest between passes, centre terms
    folded in on the way).  Random Terms as above: the result and every Halo
    parent equal to the oracle's."""
    rng = np.random.default_rng(5000 + 1000 * nd + 10 * seed + np.dtype(dt).itemsize)
    f, arrays, specs = _random_case(rng, dt, nd, LARGE)
    gargs = _args(arrays, specs, device=True)
    oargs = _args(arrays, specs, device=False)
    if mode == "fused":
        g = sh.mapstencil(f, *gargs)
    else:
        g = _dev(np.zeros(arrays[0].shape, dt))  # `arrays` hold the logical (unpadded) fields
        sh.mapstencil_(f, g, *gargs, flags=sh._abi.NO_FUSE)
    o = oracle.mapstencil_multi(f, tuple(oargs))
    np.testing.assert_array_equal(_host(g), o, err_msg=f"{f} {[s[:2] if s else None for s in specs]}")
    for ga, oa in zip(gargs, oargs):
        if isinstance(ga, sh.StencilArray):
            np.testing.assert_array_equal(_host(ga.parent()), oa.parent())


def test_large_leading_centre_terms(gpu, oracle):
    """Centre terms before the first stencil argument and between two:
    center(A) * 0.5 + sum(B) + center(C) + maximum(D) at 1024^2, Float64."""
    rng = np.random.default_rng(3)
    shape = (1024, 1024)
    arrs = [np.asfortranarray(rng.random(shape)) for _ in range(4)]
    specs = [None, (sh.Moore(1), sh.Wrap(), sh.Conditional()), None, (sh.Circle(2), sh.Reflect(), sh.Halo("out"))]
    f = sh.Terms(sh.Scaled(0.5, sh.center), sh.sum, sh.center, sh.maximum)
    # the first argument is the source, a StencilArray (read at its centre here)
    specs[0] = (sh.Moore(1), sh.Wrap(), sh.Conditional())
    gargs = _args(arrs, specs, device=True)
    oargs = _args(arrs, specs, device=False)
    g = sh.mapstencil(f, *gargs)
    o = oracle.mapstencil_multi(f, tuple(oargs))
    np.testing.assert_array_equal(_host(g), o)


@pytest.mark.parametrize("pad", [sh.Conditional(), sh.Halo("out")], ids=repr)
def test_switching_with_coupled_field(gpu, oracle, pad):
    """Heat with a source: mapstencil!(Terms(Laplacian(0.2), 0.05 * mean),
    S, B) for seven steps on device, both buffers (halos included) against
    the oracle."""
    rng = np.random.default_rng(7)
    U = np.asfortranarray(rng.random((70, 41)).astype(np.float32))
    Q = np.asfortranarray(rng.random((70, 41)).astype(np.float32))
    f = sh.Terms(sh.Laplacian(0.2), sh.Scaled(float(np.float32(0.05)), sh.mean))
    G = sh.SwitchingStencilArray(_dev(U), sh.VonNeumann(1), boundary=sh.Wrap(), padding=pad)
    O = sh.SwitchingStencilArray(U.copy(order="F"), sh.VonNeumann(1), boundary=sh.Wrap(), padding=pad)
    gq = sh.StencilArray(_dev(Q), sh.Window(1), boundary=sh.Reflect(), padding=pad)
    oq = sh.StencilArray(Q.copy(order="F"), sh.Window(1), boundary=sh.Reflect(), padding=pad)
    for _ in range(7):
        G = sh.mapstencil_(f, G, gq)
        O = oracle.mapstencil_multi_switching(f, O, (oq,))
    np.testing.assert_array_equal(_host(G.source), O.source)
    np.testing.assert_array_equal(_host(G.dest), O.dest)
    np.testing.assert_array_equal(_host(gq.parent()), oq.parent())


def test_life_with_mask(gpu, oracle):
    """Game of Life with a mask: LifeRule(A) * center(M) over UInt8, ten
    steps of a SwitchingStencilArray against the oracle."""
    rng = np.random.default_rng(3)
    A = np.asfortranarray((rng.random((96, 50)) < 0.35).astype(np.uint8))
    M = np.asfortranarray((rng.random((96, 50)) < 0.8).astype(np.uint8))
    f = sh.Terms(sh.LifeRule(), sh.center, combine="*")
    G = sh.SwitchingStencilArray(_dev(A), sh.Moore(1), boundary=sh.Wrap(), padding=sh.Halo())
    O = sh.SwitchingStencilArray(A.copy(order="F"), sh.Moore(1), boundary=sh.Wrap(), padding=sh.Halo())
    gm, om = _dev(M), M.copy(order="F")
    for _ in range(10):
        G = sh.mapstencil_(f, G, gm)
        O = oracle.mapstencil_multi_switching(f, O, (om,))
    np.testing.assert_array_equal(_host(G.source), O.source)


def test_multi_errors(gpu):
    """Argument checks: sizes (the reference's ArgumentError), element types,
    scaled terms that would promote, arity, plain-array folds."""
    import torch
    a = _dev(np.zeros((8, 9), np.float32))
    b = _dev(np.zeros((8, 10), np.float32))
    c = _dev(np.zeros((8, 9), np.float64))
    sa = sh.StencilArray(a, sh.Moore(1))
    with pytest.raises(ValueError, match="sizes must match"):
        sh.mapstencil(sh.Terms(sh.center, sh.center), sa, sh.StencilArray(b, sh.Moore(1)))
    with pytest.raises(sh.Unsupported):
        sh.mapstencil(sh.Terms(sh.center, sh.center), sa, sh.StencilArray(c, sh.Moore(1)))
    with pytest.raises(sh.Unsupported):  # 0.1 is no Float32: the reference result would be Float64
        sh.mapstencil(sh.Terms(sh.center, sh.Scaled(0.1, sh.sum)), sa, sh.StencilArray(a, sh.Moore(1)))
    u = _dev(np.zeros((8, 9), np.uint8))
    with pytest.raises(sh.Unsupported):
        sh.mapstencil(sh.Terms(sh.center, sh.Scaled(2.0, sh.maximum)), sh.StencilArray(u, sh.Moore(1)),
                      sh.StencilArray(u, sh.Moore(1)))
    with pytest.raises(TypeError):
        sh.mapstencil(sh.Terms(sh.center, sh.center, sh.center), sa, sa)
    with pytest.raises(TypeError):
        sh.mapstencil(sh.Terms(sh.center, sh.sum), sa, a)
    with pytest.raises(sh.Unsupported):
        sh.mapstencil(sh.Terms(*([sh.center] * 9)), *([sa] * 9))
    del torch
