"""HIP path vs the oracle, through the C ABI (libstencils_hip.so).

Integer, Bool and Float64/Float32 results must be bit-identical to the oracle
(the reference's fold order, no contraction).  BF16/F16 storage results must
be bit-identical to the oracle's f32-accumulate restatement; their error
against the FP64 oracle is checked separately.
"""
import numpy as np
import pytest

import stencils_hip as sh
from conftest import GOLDEN
from _spec import F, make_boundary, make_input, make_padding, make_stencil

pytestmark = pytest.mark.gpu


def _np_input(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    if dtype == np.bool_:
        return np.asfortranarray(rng.random(shape) < 0.4)
    if np.dtype(dtype).kind in "iu":
        info = np.iinfo(dtype)
        lo, hi = max(info.min, -50), min(info.max, 50)
        return np.asfortranarray(rng.integers(lo, hi + 1, shape).astype(dtype))
    return np.asfortranarray((rng.random(shape) * 2 - 1).astype(dtype))


def _both(f, A, st, bc, pad, oracle):
    """(GPU result, oracle result, GPU source parent, oracle source parent)."""
    ga = sh.StencilArray(A.copy(order="F"), st, boundary=bc, padding=pad)
    oa = sh.StencilArray(A.copy(order="F"), st, boundary=bc, padding=pad)
    g = sh.mapstencil(f, ga)
    o = oracle.mapstencil(f, oa)
    return g, o, ga.parent(), oa.parent()


# ------------------------------------------------------------ goldens

def _cases():
    for c in GOLDEN["cases"]:
        for pad in c["paddings"]:
            yield pytest.param(c, pad, id=f"{c['name']}-{pad}")


@pytest.mark.parametrize("case,pad", list(_cases()))
def test_golden_cases_on_gpu(gpu, case, pad):
    A = make_input(case)
    st = make_stencil(case["stencil"])
    sa = sh.StencilArray(A.copy(order="F"), st, boundary=make_boundary(case["boundary"], A.dtype),
                         padding=make_padding(pad))
    out = sh.mapstencil(F[case["f"]], sa)
    exp = np.asarray(case["expected"])
    if case.get("interior"):
        out = out[1:-1, 1:-1]
    if case["tol"] is None:
        np.testing.assert_array_equal(out, exp)
    else:
        np.testing.assert_allclose(out, exp, rtol=0, atol=case["tol"])


# ------------------------------------------------- random parity sweep

SHAPES_2D = [sh.Window(1), sh.Window(2), sh.Window(3), sh.Moore(1), sh.Moore(2),
             sh.VonNeumann(1), sh.VonNeumann(2), sh.VonNeumann(3), sh.Circle(1), sh.Circle(2),
             sh.Circle(3), sh.Positional((2, -1), (0, 0), (-1, 3), (1, 1)), sh.Cross(2),
             sh.Annulus(3), sh.Rectangle((-2, 1), (0, 2)), sh.Ordinal(2)]
BOUNDS = {"remove": lambda dt: sh.Remove(np.dtype(dt).type(0) if np.dtype(dt) != np.bool_ else np.bool_(False)),
          "remove7": lambda dt: sh.Remove(np.dtype(dt).type(7) if np.dtype(dt).kind in "iuf" else np.bool_(True)),
          "wrap": lambda dt: sh.Wrap(), "reflect": lambda dt: sh.Reflect()}
PADS = {"cond": sh.Conditional(), "halo_out": sh.Halo("out"), "halo_in": sh.Halo("in")}


@pytest.mark.parametrize("pad", list(PADS))
@pytest.mark.parametrize("bname", list(BOUNDS))
@pytest.mark.parametrize("st", SHAPES_2D, ids=repr)
def test_parity_f64_2d(gpu, oracle, st, bname, pad):
    A = _np_input((70, 45), np.float64, 1)
    bc = BOUNDS[bname](np.float64)
    for f in (sh.sum, sh.mean, sh.minimum, sh.maximum):
        g, o, gp, op_ = _both(f, A, st, bc, PADS[pad], oracle)
        np.testing.assert_array_equal(g, o, err_msg=f"{f}")
        np.testing.assert_array_equal(gp, op_)  # Halo refresh side effect
    w = np.random.default_rng(4).standard_normal(len(st))
    K = sh.Kernel(st, w)
    g, o, _, _ = _both(sh.kernelproduct, A, K, bc, PADS[pad], oracle)
    np.testing.assert_array_equal(g, o)


@pytest.mark.parametrize("bname", ["remove", "wrap", "reflect"])
@pytest.mark.parametrize("st", [sh.Window(1), sh.Circle(3), sh.VonNeumann(2), sh.Window(3),
                                sh.Positional((3, -2), (0, 0), (-3, 1))], ids=repr)
def test_parity_f32_2d(gpu, oracle, st, bname):
    A = _np_input((131, 67), np.float32, 2)
    bc = BOUNDS[bname](np.float32)
    for pad in ("cond", "halo_out"):
        for f in (sh.sum, sh.mean, sh.minimum, sh.maximum):
            g, o, _, _ = _both(f, A, st, bc, PADS[pad], oracle)
            np.testing.assert_array_equal(g, o)
        K = sh.Kernel(st, np.random.default_rng(5).standard_normal(len(st)).astype(np.float32))
        g, o, _, _ = _both(sh.kernelproduct, A, K, bc, PADS[pad], oracle)
        np.testing.assert_array_equal(g, o)


REMAINING_SHAPES = [sh.Diamond(2), sh.AngledCross(2), sh.ForwardSlash(2), sh.BackSlash(3), sh.Vertical(2),
                    sh.Horizontal(3), sh.Cardinal(2), sh.Cross(3), sh.Diamond(1, 3), sh.Vertical(1, 3),
                    sh.Horizontal(1, 3)]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("st", REMAINING_SHAPES, ids=repr)
def test_parity_remaining_shapes(gpu, oracle, st, dt):
    """The shapes outside the build-time catalogue (runtime offsets) and
    distance-weighted kernels Kernel(f, stencil) (src/stencils/kernel.jl:
    98-109), every boundary and padding, bit-identical to the oracle."""
    shape = (29, 23) if st.N == 2 else (13, 11, 9)
    A = _np_input(shape, dt, 31)
    K = sh.Kernel(lambda d: dt(1.0 / (1.0 + d)), st)
    for bname in BOUNDS:
        bc = BOUNDS[bname](dt)
        for pad in PADS.values():
            for f in (sh.sum, sh.mean, sh.minimum, sh.maximum):
                g, o, gp, op_ = _both(f, A, st, bc, pad, oracle)
                np.testing.assert_array_equal(g, o, err_msg=f"{bname} {pad} {f}")
                np.testing.assert_array_equal(gp, op_)
            g, o, _, _ = _both(sh.kernelproduct, A, K, bc, pad, oracle)
            np.testing.assert_array_equal(g, o, err_msg=f"{bname} {pad} kernelproduct")


@pytest.fixture
def int_wrap():
    """Integer sum / mean with the element type's wrapping fold stated by the
    caller (the library declines them otherwise)."""
    old = sh.set_integer_fold_wraps(True)
    yield
    sh.set_integer_fold_wraps(old)


@pytest.mark.parametrize("dt", [np.bool_, np.uint8, np.int8, np.uint16, np.int16, np.uint32,
                                np.int32, np.uint64, np.int64], ids=lambda d: np.dtype(d).name)
def test_parity_integers(gpu, oracle, int_wrap, dt):
    A = _np_input((53, 38), dt, 3)
    for st in (sh.Window(1), sh.Moore(2), sh.Circle(2), sh.Positional((0, 1), (2, -1))):
        for bname in ("remove7", "wrap", "reflect"):
            bc = BOUNDS[bname](dt)
            for pad in ("cond", "halo_out"):
                ops = [sh.sum, sh.mean, sh.minimum, sh.maximum, sh.LifeRule(), sh.LifeRule((1, 3), (0, 5))]
                for f in ops:
                    g, o, gp, op_ = _both(f, A, st, bc, PADS[pad], oracle)
                    np.testing.assert_array_equal(g, o, err_msg=f"{st} {bname} {pad} {f}")
                    np.testing.assert_array_equal(gp, op_)
                if dt != np.bool_:
                    K = sh.Kernel(st, np.arange(1, len(st) + 1).astype(dt))
                    g, o, _, _ = _both(sh.kernelproduct, A, K, bc, PADS[pad], oracle)
                    np.testing.assert_array_equal(g, o)


@pytest.mark.parametrize("dt", [np.uint8, np.int8, np.int16, np.uint32, np.int32, np.int64, np.uint64])
def test_unpinned_integer_folds_are_declined(gpu, dt):
    """Without the caller's statement of how its `sum` folds, the library
    itself (not only the Python front end) declines integer sums over
    sub-64-bit types and integer means; Bool, 64-bit sums, min/max and
    one-offset stencils stay accelerated."""
    assert not sh.integer_fold_wraps()
    A = _np_input((20, 17), dt, 30)
    big = np.dtype(dt).itemsize == 8
    for f in (sh.sum, sh.mean):
        sa = sh.StencilArray(A.copy(order="F"), sh.Moore(2), boundary=sh.Wrap())
        out = np.zeros(A.shape, dtype=np.float64 if f is sh.mean else dt, order="F")
        if f is sh.sum and big:
            assert sh.result_dtype(f, sa) == sh._abi.NUMPY_TO_STH[np.dtype(dt)]
            sh.mapstencil_(f, out, sa)  # 64-bit sums wrap under either reading
            continue
        with pytest.raises(sh.Unsupported):
            sh.result_dtype(f, sa)  # the front end's check
        with pytest.raises(sh.Unsupported, match="unpinned"):
            sh.mapstencil_(f, out, sa)  # straight to the library, no front-end check
    one = sh.StencilArray(A.copy(order="F"), sh.Positional((1, 0)), boundary=sh.Wrap())
    assert np.array_equal(sh.mapstencil(sh.sum, one), np.roll(A, -1, axis=0))
    np.testing.assert_array_equal(sh.mapstencil(sh.mean, one), np.roll(A, -1, axis=0).astype(np.float64))
    b = sh.StencilArray((A % 2 == 0).copy(order="F"), sh.Moore(2), boundary=sh.Wrap())
    assert sh.mapstencil(sh.sum, b).dtype == np.int64
    assert sh.mapstencil(sh.maximum, sh.StencilArray(A.copy(order="F"), sh.Moore(2))).dtype == dt


@pytest.mark.parametrize("n", [1, 2, 5, 17, 300, 1000])
def test_parity_1d(gpu, oracle, n):
    A = _np_input((n,), np.float64, 6)
    for st in (sh.Window(1, 1), sh.Moore(1, 1), sh.Positional((-1,), (0,), (2,))):
        if max(st.reach()) >= n:
            continue
        for bname in ("remove7", "wrap", "reflect"):
            for pad in ("cond", "halo_out", "halo_in"):
                logical = n - 2 * st.radius() if pad == "halo_in" else n
                if logical < 1 or (pad != "cond" and bname != "remove7" and st.radius() >= logical):
                    continue  # the reference throws a BoundsError here
                g, o, gp, op_ = _both(sh.mean, A, st, BOUNDS[bname](np.float64), PADS[pad], oracle)
                np.testing.assert_array_equal(g, o)
                np.testing.assert_array_equal(gp, op_)


@pytest.mark.parametrize("st", [sh.VonNeumann(1, 3), sh.Moore(1, 3), sh.Circle(2, 3),
                                sh.Window(1, 2), sh.Positional((2, 0, -1), (0, 0, 0), (0, -2, 1))],
                         ids=repr)
def test_parity_3d(gpu, oracle, st):
    A = _np_input((21, 18, 17), np.float64, 7)
    for bname in ("remove7", "wrap", "reflect"):
        for pad in ("cond", "halo_out"):
            for f in (sh.sum, sh.mean, sh.maximum):
                g, o, _, _ = _both(f, A, st, BOUNDS[bname](np.float64), PADS[pad], oracle)
                np.testing.assert_array_equal(g, o)
            if st.N == 3 and st.offsets() == sh.VonNeumann(1, 3).offsets():
                g, o, _, _ = _both(sh.Laplacian(0.15), A, st, BOUNDS[bname](np.float64), PADS[pad], oracle)
                np.testing.assert_array_equal(g, o)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_parity_wide_aligned_rows(gpu, oracle, dt):
    """Device rows that start on vector boundaries (the library's pitched
    layout) take the three-vector interior path in every compiled variant;
    results must equal the oracle bit for bit."""
    from stencils_hip import _lib
    A = _np_input((1100, 37), dt, 21)
    try:
        for pat in ("x1", "x2", "x4", "x8", ""):
            _lib.set_variant(pat)
            for st in (sh.Window(1), sh.Circle(3), sh.VonNeumann(2)):
                for bname, pad in (("remove7", "cond"), ("reflect", "halo_out"), ("wrap", "halo_in")):
                    for f in (sh.sum, sh.maximum):
                        g, o, _, _ = _both(f, A, st, BOUNDS[bname](dt), PADS[pad], oracle)
                        np.testing.assert_array_equal(g, o, err_msg=f"{pat} {st} {bname} {pad} {f}")
    finally:
        _lib.set_variant("")


@pytest.mark.parametrize("variant", ["", "stream4x4"])
@pytest.mark.parametrize("shape", [(96, 77), (1100, 23)])
@pytest.mark.parametrize("tname", ["bfloat16", "float16"])
def test_parity_storage_types(gpu, oracle, tname, shape, variant):
    """BF16/F16 storage, f32 accumulation, one rounding at the store: bit-exact
    against the oracle's restatement of the same arithmetic (the wide shape's
    interior waves take the aligned three-vector path)."""
    import torch
    tdt = getattr(torch, tname)
    A32 = _np_input(shape, np.float32, 8)
    At = torch.from_numpy(np.ascontiguousarray(A32)).to(tdt)
    raw = At.view(torch.int16).numpy().view(np.uint16)
    code = sh._abi.BF16 if tname == "bfloat16" else sh._abi.F16
    dev = torch.device("cuda")
    from stencils_hip import _lib
    _lib.set_variant(variant)  # "stream4x4": four rows per lane where a shape has that variant
    try:
        _storage_cases(torch, tdt, A32, At, raw, code, dev, oracle)
    finally:
        _lib.set_variant("")


def _storage_cases(torch, tdt, A32, At, raw, code, dev, oracle):
    for st in (sh.Window(3), sh.VonNeumann(1), sh.Circle(3), sh.Cross(1)):
        for bname, bc in (("wrap", sh.Wrap()), ("reflect", sh.Reflect()), ("remove", sh.Remove(0.0))):
            for f in (sh.sum, sh.mean, sh.maximum, sh.Laplacian(0.2)):
                if isinstance(f, sh.Laplacian) and st.offsets() != sh.VonNeumann(1).offsets():
                    continue
                src = sh.colmajor_empty(A32.shape, tdt, dev)
                src.copy_(At.to(dev))
                g = sh.mapstencil(f, sh.StencilArray(src, st, boundary=bc))
                g = g.cpu().view(torch.int16).numpy().view(np.uint16)
                pv = b"\0\0" if bname == "remove" else None
                o = oracle.map_raw(np.asfortranarray(raw), code, A32.shape, 0, st.offsets(), 2,
                                   bc.code, pv, 0, f.op(sh._abi.OPF_ACC_F32))
                np.testing.assert_array_equal(g, o.view(np.uint16), err_msg=f"{st} {bname} {f}")
            K = sh.Kernel(st, np.random.default_rng(1).random(len(st)).astype(np.float32))
            src = sh.colmajor_empty(A32.shape, tdt, dev)
            src.copy_(At.to(dev))
            g = sh.mapstencil(sh.kernelproduct, sh.StencilArray(src, K, boundary=bc))
            g = g.cpu().view(torch.int16).numpy().view(np.uint16)
            pv = b"\0\0" if bname == "remove" else None
            o = oracle.map_raw(np.asfortranarray(raw), code, A32.shape, 0, st.offsets(), 2, bc.code,
                               pv, 0, sh.kernelproduct.op(sh._abi.OPF_ACC_F32), K.weights, sh._abi.F32)
            bad = np.argwhere(g != o.view(np.uint16))
            np.testing.assert_array_equal(g, o.view(np.uint16),
                                          err_msg=f"{st} {bname} kernelproduct at {bad[:4].tolist()}")


# ---------------------------------------------------- iteration / Life

@pytest.mark.parametrize("pad", ["cond", "halo_out", "halo_in"])
@pytest.mark.parametrize("bname", ["remove", "remove7", "wrap", "reflect"])
@pytest.mark.parametrize("shape", [(256, 96), (48, 40), (2048, 64)])
def test_life_iterate(gpu, oracle, shape, bname, pad):
    A = (np.random.default_rng(9).random(shape) < 0.3).astype(np.uint8)
    bc = BOUNDS[bname](np.uint8)
    for rule in (sh.LifeRule(), sh.LifeRule((3, 6), (2, 3))):
        for graph in (False, True):
            G = sh.SwitchingStencilArray(np.asfortranarray(A), sh.Moore(1), boundary=bc, padding=PADS[pad])
            O = sh.SwitchingStencilArray(np.asfortranarray(A), sh.Moore(1), boundary=bc, padding=PADS[pad])
            G = sh.mapstencil_steps(rule, G, 9, graph=graph)
            O = oracle.mapstencil_steps(rule, O, 9)
            np.testing.assert_array_equal(G.source, O.source)
            np.testing.assert_array_equal(G.dest, O.dest)


@pytest.mark.parametrize("bname", ["remove", "remove1", "wrap", "reflect"])
@pytest.mark.parametrize("gens", ["2", "4", "8", "8nt"])
def test_life_bit_kernel_generations(gpu, oracle, monkeypatch, gens, bname):
    """Conway iterations whose steady launches advance 2, 4 or 8 generations
    with one bit per cell (8nt: with non-temporal stores): both buffers
    (halos included) against the step-by-step oracle, for a one-lane grid, a
    strip with a single active lane, three-row grids (rows reflected or
    wrapped several times over), Bool cells, and chunkings from one chunk
    per strip to the shortest."""
    monkeypatch.setenv("STH_LIFEB_NT", "1" if gens.endswith("nt") else "0")
    gens = int(gens.rstrip("nt"))
    monkeypatch.setenv("STH_LIFEB_G", str(gens))
    bc = {"remove": sh.Remove(np.uint8(0)), "remove1": sh.Remove(np.uint8(1)), "wrap": sh.Wrap(),
          "reflect": sh.Reflect()}[bname]
    rng = np.random.default_rng(90 + gens)
    cases = [((16, 40), "cond"), ((1040, 70), "halo_out"), ((2048, 33), "cond"), ((48, 37), "halo_out"),
             ((32, 3), "cond"), ((64, 4), "halo_out")]
    for waves in ("2048", "1", "100000"):
        monkeypatch.setenv("STH_LIFEB_WAVES", waves)
        for shape, pad in cases:
            A = (rng.random(shape) < 0.35).astype(np.uint8)
            for steps in (2 * gens + 3, 4 * gens + 5, 37):
                G = sh.SwitchingStencilArray(np.asfortranarray(A), sh.Moore(1), boundary=bc, padding=PADS[pad])
                O = sh.SwitchingStencilArray(np.asfortranarray(A), sh.Moore(1), boundary=bc, padding=PADS[pad])
                G = sh.mapstencil_steps(sh.LifeRule(), G, steps, graph=steps % 2 == 1)
                O = oracle.mapstencil_steps(sh.LifeRule(), O, steps)
                msg = f"G={gens} waves={waves} {shape} {pad} {steps}"
                np.testing.assert_array_equal(G.source, O.source, err_msg=msg)
                np.testing.assert_array_equal(G.dest, O.dest, err_msg=msg)
    B = rng.random((96, 50)) < 0.4
    bcb = {"remove": sh.Remove(np.bool_(False)), "remove1": sh.Remove(np.bool_(True))}.get(bname, bc)
    G = sh.SwitchingStencilArray(np.asfortranarray(B), sh.Moore(1), boundary=bcb)
    O = sh.SwitchingStencilArray(np.asfortranarray(B), sh.Moore(1), boundary=bcb)
    G = sh.mapstencil_steps(sh.LifeRule(), G, 4 * gens + 3)
    O = oracle.mapstencil_steps(sh.LifeRule(), O, 4 * gens + 3)
    np.testing.assert_array_equal(G.source, O.source)


def test_life_device_resident_swar_path(gpu, oracle):
    """Halo{:out} device buffers take the SWAR kernel; compare with the oracle."""
    import torch
    n = 1024
    A = (np.random.default_rng(10).random((n, n)) < 0.3).astype(np.uint8)
    dev = torch.device("cuda")
    t = sh.colmajor_empty((n, n), torch.uint8, dev)
    t.copy_(torch.from_numpy(np.ascontiguousarray(A)).to(dev))
    G = sh.SwitchingStencilArray(t, sh.Moore(1), boundary=sh.Wrap(), padding=sh.Halo())
    G = sh.mapstencil_steps(sh.LifeRule(), G, 33, graph=True)
    O = sh.SwitchingStencilArray(np.asfortranarray(A), sh.Moore(1), boundary=sh.Wrap(), padding=sh.Halo())
    O = oracle.mapstencil_steps(sh.LifeRule(), O, 33, nthreads=8)
    np.testing.assert_array_equal(G.source.cpu().numpy(), O.source)
    np.testing.assert_array_equal(G.dest.cpu().numpy(), O.dest)


def test_single_steps_match_iterate(gpu):
    A = np.random.default_rng(12).random((64, 48))
    S1 = sh.SwitchingStencilArray(A.copy(order="F"), sh.Window(1), boundary=sh.Reflect())
    for _ in range(5):
        S1 = sh.mapstencil_(sh.mean, S1)
    S2 = sh.SwitchingStencilArray(A.copy(order="F"), sh.Window(1), boundary=sh.Reflect())
    S2 = sh.mapstencil_steps(sh.mean, S2, 5, graph=True)
    np.testing.assert_array_equal(np.asarray(S1), np.asarray(S2))


def test_graph_cache_bound(gpu, oracle):
    """Iterations on many distinct buffers fill the per-context graph cache
    past its bound (it is then cleared); every result stays exact."""
    import torch
    dev = torch.device("cuda")
    rng = np.random.default_rng(31)
    alive = []  # distinct device buffers, so every call plans new graphs
    for i in range(300):
        A = rng.random((16 + i % 5, 12))
        t = sh.colmajor_empty(A.shape, torch.float64, dev)
        t.copy_(torch.from_numpy(A))
        G = sh.SwitchingStencilArray(t, sh.Window(1), boundary=sh.Reflect())
        alive.append(G)
        O = sh.SwitchingStencilArray(A.copy(order="F"), sh.Window(1), boundary=sh.Reflect())
        # 20 steps (graphs start at 17 launches): two graphs per call (the
        # first step with 18 replayed steps, and 1), 600 in all
        G = sh.mapstencil_steps(sh.mean, G, 20, graph=True)
        O = oracle.mapstencil_steps(sh.mean, O, 20)
        np.testing.assert_array_equal(G.source.cpu().numpy(), O.source, err_msg=f"call {i}")


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_heat_3d_iterate(gpu, oracle, dt):
    A = _np_input((40, 33, 29), dt, 13)
    for pad in ("cond", "halo_out"):
        G = sh.SwitchingStencilArray(A.copy(order="F"), sh.VonNeumann(1, 3), boundary=sh.Wrap(),
                                     padding=PADS[pad])
        O = sh.SwitchingStencilArray(A.copy(order="F"), sh.VonNeumann(1, 3), boundary=sh.Wrap(),
                                     padding=PADS[pad])
        G = sh.mapstencil_steps(sh.Laplacian(0.15), G, 21, graph=True)
        O = oracle.mapstencil_steps(sh.Laplacian(0.15), O, 21)
        np.testing.assert_array_equal(np.asarray(G), np.asarray(O))


ZM1_VARIANTS = {"float32": ["zm1v2r8d4nt", "zm1v1r16d4nt", "zm1v1r16d4c"],
                "float64": ["zm1v1r4d6nt", "zm1v1r8d4nt", "zm1v1r4d6c"],
                "bfloat16": ["zm1v2r8d4nt", "zm1v4r8d4nt", "zm1v2r8d4c"],
                "float16": ["zm1v2r8d4nt", "zm1v4r8d4nt"]}


@pytest.mark.parametrize("pad", ["cond", "halo_out"])
@pytest.mark.parametrize("bname", ["remove", "remove7", "wrap", "reflect"])
@pytest.mark.parametrize("tname", list(ZM1_VARIANTS))
def test_zmarch1_variants(gpu, oracle, tname, bname, pad):
    """Every z-marching one-step variant (forced by name) on 3-D arrays of
    reach-1 stencils, against the oracle: partial x tiles, partial row
    strips, short z chunks, two-cell axes, and widths that are not a whole
    number of a variant's vectors, where it does not apply and the
    streaming kernel runs."""
    import torch
    from stencils_hip import _lib
    tdt = getattr(torch, tname)
    storage = tname in ("bfloat16", "float16")
    code = {"bfloat16": sh._abi.BF16, "float16": sh._abi.F16}.get(tname)
    dev = torch.device("cuda")
    cases = ((sh.VonNeumann(1, 3), sh.Laplacian(0.15)), (sh.Moore(1, 3), sh.mean),
             (sh.VonNeumann(1, 3), sh.maximum))
    shapes = [(130, 37, 29), (64, 9, 5), (2, 3, 2), (196, 20, 70), (67, 21, 11)]
    rng = np.random.default_rng(77)
    try:
        for pat in ZM1_VARIANTS[tname]:
            _lib.set_variant(pat)
            for shape in shapes:
                A32 = rng.random(shape).astype(np.float32) * 2 - 1
                for st, f in cases:
                    if storage:
                        pvv = 0.75 if bname == "remove7" else 0.0
                        bc = {"remove": sh.Remove(0.0), "remove7": sh.Remove(pvv), "wrap": sh.Wrap(),
                              "reflect": sh.Reflect()}[bname]
                    else:
                        bc = BOUNDS[bname](np.dtype(tname))
                    At = torch.from_numpy(np.ascontiguousarray(A32)).to(tdt)
                    t = sh.colmajor_empty(shape, tdt, dev)
                    t.copy_(At.to(dev))
                    sa = sh.StencilArray(t, st, boundary=bc, padding=PADS[pad])
                    dest = sh.colmajor_empty(shape, tdt, dev)
                    chosen = sh.autotune(f, dest, sa)
                    v = int(pat[pat.index("v") + 1:pat.index("r")])
                    if shape[0] % v == 0:  # rows a whole number of vectors
                        assert pat in chosen, (pat, chosen, shape)
                    else:
                        assert pat not in chosen, (pat, chosen, shape)
                    if storage:
                        raw = np.asfortranarray(sa.parent().cpu().view(torch.int16).numpy().view(np.uint16))
                        pv = torch.tensor([bc.padval], dtype=tdt).view(torch.int16).numpy().tobytes() \
                            if isinstance(bc, sh.Remove) else None
                        o = oracle.map_raw(raw, code, shape, sa.halo(), st.offsets(), 3, bc.code, pv,
                                           PADS[pad].code, f.op(sh._abi.OPF_ACC_F32))
                        g = dest.cpu().view(torch.int16).numpy().view(np.uint16)
                        np.testing.assert_array_equal(g, o.view(np.uint16), err_msg=f"{pat} {shape} {st} {f}")
                    else:
                        A = At.numpy()
                        o = oracle.mapstencil(f, sh.StencilArray(np.asfortranarray(A), st, boundary=bc,
                                                                 padding=PADS[pad]))
                        np.testing.assert_array_equal(dest.cpu().numpy(), o, err_msg=f"{pat} {shape} {st} {f}")
    finally:
        _lib.set_variant("")


@pytest.mark.parametrize("pad", ["cond", "halo_out", "halo_in"])
@pytest.mark.parametrize("bname", ["remove", "remove7", "wrap", "reflect"])
@pytest.mark.parametrize("shape", [(70, 17, 35), (3, 3, 3), (130, 40, 33)])
def test_fused_3d_iterate(gpu, oracle, shape, bname, pad):
    """Iterations of 3-D radius-1 stencils advance two steps per launch in the
    steady state; both buffers (halos included) must equal the step-by-step
    oracle for step counts around the schedule's boundaries."""
    if pad == "halo_in":  # the halo lies inside the array: keep the same interior
        shape = tuple(n + 2 for n in shape)
    for dt, st, f in ((np.float32, sh.VonNeumann(1, 3), sh.Laplacian(0.15)),
                      (np.float64, sh.VonNeumann(1, 3), sh.Laplacian(0.1)),
                      (np.float32, sh.Moore(1, 3), sh.mean),
                      (np.float64, sh.Moore(1, 3), sh.maximum)):
        A = _np_input(shape, dt, sum(shape))
        bc = BOUNDS[bname](dt)
        for steps in (6, 7, 10, 13):
            G = sh.SwitchingStencilArray(A.copy(order="F"), st, boundary=bc, padding=PADS[pad])
            O = sh.SwitchingStencilArray(A.copy(order="F"), st, boundary=bc, padding=PADS[pad])
            G = sh.mapstencil_steps(f, G, steps, graph=steps % 2 == 0)
            O = oracle.mapstencil_steps(f, O, steps)
            np.testing.assert_array_equal(G.source, O.source, err_msg=f"{dt} {st} {steps}")
            np.testing.assert_array_equal(G.dest, O.dest, err_msg=f"{dt} {st} {steps}")


@pytest.mark.parametrize("pad", ["cond", "halo_out", "halo_in"])
@pytest.mark.parametrize("bname", ["remove", "remove7", "wrap", "reflect"])
@pytest.mark.parametrize("shape", [(300, 37), (3, 3), (530, 70), (17, 5), (224, 9), (225, 8)])
def test_fused_2d_iterate(gpu, oracle, shape, bname, pad):
    """Iterations of 2-D radius-1 stencils on aligned rows advance two steps
    per launch in the steady state; both buffers (halos included) must equal
    the step-by-step oracle."""
    if pad == "halo_in":
        shape = tuple(n + 2 for n in shape)
    for dt, st, f in ((np.float32, sh.Window(1), sh.mean),
                      (np.float64, sh.VonNeumann(1), sh.Laplacian(0.2)),
                      (np.float32, sh.Moore(1), sh.maximum),
                      (np.float64, sh.Window(1), sh.sum)):
        A = _np_input(shape, dt, sum(shape) + 1)
        bc = BOUNDS[bname](dt)
        for steps in (6, 7, 10, 13):
            G = sh.SwitchingStencilArray(A.copy(order="F"), st, boundary=bc, padding=PADS[pad])
            O = sh.SwitchingStencilArray(A.copy(order="F"), st, boundary=bc, padding=PADS[pad])
            G = sh.mapstencil_steps(f, G, steps, graph=steps % 2 == 0)
            O = oracle.mapstencil_steps(f, O, steps)
            np.testing.assert_array_equal(G.source, O.source, err_msg=f"{dt} {st} {steps}")
            np.testing.assert_array_equal(G.dest, O.dest, err_msg=f"{dt} {st} {steps}")


STORAGE_BCS = {"wrap": lambda: sh.Wrap(), "reflect": lambda: sh.Reflect(),
               "remove_nz": lambda: sh.Remove(0.75)}


def _storage_iterate_pair(torch, shape, dt, st, f, bc, pad, steps):
    """Both buffers after `steps` iterations with two-step launches and with
    one launch per step, from the same initial state."""
    dev = torch.device("cuda")
    A = sh.colmajor_empty(shape, dt, dev)
    A.copy_(torch.rand(shape, device=dev).to(dt))
    out = []
    for flags in (0, sh._abi.NO_FUSE):
        src = sh.colmajor_empty(shape, dt, dev)
        src.copy_(A)
        S = sh.SwitchingStencilArray(src, st, boundary=bc, padding=pad)
        S = sh.mapstencil_steps(f, S, steps, flags=flags)
        out.append((S.source.clone(), S.dest.clone()))
    return out


@pytest.mark.parametrize("pad", ["cond", "halo_out"])
@pytest.mark.parametrize("bname", list(STORAGE_BCS))
@pytest.mark.parametrize("tname", ["bfloat16", "float16"])
def test_fused_2d_iterate_storage_types(gpu, tname, bname, pad):
    """16-bit storage in 2-D: two-step launches equal the unfused path, for
    Wrap, Reflect (mirrored rim fix-ups) and Remove with a nonzero padval.
    n0 = 56 * 8 + 3 puts the right rim of a Halo{:out} pitched row in the
    previous wave's vectors."""
    import torch
    dt = getattr(torch, tname)
    for shape in ((700, 45), (56 * 8 + 3, 37)):
        out = _storage_iterate_pair(torch, shape, dt, sh.Window(1), sh.mean, STORAGE_BCS[bname](), PADS[pad], 17)
        assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1]), str(shape)


@pytest.mark.parametrize("pad", ["cond", "halo_out"])
@pytest.mark.parametrize("bname", list(STORAGE_BCS))
@pytest.mark.parametrize("tdt", ["bfloat16", "float16"])
def test_fused_3d_iterate_storage_types(gpu, tdt, bname, pad):
    """16-bit storage: the fused launches round the intermediate step to the
    storage type, so results equal the unfused step-by-step path bit for bit
    (Wrap, Reflect, Remove with a nonzero padval)."""
    import torch
    dt = getattr(torch, tdt)
    f = sh.Laplacian(0.15) if tdt == "bfloat16" else sh.mean
    for shape in ((96, 37, 41), (63, 20, 19)):
        out = _storage_iterate_pair(torch, shape, dt, sh.VonNeumann(1, 3), f, STORAGE_BCS[bname](), PADS[pad], 17)
        assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1]), str(shape)


# ------------------------------------------------------------ errors

def test_argument_errors(gpu):
    A = np.zeros((10, 10))
    with pytest.raises(ValueError, match="sizes must match"):
        sh.mapstencil_(sh.mean, np.zeros((9, 10)), sh.StencilArray(A, sh.Window(1)))
    with pytest.raises(ValueError, match="Use"):
        sh.mapstencil(sh.mean, sh.StencilArray(A, sh.Window(1), boundary=sh.Use()))
    with pytest.raises(sh.Unsupported):  # reach 3 > halo 1 (Positional signed radius)
        sh.mapstencil(sh.sum, sh.StencilArray(A, sh.Positional((-3, 0), (1, 0)), padding=sh.Halo()))
    with pytest.raises(sh.Unsupported):
        sh.mapstencil(lambda s: s, sh.StencilArray(A, sh.Window(1)))
    with pytest.raises(sh.Unsupported):
        sh.mapstencil(sh.mean, sh.StencilArray(np.zeros((5, 5, 5)), sh.Window(1, 3)))
