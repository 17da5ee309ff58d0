"""The oracle against checks that do not share its code: scipy.ndimage for the
boundary modes and reductions, closed forms (Fourier-mode decay of the heat
step, mass conservation), textbook Life patterns and a numpy-roll Life,
rounding against torch / numpy conversions, and Julia's min/max semantics.
Also plants mistakes (transposed offsets, a repeated-edge reflect) and shows
the checks catch them."""
import numpy as np
import pytest
import scipy.ndimage as ndi

import stencils_hip as sh
from stencils_hip import _abi

MODES = {"remove": ("constant", lambda: sh.Remove(0.0)), "wrap": ("wrap", sh.Wrap),
         "reflect": ("mirror", sh.Reflect)}
SHAPES = [sh.Window(1), sh.Window(2), sh.Moore(1), sh.VonNeumann(2), sh.Circle(3),
          sh.Positional((2, -1), (0, 0), (-1, 3)), sh.Cross(2), sh.Diamond(2), sh.Annulus(3)]


def footprint(st):
    R = max(st.reach())
    fp = np.zeros((2 * R + 1,) * st.N, dtype=bool)
    for o in st.offsets():
        fp[tuple(v + R for v in o)] = True
    return fp


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("st", SHAPES, ids=repr)
def test_sum_mean_max_min_vs_scipy(oracle, st, mode):
    rng = np.random.default_rng(11)
    A = np.asfortranarray(rng.random((23, 17)))
    smode, bc = MODES[mode]
    fp = footprint(st)
    sa = sh.StencilArray(A.copy(order="F"), st, boundary=bc())
    s = oracle.mapstencil(sh.sum, sa)
    ref = ndi.correlate(A, fp.astype(float), mode=smode, cval=0.0)
    np.testing.assert_allclose(s, ref, rtol=1e-13, atol=1e-13)
    m = oracle.mapstencil(sh.mean, sa)
    np.testing.assert_allclose(m, ref / len(st), rtol=1e-13, atol=1e-13)
    mx = oracle.mapstencil(sh.maximum, sa)
    np.testing.assert_array_equal(mx, ndi.maximum_filter(A, footprint=fp, mode=smode, cval=0.0))
    mn = oracle.mapstencil(sh.minimum, sa)
    np.testing.assert_array_equal(mn, ndi.minimum_filter(A, footprint=fp, mode=smode, cval=0.0))


@pytest.mark.parametrize("mode", list(MODES))
def test_integer_sum_exact_vs_scipy(oracle, mode):
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.integers(-1000, 1000, (31, 29), dtype=np.int64))
    smode, _ = MODES[mode]
    bc = {"remove": sh.Remove(np.int64(0)), "wrap": sh.Wrap(), "reflect": sh.Reflect()}[mode]
    for st in (sh.Moore(2), sh.Circle(3), sh.VonNeumann(3)):
        s = oracle.mapstencil(sh.sum, sh.StencilArray(A, st, boundary=bc))
        ref = ndi.correlate(A, footprint(st).astype(np.int64), mode=smode, cval=0)
        np.testing.assert_array_equal(s, ref)


def test_kernelproduct_vs_scipy_3d(oracle):
    rng = np.random.default_rng(5)
    A = np.asfortranarray(rng.random((9, 10, 11)))
    W = rng.random((3, 3, 3))
    K = sh.Kernel(sh.Positional(*[o for o in sh.Moore(1, 3).offsets()] + [(0, 0, 0)]),
                  np.array([W[o[0] + 1, o[1] + 1, o[2] + 1] for o in sh.Moore(1, 3).offsets()] +
                           [W[1, 1, 1]]))
    out = oracle.mapstencil(sh.kernelproduct, sh.StencilArray(A, K, boundary=sh.Wrap()))
    np.testing.assert_allclose(out, ndi.correlate(A, W, mode="wrap"), rtol=1e-13, atol=1e-13)


def test_kernel_matrix_is_correlation(oracle):
    """Kernel(W) pairs W[a, b] with the neighbour at (a - R, b - R): no flip."""
    rng = np.random.default_rng(9)
    A = np.asfortranarray(rng.random((15, 13)))
    W = rng.random((7, 7))
    out = oracle.mapstencil(sh.kernelproduct, sh.StencilArray(A, sh.Kernel(W), boundary=sh.Reflect()))
    np.testing.assert_allclose(out, ndi.correlate(A, W, mode="mirror"), rtol=1e-13, atol=1e-13)


def test_planted_mistakes_are_caught(oracle, golden):
    """Transposed offsets and a repeated-edge reflect fail the goldens."""
    win = np.asfortranarray(np.asarray(golden["cases"][-1]["input"], dtype=np.int64))
    exp = np.asarray(golden["cases"][-1]["expected"])  # Ordinal NE + NW
    good = sh.Positional((1, -1), (1, 1))
    bad = sh.Positional((-1, 1), (1, 1))  # transposed first offset
    assert np.array_equal(oracle.mapstencil(sh.sum, sh.StencilArray(win, good, boundary=sh.Remove(np.int64(0)))), exp)
    assert not np.array_equal(oracle.mapstencil(sh.sum, sh.StencilArray(win, bad, boundary=sh.Remove(np.int64(0)))), exp)
    # scipy's 'reflect' repeats the edge; the reference's Reflect does not
    x = np.asfortranarray(np.arange(1.0, 6.0))
    ours = oracle.mapstencil(sh.mean, sh.StencilArray(x, sh.Window(1, 1), boundary=sh.Reflect()))
    rep = ndi.correlate(x, np.ones(3) / 3, mode="reflect")
    assert not np.allclose(ours, rep)
    np.testing.assert_allclose(ours, ndi.correlate(x, np.ones(3) / 3, mode="mirror"), atol=1e-15)


# ------------------------------------------------------------------ Life

def life_roll(A, steps, birth=(3,), survive=(2, 3)):
    """Textbook torus Life with numpy.roll (independent of the oracle)."""
    A = (A != 0).astype(np.int64)
    for _ in range(steps):
        n = sum(np.roll(np.roll(A, i, 0), j, 1) for i in (-1, 0, 1) for j in (-1, 0, 1)
                if (i, j) != (0, 0))
        A = np.where(A == 1, np.isin(n, survive), np.isin(n, birth)).astype(np.int64)
    return A


def _life(oracle, A, steps, padding=sh.Conditional(), rule=None):
    S = sh.SwitchingStencilArray(np.asfortranarray(A.astype(np.uint8)), sh.Moore(1),
                                 boundary=sh.Wrap(), padding=padding)
    S = oracle.mapstencil_steps(rule or sh.LifeRule(), S, steps)
    return np.asarray(S)


def test_life_patterns(oracle):
    n = 16
    blinker = np.zeros((n, n), np.uint8)
    blinker[5, 4:7] = 1
    assert np.array_equal(_life(oracle, blinker, 2), blinker)
    assert not np.array_equal(_life(oracle, blinker, 1), blinker)
    block = np.zeros((n, n), np.uint8)
    block[3:5, 3:5] = 1
    assert np.array_equal(_life(oracle, block, 7), block)
    glider = np.zeros((n, n), np.uint8)
    for i, j in [(0, 1), (1, 2), (2, 0), (2, 1), (2, 2)]:
        glider[i + 4, j + 4] = 1
    # after 4 generations the glider has moved one cell down and right
    assert np.array_equal(_life(oracle, glider, 4), np.roll(np.roll(glider, 1, 0), 1, 1))
    # on the 16-torus it returns home after 64 generations
    assert np.array_equal(_life(oracle, glider, 64, padding=sh.Halo()), glider)


@pytest.mark.parametrize("padding", [sh.Conditional(), sh.Halo("out")], ids=repr)
def test_life_vs_numpy_roll(oracle, padding):
    rng = np.random.default_rng(2)
    A = (rng.random((48, 40)) < 0.3).astype(np.uint8)
    np.testing.assert_array_equal(_life(oracle, A, 60, padding), life_roll(A, 60))
    rule = sh.LifeRule(birth=(3, 6), survive=(2, 3))  # HighLife
    np.testing.assert_array_equal(_life(oracle, A, 30, padding, rule),
                                  life_roll(A, 30, (3, 6), (2, 3)))


# ------------------------------------------------------------- diffusion

@pytest.mark.parametrize("N,shape,k", [(2, (32, 24), (3, 2)), (3, (16, 12, 20), (1, 2, 3))])
def test_heat_fourier_mode_decay(oracle, N, shape, k):
    """u = prod cos(2 pi k_d x_d / n_d) decays by (1 - 4 a sum sin^2(pi k_d / n_d))^t."""
    a = 0.15
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    u0 = np.ones(shape)
    for g, n, kk in zip(grids, shape, k):
        u0 = u0 * np.cos(2 * np.pi * kk * g / n)
    S = sh.SwitchingStencilArray(np.asfortranarray(u0), sh.VonNeumann(1, N), boundary=sh.Wrap())
    t = 40
    S = oracle.mapstencil_steps(sh.Laplacian(a), S, t)
    lam = 1 - 4 * a * sum(np.sin(np.pi * kk / n) ** 2 for n, kk in zip(shape, k))
    np.testing.assert_allclose(np.asarray(S), lam ** t * u0, rtol=0, atol=1e-12)


def test_heat_conserves_mass_under_wrap(oracle):
    rng = np.random.default_rng(8)
    u0 = rng.random((20, 18, 16))
    S = sh.SwitchingStencilArray(np.asfortranarray(u0), sh.VonNeumann(1, 3), boundary=sh.Wrap())
    S = oracle.mapstencil_steps(sh.Laplacian(0.1), S, 25)
    assert abs(np.asarray(S).sum() - u0.sum()) < 1e-9


# ------------------------------------------------- rounding and min / max

def test_bf16_rounding_matches_torch(oracle):
    import torch
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 20000),
                         np.array([0.0, -0.0, np.inf, -np.inf, 1.0, 3.0e38], np.float32)])
    ref = torch.from_numpy(xs).to(torch.bfloat16).view(torch.int16).numpy().astype(np.uint16)
    ours = np.array([oracle.f32_to_bf16(float(x)) for x in xs], dtype=np.uint16)
    np.testing.assert_array_equal(ours, ref)


def test_f16_rounding_matches_numpy(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([(rng.standard_normal(20000) * 10.0 ** rng.integers(-9, 6, 20000)).astype(np.float32),
                         np.array([0.0, -0.0, np.inf, 65504.0, 65520.0, 6.0e-8, 2.9802322e-08], np.float32)])
    with np.errstate(over="ignore"):
        ref = xs.astype(np.float16).view(np.uint16)
    ours = np.array([L.orc_f32_to_f16(float(x)) for x in xs], dtype=np.uint16)
    np.testing.assert_array_equal(ours, ref)


def test_julia_min_max_semantics(oracle):
    """Julia max/min: NaN propagates; -0.0 < +0.0 (base/math.jl)."""
    A = np.array([0.0, -0.0, 1.0, np.nan, 2.0, -0.0, 0.0])
    pv = sh.Remove(-5.0)
    mx = oracle.mapstencil(sh.maximum, sh.StencilArray(A, sh.Positional((0,), (1,)), boundary=pv))
    mn = oracle.mapstencil(sh.minimum, sh.StencilArray(A, sh.Positional((0,), (1,)), boundary=pv))
    assert np.signbit(mn[0]) and not np.signbit(mx[0])          # max(+0, -0) = +0, min = -0
    assert np.isnan(mx[2]) and np.isnan(mx[3]) and np.isnan(mn[3])
    assert np.signbit(mn[5]) and not np.signbit(mx[5])          # (-0, +0)
    assert mx[6] == 0.0 and mn[6] == -5.0


def test_integer_sum_wraps_in_element_type(oracle):
    """Integer folds use `+` of the element type (wrapping).  The reference's
    tests only cover Int64 (no overflow), so small-int overflow is this
    project's stated semantics: StaticArrays' `sum` folds with `+`."""
    A = np.asfortranarray(np.full((4, 4), 100, dtype=np.int8))
    s = oracle.mapstencil(sh.sum, sh.StencilArray(A, sh.Window(1), boundary=sh.Wrap()))
    assert s.dtype == np.int8 and (s == np.int8(900 - 1024)).all()   # 900 mod 256 as Int8
    B = np.asfortranarray(np.ones((4, 4), dtype=np.bool_))
    s = oracle.mapstencil(sh.sum, sh.StencilArray(B, sh.Window(1), boundary=sh.Wrap()))
    assert s.dtype == np.int64 and (s == 9).all()       # Bool + Bool is Int
    m = oracle.mapstencil(sh.mean, sh.StencilArray(B, sh.Window(1), boundary=sh.Wrap()))
    assert m.dtype == np.float64 and (m == 1.0).all()   # mean of Bool is Float64


def test_integer_mean_is_element_type_sum_over_length(oracle):
    """Integer `mean` = `sum(a) / length(a)` over the StaticVector: the sum
    folds with the element type's wrapping `+` and is then divided in
    Float64.  Checked by brute force on tiny grids for every integer type,
    with values chosen to overflow the narrow types (so widening first, the
    other plausible reading, gives different numbers)."""
    rng = np.random.default_rng(40)
    st = sh.Moore(1)
    offs = st.offsets()
    for dt in (np.uint8, np.int8, np.uint16, np.int16, np.uint32, np.int32, np.uint64, np.int64):
        info = np.iinfo(dt)
        A = np.asfortranarray(rng.integers(info.max // 2, info.max, (7, 6), dtype=dt))
        if info.min < 0:
            A[::2] = (-A[::2] - 1).astype(dt)
        got = oracle.mapstencil(sh.mean, sh.StencilArray(A.copy(order="F"), st, boundary=sh.Wrap()))
        exp = np.zeros(A.shape)
        widened = np.zeros(A.shape)
        for i in range(A.shape[0]):
            for j in range(A.shape[1]):
                vals = [int(A[(i + dx) % A.shape[0], (j + dy) % A.shape[1]]) for dx, dy in offs]
                s = sum(vals) % (1 << (8 * A.itemsize))        # wrap to the width ...
                if info.min < 0 and s > info.max:               # ... as two's complement
                    s -= 1 << (8 * A.itemsize)
                exp[i, j] = float(s) / len(offs)
                widened[i, j] = sum(float(v) for v in vals) / len(offs)
        np.testing.assert_array_equal(got, exp, err_msg=np.dtype(dt).name)
        assert not np.array_equal(got, widened)   # the test input overflows


def test_host_declines_unpinned_integer_folds():
    """Without OPF_INT_WRAP the front end and the ABI mirror decline integer
    sum over sub-64-bit types and integer mean (L >= 2); with it, or for
    Bool, 64-bit sums, min/max and L = 1, they are accepted."""
    A8 = sh.StencilArray(np.zeros((6, 6), np.int8, order="F"), sh.Moore(2))
    old = sh.set_integer_fold_wraps(False)
    try:
        for f in (sh.sum, sh.mean):
            with pytest.raises(sh.Unsupported, match="unpinned"):
                sh.result_dtype(f, A8)
        assert sh.result_dtype(sh.maximum, A8) == sh._abi.I8
        A64 = sh.StencilArray(np.zeros((6, 6), np.int64, order="F"), sh.Moore(2))
        assert sh.result_dtype(sh.sum, A64) == sh._abi.I64
        with pytest.raises(sh.Unsupported):
            sh.result_dtype(sh.mean, A64)
        B = sh.StencilArray(np.zeros((6, 6), np.bool_, order="F"), sh.Moore(2))
        assert sh.result_dtype(sh.sum, B) == sh._abi.I64 and sh.result_dtype(sh.mean, B) == sh._abi.F64
        one = sh.StencilArray(np.zeros((6, 6), np.int8, order="F"), sh.Positional((1, 0)))
        assert sh.result_dtype(sh.sum, one) == sh._abi.I8
        sh.set_integer_fold_wraps(True)
        assert sh.result_dtype(sh.sum, A8) == sh._abi.I8 and sh.result_dtype(sh.mean, A8) == sh._abi.F64
    finally:
        sh.set_integer_fold_wraps(old)
    for dt in range(13):
        for op in range(7):
            for L in (1, 2, 25):
                declined = _abi.int_fold_declined(dt, op, 0, L)
                assert declined == (L >= 2 and dt not in _abi.FLOAT_DTYPES and dt != _abi.BOOL and
                                    ((op == _abi.OP_SUM and _abi.DTYPE_SIZE[dt] < 8) or op == _abi.OP_MEAN))
                assert not _abi.int_fold_declined(dt, op, _abi.OPF_INT_WRAP, L)


@pytest.mark.parametrize("R", [1, 2, 3])
def test_remaining_shape_tables(R):
    """Offset tables of the shapes the reference's tests do not enumerate,
    against closed forms from their definitions' docstrings and geometry
    (src/stencils/shapes.jl): counts, symmetry, the fold order
    (CartesianIndices order: dim 1 fastest), and the centre where it
    belongs."""
    def order_ok(offs):
        keys = [tuple(reversed(o)) for o in offs]
        return keys == sorted(keys)

    cases = {
        sh.Diamond(R): 2 * R * R + 2 * R + 1,          # the L1 ball in 2-D
        sh.Diamond(R, 3): (2 * R + 1) * (2 * R * R + 2 * R + 3) // 3,
        sh.ForwardSlash(R): 2 * R + 1, sh.BackSlash(R): 2 * R + 1,   # "contains 2R+1 neighbors"
        sh.AngledCross(R): 4 * R + 1,                  # both diagonals, one shared centre
        sh.Vertical(R): 2 * R + 1, sh.Horizontal(R): 2 * R + 1,
        sh.Vertical(R, 3): (2 * R + 1) ** 2, sh.Horizontal(R, 3): (2 * R + 1) ** 2,
        sh.Cross(R): 4 * R + 1,
    }
    for st, n in cases.items():
        offs = [tuple(o) for o in st.offsets()]
        assert len(offs) == n == len(set(offs)), (st, len(offs), n)
        assert order_ok(offs), st
        assert (0,) * st.N in offs, st
        assert sorted(tuple(-v for v in o) for o in offs) == sorted(offs), st  # point symmetric
        assert max(max(abs(v) for v in o) for o in offs) == R
    # which axis the bars run along: Vertical varies dim 1 (dim 2 fixed at 0),
    # Horizontal varies dim 2, and the slashes are the two diagonals
    assert all(o[1] == 0 for o in sh.Vertical(R).offsets())
    assert all(o[0] == 0 for o in sh.Horizontal(R).offsets())
    assert all(o[1] == -o[0] for o in sh.ForwardSlash(R).offsets())
    assert all(o[1] == o[0] for o in sh.BackSlash(R).offsets())
    assert set(map(tuple, sh.AngledCross(R).offsets())) == \
        set(map(tuple, sh.ForwardSlash(R).offsets())) | set(map(tuple, sh.BackSlash(R).offsets()))
