"""The oracle's multi-array gather (mapstencil(f, A, B, ...)) pinned on the CPU.

The reference's own cases (test/array.jl:312-383) are checked with the values
they assert and, where the reference only checks a size or an inequality,
against a brute-force restatement in plain Python loops over these tiny
arrays (float arithmetic in the element type, offsets in fold order, the
boundary applied per neighbour).  Random small cases compare the oracle's
composition of per-term folds with the same brute force.
"""
import itertools

import numpy as np
import pytest

import stencils_hip as sh


def _neighbour(A, I, off, bc):
    """Value of A at I + off under boundary `bc` (Remove(pv) / Wrap / Reflect),
    src/array.jl:129-179; None for the Remove padval."""
    J = [i + o for i, o in zip(I, off)]
    if all(0 <= j < s for j, s in zip(J, A.shape)):
        return A[tuple(J)]
    if isinstance(bc, sh.Remove):
        return None
    out = []
    for j, s in zip(J, A.shape):
        if isinstance(bc, sh.Wrap):
            j = j + s if j < 0 else (j - s if j >= s else j)
        else:  # Reflect: the edge is not repeated
            j = -j if j < 0 else (2 * s - 2 - j if j >= s else j)
        out.append(j)
    return A[tuple(out)]


def _fold(A, I, offs, bc, g, T):
    """g over the hood at I, left fold in offset order in element type T."""
    vals = []
    for o in offs:
        v = _neighbour(A, I, tuple(o) + (0,) * (A.ndim - len(o)), bc)
        vals.append(T(bc.padval) if v is None else T(v))
    if g is sh.center:
        return T(A[I])
    acc = vals[0]
    for v in vals[1:]:
        if g is sh.maximum:
            acc = max(acc, v) if not (np.isnan(acc) or np.isnan(v)) else T(np.nan)
        elif g is sh.minimum:
            acc = min(acc, v) if not (np.isnan(acc) or np.isnan(v)) else T(np.nan)
        else:
            acc = T(acc + v)
    if g is sh.mean:
        acc = T(acc / T(len(vals)))
    return acc


def _brute(terms, arrays, stencils, bcs, combine, T):
    with np.errstate(over="ignore"):  # integer `+` / `*` wrap
        return _brute_cells(terms, arrays, stencils, bcs, combine, T)


def _brute_cells(terms, arrays, stencils, bcs, combine, T):
    out = np.zeros(arrays[0].shape, dtype=T)
    for I in itertools.product(*(range(s) for s in arrays[0].shape)):
        acc = None
        for g, A, st, bc in zip(terms, arrays, stencils, bcs):
            coef = None
            if isinstance(g, sh.Scaled):
                coef, g = g.coef, g.g
            v = _fold(A, I, st.offsets() if st is not None else [(0,) * A.ndim], bc, g, T)
            if coef is not None:
                v = T(T(coef) * v)
            acc = v if acc is None else (T(acc * v) if combine == "*" else T(acc + v))
        out[I] = acc
    return out


def test_two_stencil_arrays_centres(oracle):
    """test/array.jl:313-324: center(a) + center(b) == A .+ B."""
    A = np.array([[1, 2, 3], [4, 5, 6], [7, 8, 9]], dtype=np.float64)
    B = np.array([[10, 20, 30], [40, 50, 60], [70, 80, 90]], dtype=np.float64)
    r = oracle.mapstencil_multi(sh.Terms(sh.center, sh.center),
                                (sh.StencilArray(np.asfortranarray(A), sh.Moore(1)),
                                 sh.StencilArray(np.asfortranarray(B), sh.Moore(1))))
    np.testing.assert_array_equal(r, A + B)


def test_two_stencil_arrays_neighbour_sums(oracle):
    """test/array.jl:326-330 (the reference checks the size): sum(neighbors(a))
    + sum(neighbors(b)) under the default Remove(0); against brute force, and
    the corner value by hand: 2 + 4 + 5 = 11 and 20 + 40 + 50 = 110."""
    A = np.array([[1, 2, 3], [4, 5, 6], [7, 8, 9]], dtype=np.float64)
    B = 10 * A
    r = oracle.mapstencil_multi(sh.Terms(sh.sum, sh.sum),
                                (sh.StencilArray(np.asfortranarray(A), sh.Moore(1)),
                                 sh.StencilArray(np.asfortranarray(B), sh.Moore(1))))
    assert r.shape == A.shape
    assert r[0, 0] == 11 + 110
    bc = sh.Remove(0.0)
    np.testing.assert_array_equal(r, _brute([sh.sum, sh.sum], [A, B], [sh.Moore(1)] * 2, [bc, bc], "+",
                                            np.float64))


def test_switching_with_stencil_array(oracle):
    """test/array.jl:333-351: three mapstencil!(ssa, sa) steps of
    c + 0.1 * sum(neighbors(r)); the reference asserts the centre grows; the
    whole state equals three brute-force steps."""
    m = np.array([[0, 0, 0], [0, 1, 0], [0, 0, 0]], dtype=np.float64)
    ro = np.array([[1, 1, 1], [1, 0, 1], [1, 1, 1]], dtype=np.float64)
    ssa = sh.SwitchingStencilArray(np.asfortranarray(m), sh.Moore(1))
    sa = sh.StencilArray(np.asfortranarray(ro), sh.Moore(1))
    f = sh.Terms(sh.center, sh.Scaled(0.1, sh.sum))
    ref = m.copy()
    bc = sh.Remove(0.0)
    for _ in range(3):
        ssa = oracle.mapstencil_multi_switching(f, ssa, (sa,))
        ref = _brute(f.terms, [ref, ro], [None, sh.Moore(1)], [bc, bc], "+", np.float64)
    res = np.asarray(ssa.source)
    assert res.shape == m.shape
    assert res[1, 1] > m[1, 1]
    np.testing.assert_array_equal(res, ref)
    # by hand: the centre sees 8 ones, so it gains 0.1 * 8 three times
    assert res[1, 1] == ((1 + 0.1 * 8.0) + 0.1 * 8.0) + 0.1 * 8.0


def test_three_stencil_arrays(oracle):
    """test/array.jl:354-368: 1 + 2 + 3 at the centres."""
    arrs = [np.full((5, 5), v, dtype=np.float64, order="F") for v in (1.0, 2.0, 3.0)]
    r = oracle.mapstencil_multi(sh.Terms(sh.center, sh.center, sh.center),
                                tuple(sh.StencilArray(a, sh.Moore(1)) for a in arrs))
    assert (r[1:4, 1:4] == 6.0).all() and (r == 6.0).all()


def test_stencil_array_and_plain_array(oracle):
    """test/array.jl:370-381: a plain array argument is read at I."""
    A = np.array([[1, 2, 3], [4, 5, 6], [7, 8, 9]], dtype=np.float64)
    B = np.array([[10, 20, 30], [40, 50, 60], [70, 80, 90]], dtype=np.float64)
    r = oracle.mapstencil_multi(sh.Terms(sh.center, sh.center),
                                (sh.StencilArray(np.asfortranarray(A), sh.Moore(1)), np.asfortranarray(B)))
    np.testing.assert_array_equal(r, A + B)


@pytest.mark.parametrize("seed", range(6))
def test_random_terms_against_brute_force(oracle, seed):
    """Mixed folds, boundaries, scaled terms and both combines over small
    Float32 / Float64 / UInt8 grids: the oracle's composition equals the
    brute-force restatement bit for bit."""
    rng = np.random.default_rng(100 + seed)
    T = [np.float32, np.float64, np.uint8][seed % 3]
    shape = (int(rng.integers(4, 8)), int(rng.integers(4, 8)))
    k = int(rng.integers(2, 5))
    pool_st = [sh.Window(1), sh.VonNeumann(1), sh.Moore(1), sh.Positional((1, 0), (0, -1), (-1, 1))]
    floats = T in (np.float32, np.float64)
    pool_g = [sh.sum, sh.mean, sh.maximum, sh.minimum, sh.center] if floats else \
        [sh.maximum, sh.minimum, sh.center]
    arrays, sas, sts, bcs, terms = [], [], [], [], []
    for i in range(k):
        if floats:
            A = (rng.random(shape) * 2 - 1).astype(T)
        else:
            A = rng.integers(0, 256, shape).astype(T)
        st = pool_st[int(rng.integers(len(pool_st)))]
        bc = [sh.Remove(T(0.5) if floats else T(7)), sh.Wrap(), sh.Reflect()][int(rng.integers(3))]
        g = pool_g[int(rng.integers(len(pool_g)))]
        if floats and rng.random() < 0.4:
            g = sh.Scaled(float(T(rng.choice([0.5, 0.1, -3.0]))), g)
        arrays.append(A)
        sas.append(sh.StencilArray(np.asfortranarray(A.copy()), st, boundary=bc))
        sts.append(st)
        bcs.append(bc)
        terms.append(g)
    combine = "*" if rng.random() < 0.3 else "+"
    f = sh.Terms(*terms, combine=combine)
    r = oracle.mapstencil_multi(f, tuple(sas))
    b = _brute(terms, arrays, sts, bcs, combine, T)
    np.testing.assert_array_equal(np.ascontiguousarray(r).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))
