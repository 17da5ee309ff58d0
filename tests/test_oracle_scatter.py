"""The oracle's scatterstencil! pinned on the CPU.

The reference's own cases (test/array.jl:385-492) with the values they
assert; order-free closed forms (counts of reaching sources for constant
values that add exactly; maxima over the cells that reach a target) against
scipy and an independent enumeration of the boundary map; and, for the
order of floating-point `+`, a plain-Python restatement of the reference's
loop (passes over columns, rows ascending, offsets in order).
"""
import numpy as np
import pytest
import scipy.ndimage as ndi

import stencils_hip as sh
from stencils_hip import Scatter


def _sa(A, st, bc, pad=None):
    return sh.StencilArray(np.asfortranarray(A), st, boundary=bc, padding=pad or sh.Conditional())


def test_reference_basic_add(oracle):
    """test/array.jl:386-403: every cell scatters 0.1 to its Moore(1)
    neighbours under Remove(0): 8, 5 and 3 contributions."""
    dest = np.zeros((5, 5), order="F")
    oracle.scatterstencil(Scatter.constant(0.1), "+", dest, _sa(np.ones((5, 5)), sh.Moore(1), sh.Remove(0.0)))
    assert dest[2, 2] == pytest.approx(0.8) and dest[0, 2] == pytest.approx(0.5)
    assert dest[2, 0] == pytest.approx(0.5) and dest[0, 0] == pytest.approx(0.3) and dest[4, 4] == pytest.approx(0.3)


def test_reference_max(oracle):
    """test/array.jl:405-419: the centre value scattered with max."""
    src = np.array([[i + j for j in range(1, 6)] for i in range(1, 6)], dtype=np.float64)
    dest = np.zeros((5, 5), order="F")
    oracle.scatterstencil(Scatter.center(), "max", dest, _sa(src, sh.Moore(1), sh.Remove(0.0)))
    assert dest[2, 2] == 8.0 and dest[0, 0] == 4.0


def test_reference_water_flow(oracle):
    """test/array.jl:421-447: water flows to lower neighbours only."""
    water = np.ones((5, 5))
    dem = np.array([[abs(i - 3) + abs(j - 3) for j in range(1, 6)] for i in range(1, 6)], dtype=np.float64)
    dest = np.zeros((5, 5), order="F")
    oracle.scatterstencil(Scatter.downhill(0.1), "+", dest, _sa(water, sh.Moore(1), sh.Remove(0.0)),
                          _sa(dem, sh.Moore(1), sh.Remove(0.0)))
    assert dest[2, 2] > 0.0 and dest[0, 0] == 0.0 and dest[4, 4] == 0.0


def test_reference_vonneumann_and_radius_2(oracle):
    """test/array.jl:449-479."""
    dest = np.zeros((5, 5), order="F")
    oracle.scatterstencil(Scatter.constant(0.25), "+", dest, _sa(np.ones((5, 5)), sh.VonNeumann(1), sh.Remove(0.0)))
    assert dest[2, 2] == 1.0 and dest[0, 2] == 0.75 and dest[0, 0] == 0.5
    dest = np.zeros((7, 7), order="F")
    oracle.scatterstencil(Scatter.constant(0.01), "+", dest, _sa(np.ones((7, 7)), sh.Moore(2), sh.Remove(0.0)))
    assert dest[3, 3] == pytest.approx(0.24)


def _footprint(st):
    R = st.radius()
    fp = np.zeros((2 * R + 1, 2 * R + 1))
    for o in st.offsets():
        fp[o[0] + R, o[1] + R] = 1
    return fp


def _targets(i, j, st, bc, n0, n1):
    """Independent enumeration of where cell (i, j) sends (0-based), per the
    reference's _scatter_to! (src/scatterstencil.jl:78-114)."""
    out = []
    for o in st.offsets():
        t = []
        for p, n in ((i + o[0], n0), (j + o[1], n1)):
            if 0 <= p < n:
                t.append(p)
            elif isinstance(bc, sh.Wrap):
                t.append(p % n)
            elif isinstance(bc, sh.Reflect):
                t.append(-p if p < 0 else 2 * n - 2 - p)
            else:
                t = None
                break
        if t is not None:
            out.append(tuple(t))
    return out


@pytest.mark.parametrize("bc", [sh.Remove(0.0), sh.Wrap(), sh.Reflect()], ids=repr)
@pytest.mark.parametrize("st", [sh.Moore(1), sh.VonNeumann(2), sh.Positional((1, 0), (0, -1), (-1, 1))], ids=repr)
def test_order_free_closed_forms(oracle, st, bc):
    """Constant values that add exactly count the sources reaching each cell
    (Remove: scipy correlate of the flipped footprint); the max of centre
    values equals the max over the enumerated sources."""
    rng = np.random.default_rng(3)
    n0, n1 = 11, 9
    A = rng.random((n0, n1))
    dest = np.zeros((n0, n1), order="F")
    oracle.scatterstencil(Scatter.constant(0.25), "+", dest, _sa(A, st, bc))
    count = np.zeros((n0, n1))
    mx = np.full((n0, n1), -np.inf)
    for i in range(n0):
        for j in range(n1):
            for t in _targets(i, j, st, bc, n0, n1):
                count[t] += 1
                mx[t] = max(mx[t], A[i, j])
    np.testing.assert_array_equal(dest, 0.25 * count)
    if isinstance(bc, sh.Remove):
        fp = _footprint(st)[::-1, ::-1]
        np.testing.assert_array_equal(count, ndi.correlate(np.ones((n0, n1)), fp, mode="constant", cval=0.0))
    if isinstance(bc, sh.Wrap):
        assert (count == len(st)).all()
    dmax = np.full((n0, n1), -1.0, order="F")
    oracle.scatterstencil(Scatter.center(), "max", dmax, _sa(A, st, bc))
    np.testing.assert_array_equal(dmax, np.maximum(mx, -1.0))


def _loop(kind, coef, op, dest, A, B, st, bc, bbc):
    """The reference's loop in plain Python floats (src/scatterstencil.jl:51-75)."""
    n0, n1 = A.shape
    R = st.radius()
    S = 2 * R + 1
    offs = st.offsets()
    for p in range(S):
        for j in range(p, n1, S):
            for i in range(n0):
                for k, o in enumerate(offs):
                    if kind == "downhill_flow":
                        x, y = i + o[0], j + o[1]
                        if 0 <= x < n0 and 0 <= y < n1:
                            bk = B[x, y]
                        elif isinstance(bbc, sh.Remove):
                            bk = bbc.padval
                        else:
                            x = x % n0 if isinstance(bbc, sh.Wrap) else (-x if x < 0 else (2 * n0 - 2 - x if x >= n0 else x))
                            y = y % n1 if isinstance(bbc, sh.Wrap) else (-y if y < 0 else (2 * n1 - 2 - y if y >= n1 else y))
                            bk = B[x, y]
                        dz = B[i, j] - bk
                        v = (A[i, j] * dz) * coef if dz > 0 else 0.0
                    else:
                        v = A[i, j] * coef
                    t = _targets(i, j, sh.Positional(o), bc, n0, n1)
                    if t:
                        dest[t[0]] = dest[t[0]] + v if op == "+" else max(dest[t[0]], v)
    return dest


@pytest.mark.parametrize("bc", [sh.Remove(0.0), sh.Wrap(), sh.Reflect()], ids=repr)
def test_addition_order_against_the_loop(oracle, bc):
    """Float64 sums of inexact values in the reference's loop order, for the
    scaled and the downhill-flow rules (Moore(1) and a radius-2 shape)."""
    rng = np.random.default_rng(8)
    n0, n1 = 10, 13
    for st in (sh.Moore(1), sh.Circle(2)):
        A = rng.random((n0, n1))
        B = rng.random((n0, n1))
        d0 = rng.random((n0, n1))
        got = np.array(d0, order="F")
        oracle.scatterstencil(Scatter.scaled(0.1), "+", got, _sa(A, st, bc))
        np.testing.assert_array_equal(got, _loop("scaled", 0.1, "+", np.array(d0), A, B, st, bc, bc))
        got = np.array(d0, order="F")
        bbc = sh.Reflect() if isinstance(bc, sh.Remove) else sh.Remove(0.5)
        oracle.scatterstencil(Scatter.downhill_flow(0.3), "+", got, _sa(A, st, bc), _sa(B, st, bbc))
        np.testing.assert_array_equal(got, _loop("downhill_flow", 0.3, "+", np.array(d0), A, B, st, bc, bbc))
