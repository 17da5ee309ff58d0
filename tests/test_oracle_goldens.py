"""The oracle against the golden vectors and known answers of the
reference's own tests (test/array.jl, test/stencils.jl)."""
import numpy as np
import pytest

import stencils_hip as sh
from conftest import GOLDEN
from _spec import F, make_boundary, make_input, make_padding, make_stencil


def _cases():
    for c in GOLDEN["cases"]:
        for pad in c["paddings"]:
            yield pytest.param(c, pad, id=f"{c['name']}-{pad}")


@pytest.mark.parametrize("case,pad", list(_cases()))
def test_mapstencil_golden(oracle, case, pad):
    A = make_input(case)
    st = make_stencil(case["stencil"])
    sa = sh.StencilArray(A.copy(order="F"), st, boundary=make_boundary(case["boundary"], A.dtype),
                         padding=make_padding(pad))
    out = oracle.mapstencil(F[case["f"]], sa)
    exp = np.asarray(case["expected"])
    if case.get("interior"):
        k = case["interior"]
        out = out[tuple(slice(k, -k) for _ in range(out.ndim))]
    assert out.shape == exp.shape
    if case["tol"] is None:
        np.testing.assert_array_equal(out, exp)
    else:
        np.testing.assert_allclose(out, exp, rtol=0, atol=case["tol"])


@pytest.mark.parametrize("case", [c for c in GOLDEN["cases"] if len(c["paddings"]) > 1],
                         ids=lambda c: c["name"])
def test_conditional_equals_halo_and_switching(oracle, case):
    """A1 == B1 == SA1 == SB1 exactly (test/array.jl:139,156,181,200,228,247)."""
    A = make_input(case)
    st = make_stencil(case["stencil"])
    bc = make_boundary(case["boundary"], A.dtype)
    outs = []
    for pad in case["paddings"]:
        sa = sh.StencilArray(A.copy(order="F"), st, boundary=bc, padding=make_padding(pad))
        outs.append(oracle.mapstencil(F[case["f"]], sa))
        if F[case["f"]] is sh.mean or A.dtype.kind == "f":
            S = sh.SwitchingStencilArray(A.copy(order="F"), st, boundary=bc, padding=make_padding(pad))
            S1 = oracle.mapstencil_steps(F[case["f"]], S, 1)
            outs.append(np.asarray(S1))
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


def test_reflect_relations(oracle):
    """test/array.jl:255-258: Reflect interior relations across paddings."""
    A = make_input(GOLDEN["cases"][0])
    st = sh.Window(1, 2)
    A1 = oracle.mapstencil(sh.mean, sh.StencilArray(A.copy(order="F"), st, boundary=sh.Reflect()))
    C = sh.StencilArray(A.copy(order="F"), st, boundary=sh.Reflect(), padding=sh.Halo("in"))
    C1 = oracle.mapstencil(sh.mean, C)
    np.testing.assert_array_equal(A1[2:-2, 2:-2], C1[1:-1, 1:-1])
    np.testing.assert_array_equal(A1[1:-1, 1:-1], A[1:-1, 1:-1])
    assert not np.array_equal(A1[1:-1, 1:-1], C1)


@pytest.mark.parametrize("case", GOLDEN["points"], ids=lambda c: c["name"])
def test_point_known_answers(oracle, case):
    """sum(neighbors(A, I)) known answers from test/stencils.jl."""
    A = make_input(dict(case, dtype="i64"))
    st = make_stencil(case["stencil"])
    out = oracle.mapstencil(sh.sum, sh.StencilArray(A, st, boundary=sh.Remove(np.int64(0))))
    assert out[tuple(case["index"])] == case["sum"]
    # the neighbour list itself, one offset at a time
    idx = tuple(case["index"])
    for o, v in zip(st.offsets(), case["neighbors"]):
        one = sh.Positional(tuple(o))
        got = oracle.mapstencil(sh.sum, sh.StencilArray(A, one, boundary=sh.Remove(np.int64(0))))
        assert got[idx] == v


@pytest.mark.parametrize("case", GOLDEN["indices"], ids=lambda c: c["boundary"])
def test_boundary_index_maps(oracle, case):
    """indices(A, (1,1)) on a 4x4 Moore(1) array (test/array.jl:5-14): each
    neighbour's value encodes its (mapped) index, Remove reads the padval."""
    n = 4
    code = np.asfortranarray(np.array([[1000 + 10 * (i + 1) + (j + 1) for j in range(n)]
                                       for i in range(n)], dtype=np.int64))
    bc = {"Remove": sh.Remove(np.int64(-1)), "Wrap": sh.Wrap(), "Reflect": sh.Reflect()}[case["boundary"]]
    for o, e in zip(sh.Moore(1).offsets(), case["expected"]):
        got = oracle.mapstencil(sh.sum, sh.StencilArray(code, sh.Positional(tuple(o)), boundary=bc))[0, 0]
        i, j = e
        if 1 <= i <= n and 1 <= j <= n:
            assert got == 1000 + 10 * i + j
        else:
            assert case["boundary"] == "Remove" and got == -1


def test_kernelproduct_closed_forms(oracle):
    """test/stencils.jl:276-277 and :296-303."""
    k1, k2 = GOLDEN["kernelproduct"]
    X = np.asfortranarray(np.arange(1, 10, dtype=np.int64).reshape(3, 3, order="F"))
    K = sh.Kernel(sh.Window(1, 2), np.asarray(k1["weights"], dtype=np.int64))
    out = oracle.mapstencil(sh.kernelproduct, sh.StencilArray(X, K, boundary=sh.Remove(np.int64(0))))
    assert out[1, 1] == k1["expected"]
    W = np.asfortranarray(np.asarray(k2["window"], dtype=np.int64))
    st = make_stencil(k2["stencil"])
    K2 = sh.Kernel(st, np.asarray(k2["weights"], dtype=np.int64))
    out = oracle.mapstencil(sh.kernelproduct, sh.StencilArray(W, K2, boundary=sh.Remove(np.int64(0))))
    assert out[tuple(k2["index"])] == k2["expected"]


@pytest.mark.parametrize("case", GOLDEN["offsets"], ids=lambda c: c["src"])
def test_offset_tables(case):
    st = make_stencil(case["stencil"])
    assert [list(o) for o in st.offsets()] == case["offsets"]


def test_trailing_dims(oracle):
    """test/array.jl:81-86 and :109-118: lower-dimensional stencils on
    higher-dimensional arrays leave the trailing index unchanged."""
    rng = np.random.default_rng(7)
    r2 = np.asfortranarray(rng.random((20, 20)))
    a = oracle.mapstencil(sh.sum, sh.StencilArray(r2.copy(order="F"), sh.Window(1, 1), boundary=sh.Remove(0.0)))
    b = oracle.mapstencil(sh.sum, sh.StencilArray(r2.copy(order="F"), sh.Vertical(1, 2), boundary=sh.Remove(0.0)))
    np.testing.assert_array_equal(a, b)
    r3 = np.asfortranarray(rng.random((12, 12, 12)))
    pos = sh.Positional((-1, -1, 0), (0, -1, 0), (1, -1, 0), (-1, 0, 0), (0, 0, 0), (1, 0, 0),
                        (-1, 1, 0), (0, 1, 0), (1, 1, 0))
    assert sh.Window(1, 2).indices((10, 10, 10)) == pos.indices((10, 10, 10))
    a = oracle.mapstencil(sh.sum, sh.StencilArray(r3.copy(order="F"), sh.Window(1, 2), boundary=sh.Remove(0.0)))
    b = oracle.mapstencil(sh.sum, sh.StencilArray(r3.copy(order="F"), pos, boundary=sh.Remove(0.0)))
    np.testing.assert_array_equal(a, b)
