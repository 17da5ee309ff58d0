"""scatterstencil! on the GPU (sth_scatter) against the oracle and the
reference's own cases (test/array.jl:385-492).  Results are bit-identical:
each dest cell receives its values in the reference loop's order."""
import numpy as np
import pytest

import stencils_hip as sh
from stencils_hip import Scatter

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    t = sh.colmajor_empty(a.shape, torch.from_numpy(np.zeros(1, a.dtype)).dtype, torch.device("cuda"))
    t.copy_(torch.from_numpy(np.ascontiguousarray(a.T)).T)
    return t


def _host(t):
    return np.asfortranarray(t.cpu().numpy())


def test_reference_cases(gpu):
    """test/array.jl:386-491 through the GPU path (host arrays)."""
    ones = lambda n: np.ones((n, n), order="F")
    dest = np.zeros((5, 5), order="F")
    sh.scatterstencil_(Scatter.constant(0.1), "+", dest, sh.StencilArray(ones(5), sh.Moore(1), boundary=sh.Remove(0.0)))
    assert dest[2, 2] == pytest.approx(0.8) and dest[0, 2] == pytest.approx(0.5)
    assert dest[2, 0] == pytest.approx(0.5) and dest[0, 0] == pytest.approx(0.3) and dest[4, 4] == pytest.approx(0.3)
    src = np.asfortranarray([[i + j for j in range(1, 6)] for i in range(1, 6)], dtype=np.float64)
    dest = np.zeros((5, 5), order="F")
    sh.scatterstencil_(Scatter.center(), max, dest, sh.StencilArray(src, sh.Moore(1), boundary=sh.Remove(0.0)))
    assert dest[2, 2] == 8.0 and dest[0, 0] == 4.0
    dem = np.asfortranarray([[abs(i - 3) + abs(j - 3) for j in range(1, 6)] for i in range(1, 6)], dtype=np.float64)
    dest = np.zeros((5, 5), order="F")
    sh.scatterstencil_(Scatter.downhill(0.1), "+", dest, sh.StencilArray(ones(5), sh.Moore(1), boundary=sh.Remove(0.0)),
                       sh.StencilArray(dem, sh.Moore(1), boundary=sh.Remove(0.0)))
    assert dest[2, 2] > 0.0 and dest[0, 0] == 0.0 and dest[4, 4] == 0.0
    dest = np.zeros((5, 5), order="F")
    sh.scatterstencil_(Scatter.constant(0.25), "+", dest,
                       sh.StencilArray(ones(5), sh.VonNeumann(1), boundary=sh.Remove(0.0)))
    assert dest[2, 2] == 1.0 and dest[0, 2] == 0.75 and dest[0, 0] == 0.5
    dest = np.zeros((7, 7), order="F")
    sh.scatterstencil_(Scatter.constant(0.01), "+", dest, sh.StencilArray(ones(7), sh.Moore(2), boundary=sh.Remove(0.0)))
    assert dest[3, 3] == pytest.approx(0.24)
    S = sh.SwitchingStencilArray(ones(5), sh.Moore(1), boundary=sh.Remove(0.0))
    res = sh.scatterstencil_(Scatter.constant(0.1), "+", S)
    assert res is not S and np.asarray(res.source)[2, 2] == pytest.approx(0.8)


RULES = [Scatter.constant(0.25), Scatter.center(), Scatter.scaled(0.1), Scatter.downhill(0.3),
         Scatter.downhill_flow(0.7)]
STENCILS = [sh.Moore(1), sh.VonNeumann(2), sh.Circle(2), sh.Positional((1, 0), (0, -2), (-1, 1), (2, 2))]
BOUNDS = [sh.Remove(0.5), sh.Wrap(), sh.Reflect()]


@pytest.mark.parametrize("shape", [(5, 7), (33, 17), (130, 70)])
@pytest.mark.parametrize("dt", [np.float64, np.float32], ids=lambda d: np.dtype(d).name)
def test_random_against_oracle(gpu, oracle, dt, shape):
    """Every rule, op, boundary and stencil, Conditional and Halo padding of
    source and second array, starting from nonzero dest values: device
    results and refreshed Halo parents equal the oracle's bit for bit."""
    rng = np.random.default_rng(shape[0] * 7 + np.dtype(dt).itemsize)
    for trial in range(12):
        f = RULES[trial % len(RULES)]
        if dt == np.float32:
            f = Scatter(f.kind, float(np.float32(f.coef)))
        op = ["+", max, min][trial % 3]
        st = STENCILS[int(rng.integers(len(STENCILS)))]
        if max(max(abs(v) for v in o) for o in st.offsets()) >= min(shape):
            continue
        bc = BOUNDS[int(rng.integers(3))]
        bc = sh.Remove(dt(0.5)) if isinstance(bc, sh.Remove) else bc
        pad = [sh.Conditional(), sh.Halo()][int(rng.integers(2))]
        A = np.asfortranarray(rng.random(shape).astype(dt))
        B = np.asfortranarray(rng.random(shape).astype(dt))
        D = np.asfortranarray(rng.random(shape).astype(dt))
        bbc = [sh.Remove(dt(0.25)), sh.Wrap(), sh.Reflect()][int(rng.integers(3))]
        gsa = sh.StencilArray(_dev(A), st, boundary=bc, padding=pad)
        osa = sh.StencilArray(A.copy(order="F"), st, boundary=bc, padding=pad)
        gaux = sh.StencilArray(_dev(B), st, boundary=bbc, padding=pad) if f.needs_aux() else None
        oaux = sh.StencilArray(B.copy(order="F"), st, boundary=bbc, padding=pad) if f.needs_aux() else None
        gd = _dev(D)
        od = D.copy(order="F")
        sh.scatterstencil_(f, op, gd, gsa, *([gaux] if gaux is not None else []))
        oracle.scatterstencil(f, op, od, osa, oaux)
        msg = f"{f} {op} {st} {bc} {pad} {bbc}"
        np.testing.assert_array_equal(_host(gd), od, err_msg=msg)
        np.testing.assert_array_equal(_host(gsa.parent()), osa.parent(), err_msg=msg)


@pytest.mark.parametrize("st", [sh.Moore(3), sh.Moore(5), sh.Diamond(7)], ids=["moore3", "moore5", "diamond7"])
@pytest.mark.parametrize("bc", BOUNDS, ids=["remove", "wrap", "reflect"])
def test_large_radius_against_oracle(gpu, oracle, st, bc):
    """Radii 3 to 7 over several tiles in both axes: band cells R deep on
    every side under Wrap and Reflect, and order tables longer than the 1024
    entries each workgroup loads alongside its box (Moore(5): 120 offsets,
    Diamond(7): 113, 11 and 15 residues)."""
    rng = np.random.default_rng(len(st.offsets()))
    shape = (200, 150)
    for f, op in [(Scatter.downhill_flow(0.7), "+"), (Scatter.center(), max), (Scatter.scaled(0.1), "+")]:
        A = np.asfortranarray(rng.random(shape))
        B = np.asfortranarray(rng.random(shape))
        D = np.asfortranarray(rng.random(shape))
        gsa = sh.StencilArray(_dev(A), st, boundary=bc)
        osa = sh.StencilArray(A.copy(order="F"), st, boundary=bc)
        gaux = [sh.StencilArray(_dev(B), st, boundary=sh.Wrap())] if f.needs_aux() else []
        oaux = sh.StencilArray(B.copy(order="F"), st, boundary=sh.Wrap()) if f.needs_aux() else None
        gd = _dev(D)
        od = D.copy(order="F")
        sh.scatterstencil_(f, op, gd, gsa, *gaux)
        oracle.scatterstencil(f, op, od, osa, oaux)
        np.testing.assert_array_equal(_host(gd), od, err_msg=f"{f} {op} {st} {bc}")


@pytest.mark.parametrize("st", [sh.Moore(1), sh.VonNeumann(2), sh.Circle(2)], ids=repr)
@pytest.mark.parametrize("bc", BOUNDS, ids=["remove", "wrap", "reflect"])
def test_large_against_oracle(gpu, oracle, st, bc):
    """Calls of 2^20 cells and more run the kernel with each residue's
    interior order compiled in at run time: every rule and op at 1024^2 in
    Float64 and Float32 against the oracle, bit for bit."""
    rng = np.random.default_rng(len(st.offsets()) * 3 + 1)
    shape = (1024, 1024)
    cases = [(Scatter.downhill_flow(0.7), "+", np.float64), (Scatter.downhill(0.3), max, np.float64),
             (Scatter.center(), min, np.float32), (Scatter.scaled(0.1), "+", np.float32),
             (Scatter.constant(0.25), "+", np.float64)]
    for f, op, dt in cases:
        if dt == np.float32:
            f = Scatter(f.kind, float(np.float32(f.coef)))
        A = np.asfortranarray(rng.random(shape).astype(dt))
        B = np.asfortranarray(rng.random(shape).astype(dt))
        D = np.asfortranarray(rng.random(shape).astype(dt))
        gsa = sh.StencilArray(_dev(A), st, boundary=bc)
        osa = sh.StencilArray(A.copy(order="F"), st, boundary=bc)
        gaux = [sh.StencilArray(_dev(B), st, boundary=sh.Wrap())] if f.needs_aux() else []
        oaux = sh.StencilArray(B.copy(order="F"), st, boundary=sh.Wrap()) if f.needs_aux() else None
        gd = _dev(D)
        od = D.copy(order="F")
        sh.scatterstencil_(f, op, gd, gsa, *gaux)
        oracle.scatterstencil(f, op, od, osa, oaux)
        np.testing.assert_array_equal(_host(gd), od, err_msg=f"{f} {op} {st} {bc}")


def test_switching_forms(gpu, oracle):
    """scatterstencil!(f, op, A::SwitchingStencilArray) clears dest and
    scatters A's source (Conditional); the explicit-source form scatters
    another StencilArray into A's dest; a Halo-padded A has a parent larger
    than its source, which the reference's size check rejects."""
    rng = np.random.default_rng(12)
    U = np.asfortranarray(rng.random((40, 29)))
    S = sh.SwitchingStencilArray(_dev(U), sh.Moore(1), boundary=sh.Wrap())
    S2 = sh.scatterstencil_(Scatter.scaled(0.125), "+", S)
    ref = np.zeros_like(U, order="F")
    oracle.scatterstencil(Scatter.scaled(0.125), "+", ref, sh.StencilArray(U.copy(order="F"), sh.Moore(1),
                                                                             boundary=sh.Wrap()))
    np.testing.assert_array_equal(_host(S2.source), ref)
    V = np.asfortranarray(rng.random((40, 29)))
    S3 = sh.scatterstencil_(Scatter.center(), max, S2, sh.StencilArray(_dev(V), sh.VonNeumann(1), boundary=sh.Reflect()))
    ref = np.zeros_like(U, order="F")
    oracle.scatterstencil(Scatter.center(), max, ref, sh.StencilArray(V.copy(order="F"), sh.VonNeumann(1),
                                                                     boundary=sh.Reflect()))
    np.testing.assert_array_equal(_host(S3.source), ref)
    H = sh.SwitchingStencilArray(_dev(U), sh.Moore(1), boundary=sh.Wrap(), padding=sh.Halo())
    with pytest.raises(ValueError, match="sizes must match"):
        sh.scatterstencil_(Scatter.scaled(0.125), "+", H)


def test_scatter_errors(gpu):
    a = _dev(np.zeros((9, 8), np.float32))
    d = _dev(np.zeros((9, 8), np.float32))
    with pytest.raises(sh.Unsupported):   # 0.1 is no Float32: the values would promote
        sh.scatterstencil_(Scatter.scaled(0.1), "+", d, sh.StencilArray(a, sh.Moore(1)))
    with pytest.raises(sh.Unsupported):   # reach 3 beyond the signed-max radius 1
        sh.scatterstencil_(Scatter.constant(1.0), "+", d, sh.StencilArray(a, sh.Positional((-3, 0), (1, 0))))
    with pytest.raises(sh.Unsupported):
        sh.scatterstencil_(Scatter.constant(1.0), "*", d, sh.StencilArray(a, sh.Moore(1)))
    with pytest.raises(sh.Unsupported):   # the second array's stencil differs
        sh.scatterstencil_(Scatter.downhill(1.0), "+", d, sh.StencilArray(a, sh.Moore(1)),
                           sh.StencilArray(a, sh.VonNeumann(1)))
    with pytest.raises(ValueError, match="sizes must match"):
        sh.scatterstencil_(Scatter.constant(1.0), "+", _dev(np.zeros((9, 9), np.float32)), sh.StencilArray(a, sh.Moore(1)))
    a3 = _dev(np.zeros((5, 5, 5), np.float32))
    with pytest.raises(sh.Unsupported):
        sh.scatterstencil_(Scatter.constant(1.0), "+", _dev(np.zeros((5, 5, 5), np.float32)),
                           sh.StencilArray(a3, sh.Moore(1, 3)))
