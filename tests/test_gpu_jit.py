"""Runtime-compiled specialisations (hipRTC) of the streaming kernel against
the oracle: user offset lists and integer / Bool element types on arrays of
2^20 cells and more (the size from which calls outside the build-time
catalogue compile one), every boundary and padding, bit for bit."""
import numpy as np
import pytest

import stencils_hip as sh

pytestmark = pytest.mark.gpu

N = (1024, 1024)  # 2^20 cells: the default size from which calls compile a specialisation


def _input(dtype, seed):
    rng = np.random.default_rng(seed)
    if dtype == np.bool_:
        return np.asfortranarray(rng.random(N) < 0.4)
    if np.dtype(dtype).kind in "iu":
        info = np.iinfo(dtype)
        return np.asfortranarray(rng.integers(max(info.min, -50), min(info.max, 50) + 1, N).astype(dtype))
    return np.asfortranarray((rng.random(N) * 2 - 1).astype(dtype))


def _remove(dt):
    return sh.Remove(np.bool_(True) if dt == np.bool_ else np.dtype(dt).type(3))


CASES = [
    # (stencil, dtype, fold)
    (sh.Positional((-2, 0), (0, -1), (0, 0), (1, 1), (2, -2), (-1, 2)), np.float32, sh.sum),
    (sh.Positional((-2, 0), (0, -1), (0, 0), (1, 1), (2, -2), (-1, 2)), np.float64, sh.mean),
    (sh.Positional((1, 0), (0, 3), (-3, -1)), np.float64, sh.maximum),
    (sh.Moore(1), np.int64, sh.sum),
    (sh.Moore(2), np.uint8, sh.maximum),
    (sh.VonNeumann(2), np.int32, sh.minimum),
    (sh.Positional((-1, -1), (1, 1), (0, 2)), np.bool_, sh.sum),
    (sh.Cross(2), np.int16, sh.maximum),
    (sh.Annulus(3), np.float32, sh.minimum),
]


@pytest.mark.parametrize("pad", ["cond", "halo"])
@pytest.mark.parametrize("bname", ["remove", "wrap", "reflect"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]!r}-{np.dtype(c[1]).name}-{getattr(c[2], 'name', c[2])!s}")
def test_specialised_against_oracle(gpu, oracle, case, bname, pad):
    st, dt, f = case
    A = _input(dt, len(st.offsets()) + np.dtype(dt).itemsize)
    bc = {"remove": _remove(dt), "wrap": sh.Wrap(), "reflect": sh.Reflect()}[bname]
    padding = sh.Conditional() if pad == "cond" else sh.Halo()
    ga = sh.StencilArray(A.copy(order="F"), st, boundary=bc, padding=padding)
    oa = sh.StencilArray(A.copy(order="F"), st, boundary=bc, padding=padding)
    g = sh.mapstencil(f, ga)
    o = oracle.mapstencil(f, oa)
    np.testing.assert_array_equal(g, o)
    np.testing.assert_array_equal(ga.parent(), oa.parent())


def test_specialisation_is_what_runs(gpu):
    """A call outside the catalogue on 2^20 cells lists a runtime-compiled
    variant in the tuner's table, and the tuner chose between it and the
    runtime-offset kernel."""
    A = _input(np.float32, 5)
    st = sh.Positional((3, 0), (0, 1), (-1, -2))
    sh.mapstencil(sh.sum, sh.StencilArray(A, st, boundary=sh.Wrap()))
    table = sh._lib.tuning_table()
    fams = [k for k in table if k.startswith("jit_")]
    assert fams, table
    assert all(table[k].startswith(("jit_", "rt_")) for k in fams), table


def test_specialised_iteration(gpu, oracle):
    """A SwitchingStencilArray iterated 9 steps through sth_iterate (graph
    replay of the module launches) equals the oracle's steps."""
    import torch
    A = _input(np.float32, 11)
    st = sh.Positional((-1, 0), (1, 0), (0, -2), (0, 2), (0, 0))
    t = sh.colmajor_empty(N, torch.float32, torch.device("cuda"))
    t.copy_(torch.from_numpy(np.ascontiguousarray(A.T)).T)
    S = sh.SwitchingStencilArray(t, st, boundary=sh.Wrap())
    S = sh.mapstencil_steps(sh.mean, S, 9, graph=True)
    O = oracle.mapstencil_steps(sh.mean, sh.SwitchingStencilArray(A.copy(order="F"), st, boundary=sh.Wrap()), 9)
    np.testing.assert_array_equal(np.asfortranarray(S.source.cpu().numpy()), np.asarray(O.source))


FUSED = [
    # (stencil, dtype, fold, shape): two steps per launch, runtime-compiled (2^21 cells and more)
    (sh.VonNeumann(1), np.int32, sh.maximum, (2048, 1536)),
    (sh.Moore(1), np.int64, sh.sum, (1536, 1536)),
    (sh.Positional((-1, 0), (0, 0), (1, 1), (0, -1)), np.float32, sh.mean, (2048, 1024)),
    (sh.VonNeumann(1, 3), np.int32, sh.minimum, (128, 128, 128)),
    (sh.Moore(1, 3), np.int16, sh.maximum, (128, 136, 128)),
]


@pytest.mark.parametrize("pad", ["cond", "halo"])
@pytest.mark.parametrize("bname", ["remove", "wrap", "reflect"])
@pytest.mark.parametrize("case", FUSED, ids=lambda c: f"{c[0]!r}-{np.dtype(c[1]).name}-{len(c[3])}d")
def test_specialised_two_step_iteration(gpu, oracle, case, bname, pad):
    """Iterations of radius-1 stencils outside the catalogue (integer types,
    a user offset list) run the two-step kernels compiled at run time: 13
    steps through sth_iterate (graph replay), both buffers equal to the
    oracle's step-by-step iteration."""
    import torch
    st, dt, f, shape = case
    rng = np.random.default_rng(len(shape) * 10 + np.dtype(dt).itemsize)
    if np.dtype(dt).kind == "f":
        A = np.asfortranarray(rng.random(shape).astype(dt))
    else:
        A = np.asfortranarray(rng.integers(-60, 60, shape).astype(dt))
    bc = {"remove": sh.Remove(np.dtype(dt).type(5)), "wrap": sh.Wrap(), "reflect": sh.Reflect()}[bname]
    t = sh.colmajor_empty(shape, torch.from_numpy(np.zeros(1, dt)).dtype, torch.device("cuda"))
    rev = tuple(reversed(range(len(shape))))
    t.copy_(torch.from_numpy(np.ascontiguousarray(A.transpose(rev))).permute(*rev))
    padding = sh.Conditional() if pad == "cond" else sh.Halo()
    S = sh.mapstencil_steps(f, sh.SwitchingStencilArray(t, st, boundary=bc, padding=padding), 13, graph=True)
    O = oracle.mapstencil_steps(f, sh.SwitchingStencilArray(A.copy(order="F"), st, boundary=bc, padding=padding), 13)
    np.testing.assert_array_equal(np.asfortranarray(S.source.cpu().numpy()), np.asarray(O.source))
    np.testing.assert_array_equal(np.asfortranarray(S.dest.cpu().numpy()), np.asarray(O.dest))
