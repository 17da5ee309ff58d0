"""sth_iterate_batch: many SwitchingStencilArrays iterated with their own
step counts in one call (one batched launch set per round), against the
oracle's step-by-step iteration of each array alone: both buffers, halos
included, bit for bit; with and without graph capture and replay."""
import numpy as np
import pytest

import stencils_hip as sh

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(12)

    def f(shape, dt):
        if dt == np.uint8:
            return np.asfortranarray((rng.random(shape) < 0.3).astype(np.uint8))
        if np.dtype(dt).kind == "i":
            return np.asfortranarray(rng.integers(-40, 40, shape).astype(dt))
        return np.asfortranarray(rng.random(shape).astype(dt))

    return [
        # (f, A host, stencil, boundary, padding, steps)
        (sh.Laplacian(0.2), f((512, 384), np.float32), sh.VonNeumann(1), sh.Wrap(), sh.Conditional(), 37),
        (sh.mean, f((300, 200), np.float64), sh.Window(1), sh.Remove(0.0), sh.Conditional(), 0),
        (sh.maximum, f((64, 80), np.float64), sh.Moore(2), sh.Reflect(), sh.Halo(), 5),
        (sh.LifeRule(), f((256, 256), np.uint8), sh.Moore(1), sh.Wrap(), sh.Halo(), 40),
        (sh.maximum, f((120, 90), np.int32), sh.VonNeumann(2), sh.Wrap(), sh.Conditional(), 17),
        (sh.Laplacian(0.1), f((2048, 1024), np.float32), sh.VonNeumann(1), sh.Remove(np.float32(0)),
         sh.Conditional(), 12),
        (sh.mean, f((100, 100), np.float32), sh.Positional((1, 0), (0, -1), (-2, 2)), sh.Wrap(),
         sh.Conditional(), 9),
        (sh.Laplacian(0.15), f((40, 30, 20), np.float64), sh.VonNeumann(1, 3), sh.Wrap(), sh.Conditional(), 6),
        (sh.minimum, f((33, 47), np.float32), sh.Circle(2), sh.Remove(np.float32(0.5)), sh.Halo(), 1),
    ]


def _device(A, st, bc, pad):
    import torch
    t = sh.colmajor_empty(A.shape, torch.from_numpy(np.zeros(1, A.dtype)).dtype, torch.device("cuda"))
    t.copy_(torch.from_numpy(np.ascontiguousarray(A.T)).T)
    return sh.SwitchingStencilArray(t, st, boundary=bc, padding=pad)


def _host(t):
    return np.asfortranarray(t.cpu().numpy())


def _check(res, oracle_arrays):
    for i, (R, O) in enumerate(zip(res, oracle_arrays)):
        np.testing.assert_array_equal(_host(R.source), np.asarray(O.source), err_msg=f"array {i} source")
        np.testing.assert_array_equal(_host(R.dest), np.asarray(O.dest), err_msg=f"array {i} dest")


def _oracle(oracle, cases, times=1):
    out = []
    for f, A, st, bc, pad, k in cases:
        O = sh.SwitchingStencilArray(A.copy(order="F"), st, boundary=bc, padding=pad)
        for _ in range(times):
            O = oracle.mapstencil_steps(f, O, k)
        out.append(O)
    return out


@pytest.mark.parametrize("mode", ["plain", "graph", "serial"])
def test_uneven_steps_against_oracle(gpu, oracle, mode):
    cases = _cases()
    arrays = [_device(A, st, bc, pad) for _, A, st, bc, pad, _ in cases]
    res = sh.mapstencil_steps_batch([(c[0], a, c[5]) for c, a in zip(cases, arrays)],
                                    graph=mode == "graph", serial=mode == "serial")
    _check(res, _oracle(oracle, cases))


def test_graph_replay(gpu, oracle):
    """Even step counts leave every array's roles as they were, so the second
    call has the same key and replays the captured schedule; the arrays then
    hold twice the steps."""
    cases = [c[:5] + (2 * (c[5] // 2 + 1),) for c in _cases()]
    arrays = [_device(A, st, bc, pad) for _, A, st, bc, pad, _ in cases]
    items = [(c[0], a, c[5]) for c, a in zip(cases, arrays)]
    res = sh.mapstencil_steps_batch(items, graph=True)
    assert all(r is a for r, a in zip(res, arrays))
    res = sh.mapstencil_steps_batch(items, graph=True)
    _check(res, _oracle(oracle, cases, times=2))


def test_matches_single_iterations(gpu):
    """Each array of the batch equals mapstencil_steps on that array alone."""
    cases = _cases()
    a1 = [_device(A, st, bc, pad) for _, A, st, bc, pad, _ in cases]
    a2 = [_device(A, st, bc, pad) for _, A, st, bc, pad, _ in cases]
    res = sh.mapstencil_steps_batch([(c[0], a, c[5]) for c, a in zip(cases, a1)])
    for c, r, a in zip(cases, res, a2):
        s = sh.mapstencil_steps(c[0], a, c[5])
        np.testing.assert_array_equal(_host(r.source), _host(s.source))
        np.testing.assert_array_equal(_host(r.dest), _host(s.dest))


def test_errors(gpu):
    A = _device(np.zeros((8, 8), np.float32), sh.Moore(1), sh.Wrap(), sh.Conditional())
    with pytest.raises(ValueError):
        sh.mapstencil_steps_batch([(sh.sum, A, -1)])
    assert sh.mapstencil_steps_batch([]) == []
