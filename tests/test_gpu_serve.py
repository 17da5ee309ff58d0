"""Serving loop (sth_server_*): every request's result is the same alone and
in the mix -- bit for bit against the same steps run by themselves
(sth_iterate) and against the oracle -- in batched and serial modes, under a
budget that makes requests wait for memory."""
import numpy as np
import pytest

import stencils_hip as sh

pytestmark = pytest.mark.gpu


def _mixed_requests(seed, n):
    """Mixed 3-D heat (f32, bf16), 2-D mean and Life requests."""
    import torch
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda")
    out = []
    for i in range(n):
        kind = i % 5
        steps = int(rng.integers(0 if i == 3 else 1, 40))
        if kind in (0, 1):
            s = int(rng.integers(6, 40))
            shape = (s, int(rng.integers(6, 40)), int(rng.integers(6, 30)))
            tdt = torch.float32 if kind == 0 else torch.bfloat16
            A = sh.colmajor_empty(shape, tdt, dev)
            A.copy_(torch.rand(shape, device=dev).to(tdt))
            S = sh.SwitchingStencilArray(A, sh.VonNeumann(1, 3), boundary=sh.Wrap())
            f = sh.Laplacian(0.15)
        elif kind == 2:
            shape = (int(rng.integers(8, 90)), int(rng.integers(8, 60)))
            A = sh.colmajor_empty(shape, torch.float64, dev)
            A.copy_(torch.rand(shape, dtype=torch.float64, device=dev))
            S = sh.SwitchingStencilArray(A, sh.Window(1), boundary=sh.Reflect())
            f = sh.mean
        elif kind == 3:
            shape = (int(rng.integers(10, 30)),) * 3
            A = sh.colmajor_empty(shape, torch.float32, dev)
            A.copy_(torch.rand(shape, device=dev))
            S = sh.SwitchingStencilArray(A, sh.VonNeumann(1, 3), boundary=sh.Remove(np.float32(0)),
                                         padding=sh.Halo())
            f = sh.Laplacian(0.15)
        else:
            shape = (64 * int(rng.integers(1, 4)), int(rng.integers(8, 50)))
            A = (rng.random(shape) < 0.35).astype(np.uint8)
            S = sh.SwitchingStencilArray(A, sh.Moore(1), boundary=sh.Wrap(), padding=sh.Halo())
            f = sh.LifeRule()
        out.append((f, S, steps, 0.002 * i))
    return out


def _alone(f, S, steps):
    """The same request run by itself: `steps` mapstencil! on copies."""
    import torch
    from stencils_hip.array import colmajor_copy, is_device
    src = colmajor_copy(S.source, torch.cuda.current_device())
    dst = colmajor_copy(S.dest, torch.cuda.current_device())
    T = sh.SwitchingStencilArray(stencil=S.stencil, boundary=S.boundary, padding=S.padding,
                                 _buffers=(src, dst))
    T = sh.mapstencil_steps(f, T, steps, graph=False)
    return T.source


def _serve(reqs, **kw):
    srv = sh.Server(**kw)
    tags = {}
    for f, S, steps, arr in reqs:
        tags[srv.submit(f, S, steps, arrival=arr)] = (f, S, steps)
    srv.start()
    srv.drain(timeout=120)
    done = srv.completions()
    stats = srv.stats()
    srv.close()
    return done, tags, stats


@pytest.mark.parametrize("serial", [False, True])
def test_served_equals_alone(gpu, serial):
    reqs = _mixed_requests(1, 15)
    done, tags, stats = _serve(reqs, hbm_budget=1 << 28, serial=serial, round_steps=4)
    assert len(done) == len(reqs) and stats["completed"] == len(reqs)
    for c in done:
        f, S, steps = tags[c.tag]
        import torch
        ref = _alone(f, S, steps)
        assert c.result.dtype == ref.dtype and torch.equal(c.result, ref), (S, steps)  # halo included
        assert c.finished >= c.admitted >= c.arrival - 1e-9


def test_served_matches_oracle(gpu, oracle):
    from stencils_hip.array import to_numpy
    reqs = [r for r in _mixed_requests(2, 10) if r[1].sth_dtype() != sh._abi.BF16]
    done, tags, _ = _serve(reqs, hbm_budget=1 << 28, round_steps=3)
    for c in done:
        f, S, steps = tags[c.tag]
        H = sh.SwitchingStencilArray(stencil=S.stencil, boundary=S.boundary, padding=S.padding,
                                     _buffers=(np.array(to_numpy(S.source), order="F"),
                                               np.array(to_numpy(S.dest), order="F")))
        H = oracle.mapstencil_steps(f, H, steps)
        got = c.result.cpu().numpy()
        assert np.array_equal(got, np.asarray(H.source)), (S, steps)  # halo included


def test_budget_admission(gpu):
    """A budget that holds about two requests: they queue for memory, all
    finish, the arena never exceeds the budget and results stay exact."""
    import torch
    dev = torch.device("cuda")
    reqs = []
    for i in range(6):
        A = sh.colmajor_empty((48, 48, 48), torch.float32, dev)
        A.copy_(torch.rand((48, 48, 48), device=dev))
        reqs.append((sh.Laplacian(0.15), sh.SwitchingStencilArray(A, sh.VonNeumann(1, 3), boundary=sh.Wrap()),
                     25 + 5 * i, 0.0))
    one = 48 * 48 * 48 * 4
    budget = 5 * one  # two requests of two buffers each
    done, tags, stats = _serve(reqs, hbm_budget=budget, round_steps=4)
    assert len(done) == 6
    assert stats["hbm_peak"] <= budget
    assert stats["hbm_peak"] >= 2 * one
    admitted = sorted(c.admitted for c in done)
    assert admitted[-1] > admitted[0]  # later requests waited for memory
    for c in done:
        f, S, steps = tags[c.tag]
        assert torch.equal(c.result, _alone(f, S, steps))


def test_large_requests_two_step_launches(gpu):
    """Requests that fill the device run their own two-step launches mixed
    with batched small ones; results (halo included) equal the step-by-step
    path alone."""
    import torch
    dev = torch.device("cuda")
    reqs = []
    for i, (shape, bc, pad, steps) in enumerate((((256, 256, 128), sh.Wrap(), sh.Conditional(), 37),
                                                 ((250, 260, 130), sh.Remove(np.float32(0.25)), sh.Halo(), 30),
                                                 ((256, 256, 128), sh.Wrap(), sh.Conditional(), 2),
                                                 ((256, 256, 128), sh.Wrap(), sh.Conditional(), 5))):
        A = sh.colmajor_empty(shape, torch.float32, dev)
        A.copy_(torch.rand(shape, device=dev))
        reqs.append((sh.Laplacian(0.15), sh.SwitchingStencilArray(A, sh.VonNeumann(1, 3), boundary=bc,
                                                                  padding=pad), steps, 0.0))
    # 2-D requests of 2^21 cells with two-step kernels: heat (f32) and Life
    # with bytes other than 0 / 1 (the first served step must normalise them)
    H = sh.colmajor_empty((2048, 1024), torch.float32, dev)
    H.copy_(torch.rand((2048, 1024), device=dev))
    reqs.append((sh.Laplacian(0.2), sh.SwitchingStencilArray(H, sh.VonNeumann(1), boundary=sh.Wrap()), 23, 0.0))
    rng = np.random.default_rng(4)
    G = (rng.random((2048, 1024)) < 0.3).astype(np.uint8) * rng.integers(1, 6, (2048, 1024)).astype(np.uint8)
    Gt = sh.colmajor_empty((2048, 1024), torch.uint8, dev)
    Gt.copy_(torch.from_numpy(G))
    reqs.append((sh.LifeRule(), sh.SwitchingStencilArray(Gt, sh.Moore(1), boundary=sh.Wrap(), padding=sh.Halo()),
                 21, 0.0))
    reqs += _mixed_requests(3, 6)
    done, tags, stats = _serve(reqs, hbm_budget=1 << 30, round_steps=4)
    assert len(done) == len(reqs)
    for c in done:
        f, S, steps = tags[c.tag]
        ref = _alone(f, S, steps)
        assert torch.equal(c.result, ref), (S, steps)  # halo included


def test_served_2d_requests_run_solo_two_step_kernels(gpu, oracle):
    """2-D requests of 2^21 cells with multi-step kernels (Float32 heat,
    Conditional; Life on a Halo-padded grid whose cells hold bytes 0..5)
    run as solo launches, most of them advancing several steps (the Halo
    interior sits on 16 bytes in the arena, so the vector kernels apply; the
    first served Life step normalises the bytes); results, halo included,
    equal the step-by-step oracle."""
    import torch
    dev = torch.device("cuda")
    rng = np.random.default_rng(4)
    H = (rng.random((2048, 1024)) * 2 - 1).astype(np.float32)
    G = (rng.random((2048, 1024)) < 0.3).astype(np.uint8) * rng.integers(1, 6, (2048, 1024)).astype(np.uint8)
    cases = [(sh.Laplacian(0.2), H, sh.VonNeumann(1), sh.Conditional(), 23),
             (sh.LifeRule(), G, sh.Moore(1), sh.Halo(), 21)]
    reqs, hosts = [], []
    for f, A, st, pad, steps in cases:
        t = sh.colmajor_empty(A.shape, torch.from_numpy(A[:1, :1]).dtype, dev)
        t.copy_(torch.from_numpy(A))
        reqs.append((f, sh.SwitchingStencilArray(t, st, boundary=sh.Wrap(), padding=pad), steps, 0.0))
        hosts.append(sh.SwitchingStencilArray(np.asfortranarray(A), st, boundary=sh.Wrap(), padding=pad))
    done, tags, stats = _serve(reqs, hbm_budget=1 << 30, round_steps=4)
    assert len(done) == 2
    assert stats["fused_launches"] >= 2 and stats["solo_launches"] < 23 + 21, stats
    for c in done:
        f, S, steps = tags[c.tag]
        k = next(i for i, r in enumerate(reqs) if r[1] is S)
        O = oracle.mapstencil_steps(f, hosts[k], steps)
        np.testing.assert_array_equal(c.result.cpu().numpy(), O.source, err_msg=f"{f}")


def test_served_integer_requests_run_runtime_compiled_two_step_kernels(gpu, oracle):
    """Int32 requests of 2^21 cells (maximum over VonNeumann(1), Wrap; a
    user offset list under Reflect) have no build-time two-step kernel: they
    run solo on the two-step kernels compiled at run time, and their results
    equal the step-by-step oracle."""
    import torch
    dev = torch.device("cuda")
    rng = np.random.default_rng(6)
    A = rng.integers(-500, 500, (2048, 1024)).astype(np.int32)
    B = rng.integers(-500, 500, (1536, 1536)).astype(np.int32)
    cases = [(sh.maximum, A, sh.VonNeumann(1), sh.Wrap(), 22),
             (sh.minimum, B, sh.Positional((-1, 1), (0, 0), (1, 0), (0, -1)), sh.Reflect(), 19)]
    reqs, hosts = [], []
    for f, X, st, bc, steps in cases:
        t = sh.colmajor_empty(X.shape, torch.int32, dev)
        t.copy_(torch.from_numpy(X))
        reqs.append((f, sh.SwitchingStencilArray(t, st, boundary=bc), steps, 0.0))
        hosts.append(sh.SwitchingStencilArray(np.asfortranarray(X), st, boundary=bc))
    done, tags, stats = _serve(reqs, hbm_budget=1 << 30, round_steps=4)
    assert len(done) == 2
    assert stats["fused_launches"] >= 2, stats
    for c in done:
        f, S, steps = tags[c.tag]
        k = next(i for i, r in enumerate(reqs) if r[1] is S)
        O = oracle.mapstencil_steps(f, hosts[k], steps)
        np.testing.assert_array_equal(c.result.cpu().numpy(), O.source, err_msg=f"{f}")


def test_invalid_request_rejected_at_submit(gpu):
    """A request the library cannot run (reach >= axis, the reference's
    ArgumentError) is rejected when submitted, through the C ABI past the
    Python-side checks; the server keeps serving the others."""
    import ctypes
    import torch
    from stencils_hip import _abi, _lib
    from stencils_hip.gather import _Call, stencil_descriptor
    dev = torch.device("cuda")
    srv = sh.Server(hbm_budget=1 << 26)
    good = sh.colmajor_empty((20, 20, 20), torch.float32, dev)
    good.copy_(torch.rand((20, 20, 20), device=dev))
    S = sh.SwitchingStencilArray(good, sh.VonNeumann(1, 3), boundary=sh.Wrap())
    tag_ok = srv.submit(sh.Laplacian(0.15), S, 5)
    bad = sh.colmajor_empty((6, 6, 1), torch.float32, dev)
    bad.zero_()
    out = sh.colmajor_empty((6, 6, 1), torch.float32, dev)
    keep = _Call()
    q = _abi.SthRequest()
    q.state = _abi.make_array(bad.data_ptr(), _abi.F32, (6, 6, 1), (1, 6, 36), 0)
    q.stencil = stencil_descriptor(sh.VonNeumann(1, 3), _abi.F32, keep)
    q.op = sh.Laplacian(0.15).op(0)
    q.boundary, q.padding, q.steps = sh.Wrap().code, _abi.CONDITIONAL, 3
    q.result = out.data_ptr()
    q.tag = 999
    with pytest.raises(ValueError, match="radius"):
        _lib.check(_lib.lib().sth_server_submit(srv._h, ctypes.byref(q)))
    srv.start()
    srv.drain(timeout=60)
    done = {c.tag: c for c in srv.completions()}
    srv.close()
    assert set(done) == {tag_ok} and done[tag_ok].status == _abi.STH_OK
    assert torch.equal(done[tag_ok].result, _alone(sh.Laplacian(0.15), S, 5))


def test_request_errors(gpu):
    import torch
    dev = torch.device("cuda")
    srv = sh.Server(hbm_budget=1 << 20)
    A = sh.colmajor_empty((64, 64, 64), torch.float32, dev)
    S = sh.SwitchingStencilArray(A, sh.VonNeumann(1, 3), boundary=sh.Wrap())
    with pytest.raises(sh.StencilsHipError):   # two 1 MiB buffers exceed the budget
        srv.submit(sh.Laplacian(0.15), S, 3)
    B = sh.colmajor_empty((4, 4, 4), torch.float32, dev)
    T = sh.SwitchingStencilArray(B, sh.VonNeumann(1, 3), boundary=sh.Wrap())
    with pytest.raises(ValueError):
        srv.submit(sh.Laplacian(0.15), T, -1)
    srv.close()
