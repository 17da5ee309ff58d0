"""Host-side logic and the C ABI surface, without a GPU: the library loads and
exports every entry point the header declares, the ctypes descriptors match
the C layout, and the Python mirror reproduces the reference's sizes,
padding layouts, result types and argument errors."""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import stencils_hip as sh
from stencils_hip import _abi, _lib

REPO = Path(__file__).resolve().parent.parent
HEADER = REPO / "include" / "stencils_hip.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sth_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.load_library()
    syms = declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(L, s), f"{s} missing from libstencils_hip.so"
        assert s in _lib.EXPORTS, f"{s} has no ctypes prototype"
    assert L.sth_version() // 10000 == 1
    assert L.sth_last_error() == b""


def test_ctypes_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "stencils_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(sth_array), offsetof(sth_array, size),
         offsetof(sth_array, stride), offsetof(sth_array, halo), offsetof(sth_array, dtype));
  printf("%zu %zu %zu %zu\n", sizeof(sth_stencil), offsetof(sth_stencil, offsets),
         offsetof(sth_stencil, weights), offsetof(sth_stencil, weights_dtype));
  printf("%zu %zu %zu %zu\n", sizeof(sth_op), offsetof(sth_op, alpha),
         offsetof(sth_op, birth_mask), offsetof(sth_op, survive_mask));
  printf("%zu %zu %zu %zu %zu\n", sizeof(sth_job), offsetof(sth_job, stencil),
         offsetof(sth_job, op), offsetof(sth_job, padval), offsetof(sth_job, padding));
  printf("%zu %zu %zu %zu\n", sizeof(sth_server_config), offsetof(sth_server_config, max_active),
         offsetof(sth_server_config, round_steps), offsetof(sth_server_config, flags));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(sth_request), offsetof(sth_request, stencil),
         offsetof(sth_request, op), offsetof(sth_request, steps), offsetof(sth_request, arrival),
         offsetof(sth_request, tag));
  printf("%zu %zu %zu %zu\n", sizeof(sth_completion), offsetof(sth_completion, arrival),
         offsetof(sth_completion, finished), offsetof(sth_completion, cells));
  printf("%zu %zu %zu %zu\n", sizeof(sth_server_stats), offsetof(sth_server_stats, hbm_in_use),
         offsetof(sth_server_stats, started), offsetof(sth_server_stats, fused_launches));
  printf("%zu %zu\n", sizeof(sth_scatter_rule), offsetof(sth_scatter_rule, coef));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(sth_term), offsetof(sth_term, stencil), offsetof(sth_term, op),
         offsetof(sth_term, padval), offsetof(sth_term, coef), offsetof(sth_term, scaled));
  return 0;
}''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(REPO / "include"), str(src), "-o", str(exe)], check=True)
    out = [list(map(int, l.split())) for l in subprocess.check_output([str(exe)]).decode().splitlines()]
    A, S, O = _abi.SthArray, _abi.SthStencil, _abi.SthOp
    assert out[0] == [ctypes.sizeof(A), A.size.offset, A.stride.offset, A.halo.offset, A.dtype.offset]
    assert out[1] == [ctypes.sizeof(S), S.offsets.offset, S.weights.offset, S.weights_dtype.offset]
    assert out[2] == [ctypes.sizeof(O), O.alpha.offset, O.birth_mask.offset, O.survive_mask.offset]
    J = _abi.SthJob
    assert out[3] == [ctypes.sizeof(J), J.stencil.offset, J.op.offset, J.padval.offset, J.padding.offset]
    C = _abi.SthServerConfig
    assert out[4] == [ctypes.sizeof(C), C.max_active.offset, C.round_steps.offset, C.flags.offset]
    Q = _abi.SthRequest
    assert out[5] == [ctypes.sizeof(Q), Q.stencil.offset, Q.op.offset, Q.steps.offset, Q.arrival.offset,
                      Q.tag.offset]
    M = _abi.SthCompletion
    assert out[6] == [ctypes.sizeof(M), M.arrival.offset, M.finished.offset, M.cells.offset]
    T = _abi.SthServerStats
    assert out[7] == [ctypes.sizeof(T), T.hbm_in_use.offset, T.started.offset, T.fused_launches.offset]
    SR = _abi.SthScatterRule
    assert out[8] == [ctypes.sizeof(SR), SR.coef.offset]
    M2 = _abi.SthTerm
    assert out[9] == [ctypes.sizeof(M2), M2.stencil.offset, M2.op.offset, M2.padval.offset, M2.coef.offset,
                      M2.scaled.offset]


def test_header_enums_match_python():
    text = HEADER.read_text()
    vals = dict(re.findall(r"\b(STH_\w+)\s*=\s*(\d+)", text))
    assert int(vals["STH_F64"]) == _abi.F64 and int(vals["STH_BF16"]) == _abi.BF16
    assert int(vals["STH_WRAP"]) == _abi.WRAP and int(vals["STH_HALO_IN"]) == _abi.HALO_IN
    assert int(vals["STH_OP_LAPLACIAN"]) == _abi.OP_LAPLACIAN
    assert int(vals["STH_ENOTSUP"]) == _abi.STH_ENOTSUP


# ------------------------------------------------------- sizes and padding

def test_sizes_1d():
    """test/array.jl:16-43."""
    r = np.random.default_rng(0).random(100)
    A = sh.StencilArray(r, sh.VonNeumann(3, 1), padding=sh.Conditional(), boundary=sh.Remove(0.0))
    B = sh.StencilArray(r, sh.Window(10, 1), padding=sh.Halo("out"), boundary=sh.Remove(0.0))
    C = sh.StencilArray(r, sh.Moore(10, 1), padding=sh.Halo("in"), boundary=sh.Remove(0.0))
    sA = sh.SwitchingStencilArray(r, sh.Window(10, 1), padding=sh.Conditional(), boundary=sh.Remove(0.0))
    sB = sh.SwitchingStencilArray(r, sh.Window(10, 1), padding=sh.Halo("out"), boundary=sh.Remove(0.0))
    sC = sh.SwitchingStencilArray(r, sh.Window(10, 1), padding=sh.Halo("in"), boundary=sh.Remove(0.0))
    assert sA.source is not sA.dest
    assert A.size() == A.parent().shape == sA.size() == sA.parent().shape == (100,)
    assert B.size() == sB.size() == (100,)
    assert B.parent().shape == sB.parent().shape == (120,)
    assert C.size() == sC.size() == (80,)
    assert C.parent().shape == sC.parent().shape == (100,)


def test_sizes_2d_3d():
    """test/array.jl:45-107."""
    r = np.random.default_rng(0).random((100, 100))
    B = sh.StencilArray(r.copy(), sh.Window(10), padding=sh.Halo("out"), boundary=sh.Remove(0.0))
    C = sh.StencilArray(r.copy(), sh.Moore(10), padding=sh.Halo("in"), boundary=sh.Remove(0.0))
    assert B.size() == (100, 100) and B.parent().shape == (120, 120)
    assert C.size() == (80, 80) and C.parent().shape == (100, 100)
    r3 = np.random.default_rng(0).random((30, 30, 30))
    B3 = sh.StencilArray(r3, sh.Window(10, 3), padding=sh.Halo("out"), boundary=sh.Remove(0.0))
    C3 = sh.StencilArray(r3, sh.Moore(10, 3), padding=sh.Halo("in"), boundary=sh.Remove(0.0))
    assert B3.size() == (30, 30, 30) and B3.parent().shape == (50, 50, 50)
    assert C3.size() == (10, 10, 10)


def test_halo_out_copies_interior():
    r = np.arange(12.0).reshape(3, 4, order="F")
    B = sh.StencilArray(r, sh.Window(1), padding=sh.Halo(), boundary=sh.Wrap())
    np.testing.assert_array_equal(np.asarray(B), r)


def test_radius_check():
    """src/array.jl:451-453: R must be smaller than every axis."""
    with pytest.raises(ValueError, match="radius"):
        sh.StencilArray(np.zeros((3, 10)), sh.Window(3))
    with pytest.raises(ValueError, match="same size"):
        sh.SwitchingStencilArray(stencil=sh.Window(1), boundary=sh.Wrap(), padding=sh.Conditional(),
                                 _buffers=(np.zeros((4, 4)), np.zeros((5, 4))))


def test_kernel_length_check():
    """src/stencils/kernel.jl:115 ArgumentError."""
    with pytest.raises(ValueError, match="does not match"):
        sh.Kernel(sh.Window(2), np.arange(9.0))
    k = sh.Kernel(np.arange(9.0).reshape(3, 3, order="F"))
    assert k.stencil == sh.Window(1, 2) and list(k.weights) == list(range(9))


# --------------------------------------------------------- result types

@pytest.mark.parametrize("dt,op,exp", [
    (_abi.F64, _abi.OP_MEAN, _abi.F64), (_abi.F32, _abi.OP_MEAN, _abi.F32),
    (_abi.I64, _abi.OP_MEAN, _abi.F64), (_abi.U8, _abi.OP_MEAN, _abi.F64),
    (_abi.BOOL, _abi.OP_SUM, _abi.I64), (_abi.I8, _abi.OP_SUM, _abi.I8),
    (_abi.U8, _abi.OP_LIFE, _abi.U8), (_abi.F32, _abi.OP_LIFE, -1),
    (_abi.F32, _abi.OP_LAPLACIAN, _abi.F32), (_abi.I32, _abi.OP_LAPLACIAN, -1),
    (_abi.BF16, _abi.OP_SUM, -1),
])
def test_out_dtype(dt, op, exp):
    assert _abi.out_dtype(dt, op) == exp


def test_out_dtype_storage_and_weights():
    assert _abi.out_dtype(_abi.BF16, _abi.OP_SUM, _abi.OPF_ACC_F32) == _abi.BF16
    assert _abi.out_dtype(_abi.BF16, _abi.OP_KERNELPRODUCT, _abi.OPF_ACC_F32, _abi.F32) == _abi.BF16
    assert _abi.out_dtype(_abi.F32, _abi.OP_KERNELPRODUCT, 0, _abi.F64) == -1   # promotes
    assert _abi.out_dtype(_abi.F64, _abi.OP_KERNELPRODUCT, 0, _abi.F64) == _abi.F64


def test_oracle_and_python_agree_on_result_types(oracle):
    L = oracle.lib()
    for dt in range(13):
        for op in range(7):
            for fl in (0, _abi.OPF_ACC_F32):
                for w in (-1, dt, _abi.F32):
                    assert L.orc_out_dtype(dt, op, fl, w, 0) == _abi.out_dtype(dt, op, fl, w)


def test_resolve_functions():
    import statistics
    assert sh.ops.resolve(sum) is sh.sum and sh.ops.resolve(np.mean) is sh.mean
    assert sh.ops.resolve(max) is sh.maximum and sh.ops.resolve(statistics.mean) is sh.mean
    with pytest.raises(sh.Unsupported):
        sh.ops.resolve(lambda s: s)


def test_padval_rules():
    from stencils_hip.gather import _padval_bytes
    a32 = np.zeros((4, 4), np.float32)
    assert _padval_bytes(sh.StencilArray(a32, sh.Window(1), boundary=sh.Remove(0.0))) == b"\0" * 4
    with pytest.raises(sh.Unsupported):   # Float64 padval promotes a Float32 array
        _padval_bytes(sh.StencilArray(a32, sh.Window(1), boundary=sh.Remove(np.float64(0))))
    ai = np.zeros((4, 4), np.int32)
    with pytest.raises(sh.Unsupported):   # a Float padval promotes an Int array
        _padval_bytes(sh.StencilArray(ai, sh.Window(1), boundary=sh.Remove(0.5)))
    with pytest.raises(sh.Unsupported):
        _padval_bytes(sh.StencilArray(ai, sh.Window(1), boundary=sh.Remove()))


def test_window3_declined():
    """Window{R,3} offsets are malformed in the reference (2-tuples)."""
    assert not sh.Window(1, 3).gpu_supported()
    assert sh.Window(1, 2).gpu_supported()


def test_positional_radius_quirk():
    """Positional radius is the signed maximum (src/stencils/positional.jl:74-76)."""
    p = sh.Positional((-3, 0), (1, 0))
    assert p.radius() == 1 and p.reach() == (3, 0)


def test_gpu_path_fails_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    A = sh.StencilArray(np.zeros((8, 8)), sh.Window(1))
    with pytest.raises(sh.BackendUnavailable):
        sh.mapstencil(sh.mean, A)
