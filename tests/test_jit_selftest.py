"""The runtime compiler builds the library's specialisations on this machine
(no GPU needed): every generator's source compiles with hipRTC.  A source
that does not compile would make calls silently keep the slower kernels."""
import ctypes

import stencils_hip as sh


def test_runtime_specialisations_compile():
    lib = sh.load_library()
    buf = ctypes.create_string_buffer(16384)
    failures = lib.sth_jit_selftest(buf, len(buf))
    assert failures == 0, buf.value.decode(errors="replace")
