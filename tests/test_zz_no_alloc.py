"""sesmd_pair_counts_noalloc (no allocation: its partial sums live in a static
device array) against sesmd_pair_counts (which allocates a counter buffer per
call), through the C ABI.  Integer counts: the bar is bit-exact.  Two calls
in a row, the second on a smaller system, both against the allocating entry
point on the same context."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_pair_counts_match(sys_medium, sys_small):
    from surfactant_enhanced_spreading_amd import engine
    L = engine.lib()
    L.sesmd_pair_counts_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    hip = ctypes.CDLL('/opt/rocm/lib/libamdhip64.so')
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    ctxs = []
    out = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(out), 16) == 0
    try:
        for s in (sys_medium, sys_small):  # the second system is the smaller one
            e = engine.Engine(s)
            e.rebuild()
            ctxs.append((e, e.pair_counts()))
        assert ctxs[1][0].n < ctxs[0][0].n
        for e, want in ctxs:
            print('sites', e.n, 'pairs', want, flush=True)
            rc = L.sesmd_pair_counts_noalloc(e.h, out)
            assert rc == 0, L.sesmd_last_error().decode()
            assert hip.hipStreamSynchronize(ctypes.c_void_p(e.stream())) == 0
            host = (ctypes.c_double * 2)()
            assert hip.hipMemcpy(ctypes.cast(host, ctypes.c_void_p), out, 16, 2) == 0  # device to host
            assert want[0] > 0 and want[1] > 0
            assert (host[0], host[1]) == want
    finally:
        hip.hipFree(out)
        for e, _ in ctxs:
            e.close()
