"""lgt_plaquette_aos_noalloc, the allocation-free plaquette, against the entry
point whose device work it shares (lgt_upload_plaquette, in FP64 and FP32
engines) and the oracle: two calls in a row on caller-owned device buffers,
the second on a smaller lattice (the static scratch keeps the first call's
partial sums beyond the second's), on a created stream and on the null
stream; a volume beyond the scratch limit is refused with a message.

Device buffers come from the HIP runtime the library itself links
(/opt/rocm/lib/libamdhip64.so.7): test-side allocation, outside the call
under test."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip_runtime():
    h = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


def test_noalloc_plaquette_matches_engine():
    import masterarbeit_amd as lgt
    lib = lgt.load_library()
    h = hip_runtime()
    cases = [(8, 8, 8, 8), (6, 4, 4, 8)]           # the second call is smaller
    vmax = max(int(np.prod(L)) for L in cases)
    dfield, dout, stream = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(dfield), 16 * 8 * vmax) == 0
    assert h.hipMalloc(C.byref(dout), 8) == 0
    assert h.hipStreamCreate(C.byref(stream)) == 0
    try:
        for k, L in enumerate(cases):
            o = Oracle(L)
            U = o.hot(k + 3)
            o.engine_sweep(U, 3.09, 1, k + 3, 0)
            U = np.ascontiguousarray(U)
            assert h.hipMemcpy(dfield, U.ctypes.data, U.nbytes, H2D) == 0
            s = stream if k == 0 else None
            rc = lib.lgt_plaquette_aos_noalloc((C.c_int * 4)(*L), dfield, dout, s)
            assert rc == 0, lib.lgt_last_error()
            assert h.hipStreamSynchronize(s) == 0
            got = C.c_double(0.0)
            assert h.hipMemcpy(C.byref(got), dout, 8, D2H) == 0
            refs = []
            for prec in (lgt.FP64, lgt.FP32):
                with lgt.Engine(L, prec, 1) as e:
                    refs.append(e.upload_plaquette(U))
            print(f"{L}: allocation-free {got.value!r}, lgt_upload_plaquette {refs}")
            assert got.value == refs[0] == refs[1]
            assert got.value == pytest.approx(o.plaquette(U, acc=True), abs=1e-15)
        too_big = (C.c_int * 4)(98, 96, 96, 96)
        assert lib.lgt_plaquette_aos_noalloc(too_big, dfield, dout, stream) == -1   # LGT_EINVAL
        assert b"scratch limit" in lib.lgt_last_error()
    finally:
        h.hipStreamSynchronize(stream)
        h.hipStreamDestroy(stream)
        h.hipFree(dfield)
        h.hipFree(dout)
