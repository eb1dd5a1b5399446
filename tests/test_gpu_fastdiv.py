"""The engine's short division (wmh::recip_rn and wmh::div_rn_fast in
minhashcuda_amd/csrc/wmh_div.h, used for ln(w) / r when every r of the table
lies in [2^-60, 2^60]) against IEEE binary32 division, on the GPU:

* recip_rn(r) == 1.0f / r for every float r in [2^-60, 2^61) -- all
  1,015,021,568 of them;
* div_rn_fast(a, b) == a / b for all 2^23 significands of b against slices
  of the significands of a, at exponent pairs spanning the operand range
  (|ln w| in [2^-25, 2^7), r in [2^-60, 2^60]).

The full sweep of one exponent pair (2^46 quotients, about 25 s) is
``python tests/test_gpu_fastdiv.py EA EB``; it was run for (0, 0), (-20, 30),
(6, -60) and (-25, 59) with no mismatch.  The helper library is built by
tests/native/build_helpers.py (called from the build entry point).
"""
import ctypes
import os
import sys

import pytest

pytestmark = pytest.mark.gpu
LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libfastdiv_check.so")


def load():
    if not os.path.exists(LIB):
        sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "native"))
        import build_helpers
        build_helpers.build()
    lib = ctypes.CDLL(LIB)
    for name in ("fastdiv_rcp_check",):
        getattr(lib, name).argtypes = [ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_ulonglong)]
        getattr(lib, name).restype = ctypes.c_int
    lib.fastdiv_pair_check.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                       ctypes.c_uint32, ctypes.POINTER(ctypes.c_ulonglong)]
    lib.fastdiv_pair_check.restype = ctypes.c_int
    return lib


@pytest.fixture(scope="module")
def lib(gpu_available):
    return load()


def bits(e):
    return (127 + e) << 23


def test_reciprocal_exhaustive(lib):
    out = (ctypes.c_ulonglong * 2)()
    assert lib.fastdiv_rcp_check(bits(-60), bits(61), out) == 0
    assert out[0] == 0, "recip_rn differs from 1/r, first at bits %08x" % out[1]


@pytest.mark.parametrize("ea,eb", [(0, 0), (-25, -60), (-25, 59), (6, -60), (6, 59),
                                   (-1, 0), (3, -7), (-13, 21)])
def test_quotient_slices(lib, ea, eb):
    """Every b significand against 2^17 a significands (2^40 quotients) taken
    from the start, the middle and the end of the a significand range."""
    out = (ctypes.c_ulonglong * 2)()
    for a0 in (0, (1 << 22) - (1 << 15), (1 << 23) - (1 << 15)):
        print("ea", ea, "eb", eb, "a0", a0, flush=True)
        assert lib.fastdiv_pair_check(ea, eb, a0, 1 << 15, out) == 0
        assert out[0] == 0, "first mismatch b sig %x a sig %x" % (out[1] >> 32, out[1] & 0xffffffff)


if __name__ == "__main__":
    import torch  # noqa: F401  (one HIP runtime in the process)
    ea, eb = int(sys.argv[1]), int(sys.argv[2])
    lib_ = load()
    out = (ctypes.c_ulonglong * 2)()
    for a0 in range(0, 1 << 23, 1 << 20):
        rc = lib_.fastdiv_pair_check(ea, eb, a0, 1 << 20, out)
        assert rc == 0
        print("a significands < %d: %d mismatches" % (a0 + (1 << 20), out[0]), flush=True)
        assert out[0] == 0
    print("div_rn_fast == a / b for all 2^46 significand pairs at exponents (%d, %d)" % (ea, eb))
