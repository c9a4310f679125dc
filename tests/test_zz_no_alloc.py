"""The allocation-free guarded pair decisions (rp_pair_decisions_guarded_noalloc) against rp_pair_decisions_guarded:
device arrays in and out (the test's own buffers from the HIP runtime the library uses), the statistics from the
library's fixed device scratch area; two calls in a row, the second smaller (the area keeps the first call's contents, which must not leak into the
second), for spherocylinders near contact and for spheres; a call above the area's size is refused."""
import ctypes

import numpy as np
import pytest

import corpus
import rampack_amd as ra

pytestmark = pytest.mark.gpu

SC = ra.Shape("spherocylinder", radius=0.5, length=5.0)
HS = ra.Shape("sphere", radius=0.5)


def hip():
    """The HIP runtime the library itself is linked against (already loaded with it), for the test's own device
    buffers."""
    ra.load_library()
    lib = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return lib


class DeviceBuffer:
    def __init__(self, lib, host: np.ndarray):
        self.lib, self.nbytes, self.ptr = lib, host.nbytes, ctypes.c_void_p()
        assert lib.hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)) == 0
        assert lib.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1) == 0  # host to device

    def read(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.lib.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.lib.hipFree(self.ptr)


def device_call(shape, rec, codes=True):
    lib = hip()
    n = len(rec)
    rec = np.ascontiguousarray(rec, dtype=np.float64).reshape(-1)
    fill_dec, fill_codes = np.full(n, 7, dtype=np.uint8), np.full(n, 255, dtype=np.uint8)
    d_rec, d_dec = DeviceBuffer(lib, rec), DeviceBuffer(lib, fill_dec)
    d_codes = DeviceBuffer(lib, fill_codes) if codes else None
    try:
        st = ra.pair_decisions_guarded_noalloc(shape, n, d_rec.ptr.value, d_dec.ptr.value,
                                               d_codes.ptr.value if codes else 0)
        return d_dec.read(fill_dec), (d_codes.read(fill_codes) if codes else None), st
    finally:
        for b in (d_rec, d_dec, d_codes):
            if b is not None:
                b.free()


@pytest.mark.parametrize("with_codes", [True, False])
def test_noalloc_matches_allocating_entry_point(with_codes):
    big = corpus.spherocylinder_contact_corpus(600000, 31, radius=0.5, length=5.0, delta=1e-6)
    small = corpus.spherocylinder_contact_corpus(70000, 32, radius=0.5, length=5.0, delta=1e-3)
    for rec in (big, small):
        dec, codes, st = device_call(SC, rec, codes=with_codes)
        want_dec, want_codes, want_st = ra.pair_decisions_guarded(SC, rec)
        print(f"{len(rec)} pairs: {st}", flush=True)
        assert np.array_equal(dec, want_dec)
        if with_codes:
            assert np.array_equal(codes, want_codes)
        assert st == want_st
        assert st["mismatches"] == 0 and st["pairs"] == len(rec)


def test_noalloc_spheres_and_limits():
    rec = corpus.sphere_contact_corpus(200000, 33, radius=0.5, delta=1e-3)
    for part in (rec, rec[:5000]):
        dec, codes, st = device_call(HS, part)
        want_dec, want_codes, want_st = ra.pair_decisions_guarded(HS, part)
        assert np.array_equal(dec, want_dec) and np.array_equal(codes, want_codes) and st == want_st
    with pytest.raises(ra.PreconditionError, match="scratch area"):
        ra.pair_decisions_guarded_noalloc(SC, ra.GUARD_NOALLOC_MAX_PAIRS + 1, 1, 1)
    assert ra.pair_decisions_guarded_noalloc(SC, 0, 0, 0)["pairs"] == 0
