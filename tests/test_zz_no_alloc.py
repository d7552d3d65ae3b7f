"""The allocation-free sort (dl_sort_pairs_u32_noalloc) against the stable
sort it varies (dl_sort_pairs_u32), through the C ABI: keys and values live
in device buffers that the test allocates with the HIP runtime, the library
sorts them in its fixed scratch array. Two calls in a row, the second
smaller (the scratch keeps the first call's contents), are both bit-exact;
a call over the maximum fails with DL_ERR_CAPACITY and allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from droplet_amd import _lib as L
from droplet_amd.renderer import sort_pairs_u32

pytestmark = pytest.mark.gpu


def _hip():
    L.load()  # libdroplet_hip.so pulls in the HIP runtime
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, hip, host: np.ndarray):
        self.hip, self.n = hip, host.size
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(host.nbytes, 4)) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, H2D) == 0

    def read(self) -> np.ndarray:
        out = np.empty(self.n, np.uint32)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc_sort(device, keys, vals, key_bits):
    hip = _hip()
    dk, dv = DeviceArray(hip, keys), DeviceArray(hip, vals)
    try:
        assert hip.hipDeviceSynchronize() == 0
        L.check(L.load().dl_sort_pairs_u32_noalloc(device.handle, dk.ptr, dv.ptr, keys.size, key_bits))
        return dk.read(), dv.read()
    finally:
        dk.free()
        dv.free()


def test_noalloc_sort_matches_the_allocating_sort(device):
    lib = L.load()
    cap = int(lib.dl_sort_pairs_u32_noalloc_max())
    assert cap == 1 << 20
    rng = np.random.default_rng(5)
    # two calls in a row, the second smaller: the scratch holds the first
    # call's ping-pong arrays and counts when the second starts
    for n, bits, kmax in ((cap, 32, 1 << 32), (300_001, 17, 1 << 17), (5000, 32, 7), (1, 8, 256)):
        keys = rng.integers(0, kmax, n, dtype=np.uint64).astype(np.uint32)
        vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        ref_k, ref_v = sort_pairs_u32(device, keys, vals, key_bits=bits)
        got_k, got_v = _noalloc_sort(device, keys, vals, bits)
        print(f"case: n={n} key_bits={bits}", flush=True)
        np.testing.assert_array_equal(got_k, ref_k)
        np.testing.assert_array_equal(got_v, ref_v)
        # stable by the low key_bits bits
        order = np.argsort(keys & np.uint32((1 << bits) - 1 if bits < 32 else 0xFFFFFFFF), kind="stable")
        np.testing.assert_array_equal(got_v, vals[order])


def test_noalloc_sort_rejects_more_than_its_scratch(device):
    lib = L.load()
    cap = int(lib.dl_sort_pairs_u32_noalloc_max())
    hip = _hip()
    big = np.zeros(cap + 1, np.uint32)
    dk, dv = DeviceArray(hip, big), DeviceArray(hip, big)
    try:
        rc = lib.dl_sort_pairs_u32_noalloc(device.handle, dk.ptr, dv.ptr, cap + 1, 32)
        assert rc == 4  # DL_ERR_CAPACITY
        assert b"at most" in lib.dl_last_error()
        assert lib.dl_sort_pairs_u32_noalloc(device.handle, None, None, 5, 32) == 1  # DL_ERR_INVALID
    finally:
        dk.free()
        dv.free()
