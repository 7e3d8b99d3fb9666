"""The stable key sort for callers that cannot allocate (upf_radix_sort_noalloc):
caller-owned device buffers, the library's fixed working array.  Two calls in
a row -- the maximum size, then a smaller one with ties and the full 32-bit
key range -- give exactly what upf_radix_sort gives, and sizes beyond the
maximum are refused."""
import ctypes as C

import numpy as np
import pytest

import unifiedparticleframeworkcuda_amd as upf

pytestmark = pytest.mark.gpu

MAX = 1 << 20


def _hip():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


def _host_sort(keys):
    L = upf.lib()
    ko = np.zeros_like(keys)
    perm = np.zeros(len(keys), np.int32)
    pu = C.POINTER(C.c_uint32)
    assert L.upf_radix_sort(len(keys), keys.ctypes.data_as(pu), ko.ctypes.data_as(pu),
                            perm.ctypes.data_as(C.POINTER(C.c_int))) == 0
    return ko, perm


def test_noalloc_sort_equals_the_allocating_sort(gpu):
    L = upf.lib()
    L.upf_radix_sort_noalloc.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    hip = _hip()
    bufs = []
    for _ in range(3):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), MAX * 4) == 0
        bufs.append(p)
    d_in, d_out, d_perm = bufs
    try:
        rng = np.random.RandomState(17)
        cases = [rng.randint(0, 1 << 30, MAX).astype(np.uint32),          # 30-bit keys, the maximum size
                 np.concatenate([rng.randint(0, 64, 40000),               # many ties
                                 rng.randint(0, 1 << 32, 37000, dtype=np.uint64)]).astype(np.uint32)]
        cases[1][:5] = [0xFFFFFFFF, 0, 0xFFFFFFFF, 0x80000000, 0]
        for keys in cases:
            n = len(keys)
            print("noalloc sort: n = %d" % n, flush=True)
            exp_k, exp_p = _host_sort(keys)
            assert hip.hipMemcpy(d_in, keys.ctypes.data_as(C.c_void_p), n * 4, H2D) == 0
            assert L.upf_radix_sort_noalloc(n, d_in, d_out, d_perm, None) == 0
            assert hip.hipDeviceSynchronize() == 0
            got_k = np.zeros(n, np.uint32)
            got_p = np.zeros(n, np.int32)
            assert hip.hipMemcpy(got_k.ctypes.data_as(C.c_void_p), d_out, n * 4, D2H) == 0
            assert hip.hipMemcpy(got_p.ctypes.data_as(C.c_void_p), d_perm, n * 4, D2H) == 0
            assert np.array_equal(got_k, exp_k)
            assert np.array_equal(got_p, exp_p)
            assert np.array_equal(got_p, np.argsort(keys, kind="stable"))
        assert L.upf_radix_sort_noalloc(MAX + 1, d_in, d_out, d_perm, None) == 2  # UPF_ERR_ARG
        assert L.upf_radix_sort_noalloc(0, d_in, d_out, d_perm, None) == 0
    finally:
        for p in bufs:
            hip.hipFree(p)
