"""The allocation-free dense LU (sov_dense_lu_dev_noalloc) against the
allocating one (sov_dense_lu) through the C ABI: the same LU factors and
permutation, bit for bit, on two calls in a row (the second smaller, so
the fixed scratch array holds the first call's swap list when it starts),
the singular status on a rank-deficient matrix, and the size limit."""
import ctypes as C

import numpy as np
import pytest

from sovereign_amd import _capi as K

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def host_lu(a):
    lib = K.load()
    n = a.shape[0]
    flat = np.asfortranarray(a, dtype=np.float64).reshape(-1, order="F").copy()
    piv = np.zeros(n, dtype=np.int32)
    rc = lib.sov_dense_lu(n, K.pf(flat), K.pi(piv))
    return rc, flat, piv


class Hip:
    """Device buffers for the test (the caller's side of the entry point),
    through the HIP runtime the library itself links."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, arr):
        assert self.rt.hipMemcpy(dst, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) == 0

    def d2h(self, arr, src):
        assert self.rt.hipMemcpy(arr.ctypes.data_as(C.c_void_p), src, arr.nbytes, 2) == 0

    def free(self, p):
        assert self.rt.hipFree(p) == 0


def device_lu(a):
    lib = K.load()
    hip = Hip()
    n = a.shape[0]
    flat = np.asfortranarray(a, dtype=np.float64).reshape(-1, order="F").copy()
    piv = np.full(n, -7, dtype=np.int32)
    da, dp = hip.alloc(flat.nbytes), hip.alloc(piv.nbytes)
    try:
        hip.h2d(da, flat)
        hip.h2d(dp, piv)
        assert hip.rt.hipDeviceSynchronize() == 0
        rc = lib.sov_dense_lu_dev_noalloc(n, da, dp)
        assert hip.rt.hipDeviceSynchronize() == 0
        hip.d2h(flat, da)
        hip.d2h(piv, dp)
    finally:
        hip.free(da)
        hip.free(dp)
    return rc, flat, piv


@pytest.fixture(scope="module")
def lib(gpu_api):
    return K.load()


def test_noalloc_lu_matches_allocating_lu_twice(lib):
    rng = np.random.default_rng(20261204)
    for n in (700, 300):  # second call smaller than the first
        say(f"noalloc LU n={n}")
        a = rng.standard_normal((n, n))
        rc_h, lu_h, piv_h = host_lu(a)
        rc_d, lu_d, piv_d = device_lu(a)
        assert rc_h == 0 and rc_d == 0
        np.testing.assert_array_equal(piv_d, piv_h)
        np.testing.assert_array_equal(lu_d, lu_h)
        # and the factors reconstruct P A
        lu = lu_d.reshape((n, n), order="F")
        L = np.tril(lu, -1) + np.eye(n)
        U = np.triu(lu)
        assert np.max(np.abs(L @ U - a[piv_d])) <= 1e-10 * n


def test_noalloc_lu_singular_and_limits(lib):
    rng = np.random.default_rng(7)
    n = 130
    a = rng.standard_normal((n, n))
    a[:, 5] = a[:, 3]  # rank deficient
    say("noalloc LU singular n=130")
    rc_h, _, _ = host_lu(a)
    rc_d, _, _ = device_lu(a)
    assert rc_h == 1 and rc_d == 1
    say("noalloc LU after a singular call n=64")
    b = rng.standard_normal((64, 64))
    rc_h, lu_h, piv_h = host_lu(b)
    rc_d, lu_d, piv_d = device_lu(b)
    assert rc_h == 0 and rc_d == 0  # the flag starts at zero on every call
    np.testing.assert_array_equal(lu_d, lu_h)
    np.testing.assert_array_equal(piv_d, piv_h)
    hip = Hip()
    small, p = hip.alloc(64), hip.alloc(64)
    try:
        # above the scratch limit, or empty: rejected before any device work
        assert lib.sov_dense_lu_dev_noalloc(40000, small, p) == K.SOV_ERR_ARG
        assert lib.sov_dense_lu_dev_noalloc(0, small, p) == K.SOV_ERR_ARG
    finally:
        hip.free(small)
        hip.free(p)
