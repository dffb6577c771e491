"""nkr_fr_ntt_dev_noalloc, the transform entry point that allocates no
device memory (tables and reordering buffer in one fixed device array per
GPU), called through the C ABI: its results equal nkr_fr_ntt_dev's for every
transform kind, on two calls in a row with the second smaller (the fixed
array keeps the first call's tables), and sizes beyond its 2^20-element
limit are rejected with an error before any device work."""
import ctypes

import numpy as np
import pytest

import nkr_g16 as N
import pyoracle as O

pytestmark = pytest.mark.gpu


def to_dev(a):
    ptr = ctypes.c_void_p()
    N._check(N.lib().nkr_dev_alloc(a.nbytes, ctypes.byref(ptr)))
    N._check(N.lib().nkr_memcpy_htod(ptr, N._p(a), a.nbytes))
    return ptr


def from_dev(ptr, like):
    out = np.empty_like(like)
    N._check(N.lib().nkr_memcpy_dtoh(N._p(out), ptr, out.nbytes))
    return out


def transform(fn, a, log_n, kind):
    d = to_dev(a)
    try:
        N._check(fn(d, log_n, kind, None))
        N._check(N.lib().nkr_device_synchronize())
        return from_dev(d, a)
    finally:
        N._check(N.lib().nkr_dev_free(d))


@pytest.mark.parametrize("sizes", [(20, 12), (16, 3)])
def test_noalloc_ntt_equals_ntt_dev(sizes):
    L = N.lib()
    for kind in (N.FFT, N.IFFT, N.COSET_FFT, N.ICOSET_FFT):
        for log_n in sizes:  # the second call is smaller than the first
            print(f"[noalloc] kind {kind} 2^{log_n}", flush=True)
            a = O.fr_random(9000 + 10 * log_n + kind, 1 << log_n)
            want = transform(L.nkr_fr_ntt_dev, a, log_n, kind)
            got = transform(L.nkr_fr_ntt_dev_noalloc, a, log_n, kind)
            assert np.array_equal(got, want), (kind, log_n)
            if log_n <= 12:
                assert np.array_equal(got, O.fft(a, kind)), (kind, log_n)


def test_noalloc_ntt_rejects_sizes_beyond_its_scratch():
    L = N.lib()
    a = O.fr_random(9100, 4)
    d = to_dev(a)
    try:
        for log_n in (21, 24, 0):
            assert L.nkr_fr_ntt_dev_noalloc(d, log_n, N.FFT, None) != 0
            assert "2^1 .. 2^20" in L.nkr_last_error().decode()
        assert L.nkr_fr_ntt_dev_noalloc(d, 2, 7, None) != 0
        assert "unknown transform kind" in L.nkr_last_error().decode()
        assert L.nkr_fr_ntt_dev_noalloc(None, 2, N.FFT, None) != 0
        # the rejected calls left the data alone
        assert np.array_equal(from_dev(d, a), a)
    finally:
        N._check(L.nkr_dev_free(d))
