"""The allocation-free FP64 inversion, qmcx_invert_transpose_noalloc, against qmcx_invert_transpose
(DiracMatrix::invert_transpose, src/QMCWaveFunctions/DiracMatrix.h:93-121) on the same matrices.

Both run the same Gauss-Jordan kernel, so the outputs must be bit-identical.  Two calls in a row on
the shared scratch array: the first large enough to need two launches (n = 384: 14 matrices per
launch), the second smaller (n = 96).  The test's device buffers and stream come from the HIP
runtime that libqmcx.so links; the call goes through the C ABI.
"""
import ctypes as C

import numpy as np
import pytest

from miniqmc_amd import qmcx

pytestmark = pytest.mark.gpu


def _mats(seed, nmat, n):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, (nmat, n, n))


class _Hip:
    """The HIP runtime libqmcx.so links (/opt/rocm), for the test's own device buffers and stream."""

    def __init__(self):
        qmcx.load()
        self.L = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: hip error {rc}")

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), C.c_size_t(nbytes)), "hipMalloc")
        self.bufs.append(p)
        return p

    def h2d(self, dst, arr):
        self.check(self.L.hipMemcpy(dst, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1), "hipMemcpy")

    def d2h(self, arr, src):
        self.check(self.L.hipMemcpy(arr.ctypes.data_as(C.c_void_p), src, C.c_size_t(arr.nbytes), 2), "hipMemcpy")

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc(a, hip, stream):
    nmat, n, _ = a.shape
    da, dp = hip.malloc(a.nbytes), hip.malloc(a.nbytes)
    dl, dph = hip.malloc(8 * nmat), hip.malloc(8 * nmat)
    hip.h2d(da, a)
    hip.h2d(dp, np.full_like(a, np.nan))
    rc = qmcx.load().qmcx_invert_transpose_noalloc(n, nmat, da, dp, dl, dph, stream)
    if rc != 0:
        raise qmcx.QmcxError(qmcx.load().qmcx_last_error().decode())
    hip.check(hip.L.hipStreamSynchronize(stream), "hipStreamSynchronize")
    p, l, ph = np.empty_like(a), np.empty(nmat), np.empty(nmat)
    hip.d2h(p, dp)
    hip.d2h(l, dl)
    hip.d2h(ph, dph)
    return p, l, ph


def test_noalloc_matches_allocating_entry_point():
    """Two calls in a row on one stream: n = 384 over 20 matrices (two launches), then n = 96 over 3."""
    hip = _Hip()
    stream = C.c_void_p()
    hip.check(hip.L.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    try:
        for n, nmat, seed in ((384, 20, 1), (96, 3, 2)):
            print(f"noalloc inversion n={n} nmat={nmat}", flush=True)
            a = _mats(seed, nmat, n)
            p_ref, l_ref, ph_ref = qmcx.invert_transpose(a)
            p, l, ph = _noalloc(a, hip, stream)
            np.testing.assert_array_equal(p, p_ref)
            np.testing.assert_array_equal(l, l_ref)
            np.testing.assert_array_equal(ph, ph_ref)
            sign, ld = np.linalg.slogdet(a)
            np.testing.assert_allclose(l, ld, rtol=1e-10)
            np.testing.assert_array_equal(ph, np.where(sign > 0, 0.0, np.pi))
    finally:
        hip.L.hipStreamDestroy(stream)
        hip.free()


def test_noalloc_rejects_order_beyond_scratch():
    L = qmcx.load()
    buf = (C.c_double * 1)()
    p = C.addressof(buf)
    rc = L.qmcx_invert_transpose_noalloc(1025, 1, p, p, p, p, None)
    assert rc == -1 and b"exceeds" in L.qmcx_last_error()   # QMCX_ERR_ARG, before any device work
