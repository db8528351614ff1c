"""Failure paths of the C ABI on the GPU, forced from the host so that every
kernel still exits normally.

* A wait limit of 0 makes every cooperative wait give up at once: the call
  reports LAIR_ERR_SYNC, the status word is reset by that report, and the next
  call on the same context succeeds with the oracle's pivots.
* Device getrs range-checks caller pivots (as the host getrs does, which the
  reference leaves to the caller: src/lapack/getrs.rs:18-20 panics on a shape
  mismatch) and reports LAIR_ERR_SHAPE for a pivot outside [0, n).
"""
import ctypes

import numpy as np
import pytest

import lair_amd as L
from lair_amd import _capi, inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEFAULT_LIMIT = 1 << 23
EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def ctx():
    c = L.Context()
    yield c
    c.lib.lair_context_set_wait_limit(c.h, DEFAULT_LIMIT)
    c.close()


def _factor(ctx, a):
    n = a.shape[0]
    da = L.DeviceArray.from_host(a)
    dp = L.DeviceArray(n, 1, np.int64)
    sing = ctypes.c_int64(-1)
    L.check(ctx.lib.lair_dgetrf_device(ctx.h, n, n, da.ptr, n, dp.ptr, 0, ctypes.byref(sing)), "getrf")
    return da, dp


def test_wait_limit_zero_reports_sync_then_recovers_getrf(ctx):
    n = 512
    a = inputs.gaussian_matrix(31, n, n)
    ref = a.copy()
    piv_o, _ = O.getrf(ref)
    lib = ctx.lib
    da = L.DeviceArray.from_host(a)
    dp = L.DeviceArray(n, 1, np.int64)
    sing = ctypes.c_int64(-1)
    L.check(lib.lair_context_set_wait_limit(ctx.h, 0), "wait limit")
    try:
        st = lib.lair_dgetrf_device(ctx.h, n, n, da.ptr, n, dp.ptr, 0, ctypes.byref(sing))
    finally:
        L.check(lib.lair_context_set_wait_limit(ctx.h, DEFAULT_LIMIT), "wait limit")
    assert st == _capi.LAIR_ERR_SYNC
    # The report reset the status word: the context is clean again.
    assert lib.lair_context_synchronize(ctx.h) == _capi.LAIR_OK
    da.upload(a)
    L.check(lib.lair_dgetrf_device(ctx.h, n, n, da.ptr, n, dp.ptr, 0, ctypes.byref(sing)), "getrf")
    assert sing.value == -1
    assert np.array_equal(dp.download().ravel(), piv_o)


def test_wait_limit_zero_reports_sync_in_getrs_and_mixed(ctx):
    n = 600
    a = inputs.gaussian_matrix(32, n, n)
    b = inputs.gaussian_matrix(33, n, 2)
    da, dp = _factor(ctx, a)
    lib = ctx.lib
    db = L.DeviceArray.from_host(b)
    L.check(lib.lair_context_set_wait_limit(ctx.h, 0), "wait limit")
    try:
        # n > 128: the dataflow triangular solve waits on the blocks above.
        st_getrs = lib.lair_dgetrs_device(ctx.h, n, 2, da.ptr, n, dp.ptr, db.ptr, 2, 0)
        info = _capi.IrInfo()
        a_dev = L.DeviceArray.from_host(a)
        x_dev = L.DeviceArray(n, 2)
        st_mixed = lib.lair_dgesv_mixed_device(ctx.h, n, 2, a_dev.ptr, n, db.ptr, 2, x_dev.ptr, 2,
                                               _capi.PREC_FP32, 10, 0.0, ctypes.byref(info))
    finally:
        L.check(lib.lair_context_set_wait_limit(ctx.h, DEFAULT_LIMIT), "wait limit")
    assert st_getrs == _capi.LAIR_ERR_SYNC
    assert st_mixed == _capi.LAIR_ERR_SYNC
    # Recovered: the same solve now succeeds.
    db.upload(b)
    L.check(lib.lair_dgetrs_device(ctx.h, n, 2, da.ptr, n, dp.ptr, db.ptr, 2, 0), "getrs")
    x = db.download()
    r = O.residual(a, x, b, nthreads=4)
    assert float(np.abs(r).max() / (np.abs(a).sum(1).max() * np.abs(x).max())) <= n * EPS


@pytest.mark.parametrize("n", [64, 300])
@pytest.mark.parametrize("bad", [-1, "n", 1 << 33])
def test_device_getrs_rejects_out_of_range_pivot(ctx, n, bad):
    a = inputs.gaussian_matrix(40 + n, n, n) + n * np.eye(n)
    b = inputs.gaussian_matrix(41 + n, n, 1)
    da, dp = _factor(ctx, a)
    piv = dp.download().ravel()
    wrong = piv.copy()
    wrong[n // 2] = n if bad == "n" else bad
    lib = ctx.lib
    db = L.DeviceArray.from_host(b)
    dbad = L.DeviceArray.from_host(wrong.reshape(-1, 1))
    st = lib.lair_dgetrs_device(ctx.h, n, 1, da.ptr, n, dbad.ptr, db.ptr, 1, 0)
    assert st == _capi.LAIR_ERR_SHAPE
    # The next call with the valid pivots succeeds.
    db.upload(b)
    L.check(lib.lair_dgetrs_device(ctx.h, n, 1, da.ptr, n, dp.ptr, db.ptr, 1, 0), "getrs")
    x = db.download()
    r = O.residual(a, x, b, nthreads=4)
    assert float(np.abs(r).max() / (np.abs(a).sum(1).max() * np.abs(x).max())) <= n * EPS
