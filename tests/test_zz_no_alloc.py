"""The allocation-free getrs (lair_dgetrs_device_noalloc) against
lair_dgetrs_device on the same factors: two calls in a row, the second with a
smaller system, each bit-identical to the allocating entry point, plus the
size limits.  Both use the same kernels; only the workspace differs (a static
device array instead of the context's grow-only buffers)."""
import ctypes

import numpy as np
import pytest

import lair_amd as L
from lair_amd import _capi, inputs

pytestmark = pytest.mark.gpu


def _factor(ctx, n, seed):
    a = inputs.gaussian_matrix(seed, n, n)
    da = L.DeviceArray.from_host(a)
    dp = L.DeviceArray(n, 1, np.int64)
    sing = ctypes.c_int64(-1)
    L.check(ctx.lib.lair_dgetrf_device(ctx.h, n, n, da.ptr, n, dp.ptr, 0, ctypes.byref(sing)), "getrf")
    return da, dp


def _solve(ctx, fn, n, nrhs, da, dp, b, trans):
    db = L.DeviceArray.from_host(b)
    st = fn(ctx.h, n, nrhs, da.ptr, n, dp.ptr, db.ptr, nrhs, trans)
    return st, db.download()


def test_noalloc_getrs_matches_allocating_getrs_twice():
    ctx = L.Context()
    try:
        lib = ctx.lib
        for n, nrhs, trans in ((3000, 16, 0), (700, 3, 1)):
            da, dp = _factor(ctx, n, 500 + n)
            b = inputs.gaussian_matrix(600 + n, n, nrhs)
            st_ref, x_ref = _solve(ctx, lib.lair_dgetrs_device, n, nrhs, da, dp, b, trans)
            st, x = _solve(ctx, lib.lair_dgetrs_device_noalloc, n, nrhs, da, dp, b, trans)
            assert st_ref == _capi.LAIR_OK and st == _capi.LAIR_OK, _capi.last_error()
            assert np.array_equal(x.view(np.int64), x_ref.view(np.int64))
    finally:
        ctx.close()


def test_noalloc_getrs_limits():
    ctx = L.Context()
    try:
        lib = ctx.lib
        st = lib.lair_dgetrs_device_noalloc(ctx.h, 16385, 1, None, 16385, None, None, 1, 0)
        assert st == _capi.LAIR_ERR_UNSUPPORTED
        st = lib.lair_dgetrs_device_noalloc(ctx.h, 100, 17, None, 100, None, None, 17, 0)
        assert st == _capi.LAIR_ERR_UNSUPPORTED
        # Small systems (exact-order kernel) take the same entry point.
        da, dp = _factor(ctx, 100, 77)
        b = inputs.gaussian_matrix(78, 100, 2)
        st_ref, x_ref = _solve(ctx, lib.lair_dgetrs_device, 100, 2, da, dp, b, 0)
        st, x = _solve(ctx, lib.lair_dgetrs_device_noalloc, 100, 2, da, dp, b, 0)
        assert st == st_ref == _capi.LAIR_OK
        assert np.array_equal(x.view(np.int64), x_ref.view(np.int64))
    finally:
        ctx.close()
