"""FP64 LU at an order where the first panels use the narrow resident leaves
(32 columns and 512 rows per workgroup while at least 12288 rows remain,
then 64 columns): the getrf.rs:46-120 elimination with the panel -> laswp ->
trsm -> gemm blocking of getrf.rs:216-322.

Checked on the host, O(n^2):
  * ipiv is a LAPACK pivot vector and partial pivoting keeps |L| <= 1;
  * P A = L U through random vectors, ||P A v - L (U v)||inf <= n eps64
    ||A||inf ||v||inf, with P built from ipiv by sequential interchanges (a
    dropped or misplaced update leaves an O(1) relative error);
  * the solve from the factors (getrs.rs:12-39) has backward error
    ||b - A x||inf / (||A||inf ||x||inf) <= n eps64 by the oracle's
    long-double residual.
"""
import ctypes

import numpy as np
import pytest

import lair_amd as L
from lair_amd import inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS = float(np.finfo(np.float64).eps)


@pytest.fixture(scope="module")
def ctx():
    c = L.Context()
    yield c
    c.close()


def plu_vector_error(a, lu, piv, nvec=2, seed=11):
    n = a.shape[0]
    order = np.arange(n)
    for i, p in enumerate(piv):
        if p != i:
            order[[i, p]] = order[[p, i]]
    rng = np.random.default_rng(seed)
    anorm = float(np.abs(a).sum(1).max())
    worst = 0.0
    step = 1024
    for _ in range(nvec):
        v = rng.standard_normal(n)
        uv = np.empty(n)
        pav = np.empty(n)
        luv = np.empty(n)
        for r0 in range(0, n, step):
            uv[r0:r0 + step] = np.triu(lu[r0:r0 + step], r0) @ v
            pav[r0:r0 + step] = a[order[r0:r0 + step]] @ v
        for r0 in range(0, n, step):
            luv[r0:r0 + step] = np.tril(lu[r0:r0 + step], r0 - 1) @ uv + uv[r0:r0 + step]
        worst = max(worst, float(np.abs(pav - luv).max() / (anorm * np.abs(v).max())))
    return worst


def test_fp64_narrow_leaves_factor_and_solve(ctx):
    # 13000 rows: panels at columns 0..640 run 32-column leaves, the rest
    # 64-column ones; the order is not a multiple of 128.
    n = 13000
    a = inputs.gaussian_matrix(9300, n, n)
    b = inputs.gaussian_matrix(9301, n, 1)
    da = L.DeviceArray.from_host(a)
    dp = L.DeviceArray(n, 1, np.int64)
    sing = ctypes.c_int64(-2)
    L.check(ctx.lib.lair_dgetrf_device(ctx.h, n, n, da.ptr, da.ld, dp.ptr, 128, ctypes.byref(sing)), "getrf")
    assert sing.value == -1
    lu = da.download()
    piv = dp.download().ravel()
    assert np.all(piv >= np.arange(n)) and np.all(piv < n)
    assert float(np.abs(np.tril(lu, -1)).max()) <= 1.0
    err = plu_vector_error(a, lu, piv)
    assert err <= n * EPS, err
    db = L.DeviceArray.from_host(b)
    L.check(ctx.lib.lair_dgetrs_device(ctx.h, n, 1, da.ptr, da.ld, dp.ptr, db.ptr, 1, 0), "getrs")
    x = db.download()
    r = O.residual(a, x, b, nthreads=8)
    be = float(np.abs(r).max() / (np.abs(a).sum(1).max() * np.abs(x).max()))
    assert be <= n * EPS, be
    da.free()
    db.free()
