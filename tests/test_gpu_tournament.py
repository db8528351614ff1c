"""Tournament-pivoting FP32 LU, the mixed solver's optional factorization.

Its pivot order differs from partial pivoting by design, so these tests check
what any LU with row interchanges must satisfy:
  * ipiv is a LAPACK pivot vector (i <= ipiv[i] < m) and P A = L U holds to
    FP32 accuracy, ||P A - L U||max <= k u32 max(|L| |U|) with k the inner
    dimension (the classic backward-error bound of Gaussian elimination),
    P built from ipiv by sequential interchanges;
  * an exactly zero column is reported as the singular column;
  * the mixed solver with tournament pivoting reaches FP64 backward error.
The shapes cover one workgroup (m <= 2048 rows), several workgroups with a
finalist round of at most 128 rows, more than 8 workgroups (finalist round of
256 rows), short last leaves and wide matrices.
"""
import ctypes

import numpy as np
import pytest

import lair_amd as L
from lair_amd import _capi, inputs

pytestmark = pytest.mark.gpu
U32 = np.finfo(np.float32).eps


@pytest.fixture(scope="module")
def ctx():
    c = L.Context()
    yield c
    c.close()


def factor(ctx, a32, prec=_capi.PREC_FP32, nb=0, pivoting=_capi.PIVOT_TOURNAMENT):
    m, n = a32.shape
    da = L.DeviceArray.from_host(a32)
    dp = L.DeviceArray(min(m, n), 1, np.int64)
    sing = ctypes.c_int64(-2)
    L.check(ctx.lib.lair_sgetrf_pivot_device(ctx.h, m, n, da.ptr, da.ld, dp.ptr, nb,
                                             pivoting, prec, ctypes.byref(sing)),
            "sgetrf_pivot")
    return da.download(), dp.download().ravel(), sing.value


def check_plu(a, lu, piv):
    m, n = a.shape
    k = min(m, n)
    assert np.all(piv >= np.arange(k)) and np.all(piv < m)
    pa = a.astype(np.float64).copy()
    for i, p in enumerate(piv):
        if p != i:
            pa[[i, p]] = pa[[p, i]]
    lu = lu.astype(np.float64)
    lo = np.tril(lu[:, :k], -1) + np.eye(m, k)
    up = np.triu(lu[:k, :])
    err = np.abs(pa - lo @ up).max()
    scale = (np.abs(lo) @ np.abs(up)).max()
    assert err <= k * U32 * scale, (err, scale)
    return np.abs(lo).max()


@pytest.mark.parametrize("m,n", [(1000, 1000), (777, 777), (300, 500), (5000, 300), (20000, 48)])
def test_tournament_lu_reconstructs(ctx, m, n):
    a = inputs.gaussian_matrix(900 + m + n, m, n).astype(np.float32)
    lu, piv, sing = factor(ctx, a)
    assert sing == -1
    lmax = check_plu(a, lu, piv)
    assert lmax < 50.0  # multipliers stay moderate on Gaussian data


def test_tournament_bf16x6_update_reconstructs(ctx):
    m = n = 1536
    a = inputs.gaussian_matrix(77, m, n).astype(np.float32)
    lu, piv, sing = factor(ctx, a, prec=_capi.PREC_BF16X6)
    assert sing == -1
    check_plu(a, lu, piv)


def test_tournament_zero_column_is_singular(ctx):
    n = 200
    a = inputs.gaussian_matrix(12, n, n).astype(np.float32)
    a[:, 37] = 0.0
    lu, piv, sing = factor(ctx, a)
    assert sing == 37
    assert np.isfinite(lu).all()


@pytest.mark.parametrize("prec", [_capi.PREC_FP32, _capi.PREC_BF16X6])
def test_mixed_solver_with_tournament_pivoting(ctx, prec):
    from test_gpu_mixed import kappa_matrix, host_berr, run_mixed
    n, nrhs = 2048, 2
    a = kappa_matrix(55, n)
    b = inputs.gaussian_matrix(56, n, nrhs)
    st, info, x = run_mixed(ctx, a, b, prec | _capi.PIVOT_TOURNAMENT)
    assert st == _capi.LAIR_OK, _capi.last_error()
    assert info.converged == 1 and info.fell_back == 0
    assert host_berr(a, x, b) <= n * np.finfo(np.float64).eps


def test_partial_pivoting_fp32_tall_panels_reconstruct(ctx):
    # m > 4096: the FP32 exact-pivoting leaf over several 2048-row workgroups.
    m, n = 9000, 160
    a = inputs.gaussian_matrix(4242, m, n).astype(np.float32)
    lu, piv, sing = factor(ctx, a, pivoting=0)
    assert sing == -1
    lmax = check_plu(a, lu, piv)
    assert lmax <= 1.0  # partial pivoting: |L| <= 1
