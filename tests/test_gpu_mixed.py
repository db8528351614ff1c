"""Mixed-precision LU + FP64 iterative refinement on the GPU.

Inputs follow the headline workload's construction at test sizes: A = U diag(s)
V^T with U, V Haar orthogonal (QR of a Gaussian with the sign fix of Mezzadri
2007) and s log-spaced from 1 to 1e6.  Every result is checked on the host with
the oracle's long-double residual: normwise backward error
||b - A x||inf / (||A||inf ||x||inf) <= n * eps64.
"""
import ctypes

import numpy as np
import pytest

import lair_amd as L
from lair_amd import _capi, inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def kappa_matrix(seed, n, kappa=1e6):
    def haar(s):
        q, r = np.linalg.qr(inputs.gaussian_matrix(s, n, n))
        return q * np.sign(np.diag(r))
    s = 10.0 ** (np.log10(kappa) * np.arange(n) / (n - 1))
    return np.ascontiguousarray((haar(seed) * s) @ haar(seed + 1).T)


def host_berr(a, x, b):
    r = O.residual(a, x, b, nthreads=8)
    return float((np.abs(r).max(0) / (np.abs(a).sum(1).max() * np.abs(x).max(0))).max())


def run_mixed(ctx, a, b, prec, max_iter=20, pad=64):
    n, nrhs = b.shape
    ld = n + pad
    da = L.DeviceArray.from_host(a, ld=ld)
    db = L.DeviceArray.from_host(b)
    dx = L.DeviceArray(n, nrhs)
    info = _capi.IrInfo()
    st = ctx.lib.lair_dgesv_mixed_device(ctx.h, n, nrhs, da.ptr, ld, db.ptr, nrhs, dx.ptr, nrhs, prec,
                                         max_iter, 0.0, ctypes.byref(info))
    return st, info, dx.download()


@pytest.fixture(scope="module")
def ctx():
    c = L.Context()
    yield c
    c.close()


@pytest.mark.parametrize("prec", [_capi.PREC_FP32, _capi.PREC_BF16X6])
@pytest.mark.parametrize("n,nrhs", [(1024, 1), (1024, 64), (2048, 4)])
def test_mixed_converges_to_fp64_backward_error(ctx, prec, n, nrhs):
    a = kappa_matrix(31 + n, n)
    b = inputs.gaussian_matrix(33 + n, n, nrhs)
    st, info, x = run_mixed(ctx, a, b, prec)
    assert st == _capi.LAIR_OK, _capi.last_error()
    assert info.converged == 1 and info.fell_back == 0
    assert info.iterations <= 10
    be = host_berr(a, x, b)
    assert be <= n * EPS
    # the device's own estimate agrees with the host check
    assert abs(info.backward_error - be) <= 0.5 * be + 1e-18


# Orders that are not a multiple of the 128-row solve block, odd leading
# dimensions (the residual's A rows lose their 16-byte alignment) and 1 to 3
# right-hand sides: the refinement step that fuses the backward solve, x += d
# and the residual covers them all.
@pytest.mark.parametrize("prec", [_capi.PREC_FP32, _capi.PREC_BF16X6])
@pytest.mark.parametrize("n,nrhs,pad", [(333, 2, 0), (1001, 3, 2), (640, 1, 1), (129, 4, 0)])
def test_mixed_ragged_shapes(ctx, prec, n, nrhs, pad):
    a = kappa_matrix(41 + n, n)
    b = inputs.gaussian_matrix(43 + n, n, nrhs)
    st, info, x = run_mixed(ctx, a, b, prec, pad=pad)
    assert st == _capi.LAIR_OK, _capi.last_error()
    assert info.converged == 1 and info.fell_back == 0
    be = host_berr(a, x, b)
    assert be <= n * EPS
    assert abs(info.backward_error - be) <= 0.5 * be + 1e-18
    # The history is the backward error after each step, ending at the
    # reported one.
    hist = [info.history[i] for i in range(info.iterations + 1)]
    assert hist[-1] == info.backward_error and all(h > n * EPS for h in hist[:-1])


def test_single_bf16_product_falls_back_to_fp64(ctx):
    n = 1024
    a = kappa_matrix(77, n)
    b = inputs.gaussian_matrix(78, n, 1)
    st, info, x = run_mixed(ctx, a, b, _capi.PREC_BF16, max_iter=6)
    assert st == _capi.LAIR_ERR_NOCONV
    assert info.fell_back == 1 and info.precision == _capi.PREC_FP64
    assert host_berr(a, x, b) <= n * EPS


def test_fp64_device_solve_and_backward_error_api(ctx):
    n, nrhs = 1500, 3
    a = inputs.gaussian_matrix(5, n, n)
    b = inputs.gaussian_matrix(6, n, nrhs)
    ld = n + 64
    da = L.DeviceArray.from_host(a, ld=ld)
    db = L.DeviceArray.from_host(b)
    dx = L.DeviceArray(n, nrhs)
    L.check(ctx.lib.lair_dgesv_device(ctx.h, n, nrhs, da.ptr, ld, db.ptr, nrhs, dx.ptr, nrhs), "gesv")
    x = dx.download()
    be = host_berr(a, x, b)
    assert be <= n * EPS
    dev = ctypes.c_double(-1.0)
    L.check(ctx.lib.lair_dbackward_error(ctx.h, n, nrhs, da.ptr, ld, db.ptr, nrhs, dx.ptr, nrhs,
                                         ctypes.byref(dev)), "berr")
    assert abs(dev.value - be) <= 0.5 * be + 1e-18


def test_profiling_counts_kernels(ctx):
    n = 512
    a = inputs.gaussian_matrix(8, n, n)
    da = L.DeviceArray.from_host(a)
    dp = L.DeviceArray(n, 1, np.int64)
    ctx.lib.lair_profile_reset(ctx.h)
    ctx.lib.lair_profile_enable(ctx.h, 0xFF)
    sing = ctypes.c_int64(-1)
    L.check(ctx.lib.lair_dgetrf_device(ctx.h, n, n, da.ptr, n, dp.ptr, 128, ctypes.byref(sing)), "getrf")
    ctx.lib.lair_profile_enable(ctx.h, 0)
    ms, cnt, fl = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
    L.check(ctx.lib.lair_profile_read(ctx.h, _capi.KCLASS["gemm"], ctypes.byref(ms), ctypes.byref(cnt),
                                      ctypes.byref(fl), None), "read")
    assert cnt.value > 0 and ms.value > 0 and fl.value > 0
    L.check(ctx.lib.lair_profile_read(ctx.h, _capi.KCLASS["panel"], ctypes.byref(ms), ctypes.byref(cnt),
                                      ctypes.byref(fl), None), "read")
    # FP64 resident leaves are 64 columns wide: two per 128-column panel.
    assert cnt.value == n // 64


@pytest.mark.parametrize("precision", ["fp32", "bf16x6", "fp64"])
def test_python_gesv_mixed(precision):
    n = 700
    a = kappa_matrix(77, n)
    b = inputs.gaussian_matrix(78, n, 3)
    x, info = L.gesv_mixed(a, b, precision=precision)
    assert info["converged"] and not info["fell_back"]
    assert host_berr(a, x, b) <= n * EPS
    x1, info1 = L.gesv_mixed(a, b[:, 0], precision=precision)
    assert x1.shape == (n,) and host_berr(a, x1, b[:, :1]) <= n * EPS
    with pytest.raises(L.ShapeError):
        L.gesv_mixed(a, b[:5])
    sing = a.copy()
    sing[:, 3] = 0.0
    with pytest.raises(L.ValueError_):
        L.gesv_mixed(sing, b, precision="fp64")
