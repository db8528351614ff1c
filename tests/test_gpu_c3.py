"""The headline mixed-precision workload at its full size, on the GPU.

n = 16384, A = U diag(s) V^T with U, V Haar orthogonal and s log-spaced from
1 to 1e6 (cond_2(A) = 1e6), B with 1 and 64 columns i.i.d. N(0,1), both built
on the device by lair_amd/lib/liblair_gen.so.  The factorization runs in FP32
and in the six-product BF16 split; FP64 iterative refinement must bring every
column to a normwise backward error ||b - A x||inf / (||A||inf ||x||inf)
<= n * eps64 without falling back to the FP64 LU.  The device backward error
is cross-checked on the host with the oracle's long-double residual for a few
columns (the whole A is downloaded once, 2 GiB).
"""
import ctypes
import os

import numpy as np
import pytest

import lair_amd as L
from lair_amd import _capi
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = 16384
EPS = np.finfo(np.float64).eps
GEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lair_amd", "lib",
                   "liblair_gen.so")


@pytest.fixture(scope="module")
def c3():
    gl = ctypes.CDLL(GEN)
    gl.lair_gen_kappa.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                  ctypes.c_uint64]
    gl.lair_gen_gaussian.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_uint64]
    ctx = L.Context()
    ld = N + 64
    da = L.DeviceArray(N, N, np.float64, ld=ld)
    assert gl.lair_gen_kappa(da.ptr, N, ld, 1e6, 16384) == 0
    db = L.DeviceArray(N, 64)
    assert gl.lair_gen_gaussian(db.ptr, N, 64, 64, 16385) == 0
    a_host = da.download()
    b_host = db.download()
    yield ctx, da, ld, b_host, a_host
    da.free()
    db.free()
    ctx.close()


@pytest.mark.parametrize("nrhs", [1, 64])
@pytest.mark.parametrize("prec", [_capi.PREC_FP32, _capi.PREC_BF16X6])
def test_c3_mixed_reaches_fp64_backward_error(c3, prec, nrhs):
    ctx, da, ld, b_all, a = c3
    b = np.ascontiguousarray(b_all[:, :nrhs])
    db = L.DeviceArray.from_host(b)
    dx = L.DeviceArray(N, nrhs)
    info = _capi.IrInfo()
    st = ctx.lib.lair_dgesv_mixed_device(ctx.h, N, nrhs, da.ptr, ld, db.ptr, nrhs, dx.ptr, nrhs, prec, 20,
                                         0.0, ctypes.byref(info))
    assert st == _capi.LAIR_OK, _capi.last_error()
    assert info.converged == 1 and info.fell_back == 0
    assert info.backward_error <= N * EPS
    dev = ctypes.c_double(-1.0)
    L.check(ctx.lib.lair_dbackward_error(ctx.h, N, nrhs, da.ptr, ld, db.ptr, nrhs, dx.ptr, nrhs,
                                         ctypes.byref(dev)), "berr")
    assert dev.value <= N * EPS
    x = dx.download()
    cols = sorted({0, nrhs // 2, nrhs - 1})
    r = O.residual(a, np.ascontiguousarray(x[:, cols]), np.ascontiguousarray(b[:, cols]), nthreads=16)
    host = (np.abs(r).max(0) / (np.abs(a).sum(1).max() * np.abs(x[:, cols]).max(0))).max()
    assert host <= N * EPS
    # The device's FP64 norms agree with the long-double host check.
    assert host <= dev.value * 1.5 + 1e-18
    db.free()
    dx.free()
