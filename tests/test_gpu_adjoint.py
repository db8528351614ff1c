"""Adjoint of X = A^-1 B (solve_backward) on the GPU.

Checks, for f(A, B) = <XBAR, A^-1 B>:
  * against the closed form in FP64 on the CPU (numpy/LAPACK):
    BBAR = A^-T XBAR, ABAR = -BBAR X^T;
  * against central finite differences along 16 seeded directions (dA, dB):
    (f(A + h dA, B + h dB) - f(A - h dA, B - h dB)) / 2h  vs  <ABAR, dA> + <BBAR, dB>,
    with f evaluated by the FP64 GPU solver.
Workload of the adjoint configuration: A = N(0,1) + 4 sqrt(n) I, B and XBAR N(0,1).
"""
import ctypes

import numpy as np
import pytest

import lair_amd as L
from lair_amd import inputs

pytestmark = pytest.mark.gpu


def make(n, nrhs, seed):
    a = inputs.gaussian_matrix(seed, n, n) + 4.0 * np.sqrt(n) * np.eye(n)
    b = inputs.gaussian_matrix(seed + 1, n, nrhs)
    xbar = inputs.gaussian_matrix(seed + 2, n, nrhs)
    return a, b, xbar


@pytest.mark.parametrize("n,nrhs", [(200, 3), (513, 64)])
def test_adjoint_matches_closed_form(n, nrhs):
    a, b, xbar = make(n, nrhs, 40 + n)
    lu = a.copy()
    piv, sing = L.getrf(lu)
    assert sing is None
    x = L.getrs(lu, piv, b)
    bbar, abar = L.solve_backward(lu, piv, x, xbar)
    bbar_ref = np.linalg.solve(a.T, xbar)
    abar_ref = -bbar_ref @ x.T
    assert np.abs(bbar - bbar_ref).max() <= 1e-12 * np.abs(bbar_ref).max()
    assert np.abs(abar - abar_ref).max() <= 1e-12 * np.abs(abar_ref).max()


def test_adjoint_c5_finite_differences_n4096():
    n, nrhs, ndir = 4096, 64, 16
    a, b, xbar = make(n, nrhs, 4096)
    ld = n + 64
    with L.Context() as ctx:
        lib = ctx.lib
        da = L.DeviceArray.from_host(a, ld=ld)
        db = L.DeviceArray.from_host(b)
        dx = L.DeviceArray(n, nrhs)
        # factors kept by the solver: factor a copy explicitly for the adjoint
        dlu = L.DeviceArray.from_host(a, ld=ld)
        dp = L.DeviceArray(n, 1, np.int64)
        sing = ctypes.c_int64(-1)
        L.check(lib.lair_dgetrf_device(ctx.h, n, n, dlu.ptr, ld, dp.ptr, 0, ctypes.byref(sing)), "getrf")
        assert sing.value == -1
        L.check(lib.lair_dgesv_device(ctx.h, n, nrhs, da.ptr, ld, db.ptr, nrhs, dx.ptr, nrhs), "gesv")
        x = dx.download()
        dxb = L.DeviceArray.from_host(xbar)
        dbb = L.DeviceArray(n, nrhs)
        dab = L.DeviceArray(n, n)
        L.check(lib.lair_dsolve_backward_device(ctx.h, n, nrhs, dlu.ptr, ld, dp.ptr, dx.ptr, nrhs, dxb.ptr, nrhs,
                                                dbb.ptr, nrhs, dab.ptr, n), "solve_backward")
        bbar = dbb.download()
        abar = dab.download()

        # closed form in FP64 on the CPU
        bbar_ref = np.linalg.solve(a.T, xbar)
        assert np.abs(bbar - bbar_ref).max() <= 1e-12 * np.abs(bbar_ref).max()
        abar_ref = -bbar_ref @ x.T
        assert np.abs(abar - abar_ref).max() <= 1e-12 * np.abs(abar_ref).max()

        def f(a_, b_):
            da.upload(a_)
            db.upload(b_)
            L.check(lib.lair_dgesv_device(ctx.h, n, nrhs, da.ptr, ld, db.ptr, nrhs, dx.ptr, nrhs), "gesv")
            return float(np.sum(xbar * dx.download()))

        h = 1e-4
        rng_seed = 777
        for k in range(ndir):
            dA = inputs.gaussian_matrix(rng_seed + 2 * k, n, n)
            dB = inputs.gaussian_matrix(rng_seed + 2 * k + 1, n, nrhs)
            fd = (f(a + h * dA, b + h * dB) - f(a - h * dA, b - h * dB)) / (2 * h)
            pred = float(np.sum(abar * dA) + np.sum(bbar * dB))
            assert abs(fd - pred) <= 1e-6 * (abs(pred) + np.sqrt(np.sum(abar ** 2)) * h), (k, fd, pred)
