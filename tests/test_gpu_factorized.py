"""The device-resident lu::Factorized (src/decomposition/lu.rs:12-171) through
the C ABI's lair_dlu_* handle and the Python mirror lu.Factorized (float64).

* p, l, u, into_pl against their definitions (lu.rs:28-72, 107-153): L and U
  are the triangles of the handle's factors, P is the permutation of the
  oracle's pivots applied by laswp, P L U reproduces A to the factorization's
  backward error, into_pl equals P @ L exactly;
* solve (lu.rs:87-98): bit-exact against the oracle's getrs on the same
  factors for n <= 128 (the reference's operation order), backward error
  <= n*eps beyond;
* is_singular (lu.rs:75-77) and the last zero-pivot column;
* multi-RHS and transposed solves and the adjoint from the handle, against
  the context-less entry points on the same factors (bit-identical: the same
  kernels, with the permutation and diagonal inverses reused);
* the reference's Shape error for a right-hand side of the wrong length.
"""
import ctypes

import numpy as np
import pytest

import lair_amd as L
from lair_amd import _capi, inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def berr(a, x, b):
    r = O.residual(a, x, b, nthreads=4)
    x2 = np.asarray(x).reshape(a.shape[0], -1)
    return float((np.abs(r).max(0) / (np.abs(a).sum(1).max() * np.abs(x2).max(0))).max())


@pytest.mark.parametrize("shape", [(5, 5), (7, 4), (4, 7), (130, 130), (300, 200), (200, 300)])
def test_factor_extraction_identities(shape):
    m, n = shape
    a = inputs.gaussian_matrix(900 + m + n, m, n)
    f = L.lu.Factorized(a)
    ref = a.copy()
    piv_o, _ = O.getrf(ref)
    assert np.array_equal(f.pivots, piv_o)
    p, l, u = f.p(), f.l(), f.u()
    k = min(m, n)
    assert p.shape == (m, m) and l.shape == (m, k) and u.shape == (k, n)
    # P from the pivots applied to 0..m by laswp (lu.rs:28-39)
    perm = np.arange(m)
    for i, q in enumerate(piv_o):
        perm[[i, q]] = perm[[q, i]]
    pe = np.zeros((m, m))
    pe[perm, np.arange(m)] = 1.0
    assert np.array_equal(p, pe)
    lu = f.lu
    assert np.array_equal(np.tril(l, -1), np.tril(lu[:, :k], -1))
    assert np.array_equal(np.diag(l), np.ones(k)) and np.all(np.triu(l, 1) == 0)
    assert np.array_equal(u, np.triu(lu[:k, :]))
    assert np.abs(lu - ref).max() <= 1e-8 * max(1.0, np.abs(ref).max())
    resid = np.abs(p @ l @ u - a).max() / np.abs(a).max()
    assert resid <= 16 * max(m, n) * EPS
    pl = L.lu.Factorized(a).into_pl()
    assert np.array_equal(pl[:, :k], p @ l)
    if n > k:
        assert np.array_equal(pl[:, k:], lu[:, k:])


@pytest.mark.parametrize("n", [1, 17, 128, 129, 600])
def test_solve_matches_reference_getrs(n):
    a = inputs.gaussian_matrix(1000 + n, n, n)
    b = inputs.gaussian_matrix(1001 + n, n, 1).ravel()
    f = L.lu.Factorized(a)
    x = f.solve(b)
    assert x.shape == b.shape and not f.is_singular()
    # The oracle's getrs on the handle's own factors: same operands, so the
    # reference's operation order (n <= 128) gives the same bits.
    xo = O.getrs(f.lu, f.pivots, b.reshape(n, 1)).ravel()
    if n <= 128:
        assert np.array_equal(x.view(np.int64), xo.view(np.int64))
    else:
        assert berr(a, x, b) <= n * EPS


def test_is_singular_and_last_zero_pivot():
    a = inputs.gaussian_matrix(77, 40, 40)
    a[:, 7] = 0.0
    a[:, 23] = 2.0 * a[:, 3]
    f = L.lu.Factorized(a)
    ref = a.copy()
    piv, sing = O.getrf(ref)
    assert f.is_singular() and f.singular == sing
    assert np.array_equal(f.pivots, piv)
    g = L.lu.Factorized(inputs.gaussian_matrix(78, 40, 40))
    assert not g.is_singular() and g.singular is None


def test_solve_shape_error():
    f = L.lu.Factorized(inputs.gaussian_matrix(79, 6, 6))
    with pytest.raises(L.ShapeError):
        f.solve(np.ones(5))


def _handle(ctx, a):
    n = a.shape[0]
    da = L.DeviceArray.from_host(a)
    h = ctypes.c_void_p()
    L.check(ctx.lib.lair_dlu_factor(ctx.h, n, n, da.ptr, n, 0, ctypes.byref(h)), "lu_factor")
    return da, h


@pytest.mark.parametrize("n,nrhs,trans", [(100, 3, 1), (700, 8, 0), (700, 5, 1), (2048, 64, 0)])
def test_handle_solve_matches_getrs_device(n, nrhs, trans):
    ctx = L.Context()
    try:
        lib = ctx.lib
        a = inputs.gaussian_matrix(1100 + n, n, n)
        b = inputs.gaussian_matrix(1101 + n, n, nrhs)
        da, h = _handle(ctx, a)
        lu_p, ld, piv_p = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_void_p()
        L.check(lib.lair_dlu_factors(h, ctypes.byref(lu_p), ctypes.byref(ld), ctypes.byref(piv_p), None), "f")
        db = L.DeviceArray.from_host(b)
        for _ in range(2):  # the second solve reuses the diagonal inverses
            db.upload(b)
            L.check(lib.lair_dlu_solve(ctx.h, h, nrhs, db.ptr, nrhs, trans), "lu_solve")
        x = db.download()
        db.upload(b)
        L.check(lib.lair_dgetrs_device(ctx.h, n, nrhs, lu_p, ld.value, piv_p, db.ptr, nrhs, trans), "getrs")
        assert np.array_equal(x.view(np.int64), db.download().view(np.int64))
        at = a.T.copy() if trans else a
        assert berr(at, x, b) <= n * EPS
        lib.lair_dlu_destroy(h)
    finally:
        ctx.close()


def test_handle_solve_backward_matches_adjoint():
    ctx = L.Context()
    try:
        lib = ctx.lib
        n, nrhs = 1000, 16
        a = inputs.gaussian_matrix(1200, n, n) + 4 * np.sqrt(n) * np.eye(n)
        x = inputs.gaussian_matrix(1201, n, nrhs)
        xb = inputs.gaussian_matrix(1202, n, nrhs)
        da, h = _handle(ctx, a)
        lu_p, ld, piv_p = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_void_p()
        L.check(lib.lair_dlu_factors(h, ctypes.byref(lu_p), ctypes.byref(ld), ctypes.byref(piv_p), None), "f")
        dx, dxb = L.DeviceArray.from_host(x), L.DeviceArray.from_host(xb)
        bb1, ab1, bb2, ab2 = (L.DeviceArray(n, nrhs), L.DeviceArray(n, n), L.DeviceArray(n, nrhs),
                              L.DeviceArray(n, n))
        L.check(lib.lair_dlu_solve_backward(ctx.h, h, nrhs, dx.ptr, nrhs, dxb.ptr, nrhs, bb1.ptr, nrhs, ab1.ptr,
                                            n), "lu_backward")
        L.check(lib.lair_dsolve_backward_device(ctx.h, n, nrhs, lu_p, ld.value, piv_p, dx.ptr, nrhs, dxb.ptr,
                                                nrhs, bb2.ptr, nrhs, ab2.ptr, n), "backward")
        assert np.array_equal(bb1.download().view(np.int64), bb2.download().view(np.int64))
        assert np.array_equal(ab1.download().view(np.int64), ab2.download().view(np.int64))
        bbar = bb1.download()
        assert np.abs(a.T @ bbar - xb).max() / (np.abs(a).sum(0).max() * np.abs(bbar).max()) <= n * EPS
        lib.lair_dlu_destroy(h)
    finally:
        ctx.close()


def test_handle_errors():
    ctx = L.Context()
    try:
        lib = ctx.lib
        a = inputs.gaussian_matrix(1300, 5, 7)
        da = L.DeviceArray.from_host(a)
        h = ctypes.c_void_p()
        L.check(lib.lair_dlu_factor(ctx.h, 5, 7, da.ptr, 7, 0, ctypes.byref(h)), "lu_factor")
        db = L.DeviceArray(5, 1)
        assert lib.lair_dlu_solve(ctx.h, h, 1, db.ptr, 1, 0) == _capi.LAIR_ERR_SHAPE
        assert lib.lair_dlu_factor(ctx.h, 0, 7, da.ptr, 7, 0, ctypes.byref(h)) < 0
        lib.lair_dlu_destroy(h)
    finally:
        ctx.close()
