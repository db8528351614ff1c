"""Complex LU on the GPU: the reference's generic getrf / getrs / solve for
A = Complex<f64> and Complex<f32> (src/scalar.rs:310-330), against the
oracle's complex restatements.

Bar: pivots bit-exact against lo_zgetrf_row_major / lo_cgetrf_row_major
(getrf.rs:46-120 with iamax's |re| + |im| key, iamax.rs:14); factors within
1e-8 (f64) or n * 64 eps32 (f32) of the oracle's, P A = L U to n eps; getrs
bit-exact against lo_zgetrs / lo_cgetrs on the same factors for n <= 128
(the reference's operation order, getrs.rs:12-39), backward error <= n eps
beyond.  Fortran-order input: the reference dispatches it to
getrf_col_major (getrf.rs:128-213); the oracle has no complex col-major
restatement, so pivots are compared with the row-major oracle's on the same
matrix (the two agree at every seeded real case: test_oracle_pins).
"""
import ctypes

import numpy as np
import pytest

import lair_amd as L
from lair_amd import _capi, inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS = {np.complex128: np.finfo(np.float64).eps, np.complex64: float(np.finfo(np.float32).eps)}


def cgauss(seed, m, n, dtype=np.complex128):
    re = inputs.gaussian_matrix(seed, m, n)
    im = inputs.gaussian_matrix(seed + 7919, m, n)
    return (re + 1j * im).astype(dtype)


def lu_berr(a, lu, piv):
    m, n = a.shape
    k = min(m, n)
    perm = O.lu_perm(m, piv)
    l = np.tril(lu[:, :k], -1) + np.eye(m, k)
    u = np.triu(lu[:k, :])
    a128 = a.astype(np.complex128)
    return np.abs(a128[perm, :] - l.astype(np.complex128) @ u.astype(np.complex128)).max() / np.abs(a128).max()


def solve_berr(a, x, b):
    a = a.astype(np.complex128)
    x2 = np.asarray(x, np.complex128).reshape(a.shape[0], -1)
    b2 = np.asarray(b, np.complex128).reshape(a.shape[0], -1)
    r = b2 - a @ x2
    return float((np.abs(r).max(0) / (np.abs(a).sum(1).max() * np.abs(x2).max(0))).max())


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_known_square_complex(dtype):  # src/lapack/getrf.rs:485-502
    a = np.array([[1 + 1j, 2 - 1j], [3 + 1j, 4 - 1j]], dtype)
    piv, sing = L.getrf(a)
    assert list(piv) == [1, 1] and sing is None
    want = np.array([[3 + 1j, 4 - 1j], [0.4 + 0.2j, 0.2 - 1.4j]])
    assert np.abs(a - want).max() <= 1e-6


@pytest.mark.parametrize("shape", [(1, 1), (2, 2), (17, 17), (64, 64), (65, 65), (129, 129), (300, 200),
                                   (200, 300), (512, 512), (1025, 1025), (2049, 1500), (1500, 2049)])
def test_zgetrf_pivots_match_oracle(shape):
    m, n = shape
    a = cgauss(5000 + 3 * m + n, m, n)
    ref = a.copy()
    piv_o, sing_o = O.zgetrf(ref)
    g = a.copy()
    piv_g, sing_g = L.getrf(g)
    assert sing_g == sing_o
    np.testing.assert_array_equal(piv_g, piv_o)
    assert np.abs(g - ref).max() <= 1e-8 * max(1.0, np.abs(ref).max())
    assert lu_berr(a, g, piv_g) <= 8 * max(m, n) * EPS[np.complex128]


@pytest.mark.parametrize("shape", [(1, 1), (16, 16), (33, 33), (64, 64), (65, 65), (129, 129), (300, 200),
                                   (200, 300), (513, 513)])
def test_cgetrf_pivots_match_oracle(shape):
    m, n = shape
    a = cgauss(6000 + 3 * m + n, m, n, np.complex64)
    ref = a.copy()
    piv_o, sing_o = O.cgetrf(ref)
    g = a.copy()
    piv_g, sing_g = L.getrf(g)
    assert sing_g == sing_o
    np.testing.assert_array_equal(piv_g, piv_o)
    e32 = EPS[np.complex64]
    assert np.abs(g - ref).max() <= 64 * max(m, n) * e32 * max(1.0, float(np.abs(ref).max()))
    assert lu_berr(a, g, piv_g) <= 8 * max(m, n) * e32


@pytest.mark.parametrize("shape", [(129, 129), (700, 500)])
def test_complex_fortran_order(shape):
    m, n = shape
    a = cgauss(7000 + m, m, n)
    ref = a.copy()
    piv_o, _ = O.zgetrf(ref)
    g = np.asfortranarray(a)
    piv_g, sing_g = L.getrf(g)
    assert sing_g is None
    np.testing.assert_array_equal(piv_g, piv_o)
    assert lu_berr(a, np.array(g), piv_g) <= 8 * max(m, n) * EPS[np.complex128]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_complex_singular_last_zero_pivot_wins(dtype):  # getrf.rs:72-73
    # Exactly zero columns stay exactly zero through every update (a column
    # made dependent in floating point leaves rounding noise instead, whose
    # pivots depend on the summation order), spread over several panels.
    n = 300
    a = cgauss(7100, n, n, dtype)
    a[:, 5] = 0
    a[:, 150] = 0
    a[:, 250] = 0
    ref = a.copy()
    piv_o, sing_o = (O.zgetrf if dtype == np.complex128 else O.cgetrf)(ref)
    g = a.copy()
    piv_g, sing_g = L.getrf(g)
    assert sing_g == sing_o == 250
    np.testing.assert_array_equal(piv_g, piv_o)


@pytest.mark.parametrize("n", [1, 17, 64, 128])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_complex_getrs_matches_reference_order(n, dtype):
    a = cgauss(7200 + n, n, n, dtype)
    b = cgauss(7300 + n, n, 3, dtype)
    g = a.copy()
    piv, _ = L.getrf(g)
    x = L.getrs(g, piv, b)
    fn = O.zgetrs if dtype == np.complex128 else O.cgetrs
    xo = fn(g, piv, b)
    assert x.dtype == dtype
    assert np.array_equal(x.view(np.uint8), xo.view(np.uint8))


@pytest.mark.parametrize("n,nrhs", [(129, 1), (600, 3), (2048, 2), (4100, 1)])
def test_complex_getrs_backward_error(n, nrhs):
    a = cgauss(7400 + n, n, n)
    b = cgauss(7500 + n, n, nrhs)
    g = a.copy()
    piv, _ = L.getrf(g)
    x = L.getrs(g, piv, b)
    assert solve_berr(a, x, b) <= n * EPS[np.complex128]


def test_complex_equation_solve_and_errors():  # src/equation.rs:32-60
    a = np.array([[1 + 1j, 2 - 1j], [3 + 1j, 4 - 1j]])
    b = np.array([3 + 0j, 7 + 2j])
    x = L.equation.solve(a, b)
    assert np.abs(a @ x - b).max() <= 1e-14
    x64 = L.equation.solve(a.astype(np.complex64), b.astype(np.complex64))
    assert x64.dtype == np.complex64 and np.abs(a @ x64 - b).max() <= 1e-5
    with pytest.raises(L.ShapeError):
        L.equation.solve(np.ones((2, 3), complex), np.ones(2, complex))
    with pytest.raises(L.ShapeError):
        L.equation.solve(a, np.ones(3, complex))
    with pytest.raises(L.ValueError_):
        L.equation.solve(np.array([[1 + 1j, 2 + 2j], [2 + 2j, 4 + 4j]]), b)


def test_complex_factorized():  # src/decomposition/lu.rs:12-98
    a = cgauss(7600, 50, 50)
    f = L.lu.Factorized(a)
    assert not f.is_singular()
    p, l, u = f.p(), f.l(), f.u()
    assert np.abs(p @ l @ u - a).max() <= 50 * 8 * EPS[np.complex128] * np.abs(a).max()
    b = cgauss(7601, 50, 1).ravel()
    x = f.solve(b)
    assert solve_berr(a, x, b) <= 50 * EPS[np.complex128]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_complex_device_entry_points(dtype):
    n, nrhs = 1000, 4
    a = cgauss(7700, n, n, dtype)
    b = cgauss(7701, n, nrhs, dtype)
    ctx = L.Context()
    try:
        lib = ctx.lib
        ld = n + 16
        da = L.DeviceArray.from_host(a, ld=ld)
        dp = L.DeviceArray(n, 1, np.int64)
        sing = ctypes.c_int64(-5)
        fgetrf = lib.lair_zgetrf_device if dtype == np.complex128 else lib.lair_cgetrf_device
        fgetrs = lib.lair_zgetrs_device if dtype == np.complex128 else lib.lair_cgetrs_device
        L.check(fgetrf(ctx.h, n, n, da.ptr, ld, dp.ptr, 0, ctypes.byref(sing)), "getrf")
        assert sing.value == -1
        piv = dp.download().ravel()
        g = a.copy()
        piv_h, _ = L.getrf(g)
        np.testing.assert_array_equal(piv, piv_h)
        db = L.DeviceArray.from_host(b)
        L.check(fgetrs(ctx.h, n, nrhs, da.ptr, ld, dp.ptr, db.ptr, nrhs), "getrs")
        x = db.download()
        assert solve_berr(a, x, b) <= 8 * n * EPS[dtype]
        # Out-of-range pivot: reported, and the next call succeeds.
        bad = piv.copy()
        bad[3] = n
        dbad = L.DeviceArray.from_host(bad.reshape(-1, 1))
        assert fgetrs(ctx.h, n, nrhs, da.ptr, ld, dbad.ptr, db.ptr, nrhs) == _capi.LAIR_ERR_SHAPE
        db.upload(b)
        L.check(fgetrs(ctx.h, n, nrhs, da.ptr, ld, dp.ptr, db.ptr, nrhs), "getrs")
        assert np.array_equal(db.download(), x)
    finally:
        ctx.close()
