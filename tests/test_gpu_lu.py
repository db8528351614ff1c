"""GPU parity of the FP64/FP32 LU path against the CPU oracle and the
reference's known answers.

Bar: pivot vectors bit-exact (integer work), L/U/x within a normwise backward
error of order n*eps (stated per test).  All calls go through the C ABI of
liblair_dev.so.
"""
import json
import os

import numpy as np
import pytest

import lair_amd as L
from lair_amd import inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = np.finfo(np.float64).eps


def lu_backward_error(a, lu, piv):
    """||P A - L U||_inf / ||A||_inf with P the forward product of the swaps."""
    m, n = a.shape
    k = min(m, n)
    pa = a.copy()
    for i, p in enumerate(piv):
        if p != i:
            pa[[i, p]] = pa[[p, i]]
    l = np.tril(lu[:, :k], -1)
    l[np.arange(k), np.arange(k)] = 1.0
    u = np.triu(lu[:k, :])
    return np.abs(pa - l @ u).sum(1).max() / max(np.abs(a).sum(1).max(), 1e-300)


def solve_backward_error(a, x, b):
    r = O.residual(a, x, b, nthreads=8)
    x2 = np.asarray(x).reshape(a.shape[0], -1)
    return float((np.abs(r).max(0) / (np.abs(a).sum(1).max() * np.abs(x2).max(0))).max())


# ---------------------------------------------------- reference known answers

A_5X5 = np.array([[1., 2., 3., 1., 2.], [2., 2., 1., 3., 3.], [3., 1., 2., 2., 1.],
                  [2., 3., 3., 1., 1.], [1., 3., 1., 3., 1.]])
LU_5X5 = np.array([
    [3., 1., 2., 2., 1.],
    [0.33333333, 2.66666667, 0.33333333, 2.33333333, 0.66666667],
    [0.33333333, 0.625, 2.125, -1.125, 1.25],
    [0.66666667, 0.875, 0.64705882, -1.64705882, -1.05882353],
    [0.66666667, 0.5, -0.23529412, -0.14285714, 2.14285714],
])


def test_known_square_row_major():  # src/lapack/getrf.rs:405-428
    a = A_5X5.copy()
    piv, sing = L.getrf(a)
    assert list(piv) == [2, 4, 2, 3, 4] and sing is None
    np.testing.assert_allclose(a, LU_5X5, atol=1e-6, rtol=0)


def test_known_square_col_major():  # getrf.rs:430-455
    base = np.array([[1., 2., 3., 2., 1.], [2., 2., 1., 3., 3.], [3., 1., 2., 3., 1.],
                     [1., 3., 2., 1., 3.], [2., 3., 1., 1., 1.]])
    a = base.T
    piv, sing = L.getrf(a)
    assert list(piv) == [2, 4, 2, 3, 4] and sing is None
    np.testing.assert_allclose(a, LU_5X5, atol=1e-6, rtol=0)


def test_known_negative_strides():  # getrf.rs:457-483
    base = np.array([[1., 3., 1., 3., 1.], [1., 1., 3., 3., 2.], [1., 2., 2., 1., 3.],
                     [3., 3., 1., 2., 2.], [2., 1., 3., 2., 1.]])
    a = base[::-1, ::-1]
    piv, sing = L.getrf(a)
    assert list(piv) == [2, 4, 2, 3, 4] and sing is None
    np.testing.assert_allclose(a, LU_5X5, atol=1e-6, rtol=0)


def test_known_singular_empty_smallest():  # getrf.rs:342-379
    a = np.array([[1., 1.], [1., 1.]])
    assert L.getrf(a)[1] == 1
    piv, sing = L.getrf(np.zeros((0, 0)))
    assert len(piv) == 0 and sing is None
    a = np.array([[3.]])
    piv, sing = L.getrf(a)
    assert list(piv) == [0] and sing is None and a[0, 0] == 3.0


def test_known_wide_tall():  # getrf.rs:504-524
    a = np.array([[1., 2., 3.], [2., 3., 4.]])
    piv, sing = L.getrf(a)
    assert list(piv) == [1, 1] and sing is None
    assert (a == np.array([[2., 3., 4.], [0.5, 0.5, 1.]])).all()
    a = np.array([[1., 2.], [1., 3.], [2., 3.]])
    piv, sing = L.getrf(a)
    assert list(piv) == [2, 1] and sing is None
    np.testing.assert_allclose(a, [[2., 3.], [0.5, 1.5], [0.5, 0.3333333333333333]], atol=1e-6, rtol=0)


def test_known_getrs():  # src/lapack/getrs.rs:46-78
    a = np.array([[1., 2.], [3., 4.]])
    piv, _ = L.getrf(a)
    assert list(piv) == [1, 1]
    np.testing.assert_allclose(L.getrs(a, piv, np.array([3., 7.])), [1., 1.], rtol=1e-8)
    a = A_5X5.copy()
    piv, _ = L.getrf(a)
    x = L.getrs(a, piv, np.array([9., 11., 9., 10., 9.]))
    assert np.all(np.abs(x - 1.0) <= EPS)


def test_known_equation_solve():  # src/equation.rs:21-31, 42-57
    x = L.equation.solve(np.array([[3., 1.], [1., 2.]]), np.array([9., 8.]))
    assert list(x) == [2.0, 3.0]
    with pytest.raises(L.ShapeError):
        L.equation.solve(np.zeros((2, 3)), np.zeros(2))
    with pytest.raises(L.ShapeError):
        L.equation.solve(np.eye(3), np.zeros(2))
    with pytest.raises(L.ValueError_):
        L.equation.solve(np.array([[1., 1.], [1., 1.]]), np.ones(2))


def test_known_lu_f32_accessors():  # src/decomposition/lu.rs:181-254
    f = L.lu.Factorized(np.array([[1, 2, 3], [2, 2, 1], [3, 1, 2]], np.float32))
    p, l, u = f.p(), f.l(), f.u()
    assert p[0, 1] == 1 and p[1, 2] == 1 and p[2, 0] == 1
    np.testing.assert_allclose([l[1, 0], l[2, 0]], [0.33333333, 0.6666667], rtol=1e-6)
    np.testing.assert_allclose([u[0, 2], u[1, 2], u[2, 2]], [2., 2.3333333, -2.2], rtol=1e-6)
    f = L.lu.Factorized(np.array([[1, 2, 3], [2, 2, 1], [3, 1, 2], [2, 3, 3]], np.float32))
    p, u = f.p(), f.u()
    assert p[0, 3] == 1 and p[1, 2] == 1 and p[2, 0] == 1 and p[3, 1] == 1
    np.testing.assert_allclose([u[0, 2], u[1, 2], u[2, 2]], [2., 1.6666666, -1.2857143], rtol=1e-6)


def test_known_into_pl():  # lu.rs:256-312
    m = np.array([[0, 0], [3, 4], [6, 8]], np.float32)
    pl = L.lu.Factorized(m).into_pl()
    assert (pl[:, :2] == np.array([[0., 0.], [0.5, 1.], [1., 0.]])).all()
    m = np.array([[0, 1, 2], [1, 2, 3], [2, 3, 4]], np.float32)
    pl = L.lu.Factorized(m).into_pl()
    assert (pl[:, :3] == np.array([[0., 1., 0.], [0.5, 0.5, 1.], [1., 0., 0.]])).all()


# ------------------------------------------------------- seeded parity ----

@pytest.mark.parametrize("shape", [(1, 1), (2, 2), (15, 15), (16, 16), (17, 17), (63, 63), (64, 64),
                                   (65, 65), (129, 129), (257, 257), (300, 200), (200, 300),
                                   (1025, 1025), (2049, 1500), (1500, 2049)])
def test_pivots_match_oracle(shape):
    m, n = shape
    a = inputs.gaussian_matrix(1000 + m * 7 + n, m, n)
    ref = a.copy()
    piv_o, sing_o = O.getrf(ref, nthreads=8)
    g = a.copy()
    piv_g, sing_g = L.getrf(g)
    assert sing_g == sing_o
    np.testing.assert_array_equal(piv_g, piv_o)
    tol = 8 * max(m, n) * EPS
    assert lu_backward_error(a, g, piv_g) <= tol
    assert np.max(np.abs(g - ref)) <= 1e-8 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("shape", [(17, 17), (129, 129), (300, 200), (200, 300), (1025, 1025), (2049, 1500)])
@pytest.mark.parametrize("layout", ["F", "neg"])
def test_pivots_match_col_major_oracle(shape, layout):
    """Non-standard layouts dispatch to getrf_col_major (getrf.rs:128-213,
    the dispatch at getrf.rs:17-26): seeded Fortran-order and negative-stride
    views against the oracle's col-major restatement on the same view."""
    m, n = shape
    a = inputs.gaussian_matrix(3000 + m * 5 + n, m, n)

    def view(x):
        if layout == "F":
            return np.asfortranarray(x)
        return np.ascontiguousarray(x[::-1, ::-1])[::-1, ::-1]

    ref = view(a)
    assert not O.is_standard_layout(m, n, ref.strides[0] // 8, ref.strides[1] // 8)
    piv_o, sing_o = O.getrf(ref, nthreads=8)
    g = view(a)
    piv_g, sing_g = L.getrf(g)
    assert sing_g == sing_o
    np.testing.assert_array_equal(piv_g, piv_o)
    assert lu_backward_error(a, np.array(g), piv_g) <= 8 * max(m, n) * EPS
    assert np.max(np.abs(g - ref)) <= 1e-8 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("shape", [(1, 1), (16, 16), (33, 33), (64, 64), (65, 65), (129, 129), (257, 257),
                                   (300, 200), (200, 300), (513, 513)])
def test_sgetrf_pivots_match_oracle(shape):
    """FP32 dense LU (getrf.rs:46-120 for f32) through lair_sgetrf against the
    oracle's f32 restatement: pivots bit-exact; factors within n * eps32."""
    m, n = shape
    a = inputs.gaussian_matrix(4000 + m * 3 + n, m, n).astype(np.float32)
    ref = a.copy()
    piv_o, sing_o = O.sgetrf(ref, nthreads=8)
    g = a.copy()
    piv_g, sing_g = L.getrf(g)
    assert sing_g == sing_o
    np.testing.assert_array_equal(piv_g, piv_o)
    eps32 = float(np.finfo(np.float32).eps)
    assert lu_backward_error(a.astype(np.float64), g.astype(np.float64), piv_g) <= 8 * max(m, n) * eps32
    assert np.max(np.abs(g - ref)) <= 64 * max(m, n) * eps32 * max(1.0, float(np.abs(ref).max()))


def test_c1_fixture_n512():
    fx = json.load(open(os.path.join(HERE, "golden", "c1_n512.json")))
    a = inputs.gaussian_matrix(fx["seed_a"], fx["n"], fx["n"])
    b = inputs.gaussian_matrix(fx["seed_b"], fx["n"], fx["nrhs"])
    assert inputs.digest(a) == fx["digest_a"] and inputs.digest(b) == fx["digest_b"]
    g = a.copy()
    piv, sing = L.getrf(g)
    assert sing is None
    np.testing.assert_array_equal(piv, np.array(fx["pivots"]))
    x = L.getrs(g, piv, b)
    assert solve_backward_error(a, x, b) <= fx["n"] * EPS
    assert lu_backward_error(a, g, piv) <= 8 * fx["n"] * EPS


@pytest.mark.parametrize("nb", [0, 64, 256, -1])  # default width, fixed widths, autotuned
def test_c2_fixture_n8192_device(nb):
    path = os.path.join(HERE, "golden", "c2_n8192.json")
    fx = json.load(open(path))
    n, nrhs = fx["n"], fx["nrhs"]
    a = inputs.gaussian_matrix(fx["seed_a"], n, n)
    b = inputs.gaussian_matrix(fx["seed_b"], n, nrhs)
    assert inputs.digest(a) == fx["digest_a"] and inputs.digest(b) == fx["digest_b"]
    with L.Context() as ctx:
        ld = n + 64
        da = L.DeviceArray.from_host(a, ld=ld)
        dp = L.DeviceArray(n, 1, np.int64)
        import ctypes
        sing = ctypes.c_int64(-1)
        L.check(ctx.lib.lair_dgetrf_device(ctx.h, n, n, da.ptr, ld, dp.ptr, nb, ctypes.byref(sing)), "getrf")
        assert sing.value == -1
        piv = dp.download().reshape(-1)
        np.testing.assert_array_equal(piv, np.array(fx["pivots"]))
        db = L.DeviceArray.from_host(b)
        L.check(ctx.lib.lair_dgetrs_device(ctx.h, n, nrhs, da.ptr, ld, dp.ptr, db.ptr, nrhs, 0), "getrs")
        ctx.synchronize()
        x = db.download()
    assert solve_backward_error(a, x, b) <= n * EPS


@pytest.mark.parametrize("n,nrhs", [(100, 3), (700, 5), (1100, 64)])
def test_getrs_trans_and_multi_rhs(n, nrhs):
    a = inputs.gaussian_matrix(77 + n, n, n)
    b = inputs.gaussian_matrix(78 + n, n, nrhs)
    g = a.copy()
    piv, _ = L.getrf(g)
    x = L.getrs(g, piv, b)
    assert solve_backward_error(a, x, b) <= n * EPS
    y = L.getrs(g, piv, b, trans=True)
    assert solve_backward_error(np.ascontiguousarray(a.T), y, b) <= n * EPS
    yo = O.getrs_trans(g, piv, b)
    np.testing.assert_allclose(y, yo, rtol=0, atol=1e-9 * np.abs(yo).max())


@pytest.mark.parametrize("n", [20500, 33000])
def test_getrs_large_caller_pivots_permutation(n):
    """getrs from caller-supplied factors composes its row permutation from the
    pivots (laswp.rs:21-39 order).  With identity factors the solve returns
    P b exactly, so the composition is checked beyond the sizes whose pivots
    (n > 20480) or permutation (n > 32768) still fit LDS."""
    rng = np.random.default_rng(n)
    piv = np.array([rng.integers(i, n) for i in range(n)], np.int64)
    b = rng.standard_normal((n, 1)).astype(np.float32)
    expect = b.copy()
    for i, q in enumerate(piv):  # getrs.rs:23, laswp.rs:28-38
        if q != i:
            expect[[i, q]] = expect[[q, i]]
    ctx = L.Context()
    lu = L.DeviceArray.from_host(np.eye(n, dtype=np.float32))
    dpiv = L.DeviceArray.from_host(piv.reshape(-1, 1))
    db = L.DeviceArray.from_host(b)
    try:
        L.check(ctx.lib.lair_sgetrs_device(ctx.h, n, 1, lu.ptr, n, dpiv.ptr, db.ptr, 1, 0), "sgetrs")
        ctx.synchronize()
        np.testing.assert_array_equal(db.download(), expect)
    finally:
        for d in (lu, dpiv, db):
            d.free()
        ctx.close()


def test_singular_last_zero_pivot_wins():  # getrf.rs:72-73
    a = np.array([[0., 0., 1.], [0., 0., 2.], [0., 0., 3.]])
    ref = a.copy()
    piv_o, sing_o = O.getrf(ref)
    piv, sing = L.getrf(a)
    assert sing == sing_o == 1
    np.testing.assert_array_equal(piv, piv_o)
    # a larger matrix with duplicated columns spread over several panels
    n = 300
    a = inputs.gaussian_matrix(5, n, n)
    a[:, 40] = 0.0
    a[:, 250] = 0.0
    ref = a.copy()
    piv_o, sing_o = O.getrf(ref)
    piv, sing = L.getrf(a)
    assert sing == sing_o == 250
    np.testing.assert_array_equal(piv, piv_o)


def test_nan_never_selected_as_pivot():  # iamax.rs:15
    a = inputs.gaussian_matrix(9, 40, 40)
    a[3, 0] = np.nan
    ref = a.copy()
    piv_o, _ = O.getrf(ref)
    g = a.copy()
    piv, _ = L.getrf(g)
    assert piv[0] == piv_o[0] and piv[0] != 3


def test_laswp_matches_oracle():  # laswp.rs:11-40
    a = inputs.gaussian_matrix(3, 10, 6)
    piv = np.array([3, 9, 2, 8, 4], np.int64)
    r1 = a.copy()
    O.laswp(r1, 1, piv)
    r2 = a.copy()
    L.laswp(6, r2, 1, piv)
    assert (r1 == r2).all()


def test_ties_pick_lowest_row():  # iamax.rs:15 strict '>' keeps the first maximum
    n = 2100  # rows spread over three cooperating workgroups
    a = inputs.gaussian_matrix(4, n, 16) * 0.5
    a[:, 0] = 0.25
    a[7, 0] = -3.0
    a[1500, 0] = 3.0
    a[2000, 0] = -3.0
    ref = a.copy()
    piv_o, _ = O.getrf(ref)
    g = a.copy()
    piv, _ = L.getrf(g)
    assert piv[0] == 7 == piv_o[0]
    np.testing.assert_array_equal(piv, piv_o)


@pytest.mark.parametrize("n,nrhs", [(300, 16), (1000, 17), (2049, 64), (129, 33)])
def test_sgetrs_many_rhs(n, nrhs):
    """FP32 getrs with 16 or more right-hand sides (the matrix-core dataflow
    triangular solve): backward error of order n * eps32 against the FP32
    factors, partial 128-row blocks and partial 16-column chunks included;
    every column equals the one-column solve to rounding."""
    a = inputs.gaussian_matrix(5100 + n, n, n).astype(np.float32)
    b = inputs.gaussian_matrix(5101 + n, n, nrhs).astype(np.float32)
    g = a.copy()
    piv, sing = L.getrf(g)
    assert sing is None
    x = L.getrs(g, piv, b)
    eps32 = float(np.finfo(np.float32).eps)
    a64, x64, b64 = a.astype(np.float64), x.astype(np.float64), b.astype(np.float64)
    r = b64 - a64 @ x64
    berr = (np.abs(r).max(0) / (np.abs(a64).sum(1).max() * np.abs(x64).max(0))).max()
    assert berr <= 8 * n * eps32
    x1 = L.getrs(g, piv, b[:, 5])
    assert np.abs(x1 - x[:, 5]).max() <= 64 * n * eps32 * np.abs(x1).max()
