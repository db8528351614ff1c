"""Pins of the oracle beyond the reference's own known answers.

* Pivots of the oracle's getrf_row_major (src/lapack/getrf.rs:46-120) and
  getrf_col_major (getrf.rs:128-213) equal LAPACK's dgetrf pivots
  (scipy.linalg.lu_factor, an independent implementation of the same partial
  pivoting rule: first maximum of |a| in the column) on seeded inputs, and the
  committed C1 (n=512) and C2 (n=8192) fixture pivots equal LAPACK's: the
  fixtures are the oracle's output, so this pins the oracle at the full C2
  size too.
* P A = L U to the factorization's backward error, P built by lu_perm
  (src/decomposition/lu.rs:28-39).
* trsm (src/blas/trsm.rs:6-22) and gemm (src/blas/gemm.rs:6-32) against brute
  force on integer-valued operands, where every rounding vanishes and the
  result must be exact: a transposed operand, a sign, a dropped alpha or an
  index slip changes it.  Also on strided and negative-stride views.
* getrs (src/lapack/getrs.rs:12-39) against a dense solve of P L U.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl

from lair_amd import inputs
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = np.finfo(np.float64).eps


def lapack_pivots(a):
    return sl.lu_factor(a, check_finite=False)[1].astype(np.int64)


@pytest.mark.parametrize("name", ["c1_n512", "c2_n8192"])
def test_fixture_pivots_equal_lapack(name):
    fx = json.load(open(os.path.join(HERE, "golden", name + ".json")))
    n = fx["n"]
    a = inputs.gaussian_matrix(fx["seed_a"], n, n)
    assert inputs.digest(a) == fx["digest_a"]
    assert np.array_equal(lapack_pivots(a), np.array(fx["pivots"], np.int64))


@pytest.mark.parametrize("shape", [(1, 1), (7, 7), (64, 64), (129, 129), (300, 200), (200, 300), (512, 512),
                                   (1025, 1025)])
@pytest.mark.parametrize("order", ["C", "F", "neg"])
def test_oracle_pivots_equal_lapack(shape, order):
    m, n = shape
    a = inputs.gaussian_matrix(2000 + m + 3 * n, m, n)
    if order == "C":
        w = a.copy()
    elif order == "F":
        w = np.asfortranarray(a)
    else:  # the same logical matrix seen through negative strides
        store = np.ascontiguousarray(a[::-1, ::-1])
        w = store[::-1, ::-1]
        assert min(m, n) == 1 or (w.strides[0] < 0 and w.strides[1] < 0)
    if min(m, n) > 1:  # which of the reference's two kernels runs
        assert O.is_standard_layout(m, n, *(s // 8 for s in w.strides)) == (order == "C")
    piv, sing = O.getrf(w, nthreads=4)
    assert sing is None
    k = min(m, n)
    assert np.array_equal(piv, lapack_pivots(a)[:k])
    perm = O.lu_perm(m, piv)
    l = np.tril(w[:, :k], -1) + np.eye(m, k)
    u = np.triu(w[:k, :])
    pa = a[perm, :]
    assert np.abs(pa - l @ u).max() <= 8 * max(m, n) * EPS * np.abs(a).max()
    # The factors agree with LAPACK's to rounding.
    lu_l = sl.lu_factor(a, check_finite=False)[0] if m == n else None
    if lu_l is not None:
        assert np.abs(lu_l - w).max() <= 1e-9 * max(1.0, np.abs(w).max())


def test_col_major_and_row_major_agree_bitwise_on_symmetric_roles():
    # getrf_col_major on A^T's storage read as column-major is getrf_row_major
    # on A: the same logical matrix, two layouts, one algorithm's pivots.
    a = inputs.gaussian_matrix(2100, 300, 300)
    r = a.copy()
    c = np.asfortranarray(a)
    pr, _ = O.getrf(r)
    pc, _ = O.getrf(c)
    assert np.array_equal(pr, pc)
    assert np.abs(r - c).max() <= 1e-10


def test_lu_perm_matches_sequential_swaps():
    rng = np.random.default_rng(5)
    m = 40
    piv = np.array([rng.integers(i, m) for i in range(30)], np.int64)
    perm = np.arange(m)
    for i, q in enumerate(piv):
        perm[[i, q]] = perm[[q, i]]
    assert np.array_equal(O.lu_perm(m, piv), perm)


def _int_matrix(rng, m, n, lo=-4, hi=5):
    return rng.integers(lo, hi, size=(m, n)).astype(np.float64)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 3, 4), (17, 9, 33), (64, 65, 63)])
@pytest.mark.parametrize("alpha", [1.0, -1.0, 3.0])
def test_gemm_brute_force_integers(m, n, k, alpha):
    rng = np.random.default_rng(m * 100 + n * 10 + k)
    a, b, c = _int_matrix(rng, m, k), _int_matrix(rng, k, n), _int_matrix(rng, m, n)
    want = c.copy()
    for i in range(m):
        for j in range(n):
            s = 0
            for p in range(k):
                s += int(a[i, p]) * int(b[p, j])
            want[i, j] = c[i, j] + alpha * s
    got = c.copy()
    O.gemm(alpha, a, b, got)
    assert np.array_equal(got, want)
    # Strided views: transposed storage for A, negative strides for B and C.
    at = np.asfortranarray(a)
    bs = np.ascontiguousarray(b[::-1, ::-1])[::-1, ::-1]
    cs = np.ascontiguousarray(c[::-1, :])[::-1, :]
    O.gemm(alpha, at, bs, cs)
    assert np.array_equal(cs, want)


@pytest.mark.parametrize("m,n", [(1, 1), (6, 2), (33, 5), (70, 17)])
def test_trsm_brute_force_integers(m, n):
    rng = np.random.default_rng(m * 7 + n)
    # Unit lower L with small integers: forward substitution stays in the
    # integers, exact while |x| < 2^53.
    l = np.tril(_int_matrix(rng, m, m, -1, 2), -1) + np.triu(_int_matrix(rng, m, m, -9, 9), 0)
    b = _int_matrix(rng, m, n)
    want = [[int(v) for v in row] for row in b]
    for j in range(n):
        for i in range(m):
            for p in range(i):
                want[i][j] -= int(l[i, p]) * want[p][j]
    assert max(abs(v) for row in want for v in row) < 2 ** 53
    got = b.copy()
    O.trsm(l, got)
    assert np.array_equal(got, np.array(want, np.float64))
    # The upper triangle and diagonal of l are never read (unit lower).
    lf = np.asfortranarray(np.tril(l, -1) + 1000 * np.triu(np.ones((m, m))))
    got2 = np.asfortranarray(b)
    O.trsm(lf, got2)
    assert np.array_equal(got2, np.array(want, np.float64))


@pytest.mark.parametrize("n,nrhs", [(1, 1), (10, 3), (257, 4)])
def test_getrs_matches_dense_solve(n, nrhs):
    a = inputs.gaussian_matrix(2200 + n, n, n)
    b = inputs.gaussian_matrix(2201 + n, n, nrhs)
    lu = a.copy()
    piv, _ = O.getrf(lu)
    x = O.getrs(lu, piv, b)
    assert np.abs(x - np.linalg.solve(a, b)).max() <= 1e-9 * np.abs(x).max()
    # LAPACK's dgetrs on the oracle's own factors and pivots.
    xl = sl.lu_solve((lu, piv.astype(np.int32)), b, check_finite=False)
    assert np.abs(x - xl).max() <= 1e-12 * np.abs(x).max()


# ---------------------------------------------------------------- complex --

def complex_gaussian(seed, m, n, dtype=np.complex128):
    re = inputs.gaussian_matrix(seed, m, n)
    im = inputs.gaussian_matrix(seed + 7919, m, n)
    return (re + 1j * im).astype(dtype)


@pytest.mark.parametrize("shape", [(1, 1), (7, 7), (64, 64), (129, 129), (300, 200), (200, 300), (512, 512)])
def test_zgetrf_pivots_equal_lapack(shape):
    """LAPACK's zgetrf picks pivots with izamax, whose key is |re| + |im|
    (dcabs1), the reference's iamax key (src/blas/iamax.rs:14)."""
    m, n = shape
    a = complex_gaussian(2300 + m + n, m, n)
    w = a.copy()
    piv, sing = O.zgetrf(w)
    assert sing is None
    k = min(m, n)
    assert np.array_equal(piv, lapack_pivots(a)[:k])
    perm = O.lu_perm(m, piv)
    l = np.tril(w[:, :k], -1) + np.eye(m, k)
    u = np.triu(w[:k, :])
    assert np.abs(a[perm, :] - l @ u).max() <= 8 * max(m, n) * EPS * np.abs(a).max()


@pytest.mark.parametrize("shape", [(1, 1), (7, 7), (33, 33), (64, 40), (40, 64)])
def test_cgetrf_pivots_equal_lapack(shape):
    m, n = shape
    a = complex_gaussian(2400 + m + n, m, n, np.complex64)
    w = a.copy()
    piv, sing = O.cgetrf(w)
    assert sing is None
    k = min(m, n)
    assert np.array_equal(piv, lapack_pivots(a)[:k])
    perm = O.lu_perm(m, piv)
    l = np.tril(w[:, :k], -1) + np.eye(m, k)
    u = np.triu(w[:k, :])
    eps32 = float(np.finfo(np.float32).eps)
    assert np.abs(a[perm, :] - l @ u).max() <= 8 * max(m, n) * eps32 * np.abs(a).max()


def test_cgetrf_square_complex_known_answer():  # src/lapack/getrf.rs:485-502 in f32
    a = np.array([[1 + 1j, 2 - 1j], [3 + 1j, 4 - 1j]], np.complex64)
    piv, sing = O.cgetrf(a)
    assert list(piv) == [1, 1] and sing is None
    want = np.array([[3 + 1j, 4 - 1j], [0.4 + 0.2j, 0.2 - 1.4j]])
    assert np.abs(a - want).max() <= 1e-6


def _gauss_int(rng, m, n, lo=-3, hi=4):
    return (rng.integers(lo, hi, size=(m, n)) + 1j * rng.integers(lo, hi, size=(m, n))).astype(np.complex128)


def _exact_getrs(lu, piv, b):
    """Exact integer arithmetic on Gaussian-integer factors whose diagonal is a
    unit (1, -1, i, -i): x = U^-1 L^-1 P b with no rounding at all."""
    n = lu.shape[0]
    L = [[(int(v.real), int(v.imag)) for v in row] for row in lu]
    x = [(int(v.real), int(v.imag)) for v in b]
    for i, p in enumerate(piv):
        x[i], x[p] = x[p], x[i]

    def mul(a, c):
        return (a[0] * c[0] - a[1] * c[1], a[0] * c[1] + a[1] * c[0])

    for i in range(n):
        for k in range(i):
            pr = mul(L[i][k], x[k])
            x[i] = (x[i][0] - pr[0], x[i][1] - pr[1])
    for i in range(n - 1, -1, -1):
        for k in range(i + 1, n):
            pr = mul(L[i][k], x[k])
            x[i] = (x[i][0] - pr[0], x[i][1] - pr[1])
        d = L[i][i]
        assert d[0] * d[0] + d[1] * d[1] == 1
        x[i] = mul(x[i], (d[0], -d[1]))  # x / d = x * conj(d) for a unit d
    return np.array([complex(*v) for v in x])


@pytest.mark.parametrize("n,dtype", [(1, np.complex128), (2, np.complex128), (9, np.complex128),
                                     (24, np.complex128), (1, np.complex64), (5, np.complex64),
                                     (10, np.complex64)])
def test_complex_getrs_exact_on_gaussian_integers(n, dtype):
    """Unit-modulus diagonal and small Gaussian-integer off-diagonal factors:
    every product, difference and quotient is exact, so the oracle must match
    exact integer arithmetic bit for bit.  A conjugated product, a wrong
    division formula, a transposed factor or a skipped interchange fails."""
    rng = np.random.default_rng(n * 31 + (dtype == np.complex64))
    lu = np.tril(_gauss_int(rng, n, n, -1, 2), -1) + np.triu(_gauss_int(rng, n, n, -1, 2), 1)
    units = np.array([1, -1, 1j, -1j])
    lu[np.arange(n), np.arange(n)] = units[rng.integers(0, 4, n)]
    piv = np.array([rng.integers(i, n) for i in range(n)], np.int64)
    b = _gauss_int(rng, n, 1).ravel()
    want = _exact_getrs(lu, piv, b)
    limit = 2 ** 23 if dtype == np.complex64 else 2 ** 52
    assert np.abs(want.real).max() < limit and np.abs(want.imag).max() < limit
    fn = O.cgetrs if dtype == np.complex64 else O.zgetrs
    got = fn(lu.astype(dtype), piv, b.astype(dtype))
    assert np.array_equal(got.astype(np.complex128), want)


@pytest.mark.parametrize("n,nrhs", [(1, 1), (10, 3), (257, 4)])
def test_zgetrs_matches_dense_solve(n, nrhs):
    a = complex_gaussian(2500 + n, n, n)
    b = complex_gaussian(2501 + n, n, nrhs)
    lu = a.copy()
    piv, _ = O.zgetrf(lu)
    x = O.zgetrs(lu, piv, b)
    assert np.abs(x - np.linalg.solve(a, b)).max() <= 1e-9 * np.abs(x).max()
    xl = sl.lu_solve((lu, piv.astype(np.int32)), b, check_finite=False)
    assert np.abs(x - xl).max() <= 1e-12 * np.abs(x).max()
