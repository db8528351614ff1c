"""Pins the CPU oracle to the reference's own known-answer tests.

Every vector below is the expectation written in the reference's unit tests
(file:line cited per test); tolerances are the reference's (approx crate:
abs 1e-6, default f64::EPSILON, rel 1e-8 / 1e-6).
"""
import numpy as np
import pytest

from oracle import oracle as O

LU_5X5 = np.array([
    [3., 1., 2., 2., 1.],
    [0.33333333, 2.66666667, 0.33333333, 2.33333333, 0.66666667],
    [0.33333333, 0.625, 2.125, -1.125, 1.25],
    [0.66666667, 0.875, 0.64705882, -1.64705882, -1.05882353],
    [0.66666667, 0.5, -0.23529412, -0.14285714, 2.14285714],
])
A_5X5 = np.array([
    [1., 2., 3., 1., 2.],
    [2., 2., 1., 3., 3.],
    [3., 1., 2., 2., 1.],
    [2., 3., 3., 1., 1.],
    [1., 3., 1., 3., 1.],
])


def test_singular():  # src/lapack/getrf.rs:342-347
    a = np.array([[1., 1.], [1., 1.]])
    _, sing = O.getrf(a)
    assert sing == 1


def test_empty():  # getrf.rs:349-355
    a = np.zeros((0, 0))
    piv, sing = O.getrf(a)
    assert len(piv) == 0 and sing is None


def test_recursive_empty():  # getrf.rs:357-362
    piv, err = O.getrf_recursive(np.zeros((0, 0)))
    assert len(piv) == 0 and err is None


def test_smallest():  # getrf.rs:364-371
    a = np.array([[3.]])
    piv, sing = O.getrf(a)
    assert list(piv) == [0] and sing is None and a[0, 0] == 3.0


def test_recursive_smallest():  # getrf.rs:373-379
    a = np.array([[3.]])
    piv, err = O.getrf_recursive(a)
    assert list(piv) == [0] and err is None and a[0, 0] == 3.0


def test_recursive_square():  # getrf.rs:381-403
    a = A_5X5.copy()
    piv, err = O.getrf_recursive(a)
    assert err is None
    assert list(piv) == [2, 4, 2, 3, 4]
    np.testing.assert_allclose(a, LU_5X5, atol=1e-6, rtol=0)


def test_square_row_major():  # getrf.rs:405-428
    a = A_5X5.copy()
    assert O.is_standard_layout(5, 5, 5, 1)
    piv, sing = O.getrf(a)
    assert list(piv) == [2, 4, 2, 3, 4] and sing is None
    np.testing.assert_allclose(a, LU_5X5, atol=1e-6, rtol=0)


def test_square_col_major():  # getrf.rs:430-455
    base = np.array([
        [1., 2., 3., 2., 1.],
        [2., 2., 1., 3., 3.],
        [3., 1., 2., 3., 1.],
        [1., 3., 2., 1., 3.],
        [2., 3., 1., 1., 1.],
    ])
    a = base.T  # swap_axes(0, 1): a view with strides (1, 5)
    assert not O.is_standard_layout(5, 5, *[s // 8 for s in a.strides])
    piv, sing = O.getrf(a)
    assert list(piv) == [2, 4, 2, 3, 4] and sing is None
    np.testing.assert_allclose(a, LU_5X5, atol=1e-6, rtol=0)


def test_square_negative_strides():  # getrf.rs:457-483
    base = np.array([
        [1., 3., 1., 3., 1.],
        [1., 1., 3., 3., 2.],
        [1., 2., 2., 1., 3.],
        [3., 3., 1., 2., 2.],
        [2., 1., 3., 2., 1.],
    ])
    a = base[::-1, ::-1]
    rs, cs = a.strides[0] // 8, a.strides[1] // 8
    assert rs < 0 and cs < 0 and not O.is_standard_layout(5, 5, rs, cs)
    piv, sing = O.getrf(a)
    assert list(piv) == [2, 4, 2, 3, 4] and sing is None
    np.testing.assert_allclose(a, LU_5X5, atol=1e-6, rtol=0)


def test_square_complex():  # getrf.rs:485-502
    a = np.array([[1 + 1j, 2 - 1j], [3 + 1j, 4 - 1j]])
    piv, sing = O.zgetrf(a)
    assert list(piv) == [1, 1] and sing is None
    exp = np.array([[3 + 1j, 4 - 1j], [0.4 + 0.2j, 0.2 - 1.4j]])
    np.testing.assert_allclose(a.real, exp.real, atol=1e-6, rtol=0)
    np.testing.assert_allclose(a.imag, exp.imag, atol=1e-6, rtol=0)


def test_wide():  # getrf.rs:504-511 (exact)
    a = np.array([[1., 2., 3.], [2., 3., 4.]])
    piv, sing = O.getrf(a)
    assert list(piv) == [1, 1] and sing is None
    assert (a == np.array([[2., 3., 4.], [0.5, 0.5, 1.]])).all()


def test_tall():  # getrf.rs:513-524
    a = np.array([[1., 2.], [1., 3.], [2., 3.]])
    piv, sing = O.getrf(a)
    assert list(piv) == [2, 1] and sing is None
    np.testing.assert_allclose(a, [[2., 3.], [0.5, 1.5], [0.5, 0.3333333333333333]], atol=1e-6, rtol=0)


def test_getrs_square_2x2():  # src/lapack/getrs.rs:46-57
    a = np.array([[1., 2.], [3., 4.]])
    piv, _ = O.getrf(a)
    assert list(piv) == [1, 1]
    x = O.getrs(a, piv, np.array([3., 7.]))
    np.testing.assert_allclose(x, [1., 1.], rtol=1e-8)


def test_getrs_square_5x5():  # getrs.rs:59-78 (abs <= f64::EPSILON)
    a = A_5X5.copy()
    piv, _ = O.getrf(a)
    x = O.getrs(a, piv, np.array([9., 11., 9., 10., 9.]))
    assert np.all(np.abs(x - 1.0) <= np.finfo(np.float64).eps)


def test_equation_doctest():  # src/equation.rs:21-31 (exact)
    st, x = O.solve(np.array([[3., 1.], [1., 2.]]), np.array([9., 8.]))
    assert st == 0 and list(x) == [2.0, 3.0]


def test_equation_errors():  # equation.rs:42-57
    assert O.solve(np.zeros((2, 3)), np.zeros(2))[0] == 1
    assert O.solve(np.eye(3), np.zeros(2))[0] == 1
    assert O.solve(np.array([[1., 1.], [1., 1.]]), np.ones(2))[0] == 2


def test_iamax():  # src/blas/iamax.rs:26-34
    assert O.iamax(np.array([1., 3., 2.])) == (1, 3.0)


def _lu_accessors_f32(m):
    """lu.rs:28-72 on top of the f32 oracle: P (via laswp of an index vector),
    L (unit lower, nrows x rank), U (rank x ncols)."""
    a = np.array(m, np.float32)
    piv, sing = O.sgetrf(a)
    nr, nc = a.shape
    rank = min(nr, nc)
    perm = O.lu_perm(nr, piv)
    p = np.zeros((nr, nr), np.float32)
    p[perm, np.arange(nr)] = 1.0
    l = np.zeros((nr, rank), np.float32)
    for i in range(nr):
        for j in range(min(i, rank)):
            l[i, j] = a[i, j]
        if i < rank:
            l[i, i] = 1.0
    u = np.zeros((rank, nc), np.float32)
    for i in range(rank):
        u[i, i:] = a[i, i:]
    return p, l, u


def test_lu_square_f32():  # src/decomposition/lu.rs:181-200
    p, l, u = _lu_accessors_f32([[1, 2, 3], [2, 2, 1], [3, 1, 2]])
    assert p[0, 1] == 1 and p[1, 2] == 1 and p[2, 0] == 1
    np.testing.assert_allclose([l[0, 0], l[1, 0], l[2, 0]], [1., 0.33333333, 0.6666667], rtol=1e-6)
    np.testing.assert_allclose([u[0, 2], u[1, 2], u[2, 2]], [2., 2.3333333, -2.2], rtol=1e-6)


def test_lu_wide_f32():  # lu.rs:202-225
    p, l, u = _lu_accessors_f32([[1, 2, 3, 1], [2, 2, 1, 3], [3, 1, 2, 2]])
    assert p.shape == (3, 3) and p[0, 1] == 1 and p[1, 2] == 1 and p[2, 0] == 1
    assert l.shape == (3, 3) and u.shape == (3, 4)
    np.testing.assert_allclose([l[0, 0], l[1, 0], l[2, 0]], [1., 0.33333333, 0.6666667], rtol=1e-6)
    np.testing.assert_allclose([u[0, 2], u[1, 2], u[2, 2]], [2., 2.3333333, -2.2], rtol=1e-6)


def test_lu_tall_f32():  # lu.rs:227-254
    p, l, u = _lu_accessors_f32([[1, 2, 3], [2, 2, 1], [3, 1, 2], [2, 3, 3]])
    assert p.shape == (4, 4)
    assert p[0, 3] == 1 and p[1, 2] == 1 and p[2, 0] == 1 and p[3, 1] == 1
    assert l.shape == (4, 3) and u.shape == (3, 3)
    np.testing.assert_allclose([l[0, 0], l[1, 0], l[2, 0], l[3, 0]], [1., 0.6666667, 0.6666667, 0.33333333], rtol=1e-6)
    np.testing.assert_allclose([u[0, 2], u[1, 2], u[2, 2]], [2., 1.6666666, -1.2857143], rtol=1e-6)
