"""ctypes front end of the CPU oracle (oracle/lair_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker, never as the thing measured
or shipped.  Each wrapper names the reference function it restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liblair_oracle.so")
_lib = None

i64 = ctypes.c_int64
dp = ctypes.POINTER(ctypes.c_double)
fp = ctypes.POINTER(ctypes.c_float)
ip = ctypes.POINTER(ctypes.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.lo_dgetrf_row_major.argtypes = [i64, i64, dp, i64, i64, ip, ip, ctypes.c_int]
        L.lo_sgetrf_row_major.argtypes = [i64, i64, fp, i64, i64, ip, ip, ctypes.c_int]
        L.lo_zgetrf_row_major.argtypes = [i64, i64, dp, i64, i64, ip, ip]
        L.lo_cgetrf_row_major.argtypes = [i64, i64, fp, i64, i64, ip, ip]
        L.lo_zgetrs.argtypes = [i64, i64, dp, i64, i64, ip, dp, i64]
        L.lo_cgetrs.argtypes = [i64, i64, fp, i64, i64, ip, fp, i64]
        L.lo_dgetrf_col_major.argtypes = [i64, i64, dp, i64, i64, ip, ip]
        L.lo_dgetrf.argtypes = [i64, i64, dp, i64, i64, ip, ip, ctypes.c_int]
        L.lo_dgetrf_recursive.argtypes = [i64, i64, dp, i64, i64, ip]
        L.lo_dgetrf_recursive.restype = i64
        L.lo_dgetrs.argtypes = [i64, i64, dp, i64, i64, ip, dp, i64]
        L.lo_sgetrs.argtypes = [i64, i64, fp, i64, i64, ip, fp, i64]
        L.lo_dgetrs_trans.argtypes = [i64, i64, dp, i64, i64, ip, dp, i64]
        L.lo_dlaswp.argtypes = [i64, dp, i64, i64, i64, i64, ip]
        L.lo_diamax.argtypes = [i64, dp, i64, ip, dp]
        L.lo_dtrsm_llu.argtypes = [dp, i64, i64, i64, i64, dp, i64, i64]
        L.lo_dgemm.argtypes = [ctypes.c_double, i64, i64, i64, dp, i64, i64, dp, i64, i64, dp, i64, i64]
        L.lo_dsolve.argtypes = [i64, i64, dp, i64, i64, i64, dp, dp]
        L.lo_dsolve.restype = ctypes.c_int
        L.lo_lu_perm.argtypes = [i64, i64, ip, ip]
        L.lo_is_standard_layout.argtypes = [i64, i64, i64, i64]
        L.lo_is_standard_layout.restype = ctypes.c_int
        L.lo_dresidual.argtypes = [i64, i64, dp, i64, dp, i64, dp, i64, dp, i64, ctypes.c_int]
        L.lo_num_threads_available.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t=dp):
    return a.ctypes.data_as(t)


def _strided(a):
    """Base pointer and element strides of a numpy view (strides may be negative)."""
    item = a.itemsize
    rs, cs = a.strides[0] // item, a.strides[1] // item
    return rs, cs


def getrf(a, nthreads=1):
    """src/lapack/getrf.rs:12-27 on a float64 view, in place; layout dispatch as
    the reference (standard layout -> row-major kernel, otherwise column-major).
    Returns (pivots int64[min(m,n)], singular or None)."""
    assert a.dtype == np.float64
    m, n = a.shape
    rs, cs = _strided(a)
    piv = np.zeros(min(m, n), np.int64)
    sing = np.zeros(1, np.int64)
    base = a.ctypes.data
    lib().lo_dgetrf(m, n, ctypes.cast(base, dp), rs, cs, _p(piv, ip), _p(sing, ip), nthreads)
    return piv, (None if sing[0] < 0 else int(sing[0]))


def sgetrf(a, nthreads=1):
    """getrf.rs:46-120 for f32, standard layout."""
    assert a.dtype == np.float32 and a.flags.c_contiguous
    m, n = a.shape
    piv = np.zeros(min(m, n), np.int64)
    sing = np.zeros(1, np.int64)
    lib().lo_sgetrf_row_major(m, n, _p(a, fp), n, 1, _p(piv, ip), _p(sing, ip), nthreads)
    return piv, (None if sing[0] < 0 else int(sing[0]))


def zgetrf(a):
    """getrf.rs:46-120 for Complex64 (interleaved), standard layout."""
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    m, n = a.shape
    piv = np.zeros(min(m, n), np.int64)
    sing = np.zeros(1, np.int64)
    lib().lo_zgetrf_row_major(m, n, ctypes.cast(a.ctypes.data, dp), n, 1, _p(piv, ip), _p(sing, ip))
    return piv, (None if sing[0] < 0 else int(sing[0]))


def cgetrf(a):
    """getrf.rs:46-120 for Complex32 (interleaved), standard layout."""
    assert a.dtype == np.complex64 and a.flags.c_contiguous
    m, n = a.shape
    piv = np.zeros(min(m, n), np.int64)
    sing = np.zeros(1, np.int64)
    lib().lo_cgetrf_row_major(m, n, ctypes.cast(a.ctypes.data, fp), n, 1, _p(piv, ip), _p(sing, ip))
    return piv, (None if sing[0] < 0 else int(sing[0]))


def zgetrs(lu, piv, b):
    """src/lapack/getrs.rs:12-39 for Complex64, each column of b (n or n x r)."""
    lu = np.ascontiguousarray(lu, np.complex128)
    x = np.array(b, np.complex128, copy=True, order="C")
    one = x.ndim == 1
    x2 = x.reshape(x.shape[0], -1)
    n = lu.shape[0]
    piv = np.ascontiguousarray(piv, np.int64)
    lib().lo_zgetrs(n, x2.shape[1], ctypes.cast(lu.ctypes.data, dp), lu.shape[1], 1, _p(piv, ip),
                    ctypes.cast(x2.ctypes.data, dp), x2.shape[1])
    return x2.reshape(-1) if one else x2


def cgetrs(lu, piv, b):
    """getrs.rs:12-39 for Complex32."""
    lu = np.ascontiguousarray(lu, np.complex64)
    x = np.array(b, np.complex64, copy=True, order="C")
    one = x.ndim == 1
    x2 = x.reshape(x.shape[0], -1)
    n = lu.shape[0]
    piv = np.ascontiguousarray(piv, np.int64)
    lib().lo_cgetrs(n, x2.shape[1], ctypes.cast(lu.ctypes.data, fp), lu.shape[1], 1, _p(piv, ip),
                    ctypes.cast(x2.ctypes.data, fp), x2.shape[1])
    return x2.reshape(-1) if one else x2


def getrf_recursive(a):
    """getrf.rs:30-40 (test-only Toledo recursion).  Returns (pivots, err) where
    err is None for Ok or the Singular row."""
    assert a.dtype == np.float64
    m, n = a.shape
    rs, cs = _strided(a)
    piv = np.zeros(m, np.int64)
    r = lib().lo_dgetrf_recursive(m, n, ctypes.cast(a.ctypes.data, dp), rs, cs, _p(piv, ip))
    return piv, (None if r == 0 else int(r - 1))


def getrs(lu, piv, b):
    """src/lapack/getrs.rs:12-39 for each column of b (n or n x r); returns x."""
    lu = np.ascontiguousarray(lu, np.float64)
    x = np.array(b, np.float64, copy=True, order="C")
    one = x.ndim == 1
    x2 = x.reshape(x.shape[0], -1)
    n = lu.shape[0]
    piv = np.ascontiguousarray(piv, np.int64)
    lib().lo_dgetrs(n, x2.shape[1], _p(lu), n, 1, _p(piv, ip), _p(x2), x2.shape[1])
    return x2.reshape(-1) if one else x2


def sgetrs(lu, piv, b):
    lu = np.ascontiguousarray(lu, np.float32)
    x = np.array(b, np.float32, copy=True, order="C")
    one = x.ndim == 1
    x2 = x.reshape(x.shape[0], -1)
    n = lu.shape[0]
    piv = np.ascontiguousarray(piv, np.int64)
    lib().lo_sgetrs(n, x2.shape[1], _p(lu, fp), n, 1, _p(piv, ip), _p(x2, fp), x2.shape[1])
    return x2.reshape(-1) if one else x2


def getrs_trans(lu, piv, b):
    """Solve A^T y = c from the factors of A (P A = L U)."""
    lu = np.ascontiguousarray(lu, np.float64)
    x = np.array(b, np.float64, copy=True, order="C")
    one = x.ndim == 1
    x2 = x.reshape(x.shape[0], -1)
    n = lu.shape[0]
    piv = np.ascontiguousarray(piv, np.int64)
    lib().lo_dgetrs_trans(n, x2.shape[1], _p(lu), n, 1, _p(piv, ip), _p(x2), x2.shape[1])
    return x2.reshape(-1) if one else x2


def laswp(a, begin, piv, ncols=None):
    """src/lapack/laswp.rs:11-40 on the rows of a float64 view, in place."""
    rs, cs = _strided(a)
    piv = np.ascontiguousarray(piv, np.int64)
    nc = a.shape[1] if ncols is None else ncols
    lib().lo_dlaswp(nc, ctypes.cast(a.ctypes.data, dp), rs, cs, begin, len(piv), _p(piv, ip))


def iamax(x):
    """src/blas/iamax.rs:6-21."""
    x = np.ascontiguousarray(x, np.float64)
    idx = np.zeros(1, np.int64)
    val = np.zeros(1, np.float64)
    lib().lo_diamax(len(x), _p(x), 1, _p(idx, ip), _p(val))
    return int(idx[0]), float(val[0])


def trsm(l, b):
    """src/blas/trsm.rs:6-22: b := L^-1 b (unit lower), in place."""
    rs, cs = _strided(l)
    brs, bcs = _strided(b)
    lib().lo_dtrsm_llu(ctypes.cast(l.ctypes.data, dp), rs, cs, b.shape[0], b.shape[1],
                       ctypes.cast(b.ctypes.data, dp), brs, bcs)


def gemm(alpha, a, b, c):
    """src/blas/gemm.rs:6-32: c += alpha * a @ b, in place."""
    ars, acs = _strided(a)
    brs, bcs = _strided(b)
    crs, ccs = _strided(c)
    lib().lo_dgemm(alpha, a.shape[0], b.shape[1], a.shape[1], ctypes.cast(a.ctypes.data, dp), ars, acs,
                   ctypes.cast(b.ctypes.data, dp), brs, bcs, ctypes.cast(c.ctypes.data, dp), crs, ccs)


def solve(a, b):
    """src/equation.rs:32-60.  Returns (status, x): 0 ok, 1 Shape, 2 Value."""
    a = np.asarray(a, np.float64)
    m, n = a.shape
    rs, cs = _strided(a)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(max(n, 1), np.float64)
    st = lib().lo_dsolve(m, n, ctypes.cast(a.ctypes.data, dp), rs, cs, len(b), _p(b), _p(x))
    return st, x[:n]


def lu_perm(nrows, piv):
    """src/decomposition/lu.rs:28-39: P[perm[i], i] = 1."""
    piv = np.ascontiguousarray(piv, np.int64)
    perm = np.zeros(nrows, np.int64)
    lib().lo_lu_perm(nrows, len(piv), _p(piv, ip), _p(perm, ip))
    return perm


def is_standard_layout(m, n, rs, cs):
    return bool(lib().lo_is_standard_layout(m, n, rs, cs))


def residual(a, x, b, nthreads=1):
    """r = b - A x with long-double accumulation (checker helper)."""
    a = np.ascontiguousarray(a, np.float64)
    x2 = np.ascontiguousarray(np.asarray(x, np.float64).reshape(a.shape[0], -1))
    b2 = np.ascontiguousarray(np.asarray(b, np.float64).reshape(a.shape[0], -1))
    r = np.zeros_like(b2)
    n, k = b2.shape
    lib().lo_dresidual(n, k, _p(a), n, _p(x2), k, _p(b2), k, _p(r), k, nthreads)
    return r
