"""lair_amd -- MI355X (gfx950) LU factor-and-solve behind vinesystems/lair's interface.

Python mirror of the reference's LU path, calling the C ABI of
lib/liblair_dev.so (include/lair_dev.h):

* ``getrf(a)``           lapack::getrf      src/lapack/getrf.rs:12-27
* ``getrs(a, p, b)``     lapack::getrs      src/lapack/getrs.rs:12-39
* ``laswp(...)``         lapack::laswp      src/lapack/laswp.rs:11-40
* ``equation.solve``     equation::solve    src/equation.rs:32-60
* ``lu.Factorized``      decomposition::lu  src/decomposition/lu.rs:12-171
* ``gesv_mixed``         FP32 / BF16 LU with FP64 iterative refinement (new)

Arrays are numpy views; their strides are passed through unchanged (negative
strides and column-major views are accepted like ndarray views).  Every call
runs on the GPU; there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import LairError, check

__all__ = ["getrf", "getrs", "laswp", "solve_backward", "gesv_mixed", "InvalidInput", "ShapeError", "ValueError_",
           "equation", "lu", "Context", "DeviceArray", "BatchedPlan", "LairError"]


class InvalidInput(Exception):
    """src/lib.rs:23-30."""


class ShapeError(InvalidInput):
    """InvalidInput::Shape."""


class ValueError_(InvalidInput):
    """InvalidInput::Value."""


def _elem_strides(a):
    return a.strides[0] // a.itemsize, a.strides[1] // a.itemsize


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def getrf(a):
    """In-place LU with partial pivoting of a 2-D float64, float32, complex128
    or complex64 view (complex pivots: first maximum of |re| + |im|).

    Returns ``(pivots, singular)``: int64 pivots of length min(m, n) (row i was
    interchanged with row pivots[i]) and the last zero-pivot column or None,
    exactly as lapack::getrf (src/lapack/getrf.rs:12-27)."""
    lib = _capi.load()
    if a.ndim != 2:
        raise ValueError("getrf expects a 2-D array")
    m, n = a.shape
    rs, cs = _elem_strides(a)
    piv = np.zeros(min(m, n), np.int64)
    sing = ctypes.c_int64(-1)
    fns = {np.dtype(np.float64): lib.lair_dgetrf, np.dtype(np.float32): lib.lair_sgetrf,
           np.dtype(np.complex128): lib.lair_zgetrf, np.dtype(np.complex64): lib.lair_cgetrf}
    fn = fns.get(a.dtype)
    if fn is None:
        raise TypeError("getrf supports float64, float32, complex128 and complex64")
    check(fn(m, n, _ptr(a), rs, cs, piv.ctypes.data_as(_capi.P64), ctypes.byref(sing)), "getrf")
    return piv, (None if sing.value < 0 else int(sing.value))


def getrs(a, p, b, trans=False):
    """Solves a x = b from the packed factors and pivots (src/lapack/getrs.rs:12-39).

    ``b`` may be 1-D (the reference's only case) or 2-D with one right-hand
    side per column; a new array is returned.  Shape mismatches raise
    ValueError where the reference panics (getrs.rs:18-20)."""
    lib = _capi.load()
    p = np.ascontiguousarray(p, np.int64)
    if a.shape[0] != len(p) or b.shape[0] != len(p) or a.shape[1] < len(p):
        raise ValueError("getrs: shape mismatch between a, p and b")
    n = len(p)
    x = np.array(b, dtype=a.dtype, copy=True)
    x2 = x.reshape(n, -1)
    lrs, lcs = _elem_strides(a)
    brs, bcs = _elem_strides(x2)
    if a.dtype in (np.complex128, np.complex64):
        if trans:
            raise ValueError("getrs: no transposed complex solve (the reference has none)")
        fn = lib.lair_zgetrs if a.dtype == np.complex128 else lib.lair_cgetrs
        check(fn(n, x2.shape[1], _ptr(a), lrs, lcs, p.ctypes.data_as(_capi.P64), _ptr(x2), brs, bcs), "getrs")
        return x
    if a.dtype == np.float64:
        fn = lib.lair_dgetrs
    elif a.dtype == np.float32:
        fn = lib.lair_sgetrs
    else:
        raise TypeError("getrs supports float64, float32, complex128 and complex64")
    check(fn(n, x2.shape[1], _ptr(a), lrs, lcs, p.ctypes.data_as(_capi.P64), _ptr(x2), brs, bcs,
             1 if trans else 0), "getrs")
    return x


def laswp(ncols, a, begin, piv):
    """Row interchanges rows i <-> piv[i] for i in begin..len(piv), forward
    order, over ncols columns of the float64 view a (src/lapack/laswp.rs:11-40)."""
    lib = _capi.load()
    piv = np.ascontiguousarray(piv, np.int64)
    rs, cs = _elem_strides(a)
    check(lib.lair_dlaswp(ncols, _ptr(a), rs, cs, begin, len(piv), piv.ctypes.data_as(_capi.P64)),
          "laswp")


def solve_backward(lu_factors, pivots, x, xbar):
    """Adjoint of X = A^-1 B from A's factors (getrf output), the solution X
    and the upstream gradient XBAR: returns (BBAR, ABAR) with
    BBAR = A^-T XBAR (transposed getrs reusing the factors and pivots) and
    ABAR = -BBAR X^T (FP64 MFMA GEMM).  X and XBAR are n or n x nrhs."""
    lib = _capi.load()
    lu_c = np.ascontiguousarray(lu_factors, np.float64)
    n = lu_c.shape[0]
    piv = np.ascontiguousarray(pivots, np.int64)
    x2 = np.ascontiguousarray(np.asarray(x, np.float64).reshape(n, -1))
    xb2 = np.ascontiguousarray(np.asarray(xbar, np.float64).reshape(n, -1))
    if x2.shape != xb2.shape or len(piv) != n or lu_c.shape != (n, n):
        raise ValueError("solve_backward: shape mismatch")
    bbar = np.zeros_like(x2)
    abar = np.zeros((n, n), np.float64)
    check(lib.lair_dsolve_backward(n, x2.shape[1], _ptr(lu_c), piv.ctypes.data_as(_capi.P64), _ptr(x2),
                                   _ptr(xb2), _ptr(bbar), _ptr(abar)), "solve_backward")
    return bbar.reshape(np.shape(xbar)), abar


_PRECISIONS = {"fp64": _capi.PREC_FP64, "fp32": _capi.PREC_FP32, "bf16x6": _capi.PREC_BF16X6,
               "bf16": _capi.PREC_BF16}


def gesv_mixed(a, b, precision="fp32", max_iter=30, tol=0.0, tournament=False):
    """x = a^-1 b with the LU in a lower precision ("fp32", "bf16x6" = six
    BF16 products per FP32 product, "bf16", or "fp64") and FP64 iterative
    refinement until every column's backward error
    ||b - a x||inf / (||a||inf ||x||inf) is at most tol (0: n * eps64).  A
    solve that does not converge within max_iter is re-done with the FP64 LU.

    b is n or n x nrhs.  Returns (x, info) with info a dict of iterations,
    converged, fell_back, precision and backward_error.  New entry point: the
    reference solves only in the matrix's own precision (src/equation.rs:32-60)."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.ndim != 2 or a.shape[0] != a.shape[1] or b.shape[0] != a.shape[0] or b.ndim not in (1, 2):
        raise ShapeError("gesv_mixed: a must be n x n and b n or n x nrhs")
    if precision not in _PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
    n = a.shape[0]
    b2 = np.ascontiguousarray(b.reshape(n, -1))
    nrhs = b2.shape[1]
    ctx = _default_context()
    da, db, dx = DeviceArray.from_host(a), DeviceArray.from_host(b2), DeviceArray(n, nrhs)
    info = _capi.IrInfo()
    prec = _PRECISIONS[precision] | (_capi.PIVOT_TOURNAMENT if tournament else 0)
    try:
        st = ctx.lib.lair_dgesv_mixed_device(ctx.h, n, nrhs, da.ptr, n, db.ptr, nrhs, dx.ptr, nrhs, prec,
                                             max_iter, tol, ctypes.byref(info))
        if st == _capi.LAIR_ERR_SINGULAR:
            raise ValueError_("`a` is a singular matrix")
        if st != _capi.LAIR_ERR_NOCONV:
            check(st, "gesv_mixed")
        x = dx.download()
    finally:
        da.free()
        db.free()
        dx.free()
    out = {"iterations": info.iterations, "converged": bool(info.converged), "fell_back": bool(info.fell_back),
           "precision": info.precision, "backward_error": info.backward_error}
    return x.reshape(b.shape), out


class BatchedPlan:
    """Plan for solving many small systems A_i X_i = B_i in one persistent,
    graph-captured launch (packed device layout of include/lair_dev.h).

    ``a``, ``b``, ``ipiv`` and ``info`` are DeviceArray buffers holding the
    packed systems; ``run()`` overwrites a with the LU factors and b with X."""

    def __init__(self, ctx, ns, nrhs, a, b, ipiv, info, precision=0, autotune=True):
        self.ctx = ctx
        lib = ctx.lib
        ns = np.ascontiguousarray(ns, np.int32)
        nrhs = np.ascontiguousarray(nrhs, np.int32)
        h = ctypes.c_void_p()
        P32 = ctypes.POINTER(ctypes.c_int32)
        check(lib.lair_gesv_batched_plan(ctx.h, precision, len(ns), ns.ctypes.data_as(P32),
                                         nrhs.ctypes.data_as(P32), a.ptr, b.ptr, ipiv.ptr, info.ptr,
                                         1 if autotune else 0, ctypes.byref(h)), "batched_plan")
        self.h = h

    def run(self):
        check(self.ctx.lib.lair_gesv_batched_run(self.h), "batched_run")

    def config(self):
        nb, grid, lds = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(self.ctx.lib.lair_gesv_batched_config(self.h, ctypes.byref(nb), ctypes.byref(grid),
                                                    ctypes.byref(lds)), "batched_config")
        return {"nb": nb.value, "grid": grid.value, "lds_elems": lds.value}

    def close(self):
        if self.h:
            self.ctx.lib.lair_gesv_batched_destroy(self.h)
            self.h = None


class equation:
    """src/equation.rs"""

    @staticmethod
    def solve(a, b):
        """x = a^-1 b; raises ShapeError / ValueError_ like InvalidInput::{Shape, Value}.
        float64 unless a or b is complex (then complex128, or complex64 when
        both are complex64)."""
        lib = _capi.load()
        a = np.asarray(a)
        b = np.asarray(b)
        if np.iscomplexobj(a) or np.iscomplexobj(b):
            dt = np.complex64 if (a.dtype == np.complex64 and b.dtype == np.complex64) else np.complex128
            fn = lib.lair_csolve if dt == np.complex64 else lib.lair_zsolve
        else:
            dt, fn = np.float64, lib.lair_dsolve
        a = np.asarray(a, dt)
        b = np.ascontiguousarray(b, dt)
        if a.ndim != 2 or b.ndim != 1:
            raise ShapeError("a must be 2-D and b 1-D")
        m, n = a.shape
        rs, cs = _elem_strides(a)
        x = np.zeros(max(n, 1), dt)
        st = fn(m, n, _ptr(a), rs, cs, len(b), _ptr(b), _ptr(x))
        if st == _capi.LAIR_ERR_SHAPE:
            raise ShapeError(_capi.last_error())
        if st == _capi.LAIR_ERR_SINGULAR:
            raise ValueError_("`a` is a singular matrix")
        check(st, "solve")
        return x[:n]


class lu:
    """src/decomposition/lu.rs"""

    class Factorized:
        """LU decomposition factors (lu.rs:12-20).  Construction factors a copy
        of the matrix on the GPU (lu.rs:156-171).  For float64 the factors stay
        on the device in a lair_lu handle (pivots, row permutation and, after
        the first solve, the inverted diagonal blocks): p, l, u, into_pl and
        solve run there.  float32 factors through lair_sgetrf and extracts on
        the host."""

        def __init__(self, a):
            a = np.asarray(a)
            self._h = None
            if a.dtype == np.float64 and a.ndim == 2 and a.size > 0:
                ctx = _default_context()
                lib = ctx.lib
                m, n = a.shape
                da = DeviceArray.from_host(np.ascontiguousarray(a))
                h = ctypes.c_void_p()
                check(lib.lair_dlu_factor(ctx.h, m, n, da.ptr, n, 0, ctypes.byref(h)), "lu_factor")
                da.free()
                self._h, self._ctx = h, ctx
                sing = ctypes.c_int64()
                check(lib.lair_dlu_info(h, None, None, ctypes.byref(sing)), "lu_info")
                self.singular = None if sing.value < 0 else int(sing.value)
                lu_p, ld, piv_p = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_void_p()
                check(lib.lair_dlu_factors(h, ctypes.byref(lu_p), ctypes.byref(ld), ctypes.byref(piv_p), None),
                      "lu_factors")
                host = np.empty((m, ld.value), np.float64)
                check(lib.lair_memcpy_d2h(ctypes.c_void_p(host.ctypes.data), lu_p, host.nbytes), "d2h")
                self.lu = np.ascontiguousarray(host[:, :n])
                self.pivots = np.empty(min(m, n), np.int64)
                check(lib.lair_memcpy_d2h(ctypes.c_void_p(self.pivots.ctypes.data), piv_p, self.pivots.nbytes),
                      "d2h")
            else:
                self.lu = np.array(a, copy=True)
                self.pivots, self.singular = getrf(self.lu)

        def __del__(self):
            try:
                if self._h:
                    self._ctx.lib.lair_dlu_destroy(self._h)
                    self._h = None
            except Exception:
                pass

        def _extract(self, which, rows, cols):
            out = DeviceArray(rows, cols)
            check(self._ctx.lib.lair_dlu_extract(self._ctx.h, self._h, which, out.ptr, cols), "lu_extract")
            h = out.download()
            out.free()
            return h

        def p(self):
            """lu.rs:28-39: P with P[perm[i], i] = 1, perm = laswp of 0..nrows."""
            nr = self.lu.shape[0]
            if self._h:
                return self._extract(_capi.LU_P, nr, nr)
            perm = np.arange(nr, dtype=np.int64)
            for i, piv in enumerate(self.pivots):
                if piv != i:
                    perm[i], perm[piv] = perm[piv], perm[i]
            p = np.zeros((nr, nr), self.lu.dtype)
            p[perm, np.arange(nr)] = 1
            return p

        def l(self):
            """lu.rs:42-57: unit lower nrows x rank."""
            nr, nc = self.lu.shape
            rank = min(nr, nc)
            if self._h:
                return self._extract(_capi.LU_L, nr, rank)
            l = np.tril(self.lu[:, :rank], -1)
            l[np.arange(rank), np.arange(rank)] = 1
            return l

        def u(self):
            """lu.rs:60-72: upper rank x ncols."""
            nr, nc = self.lu.shape
            rank = min(nr, nc)
            if self._h:
                return self._extract(_capi.LU_U, rank, nc)
            return np.triu(self.lu[:rank, :])

        def is_singular(self):
            """lu.rs:75-77."""
            return self.singular is not None

        def solve(self, b):
            """lu.rs:87-98 (b may also be 2-D: one right-hand side per column)."""
            b = np.asarray(b)
            if b.shape[0] != self.lu.shape[0]:
                raise ShapeError(f"b must have {self.lu.shape[0]} elements")
            if not self._h:
                return getrs(self.lu, self.pivots, b)
            n = self.lu.shape[0]
            x = np.ascontiguousarray(np.asarray(b, np.float64).reshape(n, -1))
            db = DeviceArray.from_host(x)
            st = self._ctx.lib.lair_dlu_solve(self._ctx.h, self._h, x.shape[1], db.ptr, x.shape[1], 0)
            if st == _capi.LAIR_ERR_SHAPE:
                raise ShapeError(_capi.last_error())
            check(st, "lu_solve")
            out = db.download()
            db.free()
            return out.reshape(b.shape)

        def into_pl(self):
            """lu.rs:107-153: P*L in the first min(m,n) columns, U's columns
            beyond them unchanged."""
            nr, nc = self.lu.shape
            k = min(nr, nc)
            if self._h:
                out = self.lu.copy()
                out[:, :k] = self._extract(_capi.LU_PL, nr, k)
                self.lu = out
                return out
            pivots = list(self.pivots) + list(range(len(self.pivots), nr))
            for i in range(len(pivots) - 1, -1, -1):
                target = pivots[i]
                if i == target:
                    continue
                pivots[i] = pivots[target]
                pivots[target] = i
            pl = self.lu[:, :k]
            total = len(pivots)
            dst = 0
            i = dst
            while True:
                src = pivots[dst]
                for c in range(min(src, k)):
                    pl[dst, c] = pl[src, c]
                if src < k:
                    pl[dst, src] = 1
                for c in range(src + 1, k):
                    pl[dst, c] = 0
                pivots[dst] = total
                if pivots[src] == total:
                    dst = i + 1
                    while dst < total and pivots[dst] == total:
                        dst += 1
                    if dst == total:
                        break
                    i = dst
                else:
                    dst = src
            return self.lu


_default = None


def _default_context():
    """A context shared by the host-array mirror (lu.Factorized)."""
    global _default
    if _default is None:
        _default = Context()
    return _default


class Context:
    """A device context (stream + workspace) of liblair_dev.so."""

    def __init__(self):
        self.lib = _capi.load()
        h = ctypes.c_void_p()
        check(self.lib.lair_context_create(ctypes.byref(h)), "context_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.lair_context_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        check(self.lib.lair_context_synchronize(self.h), "synchronize")


class DeviceArray:
    """A row-major device buffer (rows x cols, leading dimension ld elements)."""

    def __init__(self, rows, cols, dtype=np.float64, ld=None):
        self.lib = _capi.load()
        self.rows, self.cols = rows, cols
        self.dtype = np.dtype(dtype)
        self.ld = ld if ld is not None else cols
        self.nbytes = max(rows * self.ld * self.dtype.itemsize, 1)
        p = ctypes.c_void_p()
        check(self.lib.lair_malloc(ctypes.byref(p), self.nbytes), "malloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a, ld=None):
        a = np.asarray(a)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        d = cls(a.shape[0], a.shape[1], a.dtype, ld)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.asarray(a, self.dtype).reshape(self.rows, self.cols)
        if self.ld == self.cols:
            h = np.ascontiguousarray(a)
        else:
            h = np.zeros((self.rows, self.ld), self.dtype)
            h[:, :self.cols] = a
        check(self.lib.lair_memcpy_h2d(self.ptr, ctypes.c_void_p(h.ctypes.data), h.nbytes), "h2d")

    def download(self):
        h = np.empty((self.rows, self.ld), self.dtype)
        check(self.lib.lair_memcpy_d2h(ctypes.c_void_p(h.ctypes.data), self.ptr, h.nbytes), "d2h")
        return np.ascontiguousarray(h[:, :self.cols])

    def free(self):
        if self.ptr:
            self.lib.lair_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
