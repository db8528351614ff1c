"""The trailing-update GEMM kernels called directly: C -= A B, the reference's
gemm (src/blas/gemm.rs:6-32) with alpha = -1, as getrf calls it
(src/lapack/getrf.rs:289-296).

The tiles sum in their own order, so the bar is an elementwise error bound
against an FP64 product of the same operands:
  |C_gpu - C_ref| <= c (k + 2) u (|A| |B| + |C|)
with u = eps64 (FP64 tiles), eps32 (FP32 MFMA, and the six-product BF16 split,
whose three 8-bit terms carry 27 significand bits per operand), and for the
single BF16 product an extra 2^-7 |A| |B| (one 8-bit term per operand).
Shapes cover whole interior tiles, the masked right and bottom strips, K not a
multiple of the stage depth, the small-shape kernels below 256 rows, and
leading dimensions wider than the matrix.
"""
import numpy as np
import pytest

import lair_amd as L
from lair_amd import _capi

pytestmark = pytest.mark.gpu

SHAPES = [(256, 128, 32), (384, 256, 128), (300, 200, 128), (1000, 700, 96), (257, 33, 40),
          (640, 1000, 200), (100, 50, 30), (2048, 2048, 256)]


@pytest.fixture(scope="module")
def ctx():
    c = L.Context()
    yield c
    c.close()


def operands(seed, m, n, k, dtype):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((m, k)).astype(dtype)
    b = rng.standard_normal((k, n)).astype(dtype)
    c = rng.standard_normal((m, n)).astype(dtype)
    return a, b, c


def run(ctx, prec, a, b, c, pad=0):
    m, k = a.shape
    n = b.shape[1]
    da = L.DeviceArray.from_host(a, ld=k + pad)
    db = L.DeviceArray.from_host(b, ld=n + pad)
    dc = L.DeviceArray.from_host(c, ld=n + pad)
    if a.dtype == np.float64:
        st = ctx.lib.lair_dgemm_minus_device(ctx.h, m, n, k, da.ptr, k + pad, db.ptr, n + pad, dc.ptr, n + pad)
    else:
        st = ctx.lib.lair_sgemm_minus_device(ctx.h, prec, m, n, k, da.ptr, k + pad, db.ptr, n + pad, dc.ptr,
                                             n + pad)
    assert st == _capi.LAIR_OK, _capi.last_error()
    return dc.download()


def scaled_error(a, b, c, got):
    a64, b64, c64 = (x.astype(np.float64) for x in (a, b, c))
    ref = c64 - a64 @ b64
    mag = np.abs(a64) @ np.abs(b64)
    return np.abs(got.astype(np.float64) - ref), mag, np.abs(c64)


@pytest.mark.parametrize("m,n,k", SHAPES)
def test_dgemm_minus(ctx, m, n, k):
    a, b, c = operands(100 + m + n + k, m, n, k, np.float64)
    err, mag, cm = scaled_error(a, b, c, run(ctx, 0, a, b, c))
    assert (err <= 2 * (k + 2) * np.finfo(np.float64).eps * (mag + cm)).all()


@pytest.mark.parametrize("prec", [_capi.PREC_FP32, _capi.PREC_BF16X6])
@pytest.mark.parametrize("m,n,k", SHAPES)
def test_sgemm_minus(ctx, prec, m, n, k):
    a, b, c = operands(200 + m + n + k, m, n, k, np.float32)
    err, mag, cm = scaled_error(a, b, c, run(ctx, prec, a, b, c))
    u = float(np.finfo(np.float32).eps)
    assert (err <= 2 * (k + 2) * u * (mag + cm)).all()


@pytest.mark.parametrize("m,n,k", [(384, 256, 128), (300, 200, 40)])
def test_single_bf16_product_is_coarser(ctx, m, n, k):
    a, b, c = operands(300 + m, m, n, k, np.float32)
    err1, mag, cm = scaled_error(a, b, c, run(ctx, _capi.PREC_BF16, a, b, c))
    err6, _, _ = scaled_error(a, b, c, run(ctx, _capi.PREC_BF16X6, a, b, c))
    u = float(np.finfo(np.float32).eps)
    assert (err1 <= 2.0 ** -7 * mag + 2 * (k + 2) * u * (mag + cm)).all()
    # The precision argument is honoured: one 8-bit term per operand is far
    # coarser than three.
    assert err1.max() > 100 * err6.max()


@pytest.mark.parametrize("prec", [_capi.PREC_FP32, _capi.PREC_BF16X6])
def test_sgemm_minus_padded_leading_dimensions(ctx, prec):
    m, n, k = 520, 300, 128
    a, b, c = operands(400, m, n, k, np.float32)
    err, mag, cm = scaled_error(a, b, c, run(ctx, prec, a, b, c, pad=24))
    assert (err <= 2 * (k + 2) * float(np.finfo(np.float32).eps) * (mag + cm)).all()


def test_dgemm_minus_padded_leading_dimensions(ctx):
    m, n, k = 520, 300, 128
    a, b, c = operands(401, m, n, k, np.float64)
    err, mag, cm = scaled_error(a, b, c, run(ctx, 0, a, b, c, pad=40))
    assert (err <= 2 * (k + 2) * np.finfo(np.float64).eps * (mag + cm)).all()


def test_gemm_minus_argument_checks(ctx):
    lib, h = ctx.lib, ctx.h
    d = L.DeviceArray(4, 4, np.float32)
    assert lib.lair_sgemm_minus_device(h, 9, 4, 4, 4, d.ptr, 4, d.ptr, 4, d.ptr, 4) == -2
    assert lib.lair_sgemm_minus_device(h, _capi.PREC_FP32, -1, 4, 4, d.ptr, 4, d.ptr, 4, d.ptr, 4) == -3
    assert lib.lair_sgemm_minus_device(h, _capi.PREC_FP32, 4, 4, 4, d.ptr, 3, d.ptr, 4, d.ptr, 4) == -7
    assert lib.lair_dgemm_minus_device(h, 4, 4, 4, d.ptr, 4, d.ptr, 4, d.ptr, 3) == -10
    # Empty products leave C alone.
    assert lib.lair_dgemm_minus_device(h, 4, 4, 0, None, 0, None, 4, None, 4) == _capi.LAIR_OK
