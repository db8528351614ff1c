/*
 * lair_oracle.c -- CPU oracle for the LU factor-and-solve path of vinesystems/lair.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference's algorithms in
 * plain C so that the GPU library can be checked against them.  Only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it; the
 * product library never links or calls it.
 *
 * Every function cites the reference file:line it follows.  Arithmetic order is
 * kept exactly: reciprocal-then-multiply for the pivot column, separate multiply
 * and subtract (no FMA; build with -ffp-contract=off), forward row interchanges,
 * first-index strict-greater argmax.  Rust/LLVM never contracts a*b-c into an FMA
 * without an explicit mul_add, so this restatement produces the same bits as the
 * compiled reference on IEEE-754 doubles.
 *
 * Index and pivot convention: 0-based row indices in int64, identical to the
 * reference's Vec<usize> (and to LAPACK ipiv minus one).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

/* ---------------------------------------------------------------- iamax --- */

/* src/blas/iamax.rs:6-21 -- first index of the largest |x|, strict '>' from a
 * running maximum that starts at 0, so NaN and 0 are never selected and an
 * all-zero/all-NaN vector yields (0, 0). */
void lo_diamax(i64 n, const double *x, i64 incx, i64 *idx, double *val) {
  double max_val = 0.0;
  i64 max_idx = 0;
  for (i64 i = 0; i < n; ++i) {
    double v = fabs(x[incx * i]);
    if (v > max_val) {
      max_val = v;
      max_idx = i;
    }
  }
  *idx = max_idx;
  *val = max_val;
}

void lo_siamax(i64 n, const float *x, i64 incx, i64 *idx, float *val) {
  float max_val = 0.0f;
  i64 max_idx = 0;
  for (i64 i = 0; i < n; ++i) {
    float v = fabsf(x[incx * i]);
    if (v > max_val) {
      max_val = v;
      max_idx = i;
    }
  }
  *idx = max_idx;
  *val = max_val;
}

/* Complex variant: |re| + |im| (src/blas/iamax.rs:14). */
static void lo_ziamax(i64 n, const double *x, i64 incx, i64 *idx, double *val) {
  double max_val = 0.0;
  i64 max_idx = 0;
  for (i64 i = 0; i < n; ++i) {
    const double *e = x + 2 * incx * i;
    double v = fabs(e[0]) + fabs(e[1]);
    if (v > max_val) {
      max_val = v;
      max_idx = i;
    }
  }
  *idx = max_idx;
  *val = max_val;
}

/* ---------------------------------------------------------------- laswp --- */

/* src/lapack/laswp.rs:11-40 -- rows i <-> piv[i] for i in begin..npiv, in
 * forward order, over ncols columns.  Like the reference, ncols == 0 still
 * touches the first element pair once; callers never pass 0. */
void lo_dlaswp(i64 ncols, double *a, i64 rs, i64 cs, i64 begin, i64 npiv, const i64 *piv) {
  for (i64 i = begin; i < npiv; ++i) {
    i64 p = piv[i];
    if (p == i) continue;
    double *r1 = a + i * rs, *r2 = a + p * rs;
    i64 n = ncols;
    for (;;) {
      double t = *r1;
      *r1 = *r2;
      *r2 = t;
      if (n == 1) break;
      r1 += cs;
      r2 += cs;
      --n;
    }
  }
}

void lo_slaswp(i64 ncols, float *a, i64 rs, i64 cs, i64 begin, i64 npiv, const i64 *piv) {
  for (i64 i = begin; i < npiv; ++i) {
    i64 p = piv[i];
    if (p == i) continue;
    float *r1 = a + i * rs, *r2 = a + p * rs;
    i64 n = ncols;
    for (;;) {
      float t = *r1;
      *r1 = *r2;
      *r2 = t;
      if (n == 1) break;
      r1 += cs;
      r2 += cs;
      --n;
    }
  }
}

/* src/lapack/getrf.rs:324-333 */
static void swap_rows_d(i64 n, double *r1, double *r2, i64 stride) {
  while (n > 0) {
    double t = *r1;
    *r1 = *r2;
    *r2 = t;
    r1 += stride;
    r2 += stride;
    --n;
  }
}

/* ------------------------------------------------------ getrf_row_major --- */

/* src/lapack/getrf.rs:46-120 -- unblocked right-looking LU, rank-1 form.
 * Per step ul: iamax on column ul; swap the full rows when the argmax is not
 * the diagonal; a zero maximum records singular = ul (the LAST such column wins,
 * getrf.rs:72-73) and skips elimination; otherwise every row j below is scaled
 * by the reciprocal of the pivot and then a[j,k] -= a[j,ul] * a[ul,k].
 * The per-row updates of one step are independent, so splitting rows across
 * threads yields identical bits. */
void lo_dgetrf_row_major(i64 m, i64 n, double *a, i64 rs, i64 cs, i64 *piv, i64 *singular,
                         int nthreads) {
  i64 dim_min = m < n ? m : n;
  i64 sing = -1;
  for (i64 i = 0; i < dim_min; ++i) piv[i] = i;
  for (i64 ul = 0; ul < dim_min; ++ul) {
    i64 max_idx;
    double max_val;
    lo_diamax(m - ul, a + rs * ul + cs * ul, rs, &max_idx, &max_val);
    if (max_idx != 0) {
      i64 max_row = max_idx + ul;
      piv[ul] = max_row;
      swap_rows_d(n, a + ul * rs, a + max_row * rs, cs);
    }
    if (max_val == 0.0) {
      sing = max_idx + ul;
      continue;
    }
    const double pivot_recip = 1.0 / a[ul * rs + ul * cs];
    const double *urow = a + ul * rs;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1 && (m - ul) * (n - ul) > 65536)
    for (i64 j = ul + 1; j < m; ++j) {
      double *row = a + j * rs;
      row[ul * cs] *= pivot_recip;
      const double ratio = row[ul * cs];
      for (i64 k = ul + 1; k < n; ++k) {
        const double elem = ratio * urow[k * cs];
        row[k * cs] -= elem;
      }
    }
  }
  *singular = sing;
}

void lo_sgetrf_row_major(i64 m, i64 n, float *a, i64 rs, i64 cs, i64 *piv, i64 *singular,
                         int nthreads) {
  i64 dim_min = m < n ? m : n;
  i64 sing = -1;
  for (i64 i = 0; i < dim_min; ++i) piv[i] = i;
  for (i64 ul = 0; ul < dim_min; ++ul) {
    i64 max_idx;
    float max_val;
    lo_siamax(m - ul, a + rs * ul + cs * ul, rs, &max_idx, &max_val);
    if (max_idx != 0) {
      i64 max_row = max_idx + ul;
      piv[ul] = max_row;
      float *r1 = a + ul * rs, *r2 = a + max_row * rs;
      for (i64 k = 0; k < n; ++k) {
        float t = r1[k * cs];
        r1[k * cs] = r2[k * cs];
        r2[k * cs] = t;
      }
    }
    if (max_val == 0.0f) {
      sing = max_idx + ul;
      continue;
    }
    const float pivot_recip = 1.0f / a[ul * rs + ul * cs];
    const float *urow = a + ul * rs;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1 && (m - ul) * (n - ul) > 65536)
    for (i64 j = ul + 1; j < m; ++j) {
      float *row = a + j * rs;
      row[ul * cs] *= pivot_recip;
      const float ratio = row[ul * cs];
      for (i64 k = ul + 1; k < n; ++k) {
        const float elem = ratio * urow[k * cs];
        row[k * cs] -= elem;
      }
    }
  }
  *singular = sing;
}

/* Complex128 row-major LU (same file:lines; complex reciprocal and product as in
 * num-complex 0.4: 1/z = conj(z)/|z|^2 computed as (re/(re^2+im^2), -im/(re^2+im^2)),
 * (a+bi)(c+di) = (ac-bd) + (ad+bc)i).  Storage is interleaved (re, im). */
void lo_zgetrf_row_major(i64 m, i64 n, double *a, i64 rs, i64 cs, i64 *piv, i64 *singular) {
  i64 dim_min = m < n ? m : n;
  i64 sing = -1;
  for (i64 i = 0; i < dim_min; ++i) piv[i] = i;
#define ZE(r, c) (a + 2 * ((r) * rs + (c) * cs))
  for (i64 ul = 0; ul < dim_min; ++ul) {
    i64 max_idx;
    double max_val;
    lo_ziamax(m - ul, ZE(ul, ul), rs, &max_idx, &max_val);
    if (max_idx != 0) {
      i64 max_row = max_idx + ul;
      piv[ul] = max_row;
      for (i64 k = 0; k < n; ++k) {
        double *p1 = ZE(ul, k), *p2 = ZE(max_row, k);
        double t0 = p1[0], t1 = p1[1];
        p1[0] = p2[0];
        p1[1] = p2[1];
        p2[0] = t0;
        p2[1] = t1;
      }
    }
    if (max_val == 0.0) {
      sing = max_idx + ul;
      continue;
    }
    /* num-complex Inv/Div of one by z: (1*re + 0*im)/norm_sqr, (0*re - 1*im)/norm_sqr */
    double pr = ZE(ul, ul)[0], pi = ZE(ul, ul)[1];
    double ns = pr * pr + pi * pi;
    double rr = (1.0 * pr + 0.0 * pi) / ns;
    double ri = (0.0 * pr - 1.0 * pi) / ns;
    for (i64 j = ul + 1; j < m; ++j) {
      double *lj = ZE(j, ul);
      double xr = lj[0] * rr - lj[1] * ri;
      double xi = lj[0] * ri + lj[1] * rr;
      lj[0] = xr;
      lj[1] = xi;
      for (i64 k = ul + 1; k < n; ++k) {
        const double *u = ZE(ul, k);
        double er = xr * u[0] - xi * u[1];
        double ei = xr * u[1] + xi * u[0];
        double *t = ZE(j, k);
        t[0] -= er;
        t[1] -= ei;
      }
    }
  }
#undef ZE
  *singular = sing;
}

/* Complex<f32> row-major LU: lo_zgetrf_row_major in single precision (the
 * reference's generic getrf_row_major with A = Complex<f32>; iamax's key
 * |re| + |im| and every product and quotient in f32). */
void lo_cgetrf_row_major(i64 m, i64 n, float *a, i64 rs, i64 cs, i64 *piv, i64 *singular) {
  i64 dim_min = m < n ? m : n;
  i64 sing = -1;
  for (i64 i = 0; i < dim_min; ++i) piv[i] = i;
#define CE(r, c) (a + 2 * ((r) * rs + (c) * cs))
  for (i64 ul = 0; ul < dim_min; ++ul) {
    i64 max_idx = 0;
    float max_val = 0.0f;
    for (i64 i = 0; i < m - ul; ++i) {
      const float *e = CE(ul + i, ul);
      float v = fabsf(e[0]) + fabsf(e[1]);
      if (v > max_val) {
        max_val = v;
        max_idx = i;
      }
    }
    if (max_idx != 0) {
      i64 max_row = max_idx + ul;
      piv[ul] = max_row;
      for (i64 k = 0; k < n; ++k) {
        float *p1 = CE(ul, k), *p2 = CE(max_row, k);
        float t0 = p1[0], t1 = p1[1];
        p1[0] = p2[0];
        p1[1] = p2[1];
        p2[0] = t0;
        p2[1] = t1;
      }
    }
    if (max_val == 0.0f) {
      sing = max_idx + ul;
      continue;
    }
    float pr = CE(ul, ul)[0], pi = CE(ul, ul)[1];
    float ns = pr * pr + pi * pi;
    float rr = (1.0f * pr + 0.0f * pi) / ns;
    float ri = (0.0f * pr - 1.0f * pi) / ns;
    for (i64 j = ul + 1; j < m; ++j) {
      float *lj = CE(j, ul);
      float xr = lj[0] * rr - lj[1] * ri;
      float xi = lj[0] * ri + lj[1] * rr;
      lj[0] = xr;
      lj[1] = xi;
      for (i64 k = ul + 1; k < n; ++k) {
        const float *u = CE(ul, k);
        float er = xr * u[0] - xi * u[1];
        float ei = xr * u[1] + xi * u[0];
        float *t = CE(j, k);
        t[0] -= er;
        t[1] -= ei;
      }
    }
  }
#undef CE
  *singular = sing;
}

/* src/lapack/getrs.rs:12-39 for Complex<f64> (num-complex 0.4), each of the
 * nrhs columns of b (n x nrhs, row stride ldb; all strides in complex
 * elements): x = b; forward laswp; x[i] -= a[i][k] * x[k] for k < i in order
 * (product (ar xr - ai xi, ar xi + ai xr), then the parts subtracted); back
 * substitution likewise for k > i in order, then x[i] /= a[i][i] as
 * DivAssign: re = (xr dr + xi di) / |d|^2, im = (xi dr - xr di) / |d|^2 with
 * |d|^2 = dr dr + di di. */
void lo_zgetrs(i64 n, i64 nrhs, const double *lu, i64 rs, i64 cs, const i64 *piv, double *b, i64 ldb) {
  double *x = (double *)malloc(sizeof(double) * 2 * (n > 0 ? n : 1));
#define ZA(r, c) (lu + 2 * ((r) * rs + (c) * cs))
  for (i64 r = 0; r < nrhs; ++r) {
    for (i64 i = 0; i < n; ++i) {
      x[2 * i] = b[2 * (i * ldb + r)];
      x[2 * i + 1] = b[2 * (i * ldb + r) + 1];
    }
    for (i64 i = 0; i < n; ++i) {
      const i64 p = piv[i];
      if (p != i) {
        double t0 = x[2 * i], t1 = x[2 * i + 1];
        x[2 * i] = x[2 * p];
        x[2 * i + 1] = x[2 * p + 1];
        x[2 * p] = t0;
        x[2 * p + 1] = t1;
      }
    }
    for (i64 i = 0; i < n; ++i)
      for (i64 k = 0; k < i; ++k) {
        const double *e = ZA(i, k);
        double pr = e[0] * x[2 * k] - e[1] * x[2 * k + 1];
        double pi = e[0] * x[2 * k + 1] + e[1] * x[2 * k];
        x[2 * i] -= pr;
        x[2 * i + 1] -= pi;
      }
    for (i64 i = n - 1; i >= 0; --i) {
      for (i64 k = i + 1; k < n; ++k) {
        const double *e = ZA(i, k);
        double pr = e[0] * x[2 * k] - e[1] * x[2 * k + 1];
        double pi = e[0] * x[2 * k + 1] + e[1] * x[2 * k];
        x[2 * i] -= pr;
        x[2 * i + 1] -= pi;
      }
      const double *d = ZA(i, i);
      double xr = x[2 * i], xi = x[2 * i + 1];
      double ns = d[0] * d[0] + d[1] * d[1];
      x[2 * i] = (xr * d[0] + xi * d[1]) / ns;
      x[2 * i + 1] = (xi * d[0] - xr * d[1]) / ns;
    }
    for (i64 i = 0; i < n; ++i) {
      b[2 * (i * ldb + r)] = x[2 * i];
      b[2 * (i * ldb + r) + 1] = x[2 * i + 1];
    }
  }
#undef ZA
  free(x);
}

/* lo_zgetrs for Complex<f32>. */
void lo_cgetrs(i64 n, i64 nrhs, const float *lu, i64 rs, i64 cs, const i64 *piv, float *b, i64 ldb) {
  float *x = (float *)malloc(sizeof(float) * 2 * (n > 0 ? n : 1));
#define CA(r, c) (lu + 2 * ((r) * rs + (c) * cs))
  for (i64 r = 0; r < nrhs; ++r) {
    for (i64 i = 0; i < n; ++i) {
      x[2 * i] = b[2 * (i * ldb + r)];
      x[2 * i + 1] = b[2 * (i * ldb + r) + 1];
    }
    for (i64 i = 0; i < n; ++i) {
      const i64 p = piv[i];
      if (p != i) {
        float t0 = x[2 * i], t1 = x[2 * i + 1];
        x[2 * i] = x[2 * p];
        x[2 * i + 1] = x[2 * p + 1];
        x[2 * p] = t0;
        x[2 * p + 1] = t1;
      }
    }
    for (i64 i = 0; i < n; ++i)
      for (i64 k = 0; k < i; ++k) {
        const float *e = CA(i, k);
        float pr = e[0] * x[2 * k] - e[1] * x[2 * k + 1];
        float pi = e[0] * x[2 * k + 1] + e[1] * x[2 * k];
        x[2 * i] -= pr;
        x[2 * i + 1] -= pi;
      }
    for (i64 i = n - 1; i >= 0; --i) {
      for (i64 k = i + 1; k < n; ++k) {
        const float *e = CA(i, k);
        float pr = e[0] * x[2 * k] - e[1] * x[2 * k + 1];
        float pi = e[0] * x[2 * k + 1] + e[1] * x[2 * k];
        x[2 * i] -= pr;
        x[2 * i + 1] -= pi;
      }
      const float *d = CA(i, i);
      float xr = x[2 * i], xi = x[2 * i + 1];
      float ns = d[0] * d[0] + d[1] * d[1];
      x[2 * i] = (xr * d[0] + xi * d[1]) / ns;
      x[2 * i + 1] = (xi * d[0] - xr * d[1]) / ns;
    }
    for (i64 i = 0; i < n; ++i) {
      b[2 * (i * ldb + r)] = x[2 * i];
      b[2 * (i * ldb + r) + 1] = x[2 * i + 1];
    }
  }
#undef CA
  free(x);
}

/* ------------------------------------------------------ getrf_col_major --- */

/* src/blas/dot.rs:39-133 -- four split accumulators over blocks of four, summed
 * left to right, then up to three trailing products added one at a time. */
static double dot4(i64 n, const double *x, i64 incx, const double *y, i64 incy) {
  i64 rem = n;
  double sum;
  if (rem >= 4) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (;;) {
      s0 += x[0] * y[0];
      s1 += x[incx] * y[incy];
      s2 += x[2 * incx] * y[2 * incy];
      s3 += x[3 * incx] * y[3 * incy];
      x += 4 * incx;
      y += 4 * incy;
      rem -= 4;
      if (rem < 4) break;
    }
    sum = s0 + s1 + s2 + s3;
  } else {
    sum = 0.0;
  }
  if (rem > 0) {
    sum += x[0] * y[0];
    x += incx;
    y += incy;
  }
  if (rem > 1) {
    sum += x[0] * y[0];
    x += incx;
    y += incy;
  }
  if (rem > 2) sum += x[0] * y[0];
  return sum;
}

/* src/lapack/getrf.rs:128-213 -- left-looking LU taken for non-standard layouts:
 * lazy application of earlier pivots to column ul, unit-L forward substitution
 * with dot (dot.rs), gemv::notrans column-by-column update of the lower part
 * (gemv.rs:27-47: y[r] += (alpha*x[c]) * a[r,c]), iamax, pivots[ul] always
 * written, singular = ul when the chosen pivot equals zero, swap of rows ul and
 * pivot over columns 0..=ul only, scaling of the sub-column by the reciprocal. */
void lo_dgetrf_col_major(i64 m, i64 n, double *a, i64 rs, i64 cs, i64 *piv, i64 *singular) {
  i64 dim_min = m < n ? m : n;
  i64 sing = -1;
  for (i64 i = 0; i < dim_min; ++i) piv[i] = i;
  for (i64 ul = 0; ul < n; ++ul) {
    double *col = a + ul * cs;
    i64 n_upper = ul < m ? ul : m;
    for (i64 i = 0; i < n_upper; ++i) {
      i64 ip = piv[i];
      if (ip != i) {
        double t = col[i * rs];
        col[i * rs] = col[ip * rs];
        col[ip * rs] = t;
      }
    }
    for (i64 i = 1; i < n_upper; ++i) {
      double s = dot4(i, a + i * rs, cs, col, rs);
      col[i * rs] -= s;
    }
    if (ul < m) {
      /* gemv::notrans(-1, left[ul.., ..ul], col[..ul], 1, col[ul..]) */
      for (i64 c = 0; c < ul; ++c) {
        const double alpha_x = -1.0 * col[c * rs];
        for (i64 r = ul; r < m; ++r) col[r * rs] += alpha_x * a[r * rs + c * cs];
      }
      i64 max_row;
      double mv;
      lo_diamax(m - ul, col + ul * rs, rs, &max_row, &mv);
      i64 pivot_row = ul + max_row;
      piv[ul] = pivot_row;
      double pivot = col[pivot_row * rs];
      if (pivot == 0.0) {
        sing = ul;
      } else {
        double recip = 1.0 / pivot;
        if (pivot_row != ul) swap_rows_d(ul + 1, a + ul * rs, a + pivot_row * rs, cs);
        for (i64 r = ul + 1; r < m; ++r) col[r * rs] *= recip;
      }
    }
  }
  *singular = sing;
}

/* src/lapack/getrf.rs:12-27 -- dispatch on ndarray's is_standard_layout: C-order
 * (any zero-length axis, or: last axis unit stride unless its length is 1, first
 * axis stride equal to the column count unless its length is 1). */
int lo_is_standard_layout(i64 m, i64 n, i64 rs, i64 cs) {
  if (m == 0 || n == 0) return 1;
  if (n != 1 && cs != 1) return 0;
  if (m != 1 && rs != n) return 0;
  return 1;
}

void lo_dgetrf(i64 m, i64 n, double *a, i64 rs, i64 cs, i64 *piv, i64 *singular, int nthreads) {
  if (lo_is_standard_layout(m, n, rs, cs))
    lo_dgetrf_row_major(m, n, a, rs, cs, piv, singular, nthreads);
  else
    lo_dgetrf_col_major(m, n, a, rs, cs, piv, singular);
}

/* --------------------------------------------------- BLAS trsm and gemm --- */

/* src/blas/trsm.rs:6-22 -- B := L^-1 B with unit-lower L (row/col strides),
 * column by column, skipping zero entries of B. */
void lo_dtrsm_llu(const double *l, i64 lrs, i64 lcs, i64 bm, i64 bn, double *b, i64 brs, i64 bcs) {
  for (i64 j = 0; j < bn; ++j)
    for (i64 k = 0; k < bm; ++k) {
      if (b[k * brs + j * bcs] == 0.0) continue;
      for (i64 i = k + 1; i < bm; ++i) {
        double prod = b[k * brs + j * bcs] * l[i * lrs + k * lcs];
        b[i * brs + j * bcs] -= prod;
      }
    }
}

/* src/blas/gemm.rs:6-32 -- C += alpha * (A*B): a dot product accumulated from
 * zero in k order, then scaled by alpha and added. */
void lo_dgemm(double alpha, i64 m, i64 n, i64 k, const double *a, i64 ars, i64 acs,
              const double *b, i64 brs, i64 bcs, double *c, i64 crs, i64 ccs) {
  for (i64 i = 0; i < m; ++i)
    for (i64 j = 0; j < n; ++j) {
      double s = 0.0;
      for (i64 p = 0; p < k; ++p) s = s + a[i * ars + p * acs] * b[p * brs + j * bcs];
      c[i * crs + j * ccs] += alpha * s;
    }
}

/* ------------------------------------------------------ recursive LU ------ */

/* src/lapack/getrf.rs:216-322 -- Toledo recursive LU (test-only in the
 * reference).  Returns 0 when no singular pivot was met, otherwise the singular
 * row (row 0 conflated with "none", as at getrf.rs:266-267,317). */
static i64 recursive_inner(i64 m, i64 n, double *a, i64 rs, i64 cs, i64 *piv) {
  if (m == 0 || n == 0) return 0;
  if (m == 1) {
    piv[0] = 0;
    return a[0] == 0.0 ? -2 /* Singular(0) */ : 0;
  }
  if (n == 1) {
    i64 idx;
    double mv;
    lo_diamax(m, a, rs, &idx, &mv);
    piv[0] = idx;
    if (mv == 0.0) return -2;
    if (idx != 0) {
      double t = a[0];
      a[0] = a[idx * rs];
      a[idx * rs] = t;
    }
    if (mv >= DBL_MIN) {
      double r = 1.0 / a[0];
      for (i64 i = 1; i < m; ++i) a[i * rs] = r * a[i * rs]; /* blas::scal */
    } else {
      double p = a[0];
      for (i64 i = 1; i < m; ++i) a[i * rs] /= p;
    }
    return 0;
  }
  i64 left = (m < n ? m : n) / 2;
  i64 right = n - left;
  i64 r = recursive_inner(m, left, a, rs, cs, piv);
  i64 sing = (r == -2) ? 0 : r;
  lo_dlaswp(right, a + cs * left, rs, cs, 0, left, piv);
  lo_dtrsm_llu(a, rs, cs, left, right, a + cs * left, rs, cs);
  lo_dgemm(-1.0, m - left, right, left, a + rs * left, rs, cs, a + cs * left, rs, cs,
           a + rs * left + cs * left, rs, cs);
  r = recursive_inner(m - left, right, a + rs * left + cs * left, rs, cs, piv + left);
  if (r != 0) {
    i64 row = (r == -2) ? 0 : r;
    if (sing == 0) sing = left + row;
  }
  i64 min_dim = m < n ? m : n;
  for (i64 i = left; i < min_dim; ++i) piv[i] += left;
  lo_dlaswp(left, a, rs, cs, left, min_dim, piv);
  return sing;
}

/* getrf.rs:30-40; returns 0 for Ok, otherwise 1 + singular row. */
i64 lo_dgetrf_recursive(i64 m, i64 n, double *a, i64 rs, i64 cs, i64 *piv) {
  for (i64 i = 0; i < m; ++i) piv[i] = 0;
  i64 r = recursive_inner(m, n, a, rs, cs, piv);
  if (r == 0) return 0;
  return 1 + ((r == -2) ? 0 : r);
}

/* --------------------------------------------------------------- getrs ---- */

/* src/lapack/getrs.rs:12-39 -- x = b; forward laswp; unit-L forward substitution
 * (x[i] -= a[i,k]*x[k] for k < i in order); U back substitution (x[i] -=
 * a[i,k]*x[k] for k > i in order, then x[i] /= a[i,i]).  Applied to each of
 * nrhs right-hand sides (columns of the row-major n x nrhs array b, stride ldb)
 * independently. */
void lo_dgetrs(i64 n, i64 nrhs, const double *lu, i64 rs, i64 cs, const i64 *piv, double *b,
               i64 ldb) {
  double *x = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  for (i64 r = 0; r < nrhs; ++r) {
    for (i64 i = 0; i < n; ++i) x[i] = b[i * ldb + r];
    if (n > 0) lo_dlaswp(1, x, 1, 1, 0, n, piv);
    for (i64 i = 0; i < n; ++i)
      for (i64 k = 0; k < i; ++k) {
        double prod = lu[i * rs + k * cs] * x[k];
        x[i] -= prod;
      }
    for (i64 i = n - 1; i >= 0; --i) {
      for (i64 k = i + 1; k < n; ++k) {
        double prod = lu[i * rs + k * cs] * x[k];
        x[i] -= prod;
      }
      x[i] /= lu[i * rs + i * cs];
    }
    for (i64 i = 0; i < n; ++i) b[i * ldb + r] = x[i];
  }
  free(x);
}

void lo_sgetrs(i64 n, i64 nrhs, const float *lu, i64 rs, i64 cs, const i64 *piv, float *b, i64 ldb) {
  float *x = (float *)malloc(sizeof(float) * (n > 0 ? n : 1));
  for (i64 r = 0; r < nrhs; ++r) {
    for (i64 i = 0; i < n; ++i) x[i] = b[i * ldb + r];
    if (n > 0) lo_slaswp(1, x, 1, 1, 0, n, piv);
    for (i64 i = 0; i < n; ++i)
      for (i64 k = 0; k < i; ++k) {
        float prod = lu[i * rs + k * cs] * x[k];
        x[i] -= prod;
      }
    for (i64 i = n - 1; i >= 0; --i) {
      for (i64 k = i + 1; k < n; ++k) {
        float prod = lu[i * rs + k * cs] * x[k];
        x[i] -= prod;
      }
      x[i] /= lu[i * rs + i * cs];
    }
    for (i64 i = 0; i < n; ++i) b[i * ldb + r] = x[i];
  }
  free(x);
}

/* Transposed solve A^T y = c from the same factors (no reference counterpart;
 * the adjoint needs it).  With P A = L U (P the forward product of the swaps):
 * U^T w = c forward, L^T z = w backward (unit diagonal), y = P^T z by applying
 * the interchanges in reverse order. */
void lo_dgetrs_trans(i64 n, i64 nrhs, const double *lu, i64 rs, i64 cs, const i64 *piv, double *b,
                     i64 ldb) {
  double *x = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  for (i64 r = 0; r < nrhs; ++r) {
    for (i64 i = 0; i < n; ++i) x[i] = b[i * ldb + r];
    for (i64 i = 0; i < n; ++i) {
      for (i64 k = 0; k < i; ++k) x[i] -= lu[k * rs + i * cs] * x[k];
      x[i] /= lu[i * rs + i * cs];
    }
    for (i64 i = n - 1; i >= 0; --i)
      for (i64 k = i + 1; k < n; ++k) x[i] -= lu[k * rs + i * cs] * x[k];
    for (i64 i = n - 1; i >= 0; --i) {
      i64 p = piv[i];
      if (p != i) {
        double t = x[i];
        x[i] = x[p];
        x[p] = t;
      }
    }
    for (i64 i = 0; i < n; ++i) b[i * ldb + r] = x[i];
  }
  free(x);
}

/* ----------------------------------------------------- equation::solve ---- */

/* src/equation.rs:32-60 -- returns 0 on success, 1 for InvalidInput::Shape
 * (not square, or b length mismatch), 2 for InvalidInput::Value (singular).
 * a is copied (equation.rs:54) and left untouched; x receives the solution. */
int lo_dsolve(i64 m, i64 n, const double *a, i64 rs, i64 cs, i64 blen, const double *b, double *x) {
  if (m != n) return 1;
  if (blen != m) return 1;
  double *w = (double *)malloc(sizeof(double) * (n * n > 0 ? n * n : 1));
  i64 *piv = (i64 *)malloc(sizeof(i64) * (n > 0 ? n : 1));
  for (i64 i = 0; i < n; ++i)
    for (i64 j = 0; j < n; ++j) w[i * n + j] = a[i * rs + j * cs];
  i64 sing;
  lo_dgetrf(n, n, w, n, 1, piv, &sing, 1);
  int st = 0;
  if (sing >= 0) {
    st = 2;
  } else {
    for (i64 i = 0; i < n; ++i) x[i] = b[i];
    lo_dgetrs(n, 1, w, n, 1, piv, x, 1);
  }
  free(w);
  free(piv);
  return st;
}

/* ----------------------------------------------- Factorized accessors ----- */

/* src/decomposition/lu.rs:28-39 -- permutation vector obtained by applying the
 * pivots (forward laswp) to 0..nrows; P[perm[i], i] = 1. */
void lo_lu_perm(i64 nrows, i64 npiv, const i64 *piv, i64 *perm) {
  for (i64 i = 0; i < nrows; ++i) perm[i] = i;
  for (i64 i = 0; i < npiv; ++i) {
    i64 p = piv[i];
    if (p == i) continue;
    i64 t = perm[i];
    perm[i] = perm[p];
    perm[p] = t;
  }
}

/* ------------------------------------------------------ helper kernels ---- */

/* Matrix-vector residual r = b - A x in double (row-major, ldA), for checks. */
void lo_dresidual(i64 n, i64 nrhs, const double *a, i64 lda, const double *x, i64 ldx,
                  const double *b, i64 ldb, double *r, i64 ldr, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
  for (i64 i = 0; i < n; ++i)
    for (i64 c = 0; c < nrhs; ++c) {
      long double s = 0.0L;
      for (i64 k = 0; k < n; ++k) s += (long double)a[i * lda + k] * (long double)x[k * ldx + c];
      r[i * ldr + c] = (double)((long double)b[i * ldb + c] - s);
    }
}

int lo_num_threads_available(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
