// HIP kernels of the gfx950 LU factor-and-solve path.
//
//  * panel_coop: exact partial pivoting of a column leaf, rows spread over a
//    few cooperating workgroups that synchronize once per column through
//    per-workgroup sequence flags (src/lapack/getrf.rs:55-117, src/blas/iamax.rs:6-21)
//  * laswp: row interchanges over column ranges (src/lapack/laswp.rs:11-40)
//  * trsm: LDS-resident triangular solve of a block of right-hand sides
//    (src/blas/trsm.rs:6-22, and the substitutions of src/lapack/getrs.rs:24-36)
//  * gemm_minus: C -= A*B on MFMA (src/blas/gemm.rs:6-32 with alpha = -1)
//  * FP64 residual, norms and conversions for iterative refinement.
#include <climits>

#include "lair_internal.h"

namespace lair {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// Bounded wait: a cooperative panel gives up (and flags the error) after this
// many polls instead of hanging when a peer never arrives.
constexpr int kSpinLimit = 1 << 23;

template <typename T>
__device__ __forceinline__ void st_agent(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_agent(const T *p) {
  return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// (value, row) ordering of iamax: larger |x| wins, ties go to the lower row.
// Candidates only exist for |x| > 0 (strict '>' against a running maximum that
// starts at zero), so NaN and zero never win; "none" is (0, INT_MAX).
__device__ __forceinline__ void better(double &bv, int &br, double ov, int orr) {
  if (ov > bv || (ov == bv && orr < br)) {
    bv = ov;
    br = orr;
  }
}

// ---------------------------------------------------------------- panel ---

template <typename T>
struct PanelArgs {
  T *a;
  int64_t lda;
  int row0, m, col0, wcols, npiv, P;
  int64_t *ipiv;
  int64_t *sing;
  uint32_t *flags;
  uint32_t seq;
  double *cand_val;
  int32_t *cand_idx;
  T *cand_row;
  T *diag_row;
  int32_t *err;
};

// Factors columns [col0, col0+wcols) of rows [row0, m) in place.  Workgroup g
// owns rows [row0 + g*R, row0 + (g+1)*R) with R = NT*RPT; thread t keeps rows
// base + t + s*NT (s < RPT) of the leaf in registers.  Per column j:
//  1. local argmax over the rows at or below the diagonal d = row0 + j;
//  2. with P > 1 workgroups: publish (value, row, row data), raise this
//     workgroup's sequence flag, wait for every peer's flag, and reduce the P
//     candidates in the same order everywhere;
//  3. swap the diagonal row and the pivot row, record ipiv[d] (and the
//     singular column when no candidate exists), scale by the reciprocal of
//     the pivot and apply the rank-1 update with separate multiply and
//     subtract, as getrf.rs:76-88 does.
// The diagonal rows of a leaf always lie in workgroup 0 (j < W <= R).  Each
// column step is its own template instance so every register index is a
// compile-time constant.
template <typename T, int W, int RPT, int NT>
struct PanelLeaf {
  static constexpr int NWAVE = NT / 64;
  static constexpr int R = NT * RPT;
  const PanelArgs<T> &pa;
  T (&v)[RPT][W];
  double *s_wv;
  int *s_wr;
  T *s_prow;
  T *s_drow;
  int *s_win;
  int *s_bail;
  int g, t, lane, wave, base;

  template <int J>
  __device__ __forceinline__ bool step() {
#pragma clang fp contract(off)
    if (J >= pa.npiv) return true;
    const int d = pa.row0 + J;
    constexpr int par = J & 1;
    double bv = 0.0;
    int br = INT_MAX;
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      const int r = base + t + s * NT;
      if (r >= d && r < pa.m) {
        const double x = fabs((double)v[s][J]);
        if (x > bv) {
          bv = x;
          br = r;
        }
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double ov = __shfl_xor(bv, off);
      const int orr = __shfl_xor(br, off);
      better(bv, br, ov, orr);
    }
    if (lane == 0) {
      s_wv[wave] = bv;
      s_wr[wave] = br;
    }
    if (t == 0) *s_bail = 0;
    __syncthreads();
    bv = s_wv[0];
    br = s_wr[0];
#pragma unroll
    for (int q = 1; q < NWAVE; ++q) better(bv, br, s_wv[q], s_wr[q]);

    if (pa.P == 1) {
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int r = base + t + s * NT;
        if (r == br) {
#pragma unroll
          for (int w = 0; w < W; ++w) s_prow[w] = v[s][w];
        }
        if (r == d) {
#pragma unroll
          for (int w = 0; w < W; ++w) s_drow[w] = v[s][w];
        }
      }
      if (t == 0) *s_win = br;
      __syncthreads();
    } else {
      T *crow = pa.cand_row + ((size_t)par * pa.P + g) * W;
      bool wrote = false;
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int r = base + t + s * NT;
        if (r == br) {
#pragma unroll
          for (int w = 0; w < W; ++w) st_agent(crow + w, v[s][w]);
          wrote = true;
        }
        if (r == d) {
#pragma unroll
          for (int w = 0; w < W; ++w) st_agent(pa.diag_row + par * W + w, v[s][w]);
          wrote = true;
        }
      }
      if (t == 0) {
        st_agent(pa.cand_val + par * pa.P + g, bv);
        st_agent(pa.cand_idx + par * pa.P + g, br);
        wrote = true;
      }
      if (wrote) __threadfence();
      __syncthreads();
      const uint32_t target = pa.seq + (uint32_t)J + 1u;
      if (t == 0)
        __hip_atomic_store(pa.flags + g, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      if (wave == 0) {
        int spins = 0;
        bool bail = false;
        for (;;) {
          uint32_t f = target;
          if (lane < pa.P)
            f = __hip_atomic_load(pa.flags + lane, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          const bool ok = (int32_t)(f - target) >= 0;
          if (__all(ok)) break;
          if (++spins > kSpinLimit || ld_agent(pa.err) != 0) {
            bail = true;
            break;
          }
        }
        double cv = 0.0;
        int cr = INT_MAX;
        if (!bail && lane < pa.P) {
          cv = ld_agent(pa.cand_val + par * pa.P + lane);
          cr = ld_agent(pa.cand_idx + par * pa.P + lane);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          const double ov = __shfl_xor(cv, off);
          const int orr = __shfl_xor(cr, off);
          better(cv, cr, ov, orr);
        }
        if (lane == 0) {
          *s_win = cr;
          if (bail) {
            *s_bail = 1;
            st_agent(pa.err, 1);
          }
        }
      }
      __syncthreads();
      if (*s_bail) return false;
      const int win = *s_win;
      if (t < W) {
        if (win != INT_MAX) {
          const int gw = (win - pa.row0) / R;
          s_prow[t] = ld_agent(pa.cand_row + ((size_t)par * pa.P + gw) * W + t);
        }
        s_drow[t] = ld_agent(pa.diag_row + par * W + t);
      }
      __syncthreads();
    }

    const int win = *s_win;
    const bool singular = (win == INT_MAX);
    const int pr = singular ? d : win;
    if (pr != d) {
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int r = base + t + s * NT;
        if (r == d) {
#pragma unroll
          for (int w = 0; w < W; ++w) v[s][w] = s_prow[w];
        } else if (r == pr) {
#pragma unroll
          for (int w = 0; w < W; ++w) v[s][w] = s_drow[w];
        }
      }
    }
    if (g == 0 && t == 0) {
      pa.ipiv[d] = pr;
      if (singular) *pa.sing = d;
    }
    if (!singular) {
      const T recip = T(1) / s_prow[J];
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int r = base + t + s * NT;
        if (r > d && r < pa.m) {
          const T l = v[s][J] * recip;
          v[s][J] = l;
#pragma unroll
          for (int k = J + 1; k < W; ++k) {
            const T e = l * s_prow[k];
            v[s][k] = v[s][k] - e;
          }
        }
      }
    }
    __syncthreads();
    return true;
  }
};

template <int J, int W>
struct ColumnLoop {
  template <class F>
  __device__ __forceinline__ static bool run(F &f) {
    if (!f.template step<J>()) return false;
    return ColumnLoop<J + 1, W>::run(f);
  }
};
template <int W>
struct ColumnLoop<W, W> {
  template <class F>
  __device__ __forceinline__ static bool run(F &) {
    return true;
  }
};

template <typename T, int W, int RPT, int NT>
__global__ void __launch_bounds__(NT) panel_coop_kernel(PanelArgs<T> pa) {
  constexpr int NWAVE = NT / 64;
  constexpr int R = NT * RPT;
  const int g = blockIdx.x, t = threadIdx.x;
  const int base = pa.row0 + g * R;

  T v[RPT][W];
#pragma unroll
  for (int s = 0; s < RPT; ++s) {
    const int r = base + t + s * NT;
    const T *src = pa.a + (int64_t)r * pa.lda + pa.col0;
#pragma unroll
    for (int w = 0; w < W; ++w) v[s][w] = (r < pa.m && w < pa.wcols) ? src[w] : T(0);
  }

  __shared__ double s_wv[NWAVE];
  __shared__ int s_wr[NWAVE];
  __shared__ T s_prow[W];
  __shared__ T s_drow[W];
  __shared__ int s_win;
  __shared__ int s_bail;

  PanelLeaf<T, W, RPT, NT> leaf{pa, v, s_wv, s_wr, s_prow, s_drow, &s_win, &s_bail,
                                g, t, t & 63, t >> 6, base};
  if (!ColumnLoop<0, W>::run(leaf)) return;

#pragma unroll
  for (int s = 0; s < RPT; ++s) {
    const int r = base + t + s * NT;
    if (r < pa.m) {
      T *dst = pa.a + (int64_t)r * pa.lda + pa.col0;
#pragma unroll
      for (int w = 0; w < W; ++w)
        if (w < pa.wcols) dst[w] = v[s][w];
    }
  }
}

template <typename T>
hipError_t launch_panel(lair_context *ctx, T *a, int64_t lda, int row0, int m, int col0, int wcols,
                        int64_t *ipiv) {
  const int rows = m - row0;
  if (rows <= 0 || wcols <= 0) return hipSuccess;
  PanelArgs<T> pa;
  pa.a = a;
  pa.lda = lda;
  pa.row0 = row0;
  pa.m = m;
  pa.col0 = col0;
  pa.wcols = wcols;
  pa.npiv = wcols < rows ? wcols : rows;
  pa.ipiv = ipiv;
  pa.sing = ctx->sing;
  pa.flags = ctx->flags;
  pa.seq = ctx->seq;
  pa.cand_val = ctx->cand_val;
  pa.cand_idx = ctx->cand_idx;
  pa.cand_row = reinterpret_cast<T *>(ctx->cand_row);
  pa.diag_row = reinterpret_cast<T *>(ctx->diag_row);
  pa.err = ctx->err;
  constexpr int NT = 256;
  int rpt;
  if (rows <= NT)
    rpt = 1;
  else if (rows <= 2 * NT)
    rpt = 2;
  else
    rpt = 4;
  const int R = NT * rpt;
  pa.P = (rows + R - 1) / R;
  if (pa.P > kMaxPanelWG) return hipErrorInvalidValue;
  const int slot = prof_begin(ctx);
  if (rpt == 1)
    panel_coop_kernel<T, kLeafW, 1, NT><<<pa.P, NT, 0, ctx->stream>>>(pa);
  else if (rpt == 2)
    panel_coop_kernel<T, kLeafW, 2, NT><<<pa.P, NT, 0, ctx->stream>>>(pa);
  else
    panel_coop_kernel<T, kLeafW, 4, NT><<<pa.P, NT, 0, ctx->stream>>>(pa);
  const double fl = (double)pa.npiv * (double)rows * (double)wcols;
  prof_end(ctx, slot, LAIR_KCLASS_PANEL, fl);
  ctx->seq += (uint32_t)kLeafW;
  return hipGetLastError();
}

template hipError_t launch_panel<double>(lair_context *, double *, int64_t, int, int, int, int,
                                         int64_t *);
template hipError_t launch_panel<float>(lair_context *, float *, int64_t, int, int, int, int,
                                        int64_t *);

// ---------------------------------------------------------------- laswp ---

template <typename T>
__global__ void laswp_kernel(T *a, int64_t lda, int c0, int n1, int s1, int total,
                             const int64_t *ipiv, int k0, int k1, int rev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int c = (i < n1) ? c0 + i : s1 + (i - n1);
  if (!rev) {
    for (int k = k0; k < k1; ++k) {
      const int64_t p = ipiv[k];
      if (p != k) {
        T *x = a + (int64_t)k * lda + c;
        T *y = a + p * lda + c;
        const T tmp = *x;
        *x = *y;
        *y = tmp;
      }
    }
  } else {
    for (int k = k1 - 1; k >= k0; --k) {
      const int64_t p = ipiv[k];
      if (p != k) {
        T *x = a + (int64_t)k * lda + c;
        T *y = a + p * lda + c;
        const T tmp = *x;
        *x = *y;
        *y = tmp;
      }
    }
  }
}

template <typename T>
hipError_t launch_laswp(lair_context *ctx, T *a, int64_t lda, int c0, int c1, int s0, int s1,
                        const int64_t *ipiv, int k0, int k1, bool rev) {
  if (s0 < c0) s0 = c0;
  if (s1 > c1) s1 = c1;
  if (s1 < s0) s1 = s0;
  const int n1 = s0 - c0;
  const int total = n1 + (c1 - s1);
  if (total <= 0 || k1 <= k0) return hipSuccess;
  const int slot = prof_begin(ctx);
  laswp_kernel<T><<<(total + 255) / 256, 256, 0, ctx->stream>>>(a, lda, c0, n1, s1, total, ipiv,
                                                                 k0, k1, rev ? 1 : 0);
  prof_end(ctx, slot, LAIR_KCLASS_LASWP, 0.0);
  return hipGetLastError();
}

template hipError_t launch_laswp<double>(lair_context *, double *, int64_t, int, int, int, int,
                                         const int64_t *, int, int, bool);
template hipError_t launch_laswp<float>(lair_context *, float *, int64_t, int, int, int, int,
                                        const int64_t *, int, int, bool);

// ----------------------------------------------------------------- trsm ---

// One workgroup solves a kb x 64 block of right-hand sides held in LDS.
// Forward substitution (effective lower triangle) or backward (effective
// upper).  eff(i,p) = trans ? tri[p][i] : tri[i][p].  After x_i is final its
// contribution is subtracted from every later row (right-looking order);
// a non-unit diagonal uses true division like getrs.rs:35.
template <typename T>
__global__ void __launch_bounds__(256)
    trsm_kernel(int kb, int ncols, const T *tri, int64_t ldt, T *b, int64_t ldb, int forward,
                int trans, int unit) {
#pragma clang fp contract(off)
  extern __shared__ unsigned char smem_raw[];
  T *xs = reinterpret_cast<T *>(smem_raw);  // [kb][64]
  const int t = threadIdx.x;
  const int col = t & 63, rg = t >> 6;
  const int c = blockIdx.x * 64 + col;
  const bool active = c < ncols;
  for (int i = rg; i < kb; i += 4) xs[i * 64 + col] = active ? b[(int64_t)i * ldb + c] : T(0);
  __syncthreads();
  if (forward) {
    for (int i = 0; i < kb; ++i) {
      if (!unit) {
        if (rg == 0) xs[i * 64 + col] = xs[i * 64 + col] / tri[(int64_t)i * ldt + i];
        __syncthreads();
      }
      const T xi = xs[i * 64 + col];
      for (int r = i + 1 + rg; r < kb; r += 4) {
        const T l = trans ? tri[(int64_t)i * ldt + r] : tri[(int64_t)r * ldt + i];
        const T e = l * xi;
        xs[r * 64 + col] = xs[r * 64 + col] - e;
      }
      __syncthreads();
    }
  } else {
    for (int i = kb - 1; i >= 0; --i) {
      if (!unit) {
        if (rg == 0) xs[i * 64 + col] = xs[i * 64 + col] / tri[(int64_t)i * ldt + i];
        __syncthreads();
      }
      const T xi = xs[i * 64 + col];
      for (int r = rg; r < i; r += 4) {
        const T u = trans ? tri[(int64_t)i * ldt + r] : tri[(int64_t)r * ldt + i];
        const T e = u * xi;
        xs[r * 64 + col] = xs[r * 64 + col] - e;
      }
      __syncthreads();
    }
  }
  if (active)
    for (int i = rg; i < kb; i += 4) b[(int64_t)i * ldb + c] = xs[i * 64 + col];
}

template <typename T>
hipError_t launch_trsm(lair_context *ctx, bool lower, bool trans, bool unit, int kb, int ncols,
                       const T *tri, int64_t ldt, T *b, int64_t ldb) {
  if (kb <= 0 || ncols <= 0) return hipSuccess;
  const size_t shm = (size_t)kb * 64 * sizeof(T);
  if (shm > 160 * 1024) return hipErrorInvalidValue;
  const bool forward = (lower && !trans) || (!lower && trans);
  const int slot = prof_begin(ctx);
  trsm_kernel<T><<<(ncols + 63) / 64, 256, shm, ctx->stream>>>(kb, ncols, tri, ldt, b, ldb,
                                                               forward ? 1 : 0, trans ? 1 : 0,
                                                               unit ? 1 : 0);
  prof_end(ctx, slot, LAIR_KCLASS_TRSM, (double)kb * kb * ncols);
  return hipGetLastError();
}

// Small systems (n <= kExactOrderMax): getrs in the reference's exact
// operation order (getrs.rs:23-36): forward interchanges, x[i] -= a[i,k]*x[k]
// for k = 0..i-1 in order, then x[i] -= a[i,k]*x[k] for k = i+1..n-1 in order
// followed by x[i] /= a[i,i].  One thread per right-hand side, x in LDS.
template <typename T>
__global__ void __launch_bounds__(64)
    getrs_exact_kernel(int n, int nrhs, const T *lu, int64_t lda, const int64_t *ipiv, T *b,
                       int64_t ldb) {
#pragma clang fp contract(off)
  extern __shared__ unsigned char smem_raw[];
  T *xs = reinterpret_cast<T *>(smem_raw);  // [n][64]
  const int t = threadIdx.x;
  const int c = blockIdx.x * 64 + t;
  if (c >= nrhs) return;
  for (int i = 0; i < n; ++i) xs[i * 64 + t] = b[(int64_t)i * ldb + c];
  for (int i = 0; i < n; ++i) {
    const int p = (int)ipiv[i];
    if (p != i) {
      const T tmp = xs[i * 64 + t];
      xs[i * 64 + t] = xs[p * 64 + t];
      xs[p * 64 + t] = tmp;
    }
  }
  for (int i = 0; i < n; ++i) {
    T xi = xs[i * 64 + t];
    for (int k = 0; k < i; ++k) {
      const T prod = lu[(int64_t)i * lda + k] * xs[k * 64 + t];
      xi = xi - prod;
    }
    xs[i * 64 + t] = xi;
  }
  for (int i = n - 1; i >= 0; --i) {
    T xi = xs[i * 64 + t];
    for (int k = i + 1; k < n; ++k) {
      const T prod = lu[(int64_t)i * lda + k] * xs[k * 64 + t];
      xi = xi - prod;
    }
    xs[i * 64 + t] = xi / lu[(int64_t)i * lda + i];
  }
  for (int i = 0; i < n; ++i) b[(int64_t)i * ldb + c] = xs[i * 64 + t];
}

template <typename T>
hipError_t launch_getrs_exact(lair_context *ctx, int n, int nrhs, const T *lu, int64_t lda,
                              const int64_t *ipiv, T *b, int64_t ldb) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  if (n > kExactOrderMax) return hipErrorInvalidValue;
  const size_t shm = (size_t)n * 64 * sizeof(T);
  const int slot = prof_begin(ctx);
  getrs_exact_kernel<T><<<(nrhs + 63) / 64, 64, shm, ctx->stream>>>(n, nrhs, lu, lda, ipiv, b, ldb);
  prof_end(ctx, slot, LAIR_KCLASS_GETRS, 2.0 * n * n * nrhs);
  return hipGetLastError();
}

template hipError_t launch_getrs_exact<double>(lair_context *, int, int, const double *, int64_t,
                                               const int64_t *, double *, int64_t);
template hipError_t launch_getrs_exact<float>(lair_context *, int, int, const float *, int64_t,
                                              const int64_t *, float *, int64_t);

template hipError_t launch_trsm<double>(lair_context *, bool, bool, bool, int, int, const double *,
                                        int64_t, double *, int64_t);
template hipError_t launch_trsm<float>(lair_context *, bool, bool, bool, int, int, const float *,
                                       int64_t, float *, int64_t);

// ----------------------------------------------------------------- gemm ---

// 64x64 output tile per workgroup of four waves (2x2 waves, 32x32 each, as
// 2x2 MFMA 16x16 blocks).  Operand lane layout of the 16x16xK MFMAs: lane l
// supplies A[l%16][k-group l/16] and B[k-group l/16][l%16].  Output layout:
// FP64 16x16x4 -> C[l/16 + 4r][l%16]; FP32/BF16 -> C[4(l/16) + r][l%16].
template <bool TRANSA>
__global__ void __launch_bounds__(256)
    gemm_f64_kernel(int M, int N, int K, const double *__restrict__ A, int64_t lda,
                    const double *__restrict__ B, int64_t ldb, double *__restrict__ C, int64_t ldc) {
  constexpr int BK = 16;
  __shared__ double As[64][BK + 1];
  __shared__ double Bs[BK][64 + 2];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int bm = blockIdx.y * 64, bn = blockIdx.x * 64;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
    for (int e = t; e < 64 * BK; e += 256) {
      int i, k;
      if (!TRANSA) {
        i = e / BK;
        k = e % BK;
      } else {
        k = e / 64;
        i = e % 64;
      }
      const int gi = bm + i, gk = k0 + k;
      double val = 0.0;
      if (gi < M && gk < K) val = TRANSA ? A[(int64_t)gk * lda + gi] : A[(int64_t)gi * lda + gk];
      As[i][k] = val;
    }
#pragma unroll
    for (int e = t; e < 64 * BK; e += 256) {
      const int k = e / 64, j = e % 64;
      const int gk = k0 + k, gj = bn + j;
      Bs[k][j] = (gk < K && gj < N) ? B[(int64_t)gk * ldb + gj] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double av[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) av[i] = As[wm + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 2; ++j) bv[j] = Bs[kk + (lane >> 4)][wn + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = bm + wm + i * 16 + (lane >> 4) + 4 * r;
        const int gj = bn + wn + j * 16 + (lane & 15);
        if (gi < M && gj < N) {
          double *cp = C + (int64_t)gi * ldc + gj;
          *cp = *cp - acc[i][j][r];
        }
      }
}

template <bool TRANSA>
__global__ void __launch_bounds__(256)
    gemm_f32_kernel(int M, int N, int K, const float *__restrict__ A, int64_t lda,
                    const float *__restrict__ B, int64_t ldb, float *__restrict__ C, int64_t ldc) {
  constexpr int BK = 16;
  __shared__ float As[64][BK + 1];
  __shared__ float Bs[BK][64 + 4];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int bm = blockIdx.y * 64, bn = blockIdx.x * 64;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
    for (int e = t; e < 64 * BK; e += 256) {
      int i, k;
      if (!TRANSA) {
        i = e / BK;
        k = e % BK;
      } else {
        k = e / 64;
        i = e % 64;
      }
      const int gi = bm + i, gk = k0 + k;
      float val = 0.f;
      if (gi < M && gk < K) val = TRANSA ? A[(int64_t)gk * lda + gi] : A[(int64_t)gi * lda + gk];
      As[i][k] = val;
    }
#pragma unroll
    for (int e = t; e < 64 * BK; e += 256) {
      const int k = e / 64, j = e % 64;
      const int gk = k0 + k, gj = bn + j;
      Bs[k][j] = (gk < K && gj < N) ? B[(int64_t)gk * ldb + gj] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float av[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) av[i] = As[wm + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 2; ++j) bv[j] = Bs[kk + (lane >> 4)][wn + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = bm + wm + i * 16 + 4 * (lane >> 4) + r;
        const int gj = bn + wn + j * 16 + (lane & 15);
        if (gi < M && gj < N) {
          float *cp = C + (int64_t)gi * ldc + gj;
          *cp = *cp - acc[i][j][r];
        }
      }
}

// Splits x into NS bf16 terms x ~= t0 + t1 + t2 (each remainder is exact in
// FP32, so the terms carry 8 + 8 + 8 significand bits).
template <int NS>
__device__ __forceinline__ void split_bf16(float x, __bf16 *out) {
  float r = x;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const __bf16 h = (__bf16)r;
    out[s] = h;
    r = r - (float)h;
  }
}

// BF16 trailing update with NS-way splits of both operands.  NS = 3 keeps the
// six products whose magnitude is at least 2^-16 of the leading one
// (hh, hm, mh, mm, hl, lh), accumulated smallest-first in FP32 by the MFMA.
template <int NS, bool TRANSA>
__global__ void __launch_bounds__(256)
    gemm_bf16_kernel(int M, int N, int K, const float *__restrict__ A, int64_t lda,
                     const float *__restrict__ B, int64_t ldb, float *__restrict__ C, int64_t ldc) {
  constexpr int BK = 32;
  constexpr int PADK = BK + 8;
  __shared__ __bf16 As[NS][64][PADK];  // [split][m][k]
  __shared__ __bf16 Bs[NS][64][PADK];  // [split][n][k] (transposed)
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int bm = blockIdx.y * 64, bn = blockIdx.x * 64;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
    for (int e = t; e < 64 * BK; e += 256) {
      int i, k;
      if (!TRANSA) {
        i = e / BK;
        k = e % BK;
      } else {
        k = e / 64;
        i = e % 64;
      }
      const int gi = bm + i, gk = k0 + k;
      float val = 0.f;
      if (gi < M && gk < K) val = TRANSA ? A[(int64_t)gk * lda + gi] : A[(int64_t)gi * lda + gk];
      __bf16 parts[NS];
      split_bf16<NS>(val, parts);
#pragma unroll
      for (int s = 0; s < NS; ++s) As[s][i][k] = parts[s];
    }
#pragma unroll
    for (int e = t; e < 64 * BK; e += 256) {
      const int k = e / 64, j = e % 64;
      const int gk = k0 + k, gj = bn + j;
      const float val = (gk < K && gj < N) ? B[(int64_t)gk * ldb + gj] : 0.f;
      __bf16 parts[NS];
      split_bf16<NS>(val, parts);
#pragma unroll
      for (int s = 0; s < NS; ++s) Bs[s][j][k] = parts[s];
    }
    __syncthreads();
    bf8 av[NS][2], bv[NS][2];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
        av[s][i] = *reinterpret_cast<const bf8 *>(&As[s][wm + i * 16 + (lane & 15)][(lane >> 4) * 8]);
#pragma unroll
      for (int j = 0; j < 2; ++j)
        bv[s][j] = *reinterpret_cast<const bf8 *>(&Bs[s][wn + j * 16 + (lane & 15)][(lane >> 4) * 8]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        if (NS == 3) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[2][i], bv[0][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0][i], bv[2][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[1][i], bv[1][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[1][i], bv[0][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0][i], bv[1][j], acc[i][j], 0, 0, 0);
        }
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0][i], bv[0][j], acc[i][j], 0, 0, 0);
      }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = bm + wm + i * 16 + 4 * (lane >> 4) + r;
        const int gj = bn + wn + j * 16 + (lane & 15);
        if (gi < M && gj < N) {
          float *cp = C + (int64_t)gi * ldc + gj;
          *cp = *cp - acc[i][j][r];
        }
      }
}

template <>
hipError_t launch_gemm_minus<double>(lair_context *ctx, int mode, int M, int N, int K,
                                     const double *A, int64_t lda, bool transA, const double *B,
                                     int64_t ldb, double *C, int64_t ldc) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  (void)mode;
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  const int slot = prof_begin(ctx);
  if (transA)
    gemm_f64_kernel<true><<<grid, 256, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C, ldc);
  else
    gemm_f64_kernel<false><<<grid, 256, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C, ldc);
  prof_end(ctx, slot, LAIR_KCLASS_GEMM, 2.0 * M * N * K);
  return hipGetLastError();
}

template <>
hipError_t launch_gemm_minus<float>(lair_context *ctx, int mode, int M, int N, int K,
                                    const float *A, int64_t lda, bool transA, const float *B,
                                    int64_t ldb, float *C, int64_t ldc) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  const int slot = prof_begin(ctx);
  hipStream_t st = ctx->stream;
  if (mode == kGemmBF16x6) {
    if (transA)
      gemm_bf16_kernel<3, true><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
    else
      gemm_bf16_kernel<3, false><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
  } else if (mode == kGemmBF16) {
    if (transA)
      gemm_bf16_kernel<1, true><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
    else
      gemm_bf16_kernel<1, false><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
  } else {
    if (transA)
      gemm_f32_kernel<true><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
    else
      gemm_f32_kernel<false><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
  }
  prof_end(ctx, slot, LAIR_KCLASS_GEMM, 2.0 * M * N * K);
  return hipGetLastError();
}

// ------------------------------------------------------ fills, copies ---

template <typename T>
__global__ void fill_kernel(T *p, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_i64(lair_context *ctx, int64_t *p, int64_t n, int64_t v) {
  if (n <= 0) return hipSuccess;
  fill_kernel<int64_t><<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_u32(lair_context *ctx, uint32_t *p, int64_t n, uint32_t v) {
  if (n <= 0) return hipSuccess;
  fill_kernel<uint32_t><<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_f64(lair_context *ctx, double *p, int64_t n, double v) {
  if (n <= 0) return hipSuccess;
  fill_kernel<double><<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(p, n, v);
  return hipGetLastError();
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *bits, double v) {
  // Non-negative doubles order like their bit patterns; NaN maps above +inf.
  atomicMax(bits, (unsigned long long)__double_as_longlong(v));
}

// One wave per row: y = (float) a, and ||row||_1 folded into the norm.
__global__ void convert_d2s_norm_kernel(int n, int ncols, const double *a, int64_t lda, float *a32,
                                        int64_t lda32, unsigned long long *norm_bits) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= n) return;
  const double *src = a + (int64_t)row * lda;
  float *dst = a32 + (int64_t)row * lda32;
  double s = 0.0;
  for (int c = lane; c < ncols; c += 64) {
    const double x = src[c];
    s += fabs(x);
    if (dst) dst[c] = (float)x;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) atomic_max_nonneg(norm_bits, s);
}

hipError_t launch_convert_d2s_norm(lair_context *ctx, int n, int ncols, const double *a,
                                   int64_t lda, float *a32, int64_t lda32,
                                   unsigned long long *norm_bits) {
  if (n <= 0) return hipSuccess;
  const int slot = prof_begin(ctx);
  convert_d2s_norm_kernel<<<(n + 3) / 4, 256, 0, ctx->stream>>>(n, ncols, a, lda, a32, lda32,
                                                                norm_bits);
  prof_end(ctx, slot, LAIR_KCLASS_CONVERT, 0.0);
  return hipGetLastError();
}

hipError_t launch_rows_norm(lair_context *ctx, int n, int ncols, const double *a, int64_t lda,
                            unsigned long long *norm_bits) {
  return launch_convert_d2s_norm(ctx, n, ncols, a, lda, nullptr, 0, norm_bits);
}

template <typename S, typename D>
__global__ void copy2d_kernel(int rows, int cols, const S *x, int64_t ldx, D *y, int64_t ldy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)rows * cols;
  if (i >= total) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  y[(int64_t)r * ldy + c] = (D)x[(int64_t)r * ldx + c];
}

hipError_t launch_copy_d2s(lair_context *ctx, int rows, int cols, const double *x, int64_t ldx,
                           float *y, int64_t ldy) {
  const int64_t total = (int64_t)rows * cols;
  if (total <= 0) return hipSuccess;
  copy2d_kernel<double, float><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(
      rows, cols, x, ldx, y, ldy);
  return hipGetLastError();
}
hipError_t launch_copy_s2d(lair_context *ctx, int rows, int cols, const float *x, int64_t ldx,
                           double *y, int64_t ldy) {
  const int64_t total = (int64_t)rows * cols;
  if (total <= 0) return hipSuccess;
  copy2d_kernel<float, double><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(
      rows, cols, x, ldx, y, ldy);
  return hipGetLastError();
}
hipError_t launch_copy_d2d(lair_context *ctx, int rows, int cols, const double *x, int64_t ldx,
                           double *y, int64_t ldy) {
  const int64_t total = (int64_t)rows * cols;
  if (total <= 0) return hipSuccess;
  copy2d_kernel<double, double><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(
      rows, cols, x, ldx, y, ldy);
  return hipGetLastError();
}

__global__ void axpy_s2d_kernel(int rows, int cols, const float *d, int64_t ldd, double *x,
                                int64_t ldx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)rows * cols;
  if (i >= total) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  x[(int64_t)r * ldx + c] += (double)d[(int64_t)r * ldd + c];
}

hipError_t launch_axpy_s2d(lair_context *ctx, int rows, int cols, const float *d, int64_t ldd,
                           double *x, int64_t ldx) {
  const int64_t total = (int64_t)rows * cols;
  if (total <= 0) return hipSuccess;
  axpy_s2d_kernel<<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(rows, cols, d, ldd, x,
                                                                            ldx);
  return hipGetLastError();
}

// R = B - A X for up to 4 right-hand sides in one pass over A: one wave per
// row, FP64 accumulation.
template <int NR>
__global__ void residual_kernel(int n, const double *a, int64_t lda, const double *x, int64_t ldx,
                                const double *b, int64_t ldb, double *r, int64_t ldr) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= n) return;
  const double *arow = a + (int64_t)row * lda;
  double s[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) s[q] = 0.0;
  for (int c = lane; c < n; c += 64) {
    const double av = arow[c];
#pragma unroll
    for (int q = 0; q < NR; ++q) s[q] += av * x[(int64_t)c * ldx + q];
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s[q] += __shfl_xor(s[q], off);
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NR; ++q) r[(int64_t)row * ldr + q] = b[(int64_t)row * ldb + q] - s[q];
  }
}

hipError_t launch_residual(lair_context *ctx, int n, int nrhs, const double *a, int64_t lda,
                           const double *x, int64_t ldx, const double *b, int64_t ldb, double *r,
                           int64_t ldr) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  if (nrhs <= 4) {
    const int slot = prof_begin(ctx);
    const unsigned grid = (n + 3) / 4;
    for (int q0 = 0; q0 < nrhs; q0 += 4) {
      const int nr = nrhs - q0 < 4 ? nrhs - q0 : 4;
      switch (nr) {
        case 1:
          residual_kernel<1><<<grid, 256, 0, ctx->stream>>>(n, a, lda, x + q0, ldx, b + q0, ldb,
                                                             r + q0, ldr);
          break;
        case 2:
          residual_kernel<2><<<grid, 256, 0, ctx->stream>>>(n, a, lda, x + q0, ldx, b + q0, ldb,
                                                             r + q0, ldr);
          break;
        case 3:
          residual_kernel<3><<<grid, 256, 0, ctx->stream>>>(n, a, lda, x + q0, ldx, b + q0, ldb,
                                                             r + q0, ldr);
          break;
        default:
          residual_kernel<4><<<grid, 256, 0, ctx->stream>>>(n, a, lda, x + q0, ldx, b + q0, ldb,
                                                             r + q0, ldr);
      }
    }
    prof_end(ctx, slot, LAIR_KCLASS_RESIDUAL, 2.0 * n * (double)n * nrhs);
    return hipGetLastError();
  }
  // Many right-hand sides: R = B, then R -= A X on the FP64 MFMA path.
  hipError_t e = launch_copy_d2d(ctx, n, nrhs, b, ldb, r, ldr);
  if (e != hipSuccess) return e;
  return launch_gemm_minus<double>(ctx, kGemmF64, n, nrhs, n, a, lda, false, x, ldx, r, ldr);
}

__global__ void colmax_kernel(int rows, int cols, const double *x, int64_t ldx,
                              unsigned long long *bits) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)rows * cols;
  if (i >= total) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  atomic_max_nonneg(bits + c, fabs(x[(int64_t)r * ldx + c]));
}

hipError_t launch_colmax(lair_context *ctx, int rows, int cols, const double *x, int64_t ldx,
                         unsigned long long *bits) {
  const int64_t total = (int64_t)rows * cols;
  if (total <= 0) return hipSuccess;
  colmax_kernel<<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(rows, cols, x, ldx, bits);
  return hipGetLastError();
}

}  // namespace lair
