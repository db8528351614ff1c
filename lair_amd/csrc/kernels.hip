// HIP kernels of the gfx950 LU factor-and-solve path.
//
//  * panel_coop: exact partial pivoting of a column leaf, rows spread over a
//    few cooperating workgroups that synchronize once per column through
//    per-workgroup sequence flags (src/lapack/getrf.rs:55-117, src/blas/iamax.rs:6-21)
//  * laswp: row interchanges over column ranges (src/lapack/laswp.rs:11-40)
//  * getrs_exact: the reference's substitution order for small systems
//    (src/lapack/getrs.rs:23-36)
//  * gemm_minus: C -= A*B on MFMA (src/blas/gemm.rs:6-32 with alpha = -1)
//  * FP64 residual, norms and conversions for iterative refinement.
#include <climits>

#include <algorithm>

#include "device_util.h"
#include "lair_internal.h"

namespace lair {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));


// (value, row) ordering of iamax: larger |x| wins, ties go to the lower row.
// Candidates only exist for |x| > 0 (strict '>' against a running maximum that
// starts at zero), so NaN and zero never win; "none" is (0, INT_MAX).
__device__ __forceinline__ void better(double &bv, int &br, double ov, int orr) {
  if (ov > bv || (ov == bv && orr < br)) {
    bv = ov;
    br = orr;
  }
}

// Row segment I/O for the panel kernels: W consecutive elements of one row,
// with 16-byte vector accesses when every row segment is 16-byte aligned
// (vec, uniform) and whole; element accesses with zero fill otherwise.
template <typename T, int W>
__device__ __forceinline__ void load_row(T (&x)[W], const T *src, bool valid, int ncols, bool vec) {
  constexpr int E = 16 / sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(E)));
  if (valid && vec && ncols >= W) {
#pragma unroll
    for (int q = 0; q < W / E; ++q) {
      const vec_t u = reinterpret_cast<const vec_t *>(src)[q];
#pragma unroll
      for (int e = 0; e < E; ++e) x[q * E + e] = u[e];
    }
  } else {
#pragma unroll
    for (int w = 0; w < W; ++w) x[w] = (valid && w < ncols) ? src[w] : T(0);
  }
}

template <typename T, int W>
__device__ __forceinline__ void store_row(T *dst, const T (&x)[W], int ncols, bool vec) {
  constexpr int E = 16 / sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(E)));
  if (vec && ncols >= W) {
#pragma unroll
    for (int q = 0; q < W / E; ++q) {
      vec_t u;
#pragma unroll
      for (int e = 0; e < E; ++e) u[e] = x[q * E + e];
      reinterpret_cast<vec_t *>(dst)[q] = u;
    }
  } else {
#pragma unroll
    for (int w = 0; w < W; ++w)
      if (w < ncols) dst[w] = x[w];
  }
}

template <typename T>
__device__ __forceinline__ bool rows_aligned(const T *a, int64_t lda, int col0) {
  return ((lda * (int64_t)sizeof(T)) % 16 == 0) && ((reinterpret_cast<uintptr_t>(a + col0) & 15) == 0);
}

// ---------------------------------------------------------------- panel ---

// Words per exchange record: value (2), row (1), row data (W * words per element).
constexpr int kRecWords = 64;

template <typename T>
struct PanelArgs {
  T *a;
  int64_t lda;
  int row0, m, col0, wcols, npiv, P;
  int64_t *ipiv;
  int64_t *sing;
  unsigned long long *rec;   // [2][kMaxPanelWG][kRecWords] tagged words
  unsigned long long *drec;  // [2][kRecWords] tagged words of the diagonal row
  uint32_t seq;
  int32_t *err;
  int wait_limit;  // polls before a wait gives up (0: every wait gives up at once)
  int stride;  // launched slots per workgroup; only slot 0 of each group works
  int rest1;   // panel columns [col0 + wcols, rest1) receive the leaf's
               // interchanges and their U rows (L11^-1 A12) in this launch;
  int left0;   // columns [left0, col0) receive the interchanges only
};

// Factors columns [col0, col0+wcols) of rows [row0, m) in place.  Workgroup g
// owns rows [row0 + g*R, row0 + (g+1)*R) with R = NT*RPT; thread t keeps rows
// base + t + s*NT (s < RPT) of the leaf in registers.  Per column j:
//  1. argmax over the rows at or below the diagonal d = row0 + j: per thread,
//     per wave (DPP) and across the waves of the workgroup (LDS);
//  2. with P > 1 workgroups: every workgroup publishes one record (value,
//     row, row data) whose 64-bit words each carry the column's tag, so a
//     reader that sees every tag also has every datum -- one round trip;
//     all records are reduced in the same order everywhere;
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
  static constexpr int WPE = sizeof(T) / 4;  // 32-bit words per element
  static constexpr int L = 3 + W * WPE;      // record length in words
  static_assert(L <= kRecWords, "record too long");
  const PanelArgs<T> &pa;
  T (&v)[RPT][W];
  double (*s_wv)[NWAVE];      // [2][NWAVE] wave maxima, by column parity
  int (*s_wr)[NWAVE];         // [2][NWAVE] wave rows
  T (*s_wrow)[NWAVE][W];      // [2][NWAVE][W] wave winners' rows
  T (*s_prow)[W];             // [2][W] pivot row
  T (*s_drow)[W];             // [2][W] diagonal row before the interchange
  int *s_win;                 // [2]
  int *s_piv;                 // [W] the leaf's pivot rows (workgroup 0)
  uint32_t *s_rec;            // [P][L] record words, then [W*WPE] diagonal words
  int *s_bail;
  int g, t, lane, wave, base;

  template <int J>
  __device__ __forceinline__ bool exchange(double bv, int br, int qw) {
    constexpr int par = J & 1;
    const uint32_t tg = pa.seq + (uint32_t)J + 1u;
    if (wave == 0 && lane < L) {
      uint32_t data;
      if (lane == 0)
        data = word_of<double>(bv, 0);
      else if (lane == 1)
        data = word_of<double>(bv, 1);
      else if (lane == 2)
        data = (uint32_t)br;
      else
        data = br == INT_MAX ? 0u : word_of<T>(s_wrow[par][qw][(lane - 3) / WPE], (lane - 3) % WPE);
      st_agent(pa.rec + ((size_t)par * pa.P + g) * kRecWords + lane, tagged(tg, data));
    }
    if (g == 0 && wave == NWAVE - 1 && lane < W * WPE)
      st_agent(pa.drec + (size_t)par * kRecWords + lane,
               tagged(tg, word_of<T>(s_drow[par][lane / WPE], lane % WPE)));
    if (wave == 0) {
      const int total = pa.P * L;
      int spins = 0;
      bool bail = false;
      for (;;) {
        if (spins >= pa.wait_limit) {
          bail = true;
          break;
        }
        bool ok = true;
        for (int idx = lane; idx < total; idx += 64) {
          const int q = idx / L, w = idx % L;
          const unsigned long long x = ld_agent(pa.rec + ((size_t)par * pa.P + q) * kRecWords + w);
          ok = ok && (uint32_t)(x >> 32) == tg;
          s_rec[idx] = (uint32_t)x;
        }
        if (lane < W * WPE) {
          const unsigned long long x = ld_agent(pa.drec + (size_t)par * kRecWords + lane);
          ok = ok && (uint32_t)(x >> 32) == tg;
          s_rec[total + lane] = (uint32_t)x;
        }
        if (__all(ok)) break;
        if ((++spins & 255) == 0 && ld_agent(pa.err) != 0) {
          bail = true;
          break;
        }
      }
      double cv = 0.0;
      int cr = INT_MAX;
      if (!bail && lane < pa.P) {
        const uint32_t *rw = s_rec + lane * L;
        cv = from_words<double>(rw);
        cr = (int)rw[2];
      }
      double wv;
      int wr;
      wave_argmax(cv, cr, wv, wr);
      if (wr != INT_MAX && lane < W) {
        const int q = (wr - pa.row0) / R;
        s_prow[par][lane] = from_words<T>(s_rec + q * L + 3 + lane * WPE);
      }
      if (lane < W) s_drow[par][lane] = from_words<T>(s_rec + total + lane * WPE);
      if (lane == 0) {
        s_win[par] = wr;
        *s_bail = bail ? 1 : 0;
        if (bail) atomic_or_agent(pa.err, kErrWait);
      }
    }
    __syncthreads();
    return *s_bail == 0;
  }

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
    double wv;
    int wr;
    wave_argmax(bv, br, wv, wr);
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      const int r = base + t + s * NT;
      if (r == wr) {
#pragma unroll
        for (int w = 0; w < W; ++w) s_wrow[par][wave][w] = v[s][w];
      }
      if (r == d) {
#pragma unroll
        for (int w = 0; w < W; ++w) s_drow[par][w] = v[s][w];
      }
    }
    if (lane == 0) {
      s_wv[par][wave] = wv;
      s_wr[par][wave] = wr;
    }
    __syncthreads();
    bv = s_wv[par][0];
    br = s_wr[par][0];
    int qw = 0;
#pragma unroll
    for (int q = 1; q < NWAVE; ++q) {
      const double ov = s_wv[par][q];
      const int orr = s_wr[par][q];
      if (ov > bv || (ov == bv && orr < br)) {
        bv = ov;
        br = orr;
        qw = q;
      }
    }
    const T *prow;
    int win;
    if (pa.P == 1) {
      win = br;
      prow = s_wrow[par][qw];
    } else {
      if (!exchange<J>(bv, br, qw)) {
        // Given up: leave in-range (identity) pivots for the columns not
        // decided, so the interchanges queued after this launch stay in
        // bounds.  Thread 0 of workgroup 0 writes every pivot of the leaf.
        if (g == 0 && t == 0)
          for (int jj = J; jj < pa.npiv; ++jj) pa.ipiv[pa.row0 + jj] = pa.row0 + jj;
        return false;
      }
      win = s_win[par];
      prow = s_prow[par];
    }
    const bool singular = (win == INT_MAX);
    const int pr = singular ? d : win;
    if (pr != d) {
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int r = base + t + s * NT;
        if (r == d) {
#pragma unroll
          for (int w = 0; w < W; ++w) v[s][w] = prow[w];
        } else if (r == pr) {
#pragma unroll
          for (int w = 0; w < W; ++w) v[s][w] = s_drow[par][w];
        }
      }
    }
    if (g == 0 && t == 0) {
      s_piv[J] = pr;
      pa.ipiv[d] = pr;
      if (singular) *pa.sing = d;
    }
    if (!singular) {
      const T recip = T(1) / prow[J];
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int r = base + t + s * NT;
        if (r > d && r < pa.m) {
          const T l = v[s][J] * recip;
          v[s][J] = l;
#pragma unroll
          for (int k = J + 1; k < W; ++k) {
            const T e = l * prow[k];
            v[s][k] = v[s][k] - e;
          }
        }
      }
    }
    // No trailing barrier: the next column uses the other parity's buffers,
    // and the one after that is written only after the next column's first
    // barrier, which every thread reaches after finishing this update.
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
  constexpr int L = PanelLeaf<T, W, RPT, NT>::L;
  if (blockIdx.x % pa.stride != 0) return;  // idle slot (placement only)
  const int g = blockIdx.x / pa.stride, t = threadIdx.x;
  const int base = pa.row0 + g * R;

  T v[RPT][W];
  const bool vec = rows_aligned(pa.a, pa.lda, pa.col0);
#pragma unroll
  for (int s = 0; s < RPT; ++s) {
    const int r = base + t + s * NT;
    load_row<T, W>(v[s], pa.a + (int64_t)r * pa.lda + pa.col0, r < pa.m, pa.wcols, vec);
  }

  __shared__ double s_wv[2][NWAVE];
  __shared__ int s_wr[2][NWAVE];
  __shared__ T s_wrow[2][NWAVE][W];
  __shared__ T s_prow[2][W];
  __shared__ T s_drow[2][W];
  __shared__ int s_win[2];
  __shared__ int s_piv[W];
  __shared__ uint32_t s_rec[kMaxPanelWG * L + W * (sizeof(T) / 4)];
  __shared__ int s_bail;

  PanelLeaf<T, W, RPT, NT> leaf{pa, v, s_wv, s_wr, s_wrow, s_prow, s_drow, s_win, s_piv, s_rec, &s_bail,
                                g, t, t & 63, t >> 6, base};
  if (!ColumnLoop<0, W>::run(leaf)) return;

#pragma unroll
  for (int s = 0; s < RPT; ++s) {
    const int r = base + t + s * NT;
    if (r < pa.m) store_row<T, W>(pa.a + (int64_t)r * pa.lda + pa.col0, v[s], pa.wcols, vec);
  }
  if (g != 0 || (pa.rest1 <= pa.col0 + pa.wcols && pa.left0 >= pa.col0)) return;

  // Rest of the panel (workgroup 0 only; no other workgroup touches these
  // columns during the launch).  The leaf's npiv interchanges compose into a
  // row map (rows[e] receives the old contents of src[e]); the first npiv
  // entries are the leaf's diagonal rows, which become U12 = L11^-1 A12
  // (getrf.rs:270-278: laswp, then trsm with the unit-lower L11).
  __shared__ int s_mrow[2 * W], s_msrc[2 * W], s_mn;
  __shared__ T s_l11[W][W];
  __shared__ T s_buf[2 * W][64];
  const int npiv = pa.npiv;
  if (t < npiv) {
#pragma unroll
    for (int w = 0; w < W; ++w) s_l11[t][w] = v[0][w];  // row col0 + t lives in thread t
  }
  if (t == 0) {
    int nx = npiv;
    for (int e = 0; e < npiv; ++e) {
      s_mrow[e] = pa.col0 + e;
      s_msrc[e] = pa.col0 + e;
    }
    for (int e = 0; e < npiv; ++e) {
      const int p = s_piv[e];
      const int k = pa.col0 + e;
      if (p == k) continue;
      int ep = -1;
      if (p < pa.col0 + npiv) {
        ep = p - pa.col0;
      } else {
        for (int q = npiv; q < nx; ++q)
          if (s_mrow[q] == p) ep = q;
        if (ep < 0) {
          ep = nx++;
          s_mrow[ep] = p;
          s_msrc[ep] = p;
        }
      }
      const int tmp = s_msrc[e];
      s_msrc[e] = s_msrc[ep];
      s_msrc[ep] = tmp;
    }
    s_mn = nx;
  }
  __syncthreads();
  const int nent = s_mn;
  const int c0 = pa.col0 + pa.wcols;
  const int ql = t & 63, eg = t >> 6;
  // Columns [left0, col0) (interchanges only), then [c0, rest1).
  const int nleft = pa.left0 < pa.col0 ? pa.col0 - pa.left0 : 0;
  const int ntot = nleft + (pa.rest1 > c0 ? pa.rest1 - c0 : 0);
  for (int cb = 0; cb < ntot; cb += 64) {
    const int qi = cb + ql;
    const bool act = qi < ntot;
    const int q = qi < nleft ? pa.left0 + qi : c0 + (qi - nleft);
    const bool solve = qi >= nleft;
    for (int e = eg; e < nent; e += NWAVE)
      if (act) s_buf[e][ql] = pa.a[(int64_t)s_msrc[e] * pa.lda + q];
    __syncthreads();
    if (eg == 0 && act && solve) {
      for (int i = 1; i < npiv; ++i) {
        T acc = s_buf[i][ql];
        for (int p = 0; p < i; ++p) acc -= s_l11[i][p] * s_buf[p][ql];
        s_buf[i][ql] = acc;
      }
    }
    __syncthreads();
    for (int e = eg; e < nent; e += NWAVE)
      if (act && ((solve && e < npiv) || s_msrc[e] != s_mrow[e]))
        pa.a[(int64_t)s_mrow[e] * pa.lda + q] = s_buf[e][ql];
    __syncthreads();
  }
}

template <typename T>
hipError_t launch_panel(lair_context *ctx, T *a, int64_t lda, int row0, int m, int col0, int wcols,
                        int64_t *ipiv, int rest1, int left0) {
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
  pa.rec = ctx->rec;
  pa.drec = ctx->drec;
  pa.seq = ctx->seq;
  pa.err = ctx->err;
  pa.wait_limit = ctx->wait_limit;
  pa.stride = ctx->panel_xcd_stride;
  pa.rest1 = rest1;
  pa.left0 = left0;
  if ((rest1 > col0 + wcols || left0 < col0) && row0 != col0) return hipErrorInvalidValue;
  // Rows per workgroup: 256 threads x {1, 2, 4} rows, or 512 threads x 4 rows
  // for tall panels (fewer workgroups take part in each column's exchange).
  int nt, rpt;
  if (rows <= 256) {
    nt = 256;
    rpt = 1;
  } else if (rows <= 512) {
    nt = 256;
    rpt = 2;
  } else if (rows <= 1024) {
    nt = 256;
    rpt = 4;
  } else {
    nt = 512;
    rpt = 4;
  }
  const int R = nt * rpt;
  pa.P = (rows + R - 1) / R;
  if (pa.P > kMaxPanelWG) return hipErrorInvalidValue;
  const int slot = prof_begin(ctx, LAIR_KCLASS_PANEL);
  const unsigned grid = (unsigned)(pa.P * pa.stride);
  if (nt == 512)
    panel_coop_kernel<T, kLeafW, 4, 512><<<grid, 512, 0, ctx->stream>>>(pa);
  else if (rpt == 1)
    panel_coop_kernel<T, kLeafW, 1, 256><<<grid, 256, 0, ctx->stream>>>(pa);
  else if (rpt == 2)
    panel_coop_kernel<T, kLeafW, 2, 256><<<grid, 256, 0, ctx->stream>>>(pa);
  else
    panel_coop_kernel<T, kLeafW, 4, 256><<<grid, 256, 0, ctx->stream>>>(pa);
  const double fl = (double)pa.npiv * (double)rows * (double)wcols;
  prof_end(ctx, slot, LAIR_KCLASS_PANEL, fl, 2.0 * sizeof(T) * rows * wcols);
  ctx->seq += (uint32_t)kLeafW;
  return hipGetLastError();
}

template hipError_t launch_panel<double>(lair_context *, double *, int64_t, int, int, int, int,
                                         int64_t *, int, int);
template hipError_t launch_panel<float>(lair_context *, float *, int64_t, int, int, int, int,
                                        int64_t *, int, int);

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
  const int slot = prof_begin(ctx, LAIR_KCLASS_LASWP);
  laswp_kernel<T><<<(total + 255) / 256, 256, 0, ctx->stream>>>(a, lda, c0, n1, s1, total, ipiv,
                                                                 k0, k1, rev ? 1 : 0);
  prof_end(ctx, slot, LAIR_KCLASS_LASWP, 0.0, 4.0 * sizeof(T) * (double)(k1 - k0) * total);
  return hipGetLastError();
}

template hipError_t launch_laswp<double>(lair_context *, double *, int64_t, int, int, int, int,
                                         const int64_t *, int, int, bool);
template hipError_t launch_laswp<float>(lair_context *, float *, int64_t, int, int, int, int,
                                        const int64_t *, int, int, bool);

// Small systems (n <= kExactOrderMax): getrs in the reference's exact
// operation order (getrs.rs:23-36): forward interchanges, x[i] -= a[i,k]*x[k]
// for k = 0..i-1 in order, then x[i] -= a[i,k]*x[k] for k = i+1..n-1 in order
// followed by x[i] /= a[i,i].  One thread per right-hand side, x in LDS.
template <typename T>
__global__ void __launch_bounds__(64)
    getrs_exact_kernel(int n, int nrhs, const T *lu, int64_t lda, const int64_t *ipiv, T *b,
                       int64_t ldb, int32_t *status) {
#pragma clang fp contract(off)
  extern __shared__ unsigned char smem_raw[];
  T *xs = reinterpret_cast<T *>(smem_raw);  // [n][64]
  const int t = threadIdx.x;
  const int c = blockIdx.x * 64 + t;
  if (c >= nrhs) return;
  for (int i = 0; i < n; ++i) xs[i * 64 + t] = b[(int64_t)i * ldb + c];
  for (int i = 0; i < n; ++i) {
    const int64_t p64 = ipiv[i];
    if (p64 < 0 || p64 >= n) {  // outside [0, n): skipped and reported
      if (t == 0) atomic_or_agent(status, kErrPivot);
      continue;
    }
    const int p = (int)p64;
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
  const int slot = prof_begin(ctx, LAIR_KCLASS_GETRS);
  getrs_exact_kernel<T><<<(nrhs + 63) / 64, 64, shm, ctx->stream>>>(n, nrhs, lu, lda, ipiv, b, ldb,
                                                                     ctx->err);
  prof_end(ctx, slot, LAIR_KCLASS_GETRS, 2.0 * n * n * nrhs);
  return hipGetLastError();
}

template hipError_t launch_getrs_exact<double>(lair_context *, int, int, const double *, int64_t,
                                               const int64_t *, double *, int64_t);
template hipError_t launch_getrs_exact<float>(lair_context *, int, int, const float *, int64_t,
                                              const int64_t *, float *, int64_t);

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
                                     int64_t ldb, double *C, int64_t ldc, int kclass) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  (void)mode;
  if (M >= 256 && N >= 32 && K >= 32) {
    const int slot = prof_begin(ctx, kclass);
    hipError_t e = launch_dgemm_tile(ctx, M, N, K, A, lda, transA, B, ldb, C, ldc);
    prof_end(ctx, slot, kclass, 2.0 * M * N * K,
             (double)sizeof(double) * (2.0 * M * N + (double)M * K + (double)K * N));
    return e;
  }
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  const int slot = prof_begin(ctx, kclass);
  if (transA)
    gemm_f64_kernel<true><<<grid, 256, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C, ldc);
  else
    gemm_f64_kernel<false><<<grid, 256, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C, ldc);
  prof_end(ctx, slot, kclass, 2.0 * M * N * K,
           (double)sizeof(double) * (2.0 * M * N + (double)M * K + (double)K * N));
  return hipGetLastError();
}

template <>
hipError_t launch_gemm_minus<float>(lair_context *ctx, int mode, int M, int N, int K,
                                    const float *A, int64_t lda, bool transA, const float *B,
                                    int64_t ldb, float *C, int64_t ldc, int kclass) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  if (M >= 256 && N >= 32 && K >= 32 && !(transA && mode != kGemmF32)) {
    const int slot = prof_begin(ctx, kclass);
    hipError_t e;
    if (mode == kGemmF32)
      e = launch_sgemm_tile(ctx, M, N, K, A, lda, transA, B, ldb, C, ldc);
    else
      e = launch_bgemm_tile(ctx, mode == kGemmBF16x6 ? 3 : 1, M, N, K, A, lda, B, ldb, C, ldc);
    prof_end(ctx, slot, kclass, 2.0 * M * N * K,
             (double)sizeof(float) * (2.0 * M * N + (double)M * K + (double)K * N));
    return e;
  }
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  const int slot = prof_begin(ctx, kclass);
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
  prof_end(ctx, slot, kclass, 2.0 * M * N * K,
           (double)sizeof(float) * (2.0 * M * N + (double)M * K + (double)K * N));
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

// One wave: the words and the status word to host memory, every lane's
// stores completed and written back (system-scope release fence, which waits
// for the whole wave's stores), then the sequence number.
__global__ void __launch_bounds__(64) publish_words_kernel(const unsigned long long *src, int nw,
                                                           const int32_t *err, unsigned long long *dst,
                                                           unsigned long long seq) {
  const int l = threadIdx.x;
  for (int i = l; i < nw; i += 64) dst[2 + i] = src[i];
  if (l == 0) dst[1] = (unsigned long long)(uint32_t)*err;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
  if (l == 0) __hip_atomic_store(dst, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

hipError_t launch_publish_words(lair_context *ctx, const unsigned long long *src, int nw,
                                unsigned long long seq) {
  if (!ctx->pub_dev || nw < 0 || nw > 4096) return hipErrorInvalidValue;
  publish_words_kernel<<<1, 64, 0, ctx->stream>>>(src, nw, ctx->err, ctx->pub_dev, seq);
  return hipGetLastError();
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *bits, double v) {
  // Non-negative doubles order like their bit patterns; NaN maps above +inf.
  atomicMax(bits, (unsigned long long)__double_as_longlong(v));
}

// One wave per row: y = (float) a, and ||row||_1 folded into the norm.
// When both rows are 16-byte aligned (vec), a lane takes 4 consecutive
// elements per step (two 16-byte loads, one 16-byte store), two steps in
// flight; the tail and unaligned rows go element by element.
__global__ void convert_d2s_norm_kernel(int n, int ncols, const double *a, int64_t lda, float *a32,
                                        int64_t lda32, unsigned long long *norm_bits, int vec) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= n) return;
  const double *src = a + (int64_t)row * lda;
  float *dst = a32 ? a32 + (int64_t)row * lda32 : nullptr;
  double s = 0.0;
  int c0 = 0;
  if (vec) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int nv = ncols / 512 * 512;  // whole steps of 2 x 64 lanes x 4 elements
    for (int c = 4 * lane; c < nv; c += 512) {
      d2 x[4];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        x[2 * u] = *reinterpret_cast<const d2 *>(src + c + 256 * u);
        x[2 * u + 1] = *reinterpret_cast<const d2 *>(src + c + 256 * u + 2);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        s += fabs(x[2 * u].x);
        s += fabs(x[2 * u].y);
        s += fabs(x[2 * u + 1].x);
        s += fabs(x[2 * u + 1].y);
        if (dst) {
          f4 y;
          y.x = (float)x[2 * u].x;
          y.y = (float)x[2 * u].y;
          y.z = (float)x[2 * u + 1].x;
          y.w = (float)x[2 * u + 1].y;
          *reinterpret_cast<f4 *>(dst + c + 256 * u) = y;
        }
      }
    }
    c0 = nv;
  }
  for (int c = c0 + lane; c < ncols; c += 64) {
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
  const int slot = prof_begin(ctx, LAIR_KCLASS_CONVERT);
  const bool vec = reinterpret_cast<uintptr_t>(a) % 16 == 0 && lda % 2 == 0 &&
                   (!a32 || (reinterpret_cast<uintptr_t>(a32) % 16 == 0 && lda32 % 4 == 0));
  convert_d2s_norm_kernel<<<(n + 3) / 4, 256, 0, ctx->stream>>>(n, ncols, a, lda, a32, lda32,
                                                                norm_bits, vec ? 1 : 0);
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

// y (cols x rows) = x^T (x is rows x cols), 32 x 32 tiles through LDS.
// P (P[perm[j]][j] = 1, lu.rs:28-39), L (unit lower, lu.rs:42-57), U (upper,
// lu.rs:60-72) or P*L (row perm[i] holds L's row i, lu.rs:107-153), one
// output element per thread.
__global__ void lu_extract_kernel(int which, int m, int n, const double *lu, int64_t lda, const int *perm,
                                  double *out, int64_t ldo, int rows, int cols) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)rows * cols) return;
  const int i = (int)(idx / cols), j = (int)(idx % cols);
  const double l = j < i ? lu[(int64_t)i * lda + j] : (j == i ? 1.0 : 0.0);
  if (which == LAIR_LU_P) {
    out[(int64_t)i * ldo + j] = perm[j] == i ? 1.0 : 0.0;
  } else if (which == LAIR_LU_L) {
    out[(int64_t)i * ldo + j] = l;
  } else if (which == LAIR_LU_U) {
    out[(int64_t)i * ldo + j] = j >= i ? lu[(int64_t)i * lda + j] : 0.0;
  } else {
    out[(int64_t)perm[i] * ldo + j] = l;
  }
}

hipError_t launch_lu_extract(lair_context *ctx, int which, int m, int n, const double *lu, int64_t lda,
                             const int *perm, double *out, int64_t ldo) {
  const int k = m < n ? m : n;
  const int rows = which == LAIR_LU_U ? k : m;
  const int cols = which == LAIR_LU_P ? m : (which == LAIR_LU_U ? n : k);
  const int64_t total = (int64_t)rows * cols;
  if (total <= 0) return hipSuccess;
  lu_extract_kernel<<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(which, m, n, lu, lda, perm, out,
                                                                               ldo, rows, cols);
  return hipGetLastError();
}

__global__ void transpose_d_kernel(int rows, int cols, const double *x, int64_t ldx, double *y,
                                   int64_t ldy) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int gr = by + r, gc = bx + tx;
    tile[r][tx] = (gr < rows && gc < cols) ? x[(int64_t)gr * ldx + gc] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int gc = bx + r, gr = by + tx;  // y[gc][gr] = x[gr][gc]
    if (gc < cols && gr < rows) y[(int64_t)gc * ldy + gr] = tile[tx][r];
  }
}

hipError_t launch_transpose_d(lair_context *ctx, int rows, int cols, const double *x, int64_t ldx,
                              double *y, int64_t ldy) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  dim3 grid((cols + 31) / 32, (rows + 31) / 32);
  transpose_d_kernel<<<grid, 256, 0, ctx->stream>>>(rows, cols, x, ldx, y, ldy);
  return hipGetLastError();
}

__global__ void axpy_s2d_kernel(int rows, int cols, const float *d, int64_t ldd, double *x,
                                int64_t ldx, unsigned long long *zero, int nz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)rows * cols;
  if (zero && i < nz) zero[i] = 0ull;
  if (i >= total) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  x[(int64_t)r * ldx + c] += (double)d[(int64_t)r * ldd + c];
}

hipError_t launch_axpy_s2d(lair_context *ctx, int rows, int cols, const float *d, int64_t ldd,
                           double *x, int64_t ldx, unsigned long long *zero, int nz) {
  const int64_t total = std::max<int64_t>((int64_t)rows * cols, nz);
  if (total <= 0) return hipSuccess;
  axpy_s2d_kernel<<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(rows, cols, d, ldd, x,
                                                                            ldx, zero, nz);
  return hipGetLastError();
}

// R = B - A X for up to 4 right-hand sides in one pass over A: one wave per
// row, FP64 accumulation.  With rbits / xbits, the column maxima of |R| and
// |X| are folded in as ordered bit patterns (NaN above +inf; the words start
// at zero): each row's |r| by one atomic, |X| by the first wave, which reads
// all of X anyway.
template <int NR>
__global__ void __launch_bounds__(256)
    residual_kernel(int n, const double *a, int64_t lda, const double *x, int64_t ldx, const double *b,
                    int64_t ldb, double *r, int64_t ldr, unsigned long long *rbits, unsigned long long *xbits) {
  __shared__ unsigned long long s_rm[4][NR];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = blockIdx.x * 4 + wave;
  const bool has = row < n;
  const bool xmax = xbits && row == 0;
  double s[NR];
  unsigned long long xm[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    s[q] = 0.0;
    xm[q] = 0ull;
  }
  if (has) {
    const double *arow = a + (int64_t)row * lda;
    if (xmax) {
      for (int c = lane; c < n; c += 64) {
        const double av = arow[c];
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          const double xv = x[(int64_t)c * ldx + q];
          s[q] += av * xv;
          const unsigned long long xb = (unsigned long long)__double_as_longlong(fabs(xv));
          xm[q] = xb > xm[q] ? xb : xm[q];
        }
      }
#pragma unroll
      for (int q = 0; q < NR; ++q) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          const unsigned long long o = __shfl_xor(xm[q], off);
          xm[q] = o > xm[q] ? o : xm[q];
        }
      }
    } else {
      for (int c = lane; c < n; c += 64) {
        const double av = arow[c];
#pragma unroll
        for (int q = 0; q < NR; ++q) s[q] += av * x[(int64_t)c * ldx + q];
      }
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) s[q] += __shfl_xor(s[q], off);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      unsigned long long m = 0ull;  // a missing row adds nothing to the maximum
      if (has) {
        const double rv = b[(int64_t)row * ldb + q] - s[q];
        r[(int64_t)row * ldr + q] = rv;
        m = (unsigned long long)__double_as_longlong(fabs(rv));
      }
      if (rbits) s_rm[wave][q] = m;
      if (xmax) atomicMax(xbits + q, xm[q]);
    }
  }
  if (rbits) {
    // One atomic per workgroup and column: the largest of its four rows.
    __syncthreads();
    if (threadIdx.x < NR) {
      const int q = threadIdx.x;
      unsigned long long m = s_rm[0][q];
#pragma unroll
      for (int w = 1; w < 4; ++w) m = s_rm[w][q] > m ? s_rm[w][q] : m;
      atomicMax(rbits + q, m);
    }
  }
}

// R = B + sum over the S slices W_y (in order y = 0 .. S-1).
__global__ void __launch_bounds__(256) splitk_sum_kernel(int n, int nrhs, int S, const double *__restrict__ W,
                                                         int64_t wstride, const double *__restrict__ b, int64_t ldb,
                                                         double *__restrict__ r, int64_t ldr) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n * nrhs) return;
  const int i = (int)(e / nrhs), c = (int)(e % nrhs);
  double v = b[(int64_t)i * ldb + c];
  for (int y = 0; y < S; ++y) v += W[(int64_t)y * wstride + e];
  r[(int64_t)i * ldr + c] = v;
}

hipError_t launch_residual(lair_context *ctx, int n, int nrhs, const double *a, int64_t lda,
                           const double *x, int64_t ldx, const double *b, int64_t ldb, double *r,
                           int64_t ldr, unsigned long long *rbits, unsigned long long *xbits) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  if (nrhs <= 4) {
    const int slot = prof_begin(ctx, LAIR_KCLASS_RESIDUAL);
    const unsigned grid = (n + 3) / 4;
    for (int q0 = 0; q0 < nrhs; q0 += 4) {
      const int nr = nrhs - q0 < 4 ? nrhs - q0 : 4;
      unsigned long long *rb = rbits ? rbits + q0 : nullptr, *xb = xbits ? xbits + q0 : nullptr;
      switch (nr) {
        case 1:
          residual_kernel<1><<<grid, 256, 0, ctx->stream>>>(n, a, lda, x + q0, ldx, b + q0, ldb,
                                                             r + q0, ldr, rb, xb);
          break;
        case 2:
          residual_kernel<2><<<grid, 256, 0, ctx->stream>>>(n, a, lda, x + q0, ldx, b + q0, ldb,
                                                             r + q0, ldr, rb, xb);
          break;
        case 3:
          residual_kernel<3><<<grid, 256, 0, ctx->stream>>>(n, a, lda, x + q0, ldx, b + q0, ldb,
                                                             r + q0, ldr, rb, xb);
          break;
        default:
          residual_kernel<4><<<grid, 256, 0, ctx->stream>>>(n, a, lda, x + q0, ldx, b + q0, ldb,
                                                             r + q0, ldr, rb, xb);
      }
    }
    prof_end(ctx, slot, LAIR_KCLASS_RESIDUAL, 2.0 * n * (double)n * nrhs);
    return hipGetLastError();
  }
  // Many right-hand sides: R = B - A X on the FP64 MFMA path, split over K
  // so that the few tiles of R still fill the chip: slices W_y = -A_y X_y,
  // then R = B + sum_y W_y in a fixed order.
  const int tiles = ((n + 127) / 128) * ((nrhs + 127) / 128);
  int S = (2 * 256 + tiles - 1) / tiles;
  S = std::max(1, std::min(S, std::min(16, (n + 255) / 256)));
  const int64_t wstride = (int64_t)n * nrhs;
  hipError_t e = ensure(ctx->ws_split, sizeof(double) * (size_t)wstride * S);
  if (e != hipSuccess) return e;
  double *W = static_cast<double *>(ctx->ws_split.ptr);
  const int slot = prof_begin(ctx, LAIR_KCLASS_RESIDUAL);
  e = launch_dgemm_splitk_store(ctx, n, nrhs, n, S, a, lda, x, ldx, W, nrhs, wstride);
  if (e != hipSuccess) return e;
  const int64_t total = (int64_t)n * nrhs;
  splitk_sum_kernel<<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(n, nrhs, S, W, wstride, b, ldb, r,
                                                                              ldr);
  prof_end(ctx, slot, LAIR_KCLASS_RESIDUAL, 2.0 * n * (double)n * nrhs,
           8.0 * ((double)n * n + (2.0 * S + 2.0) * n * nrhs));
  if (rbits) e = launch_colmax(ctx, n, nrhs, r, ldr, rbits);
  if (e == hipSuccess && xbits) e = launch_colmax(ctx, n, nrhs, x, ldx, xbits);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

// Column maxima of |x| as ordered bit patterns (NaN above +inf): a
// grid-stride pass per column, a workgroup reduction of the bit patterns in
// LDS, then one atomic per workgroup and column.
__global__ void __launch_bounds__(256) colmax_kernel(int rows, int cols, const double *x, int64_t ldx,
                                                     unsigned long long *bits) {
  __shared__ unsigned long long red[256];
  const int c = blockIdx.y;
  unsigned long long m = 0ull;
  for (int r = blockIdx.x * 256 + threadIdx.x; r < rows; r += gridDim.x * 256) {
    const unsigned long long v = (unsigned long long)__double_as_longlong(fabs(x[(int64_t)r * ldx + c]));
    m = v > m ? v : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const unsigned long long o = red[threadIdx.x + off];
      if (o > red[threadIdx.x]) red[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(bits + c, red[0]);
}

hipError_t launch_colmax(lair_context *ctx, int rows, int cols, const double *x, int64_t ldx,
                         unsigned long long *bits) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  const int gx = std::min((rows + 255) / 256, 128);
  colmax_kernel<<<dim3(gx, cols), 256, 0, ctx->stream>>>(rows, cols, x, ldx, bits);
  return hipGetLastError();
}


// ------------------------------------------------------ tournament panel ---
//
// FP32 leaf of the mixed-precision factorization with tournament pivoting
// (communication-avoiding LU, Grigori, Demmel and Xiang 2011): the 16 pivot
// rows of a leaf are chosen by a reduction tree of small partial-pivoting
// eliminations instead of one global argmax per column.
//  1. every wave eliminates its own rows (K per lane) with partial pivoting
//     and nominates the 16 rows it pivoted on;
//  2. wave 0 repeats the elimination on the original values of the
//     workgroup's 8 x 16 nominees and keeps 16;
//  3. with P > 1 workgroups, each publishes its 16 nominees (tagged words,
//     one round trip) and every workgroup eliminates the P x 16 finalists
//     the same way, which fixes the pivot rows, their order, L11 and U11;
//  4. the interchanges are expressed as a LAPACK pivot vector, the pivot rows
//     take the diagonal positions and every other row gets L = a U11^-1 from
//     its original values.
// The result is an LU factorization of the row-permuted leaf, as with
// partial pivoting, but the pivot order generally differs from it: this
// kernel serves only the mixed-precision solver, whose answer is refined in
// FP64 against the original matrix.

struct CaluArgs {
  float *a;
  int64_t lda;
  int row0, m, col0, wcols, npiv, P;
  int64_t *ipiv;
  int64_t *sing;
  unsigned long long *rec;   // [kCaluMaxWG][kCaluRec] tagged words
  unsigned long long *drec;  // [kCaluDiag] tagged words
  uint32_t tag;
  int32_t *err;
  int wait_limit;  // polls before a wait gives up (0: every wait gives up at once)
  int rest1, left0;
};

// Partial-pivoting elimination of the rows held by one wave (K per lane,
// ids[k] = row id or INT_MAX for an empty slot) over nsteps <= 16 columns.
// Writes the chosen row ids in pivot order to out_ids, their slots
// (lane + 64 k) to out_slot and, when out_rows is given, each pivot row as it
// stood when chosen (multipliers left of the diagonal, U from it on).
// Ties go to the lowest row id; a column without a nonzero candidate picks
// the lowest remaining row and eliminates nothing.
template <int K, int J>
__device__ __forceinline__ void wave_gepp_step(float (&v)[K][kCaluW], const int (&ids)[K], bool (&used)[K],
                                               int lane, float *s_p, int *out_ids, int *out_slot,
                                               float (*out_rows)[kCaluW]) {
  double bv = 0.0;
  int bid = INT_MAX;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!used[k]) {
      const double x = fabs((double)v[k][J]);
      if (bid == INT_MAX || x > bv || (x == bv && ids[k] < bid)) {
        bv = x;
        bid = ids[k];
      }
    }
  }
  double wv;
  int wid;
  wave_argmax(bv, bid, wv, wid);
  if (wid == INT_MAX) {
    if (lane == 0) {
      out_ids[J] = INT_MAX;
      out_slot[J] = -1;
    }
    return;
  }
  // The winning row (row ids are unique) copies itself to the wave's LDS row;
  // LDS operations of one wave complete in order, so the other lanes read it
  // back below.
  int wslot = -1;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool win = !used[k] && ids[k] == wid;
    if (win) {
#pragma unroll
      for (int w = 0; w < kCaluW; ++w) s_p[w] = v[k][w];
      used[k] = true;
      wslot = lane + 64 * k;
    }
  }
  float p[kCaluW];
#pragma unroll
  for (int w = 0; w < kCaluW; ++w) p[w] = s_p[w];
  if (wslot >= 0) {
    out_ids[J] = wid;
    out_slot[J] = wslot;
    if (out_rows != nullptr) {
#pragma unroll
      for (int w = 0; w < kCaluW; ++w) out_rows[J][w] = p[w];
    }
  }
  if (wv != 0.0) {
    const float rp = 1.0f / p[J];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!used[k]) {
        const float l = v[k][J] * rp;
        v[k][J] = l;
#pragma unroll
        for (int w = J + 1; w < kCaluW; ++w) v[k][w] = v[k][w] - l * p[w];
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
}

template <int K, int J = 0>
__device__ __forceinline__ void wave_gepp_steps(float (&v)[K][kCaluW], const int (&ids)[K],
                                                bool (&used)[K], int lane, int nsteps, float *s_p,
                                                int *out_ids, int *out_slot,
                                                float (*out_rows)[kCaluW]) {
  if constexpr (J < kCaluW) {
    if (J >= nsteps) return;
    wave_gepp_step<K, J>(v, ids, used, lane, s_p, out_ids, out_slot, out_rows);
    wave_gepp_steps<K, J + 1>(v, ids, used, lane, nsteps, s_p, out_ids, out_slot, out_rows);
  }
}

// Partial-pivoting elimination of the rows held by one wave (K per lane,
// ids[k] = row id or INT_MAX for an empty slot) over nsteps <= 16 columns.
// Writes the chosen row ids in pivot order to out_ids, their slots
// (lane + 64 k) to out_slot and, when out_rows is given, each pivot row as it
// stood when chosen (multipliers left of the diagonal, U from it on).
// Ties go to the lowest row id; a column without a nonzero candidate picks
// the lowest remaining row and eliminates nothing.
template <int K>
__device__ __forceinline__ void wave_gepp(float (&v)[K][kCaluW], const int (&ids)[K], int lane,
                                          int nsteps, float *s_p, int *out_ids, int *out_slot,
                                          float (*out_rows)[kCaluW]) {
  bool used[K];
#pragma unroll
  for (int k = 0; k < K; ++k) used[k] = ids[k] == INT_MAX;
  wave_gepp_steps<K>(v, ids, used, lane, nsteps, s_p, out_ids, out_slot, out_rows);
}

// Spins on tagged words [0, total) of src until each carries tag, storing
// the data halves to dst.  Thread t of nt takes words t, t + nt, ... (at most
// kGatherMax each, all loads in flight together).  Returns false (for this
// thread) when the bounded wait expires.
constexpr int kGatherMax = 9;
static_assert(kCaluMaxWG * kCaluRec <= kGatherMax * 512, "gather capacity");
__device__ __forceinline__ bool gather_tagged(const unsigned long long *src, int total, uint32_t tag,
                                              uint32_t *dst, int t, int nt, const int32_t *err,
                                              int limit) {
  bool got[kGatherMax];
#pragma unroll
  for (int q = 0; q < kGatherMax; ++q) got[q] = t + q * nt >= total;
  int spins = 0;
  for (;;) {
    if (spins >= limit) return false;
    bool all = true;
#pragma unroll
    for (int q = 0; q < kGatherMax; ++q) {
      if (!got[q]) {
        const int idx = t + q * nt;
        const unsigned long long x = ld_agent(src + idx);
        if ((uint32_t)(x >> 32) == tag) {
          dst[idx] = (uint32_t)x;
          got[q] = true;
        } else {
          all = false;
        }
      }
    }
    if (all) return true;
    if ((++spins & 255) == 0 && ld_agent(err) != 0) return false;
  }
}

template <int K>
__global__ void __launch_bounds__(512) calu_leaf_kernel(CaluArgs ca) {
  constexpr int NT = 512, NWAVE = NT / 64, W = kCaluW;
  constexpr int R = NT * K;
  const int g = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int base = ca.row0 + g * R;
  const int npiv = ca.npiv;

  __shared__ int s_nid[NWAVE][W], s_nslot[NWAVE][W];
  __shared__ float s_prow[NWAVE][W];               // each wave's pivot row
  __shared__ float s_nval[NWAVE * W][W];           // nominees' original values
  __shared__ int s_wid[W], s_wslot[W];             // workgroup nominees
  __shared__ uint32_t s_fin[kCaluMaxWG * kCaluRec];  // finalists (ids, values)
  __shared__ float s_diag[W][W];                   // original diagonal rows
  __shared__ int s_fid[W], s_fslot[W];             // final pivot rows, in order
  __shared__ float s_frow[W][W];                   // their L11 \ U11 rows
  __shared__ int s_piv[W];
  __shared__ int s_mrow[2 * W], s_msrc[2 * W], s_mn;
  __shared__ int s_bail;

  const bool vec = rows_aligned(ca.a, ca.lda, ca.col0);
  // 1. Local elimination on a register copy of this thread's rows.
  {
    float v[K][W];
    int ids[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int r = base + t + k * NT;
      ids[k] = r < ca.m ? r : INT_MAX;
      load_row<float, W>(v[k], ca.a + (int64_t)r * ca.lda + ca.col0, r < ca.m, ca.wcols, vec);
    }
    wave_gepp<K>(v, ids, lane, npiv, s_prow[wave], s_nid[wave], s_nslot[wave], nullptr);
  }
  if (lane == 0)
    for (int j = npiv; j < W; ++j) s_nid[wave][j] = INT_MAX;
  __syncthreads();
  // Nominees' original values (still unmodified in memory).
  for (int i = lane; i < W * W; i += 64) {
    const int e = i / W, w = i % W;
    const int id = s_nid[wave][e];
    s_nval[wave * W + e][w] =
        (id != INT_MAX && w < ca.wcols) ? ca.a[(int64_t)id * ca.lda + ca.col0 + w] : 0.0f;
  }
  if (g == 0 && wave == NWAVE - 1) {
    for (int i = lane; i < W * W; i += 64) {
      const int e = i / W, w = i % W;
      const int r = ca.row0 + e;
      s_diag[e][w] = (r < ca.m && w < ca.wcols) ? ca.a[(int64_t)r * ca.lda + ca.col0 + w] : 0.0f;
    }
  }
  __syncthreads();
  // 2. Workgroup round: 8 x 16 nominees, 2 per lane.
  if (wave == 0) {
    float v[2][W];
    int ids[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int e = lane + 64 * k;
      ids[k] = s_nid[e / W][e % W];
#pragma unroll
      for (int w = 0; w < W; ++w) v[k][w] = s_nval[e][w];
    }
    if (ca.P == 1) {
      wave_gepp<2>(v, ids, lane, npiv, s_prow[0], s_fid, s_fslot, s_frow);
    } else {
      wave_gepp<2>(v, ids, lane, npiv, s_prow[0], s_wid, s_wslot, nullptr);
      if (lane == 0)
        for (int j = npiv; j < W; ++j) s_wid[j] = INT_MAX;
    }
  }
  __syncthreads();
  if (ca.P > 1) {
    // 3. Publish the workgroup's nominees (and, from workgroup 0, the
    // original diagonal rows); gather everyone's.
    if (wave == 0) {
      for (int i = lane; i < kCaluRec; i += 64) {
        uint32_t data;
        if (i < W) {
          data = (uint32_t)s_wid[i];
        } else {
          const int e = (i - W) / W, w = (i - W) % W;
          data = s_wid[e] == INT_MAX ? 0u : __float_as_uint(s_nval[s_wslot[e]][w]);
        }
        st_agent(ca.rec + (size_t)g * kCaluRec + i, tagged(ca.tag, data));
      }
    }
    if (g == 0 && wave == 1) {
      for (int i = lane; i < kCaluDiag; i += 64)
        st_agent(ca.drec + i, tagged(ca.tag, __float_as_uint(s_diag[i / W][i % W])));
    }
    if (t == 0) s_bail = 0;
    __syncthreads();
    bool ok = gather_tagged(ca.rec, ca.P * kCaluRec, ca.tag, s_fin, t, NT, ca.err, ca.wait_limit);
    if (ok && g != 0)
      ok = gather_tagged(ca.drec, kCaluDiag, ca.tag, reinterpret_cast<uint32_t *>(&s_diag[0][0]), t, NT,
                         ca.err, ca.wait_limit);
    if (!ok) {
      s_bail = 1;
      atomic_or_agent(ca.err, kErrWait);
    }
    __syncthreads();
    if (s_bail) {
      // Given up: identity pivots keep the interchanges queued after this
      // launch in bounds.
      if (g == 0 && t == 0)
        for (int i = 0; i < ca.npiv; ++i) ca.ipiv[ca.col0 + i] = ca.col0 + i;
      return;
    }
    // Final round on the P x 16 finalists.
    if (wave == 0) {
      const int nf = ca.P * W;
      if (nf <= 128) {
        float v[2][W];
        int ids[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          const int e = lane + 64 * k;
          const uint32_t *rw = s_fin + (e / W) * kCaluRec;
          ids[k] = e < nf ? (int)rw[e % W] : INT_MAX;
#pragma unroll
          for (int w = 0; w < W; ++w) v[k][w] = e < nf ? __uint_as_float(rw[W + (e % W) * W + w]) : 0.0f;
        }
        wave_gepp<2>(v, ids, lane, npiv, s_prow[0], s_fid, s_fslot, s_frow);
      } else {
        float v[4][W];
        int ids[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int e = lane + 64 * k;
          const uint32_t *rw = s_fin + (e / W) * kCaluRec;
          ids[k] = e < nf ? (int)rw[e % W] : INT_MAX;
#pragma unroll
          for (int w = 0; w < W; ++w) v[k][w] = e < nf ? __uint_as_float(rw[W + (e % W) * W + w]) : 0.0f;
        }
        wave_gepp<4>(v, ids, lane, npiv, s_prow[0], s_fid, s_fslot, s_frow);
      }
    }
    __syncthreads();
  }
  // 4. Pivot vector and row map: position col0 + i receives row s_fid[i]
  // (sequential interchanges, LAPACK convention).
  if (t == 0) {
    int nx = npiv;
    for (int e = 0; e < npiv; ++e) {
      s_mrow[e] = ca.col0 + e;
      s_msrc[e] = ca.col0 + e;
    }
    for (int i = 0; i < npiv; ++i) {
      const int want = s_fid[i];
      int q = -1;  // current position of row `want`
      for (int e = 0; e < nx; ++e)
        if (s_msrc[e] == want) q = e;
      if (q < 0) {  // untouched so far: it sits at its own position
        q = nx++;
        s_mrow[q] = want;
        s_msrc[q] = want;
      }
      s_piv[i] = s_mrow[q];
      const int tmp = s_msrc[i];
      s_msrc[i] = s_msrc[q];
      s_msrc[q] = tmp;
    }
    s_mn = nx;
    if (g == 0) {
      int64_t sg = -1;
      for (int i = 0; i < npiv; ++i) {
        ca.ipiv[ca.col0 + i] = s_piv[i];
        if (s_frow[i][i] == 0.0f) sg = ca.col0 + i;
      }
      if (sg >= 0) *ca.sing = sg;
    }
  }
  __syncthreads();
  // 5. Final values of this workgroup's rows: diagonal positions take the
  // pivot rows' L11 \ U11; every other position gets L = a U11^-1 from the
  // original values of the row it now holds.
  const int nent = s_mn;
#pragma unroll 1
  for (int k = 0; k < K; ++k) {
    const int r = base + t + k * NT;
    if (r >= ca.m) continue;
    float x[W];
    if (r < ca.col0 + npiv) {
#pragma unroll
      for (int w = 0; w < W; ++w) x[w] = s_frow[r - ca.col0][w];
    } else {
      int src = r;
      for (int e = npiv; e < nent; ++e)
        if (s_mrow[e] == r) src = s_msrc[e];
      if (src != r) {  // a displaced diagonal row
#pragma unroll
        for (int w = 0; w < W; ++w) x[w] = s_diag[src - ca.col0][w];
      } else {
        load_row<float, W>(x, ca.a + (int64_t)r * ca.lda + ca.col0, true, ca.wcols, vec);
      }
      // A zero pivot leaves its column unscaled and out of the later
      // updates, as getrf.rs:72-73 does for an all-zero column.
      bool live[W];
#pragma unroll
      for (int j = 0; j < W; ++j) {
        live[j] = false;
        if (j < npiv) {
          float acc = x[j];
#pragma unroll
          for (int i = 0; i < j; ++i)
            if (live[i]) acc = acc - x[i] * s_frow[i][j];
          const float ujj = s_frow[j][j];
          live[j] = ujj != 0.0f;
          x[j] = live[j] ? acc / ujj : acc;
        }
      }
      // Columns past npiv (short leaves): the U part of the update.
#pragma unroll
      for (int w = 0; w < W; ++w) {
        if (w >= npiv && w < ca.wcols) {
          float acc = x[w];
#pragma unroll
          for (int i = 0; i < W; ++i)
            if (i < npiv && live[i]) acc = acc - x[i] * s_frow[i][w];
          x[w] = acc;
        }
      }
    }
    store_row<float, W>(ca.a + (int64_t)r * ca.lda + ca.col0, x, ca.wcols, vec);
  }
  if (g != 0 || (ca.rest1 <= ca.col0 + ca.wcols && ca.left0 >= ca.col0)) return;
  // 6. Rest of the panel (workgroup 0; see panel_coop_kernel): interchanges
  // on [left0, col0) and [col0 + wcols, rest1), U12 = L11^-1 A12 on the latter.
  __shared__ float s_buf[2 * W][64];
  const int c0 = ca.col0 + ca.wcols;
  const int ql = t & 63, eg = t >> 6;
  const int nleft = ca.left0 < ca.col0 ? ca.col0 - ca.left0 : 0;
  const int ntot = nleft + (ca.rest1 > c0 ? ca.rest1 - c0 : 0);
  for (int cb = 0; cb < ntot; cb += 64) {
    const int qi = cb + ql;
    const bool act = qi < ntot;
    const int q = qi < nleft ? ca.left0 + qi : c0 + (qi - nleft);
    const bool solve = qi >= nleft;
    for (int e = eg; e < nent; e += NWAVE)
      if (act) s_buf[e][ql] = ca.a[(int64_t)s_msrc[e] * ca.lda + q];
    __syncthreads();
    if (eg == 0 && act && solve) {
      for (int i = 1; i < npiv; ++i) {
        float acc = s_buf[i][ql];
        for (int p = 0; p < i; ++p) acc -= s_frow[i][p] * s_buf[p][ql];
        s_buf[i][ql] = acc;
      }
    }
    __syncthreads();
    for (int e = eg; e < nent; e += NWAVE)
      if (act && ((solve && e < npiv) || s_msrc[e] != s_mrow[e]))
        ca.a[(int64_t)s_mrow[e] * ca.lda + q] = s_buf[e][ql];
    __syncthreads();
  }
}

hipError_t launch_panel_calu(lair_context *ctx, float *a, int64_t lda, int c, int m, int wcols,
                             int64_t *ipiv, int rest1, int left0) {
  const int rows = m - c;
  if (rows <= 0 || wcols <= 0) return hipSuccess;
  if (wcols > kCaluW) return hipErrorInvalidValue;
  CaluArgs ca;
  ca.a = a;
  ca.lda = lda;
  ca.row0 = c;
  ca.m = m;
  ca.col0 = c;
  ca.wcols = wcols;
  ca.npiv = wcols < rows ? wcols : rows;
  ca.ipiv = ipiv;
  ca.sing = ctx->sing;
  ca.rec = ctx->crec;
  ca.drec = ctx->crec + (size_t)kCaluMaxWG * kCaluRec;
  ca.tag = ctx->seq + 1u;
  ca.err = ctx->err;
  ca.wait_limit = ctx->wait_limit;
  ca.rest1 = rest1;
  ca.left0 = left0;
  constexpr int K = 4, R = 512 * K;
  ca.P = (rows + R - 1) / R;
  if (ca.P > kCaluMaxWG) return hipErrorInvalidValue;
  const int slot = prof_begin(ctx, LAIR_KCLASS_PANEL);
  calu_leaf_kernel<K><<<ca.P, 512, 0, ctx->stream>>>(ca);
  prof_end(ctx, slot, LAIR_KCLASS_PANEL, (double)ca.npiv * rows * wcols, 8.0 * rows * wcols);
  ctx->seq += (uint32_t)kLeafW;
  return hipGetLastError();
}

}  // namespace lair
