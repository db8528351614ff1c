// Row interchanges as gathers, blocked triangular solves and the dataflow
// triangular solve used by getrs.
//
//  * rowmap + rowgather: the interchanges ipiv[k0..k1) applied in forward
//    order (src/lapack/laswp.rs:21-39) compose to a permutation of at most
//    2*(k1-k0) rows; one wave composes it, then every column stripe moves the
//    affected rows in one read-all/write-all pass.  The same map updates the
//    running row permutation perm (perm[i] = original row now at row i,
//    src/decomposition/lu.rs:28-39).
//  * trsm_llu: B := L^-1 B for a unit lower L (src/blas/trsm.rs:6-22) with the
//    B stripe in LDS, 16-row diagonal sub-blocks solved per column and the rows
//    below updated in parallel.
//  * trtri_diag: inverses of the diagonal blocks of L (unit) and U.
//  * trsv_dataflow: block-row workgroups that wait for the solved values of
//    the blocks they depend on; those values are published as tagged 64-bit
//    words (tag = the solve's epoch), so a reader that sees its epoch in a
//    word also has the value, with no flags and no fences.  Tiles are fetched
//    before waiting, so the sequential chain only carries one round trip, one
//    tile product and one diagonal-inverse product per block (the
//    substitutions of src/lapack/getrs.rs:24-36, blocked).
#include <algorithm>
#include <climits>

#include "device_util.h"
#include "lair_internal.h"

namespace lair {

template <typename T>
__device__ __forceinline__ void st_sc1(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_sc1(const T *p) {
  return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


// Allows a kernel to use up to the full 160 KB of LDS per workgroup.
// The attribute call's status is consumed here so that it cannot surface as
// the error of a later launch.
template <typename F>
static void allow_large_lds(F *kernel) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  (void)hipGetLastError();
}

// ------------------------------------------------------------- rowmap ---

// rows[e] <- contents of src[e] (before the interchanges), e < *count, for
// the interchanges ipiv[k0 .. k1) applied in order (k swaps with ipiv[k]).
// The rows involved are the block's rows k0 .. k1 - 1 and the distinct pivot
// rows outside it.  Each thread takes one such row r and finds its source by
// tracing r back through the interchanges, last to first (a swap of k and p
// maps k to p and p to k), so the rows are resolved independently of each
// other.
__global__ void __launch_bounds__(256)
    rowmap_kernel(const int64_t *ipiv, int k0, int k1, int *rows, int *src, int *count) {
  __shared__ int s_piv[kMaxSwapRows / 2], s_out[kMaxSwapRows / 2];
  __shared__ int s_wcnt[4];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int kb = k1 - k0;
  for (int e = t; e < kb; e += 256) s_piv[e] = (int)ipiv[k0 + e];
  __syncthreads();
  // Pivot rows outside the block, each listed once (first occurrence).
  int p = 0;
  bool first = false;
  if (t < kb) {
    p = s_piv[t];
    first = p < k0 || p >= k1;
    for (int e = 0; e < t && first; ++e) first = s_piv[e] != p;
  }
  const unsigned long long m = __ballot(first);
  if (lane == 0) s_wcnt[wv] = __popcll(m);
  __syncthreads();
  int base = 0;
  for (int w = 0; w < wv; ++w) base += s_wcnt[w];
  if (first) s_out[base + __popcll(m & ((1ull << lane) - 1))] = p;
  const int nout = s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3];
  __syncthreads();
  const int total = kb + nout;
  for (int idx = t; idx < total; idx += 256) {
    const int r = idx < kb ? k0 + idx : s_out[idx - kb];
    int x = r;
#pragma unroll 8
    for (int e = kb - 1; e >= 0; --e) {
      const int pe = s_piv[e], ke = k0 + e;
      x = x == ke ? pe : (x == pe ? ke : x);
    }
    rows[idx] = r;
    src[idx] = x;
  }
  if (t == 0) *count = total;
}

// Columns [c0, c0+n1) and [s1, s1 + total-n1): 32 per workgroup.
template <typename T>
__global__ void __launch_bounds__(256)
    rowgather_kernel(T *a, int64_t lda, const int *rows, const int *src, const int *count, int c0,
                     int n1, int s1, int total) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *buf = reinterpret_cast<T *>(smem_raw);  // [nr][32]
  const int t = threadIdx.x, cl = t & 31, rg = t >> 5;
  const int i = blockIdx.x * 32 + cl;
  const bool act = i < total;
  const int c = act ? ((i < n1) ? c0 + i : s1 + (i - n1)) : 0;
  const int nr = *count;
  // Batches of 8 independent loads per thread.
  for (int e0 = rg; e0 < nr; e0 += 64) {
    T v[8];
    int sr[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = e0 + 8 * q;
      sr[q] = e < nr ? src[e] : -1;
      const bool mv = sr[q] >= 0 && sr[q] != rows[e];
      v[q] = (act && mv) ? a[(int64_t)sr[q] * lda + c] : T(0);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = e0 + 8 * q;
      if (e < nr) buf[e * 32 + cl] = v[q];
    }
  }
  __syncthreads();
  for (int e = rg; e < nr; e += 8) {
    const int s = src[e];
    if (act && s != rows[e]) a[(int64_t)rows[e] * lda + c] = buf[e * 32 + cl];
  }
}

__global__ void __launch_bounds__(256)
    permgather_kernel(int *perm, const int *rows, const int *src, const int *count) {
  __shared__ int tmp[kMaxSwapRows];
  const int nr = *count;
  // Entries with rows[e] == src[e] do nothing (and may share their row with
  // an entry that moves one).
  for (int e = threadIdx.x; e < nr; e += 256)
    if (src[e] != rows[e]) tmp[e] = perm[src[e]];
  __syncthreads();
  for (int e = threadIdx.x; e < nr; e += 256)
    if (src[e] != rows[e]) perm[rows[e]] = tmp[e];
}

// The gather of a prepared row map over columns [c0, c1) minus [s0, s1).
template <typename T>
static hipError_t rowgather(lair_context *ctx, T *a, int64_t lda, int c0, int c1, int s0, int s1,
                            const int *rows, const int *src, const int *count, int nmax, int *perm) {
  if (s0 < c0) s0 = c0;
  if (s1 > c1) s1 = c1;
  if (s1 < s0) s1 = s0;
  const int n1 = s0 - c0;
  const int total = n1 + (c1 - s1);
  if (total > 0) {
    const size_t shm = (size_t)nmax * 32 * sizeof(T);
    if (shm > 160 * 1024) return hipErrorInvalidValue;
    static bool once = (allow_large_lds(rowgather_kernel<T>), true);
    (void)once;
    rowgather_kernel<T><<<(total + 31) / 32, 256, shm, ctx->stream>>>(a, lda, rows, src, count, c0, n1, s1,
                                                                       total);
  }
  if (perm) permgather_kernel<<<1, 256, 0, ctx->stream>>>(perm, rows, src, count);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rowswap(lair_context *ctx, T *a, int64_t lda, int c0, int c1, int s0, int s1,
                          const int64_t *ipiv, int k0, int k1, int *perm) {
  if (k1 <= k0) return hipSuccess;
  if (k1 - k0 > kMaxSwapRows / 2) return hipErrorInvalidValue;
  const int slot = prof_begin(ctx, LAIR_KCLASS_LASWP);
  rowmap_kernel<<<1, 256, 0, ctx->stream>>>(ipiv, k0, k1, ctx->swap_rows, ctx->swap_src,
                                           ctx->swap_count);
  hipError_t e = rowgather<T>(ctx, a, lda, c0, c1, s0, s1, ctx->swap_rows, ctx->swap_src, ctx->swap_count,
                              2 * (k1 - k0), perm);
  prof_end(ctx, slot, LAIR_KCLASS_LASWP, 0.0, 4.0 * sizeof(T) * (double)(k1 - k0) * (c1 - c0));
  return e;
}

template <typename T>
hipError_t launch_rowswap_map(lair_context *ctx, T *a, int64_t lda, int c0, int c1, int s0, int s1,
                              const int *rows, const int *src, const int *count, int nmax, int *perm) {
  if (nmax <= 0) return hipSuccess;
  if (nmax > kMaxSwapRows) return hipErrorInvalidValue;
  const int slot = prof_begin(ctx, LAIR_KCLASS_LASWP);
  hipError_t e = rowgather<T>(ctx, a, lda, c0, c1, s0, s1, rows, src, count, nmax, perm);
  prof_end(ctx, slot, LAIR_KCLASS_LASWP, 0.0, 2.0 * sizeof(T) * (double)nmax * (c1 - c0));
  return e;
}

template hipError_t launch_rowswap_map<double>(lair_context *, double *, int64_t, int, int, int, int,
                                               const int *, const int *, const int *, int, int *);
template hipError_t launch_rowswap_map<float>(lair_context *, float *, int64_t, int, int, int, int,
                                              const int *, const int *, const int *, int, int *);
template hipError_t launch_rowswap_map<zcplx>(lair_context *, zcplx *, int64_t, int, int, int, int,
                                              const int *, const int *, const int *, int, int *);
template hipError_t launch_rowswap_map<ccplx>(lair_context *, ccplx *, int64_t, int, int, int, int,
                                              const int *, const int *, const int *, int, int *);

template hipError_t launch_rowswap<double>(lair_context *, double *, int64_t, int, int, int, int,
                                           const int64_t *, int, int, int *);
template hipError_t launch_rowswap<float>(lair_context *, float *, int64_t, int, int, int, int,
                                          const int64_t *, int, int, int *);
template hipError_t launch_rowswap<zcplx>(lair_context *, zcplx *, int64_t, int, int, int, int,
                                          const int64_t *, int, int, int *);
template hipError_t launch_rowswap<ccplx>(lair_context *, ccplx *, int64_t, int, int, int, int,
                                          const int64_t *, int, int, int *);

// perm from the full pivot vector (for factors supplied by the caller):
// one thread applies the n interchanges to an LDS index array.  When the
// pivots fit beside it (n <= kPermPivLdsMax) they are staged in LDS first, so
// the serial loop touches LDS only.  A pivot outside [0, n) (compared as the
// full int64) is skipped rather than followed out of bounds and reported
// through the status word (kErrPivot).
constexpr int kPermPivLdsMax = 20480;  // 2 n ints in 160 KB
__global__ void __launch_bounds__(256) perm_from_ipiv_kernel(const int64_t *ipiv, int n, int *perm,
                                                             int32_t *status) {
  extern __shared__ int sp[];  // [n] perm when it fits, then [n] pivots when both fit
  const bool lds = n <= kPermLdsMax;
  const bool lds_piv = n <= kPermPivLdsMax;
  int *p = lds ? sp : perm;
  int *piv = sp + n;
  bool bad = false;
  for (int i = threadIdx.x; i < n; i += 256) {
    p[i] = i;
    const int64_t q = ipiv[i];
    const bool ok = q >= 0 && q < n;
    bad = bad || !ok;
    if (lds_piv) piv[i] = ok ? (int)q : -1;
  }
  if (bad) atomic_or_agent(status, kErrPivot);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (lds_piv) {
      int q = piv[0];
      for (int i = 0; i < n; ++i) {
        const int qn = i + 1 < n ? piv[i + 1] : 0;  // next pivot, read ahead
        if (q != i && (unsigned)q < (unsigned)n) {
          const int tmp = p[i];
          p[i] = p[q];
          p[q] = tmp;
        }
        q = qn;
      }
    } else {
      for (int i = 0; i < n; ++i) {
        const int64_t q64 = ipiv[i];
        const int q = (q64 >= 0 && q64 < n) ? (int)q64 : -1;
        if (q != i && (unsigned)q < (unsigned)n) {
          const int tmp = p[i];
          p[i] = p[q];
          p[q] = tmp;
        }
      }
    }
  }
  __syncthreads();
  if (lds)
    for (int i = threadIdx.x; i < n; i += 256) perm[i] = p[i];
}

hipError_t launch_perm_from_ipiv(lair_context *ctx, const int64_t *ipiv, int n, int *perm) {
  if (n <= 0) return hipSuccess;
  const size_t shm = n <= kPermPivLdsMax ? 2 * (size_t)n * sizeof(int)
                                         : (n <= kPermLdsMax ? (size_t)n * sizeof(int) : 0);
  static bool once = (allow_large_lds(perm_from_ipiv_kernel), true);
  (void)once;
  perm_from_ipiv_kernel<<<1, 256, shm, ctx->stream>>>(ipiv, n, perm, ctx->err);
  return hipGetLastError();
}

__global__ void iota_kernel(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

hipError_t launch_iota(lair_context *ctx, int *p, int n) {
  if (n <= 0) return hipSuccess;
  iota_kernel<<<(n + 255) / 256, 256, 0, ctx->stream>>>(p, n);
  return hipGetLastError();
}

// y[i][q] = (D) x[perm[i]][q]  (gather) or y[perm[i]][q] = (D) x[i][q] (scatter);
// a null perm copies.
template <typename S, typename D>
__global__ void permute_rows_kernel(int rows, int cols, const int *perm, const S *x, int64_t ldx,
                                    D *y, int64_t ldy, int scatter, unsigned long long *zero, int nz) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (zero && idx < nz) zero[idx] = 0ull;
  if (idx >= (int64_t)rows * cols) return;
  const int i = (int)(idx / cols), q = (int)(idx % cols);
  const int pi = perm ? perm[i] : i;
  if (!scatter)
    y[(int64_t)i * ldy + q] = (D)x[(int64_t)pi * ldx + q];
  else
    y[(int64_t)pi * ldy + q] = (D)x[(int64_t)i * ldx + q];
}

template <typename S, typename D>
hipError_t launch_permute_rows(lair_context *ctx, int rows, int cols, const int *perm, const S *x,
                               int64_t ldx, D *y, int64_t ldy, bool scatter, unsigned long long *zero,
                               int nz) {
  const int64_t total = std::max<int64_t>((int64_t)rows * cols, zero ? nz : 0);
  if (total <= 0) return hipSuccess;
  permute_rows_kernel<S, D><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(
      rows, cols, perm, x, ldx, y, ldy, scatter ? 1 : 0, zero, nz);
  return hipGetLastError();
}

template hipError_t launch_permute_rows<double, double>(lair_context *, int, int, const int *,
                                                        const double *, int64_t, double *, int64_t,
                                                        bool, unsigned long long *, int);
template hipError_t launch_permute_rows<double, float>(lair_context *, int, int, const int *,
                                                       const double *, int64_t, float *, int64_t,
                                                       bool, unsigned long long *, int);
template hipError_t launch_permute_rows<float, float>(lair_context *, int, int, const int *,
                                                      const float *, int64_t, float *, int64_t,
                                                      bool, unsigned long long *, int);

// ------------------------------------------------------------ trsm_llu ---

template <typename T>
__global__ void __launch_bounds__(256)
    trsm_llu_kernel(int kb, int ncols, const T *__restrict__ L, int64_t ldl, T *__restrict__ B,
                    int64_t ldb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *xs = reinterpret_cast<T *>(smem_raw);  // [kb][64]
  const int t = threadIdx.x, col = t & 63, rg = t >> 6;
  const int c = blockIdx.x * 64 + col;
  const bool act = c < ncols;
  for (int i = rg; i < kb; i += 4) xs[i * 64 + col] = act ? B[(int64_t)i * ldb + c] : T(0);
  __syncthreads();
  for (int s0 = 0; s0 < kb; s0 += 16) {
    const int sb = kb - s0 < 16 ? kb - s0 : 16;
    if (rg == 0) {
      for (int i = 1; i < sb; ++i) {
        T acc = xs[(s0 + i) * 64 + col];
        const T *lr = L + (int64_t)(s0 + i) * ldl + s0;
        for (int p = 0; p < i; ++p) acc -= lr[p] * xs[(s0 + p) * 64 + col];
        xs[(s0 + i) * 64 + col] = acc;
      }
    }
    __syncthreads();
    for (int r = s0 + sb + rg; r < kb; r += 4) {
      T acc = xs[r * 64 + col];
      const T *lr = L + (int64_t)r * ldl + s0;
#pragma unroll 4
      for (int p = 0; p < sb; ++p) acc -= lr[p] * xs[(s0 + p) * 64 + col];
      xs[r * 64 + col] = acc;
    }
    __syncthreads();
  }
  if (act)
    for (int i = rg; i < kb; i += 4) B[(int64_t)i * ldb + c] = xs[i * 64 + col];
}

// B := L^-1 B for kb <= 128 with 16 columns of B per workgroup, so that a
// 128-column strip (the next panel's columns) spreads over 8 workgroups and a
// wide strip over many.  L (strictly lower part) and the workgroup's columns
// are staged in LDS.  Diagonal 16-row blocks are solved with one thread per
// column; the rows below each block are then updated by all threads.  Every
// element receives its subtractions b -= b[k] * L[i][k] one at a time in
// increasing k (src/blas/trsm.rs:13-21 applies them in the same order).

template <typename T>
__global__ void __launch_bounds__(256)
    trsm_llu16_kernel(int kb, int ncols, const T *__restrict__ L, int64_t ldl, T *__restrict__ B,
                      int64_t ldb) {
#pragma clang fp contract(off)
  constexpr int CW = kTrsmCW, LX = CW + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int LL = kb + 1;
  T *sL = reinterpret_cast<T *>(smem_raw);  // [kb][kb + 1]
  T *sX = sL + (size_t)kb * LL;             // [kb][CW + 1]
  const int t = threadIdx.x;
  const int c0 = blockIdx.x * CW;
  const int nc = ncols - c0 < CW ? ncols - c0 : CW;
  for (int e = t; e < kb * kb; e += 256) {
    const int r = e / kb, c = e % kb;
    if (c < r) sL[r * LL + c] = L[(int64_t)r * ldl + c];
  }
  for (int e = t; e < kb * CW; e += 256) {
    const int r = e / CW, c = e % CW;
    sX[r * LX + c] = c < nc ? B[(int64_t)r * ldb + c0 + c] : T(0);
  }
  __syncthreads();
  for (int b0 = 0; b0 < kb; b0 += 16) {
    const int bs = kb - b0 < 16 ? kb - b0 : 16;
    if (t < CW) {
      for (int i = 1; i < bs; ++i) {
        T acc = sX[(b0 + i) * LX + t];
        for (int p = 0; p < i; ++p) {
          const T e = sX[(b0 + p) * LX + t] * sL[(b0 + i) * LL + b0 + p];
          acc = acc - e;
        }
        sX[(b0 + i) * LX + t] = acc;
      }
    }
    __syncthreads();
    const int rows = kb - b0 - bs;
    for (int e = t; e < rows * CW; e += 256) {
      const int r = b0 + bs + e / CW, c = e % CW;
      T acc = sX[r * LX + c];
      for (int p = 0; p < bs; ++p) {
        const T pr = sX[(b0 + p) * LX + c] * sL[r * LL + b0 + p];
        acc = acc - pr;
      }
      sX[r * LX + c] = acc;
    }
    __syncthreads();
  }
  for (int e = t; e < kb * CW; e += 256) {
    const int r = e / CW, c = e % CW;
    if (c < nc) B[(int64_t)r * ldb + c0 + c] = sX[r * LX + c];
  }
}

// Interchanges then U12 = L11^-1 A12 for columns [c0, c1) of a panel
// [k, k + kb) (kb <= 128): the laswp + trsm steps of recursive_inner
// (src/lapack/getrf.rs:270-278) fused per 16-column strip.  The strip's rows
// named by the row map (rows[e] receives the old contents of src[e]; the
// first kb entries are rows k .. k + kb - 1 in order) are read once into LDS;
// the rows moved below the panel are written back at once, and the top kb
// rows are solved in LDS and written as U12.  Diagonal 16-row blocks are
// solved in registers, 16 lanes per column; the rows below each block are
// then updated by all threads.  Every element receives its subtractions
// b -= b[p] * L[i][p] one at a time in increasing p (src/blas/trsm.rs:13-21
// applies them in the same order).
template <typename T>
__global__ void __launch_bounds__(256)
    swap_trsm_kernel(T *__restrict__ a, int64_t lda, int k, int kb, int c0, int ncols,
                     const int *__restrict__ rows, const int *__restrict__ src,
                     const int *__restrict__ count, int nmax) {
#pragma clang fp contract(off)
  constexpr int CW = kTrsmCW, LX = CW + 1, RG = 256 / CW;
  static_assert(CW == 16, "the diagonal solve maps 4 waves x 4 columns x 16 rows onto 256 threads");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *sL = reinterpret_cast<T *>(smem_raw);             // strictly lower L11, row r at r(r-1)/2
  T *sX = sL + ((size_t)kb * (kb - 1) / 2 + 1);         // [nmax][CW + 1] map rows' old values
  int *s_rows = reinterpret_cast<int *>(sX + (size_t)nmax * LX);  // [nmax]
  int *s_src = s_rows + nmax;                            // [nmax]
  const int t = threadIdx.x;
  const int cs = c0 + blockIdx.x * CW;
  const int nc = c0 + ncols - cs < CW ? c0 + ncols - cs : CW;
  const int nr = *count;
  for (int e = t; e < nr; e += 256) {
    s_rows[e] = rows[e];
    s_src[e] = src[e];
  }
  // L11 in batches of 16 independent loads per thread (the square is read,
  // its strictly lower part kept).
  const T *L = a + (int64_t)k * lda + k;
  for (int b = 0; b < kb * kb; b += 256 * 16) {
    T v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int idx = b + q * 256 + t, r = idx / kb, p = idx % kb;
      v[q] = (idx < kb * kb && p < r) ? L[(int64_t)r * lda + p] : T(0);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int idx = b + q * 256 + t, r = idx / kb, p = idx % kb;
      if (idx < kb * kb && p < r) sL[r * (r - 1) / 2 + p] = v[q];
    }
  }
  __syncthreads();
  const int c = t % CW, rg = t / CW;
  for (int e0 = rg; e0 < nr; e0 += RG * 16) {
    T v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = e0 + q * RG;
      v[q] = (e < nr && c < nc) ? a[(int64_t)s_src[e] * lda + cs + c] : T(0);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = e0 + q * RG;
      if (e < nr) sX[e * LX + c] = v[q];
    }
  }
  __syncthreads();
  // Rows moved below the panel (entries kb.. that move something).
  for (int e = kb + rg; e < nr; e += RG)
    if (c < nc && s_rows[e] != s_src[e]) a[(int64_t)s_rows[e] * lda + cs + c] = sX[e * LX + c];
  for (int b0 = 0; b0 < kb; b0 += 16) {
    const int bs = kb - b0 < 16 ? kb - b0 : 16;
    {
      // Diagonal block: 16 lanes per column, lane rr holding row b0 + rr of
      // column cc.  At step i row i is final; it is broadcast to the column's
      // lanes and the rows below it subtract their term.
      const int l = t & 63;
      const int cc = (t >> 6) * 4 + (l >> 4), rr = l & 15;
      T x = rr < bs ? sX[(b0 + rr) * LX + cc] : T(0);
      T lrow[15];
#pragma unroll
      for (int i = 0; i < 15; ++i)
        lrow[i] = (i < rr && rr < bs) ? sL[(b0 + rr) * (b0 + rr - 1) / 2 + b0 + i] : T(0);
#pragma unroll
      for (int i = 0; i < 15; ++i) {
        const T xi = shfl_t(x, (l & ~15) | i);
        if (i < rr && rr < bs) {
          const T e = xi * lrow[i];
          x = x - e;
        }
      }
      if (rr < bs) sX[(b0 + rr) * LX + cc] = x;
    }
    __syncthreads();
    T xb[16];
#pragma unroll
    for (int p = 0; p < 16; ++p) xb[p] = p < bs ? sX[(b0 + p) * LX + c] : T(0);
    for (int r = b0 + bs + rg; r < kb; r += RG) {
      const T *lr = sL + r * (r - 1) / 2 + b0;
      T acc = sX[r * LX + c];
#pragma unroll
      for (int p = 0; p < 16; ++p) {
        if (p < bs) {
          const T pr = xb[p] * lr[p];
          acc = acc - pr;
        }
      }
      sX[r * LX + c] = acc;
    }
    __syncthreads();
  }
  for (int r = rg; r < kb; r += RG)
    if (c < nc) a[(int64_t)(k + r) * lda + cs + c] = sX[r * LX + c];
}

template <typename T>
hipError_t launch_swap_trsm(lair_context *ctx, T *a, int64_t lda, int k, int kb, int c0, int c1,
                            const int *rows, const int *src, const int *count, int nmax) {
  if (c1 <= c0 || kb <= 0) return hipSuccess;
  if (kb > kTrsmMaxKb || nmax > kMaxSwapRows || nmax < kb) return hipErrorInvalidValue;
  const size_t shm = sizeof(T) * ((size_t)kb * (kb - 1) / 2 + 1 + (size_t)nmax * (kTrsmCW + 1)) +
                     2 * sizeof(int) * (size_t)nmax + 16;
  if (shm > 160 * 1024) return hipErrorInvalidValue;
  static bool once = (allow_large_lds(swap_trsm_kernel<T>), true);
  (void)once;
  const int slot = prof_begin(ctx, LAIR_KCLASS_TRSM);
  swap_trsm_kernel<T><<<(c1 - c0 + kTrsmCW - 1) / kTrsmCW, 256, shm, ctx->stream>>>(a, lda, k, kb, c0, c1 - c0,
                                                                                      rows, src, count, nmax);
  prof_end(ctx, slot, LAIR_KCLASS_TRSM, (double)kb * kb * (c1 - c0),
           sizeof(T) * (2.0 * (kb + nmax) * (c1 - c0) + 0.5 * kb * kb));
  return hipGetLastError();
}

template hipError_t launch_swap_trsm<double>(lair_context *, double *, int64_t, int, int, int, int,
                                             const int *, const int *, const int *, int);
template hipError_t launch_swap_trsm<float>(lair_context *, float *, int64_t, int, int, int, int,
                                            const int *, const int *, const int *, int);
template hipError_t launch_swap_trsm<zcplx>(lair_context *, zcplx *, int64_t, int, int, int, int,
                                            const int *, const int *, const int *, int);
template hipError_t launch_swap_trsm<ccplx>(lair_context *, ccplx *, int64_t, int, int, int, int,
                                            const int *, const int *, const int *, int);

hipError_t launch_rowmap(lair_context *ctx, const int64_t *ipiv, int k0, int k1) {
  if (k1 <= k0) return hipSuccess;
  if (k1 - k0 > kMaxSwapRows / 2) return hipErrorInvalidValue;
  rowmap_kernel<<<1, 256, 0, ctx->stream>>>(ipiv, k0, k1, ctx->swap_rows, ctx->swap_src, ctx->swap_count);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_trsm_llu(lair_context *ctx, int kb, int ncols, const T *L, int64_t ldl, T *B,
                           int64_t ldb) {
  if (kb <= 0 || ncols <= 0) return hipSuccess;
  if (kb <= kTrsmMaxKb) {
    const size_t shm = sizeof(T) * ((size_t)kb * (kb + 1) + (size_t)kb * (kTrsmCW + 1));
    if (shm > 160 * 1024) return hipErrorInvalidValue;
    static bool once16 = (allow_large_lds(trsm_llu16_kernel<T>), true);
    (void)once16;
    const int slot = prof_begin(ctx, LAIR_KCLASS_TRSM);
    trsm_llu16_kernel<T><<<(ncols + kTrsmCW - 1) / kTrsmCW, 256, shm, ctx->stream>>>(kb, ncols, L, ldl, B,
                                                                                      ldb);
    prof_end(ctx, slot, LAIR_KCLASS_TRSM, (double)kb * kb * ncols,
             sizeof(T) * (2.0 * kb * ncols + 0.5 * kb * kb));
    return hipGetLastError();
  }
  const size_t shm = (size_t)kb * 64 * sizeof(T);
  if (shm > 160 * 1024) return hipErrorInvalidValue;
  static bool once = (allow_large_lds(trsm_llu_kernel<T>), true);
  (void)once;
  const int slot = prof_begin(ctx, LAIR_KCLASS_TRSM);
  trsm_llu_kernel<T><<<(ncols + 63) / 64, 256, shm, ctx->stream>>>(kb, ncols, L, ldl, B, ldb);
  prof_end(ctx, slot, LAIR_KCLASS_TRSM, (double)kb * kb * ncols,
           sizeof(T) * (2.0 * kb * ncols + 0.5 * kb * kb));
  return hipGetLastError();
}

template hipError_t launch_trsm_llu<double>(lair_context *, int, int, const double *, int64_t,
                                            double *, int64_t);
template hipError_t launch_trsm_llu<float>(lair_context *, int, int, const float *, int64_t, float *,
                                           int64_t);
template hipError_t launch_trsm_llu<zcplx>(lair_context *, int, int, const zcplx *, int64_t, zcplx *,
                                           int64_t);
template hipError_t launch_trsm_llu<ccplx>(lair_context *, int, int, const ccplx *, int64_t, ccplx *,
                                           int64_t);

// ---------------------------------------------------------- trtri_diag ---

// blockIdx.x = diagonal block; blockIdx.y = 0: unit lower L -> linv,
// 1: upper U -> uinv.  Each inverse is stored dense BS x BS (row-major).
template <typename T, int BS>
__global__ void __launch_bounds__(BS)
    trtri_diag_kernel(const T *__restrict__ lu, int64_t lda, int n, T *__restrict__ linv,
                      T *__restrict__ uinv) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *M = reinterpret_cast<T *>(smem_raw);  // [BS][BS+1]
  T *X = M + BS * (BS + 1);                // [BS][BS+1]
  constexpr int LD = BS + 1;
  const int b = blockIdx.x, j = threadIdx.x;
  const bool upper = blockIdx.y == 1;
  const int r0 = b * BS;
  const int nb = n - r0 < BS ? n - r0 : BS;
  for (int e = j; e < BS * BS; e += BS) {
    const int r = e / BS, c = e % BS;
    M[r * LD + c] = (r < nb && c < nb) ? lu[(int64_t)(r0 + r) * lda + r0 + c] : T(0);
    X[r * LD + c] = T(0);
  }
  __syncthreads();
  // Column j of the inverse by substitution; each entry's dot product runs
  // over four accumulators, eight terms per step with their LDS loads issued
  // together (the loads, not the multiply-adds, set the time).
  auto dot = [&](int i, int p0, int p1) -> T {  // sum over p in [p0, p1) of M[i][p] X[p][j]
    T a[4] = {T(0), T(0), T(0), T(0)};
    int p = p0;
    for (; p + 8 <= p1; p += 8) {
      T m[8], x[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        m[q] = M[i * LD + p + q];
        x[q] = X[(p + q) * LD + j];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) a[q & 3] += m[q] * x[q];
    }
    for (; p < p1; ++p) a[p & 3] += M[i * LD + p] * X[p * LD + j];
    return (a[0] + a[1]) + (a[2] + a[3]);
  };
  if (j < nb) {
    if (!upper) {
      X[j * LD + j] = T(1);
      for (int i = j + 1; i < nb; ++i) X[i * LD + j] = -dot(i, j, i);
    } else {
      X[j * LD + j] = T(1) / M[j * LD + j];
      for (int i = j - 1; i >= 0; --i) X[i * LD + j] = -dot(i, i + 1, j + 1) / M[i * LD + i];
    }
  }
  __syncthreads();
  T *out = (upper ? uinv : linv) + (size_t)b * BS * BS;
  for (int e = j; e < BS * BS; e += BS) out[e] = X[(e / BS) * LD + e % BS];
}

template <typename T>
hipError_t launch_trtri_diag(lair_context *ctx, const T *lu, int64_t lda, int n, T *linv, T *uinv) {
  if (n <= 0) return hipSuccess;
  constexpr int BS = TrsvBlock<T>::BS;
  const int nblk = (n + BS - 1) / BS;
  const size_t shm = 2 * (size_t)BS * (BS + 1) * sizeof(T);
  static bool once = (allow_large_lds(trtri_diag_kernel<T, BS>), true);
  (void)once;
  const int slot = prof_begin(ctx, LAIR_KCLASS_GETRS);
  trtri_diag_kernel<T, BS><<<dim3(nblk, 2), BS, shm, ctx->stream>>>(lu, lda, n, linv, uinv);
  prof_end(ctx, slot, LAIR_KCLASS_GETRS, (double)nblk * BS * BS * BS / 3.0);
  return hipGetLastError();
}

template hipError_t launch_trtri_diag<double>(lair_context *, const double *, int64_t, int, double *,
                                              double *);
template hipError_t launch_trtri_diag<float>(lair_context *, const float *, int64_t, int, float *,
                                             float *);

// ------------------------------------------------------ trsv_dataflow ---

template <typename T>
struct TrsvArgs {
  const T *lu;
  int64_t lda;
  const T *dinv;  // inverses of the diagonal blocks used by this solve
  const T *xin;   // right-hand sides (may alias x)
  T *x;           // n x nrhs solution
  int64_t ldx;
  int n, nrhs, nblk, nchunks;
  unsigned long long *xt;  // [n][nrhs][words per element] tagged solved values
  uint32_t epoch;
  int forward, trans, dinv_trans;
  int32_t *err;
  int wait_limit;  // polls before a wait gives up (0: every wait gives up at once)
};

// The refinement step fused with the backward solve of the correction
// d = U^-1 y (FP32): the workgroup that solves block i of d also forms
// x_i + d_i in FP64, stores it and publishes it as tagged words (two per
// value, tag = the solve's epoch); residual workgroups stream the FP64 A by
// row blocks (48 rows for one right-hand side, else 24) and consume those
// values column block by column block in the order the solve finishes them,
// so the residual's pass over A overlaps the solve's chain.  The column maxima of |x + d| and |r| fold into
// xbits and rbits (ordered bit patterns, zero at launch).
struct IrFuse {
  double *x;
  int64_t ldx;
  int x_zero;                // 1: x starts at zero (x is not read)
  const double *a;
  int64_t lda;
  const double *b;
  int64_t ldb;
  double *r;
  int64_t ldr;
  unsigned long long *xw;    // [n][nrhs][2] tagged words of x + d
  unsigned long long *xbits, *rbits;
};

// Effective matrix E: forward & !trans -> L (strictly lower part of lu);
// !forward & !trans -> U; forward & trans -> U^T; !forward & trans -> L^T.
// Workgroup (block row i, column chunk ch) computes
//   x_i = D_i^-1 (x_i - sum_j E_ij x_j)
// over the blocks j already solved (j < i forward, j > i backward).  Tiles
// move global -> registers (16-byte lane-contiguous loads, issued before the
// wait for the previous block) -> LDS -> row segments per thread.
template <typename T, int BS, int RC, bool FUSE>
__device__ __forceinline__ void trsv_block(const TrsvArgs<T> &p, int wid, const IrFuse &f) {
  constexpr int WPE = sizeof(T) / 4;  // 32-bit words per element
  constexpr int TPR = 256 / BS;   // threads per row in the product
  constexpr int SEG = BS / TPR;   // columns per thread
  constexpr int VEC = 16 / sizeof(T);           // elements per 16-byte load
  constexpr int NV = BS * BS / VEC / 256;       // 16-byte loads per thread per tile
  constexpr int LDT = BS + VEC;                 // padded LDS row
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *tile = reinterpret_cast<T *>(smem_raw);    // [BS][LDT]
  __shared__ T xs[BS][RC];
  __shared__ T ys[BS][RC];
  __shared__ int s_bail;
  const int t = threadIdx.x;
  const int r = t / TPR, h = t % TPR;
  const int ib = wid / p.nchunks, ch = wid % p.nchunks;
  const int i = p.forward ? ib : p.nblk - 1 - ib;
  const int r0 = i * BS;
  const int c0 = ch * RC;
  const int ncols = p.nrhs - c0 < RC ? p.nrhs - c0 : RC;
  const int gr = r0 + r;
  const bool rowok = gr < p.n;

  // Register staging of one BS x BS block of a row-major matrix (base, ld) at
  // block coordinates (rb, cb), optionally transposed; out-of-range is zero.
  // Element e of load k covers row (t*VEC + k*256*VEC) / BS.
  auto fetch = [&](const T *base, int64_t ld, int rb, int cb, vec_t (&v)[NV]) {
    const bool aligned = ((ld * sizeof(T)) % 16 == 0) && (reinterpret_cast<uintptr_t>(base) % 16 == 0);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int lin = (k * 256 + t) * VEC;
      const int rr = lin / BS, cc = lin % BS;
      const int grow = rb + rr, gcol = cb + cc;
      if (aligned && grow < p.n && gcol + VEC <= p.n) {
        v[k] = *reinterpret_cast<const vec_t *>(base + (int64_t)grow * ld + gcol);
      } else {
#pragma unroll
        for (int q = 0; q < VEC; ++q)
          v[k][q] = (grow < p.n && gcol + q < p.n) ? base[(int64_t)grow * ld + gcol + q] : T(0);
      }
    }
  };
  // Stores the staged block into LDS as E (row r, column c), transposing when
  // the staged block is stored M[c][r].
  auto stash = [&](const vec_t (&v)[NV], bool transposed) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int lin = (k * 256 + t) * VEC;
      const int rr = lin / BS, cc = lin % BS;
      if (!transposed) {
        *reinterpret_cast<vec_t *>(tile + rr * LDT + cc) = v[k];
      } else {
#pragma unroll
        for (int q = 0; q < VEC; ++q) tile[(cc + q) * LDT + rr] = v[k][q];
      }
    }
  };

  // Diagonal inverse D_i (dense BS x BS, row-major): row segment per thread.
  T dreg[SEG];
  {
    vec_t v[NV];
    const T *D = p.dinv + (size_t)i * BS * BS;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int lin = (k * 256 + t) * VEC;
      v[k] = *reinterpret_cast<const vec_t *>(D + lin);
    }
    stash(v, p.dinv_trans != 0);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SEG; ++s) dreg[s] = tile[r * LDT + h * SEG + s];
    __syncthreads();
  }
  T acc[RC];
#pragma unroll
  for (int q = 0; q < RC; ++q) acc[q] = T(0);
  if (t == 0) s_bail = 0;
  // Fused refinement: this block's rows of x, read while the chain runs.
  double xold[RC];
#pragma unroll
  for (int q = 0; q < RC; ++q)
    xold[q] = (FUSE && !f.x_zero && h == 0 && rowok && q < ncols) ? f.x[(int64_t)gr * f.ldx + c0 + q] : 0.0;

  const int nb = p.forward ? i : p.nblk - 1 - i;  // number of dependency blocks
  auto issue = [&](int d, vec_t (&v)[NV]) {
    const int j = p.forward ? d : p.nblk - 1 - d;
    if (!p.trans)
      fetch(p.lu, p.lda, r0, j * BS, v);   // E_ij = M[r0.., jBS..]
    else
      fetch(p.lu, p.lda, j * BS, r0, v);   // E_ij = (M[jBS.., r0..])^T
  };
  vec_t stage[NV];
  if (nb > 0) issue(0, stage);
  for (int d = 0; d < nb; ++d) {
    const int j = p.forward ? d : p.nblk - 1 - d;
    const int cb = j * BS;
    stash(stage, p.trans != 0);
    if (d + 1 < nb) issue(d + 1, stage);
    // The solved values of block j: every thread polls its own tagged words
    // (bounded waits), so the values arrive with the first load that sees
    // them.
    {
      bool bail = false;
      for (int e = t; e < BS * RC; e += 256) {
        const int rr = e / RC, q = e % RC;
        const int gi = cb + rr;
        T v = T(0);
        if (gi < p.n && q < ncols) {
          const unsigned long long *w = p.xt + ((int64_t)gi * p.nrhs + c0 + q) * WPE;
          int spins = 0;
          for (;;) {
            if (spins >= p.wait_limit) {
              bail = true;
              break;
            }
            uint32_t part[WPE];
            bool ok = true;
#pragma unroll
            for (int k = 0; k < WPE; ++k) {
              const unsigned long long x = ld_sc1(w + k);
              ok = ok && (uint32_t)(x >> 32) == p.epoch;
              part[k] = (uint32_t)x;
            }
            if (ok) {
              v = from_words<T>(part);
              break;
            }
            if ((++spins & 63) == 0 && ld_sc1(p.err) != 0) {
              bail = true;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
        xs[rr][q] = v;
      }
      if (bail) {
        atomic_or_agent(p.err, kErrWait);
        s_bail = 1;
      }
    }
    __syncthreads();
    if (s_bail) return;
#pragma unroll
    for (int s = 0; s < SEG; ++s) {
      const int c = h * SEG + s;
      const T e = tile[r * LDT + c];
#pragma unroll
      for (int q = 0; q < RC; ++q) acc[q] += e * xs[c][q];
    }
    __syncthreads();
  }
  // Combine the TPR partial sums of each row (consecutive lanes).
#pragma unroll
  for (int q = 0; q < RC; ++q) {
#pragma unroll
    for (int off = TPR / 2; off >= 1; off >>= 1) acc[q] += __shfl_xor(acc[q], off);
  }
  if (h == 0) {
#pragma unroll
    for (int q = 0; q < RC; ++q) {
      T bval = (rowok && q < ncols) ? p.xin[(int64_t)gr * p.ldx + c0 + q] : T(0);
      ys[r][q] = bval - acc[q];
    }
  }
  __syncthreads();
  T out[RC];
#pragma unroll
  for (int q = 0; q < RC; ++q) out[q] = T(0);
#pragma unroll
  for (int s = 0; s < SEG; ++s) {
    const int c = h * SEG + s;
#pragma unroll
    for (int q = 0; q < RC; ++q) out[q] += dreg[s] * ys[c][q];
  }
#pragma unroll
  for (int q = 0; q < RC; ++q) {
#pragma unroll
    for (int off = TPR / 2; off >= 1; off >>= 1) out[q] += __shfl_xor(out[q], off);
  }
  if (h == 0 && rowok) {
#pragma unroll
    for (int q = 0; q < RC; ++q) {
      if (q < ncols) {
        p.x[(int64_t)gr * p.ldx + c0 + q] = out[q];
        unsigned long long *w = p.xt + ((int64_t)gr * p.nrhs + c0 + q) * WPE;
#pragma unroll
        for (int k = 0; k < WPE; ++k) st_sc1(w + k, tagged(p.epoch, word_of<T>(out[q], k)));
      }
    }
  }
  if constexpr (FUSE) {
    // x + d in FP64 (the chain's words above go first), published for the
    // residual workgroups; the block's maxima of |x + d| by one atomic per
    // column.
    __shared__ unsigned long long s_xm[4][RC];
    unsigned long long xm[RC];
#pragma unroll
    for (int q = 0; q < RC; ++q) {
      xm[q] = 0ull;
      if (h == 0 && rowok && q < ncols) {
        const double xn = xold[q] + (double)out[q];
        f.x[(int64_t)gr * f.ldx + c0 + q] = xn;
        unsigned long long *w = f.xw + ((int64_t)gr * p.nrhs + c0 + q) * 2;
        st_sc1(w, tagged(p.epoch, word_of<double>(xn, 0)));
        st_sc1(w + 1, tagged(p.epoch, word_of<double>(xn, 1)));
        xm[q] = (unsigned long long)__double_as_longlong(fabs(xn));
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(xm[q], off);
        xm[q] = o > xm[q] ? o : xm[q];
      }
    }
    if ((t & 63) == 0) {
#pragma unroll
      for (int q = 0; q < RC; ++q) s_xm[t >> 6][q] = xm[q];
    }
    __syncthreads();
    if (t < RC && t < ncols) {
      unsigned long long m = s_xm[0][t];
#pragma unroll
      for (int w = 1; w < 4; ++w) m = s_xm[w][t] > m ? s_xm[w][t] : m;
      atomicMax(f.xbits + c0 + t, m);
    }
  }
}

template <typename T, int BS, int RC>
__global__ void __launch_bounds__(256) trsv_dataflow_kernel(TrsvArgs<T> p) {
  trsv_block<T, BS, RC, false>(p, blockIdx.x, IrFuse{});
}

// Residual workgroup rw of the fused refinement step: rows
// 3 RPW rw .. 3 RPW (rw + 1) - 1 of r = b - A (x + d).  Wave 0 only waits for
// the values: per column block (in the backward solve's order, last block
// first) every lane reads the words of its two columns, lane 0 alone polls
// while some are missing, and the values go to LDS (by block parity).
// Waves 1..3 hold RPW rows each, two columns per lane, and keep the next
// block's A loads in flight across the workgroup barrier
// (which waits for LDS only), so HBM latency overlaps the chain.  The
// per-lane partial sums of a wave's rows are reduced by a butterfly that
// halves the rows at each step (the top lane bits select the row).
template <int RC>
__device__ __forceinline__ void ir_residual_block(const TrsvArgs<float> &p, const IrFuse &f, int rw) {
  constexpr int BS = TrsvBlock<float>::BS;  // 2 columns per lane
  constexpr int RPW = RC == 1 ? 16 : 8;      // rows per worker wave
  constexpr int LOGR = RPW == 16 ? 4 : 3;
  static_assert(BS == 128, "two columns per lane");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double *sx = reinterpret_cast<double *>(smem_raw);  // [2][BS][RC]
  __shared__ unsigned long long s_rm[4][RC];
  __shared__ int s_bail[2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = p.n, nq = p.nrhs, nblk = p.nblk;
  const int row0 = rw * 3 * RPW + (wave - 1) * RPW;

  // Worker waves: A's values of block d, columns cb + lane and cb + 64 + lane
  // of the wave's rows (each load instruction reads 512 contiguous bytes).
  // Rows and columns past the end are clamped to the last one, so the loads
  // carry no branch (the compiler can then wait for one block's loads while
  // the next block's are in flight); the products mask those columns.
  auto load_a = [&](int d, double (&a0)[RPW], double (&a1)[RPW]) {
    const int cb = (nblk - 1 - d) * BS;
    const int c0 = min(cb + lane, n - 1), c1 = min(cb + 64 + lane, n - 1);
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const double *ap = f.a + (int64_t)min(row0 + i, n - 1) * f.lda;
      a0[i] = ap[c0];
      a1[i] = ap[c1];
    }
  };
  // Wave 0: block d's values of x + d into LDS; returns false if a wait gave up.
  auto wait_x = [&](int d) -> bool {
    const int cb = (nblk - 1 - d) * BS;
    double *dst = sx + (d & 1) * BS * RC;
    bool bail = false;
    int spins = 0;
    auto ready = [&](const unsigned long long *w, double *v) -> bool {
      const unsigned long long w0 = ld_sc1(w), w1 = ld_sc1(w + 1);
      if ((uint32_t)(w0 >> 32) != p.epoch || (uint32_t)(w1 >> 32) != p.epoch) return false;
      const uint32_t part[2] = {(uint32_t)w0, (uint32_t)w1};
      *v = from_words<double>(part);
      return true;
    };
    auto spin = [&]() -> bool {  // false once the wait gives up
      if (spins >= p.wait_limit) return false;
      if ((++spins & 63) == 0 && ld_sc1(p.err) != 0) return false;
      __builtin_amdgcn_s_sleep(2);
      return true;
    };
    // Every lane reads the words of its two columns (one round trip when the
    // block is already solved); while some are missing, lane 0 polls the
    // first missing one and then the lanes read theirs again.
    double v[2][RC];
    bool have[2][RC];
#pragma unroll
    for (int k2 = 0; k2 < 2; ++k2)
#pragma unroll
      for (int q = 0; q < RC; ++q) {
        v[k2][q] = 0.0;
        have[k2][q] = !(cb + 64 * k2 + lane < n && q < nq);
      }
    for (;;) {
      bool all = true;
#pragma unroll
      for (int k2 = 0; k2 < 2; ++k2)
#pragma unroll
        for (int q = 0; q < RC; ++q)
          if (!have[k2][q]) {
            have[k2][q] = ready(f.xw + ((int64_t)(cb + 64 * k2 + lane) * nq + q) * 2, &v[k2][q]);
            all = all && have[k2][q];
          }
      const unsigned long long missing = __ballot(!all);
      if (!missing) break;
      // The first missing word of the first lane that lacks one.
      const int ml = __ffsll((long long)missing) - 1;
      int mw = 0;
#pragma unroll
      for (int k2 = 1; k2 >= 0; --k2)
#pragma unroll
        for (int q = RC - 1; q >= 0; --q)
          if (!have[k2][q]) mw = (cb + 64 * k2 + lane) * nq + q;
      mw = __shfl(mw, ml);
      if (lane == 0) {
        double tmp;
        while (!ready(f.xw + (int64_t)mw * 2, &tmp))
          if (!spin()) {
            bail = true;
            break;
          }
      }
      if (__shfl((int)bail, 0)) {
        bail = true;
        break;
      }
    }
#pragma unroll
    for (int k2 = 0; k2 < 2; ++k2)
#pragma unroll
      for (int q = 0; q < RC; ++q) dst[(64 * k2 + lane) * RC + q] = bail ? 0.0 : v[k2][q];
    return !__any(bail);
  };

  if (threadIdx.x == 0) s_bail[0] = s_bail[1] = 0;
  double acc[RPW][RC];
#pragma unroll
  for (int i = 0; i < RPW; ++i)
#pragma unroll
    for (int q = 0; q < RC; ++q) acc[i][q] = 0.0;
  double ba0[RPW], ba1[RPW], bb0[RPW], bb1[RPW];
  if (wave > 0) load_a(0, ba0, ba1);
  // One block: wave 0 waits for its values while the workers issue the next
  // block's loads; after the barrier the workers multiply.  The bail word
  // alternates with the LDS buffer, so every wave leaves at the same block.
  auto step = [&](int d, double (&c0)[RPW], double (&c1)[RPW], double (&n0)[RPW], double (&n1)[RPW]) -> bool {
    if (wave == 0) {
      const bool ok = wait_x(d);
      if (lane == 0) s_bail[d & 1] = ok ? 0 : 1;
    } else if (d + 1 < nblk) {
      load_a(d + 1, n0, n1);
    }
    __syncthreads();
    if (s_bail[d & 1]) return false;
    if (wave > 0) {
      const double *xs = sx + (d & 1) * BS * RC + lane * RC;
      const int cb = (nblk - 1 - d) * BS;
      const bool ok0 = cb + lane < n, ok1 = cb + 64 + lane < n;
#pragma unroll
      for (int q = 0; q < RC; ++q) {
        const double x0 = xs[q], x1 = xs[64 * RC + q];
#pragma unroll
        for (int i = 0; i < RPW; ++i)
          acc[i][q] += (ok0 ? c0[i] : 0.0) * x0 + (ok1 ? c1[i] : 0.0) * x1;
      }
    }
    return true;
  };
  bool ok = true;
  for (int d = 0; d < nblk && ok; d += 2) {
    ok = step(d, ba0, ba1, bb0, bb1);
    if (ok && d + 1 < nblk) ok = step(d + 1, bb0, bb1, ba0, ba1);
  }
  if (!ok && threadIdx.x == 0) atomic_or_agent(p.err, kErrWait);
  unsigned long long rm[RC];
#pragma unroll
  for (int q = 0; q < RC; ++q) rm[q] = 0ull;
  if (wave > 0) {
#pragma unroll
    for (int q = 0; q < RC; ++q) {
      // Butterfly over lane bits 5, 4, ...: a lane keeps half of its rows,
      // summed with its partner's; then the lanes that share a row sum.
#pragma unroll
      for (int h = RPW / 2, o = 32; h >= 1; h >>= 1, o >>= 1) {
        const bool up = (lane & o) != 0;
#pragma unroll
        for (int m = 0; m < h; ++m) {
          const double send = up ? acc[m][q] : acc[m + h][q];
          const double keep = up ? acc[m + h][q] : acc[m][q];
          acc[m][q] = keep + __shfl_xor(send, o);
        }
      }
#pragma unroll
      for (int o = 32 >> LOGR; o >= 1; o >>= 1) acc[0][q] += __shfl_xor(acc[0][q], o);
    }
    int ri = 0;
#pragma unroll
    for (int s = 0; s < LOGR; ++s) ri += ((lane >> (5 - s)) & 1) * ((RPW / 2) >> s);
    const int row = row0 + ri;
    const bool owner = (lane & ((64 / RPW) - 1)) == 0;
#pragma unroll
    for (int q = 0; q < RC; ++q) {
      if (ok && owner && row < n && q < nq) {
        const double rv = f.b[(int64_t)row * f.ldb + q] - acc[0][q];
        f.r[(int64_t)row * f.ldr + q] = rv;
        rm[q] = (unsigned long long)__double_as_longlong(fabs(rv));
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(rm[q], off);
        rm[q] = o > rm[q] ? o : rm[q];
      }
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < RC; ++q) s_rm[wave][q] = rm[q];
  }
  __syncthreads();
  if ((int)threadIdx.x < RC && (int)threadIdx.x < nq) {
    const int q = threadIdx.x;
    unsigned long long m = s_rm[0][q];
#pragma unroll
    for (int w = 1; w < 4; ++w) m = s_rm[w][q] > m ? s_rm[w][q] : m;
    atomicMax(f.rbits + q, m);
  }
}

// The fused refinement step: workgroups 0 .. nblk - 1 run the backward
// solve (and x += d), the rest the residual.  The solve's workgroups come
// first in dispatch order and wait only on each other, the residual's wait
// only on the solve, so every wait ends whichever workgroups are resident.
template <int RC>
__global__ void __launch_bounds__(256) trsv_ir_kernel(TrsvArgs<float> p, IrFuse f) {
  constexpr int BS = TrsvBlock<float>::BS;
  const int ntri = p.nblk * p.nchunks;
  if ((int)blockIdx.x < ntri)
    trsv_block<float, BS, RC, true>(p, blockIdx.x, f);
  else
    ir_residual_block<RC>(p, f, blockIdx.x - ntri);
}

// Many right-hand sides in FP32 (nrhs >= 16, not transposed): the same
// dataflow as trsv_dataflow_kernel (one workgroup per 128-row block and
// 16-column chunk, solved values published as tagged words), with the block
// products on matrix cores.  Wave w owns rows 32w .. 32w + 31 of the block
// as two 16 x 16 v_mfma_f32_16x16x4f32 accumulators, so a workgroup needs
// few registers and two fit on a CU (all 512 workgroups of a 64-column solve
// at n = 16384 are resident at once).
//  * E_ij tiles (128 x 128) are staged through registers into LDS, 16-byte
//    chunk c of row r at chunk position c ^ (r & 31); MFMA step s takes
//    k = 16 t + 4 g + s from lane group g, so a lane's four k values are one
//    16-byte LDS read and the 16 rows read together are on distinct banks.
//  * -X_j is staged transposed ([column][k]), so acc = b - sum E_ij X_j.
//  * D_i^-1 (the inverted diagonal block) stays in registers in the MFMA
//    operand layout; y goes through LDS as the second operand of x = D y.
__global__ void __launch_bounds__(256, 2) trsm_mfma_kernel(TrsvArgs<float> p) {
  constexpr int BS = 128, RC = 16, XP = BS + 4;
  typedef float f4 __attribute__((ext_vector_type(4)));
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float *tile = reinterpret_cast<float *>(smem_raw);  // [BS][BS], swizzled 16-byte chunks
  __shared__ __align__(16) float xs[RC][XP];          // B operand, transposed
  __shared__ int s_bail;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int g = lane >> 4, lc = lane & 15;
  const int ib = blockIdx.x / p.nchunks, ch = blockIdx.x % p.nchunks;
  const int i = p.forward ? ib : p.nblk - 1 - ib;
  const int r0 = i * BS, c0 = ch * RC;
  const int ncols = p.nrhs - c0 < RC ? p.nrhs - c0 : RC;
  auto swz = [](int r, int c4) { return r * BS + ((c4 ^ (r & 31)) << 2); };
  // Rows of this lane's accumulator element (rb, r).
  auto arow = [&](int rb, int r) { return 32 * wave + 16 * rb + 4 * g + r; };

  // Staging of an E tile: thread t holds rows t/32 + 8q, chunk t % 32.
  f4 stage[16];
  const bool aligned = ((p.lda * (int64_t)sizeof(float)) % 16 == 0) && (reinterpret_cast<uintptr_t>(p.lu) % 16 == 0);
  auto fetch = [&](int cb) {
    const int cc = (t & 31) * 4;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int grow = r0 + (t >> 5) + 8 * q, gcol = cb + cc;
      if (aligned && grow < p.n && gcol + 4 <= p.n) {
        stage[q] = *reinterpret_cast<const f4 *>(p.lu + (int64_t)grow * p.lda + gcol);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          stage[q][e] = (grow < p.n && gcol + e < p.n) ? p.lu[(int64_t)grow * p.lda + gcol + e] : 0.0f;
      }
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < 16; ++q) *reinterpret_cast<f4 *>(tile + swz((t >> 5) + 8 * q, t & 31)) = stage[q];
  };

  // D_i^-1 in the MFMA A-operand layout: dreg[rb][s] = D[row(rb)][16 s + 4 g .. + 3].
  f4 dreg[2][8];
  {
    const float *D = p.dinv + (size_t)i * BS * BS;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int s = 0; s < 8; ++s)
        dreg[rb][s] = *reinterpret_cast<const f4 *>(D + (32 * wave + 16 * rb + lc) * BS + 16 * s + 4 * g);
  }
  f4 acc[2];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gi = r0 + arow(rb, r);
      acc[rb][r] = (gi < p.n && lc < ncols) ? p.xin[(int64_t)gi * p.ldx + c0 + lc] : 0.0f;
    }
  if (t == 0) s_bail = 0;

  const int nb = p.forward ? i : p.nblk - 1 - i;  // dependency blocks
  auto dep = [&](int d) { return p.forward ? d : p.nblk - 1 - d; };
  if (nb > 0) fetch(dep(0) * BS);
  for (int d = 0; d < nb; ++d) {
    const int cb = dep(d) * BS;
    __syncthreads();  // the previous product is done with tile and xs
    stash();
    if (d + 1 < nb) fetch(dep(d + 1) * BS);
    // The solved values of block j: every thread polls its own tagged words
    // (bounded waits).
    {
      bool bail = false;
#pragma unroll
      for (int e0 = 0; e0 < BS * RC; e0 += 256) {
        const int e = e0 + t, kr = e / RC, q = e % RC;
        const int gi = cb + kr;
        float v = 0.0f;
        if (gi < p.n && q < ncols) {
          const unsigned long long *w = p.xt + ((int64_t)gi * p.nrhs + c0 + q);
          int spins = 0;
          for (;;) {
            if (spins >= p.wait_limit) {
              bail = true;
              break;
            }
            const unsigned long long x = ld_sc1(w);
            if ((uint32_t)(x >> 32) == p.epoch) {
              v = __uint_as_float((uint32_t)x);
              break;
            }
            if ((++spins & 63) == 0 && ld_sc1(p.err) != 0) {
              bail = true;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
        xs[q][kr] = -v;
      }
      if (bail) {
        atomic_or_agent(p.err, kErrWait);
        s_bail = 1;
      }
    }
    __syncthreads();
    if (s_bail) return;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const f4 bq = *reinterpret_cast<const f4 *>(&xs[lc][16 * s + 4 * g]);
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const f4 aq = *reinterpret_cast<const f4 *>(tile + swz(32 * wave + 16 * rb + lc, 4 * s + g));
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) acc[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(aq[ks], bq[ks], acc[rb], 0, 0, 0);
      }
    }
  }
  __syncthreads();  // the last product is done with xs
  // y (this block's right-hand sides minus the dependencies) as the B operand.
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int r = 0; r < 4; ++r) xs[lc][arow(rb, r)] = acc[rb][r];
  __syncthreads();
  f4 out[2];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) out[rb] = (f4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const f4 bq = *reinterpret_cast<const f4 *>(&xs[lc][16 * s + 4 * g]);
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
        out[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(dreg[rb][s][ks], bq[ks], out[rb], 0, 0, 0);
  }
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gi = r0 + arow(rb, r);
      if (gi < p.n && lc < ncols) {
        p.x[(int64_t)gi * p.ldx + c0 + lc] = out[rb][r];
        st_sc1(p.xt + ((int64_t)gi * p.nrhs + c0 + lc), tagged(p.epoch, __float_as_uint(out[rb][r])));
      }
    }
}

template <typename T>
static hipError_t trsv_launch(lair_context *ctx, const T *lu, int64_t lda, const T *dinv,
                              bool dinv_trans, const T *xin, T *x, int64_t ldx, int n, int nrhs,
                              bool forward, bool trans, IrFuse *fz) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  constexpr int BS = TrsvBlock<T>::BS;
  // Right-hand sides per workgroup: 4 columns keep each chain link short
  // (the link's products and reductions grow with the chunk width), and the
  // chunks run as independent chains.  FP32 with 16 or more right-hand sides
  // (not transposed) runs the matrix-core kernel in chunks of 16.
  const bool mfma = std::is_same<T, float>::value && nrhs >= 16 && !trans && !dinv_trans;
  const int rc = mfma ? 16 : (nrhs == 1 ? 1 : 4);
  const int nblk = (n + BS - 1) / BS;
  const int nchunks = (nrhs + rc - 1) / rc;
  // Tagged words of the solved values; zeroed whenever they are (re)allocated
  // or the epoch would wrap, so that no stale word carries a live tag.  A
  // caller-provided buffer (xt_override) was zeroed by the caller.
  unsigned long long *xt = ctx->xt_override;
  if (!xt) {
    // The fused refinement step also needs two words per value of x + d.
    const size_t xt_bytes = sizeof(unsigned long long) * (size_t)n * nrhs * (sizeof(T) / 4 + (fz ? 2 : 0));
    const bool grow = ctx->ws_flags.bytes < xt_bytes;
    hipError_t e = ensure(ctx->ws_flags, xt_bytes);
    if (e != hipSuccess) return e;
    if (grow) ctx->ws_flags_zeroed = 0;
    if (ctx->ws_flags_zeroed < ctx->ws_flags.bytes || ctx->tri_epoch > 0x70000000u) {
      e = launch_fill_u32(ctx, static_cast<uint32_t *>(ctx->ws_flags.ptr),
                          (int64_t)(ctx->ws_flags.bytes / sizeof(uint32_t)), 0u);
      if (e != hipSuccess) return e;
      ctx->ws_flags_zeroed = ctx->ws_flags.bytes;
      ctx->tri_epoch = 0;
    }
    xt = static_cast<unsigned long long *>(ctx->ws_flags.ptr);
  } else if (ctx->tri_epoch > 0x70000000u) {
    ctx->tri_epoch = 0;
  }
  TrsvArgs<T> a;
  a.lu = lu;
  a.lda = lda;
  a.dinv = dinv;
  a.xin = xin;
  a.x = x;
  a.ldx = ldx;
  a.n = n;
  a.nrhs = nrhs;
  a.nblk = nblk;
  a.nchunks = nchunks;
  a.xt = xt;
  a.epoch = ++ctx->tri_epoch;
  a.forward = forward ? 1 : 0;
  a.trans = trans ? 1 : 0;
  a.dinv_trans = dinv_trans ? 1 : 0;
  a.err = ctx->err;
  a.wait_limit = ctx->wait_limit;
  const unsigned grid = (unsigned)(nblk * nchunks);
  const size_t shm = (size_t)BS * (BS + 16 / sizeof(T)) * sizeof(T);
  static bool once = (allow_large_lds(trsv_dataflow_kernel<T, BS, 1>),
                      allow_large_lds(trsv_dataflow_kernel<T, BS, 4>), true);
  (void)once;
  if constexpr (std::is_same<T, float>::value) {
    if (fz) {
      static bool once_f = (allow_large_lds(trsv_ir_kernel<1>), allow_large_lds(trsv_ir_kernel<4>), true);
      (void)once_f;
      fz->xw = xt + (size_t)n * nrhs;
      // Residual workgroups: 48 rows each for one right-hand side, else 24.
      const int rrows = rc == 1 ? 48 : 24;
      const unsigned gridf = grid + (unsigned)((n + rrows - 1) / rrows);
      const int slot = prof_begin(ctx, LAIR_KCLASS_GETRS);
      if (rc == 1)
        trsv_ir_kernel<1><<<gridf, 256, shm, ctx->stream>>>(a, *fz);
      else
        trsv_ir_kernel<4><<<gridf, 256, shm, ctx->stream>>>(a, *fz);
      // The solve's bytes plus one pass over the FP64 A.
      prof_end(ctx, slot, LAIR_KCLASS_GETRS, (double)n * n * nrhs * 3,
               sizeof(T) * 0.5 * (double)n * n + 8.0 * n * n + 32.0 * n * nrhs);
      return hipGetLastError();
    }
  }
  const int slot = prof_begin(ctx, LAIR_KCLASS_GETRS);
  if constexpr (std::is_same<T, float>::value) {
    if (mfma) {
      static bool once_m = (allow_large_lds(trsm_mfma_kernel), true);
      (void)once_m;
      trsm_mfma_kernel<<<grid, 256, (size_t)BS * BS * sizeof(float), ctx->stream>>>(a);
      prof_end(ctx, slot, LAIR_KCLASS_GETRS, (double)n * n * nrhs,
               sizeof(T) * (0.5 * (double)n * n + 2.0 * n * nrhs));
      return hipGetLastError();
    }
  }
  if (rc == 1)
    trsv_dataflow_kernel<T, BS, 1><<<grid, 256, shm, ctx->stream>>>(a);
  else
    trsv_dataflow_kernel<T, BS, 4><<<grid, 256, shm, ctx->stream>>>(a);
  prof_end(ctx, slot, LAIR_KCLASS_GETRS, (double)n * n * nrhs,
           sizeof(T) * (0.5 * (double)n * n + 2.0 * n * nrhs));
  return hipGetLastError();
}

template <typename T>
hipError_t launch_trsv_dataflow(lair_context *ctx, const T *lu, int64_t lda, const T *dinv,
                                bool dinv_trans, const T *xin, T *x, int64_t ldx, int n, int nrhs,
                                bool forward, bool trans) {
  return trsv_launch<T>(ctx, lu, lda, dinv, dinv_trans, xin, x, ldx, n, nrhs, forward, trans, nullptr);
}

hipError_t launch_trsv_ir(lair_context *ctx, const float *lu, int64_t lda, const float *uinv,
                          const float *y, float *d, int n, int nrhs, double *x, int64_t ldx,
                          bool x_zero, const double *a, int64_t lda64, const double *b, int64_t ldb,
                          double *r, int64_t ldr, unsigned long long *xbits,
                          unsigned long long *rbits) {
  if (nrhs > 4) return hipErrorInvalidValue;
  IrFuse fz{};
  fz.x = x;
  fz.ldx = ldx;
  fz.x_zero = x_zero ? 1 : 0;
  fz.a = a;
  fz.lda = lda64;
  fz.b = b;
  fz.ldb = ldb;
  fz.r = r;
  fz.ldr = ldr;
  fz.xbits = xbits;
  fz.rbits = rbits;
  return trsv_launch<float>(ctx, lu, lda, uinv, false, y, d, nrhs, n, nrhs, false, false, &fz);
}

template hipError_t launch_trsv_dataflow<double>(lair_context *, const double *, int64_t,
                                                 const double *, bool, const double *, double *,
                                                 int64_t, int, int, bool, bool);
template hipError_t launch_trsv_dataflow<float>(lair_context *, const float *, int64_t,
                                                const float *, bool, const float *, float *, int64_t,
                                                int, int, bool, bool);

}  // namespace lair
