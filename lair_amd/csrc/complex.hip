// Complex LU pieces that the real path does not share: the trailing update
// on the real MFMA GEMM through an embedding of U12, and the complex getrs.
//
// Trailing update A22 -= L21 U12 (src/lapack/getrf.rs:289-296, gemm.rs:6-32)
// in complex arithmetic, as one real GEMM: with interleaved storage, row i of
// L21 (K complex) is a real row of 2K values (re, im, re, im, ...) and row i
// of A22 one of 2N values, so
//   A22_int (M x 2N) -= L21_int (M x 2K) * E (2K x 2N)
// with E[2p] = (u.re, u.im) per column (row p of U12 as stored) and
// E[2p + 1] = (-u.im, u.re).  Same flop count as the complex product; L21 and
// A22 are read in place, only U12 (K rows) is copied.  E is indexed by the
// absolute column of the matrix, so the two streams of the look-ahead, which
// update disjoint column ranges of the same panel, never share a slot.
//
// getrs (src/lapack/getrs.rs:12-39): x = P b; unit-lower forward
// substitution x[i] -= a[i][k] x[k] for k < i in increasing k; back
// substitution x[i] -= a[i][k] x[k] for k > i in increasing k, then
// x[i] /= a[i][i] (num-complex DivAssign).  One workgroup per right-hand
// side; rows i = t + 1024 s live in registers of thread t.  The forward pass
// goes by 64-row blocks: the wave owning the block solves it (each row's
// subtractions in increasing k), the block's values go through LDS and every
// row below subtracts the block's 64 terms in increasing k, so each row sees
// exactly the reference's sequence of operations.  The back substitution
// takes its terms in increasing k, which needs every later unknown first: for
// n <= kExactOrderMax one wave runs it row by row in that order (bit-identical
// to the reference); beyond, the same 64-row blocking runs bottom-up (the
// terms of a row come in decreasing block order: a backward-stable reordering
// of the sums).
#include <algorithm>

#include "lair_internal.h"

namespace lair {

template <typename R>
__global__ void __launch_bounds__(256)
    cembed_kernel(int K, int N, const cplx<R> *__restrict__ B, int64_t ldb, R *__restrict__ E, int64_t lde) {
  const int j = blockIdx.x * 256 + threadIdx.x, p = blockIdx.y;
  if (j >= N) return;
  const cplx<R> u = B[(int64_t)p * ldb + j];
  R *e0 = E + (int64_t)(2 * p) * lde + 2 * j;
  R *e1 = e0 + lde;
  e0[0] = u.re;
  e0[1] = u.im;
  e1[0] = -u.im;
  e1[1] = u.re;
}

template <typename T>
hipError_t launch_cgemm_minus(lair_context *ctx, int mode, int M, int N, int K, const T *A, int64_t lda,
                              const T *B, int64_t ldb, T *C, int64_t ldc, int col, int kclass) {
  typedef typename cplx_traits<T>::real R;
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  R *E = static_cast<R *>(ctx->ws_cexp.ptr);
  const int64_t lde = ctx->cexp_ld;
  if (!E || K > kComplexMaxNb || col < 0 || 2 * ((int64_t)col + N) > lde ||
      ctx->ws_cexp.bytes < sizeof(R) * (size_t)(2 * kComplexMaxNb) * (size_t)lde)
    return hipErrorInvalidValue;
  R *Ec = E + 2 * (int64_t)col;
  cembed_kernel<R><<<dim3((N + 255) / 256, K), 256, 0, ctx->stream>>>(K, N, B, ldb, Ec, lde);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_gemm_minus<R>(ctx, mode, M, 2 * N, 2 * K, reinterpret_cast<const R *>(A), 2 * lda, false, Ec, lde,
                              reinterpret_cast<R *>(C), 2 * ldc, kclass);
}

template <typename T, int RPT>
__global__ void __launch_bounds__(1024)
    cgetrs_kernel(int n, const T *__restrict__ lu, int64_t lda, const int *__restrict__ perm, T *__restrict__ b,
                  int64_t ldb) {
#pragma clang fp contract(off)
  constexpr int NT = 1024;
  constexpr int LB = RPT >= 16 ? 4 : 8;  // loads in flight per batch
  __shared__ T xs[64];
  __shared__ T xall[kExactOrderMax];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int r = blockIdx.x;
  T x[RPT];
#pragma unroll
  for (int s = 0; s < RPT; ++s) {
    const int i = t + NT * s;
    x[s] = i < n ? b[(int64_t)perm[i] * ldb + r] : T(0);
  }
  const int nblk = (n + 63) / 64;

  // Forward substitution with unit lower L.
  for (int bk = 0; bk < nblk; ++bk) {
    const int k0 = bk * 64, kn = min(64, n - k0);
    if (wave == bk % 16) {
      const int os = k0 / NT;
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        if (s == os) {
          const int i = k0 + lane;
          const T *li = lu + (int64_t)(lane < kn ? i : k0) * lda + k0;
          T xv = x[s];
          for (int kc = 0; kc < kn; kc += LB) {
            T lv[LB];
#pragma unroll
            for (int q = 0; q < LB; ++q) lv[q] = (kc + q < lane && lane < kn) ? li[kc + q] : T(0);
#pragma unroll
            for (int q = 0; q < LB; ++q) {
              const int kk = kc + q;
              if (kk < kn) {
                const T xk = readlane_t(xv, kk);
                if (kk < lane && lane < kn) {
                  const T e = lv[q] * xk;
                  xv = xv - e;
                }
              }
            }
          }
          x[s] = xv;
          if (lane < kn) xs[lane] = xv;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      const int i = t + NT * s;
      if (i >= k0 + kn && i < n) {
        const T *li = lu + (int64_t)i * lda + k0;
        T xv = x[s];
        for (int kc = 0; kc < kn; kc += LB) {
          T lv[LB];
#pragma unroll
          for (int q = 0; q < LB; ++q) lv[q] = kc + q < kn ? li[kc + q] : T(0);
#pragma unroll
          for (int q = 0; q < LB; ++q) {
            if (kc + q < kn) {
              const T e = lv[q] * xs[kc + q];
              xv = xv - e;
            }
          }
        }
        x[s] = xv;
      }
    }
    __syncthreads();
  }

  // Back substitution with U.
  if (n <= kExactOrderMax) {
    if (t < n) xall[t] = x[0];
    __syncthreads();
    if (wave == 0) {
      for (int i = n - 1; i >= 0; --i) {
        const T *ui = lu + (int64_t)i * lda;
        const int c0 = i + 1 + lane, c1 = i + 65 + lane;
        const T u0 = c0 < n ? ui[c0] : T(0);
        const T u1 = c1 < n ? ui[c1] : T(0);
        T xi = xall[i];
        for (int k = i + 1; k < n; ++k) {
          const int q = k - i - 1;
          const T u = q < 64 ? readlane_t(u0, q) : readlane_t(u1, q - 64);
          const T e = u * xall[k];
          xi = xi - e;
        }
        const T xn = cplx_div(xi, ui[i]);
        if (lane == 0) xall[i] = xn;
      }
    }
    __syncthreads();
    if (t < n) x[0] = xall[t];
  } else {
    for (int bk = nblk - 1; bk >= 0; --bk) {
      const int k0 = bk * 64, kn = min(64, n - k0);
      if (wave == bk % 16) {
        const int os = k0 / NT;
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
          if (s == os) {
            const int i = k0 + lane;
            const T *ui = lu + (int64_t)(lane < kn ? i : k0) * lda + k0;
            const T dii = lane < kn ? ui[lane] : T(1);
            T xv = x[s];
            for (int kc = kn - 1; kc >= 0; kc -= LB) {
              T uv[LB];
#pragma unroll
              for (int q = 0; q < LB; ++q) {
                const int kk = kc - q;
                uv[q] = (kk >= 0 && kk > lane && lane < kn) ? ui[kk] : T(0);
              }
#pragma unroll
              for (int q = 0; q < LB; ++q) {
                const int kk = kc - q;
                if (kk >= 0) {
                  if (lane == kk) xv = cplx_div(xv, dii);
                  const T xk = readlane_t(xv, kk);
                  if (lane < kk) {
                    const T e = uv[q] * xk;
                    xv = xv - e;
                  }
                }
              }
            }
            x[s] = xv;
            if (lane < kn) xs[lane] = xv;
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        const int i = t + NT * s;
        if (i < k0) {
          const T *ui = lu + (int64_t)i * lda + k0;
          T xv = x[s];
          for (int kc = 0; kc < kn; kc += LB) {
            T uv[LB];
#pragma unroll
            for (int q = 0; q < LB; ++q) uv[q] = kc + q < kn ? ui[kc + q] : T(0);
#pragma unroll
            for (int q = 0; q < LB; ++q) {
              if (kc + q < kn) {
                const T e = uv[q] * xs[kc + q];
                xv = xv - e;
              }
            }
          }
          x[s] = xv;
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int s = 0; s < RPT; ++s) {
    const int i = t + NT * s;
    if (i < n) b[(int64_t)i * ldb + r] = x[s];
  }
}

template <typename T>
hipError_t launch_cgetrs(lair_context *ctx, int n, int nrhs, const T *lu, int64_t lda, const int *perm, T *b,
                         int64_t ldb) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  if (n > kComplexGetrsMax) return hipErrorInvalidValue;
  const int slot = prof_begin(ctx, LAIR_KCLASS_GETRS);
  const dim3 grid(nrhs), block(1024);
  hipStream_t st = ctx->stream;
  if (n <= 1024) cgetrs_kernel<T, 1><<<grid, block, 0, st>>>(n, lu, lda, perm, b, ldb);
  else if (n <= 2048) cgetrs_kernel<T, 2><<<grid, block, 0, st>>>(n, lu, lda, perm, b, ldb);
  else if (n <= 4096) cgetrs_kernel<T, 4><<<grid, block, 0, st>>>(n, lu, lda, perm, b, ldb);
  else if (n <= 8192) cgetrs_kernel<T, 8><<<grid, block, 0, st>>>(n, lu, lda, perm, b, ldb);
  else cgetrs_kernel<T, 16><<<grid, block, 0, st>>>(n, lu, lda, perm, b, ldb);
  prof_end(ctx, slot, LAIR_KCLASS_GETRS, 8.0 * (double)n * n * nrhs,
           (double)sizeof(T) * ((double)n * n * nrhs + 2.0 * n * nrhs));
  return hipGetLastError();
}

template hipError_t launch_cgemm_minus<zcplx>(lair_context *, int, int, int, int, const zcplx *, int64_t,
                                              const zcplx *, int64_t, zcplx *, int64_t, int, int);
template hipError_t launch_cgemm_minus<ccplx>(lair_context *, int, int, int, int, const ccplx *, int64_t,
                                              const ccplx *, int64_t, ccplx *, int64_t, int, int);
template hipError_t launch_cgetrs<zcplx>(lair_context *, int, int, const zcplx *, int64_t, const int *, zcplx *,
                                         int64_t);
template hipError_t launch_cgetrs<ccplx>(lair_context *, int, int, const ccplx *, int64_t, const int *, ccplx *,
                                         int64_t);

}  // namespace lair
