// Trailing-matrix update C -= A * B on gfx950 matrix cores (src/blas/gemm.rs:6-32
// with alpha = -1, as called from src/lapack/getrf.rs:289-296), tiled for the
// blocked LU: C is the trailing matrix (row-major, large M and N) and K is the
// panel width (64..256).
//
// One workgroup of four waves computes a 128 x 128 tile of C; each wave owns
// a 64 x 64 quadrant as 4 x 4 blocks of 16 x 16 MFMA accumulators.  K is
// consumed in stages held in LDS; while a stage is multiplied, the next one is
// fetched from global memory into registers (16-byte loads), then written to
// the other LDS buffer.  The C tile is read, updated and written once, after
// the last stage, so each launch moves C through HBM exactly once.
//
// Operand lane layouts (16 x 16 x K MFMAs): lane l supplies A[l%16][k-group
// l/16] and B[k-group l/16][l%16]; FP64 16x16x4 returns C[l/16 + 4r][l%16],
// FP32 16x16x4 and BF16 16x16x32 return C[4(l/16) + r][l%16].
#include "lair_internal.h"

namespace lair {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kTileM = 128;
constexpr int kTileN = 128;

// Workgroup -> tile mapping.  Consecutive workgroups are dispatched to
// different XCDs (round-robin over 8); grouping tiles so that the 8 tiles
// sharing an A row-panel run on one XCD lets that XCD's L2 serve the panel.
__device__ __forceinline__ void tile_coords(int nbm, int nbn, int &tm, int &tn) {
  const int id = blockIdx.x;
  const int total = nbm * nbn;
  int lin = id;
  if (total % 8 == 0) {
    const int xcd = id % 8, slot = id / 8;
    lin = xcd * (total / 8) + slot;
  }
  tm = lin / nbn;
  tn = lin % nbn;
}

// ------------------------------------------------------------------ FP64 --

template <bool TRANSA>
__global__ void __launch_bounds__(256)
    dgemm_tile_kernel(int M, int N, int K, const double *__restrict__ A, int64_t lda,
                      const double *__restrict__ B, int64_t ldb, double *__restrict__ C, int64_t ldc) {
  constexpr int BK = 16;
  constexpr int PA = kTileM + 4;  // padded row of the k-major A stage
  constexpr int PB = kTileN + 4;
  __shared__ double As[2][BK][PA];  // As[k][i]
  __shared__ double Bs[2][BK][PB];  // Bs[k][j]
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int nbm = (M + kTileM - 1) / kTileM, nbn = (N + kTileN - 1) / kTileN;
  int tm, tn;
  tile_coords(nbm, nbn, tm, tn);
  const int bm = tm * kTileM, bn = tn * kTileN;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;

  // Global fetch mapping: A (not transposed): thread -> row t/2, k (t%2)*8..+8;
  // A transposed (stored K x M): thread -> k t/16, rows (t%16)*8..+8;
  // B: thread -> k t/16, cols (t%16)*8..+8.
  double ra[8], rb[8];
  auto fetch = [&](int k0) {
    if (!TRANSA) {
      const int i = t >> 1, kk = (t & 1) * 8;
      const int gi = bm + i;
#pragma unroll
      for (int q = 0; q < 8; q += 2) {
        const int gk = k0 + kk + q;
        if (gi < M && gk + 1 < K) {
          const d2 v = *reinterpret_cast<const d2 *>(A + (int64_t)gi * lda + gk);
          ra[q] = v[0];
          ra[q + 1] = v[1];
        } else {
          ra[q] = (gi < M && gk < K) ? A[(int64_t)gi * lda + gk] : 0.0;
          ra[q + 1] = (gi < M && gk + 1 < K) ? A[(int64_t)gi * lda + gk + 1] : 0.0;
        }
      }
    } else {
      const int kk = t >> 4, i0 = (t & 15) * 8;
      const int gk = k0 + kk;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int gi = bm + i0 + q;
        ra[q] = (gk < K && gi < M) ? A[(int64_t)gk * lda + gi] : 0.0;
      }
    }
    const int kb = t >> 4, j0 = (t & 15) * 8;
    const int gk = k0 + kb;
#pragma unroll
    for (int q = 0; q < 8; q += 2) {
      const int gj = bn + j0 + q;
      if (gk < K && gj + 1 < N) {
        const d2 v = *reinterpret_cast<const d2 *>(B + (int64_t)gk * ldb + gj);
        rb[q] = v[0];
        rb[q + 1] = v[1];
      } else {
        rb[q] = (gk < K && gj < N) ? B[(int64_t)gk * ldb + gj] : 0.0;
        rb[q + 1] = (gk < K && gj + 1 < N) ? B[(int64_t)gk * ldb + gj + 1] : 0.0;
      }
    }
  };
  auto stash = [&](int buf) {
    if (!TRANSA) {
      const int i = t >> 1, kk = (t & 1) * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) As[buf][kk + q][i] = ra[q];
    } else {
      const int kk = t >> 4, i0 = (t & 15) * 8;
#pragma unroll
      for (int q = 0; q < 8; q += 2) *reinterpret_cast<d2 *>(&As[buf][kk][i0 + q]) = (d2){ra[q], ra[q + 1]};
    }
    const int kb = t >> 4, j0 = (t & 15) * 8;
#pragma unroll
    for (int q = 0; q < 8; q += 2) *reinterpret_cast<d2 *>(&Bs[buf][kb][j0 + q]) = (d2){rb[q], rb[q + 1]};
  };

  d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};

  const int nstage = (K + BK - 1) / BK;
  fetch(0);
  stash(0);
  __syncthreads();
  for (int s = 0; s < nstage; ++s) {
    const int buf = s & 1;
    if (s + 1 < nstage) fetch((s + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = As[buf][kk + (lane >> 4)][wm + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[buf][kk + (lane >> 4)][wn + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    if (s + 1 < nstage) stash(buf ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
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

hipError_t launch_dgemm_tile(lair_context *ctx, int M, int N, int K, const double *A, int64_t lda,
                             bool transA, const double *B, int64_t ldb, double *C, int64_t ldc) {
  const int nbm = (M + kTileM - 1) / kTileM, nbn = (N + kTileN - 1) / kTileN;
  const unsigned grid = (unsigned)(nbm * nbn);
  if (transA)
    dgemm_tile_kernel<true><<<grid, 256, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C, ldc);
  else
    dgemm_tile_kernel<false><<<grid, 256, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C, ldc);
  return hipGetLastError();
}

}  // namespace lair
