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
#include <algorithm>

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
__device__ __forceinline__ void tile_coords(int id, int nbm, int nbn, int &tm, int &tn) {
  const int total = nbm * nbn;
  int lin = id;
  if (total % 8 == 0) {
    const int xcd = id % 8, slot = id / 8;
    lin = xcd * (total / 8) + slot;
  }
  tm = lin / nbn;
  tn = lin % nbn;
}

// ------------------------------------------------------- FP64 and FP32 --

template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
  typedef d4 acc_t;
  typedef d2 pair_t;
  static __device__ __forceinline__ d4 op(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of accumulator element r of lane l within the 16 x 16 block
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma16<float> {
  typedef f4 acc_t;
  typedef float pair_t __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ f4 op(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// STORE = false: C -= op(A) B (trailing update); STORE = true: C = -op(A) B
// (C is only written, e.g. the adjoint's A-gradient).
//
// TPW > 1 (capped grids): a workgroup of 256 * TPW threads computes TPW
// horizontally adjacent tiles at once (one group of four waves each, own LDS
// stages), so one workgroup per CU has the occupancy of TPW workgroups.
template <typename T, bool TRANSA, bool STORE, int TPW>
__global__ void __launch_bounds__(256 * TPW, TPW == 1 ? 2 : 1)
    gemm_tile_kernel(int M, int N, int K, const T *__restrict__ A, int64_t lda,
                     const T *__restrict__ B, int64_t ldb, T *__restrict__ C, int64_t ldc) {
  typedef typename Mfma16<T>::acc_t acc_t;
  typedef typename Mfma16<T>::pair_t d2;
  constexpr int BK = 16;
  constexpr int PA = BK + 1;      // padded row of the A stage (16 lanes read 16 rows)
  constexpr int PB = kTileN + 4;
  __shared__ T As_all[TPW][2][kTileM][PA];  // As[i][k]
  __shared__ T Bs_all[TPW][2][BK][PB];      // Bs[k][j]
  const int h = threadIdx.x >> 8;
  auto &As = As_all[h];
  auto &Bs = Bs_all[h];
  const int t = threadIdx.x & 255, lane = t & 63, wave = t >> 6;
  const int nbm = (M + kTileM - 1) / kTileM, nbn = (N + kTileN - 1) / kTileN;
  const int nbnp = (nbn + TPW - 1) / TPW;
  const int ntile = nbm * nbnp;
  // A capped grid loops over the tiles (same XCD grouping: gridDim.x is a
  // multiple of 8 whenever it is below the tile count).  A tile past the
  // right edge (tn >= nbn) has every load and store masked.
  for (int id = blockIdx.x; id < ntile; id += gridDim.x) {
  int tm, tnp;
  tile_coords(id, nbm, nbnp, tm, tnp);
  const int tn = tnp * TPW + h;
  const int bm = tm * kTileM, bn = tn * kTileN;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;

  // Global fetch mapping: A (not transposed): thread -> row t/2, k (t%2)*8..+8;
  // A transposed (stored K x M): thread -> k t/16, rows (t%16)*8..+8;
  // B: thread -> k t/16, cols (t%16)*8..+8.
  T ra[8], rb[8];
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
          ra[q] = (gi < M && gk < K) ? A[(int64_t)gi * lda + gk] : T(0);
          ra[q + 1] = (gi < M && gk + 1 < K) ? A[(int64_t)gi * lda + gk + 1] : T(0);
        }
      }
    } else {
      const int kk = t >> 4, i0 = (t & 15) * 8;
      const int gk = k0 + kk;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int gi = bm + i0 + q;
        ra[q] = (gk < K && gi < M) ? A[(int64_t)gk * lda + gi] : T(0);
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
        rb[q] = (gk < K && gj < N) ? B[(int64_t)gk * ldb + gj] : T(0);
        rb[q + 1] = (gk < K && gj + 1 < N) ? B[(int64_t)gk * ldb + gj + 1] : T(0);
      }
    }
  };
  auto stash = [&](int buf) {
    if (!TRANSA) {
      const int i = t >> 1, kk = (t & 1) * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) As[buf][i][kk + q] = ra[q];
    } else {
      const int kk = t >> 4, i0 = (t & 15) * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) As[buf][i0 + q][kk] = ra[q];
    }
    const int kb = t >> 4, j0 = (t & 15) * 8;
#pragma unroll
    for (int q = 0; q < 8; q += 2) *reinterpret_cast<d2 *>(&Bs[buf][kb][j0 + q]) = (d2){rb[q], rb[q + 1]};
  };

  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (acc_t){T(0), T(0), T(0), T(0)};

  const int nstage = (K + BK - 1) / BK;
  fetch(0);
  stash(0);
  __syncthreads();
  for (int s = 0; s < nstage; ++s) {
    const int buf = s & 1;
    if (s + 1 < nstage) fetch((s + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = As[buf][wm + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[buf][kk + (lane >> 4)][wn + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = Mfma16<T>::op(av[i], bv[j], acc[i][j]);
    }
    if (s + 1 < nstage) stash(buf ^ 1);
    __syncthreads();
  }

  // Epilogue: per 16-row block, all 16 loads of C are issued before any store.
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    T cv[4][4];
    if (!STORE) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = bm + wm + i * 16 + Mfma16<T>::row(lane, r);
          const int gj = bn + wn + j * 16 + (lane & 15);
          cv[j][r] = (gi < M && gj < N) ? C[(int64_t)gi * ldc + gj] : T(0);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = bm + wm + i * 16 + Mfma16<T>::row(lane, r);
        const int gj = bn + wn + j * 16 + (lane & 15);
        if (gi < M && gj < N) C[(int64_t)gi * ldc + gj] = STORE ? -acc[i][j][r] : cv[j][r] - acc[i][j][r];
      }
  }
  }
}

hipError_t launch_dgemm_tile(lair_context *ctx, int M, int N, int K, const double *A, int64_t lda,
                             bool transA, const double *B, int64_t ldb, double *C, int64_t ldc,
                             bool store) {
  const int nbm = (M + kTileM - 1) / kTileM, nbn = (N + kTileN - 1) / kTileN;
  hipStream_t st = ctx->stream;
  if (ctx->bulk_wg_cap > 0 && !transA && !store) {
    const int ntile = nbm * ((nbn + 1) / 2);
    const unsigned grid = (unsigned)std::min(ntile, ctx->bulk_wg_cap);
    gemm_tile_kernel<double, false, false, 2><<<grid, 512, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
    return hipGetLastError();
  }
  const unsigned grid = (unsigned)(nbm * nbn);
  if (store) {
    if (transA)
      gemm_tile_kernel<double, true, true, 1><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
    else
      gemm_tile_kernel<double, false, true, 1><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
  } else {
    if (transA)
      gemm_tile_kernel<double, true, false, 1><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
    else
      gemm_tile_kernel<double, false, false, 1><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
  }
  return hipGetLastError();
}

hipError_t launch_sgemm_tile(lair_context *ctx, int M, int N, int K, const float *A, int64_t lda,
                             bool transA, const float *B, int64_t ldb, float *C, int64_t ldc) {
  const int nbm = (M + kTileM - 1) / kTileM, nbn = (N + kTileN - 1) / kTileN;
  if (ctx->bulk_wg_cap > 0 && !transA) {
    const int ntile = nbm * ((nbn + 3) / 4);
    const unsigned grid = (unsigned)std::min(ntile, ctx->bulk_wg_cap);
    gemm_tile_kernel<float, false, false, 4><<<grid, 1024, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C,
                                                                              ldc);
    return hipGetLastError();
  }
  const unsigned grid = (unsigned)(nbm * nbn);
  if (transA)
    gemm_tile_kernel<float, true, false, 1><<<grid, 256, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C, ldc);
  else
    gemm_tile_kernel<float, false, false, 1><<<grid, 256, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C, ldc);
  return hipGetLastError();
}

// ------------------------------------------------------- BF16 splits ------

// Splits x into NS bf16 terms whose sum carries 8*NS significand bits; each
// remainder x - (t0 + ... + ts) is exact in FP32.
template <int NS>
__device__ __forceinline__ void split3(float x, __bf16 *out) {
  float r = x;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const __bf16 h = (__bf16)r;
    out[s] = h;
    r = r - (float)h;
  }
}

// C -= A B with FP32 operands split into NS bf16 terms.  NS = 3 keeps the six
// products whose magnitude is at least 2^-16 of the leading one (hh, hm, mh,
// mm, hl, lh), issued smallest first into the FP32 MFMA accumulators.
// A is M x K row-major; B is K x N row-major.  Stage: BK = 32 (one
// 16x16x32 MFMA step).  LDS holds the split planes [s][i][k] of A and
// [s][j][k] of B (k contiguous: each lane reads 8 consecutive k); the next
// stage's raw FP32 values are prefetched into registers during the MFMAs.
template <int NS, int TPW>
__global__ void __launch_bounds__(256 * TPW, TPW == 1 ? 2 : 1)
    bgemm_tile_kernel(int M, int N, int K, const float *__restrict__ A, int64_t lda,
                      const float *__restrict__ B, int64_t ldb, float *__restrict__ C, int64_t ldc) {
  constexpr int BK = 32;
  constexpr int PK = BK + 8;  // padded k extent (bf16 elements)
  __shared__ __bf16 As_all[TPW][NS][kTileM][PK];
  __shared__ __bf16 Bs_all[TPW][NS][kTileN][PK];
  const int h = threadIdx.x >> 8;
  auto &As = As_all[h];
  auto &Bs = Bs_all[h];
  const int t = threadIdx.x & 255, lane = t & 63, wave = t >> 6;
  const int nbm = (M + kTileM - 1) / kTileM, nbn = (N + kTileN - 1) / kTileN;
  const int nbnp = (nbn + TPW - 1) / TPW;
  const int ntile = nbm * nbnp;
  // A capped grid loops over the tiles (same XCD grouping: gridDim.x is a
  // multiple of 8 whenever it is below the tile count).
  for (int id = blockIdx.x; id < ntile; id += gridDim.x) {
  int tm, tnp;
  tile_coords(id, nbm, nbnp, tm, tnp);
  const int tn = tnp * TPW + h;
  const int bm = tm * kTileM, bn = tn * kTileN;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  // A: thread -> row t/2, k (t%2)*16 .. +16.  B: thread -> j 4*(t%32) .. +4,
  // k rows 2*(t/32) + {0, 1} and + 16.
  float ra[16], rb[16];
  auto fetch = [&](int k0) {
    const int i = t >> 1, kk = (t & 1) * 16;
    const int gi = bm + i;
#pragma unroll
    for (int q = 0; q < 16; q += 4) {
      const int gk = k0 + kk + q;
      if (gi < M && gk + 3 < K) {
        const f4 v = *reinterpret_cast<const f4 *>(A + (int64_t)gi * lda + gk);
        ra[q] = v[0];
        ra[q + 1] = v[1];
        ra[q + 2] = v[2];
        ra[q + 3] = v[3];
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) ra[q + e] = (gi < M && gk + e < K) ? A[(int64_t)gi * lda + gk + e] : 0.f;
      }
    }
    const int j0 = (t & 31) * 4, kp = (t >> 5) * 2;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int gk = k0 + kp + e + 16 * h;
        const int gj = bn + j0;
        float *dst = rb + (h * 2 + e) * 4;
        if (gk < K && gj + 3 < N) {
          const f4 v = *reinterpret_cast<const f4 *>(B + (int64_t)gk * ldb + gj);
          dst[0] = v[0];
          dst[1] = v[1];
          dst[2] = v[2];
          dst[3] = v[3];
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) dst[q] = (gk < K && gj + q < N) ? B[(int64_t)gk * ldb + gj + q] : 0.f;
        }
      }
  };
  auto stash = [&]() {
    const int i = t >> 1, kk = (t & 1) * 16;
#pragma unroll
    for (int q = 0; q < 16; q += 8) {
      bf8 part[NS];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        __bf16 sp[NS];
        split3<NS>(ra[q + e], sp);
#pragma unroll
        for (int s = 0; s < NS; ++s) part[s][e] = sp[s];
      }
#pragma unroll
      for (int s = 0; s < NS; ++s) *reinterpret_cast<bf8 *>(&As[s][i][kk + q]) = part[s];
    }
    const int j0 = (t & 31) * 4, kp = (t >> 5) * 2;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        __bf16 s0[NS], s1[NS];
        split3<NS>(rb[(h * 2 + 0) * 4 + q], s0);
        split3<NS>(rb[(h * 2 + 1) * 4 + q], s1);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
          *reinterpret_cast<bf2 *>(&Bs[s][j0 + q][kp + 16 * h]) = (bf2){s0[s], s1[s]};
        }
      }
  };

  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};

  const int nstage = (K + BK - 1) / BK;
  fetch(0);
  for (int s = 0; s < nstage; ++s) {
    __syncthreads();
    stash();
    __syncthreads();
    if (s + 1 < nstage) fetch((s + 1) * BK);
    bf8 av[NS][4], bv[NS][4];
#pragma unroll
    for (int p = 0; p < NS; ++p) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        av[p][i] = *reinterpret_cast<const bf8 *>(&As[p][wm + i * 16 + (lane & 15)][(lane >> 4) * 8]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        bv[p][j] = *reinterpret_cast<const bf8 *>(&Bs[p][wn + j * 16 + (lane & 15)][(lane >> 4) * 8]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (NS == 3) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[2][i], bv[0][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0][i], bv[2][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[1][i], bv[1][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[1][i], bv[0][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0][i], bv[1][j], acc[i][j], 0, 0, 0);
        }
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0][i], bv[0][j], acc[i][j], 0, 0, 0);
      }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float cv[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = bm + wm + i * 16 + 4 * (lane >> 4) + r;
        const int gj = bn + wn + j * 16 + (lane & 15);
        cv[j][r] = (gi < M && gj < N) ? C[(int64_t)gi * ldc + gj] : 0.f;
      }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = bm + wm + i * 16 + 4 * (lane >> 4) + r;
        const int gj = bn + wn + j * 16 + (lane & 15);
        if (gi < M && gj < N) C[(int64_t)gi * ldc + gj] = cv[j][r] - acc[i][j][r];
      }
  }
  }
}

hipError_t launch_bgemm_tile(lair_context *ctx, int ns, int M, int N, int K, const float *A,
                             int64_t lda, const float *B, int64_t ldb, float *C, int64_t ldc) {
  const int nbm = (M + kTileM - 1) / kTileM, nbn = (N + kTileN - 1) / kTileN;
  if (ctx->bulk_wg_cap > 0) {
    const int ntile = nbm * ((nbn + 1) / 2);
    const unsigned grid = (unsigned)std::min(ntile, ctx->bulk_wg_cap);
    if (ns == 3)
      bgemm_tile_kernel<3, 2><<<grid, 512, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C, ldc);
    else
      bgemm_tile_kernel<1, 2><<<grid, 512, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C, ldc);
    return hipGetLastError();
  }
  const unsigned grid = (unsigned)(nbm * nbn);
  if (ns == 3)
    bgemm_tile_kernel<3, 1><<<grid, 256, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C, ldc);
  else
    bgemm_tile_kernel<1, 1><<<grid, 256, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, C, ldc);
  return hipGetLastError();
}

}  // namespace lair
