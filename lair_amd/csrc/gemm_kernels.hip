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

// Raw buffer access: a wave-uniform resource (scalar registers), a 32-bit
// per-lane byte offset and a wave-uniform byte offset.  Keeps the many
// elements of a tile at one vector register of address.
typedef unsigned int u2v __attribute__((ext_vector_type(2)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
constexpr int kBufWord3 = 0x00020000;  // raw buffer, 32-bit data format (gfx9)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, 0x7fffffff, kBufWord3);
}
template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <typename T>
__device__ __forceinline__ void buf_st(T x, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2v, x), r, voff, soff, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, voff, soff, 0);
}
// Two consecutive elements (8 or 16 bytes).
template <typename T, typename P2>
__device__ __forceinline__ P2 buf_ld2(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(P2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  else
    return __builtin_bit_cast(P2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// STORE = false: C -= op(A) B (trailing update); STORE = true: C = -op(A) B
// (C is only written, e.g. the adjoint's A-gradient).
//
// The accumulators start from C (or zero) and the MFMAs add (-A) B: A is
// negated as it is staged in LDS.  So C is read at the start of a tile, its
// latency overlapping the first stages, and each accumulator is stored once at
// the end with no second copy of C in registers.
//
// k order inside a stage of BK = 16: MFMA step s (0..3) takes k = 4 g + s
// from the lane group g = lane / 16 (the MFMA's k index within the step), so
// a lane's four k values of one 16 x 16 block are contiguous in LDS: A is
// staged as [row][k] and B transposed as [col][k], and each block costs one
// (FP32) or two (FP64) 16-byte LDS reads per lane.  The row pitch P (20
// floats, 18 doubles) puts the 16 rows that 16 lanes read at once on
// distinct 16-byte bank groups.
//
// Memory access goes through raw buffer resources based at the tile (scalar
// registers) with one 32-bit lane offset and uniform element offsets, so the
// 64 C elements of a lane cost no address registers.  Interior tiles with K a
// multiple of BK run unmasked; the others (MASKED) load from clamped, always
// valid addresses and select zero outside the matrix.
constexpr int kBK = 16;

template <typename T, bool TRANSA, bool STORE, bool MASKED>
__device__ __forceinline__ void gemm_tile(int M, int N, int K, const T *__restrict__ A, int64_t lda,
                                          const T *__restrict__ B, int64_t ldb, T *__restrict__ C, int64_t ldc,
                                          int bm, int bn, bool ghost,
                                          T (*As)[kTileM][kBK + (sizeof(T) == 4 ? 4 : 2)],
                                          T (*Bs)[kTileN][kBK + (sizeof(T) == 4 ? 4 : 2)]) {
  typedef typename Mfma16<T>::acc_t acc_t;
  typedef typename Mfma16<T>::pair_t d2;
  typedef T v16 __attribute__((ext_vector_type(16 / sizeof(T))));  // 16 bytes
  constexpr int BK = kBK;
  constexpr int VW = 16 / sizeof(T);  // elements per 16-byte access
  const int t = threadIdx.x & 255, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane((threadIdx.x & 255) >> 6);
  const int wm = (wv >> 1) * 64, wn = (wv & 1) * 64;

  // C: this lane's element (i, j, r) is at row i*16 + r*rstep, column j*16
  // of the wave's quadrant, plus the lane offset.
  const int lr0 = Mfma16<T>::row(lane, 0), rstep = Mfma16<T>::row(0, 1);
  const int mrem = M - (bm + wm + lr0), nrem = N - (bn + wn + (lane & 15));
  const int coff = (lr0 * (int)ldc + (lane & 15)) * (int)sizeof(T);
  auto csoff = [&](int i, int j, int r) { return ((i * 16 + r * rstep) * (int)ldc + j * 16) * (int)sizeof(T); };
  const __amdgpu_buffer_rsrc_t crs = buf_rsrc(C + (int64_t)(bm + wm) * ldc + (bn + wn));
  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (STORE) {
          acc[i][j][r] = T(0);
        } else if (!MASKED) {
          acc[i][j][r] = buf_ld<T>(crs, coff, csoff(i, j, r));
        } else {
          // Clamped into the matrix: an element outside it is never stored,
          // so what it accumulates does not matter.
          const int gi = min(bm + wm + lr0 + i * 16 + r * rstep, M - 1);
          const int gj = min(bn + wn + (lane & 15) + j * 16, N - 1);
          acc[i][j][r] = C[(int64_t)gi * ldc + gj];
        }
      }

  // Global fetch mapping: A (not transposed): thread -> row t/2, k (t%2)*8..+8;
  // A transposed (stored K x M): thread -> k t/16, rows (t%16)*8..+8;
  // B: thread -> column t%128, k (t/128)*8..+8 (coalesced along the row).
  T ra[8], rb[8];
  auto fetch = [&](int k0) {
    const int j = t & 127, kb = (t >> 7) * 8;
    if (!MASKED) {
      const __amdgpu_buffer_rsrc_t brs = buf_rsrc(B + (int64_t)k0 * ldb + bn);
      const int bvo = (kb * (int)ldb + j) * (int)sizeof(T);
#pragma unroll
      for (int q = 0; q < 8; ++q) rb[q] = buf_ld<T>(brs, bvo, q * (int)ldb * (int)sizeof(T));
      if (!TRANSA) {
        const int i = t >> 1, kk = (t & 1) * 8;
        const __amdgpu_buffer_rsrc_t ars = buf_rsrc(A + (int64_t)bm * lda + k0);
        const int avo = (i * (int)lda + kk) * (int)sizeof(T);
#pragma unroll
        for (int q = 0; q < 8; q += 2) {
          const d2 v = buf_ld2<T, d2>(ars, avo, q * (int)sizeof(T));
          ra[q] = v[0];
          ra[q + 1] = v[1];
        }
      } else {
        const int kk = t >> 4, i0 = (t & 15) * 8;
        const __amdgpu_buffer_rsrc_t ars = buf_rsrc(A + (int64_t)k0 * lda + bm);
        const int avo = (kk * (int)lda + i0) * (int)sizeof(T);
#pragma unroll
        for (int q = 0; q < 8; ++q) ra[q] = buf_ld<T>(ars, avo, q * (int)sizeof(T));
      }
    } else {
      // Clamped indices keep every load inside the matrix.  Rows of A and
      // columns of B outside it only reach accumulators that are never
      // stored; k beyond K is zeroed in both operands.
      const int cj = min(bn + j, N - 1);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int gk = k0 + kb + q;
        const T v = B[(int64_t)min(gk, K - 1) * ldb + cj];
        rb[q] = gk < K ? v : T(0);
      }
      if (!TRANSA) {
        const int kk = (t & 1) * 8;
        const T *ar = A + (int64_t)min(bm + (t >> 1), M - 1) * lda;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int gk = k0 + kk + q;
          const T v = ar[min(gk, K - 1)];
          ra[q] = gk < K ? v : T(0);
        }
      } else {
        const int gk = k0 + (t >> 4), i0 = (t & 15) * 8;
        const T *ak = A + (int64_t)min(gk, K - 1) * lda;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const T v = ak[min(bm + i0 + q, M - 1)];
          ra[q] = gk < K ? v : T(0);
        }
      }
    }
  };
  auto stash = [&](int buf) {
    if (!TRANSA) {
      const int i = t >> 1, kk = (t & 1) * 8;
#pragma unroll
      for (int q = 0; q < 8; q += VW) {
        v16 v;
#pragma unroll
        for (int e = 0; e < VW; ++e) v[e] = -ra[q + e];
        *reinterpret_cast<v16 *>(&As[buf][i][kk + q]) = v;
      }
    } else {
      const int kk = t >> 4, i0 = (t & 15) * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) As[buf][i0 + q][kk] = -ra[q];
    }
    const int j = t & 127, kb = (t >> 7) * 8;
#pragma unroll
    for (int q = 0; q < 8; q += VW) {
      v16 v;
#pragma unroll
      for (int e = 0; e < VW; ++e) v[e] = rb[q + e];
      *reinterpret_cast<v16 *>(&Bs[buf][j][kb + q]) = v;
    }
  };

  const int nstage = (K + BK - 1) / BK;
  const int kq = 4 * (lane >> 4);  // this lane's first k within a stage
  fetch(0);
  stash(0);
  __syncthreads();
  for (int s = 0; s < nstage; ++s) {
    const int buf = s & 1;
    if (s + 1 < nstage) fetch((s + 1) * BK);
    // Two column blocks of B at a time, A re-read per pair.
#pragma unroll
    for (int jp = 0; jp < 4; jp += 2) {
      T bq[2][4];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const T *src = &Bs[buf][wn + (jp + j) * 16 + (lane & 15)][kq];
#pragma unroll
        for (int q = 0; q < 4; q += VW) {
          const v16 v = *reinterpret_cast<const v16 *>(src + q);
#pragma unroll
          for (int e = 0; e < VW; ++e) bq[j][q + e] = v[e];
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        T aq[4];
        const T *src = &As[buf][wm + i * 16 + (lane & 15)][kq];
#pragma unroll
        for (int q = 0; q < 4; q += VW) {
          const v16 v = *reinterpret_cast<const v16 *>(src + q);
#pragma unroll
          for (int e = 0; e < VW; ++e) aq[q + e] = v[e];
        }
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][jp + j] = Mfma16<T>::op(aq[ks], bq[j][ks], acc[i][jp + j]);
      }
    }
    if (s + 1 < nstage) stash(buf ^ 1);
    __syncthreads();
  }

  // Offsets and bounds are recomputed from opaque copies rather than kept
  // (64 scalar registers, 64 masks) across the K loop.
  int ldc_e = (int)ldc, mrem_e = mrem, nrem_e = nrem;
  asm volatile("" : "+s"(ldc_e));
  asm volatile("" : "+v"(mrem_e), "+v"(nrem_e));
  if (ghost) return;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (!MASKED || (i * 16 + r * rstep < mrem_e && j * 16 < nrem_e))
          buf_st<T>(acc[i][j][r], crs, coff, ((i * 16 + r * rstep) * ldc_e + j * 16) * (int)sizeof(T));
      }
}

// TPW > 1 (capped grids): a workgroup of 256 * TPW threads computes TPW
// horizontally adjacent tiles at once (one group of four waves each, own LDS
// stages), so one workgroup per CU has the occupancy of TPW workgroups.
// MASKED = false: every tile lies inside M x N and K is a multiple of kBK; a
// group whose tile would lie past the right edge computes the last tile of
// its row again (keeping the workgroup's barriers) and stores nothing.
template <typename T, bool TRANSA, bool STORE, int TPW, bool MASKED>
__global__ void __launch_bounds__(256 * TPW, TPW == 1 ? 2 : 1)
    gemm_tile_kernel(int M, int N, int K, const T *__restrict__ A, int64_t lda,
                     const T *__restrict__ B, int64_t ldb, T *__restrict__ C, int64_t ldc) {
  constexpr int P = kBK + (sizeof(T) == 4 ? 4 : 2);
  __shared__ __align__(16) T As_all[TPW][2][kTileM][P];  // As[i][k] = -A[i][k]
  __shared__ __align__(16) T Bs_all[TPW][2][kTileN][P];  // Bs[j][k] = B[k][j]
  const int h = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8);
  const int nbm = (M + kTileM - 1) / kTileM, nbn = (N + kTileN - 1) / kTileN;
  const int nbnp = (nbn + TPW - 1) / TPW;
  const int ntile = nbm * nbnp;
  // A capped grid loops over the tiles (same XCD grouping: gridDim.x is a
  // multiple of 8 whenever it is below the tile count).
  for (int id = blockIdx.x; id < ntile; id += gridDim.x) {
    int tm, tnp;
    tile_coords(id, nbm, nbnp, tm, tnp);
    const int tn = tnp * TPW + h;
    const bool ghost = tn >= nbn;
    const int bm = tm * kTileM, bn = (ghost && !MASKED ? nbn - 1 : tn) * kTileN;
    gemm_tile<T, TRANSA, STORE, MASKED>(M, N, K, A, lda, B, ldb, C, ldc, bm, bn, ghost, As_all[h], Bs_all[h]);
  }
}

// Split-K products for tall-skinny C with a long K (the many-right-hand-side
// residual): workgroup (tile, y) writes W_y = -A[:, y Ks : (y+1) Ks] B[y Ks :
// (y+1) Ks, :] into its own slice W + y * wstride, so that every CU has work
// although C has few tiles; the caller sums the slices.
template <typename T>
__global__ void __launch_bounds__(256, 2)
    gemm_splitk_kernel(int M, int N, int K, int Ks, const T *__restrict__ A, int64_t lda, const T *__restrict__ B,
                       int64_t ldb, T *__restrict__ W, int64_t ldw, int64_t wstride) {
  constexpr int P = kBK + (sizeof(T) == 4 ? 4 : 2);
  __shared__ __align__(16) T As[2][kTileM][P];
  __shared__ __align__(16) T Bs[2][kTileN][P];
  const int y = blockIdx.y;
  const int k0 = y * Ks, kn = min(Ks, K - k0);
  if (kn <= 0) return;  // uniform over the workgroup
  const int nbn = (N + kTileN - 1) / kTileN;
  const int tm = blockIdx.x / nbn, tn = blockIdx.x % nbn;
  gemm_tile<T, false, true, true>(M, N, kn, A + k0, lda, B + (int64_t)k0 * ldb, ldb, W + (int64_t)y * wstride, ldw,
                                  tm * kTileM, tn * kTileN, false, As, Bs);
}

hipError_t launch_dgemm_splitk_store(lair_context *ctx, int M, int N, int K, int S, const double *A, int64_t lda,
                                     const double *B, int64_t ldb, double *W, int64_t ldw, int64_t wstride) {
  if (M <= 0 || N <= 0 || K <= 0 || S <= 0) return hipSuccess;
  const int Ks = ((K + S - 1) / S + kBK - 1) / kBK * kBK;
  const unsigned tiles = (unsigned)(((M + kTileM - 1) / kTileM) * ((N + kTileN - 1) / kTileN));
  gemm_splitk_kernel<double><<<dim3(tiles, (unsigned)S), 256, 0, ctx->stream>>>(M, N, K, Ks, A, lda, B, ldb, W, ldw,
                                                                                 wstride);
  return hipGetLastError();
}

// Splits C into the region of whole tiles (unmasked kernel, launched by
// `interior`) and the right and bottom strips (masked kernel, one tile per
// workgroup, launched by `masked`); K not a multiple of the stage depth runs
// masked throughout.  The 32-bit buffer offsets need 128 rows of C, A and B
// within 2 GiB.
template <typename T, bool TRANSA, typename FI, typename FM>
static hipError_t split_tiles(int M, int N, int K, int kstep, const T *A, int64_t lda, const T *B, int64_t ldb,
                              T *C, int64_t ldc, FI interior, FM masked) {
  const int64_t big = std::max(std::max(lda, ldb), ldc);
  if (big > ((int64_t)1 << 31) / ((int64_t)kTileM * (int64_t)sizeof(T))) return hipErrorInvalidValue;
  if (K % kstep != 0) {
    masked(M, N, A, B, C);
    return hipGetLastError();
  }
  const int Mf = M / kTileM * kTileM, Nf = N / kTileN * kTileN;
  if (Mf > 0 && Nf > 0) interior(Mf, Nf);
  // Right strip (rows of the interior), then the bottom strip (all columns).
  masked(Mf, N - Nf, A, B + Nf, C + Nf);
  masked(M - Mf, N, TRANSA ? A + Mf : A + (int64_t)Mf * lda, B, C + (int64_t)Mf * ldc);
  return hipGetLastError();
}

template <typename T, bool TRANSA, bool STORE, typename F>
static hipError_t gemm_split(lair_context *ctx, int M, int N, int K, const T *A, int64_t lda, const T *B,
                             int64_t ldb, T *C, int64_t ldc, F interior) {
  hipStream_t st = ctx->stream;
  auto masked = [&](int m, int n, const T *a, const T *b, T *c) {
    if (m <= 0 || n <= 0) return;
    const unsigned g = (unsigned)(((m + kTileM - 1) / kTileM) * ((n + kTileN - 1) / kTileN));
    gemm_tile_kernel<T, TRANSA, STORE, 1, true><<<g, 256, 0, st>>>(m, n, K, a, lda, b, ldb, c, ldc);
  };
  return split_tiles<T, TRANSA>(M, N, K, kBK, A, lda, B, ldb, C, ldc, interior, masked);
}

hipError_t launch_dgemm_tile(lair_context *ctx, int M, int N, int K, const double *A, int64_t lda,
                             bool transA, const double *B, int64_t ldb, double *C, int64_t ldc,
                             bool store) {
  hipStream_t st = ctx->stream;
  auto run = [&](auto ta, auto sto) {
    constexpr bool TA = decltype(ta)::value, STO = decltype(sto)::value;
    return gemm_split<double, TA, STO>(ctx, M, N, K, A, lda, B, ldb, C, ldc, [&](int m, int n) {
      const int nbm = m / kTileM, nbn = n / kTileN;
      if (ctx->bulk_wg_cap > 0 && !TA && !STO) {
        const int ntile = nbm * ((nbn + 1) / 2);
        const unsigned grid = (unsigned)std::min(ntile, ctx->bulk_wg_cap);
        gemm_tile_kernel<double, false, false, 2, false><<<grid, 512, 0, st>>>(m, n, K, A, lda, B, ldb, C, ldc);
      } else {
        gemm_tile_kernel<double, TA, STO, 1, false><<<(unsigned)(nbm * nbn), 256, 0, st>>>(m, n, K, A, lda, B, ldb,
                                                                                           C, ldc);
      }
    });
  };
  if (transA) return store ? run(std::true_type{}, std::true_type{}) : run(std::true_type{}, std::false_type{});
  return store ? run(std::false_type{}, std::true_type{}) : run(std::false_type{}, std::false_type{});
}

hipError_t launch_sgemm_tile(lair_context *ctx, int M, int N, int K, const float *A, int64_t lda,
                             bool transA, const float *B, int64_t ldb, float *C, int64_t ldc) {
  hipStream_t st = ctx->stream;
  auto run = [&](auto ta) {
    constexpr bool TA = decltype(ta)::value;
    return gemm_split<float, TA, false>(ctx, M, N, K, A, lda, B, ldb, C, ldc, [&](int m, int n) {
      const int nbm = m / kTileM, nbn = n / kTileN;
      if (ctx->bulk_wg_cap > 0 && !TA) {
        const int ntile = nbm * ((nbn + 3) / 4);
        const unsigned grid = (unsigned)std::min(ntile, ctx->bulk_wg_cap);
        gemm_tile_kernel<float, false, false, 4, false><<<grid, 1024, 0, st>>>(m, n, K, A, lda, B, ldb, C, ldc);
      } else {
        gemm_tile_kernel<float, TA, false, 1, false><<<(unsigned)(nbm * nbn), 256, 0, st>>>(m, n, K, A, lda, B, ldb,
                                                                                            C, ldc);
      }
    });
  };
  return transA ? run(std::true_type{}) : run(std::false_type{});
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
// mm, hl, lh), issued smallest first.  A is M x K row-major; B is K x N
// row-major.
//
// Same plan as gemm_tile: a 128 x 128 tile per group of four waves, the
// accumulators start from C and the MFMAs add (-A) B (A is negated as it is
// split), buffer addressing, unmasked interior tiles and masked edge strips.
// Each wave owns a 64 x 64 quadrant as 2 x 2 blocks of 32 x 32 accumulators
// (v_mfma_f32_32x32x16_bf16: lane l supplies A[l%32][8(l/32) .. +8] and
// B[8(l/32) .. +8][l%32] and holds C[8(r/4) + 4(l/32) + r%4][l%32]).  A stage
// is BK = 16 (one MFMA k-step); LDS holds the split planes of A as [i][k] and
// of B as [j][k], double buffered, at a pitch of 24 bf16 (48 bytes: the 16
// rows that 16 lanes read at once fall on distinct 16-byte bank groups).  The
// FP32 values of the next stage are fetched into registers during the MFMAs
// and split when they are written to the other buffer.
constexpr int kBBK = 16;
constexpr int kBP = kBBK + 8;

typedef float f16v __attribute__((ext_vector_type(16)));

template <int NS, int NSL, bool MASKED>
__device__ __forceinline__ void bgemm_tile(int M, int N, int K, const float *__restrict__ A, int64_t lda,
                                           const float *__restrict__ B, int64_t ldb, float *__restrict__ C,
                                           int64_t ldc, int bm, int bn, bool ghost,
                                           __bf16 (*As)[NSL][kTileM][kBP], __bf16 (*Bs)[NSL][kTileN][kBP]) {
  const int t = threadIdx.x & 255, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane((threadIdx.x & 255) >> 6);
  const int wm = (wv >> 1) * 64, wn = (wv & 1) * 64;

  // C element (i, j, r) of this lane: quadrant row i*32 + 8(r/4) + r%4 past
  // the lane's first row 4(l/32), column j*32 past the lane's column l%32.
  const int lr0 = 4 * (lane >> 5), lc = lane & 31;
  const int mrem = M - (bm + wm + lr0), nrem = N - (bn + wn + lc);
  const int coff = (lr0 * (int)ldc + lc) * 4;
  auto crow = [](int i, int r) { return i * 32 + 8 * (r >> 2) + (r & 3); };
  const __amdgpu_buffer_rsrc_t crs = buf_rsrc(C + (int64_t)(bm + wm) * ldc + (bn + wn));
  f16v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        if (!MASKED) {
          acc[i][j][r] = buf_ld<float>(crs, coff, (crow(i, r) * (int)ldc + j * 32) * 4);
        } else {
          const int gi = min(bm + wm + lr0 + crow(i, r), M - 1);
          const int gj = min(bn + wn + lc + j * 32, N - 1);
          acc[i][j][r] = C[(int64_t)gi * ldc + gj];
        }
      }

  // Global fetch: A: thread -> row t/2, k (t%2)*8 .. +8 (two 16-byte loads);
  // B: thread -> column t%128, k (t/128)*8 .. +8 (coalesced along the row).
  float ra[8], rb[8];
  auto fetch = [&](int k0) {
    const int i = t >> 1, kk = (t & 1) * 8;
    const int j = t & 127, kb = (t >> 7) * 8;
    if (!MASKED) {
      const __amdgpu_buffer_rsrc_t ars = buf_rsrc(A + (int64_t)bm * lda + k0);
      const int avo = (i * (int)lda + kk) * 4;
#pragma unroll
      for (int q = 0; q < 8; q += 4) {
        const f4 v = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(ars, avo, q * 4, 0));
#pragma unroll
        for (int e = 0; e < 4; ++e) ra[q + e] = v[e];
      }
      const __amdgpu_buffer_rsrc_t brs = buf_rsrc(B + (int64_t)k0 * ldb + bn);
      const int bvo = (kb * (int)ldb + j) * 4;
#pragma unroll
      for (int q = 0; q < 8; ++q) rb[q] = buf_ld<float>(brs, bvo, q * (int)ldb * 4);
    } else {
      // Clamped indices keep every load inside the matrix; rows of A and
      // columns of B outside it only reach accumulators that are never
      // stored, and k beyond K is zeroed in both operands.
      const float *ar = A + (int64_t)min(bm + i, M - 1) * lda;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int gk = k0 + kk + q;
        const float v = ar[min(gk, K - 1)];
        ra[q] = gk < K ? v : 0.f;
      }
      const int cj = min(bn + j, N - 1);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int gk = k0 + kb + q;
        const float v = B[(int64_t)min(gk, K - 1) * ldb + cj];
        rb[q] = gk < K ? v : 0.f;
      }
    }
  };
  auto stash = [&](int buf) {
    const int i = t >> 1, kk = (t & 1) * 8;
    const int j = t & 127, kb = (t >> 7) * 8;
    bf8 pa[NS], pb[NS];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      __bf16 sa[NS], sb[NS];
      split3<NS>(-ra[e], sa);
      split3<NS>(rb[e], sb);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        pa[s][e] = sa[s];
        pb[s][e] = sb[s];
      }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      *reinterpret_cast<bf8 *>(&As[buf][s][i][kk]) = pa[s];
      *reinterpret_cast<bf8 *>(&Bs[buf][s][j][kb]) = pb[s];
    }
  };

  const int nstage = (K + kBBK - 1) / kBBK;
  const int lrow = lane & 31, lk = (lane >> 5) * 8;
  fetch(0);
  stash(0);
  __syncthreads();
  for (int s = 0; s < nstage; ++s) {
    const int buf = s & 1;
    if (s + 1 < nstage) fetch((s + 1) * kBBK);
    bf8 av[NS][2], bv[NS][2];
#pragma unroll
    for (int p = 0; p < NS; ++p)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        av[p][i] = *reinterpret_cast<const bf8 *>(&As[buf][p][wm + i * 32 + lrow][lk]);
        bv[p][i] = *reinterpret_cast<const bf8 *>(&Bs[buf][p][wn + i * 32 + lrow][lk]);
      }
    // Products smallest first, each over the four independent blocks.
#pragma unroll
    for (int p = 0; p < (NS == 3 ? 6 : 1); ++p) {
      const int qa = NS == 3 ? (p == 0 ? 2 : p == 1 ? 0 : p == 2 ? 1 : p == 3 ? 1 : 0) : 0;
      const int qb = NS == 3 ? (p == 0 ? 0 : p == 1 ? 2 : p == 2 ? 1 : p == 3 ? 0 : p == 4 ? 1 : 0) : 0;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[qa][i], bv[qb][j], acc[i][j], 0, 0, 0);
    }
    if (s + 1 < nstage) stash(buf ^ 1);
    __syncthreads();
  }

  int ldc_e = (int)ldc, mrem_e = mrem, nrem_e = nrem;
  asm volatile("" : "+s"(ldc_e));
  asm volatile("" : "+v"(mrem_e), "+v"(nrem_e));
  if (ghost) return;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        if (!MASKED || (crow(i, r) < mrem_e && j * 32 < nrem_e))
          buf_st<float>(acc[i][j][r], crs, coff, (crow(i, r) * ldc_e + j * 32) * 4);
      }
}

// TPW = 2 (capped grids): two horizontally adjacent tiles per workgroup, as
// gemm_tile_kernel.  NSL >= 2 planes are allocated so that the workgroup's
// LDS (96 KB or more) keeps it alone on its CU.
template <int NS, int TPW, bool MASKED>
__global__ void __launch_bounds__(256 * TPW, TPW == 1 ? 2 : 1)
    bgemm_tile_kernel(int M, int N, int K, const float *__restrict__ A, int64_t lda,
                      const float *__restrict__ B, int64_t ldb, float *__restrict__ C, int64_t ldc) {
  constexpr int NSL = (TPW > 1 && NS < 2) ? 2 : NS;
  __shared__ __align__(16) __bf16 As_all[TPW][2][NSL][kTileM][kBP];
  __shared__ __align__(16) __bf16 Bs_all[TPW][2][NSL][kTileN][kBP];
  const int h = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8);
  const int nbm = (M + kTileM - 1) / kTileM, nbn = (N + kTileN - 1) / kTileN;
  const int nbnp = (nbn + TPW - 1) / TPW;
  const int ntile = nbm * nbnp;
  for (int id = blockIdx.x; id < ntile; id += gridDim.x) {
    int tm, tnp;
    tile_coords(id, nbm, nbnp, tm, tnp);
    const int tn = tnp * TPW + h;
    const bool ghost = tn >= nbn;
    const int bm = tm * kTileM, bn = (ghost && !MASKED ? nbn - 1 : tn) * kTileN;
    bgemm_tile<NS, NSL, MASKED>(M, N, K, A, lda, B, ldb, C, ldc, bm, bn, ghost, As_all[h], Bs_all[h]);
  }
}

hipError_t launch_bgemm_tile(lair_context *ctx, int ns, int M, int N, int K, const float *A,
                             int64_t lda, const float *B, int64_t ldb, float *C, int64_t ldc) {
  hipStream_t st = ctx->stream;
  auto run = [&](auto nsc) {
    constexpr int NS = decltype(nsc)::value;
    auto masked = [&](int m, int n, const float *a, const float *b, float *c) {
      if (m <= 0 || n <= 0) return;
      const unsigned g = (unsigned)(((m + kTileM - 1) / kTileM) * ((n + kTileN - 1) / kTileN));
      bgemm_tile_kernel<NS, 1, true><<<g, 256, 0, st>>>(m, n, K, a, lda, b, ldb, c, ldc);
    };
    auto interior = [&](int m, int n) {
      const int nbm = m / kTileM, nbn = n / kTileN;
      if (ctx->bulk_wg_cap > 0) {
        const int ntile = nbm * ((nbn + 1) / 2);
        const unsigned grid = (unsigned)std::min(ntile, ctx->bulk_wg_cap);
        bgemm_tile_kernel<NS, 2, false><<<grid, 512, 0, st>>>(m, n, K, A, lda, B, ldb, C, ldc);
      } else {
        bgemm_tile_kernel<NS, 1, false><<<(unsigned)(nbm * nbn), 256, 0, st>>>(m, n, K, A, lda, B, ldb, C, ldc);
      }
    };
    return split_tiles<float, false>(M, N, K, kBBK, A, lda, B, ldb, C, ldc, interior, masked);
  };
  return ns == 3 ? run(std::integral_constant<int, 3>{}) : run(std::integral_constant<int, 1>{});
}

}  // namespace lair
