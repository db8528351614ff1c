// Internal declarations of the gfx950 LU library: context, workspace and the
// kernel launchers shared by the host driver.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lair_dev.h"

namespace lair {

// Width of one cooperative panel leaf (columns factored with exact partial
// pivoting before the in-panel update).
constexpr int kLeafW = 16;
// Systems up to this order are solved in the reference's exact operation
// order by getrs (one block, one thread per right-hand side).
constexpr int kExactOrderMax = 128;
// Maximum number of workgroups in a cooperative panel launch.
constexpr int kMaxPanelWG = 64;

// Trailing-update arithmetic.
enum GemmMode : int {
  kGemmF64 = 0,    // v_mfma_f64_16x16x4_f64
  kGemmF32 = 1,    // v_mfma_f32_16x16x4_f32
  kGemmBF16x6 = 2, // six v_mfma_f32_16x16x32_bf16 products of 3-way bf16 splits
  kGemmBF16 = 3,   // one v_mfma_f32_16x16x32_bf16 product
};

void set_error(const std::string &msg);

struct Buffer {
  void *ptr = nullptr;
  size_t bytes = 0;
};

struct ProfRecord {
  int kclass;
  double flops;
  hipEvent_t start, stop;
};

}  // namespace lair

struct lair_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;  // serializes calls that share this context's workspace

  // Cooperative panel synchronization state (device memory).
  uint32_t *flags = nullptr;   // one monotonic sequence word per workgroup
  uint32_t seq = 0;            // next sequence base
  double *cand_val = nullptr;  // [2][kMaxPanelWG]
  int32_t *cand_idx = nullptr; // [2][kMaxPanelWG]
  double *cand_row = nullptr;  // [2][kMaxPanelWG][kLeafW] (holds float rows too)
  double *diag_row = nullptr;  // [2][kLeafW]
  int32_t *err = nullptr;      // set when a bounded wait expires
  int64_t *sing = nullptr;     // last zero-pivot column (device)
  double *norms = nullptr;     // scratch for norm reductions (uint64-ordered doubles)

  // Grow-only workspaces.
  lair::Buffer ws_a32, ws_lu64, ws_ipiv, ws_x32, ws_r, ws_tmp, ws_rhs;

  bool profiling = false;
  std::vector<lair::ProfRecord> prof;
  std::vector<hipEvent_t> event_pool;
  size_t event_next = 0;
  double prof_ms[LAIR_KCLASS_COUNT] = {};
  int64_t prof_n[LAIR_KCLASS_COUNT] = {};
  double prof_flops[LAIR_KCLASS_COUNT] = {};
};

namespace lair {

// Records (when profiling) a start event for a kernel of class kclass.
int prof_begin(lair_context *ctx);
void prof_end(lair_context *ctx, int slot, int kclass, double flops);

hipError_t ensure(Buffer &b, size_t bytes);

// ---- kernel launchers (kernels.hip) ------------------------------------

template <typename T>
hipError_t launch_panel(lair_context *ctx, T *a, int64_t lda, int row0, int m, int col0, int wcols,
                        int64_t *ipiv);

// Row interchanges ipiv[k0..k1) (forward, or reverse when rev) applied to the
// columns [c0, c1) minus [s0, s1) of a (row-major, lda).
template <typename T>
hipError_t launch_laswp(lair_context *ctx, T *a, int64_t lda, int c0, int c1, int s0, int s1,
                        const int64_t *ipiv, int k0, int k1, bool rev);

// B := op(Tri)^-1 B.  lower/upper refers to the stored triangle; trans uses
// its transpose; unit skips the diagonal (true division otherwise).
template <typename T>
hipError_t launch_trsm(lair_context *ctx, bool lower, bool trans, bool unit, int kb, int ncols,
                       const T *tri, int64_t ldt, T *b, int64_t ldb);

template <typename T>
hipError_t launch_getrs_exact(lair_context *ctx, int n, int nrhs, const T *lu, int64_t lda,
                              const int64_t *ipiv, T *b, int64_t ldb);

// C[M x N] -= op(A)[M x K] * B[K x N]; transA reads A stored as K x M.
template <typename T>
hipError_t launch_gemm_minus(lair_context *ctx, int mode, int M, int N, int K, const T *A,
                             int64_t lda, bool transA, const T *B, int64_t ldb, T *C, int64_t ldc);

hipError_t launch_fill_i64(lair_context *ctx, int64_t *p, int64_t n, int64_t v);
hipError_t launch_fill_u32(lair_context *ctx, uint32_t *p, int64_t n, uint32_t v);
hipError_t launch_fill_f64(lair_context *ctx, double *p, int64_t n, double v);

// Copies an FP64 matrix into FP32 storage and accumulates row |sums| maxima
// (||A||inf) into *norm_bits (uint64-ordered non-negative double).
hipError_t launch_convert_d2s_norm(lair_context *ctx, int n, int ncols, const double *a, int64_t lda,
                                   float *a32, int64_t lda32, unsigned long long *norm_bits);
hipError_t launch_rows_norm(lair_context *ctx, int n, int ncols, const double *a, int64_t lda,
                            unsigned long long *norm_bits);
// 2-D copies between precisions / same precision.
hipError_t launch_copy_d2s(lair_context *ctx, int rows, int cols, const double *x, int64_t ldx,
                           float *y, int64_t ldy);
hipError_t launch_copy_s2d(lair_context *ctx, int rows, int cols, const float *x, int64_t ldx,
                           double *y, int64_t ldy);
hipError_t launch_copy_d2d(lair_context *ctx, int rows, int cols, const double *x, int64_t ldx,
                           double *y, int64_t ldy);
// x += (double) d
hipError_t launch_axpy_s2d(lair_context *ctx, int rows, int cols, const float *d, int64_t ldd,
                           double *x, int64_t ldx);
// R = B - A X (FP64), nrhs columns.
hipError_t launch_residual(lair_context *ctx, int n, int nrhs, const double *a, int64_t lda,
                           const double *x, int64_t ldx, const double *b, int64_t ldb, double *r,
                           int64_t ldr);
// per-column max |x| -> bits[col] (atomic max on uint64-ordered doubles).
hipError_t launch_colmax(lair_context *ctx, int rows, int cols, const double *x, int64_t ldx,
                         unsigned long long *bits);

}  // namespace lair
