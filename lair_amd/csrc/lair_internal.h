// Internal declarations of the gfx950 LU library: context, workspace and the
// kernel launchers shared by the host driver.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lair_dev.h"
#include "device_util.h"

namespace lair {

// Complex<f64> and Complex<f32> elements (interleaved re, im).
typedef cplx<double> zcplx;
typedef cplx<float> ccplx;

// Width of one cooperative panel leaf (columns factored with exact partial
// pivoting before the in-panel update).
constexpr int kLeafW = 16;
// Systems up to this order are solved in the reference's exact operation
// order by getrs (one block, one thread per right-hand side).
constexpr int kExactOrderMax = 128;
// Maximum number of workgroups in a cooperative panel launch.
constexpr int kMaxPanelWG = 64;
// ... and in a tournament-pivoting leaf (2048 rows each), whose records hold
// a workgroup's 16 nominee ids and values, plus the 16 original diagonal rows.
constexpr int kCaluMaxWG = 16;
constexpr int kCaluW = 16;
constexpr int kCaluRec = kCaluW + kCaluW * kCaluW;
constexpr int kCaluDiag = kCaluW * kCaluW;
constexpr int kCaluWords = kCaluMaxWG * kCaluRec + kCaluDiag;
// CUs kept free of the look-ahead stream's bulk updates for the critical path.
constexpr int kReservedCUs = 32;
// Row interchanges of one panel touch at most 2 * (panel width) rows.
constexpr int kMaxSwapRows = 512;
// Largest n whose row permutation is composed in LDS.
constexpr int kPermLdsMax = 32768;
// Tagged words per workgroup slot of the resident panel's row exchange
// (>= leaf width * 32-bit words per element).
constexpr int kRpRowWords = 128;
// Columns of B per workgroup in the triangular solves with L11, and the
// widest panel those kernels take.
constexpr int kTrsmCW = 16;
constexpr int kTrsmMaxKb = 128;
// Tagged header words per workgroup slot of the resident panel's exchange:
// real elements use 4, complex ones 1 + 2 * (32-bit words per element), in
// slots of kRpHdrWordsMax.  The two kinds live in separate regions of the
// header buffer (a real header's first word carries only a 16-bit tag, which
// a complex word could match): real at 0, complex at kRpHdrComplexOff.
constexpr int kRpHdrWords = 4;
constexpr int kRpHdrWordsMax = 16;
constexpr int kRpHdrComplexOff = 2 * kMaxPanelWG * kRpHdrWords;
constexpr int kRpHdrTotal = kRpHdrComplexOff + 2 * kMaxPanelWG * kRpHdrWordsMax;  // 64-bit words
// Widest complex panel (the triangular solve with L11 keeps L11 in LDS),
// and the largest order of the complex getrs (unknowns of a right-hand side
// held in registers of one 1024-thread workgroup, 16 per thread).
constexpr int kComplexMaxNb = 64;
constexpr int kComplexGetrsMax = 16384;

// Complex trailing update through the real GEMM (complex.hip): the real
// embedding of U12 (K <= kComplexMaxNb rows) goes to columns [col, col + N)
// of ctx->ws_cexp, sized by the caller before the factorization starts.
template <typename T>
hipError_t launch_cgemm_minus(lair_context *ctx, int mode, int M, int N, int K, const T *A, int64_t lda,
                              const T *B, int64_t ldb, T *C, int64_t ldc, int col, int kclass);
// Complex getrs (n <= kComplexGetrsMax): b := U^-1 L^-1 (b permuted by perm).
template <typename T>
hipError_t launch_cgetrs(lair_context *ctx, int n, int nrhs, const T *lu, int64_t lda, const int *perm, T *b,
                         int64_t ldb);

// Diagonal block size of the dataflow triangular solve.
template <typename T>
struct TrsvBlock;
template <>
struct TrsvBlock<double> {
  static constexpr int BS = 64;
};
template <>
struct TrsvBlock<float> {
  static constexpr int BS = 128;
};

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
  double bytes;
  hipEvent_t start, stop;
};

}  // namespace lair

struct lair_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;  // serializes calls that share this context's workspace

  // Cooperative panel synchronization state (device memory).
  // Tagged exchange records: [2][kMaxPanelWG][64] and [2][64] 64-bit words.
  unsigned long long *rec = nullptr;
  unsigned long long *drec = nullptr;
  unsigned long long *crec = nullptr;  // tournament leaf records (tagged words)
  uint32_t seq = 0;            // tag base of the next panel launch
  int panel_xcd_stride = 1;    // launched slots per panel workgroup (see launch_panel)
  int32_t *err = nullptr;      // set when a bounded wait expires
  // Polls a cooperative kernel's bounded wait makes before it gives up and
  // sets *err (0: every wait gives up at once).
  int32_t wait_limit = 1 << 23;
  // Resident panel (rpanel.hip): header and row slots, tag base, grid cap.
  unsigned long long *rp_hdr = nullptr;   // kRpHdrTotal words
  unsigned long long *rp_rowd = nullptr;  // [2][kMaxPanelWG][kRpRowWords]
  uint32_t rp_seq = 1;
  int rp_max_wg = lair::kMaxPanelWG;
  int64_t *sing = nullptr;     // last zero-pivot column (device)
  double *norms = nullptr;     // scratch for norm reductions (uint64-ordered doubles)
  // Pinned host words [sequence, status, words...] written by
  // launch_publish_words (device address pub_dev), and the last sequence.
  unsigned long long *pub_host = nullptr, *pub_dev = nullptr;
  unsigned long long pub_seq = 0;

  // Row-interchange map of the current panel (device); the *_alt set belongs
  // to the look-ahead stream (swapped in by lair::SideStream).
  int *swap_rows = nullptr, *swap_src = nullptr, *swap_count = nullptr;
  int *swap_rows_alt = nullptr, *swap_src_alt = nullptr, *swap_count_alt = nullptr;
  // Row maps written by single-leaf panels, by panel parity (read by both
  // streams while the next panel writes the other one).
  int *pmap_rows[2] = {nullptr, nullptr}, *pmap_src[2] = {nullptr, nullptr};
  int *pmap_count[2] = {nullptr, nullptr};

  // Look-ahead: the trailing update of panel k runs on stream2 while the
  // critical path (next panel's columns, next panel) continues on stream.
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_panel = nullptr, ev_bulk = nullptr;
  int num_cu = 256;
  int bulk_wg_cap = 0;  // > 0: tile GEMMs launch at most this many workgroups
  int panel_wgs = 0;    // workgroups of the widest resident panel of the current factorization

  // Grow-only workspaces.
  lair::Buffer ws_a32, ws_lu64, ws_ipiv, ws_x32, ws_r, ws_tmp, ws_rhs;
  lair::Buffer ws_perm, ws_linv, ws_uinv, ws_flags, ws_tmp2;
  // Complex trailing update: the real embedding of U12, 2 * kComplexMaxNb
  // rows of cexp_ld real values (two per column of the matrix).
  lair::Buffer ws_cexp;
  int64_t cexp_ld = 0;
  // Split-K slices of the many-right-hand-side residual.
  lair::Buffer ws_split;
  // When set, the triangular solves publish their tagged values here instead
  // of in ws_flags (caller-provided memory, zeroed by the caller).
  unsigned long long *xt_override = nullptr;
  // Device address, on this context's device, of the allocation-free getrs's
  // scratch array (driver.hip).
  void *noalloc_scratch = nullptr;
  size_t ws_flags_zeroed = 0;
  uint32_t tri_epoch = 0;

  // Panel widths chosen by the getrf autotuner, per (m, n, element size, mode).
  std::map<std::array<int64_t, 4>, int> nb_cache;

  int32_t prof_mask = 0;  // bit k: record kernel class k
  std::vector<lair::ProfRecord> prof;
  std::vector<hipEvent_t> event_pool;
  size_t event_next = 0;
  double prof_ms[LAIR_KCLASS_COUNT] = {};
  int64_t prof_n[LAIR_KCLASS_COUNT] = {};
  double prof_flops[LAIR_KCLASS_COUNT] = {};
  double prof_bytes[LAIR_KCLASS_COUNT] = {};
};

namespace lair {

// Redirects the context's launches to stream2 (and its row-map buffers) for
// the lifetime of the object.
// Its tile GEMMs run on at most (CUs - kReservedCUs) workgroups, one per CU,
// so the reserved CUs stay free for the critical path.
struct SideStream {
  lair_context *ctx;
  explicit SideStream(lair_context *c) : ctx(c) {
    swap();
    ctx->bulk_wg_cap = (ctx->num_cu - std::max(kReservedCUs, ctx->panel_wgs)) / 8 * 8;
  }
  ~SideStream() {
    ctx->bulk_wg_cap = 0;
    swap();
  }
  void swap() {
    std::swap(ctx->stream, ctx->stream2);
    std::swap(ctx->swap_rows, ctx->swap_rows_alt);
    std::swap(ctx->swap_src, ctx->swap_src_alt);
    std::swap(ctx->swap_count, ctx->swap_count_alt);
  }
};

// Records (when profiling) a start event for a kernel of class kclass.
int prof_begin(lair_context *ctx, int kclass);
void prof_end(lair_context *ctx, int slot, int kclass, double flops, double bytes = 0.0);

hipError_t ensure(Buffer &b, size_t bytes);

// ---- kernel launchers (kernels.hip) ------------------------------------

template <typename T>
hipError_t launch_panel(lair_context *ctx, T *a, int64_t lda, int row0, int m, int col0, int wcols,
                        int64_t *ipiv, int rest1, int left0);
// Resident-panel leaf (rpanel.hip): exact partial pivoting of columns
// [col0, col0 + wcols) over rows [row0, m), wcols <= nbw, rows kept in LDS.
// When map_rows is given, the leaf also writes the row map of its
// interchanges (2 * npiv gather entries, see launch_rowswap_map).
template <typename T>
hipError_t launch_rpanel(lair_context *ctx, T *a, int64_t lda, int row0, int m, int col0, int wcols,
                         int nbw, int64_t *ipiv, int *map_rows = nullptr, int *map_src = nullptr,
                         int *map_count = nullptr);
// Widest leaf width <= want (128, 64, 32 or 16) the resident panel supports
// for `rows` rows, or 0.
template <typename T>
int rpanel_leaf_width(lair_context *ctx, int want, int rows);
// Workgroups of a resident leaf of width nbw over `rows` rows (0: unsupported).
template <typename T>
int rpanel_grid(lair_context *ctx, int nbw, int rows);
// Tournament-pivoting FP32 leaf (mixed-precision factorization only).
hipError_t launch_panel_calu(lair_context *ctx, float *a, int64_t lda, int c, int m, int wcols,
                             int64_t *ipiv, int rest1, int left0);

// Row interchanges ipiv[k0..k1) (forward, or reverse when rev) applied to the
// columns [c0, c1) minus [s0, s1) of a (row-major, lda).
template <typename T>
hipError_t launch_laswp(lair_context *ctx, T *a, int64_t lda, int c0, int c1, int s0, int s1,
                        const int64_t *ipiv, int k0, int k1, bool rev);

template <typename T>
hipError_t launch_getrs_exact(lair_context *ctx, int n, int nrhs, const T *lu, int64_t lda,
                              const int64_t *ipiv, T *b, int64_t ldb);

// C[M x N] -= op(A)[M x K] * B[K x N]; transA reads A stored as K x M.
template <typename T>
hipError_t launch_gemm_minus(lair_context *ctx, int mode, int M, int N, int K, const T *A,
                             int64_t lda, bool transA, const T *B, int64_t ldb, T *C, int64_t ldc,
                             int kclass = LAIR_KCLASS_GEMM);

// Interchanges ipiv[k0..k1) on columns [c0,c1) minus [s0,s1), as one gather;
// perm (optional) is updated with the same row map.
template <typename T>
hipError_t launch_rowswap(lair_context *ctx, T *a, int64_t lda, int c0, int c1, int s0, int s1,
                          const int64_t *ipiv, int k0, int k1, int *perm);
// As launch_rowswap, with the row map given (rows[e] receives the old
// contents of src[e] for e < *count <= nmax; entries with rows[e] == src[e]
// do nothing), e.g. the one a resident leaf wrote.
template <typename T>
hipError_t launch_rowswap_map(lair_context *ctx, T *a, int64_t lda, int c0, int c1, int s0, int s1,
                              const int *rows, const int *src, const int *count, int nmax, int *perm);
hipError_t launch_perm_from_ipiv(lair_context *ctx, const int64_t *ipiv, int n, int *perm);
hipError_t launch_iota(lair_context *ctx, int *p, int n);
template <typename S, typename D>
hipError_t launch_permute_rows(lair_context *ctx, int rows, int cols, const int *perm, const S *x,
                               int64_t ldx, D *y, int64_t ldy, bool scatter,
                               unsigned long long *zero = nullptr, int nz = 0);
// B := L^-1 B (unit lower L, kb <= 256).
template <typename T>
hipError_t launch_trsm_llu(lair_context *ctx, int kb, int ncols, const T *L, int64_t ldl, T *B,
                           int64_t ldb);
// Interchanges (row map, first kb entries = rows k .. k + kb - 1) and
// U12 = L11^-1 A12 on columns [c0, c1) of panel [k, k + kb), kb <= 128, in one
// pass per 16-column strip (the rows moved below the panel are written too).
template <typename T>
hipError_t launch_swap_trsm(lair_context *ctx, T *a, int64_t lda, int k, int kb, int c0, int c1,
                            const int *rows, const int *src, const int *count, int nmax);
// Composes ipiv[k0..k1) into the context's current row-map buffers
// (ctx->swap_rows / swap_src / swap_count, 2 * (k1 - k0) entries at most).
hipError_t launch_rowmap(lair_context *ctx, const int64_t *ipiv, int k0, int k1);
// Inverses of the diagonal blocks (TrsvBlock<T>::BS) of L (unit) and U.
template <typename T>
hipError_t launch_trtri_diag(lair_context *ctx, const T *lu, int64_t lda, int n, T *linv, T *uinv);
template <typename T>
hipError_t launch_trsv_dataflow(lair_context *ctx, const T *lu, int64_t lda, const T *dinv,
                                bool dinv_trans, const T *xin, T *x, int64_t ldx, int n, int nrhs,
                                bool forward, bool trans);
// One refinement step's second half in one launch (nrhs <= 4): the backward
// solve d = U^-1 y in FP32 (y and d with leading dimension nrhs), x += d in
// FP64 (x taken as zero when x_zero), r = b - A x in FP64 over the FP64 A,
// and the column maxima of |x| and |r| into xbits / rbits (ordered bit
// patterns that must be zero at launch).
hipError_t launch_trsv_ir(lair_context *ctx, const float *lu, int64_t lda, const float *uinv,
                          const float *y, float *d, int n, int nrhs, double *x, int64_t ldx,
                          bool x_zero, const double *a, int64_t lda64, const double *b, int64_t ldb,
                          double *r, int64_t ldr, unsigned long long *xbits,
                          unsigned long long *rbits);

// 128x128-tile MFMA trailing updates (gemm_kernels.hip).
hipError_t launch_dgemm_tile(lair_context *ctx, int M, int N, int K, const double *A, int64_t lda,
                             bool transA, const double *B, int64_t ldb, double *C, int64_t ldc,
                             bool store = false);
hipError_t launch_sgemm_tile(lair_context *ctx, int M, int N, int K, const float *A, int64_t lda,
                             bool transA, const float *B, int64_t ldb, float *C, int64_t ldc);
// W_y = -A[:, y Ks:(y+1) Ks] B[y Ks:(y+1) Ks, :] for y < S (Ks = K / S rounded
// up to 16), each into W + y * wstride (leading dimension ldw).
hipError_t launch_dgemm_splitk_store(lair_context *ctx, int M, int N, int K, int S, const double *A, int64_t lda,
                                     const double *B, int64_t ldb, double *W, int64_t ldw, int64_t wstride);
// ns = 3: six-product bf16 split (FP32-level); ns = 1: single bf16 product.
hipError_t launch_bgemm_tile(lair_context *ctx, int ns, int M, int N, int K, const float *A,
                             int64_t lda, const float *B, int64_t ldb, float *C, int64_t ldc);

// Persistent batched solver (batched.hip): device metadata and launch shape.
struct BatchLaunch {
  int count = 0;
  const int *n = nullptr;
  const int *nrhs = nullptr;
  const int64_t *a_off = nullptr;
  const int64_t *b_off = nullptr;
  const int64_t *p_off = nullptr;
  const int *order = nullptr;
  int *queue = nullptr;
  int lds_max_elems = 0;  // largest augmented matrix solved in LDS
  int nb = 16;            // panel width (8 or 16)
  int grid = 256;         // persistent workgroups
  int occ = 1;            // register budget for this many waves per SIMD (1: the compiler's choice)
};
template <typename T>
hipError_t launch_batched_gesv(lair_context *ctx, const BatchLaunch &bl, T *a, T *b, int64_t *ipiv,
                               int *info);
size_t batched_lds_bytes_for(int elem_size, int nb, int lds_max_elems);
void batched_prepare();

// Copies nw words and the status word to ctx->pub_dev[2..] and [1], then
// writes seq to ctx->pub_dev[0] (system-scope release) for the host to poll.
hipError_t launch_publish_words(lair_context *ctx, const unsigned long long *src, int nw,
                                unsigned long long seq);
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
// Factor extraction from packed LU factors (m x n, lda) and the row
// permutation: which = LAIR_LU_P (m x m), LAIR_LU_L (m x k), LAIR_LU_U
// (k x n) or LAIR_LU_PL (m x k), k = min(m, n), into out (ldo).
hipError_t launch_lu_extract(lair_context *ctx, int which, int m, int n, const double *lu, int64_t lda,
                             const int *perm, double *out, int64_t ldo);
// y (cols x rows) = x^T.
hipError_t launch_transpose_d(lair_context *ctx, int rows, int cols, const double *x, int64_t ldx,
                              double *y, int64_t ldy);
// x += (double) d
// x += d; words zero[0 .. nz) are cleared on the way (nullptr: none).
hipError_t launch_axpy_s2d(lair_context *ctx, int rows, int cols, const float *d, int64_t ldd,
                           double *x, int64_t ldx, unsigned long long *zero = nullptr, int nz = 0);
// R = B - A X (FP64), nrhs columns.
// r = b - A x; with rbits / xbits (cleared beforehand), also the column
// maxima of |r| and |x| as ordered bit patterns.
hipError_t launch_residual(lair_context *ctx, int n, int nrhs, const double *a, int64_t lda,
                           const double *x, int64_t ldx, const double *b, int64_t ldb, double *r,
                           int64_t ldr, unsigned long long *rbits = nullptr,
                           unsigned long long *xbits = nullptr);
// per-column max |x| -> bits[col] (atomic max on uint64-ordered doubles).
hipError_t launch_colmax(lair_context *ctx, int rows, int cols, const double *x, int64_t ldx,
                         unsigned long long *bits);

}  // namespace lair
