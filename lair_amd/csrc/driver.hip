// Host driver of the gfx950 LU path: context, blocked factorization loop,
// triangular solves, mixed-precision refinement and the C ABI.
//
// Reference interfaces mirrored here (vinesystems/lair 0.8.0):
//   lapack::getrf   src/lapack/getrf.rs:12-27   -> lair_dgetrf / lair_sgetrf
//   lapack::getrs   src/lapack/getrs.rs:12-39   -> lair_dgetrs / lair_sgetrs
//   lapack::laswp   src/lapack/laswp.rs:11-40   -> lair_dlaswp
//   equation::solve src/equation.rs:32-60       -> lair_dsolve
#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <vector>

#include "lair_internal.h"

namespace lair {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

#define LAIR_HIP(call)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));                    \
      return e_ == hipErrorOutOfMemory ? LAIR_ERR_NOMEM : LAIR_ERR_DEVICE;             \
    }                                                                                  \
  } while (0)

hipError_t ensure(Buffer &b, size_t bytes) {
  if (b.bytes >= bytes && b.ptr) return hipSuccess;
  if (b.ptr) {
    hipError_t e = hipFree(b.ptr);
    if (e != hipSuccess) return e;
    b.ptr = nullptr;
    b.bytes = 0;
  }
  size_t want = std::max<size_t>(bytes, 256);
  hipError_t e = hipMalloc(&b.ptr, want);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    return e;
  }
  b.bytes = want;
  return hipSuccess;
}

int prof_begin(lair_context *ctx, int kclass) {
  if (!(ctx->prof_mask & (1 << kclass))) return -1;
  if (ctx->event_next + 2 > ctx->event_pool.size()) {
    for (int i = 0; i < 64; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      ctx->event_pool.push_back(e);
    }
  }
  const int slot = (int)ctx->event_next;
  ctx->event_next += 2;
  hipEventRecord(ctx->event_pool[slot], ctx->stream);
  return slot;
}

void prof_end(lair_context *ctx, int slot, int kclass, double flops, double bytes) {
  if (slot < 0) return;
  hipEventRecord(ctx->event_pool[slot + 1], ctx->stream);
  ctx->prof.push_back({kclass, flops, bytes, ctx->event_pool[slot], ctx->event_pool[slot + 1]});
}

static void prof_collect(lair_context *ctx) {
  if (ctx->prof.empty()) return;
  hipStreamSynchronize(ctx->stream);
  for (auto &r : ctx->prof) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.start, r.stop) == hipSuccess) ctx->prof_ms[r.kclass] += ms;
    ctx->prof_n[r.kclass] += 1;
    ctx->prof_flops[r.kclass] += r.flops;
    ctx->prof_bytes[r.kclass] += r.bytes;
  }
  ctx->prof.clear();
  ctx->event_next = 0;
}

static int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Leading dimension for device copies: a multiple of 64 elements, nudged off
// multiples of 4096 elements so rows do not all start on the same channel.
template <typename T>
static int64_t pad_ld(int64_t n) {
  int64_t ld = round_up(std::max<int64_t>(n, 1), 64);
  if (ld % 4096 == 0) ld += 64;
  return ld;
}

// Default panel width.  FP64 keeps 64-column panels (one 64-column leaf, no
// in-panel update) up to k = 10240: measured 1.94 against 2.30 ms at n=512,
// 38.4 against 43.4 ms at n=8192, 56.2 against 57.9 ms at n=10240; from 12288
// on the 128-column panels' fewer bulk passes win (78.6 against 82.3 ms at
// 12288, 134 against 154 ms at 16384).
static int choose_nb(int64_t m, int64_t n, bool fp64) {
  const int64_t k = std::min(m, n);
  if (k <= 256 || (fp64 && k <= 10240)) return 64;
  return 128;
}

static lair_status hip_status(hipError_t e, const char *what) {
  if (e == hipSuccess) return LAIR_OK;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? LAIR_ERR_NOMEM : LAIR_ERR_DEVICE;
}

// Synchronizes the context's stream and reports what the device status word
// recorded since the last check: a caller pivot outside [0, n) as
// LAIR_ERR_SHAPE, a bounded wait that gave up as LAIR_ERR_SYNC.  The word is
// sticky (kernels that see it give up their own waits at once) until it is
// read here and reset.
constexpr int32_t kErrPivotBit = 2;  // device_util.h kErrPivot
static lair_status sync_status(lair_context *ctx) {
  int32_t err = 0;
  hipError_t e = hipMemcpyAsync(&err, ctx->err, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_status(e, "synchronize");
  if (!err) return LAIR_OK;
  e = launch_fill_u32(ctx, reinterpret_cast<uint32_t *>(ctx->err), 1, 0u);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_status(e, "reset");
  if (err & kErrPivotBit) {
    set_error("pivot index out of range");
    return LAIR_ERR_SHAPE;
  }
  set_error("a cooperative kernel's bounded wait gave up (LAIR_ERR_SYNC)");
  return LAIR_ERR_SYNC;
}

// Blocked right-looking LU with partial pivoting (the panel -> laswp -> trsm
// -> gemm sequence of the reference's recursive_inner, getrf.rs:216-322, with
// a fixed panel width nb).  Each panel of nb columns is factored as leaves of
// kLeafW columns with exact partial pivoting; the leaf's interchanges are
// applied to the rest of the panel, the leaf's U rows solved and the rest of
// the panel updated before the next leaf.  After the panel, its interchanges
// are applied to the columns left and right of it (one gather, which also
// updates the row permutation perm when given), U12 = L11^-1 A12 and
// A22 -= L21 U12.
template <typename T>
static lair_status getrf_impl(lair_context *ctx, int64_t m64, int64_t n64, T *a, int64_t lda,
                              int64_t *ipiv, int64_t nb, int mode, int64_t *singular,
                              int *perm = nullptr, bool tournament = false) {
  if (m64 < 0) return -2;
  if (n64 < 0) return -3;
  if (lda < std::max<int64_t>(n64, 1)) return -5;
  if (m64 > (int64_t)kMaxPanelWG * 2048 || n64 > INT32_MAX) {
    set_error("matrix too large for the cooperative panel");
    return LAIR_ERR_UNSUPPORTED;
  }
  const int m = (int)m64, n = (int)n64;
  const int kmax = std::min(m, n);
  if (nb <= 0) nb = choose_nb(m, n, std::is_same<T, double>::value);
  if (nb > 256) nb = 256;
  if (nb < kLeafW) nb = kLeafW;
  nb = (nb / kLeafW) * kLeafW;
  constexpr bool kComplex = cplx_traits<T>::is;
  if constexpr (kComplex) {
    // Complex panels: resident leaves only, at most kComplexMaxNb wide; the
    // real embedding of U12 for the trailing updates is sized here, before
    // either stream of the look-ahead uses it.
    nb = std::min<int64_t>(nb, kComplexMaxNb);
    if (kmax > 0 && rpanel_leaf_width<T>(ctx, std::max(std::min<int>((int)nb, kmax), kLeafW), m) == 0) {
      set_error("too many rows for the complex resident panel");
      return LAIR_ERR_UNSUPPORTED;
    }
    typedef typename cplx_traits<T>::real R;
    const int64_t lde = 2 * (((int64_t)n + 63) / 64 * 64);
    LAIR_HIP(ensure(ctx->ws_cexp, sizeof(R) * (size_t)(2 * kComplexMaxNb) * (size_t)lde));
    ctx->cexp_ld = lde;
  }
  // C[M x N] -= A[M x K] B[K x N] inside the matrix (leading dimension lda);
  // col is the column of C's first element.
  auto gemm_minus = [&](int M, int N, int K, const T *A, const T *B, T *C, int col, int kclass) -> hipError_t {
    if constexpr (kComplex)
      return launch_cgemm_minus<T>(ctx, mode, M, N, K, A, lda, B, lda, C, lda, col, kclass);
    else
      return launch_gemm_minus<T>(ctx, mode, M, N, K, A, lda, false, B, lda, C, lda, kclass);
  };
  hipError_t e = launch_fill_i64(ctx, ctx->sing, 1, -1);
  if (e != hipSuccess) return hip_status(e, "fill");
  if (perm) {
    e = launch_iota(ctx, perm, m);
    if (e != hipSuccess) return hip_status(e, "iota");
  }
  // Keep exchange tags far from wrap-around (a reused tag could match a stale
  // record): clear the records and restart the tags.
  if (ctx->seq > 0xF0000000u) {
    e = launch_fill_u32(ctx, reinterpret_cast<uint32_t *>(ctx->rec), 4 * kMaxPanelWG * 64, 0u);
    if (e != hipSuccess) return hip_status(e, "fill");
    e = launch_fill_u32(ctx, reinterpret_cast<uint32_t *>(ctx->drec), 4 * 64, 0u);
    if (e != hipSuccess) return hip_status(e, "fill");
    e = launch_fill_u32(ctx, reinterpret_cast<uint32_t *>(ctx->crec), 2 * kCaluWords, 0u);
    if (e != hipSuccess) return hip_status(e, "fill");
    ctx->seq = 1;
  }
  // Panels that are a single resident leaf leave their row map in
  // ctx->pmap_*[parity of the panel index]; the others are composed from the
  // pivots by rowmap_kernel.
  bool pmap_ok[2] = {false, false};
  auto parity = [&](int k) { return (k / (int)nb) & 1; };
  // Leaf width of the panel at column k: the widest resident leaf, or, for
  // FP64, half of it while at least kNarrowLeafRows rows remain.  There the
  // FP64 bulk update (MFMA-bound) sets the pace; a leaf half as wide keeps
  // twice the rows per workgroup, so half as many CUs are reserved for it
  // and the bulk update gets the rest (the FP32 leaf chain is the critical
  // path at every size, and keeps its full width).
  constexpr int kNarrowLeafRows = 12288;
  auto leaf_width = [&](int k, int kb) -> int {
    if (tournament) return 0;
    int lw = rpanel_leaf_width<T>(ctx, kComplex ? std::max(kb, kLeafW) : kb, m - k);
    if (std::is_same<T, double>::value && lw > kLeafW && m - k >= kNarrowLeafRows) {
      const int lw2 = rpanel_leaf_width<T>(ctx, lw / 2, m - k);
      if (lw2 > 0) lw = lw2;
    }
    return lw;
  };
  // Factors the panel of columns [k, k + kb) on the context's stream.
  auto factor_panel = [&](int k, int kb) -> hipError_t {
    // Resident leaves (rows kept in LDS for the whole leaf).  A leaf narrower
    // than the panel is followed by its interchanges on the panel's other
    // columns, the U rows right of it (trsm) and the update below (gemm).
    // (A complex panel narrower than the narrowest leaf goes to a leaf of
    // that width with fewer columns.)
    const int lw = leaf_width(k, kb);
    const int pp = parity(k);
    pmap_ok[pp] = lw >= kb;
    if (lw > 0) {
      for (int c = k; c < k + kb; c += lw) {
        const int w = std::min(lw, k + kb - c);
        hipError_t e2 = pmap_ok[pp] ? launch_rpanel<T>(ctx, a, lda, c, m, c, w, lw, ipiv, ctx->pmap_rows[pp],
                                                       ctx->pmap_src[pp], ctx->pmap_count[pp])
                                    : launch_rpanel<T>(ctx, a, lda, c, m, c, w, lw, ipiv);
        if (e2 != hipSuccess) return e2;
        const int npv = std::min(w, m - c);
        if (npv <= 0) continue;
        if (c > k) {
          e2 = launch_rowswap<T>(ctx, a, lda, k, c, k, k, ipiv, c, c + npv, nullptr);
          if (e2 != hipSuccess) return e2;
        }
        const int c1 = c + w, rest = k + kb - c1;
        if (rest > 0) {
          e2 = launch_rowswap<T>(ctx, a, lda, c1, k + kb, c1, c1, ipiv, c, c + npv, nullptr);
          if (e2 == hipSuccess)
            e2 = launch_trsm_llu<T>(ctx, npv, rest, a + (int64_t)c * lda + c, lda, a + (int64_t)c * lda + c1,
                                    lda);
          if (e2 == hipSuccess && m > c + npv)
            e2 = gemm_minus(m - (c + npv), rest, npv, a + (int64_t)(c + npv) * lda + c, a + (int64_t)c * lda + c1,
                            a + (int64_t)(c + npv) * lda + c1, c1, LAIR_KCLASS_GEMM_PANEL);
          if (e2 != hipSuccess) return e2;
        }
      }
      return hipSuccess;
    }
    if constexpr (kComplex) {
      return hipErrorInvalidValue;  // resident leaves are checked on entry
    } else {
      for (int c = k; c < k + kb; c += kLeafW) {
        const int w = std::min(kLeafW, k + kb - c);
        // The leaf kernel also applies its interchanges to the other columns
        // of the panel and solves the leaf's U rows right of it.
        hipError_t e2;
        if constexpr (std::is_same<T, float>::value) {
          if (tournament && m - c <= kCaluMaxWG * 2048)
            e2 = launch_panel_calu(ctx, a, lda, c, m, w, ipiv, k + kb, k);
          else
            e2 = launch_panel<T>(ctx, a, lda, c, m, c, w, ipiv, k + kb, k);
        } else {
          e2 = launch_panel<T>(ctx, a, lda, c, m, c, w, ipiv, k + kb, k);
        }
        if (e2 != hipSuccess) return e2;
        const int rest = k + kb - (c + w);
        if (rest > 0 && m > c + w) {
          e2 = gemm_minus(m - (c + w), rest, w, a + (int64_t)(c + w) * lda + c, a + (int64_t)c * lda + c + w,
                          a + (int64_t)(c + w) * lda + c + w, c + w, LAIR_KCLASS_GEMM_PANEL);
          if (e2 != hipSuccess) return e2;
        }
      }
      return hipSuccess;
    }
  };
  // Panel [k, k + kb)'s interchanges on columns [c0, c1) (and perm).
  auto swap_cols = [&](int k, int kb, int c0, int c1, int *pm) -> hipError_t {
    const int pp = parity(k);
    if (pmap_ok[pp])
      return launch_rowswap_map<T>(ctx, a, lda, c0, c1, c0, c0, ctx->pmap_rows[pp], ctx->pmap_src[pp],
                                   ctx->pmap_count[pp], 2 * std::min(kb, m - k), pm);
    return launch_rowswap<T>(ctx, a, lda, c0, c1, c0, c0, ipiv, k, k + kb, pm);
  };
  // Brings columns [c0, c1) up to date with panel [k, k + kb): interchanges,
  // U12 = L11^-1 A12 (fused for panels up to kTrsmMaxKb wide), A22 -= L21 U12.
  auto update_cols = [&](int k, int kb, int c0, int c1) -> hipError_t {
    if (c1 <= c0) return hipSuccess;
    hipError_t e2;
    if (kb <= kTrsmMaxKb) {
      const int pp = parity(k);
      if (pmap_ok[pp]) {
        e2 = launch_swap_trsm<T>(ctx, a, lda, k, kb, c0, c1, ctx->pmap_rows[pp], ctx->pmap_src[pp],
                                 ctx->pmap_count[pp], 2 * kb);
      } else {
        e2 = launch_rowmap(ctx, ipiv, k, k + kb);
        if (e2 == hipSuccess)
          e2 = launch_swap_trsm<T>(ctx, a, lda, k, kb, c0, c1, ctx->swap_rows, ctx->swap_src,
                                   ctx->swap_count, 2 * kb);
      }
    } else {
      e2 = swap_cols(k, kb, c0, c1, nullptr);
      if (e2 == hipSuccess)
        e2 = launch_trsm_llu<T>(ctx, kb, c1 - c0, a + (int64_t)k * lda + k, lda, a + (int64_t)k * lda + c0,
                                lda);
    }
    if (e2 != hipSuccess) return e2;
    if (m > k + kb)
      e2 = gemm_minus(m - (k + kb), c1 - c0, kb, a + (int64_t)(k + kb) * lda + k, a + (int64_t)k * lda + c0,
                      a + (int64_t)(k + kb) * lda + c0, c0, LAIR_KCLASS_GEMM);
    return e2;
  };
  // A bulk update leaves the concurrent panel's leaf workgroups their CUs:
  // a resident leaf needs all its workgroups running at once.
  auto reserve_for = [&](int k) {
    const int kb = std::min<int>((int)nb, kmax - k);
    const int lw = k < kmax ? leaf_width(k, kb) : 0;
    ctx->panel_wgs = lw > 0 ? rpanel_grid<T>(ctx, lw, m - k) : 0;
  };
  // Look-ahead of depth one.  Stream 1 (the context's stream) carries the
  // critical path: panel k+1's columns are updated first, then panel k+1 is
  // factored.  Stream 2 meanwhile applies panel k to everything else: the
  // columns left of it (interchanges only, and perm) and the columns right of
  // panel k+1 (interchanges, trsm, gemm).  Stream 1 waits for stream 2's
  // update of panel k before it touches panel k+2's columns.
  if (kmax > 0) {
    e = factor_panel(0, std::min<int>((int)nb, kmax));
    if (e != hipSuccess) return hip_status(e, "panel");
  }
  for (int k = 0; k < kmax; k += (int)nb) {
    const int kb = std::min<int>((int)nb, kmax - k);
    const bool has_next = k + kb < kmax;
    const int nxt1 = has_next ? std::min<int>(k + kb + (int)nb, n) : k + kb;
    LAIR_HIP(hipEventRecord(ctx->ev_panel, ctx->stream));
    reserve_for(k + kb);
    {
      SideStream side(ctx);
      LAIR_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_panel, 0));
      // Left columns: interchanges only (and the row permutation).
      e = swap_cols(k, kb, 0, k, perm);
      if (e != hipSuccess) return hip_status(e, "rowswap");
      e = update_cols(k, kb, nxt1, n);
      if (e != hipSuccess) return hip_status(e, "update");
      LAIR_HIP(hipEventRecord(ctx->ev_bulk, ctx->stream));
    }
    if (has_next) {
      e = update_cols(k, kb, k + kb, nxt1);
      if (e != hipSuccess) return hip_status(e, "update");
      e = factor_panel(k + kb, std::min<int>((int)nb, kmax - (k + kb)));
      if (e != hipSuccess) return hip_status(e, "panel");
    }
    // Panel k+1's update of the columns right of it reads what stream 2 wrote.
    LAIR_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_bulk, 0));
  }
  ctx->panel_wgs = 0;
  if (singular) {
    LAIR_HIP(hipMemcpyAsync(singular, ctx->sing, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    return sync_status(ctx);
  }
  return LAIR_OK;
}

// Panel-width autotuner: each candidate width in {64, 128, 256} factors a
// scratch copy of the matrix once (timed with events on the context's
// stream); the fastest is kept per (m, n, element size, mode) for the life of
// the context.
template <typename T>
static lair_status tune_nb(lair_context *ctx, int64_t m, int64_t n, const T *a, int64_t lda, int mode,
                           int64_t *nb_out) {
  const std::array<int64_t, 4> key{m, n, (int64_t)sizeof(T), (int64_t)mode};
  auto it = ctx->nb_cache.find(key);
  if (it != ctx->nb_cache.end()) {
    *nb_out = it->second;
    return LAIR_OK;
  }
  if (m <= 0 || n <= 0 || lda < n) {
    *nb_out = 0;
    return LAIR_OK;
  }
  const size_t bytes = sizeof(T) * (size_t)m * lda;
  void *scratch = nullptr, *sp = nullptr;
  LAIR_HIP(hipMalloc(&scratch, bytes));
  if (hipMalloc(&sp, sizeof(int64_t) * (size_t)std::min(m, n)) != hipSuccess) {
    hipFree(scratch);
    return hip_status(hipErrorOutOfMemory, "autotune");
  }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  int best = 128;
  float best_ms = 1e30f;
  lair_status st = LAIR_OK;
  for (int cand : {64, 128, 256}) {
    if (cand > std::min(m, n) && cand != 64) continue;
    hipMemcpyAsync(scratch, a, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    hipEventRecord(e0, ctx->stream);
    st = getrf_impl<T>(ctx, m, n, static_cast<T *>(scratch), lda, static_cast<int64_t *>(sp), cand, mode,
                       nullptr);
    if (st != LAIR_OK) break;
    hipEventRecord(e1, ctx->stream);
    st = sync_status(ctx);
    if (st != LAIR_OK) break;
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms < best_ms) {
      best_ms = ms;
      best = cand;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(scratch);
  hipFree(sp);
  if (st != LAIR_OK) return st;
  ctx->nb_cache[key] = best;
  *nb_out = best;
  return LAIR_OK;
}

// Solve-side data derived from a factorization: the row permutation and the
// inverses of the diagonal blocks of L and U.
template <typename T>
struct SolveAux {
  const int *perm = nullptr;
  const T *linv = nullptr;
  const T *uinv = nullptr;
};

// Device buffers of a getrs that the caller provides instead of the
// context's grow-only workspace (the allocation-free entry point).
template <typename T>
struct GetrsScratch {
  int *perm = nullptr;
  T *linv = nullptr, *uinv = nullptr, *tmp = nullptr;
};

// Builds SolveAux in the context's workspace (or in sc's buffers).  perm is
// taken from perm_in when the factorization produced it, otherwise composed
// from ipiv.
template <typename T>
static lair_status make_aux(lair_context *ctx, int n, const T *lu, int64_t lda, const int64_t *ipiv,
                            const int *perm_in, SolveAux<T> *aux,
                            const GetrsScratch<T> *sc = nullptr) {
  constexpr int BS = TrsvBlock<T>::BS;
  const int nblk = (n + BS - 1) / BS;
  if (!sc) {
    LAIR_HIP(ensure(ctx->ws_linv, sizeof(T) * (size_t)nblk * BS * BS));
    LAIR_HIP(ensure(ctx->ws_uinv, sizeof(T) * (size_t)nblk * BS * BS));
  }
  T *linv = sc ? sc->linv : static_cast<T *>(ctx->ws_linv.ptr);
  T *uinv = sc ? sc->uinv : static_cast<T *>(ctx->ws_uinv.ptr);
  hipError_t e;
  if (!perm_in) {
    if (!sc) LAIR_HIP(ensure(ctx->ws_perm, sizeof(int) * (size_t)n));
    int *perm = sc ? sc->perm : static_cast<int *>(ctx->ws_perm.ptr);
    e = launch_perm_from_ipiv(ctx, ipiv, n, perm);
    if (e != hipSuccess) return hip_status(e, "perm");
    perm_in = perm;
  }
  e = launch_trtri_diag<T>(ctx, lu, lda, n, linv, uinv);
  if (e != hipSuccess) return hip_status(e, "trtri");
  aux->perm = perm_in;
  aux->linv = linv;
  aux->uinv = uinv;
  return LAIR_OK;
}

// Dataflow getrs: A X = B as gather (P b), L forward, U backward;
// A^T X = B as U^T forward, L^T backward, scatter (P^T).
template <typename T>
static lair_status getrs_aux(lair_context *ctx, int n, int nrhs, const T *lu, int64_t lda,
                             const SolveAux<T> &aux, T *b, int64_t ldb, int trans,
                             const GetrsScratch<T> *sc = nullptr) {
  if (!sc) LAIR_HIP(ensure(ctx->ws_tmp2, sizeof(T) * (size_t)n * nrhs));
  T *t = sc ? sc->tmp : static_cast<T *>(ctx->ws_tmp2.ptr);
  hipError_t e;
  if (!trans) {
    e = launch_permute_rows<T, T>(ctx, n, nrhs, aux.perm, b, ldb, t, nrhs, false);
    if (e != hipSuccess) return hip_status(e, "permute");
    e = launch_trsv_dataflow<T>(ctx, lu, lda, aux.linv, false, t, t, nrhs, n, nrhs, true, false);
    if (e != hipSuccess) return hip_status(e, "trsv");
    if (ldb != nrhs) {
      e = launch_trsv_dataflow<T>(ctx, lu, lda, aux.uinv, false, t, t, nrhs, n, nrhs, false, false);
      if (e != hipSuccess) return hip_status(e, "trsv");
      e = launch_permute_rows<T, T>(ctx, n, nrhs, nullptr, t, nrhs, b, ldb, false);
    } else {
      e = launch_trsv_dataflow<T>(ctx, lu, lda, aux.uinv, false, t, b, ldb, n, nrhs, false, false);
    }
    if (e != hipSuccess) return hip_status(e, "trsv");
  } else {
    e = launch_permute_rows<T, T>(ctx, n, nrhs, nullptr, b, ldb, t, nrhs, false);
    if (e != hipSuccess) return hip_status(e, "copy");
    e = launch_trsv_dataflow<T>(ctx, lu, lda, aux.uinv, true, t, t, nrhs, n, nrhs, true, true);
    if (e != hipSuccess) return hip_status(e, "trsv");
    e = launch_trsv_dataflow<T>(ctx, lu, lda, aux.linv, true, t, t, nrhs, n, nrhs, false, true);
    if (e != hipSuccess) return hip_status(e, "trsv");
    e = launch_permute_rows<T, T>(ctx, n, nrhs, aux.perm, t, nrhs, b, ldb, true);
    if (e != hipSuccess) return hip_status(e, "scatter");
  }
  return LAIR_OK;
}

// getrs (getrs.rs:12-39 generalized to nrhs columns and to A^T).  Systems up
// to kExactOrderMax use the reference's exact operation order; larger ones
// the dataflow solve.
template <typename T>
static lair_status getrs_impl(lair_context *ctx, int64_t n64, int64_t nrhs64, const T *lu,
                              int64_t lda, const int64_t *ipiv, T *b, int64_t ldb, int trans,
                              const GetrsScratch<T> *sc = nullptr) {
  if (n64 < 0) return -2;
  if (nrhs64 < 0) return -3;
  if (n64 == 0 || nrhs64 == 0) return LAIR_OK;
  const int n = (int)n64, nrhs = (int)nrhs64;
  if (!trans && n <= kExactOrderMax) {
    hipError_t e = launch_getrs_exact<T>(ctx, n, nrhs, lu, lda, ipiv, b, ldb);
    return hip_status(e, "getrs");
  }
  SolveAux<T> aux;
  lair_status st = make_aux<T>(ctx, n, lu, lda, ipiv, nullptr, &aux, sc);
  if (st != LAIR_OK) return st;
  return getrs_aux<T>(ctx, n, nrhs, lu, lda, aux, b, ldb, trans, sc);
}

static double bits_to_double(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof d);
  return d;
}

// max over columns of ||b - A x||inf / (||A||inf ||x||inf) with r = b - Ax
// already in r.  norms must hold 1 + 2*nrhs words; norms[0] = ||A||inf bits.
static lair_status backward_error_read(lair_context *ctx, int nrhs, const unsigned long long *norms,
                                       double *berr);

static lair_status backward_error_from(lair_context *ctx, int n, int nrhs, const double *r,
                                       int64_t ldr, const double *x, int64_t ldx,
                                       unsigned long long *norms, double *berr) {
  hipError_t e = launch_fill_f64(ctx, reinterpret_cast<double *>(norms + 1), 2 * nrhs, 0.0);
  if (e != hipSuccess) return hip_status(e, "fill");
  e = launch_colmax(ctx, n, nrhs, r, ldr, norms + 1);
  if (e != hipSuccess) return hip_status(e, "colmax");
  e = launch_colmax(ctx, n, nrhs, x, ldx, norms + 1 + nrhs);
  if (e != hipSuccess) return hip_status(e, "colmax");
  return backward_error_read(ctx, nrhs, norms, berr);
}

// The worst column's ||r||/(||A|| ||x||) from the words [||A||, ||r_c||...,
// ||x_c||...] (ordered bit patterns) read back to the host.
static double berr_from_words(const unsigned long long *h, int nrhs) {
  const double na = bits_to_double(h[0]);
  double worst = 0.0;
  for (int c = 0; c < nrhs; ++c) {
    const double nr = bits_to_double(h[1 + c]);
    const double nx = bits_to_double(h[1 + nrhs + c]);
    double be;
    if (nr == 0.0)
      be = 0.0;
    else if (na == 0.0 || nx == 0.0)
      be = INFINITY;
    else
      be = nr / (na * nx);
    if (!(be <= worst)) worst = be;  // NaN propagates
  }
  return worst;
}

// The same from norms on the device; synchronizes the stream.
static lair_status backward_error_read(lair_context *ctx, int nrhs, const unsigned long long *norms,
                                       double *berr) {
  std::vector<unsigned long long> h(1 + 2 * (size_t)nrhs);
  LAIR_HIP(hipMemcpyAsync(h.data(), norms, h.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  lair_status st = sync_status(ctx);
  if (st != LAIR_OK) return st;
  *berr = berr_from_words(h.data(), nrhs);
  return LAIR_OK;
}

// The same without a stream synchronization: a one-wave kernel copies the
// norm words and the status word into pinned host memory and then writes a
// fresh sequence number (system-scope release); the host polls that number.
// After kPublishSpinMs without it, the host synchronizes the stream instead
// (which also reports a failed launch).  A set status word is reported (and
// reset) by sync_status.
constexpr double kPublishSpinMs = 20.0;
static lair_status backward_error_published(lair_context *ctx, int nrhs, const unsigned long long *norms,
                                            double *berr) {
  if (!ctx->pub_host) return backward_error_read(ctx, nrhs, norms, berr);
  const int nw = 1 + 2 * nrhs;
  const unsigned long long seq = ++ctx->pub_seq;
  LAIR_HIP(launch_publish_words(ctx, norms, nw, seq));
  const auto t0 = std::chrono::steady_clock::now();
  while (__atomic_load_n(&ctx->pub_host[0], __ATOMIC_ACQUIRE) != seq) {
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (ms > kPublishSpinMs) {
      LAIR_HIP(hipStreamSynchronize(ctx->stream));
      if (__atomic_load_n(&ctx->pub_host[0], __ATOMIC_ACQUIRE) != seq) {
        set_error("published norms did not arrive");
        return LAIR_ERR_DEVICE;
      }
      break;
    }
  }
  if (ctx->pub_host[1] != 0) return sync_status(ctx);
  *berr = berr_from_words(ctx->pub_host + 2, nrhs);
  return LAIR_OK;
}

}  // namespace lair

using namespace lair;

// ---------------------------------------------- allocation-free getrs ---

// Scratch of lair_dgetrs_device_noalloc: the row permutation, the inverses of
// the diagonal blocks of L and U, a staging copy of the right-hand sides and
// the triangular solves' tagged values, for n <= kNoAllocMaxN and
// nrhs <= kNoAllocMaxRhs (22.1 MiB).  One copy per GPU, for the life of the
// process; host code uses the per-device address from hipGetSymbolAddress,
// never the array's name.
constexpr int64_t kNoAllocMaxN = 16384;
constexpr int64_t kNoAllocMaxRhs = 16;
constexpr size_t kNoAllocPermBytes = sizeof(int) * kNoAllocMaxN;
constexpr size_t kNoAllocInvBytes = sizeof(double) * kNoAllocMaxN * TrsvBlock<double>::BS;
constexpr size_t kNoAllocTmpBytes = sizeof(double) * kNoAllocMaxN * kNoAllocMaxRhs;
constexpr size_t kNoAllocXtBytes = 2 * sizeof(unsigned long long) * kNoAllocMaxN * kNoAllocMaxRhs;
constexpr size_t kNoAllocBytes = kNoAllocPermBytes + 2 * kNoAllocInvBytes + kNoAllocTmpBytes + kNoAllocXtBytes;
static_assert(kNoAllocPermBytes % 256 == 0 && kNoAllocInvBytes % 256 == 0 && kNoAllocTmpBytes % 256 == 0,
              "scratch regions stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

static std::mutex g_scratch_mu;
static std::map<int, void *> g_scratch_ptr;

// The scratch array's address on `device`, which must be the current device.
static lair_status noalloc_scratch_address(int device, void **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return LAIR_OK;
  }
  void *p = nullptr;
  LAIR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = p;
  *out = p;
  return LAIR_OK;
}

// ------------------------------------------------------------------ C ABI

static std::mutex g_default_mu;
static lair_context *g_default = nullptr;

static lair_context *default_ctx() {
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default) {
    if (lair_context_create(&g_default) != LAIR_OK) g_default = nullptr;
  }
  return g_default;
}

// ---- host-pointer entry points ------------------------------------------

// Strided host views to and from dense row-major staging.  Complex elements
// are copied as their two real parts, so a caller's array needs only the
// alignment of the real type.
template <typename T>
static void gather(int64_t m, int64_t n, const T *a, int64_t rs, int64_t cs, T *dst, int64_t ld) {
  if constexpr (cplx_traits<T>::is) {
    typedef typename cplx_traits<T>::real R;
    const R *ar = reinterpret_cast<const R *>(a);
    R *dr = reinterpret_cast<R *>(dst);
    for (int64_t i = 0; i < m; ++i)
      for (int64_t j = 0; j < n; ++j) {
        dr[2 * (i * ld + j)] = ar[2 * (i * rs + j * cs)];
        dr[2 * (i * ld + j) + 1] = ar[2 * (i * rs + j * cs) + 1];
      }
  } else {
    for (int64_t i = 0; i < m; ++i)
      for (int64_t j = 0; j < n; ++j) dst[i * ld + j] = a[i * rs + j * cs];
  }
}
template <typename T>
static void scatter(int64_t m, int64_t n, const T *src, int64_t ld, T *a, int64_t rs, int64_t cs) {
  if constexpr (cplx_traits<T>::is) {
    typedef typename cplx_traits<T>::real R;
    const R *sr = reinterpret_cast<const R *>(src);
    R *ar = reinterpret_cast<R *>(a);
    for (int64_t i = 0; i < m; ++i)
      for (int64_t j = 0; j < n; ++j) {
        ar[2 * (i * rs + j * cs)] = sr[2 * (i * ld + j)];
        ar[2 * (i * rs + j * cs) + 1] = sr[2 * (i * ld + j) + 1];
      }
  } else {
    for (int64_t i = 0; i < m; ++i)
      for (int64_t j = 0; j < n; ++j) a[i * rs + j * cs] = src[i * ld + j];
  }
}

// Trailing-update arithmetic of the factorization of T: the real type's.
template <typename T>
static constexpr int gemm_mode_of() {
  return sizeof(typename cplx_traits<T>::real) == 8 ? kGemmF64 : kGemmF32;
}

// Complex getrs (getrs.rs:12-39 for Complex<R>): permutation from the pivots
// (range-checked on the device), then the substitutions of launch_cgetrs.
template <typename T>
static lair_status cgetrs_impl(lair_context *ctx, int64_t n64, int64_t nrhs64, const T *lu, int64_t lda,
                               const int64_t *ipiv, T *b, int64_t ldb) {
  if (n64 < 0) return -2;
  if (nrhs64 < 0) return -3;
  if (n64 == 0 || nrhs64 == 0) return LAIR_OK;
  if (lda < n64) return -5;
  if (ldb < nrhs64) return -8;
  if (n64 > kComplexGetrsMax || nrhs64 > INT32_MAX) {
    set_error("complex getrs supports n <= 16384");
    return LAIR_ERR_UNSUPPORTED;
  }
  const int n = (int)n64;
  LAIR_HIP(ensure(ctx->ws_perm, sizeof(int) * n));
  int *perm = static_cast<int *>(ctx->ws_perm.ptr);
  hipError_t e = launch_perm_from_ipiv(ctx, ipiv, n, perm);
  if (e == hipSuccess) e = launch_cgetrs<T>(ctx, n, (int)nrhs64, lu, lda, perm, b, ldb);
  return hip_status(e, "getrs");
}

template <typename T>
static lair_status host_getrf(int64_t m, int64_t n, T *a, int64_t rs, int64_t cs, int64_t *ipiv,
                              int64_t *singular) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (m > 0 && n > 0 && !a) return -3;
  const int64_t kmin = std::min(m, n);
  if (kmin > 0 && !ipiv) return -6;
  if (!singular) return -7;
  *singular = -1;
  if (m == 0 || n == 0) return LAIR_OK;
  lair_context *ctx = default_ctx();
  if (!ctx) return LAIR_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(ctx->mu);
  const int64_t ld = pad_ld<T>(n);
  std::vector<T> h((size_t)(m * ld));
  gather(m, n, a, rs, cs, h.data(), ld);
  LAIR_HIP(ensure(ctx->ws_tmp, sizeof(T) * m * ld));
  LAIR_HIP(ensure(ctx->ws_ipiv, sizeof(int64_t) * kmin));
  T *da = static_cast<T *>(ctx->ws_tmp.ptr);
  int64_t *dp = static_cast<int64_t *>(ctx->ws_ipiv.ptr);
  LAIR_HIP(hipMemcpyAsync(da, h.data(), sizeof(T) * m * ld, hipMemcpyHostToDevice, ctx->stream));
  lair_status st = getrf_impl<T>(ctx, m, n, da, ld, dp, 0, gemm_mode_of<T>(), singular);
  if (st != LAIR_OK) return st;
  LAIR_HIP(hipMemcpyAsync(h.data(), da, sizeof(T) * m * ld, hipMemcpyDeviceToHost, ctx->stream));
  LAIR_HIP(hipMemcpyAsync(ipiv, dp, sizeof(int64_t) * kmin, hipMemcpyDeviceToHost, ctx->stream));
  st = sync_status(ctx);
  if (st != LAIR_OK) return st;
  scatter(m, n, h.data(), ld, a, rs, cs);
  return LAIR_OK;
}


template <typename T>
static lair_status host_getrs(int64_t n, int64_t nrhs, const T *lu, int64_t lrs, int64_t lcs,
                              const int64_t *ipiv, T *b, int64_t brs, int64_t bcs, int32_t trans) {
  if (n < 0) return -1;
  if (nrhs < 0) return -2;
  if (n > 0 && (!lu || !ipiv)) return -3;
  for (int64_t i = 0; i < n; ++i)
    if (ipiv[i] < 0 || ipiv[i] >= n) {
      set_error("pivot index out of range");
      return LAIR_ERR_SHAPE;
    }
  if (n == 0 || nrhs == 0) return LAIR_OK;
  lair_context *ctx = default_ctx();
  if (!ctx) return LAIR_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(ctx->mu);
  const int64_t ld = pad_ld<T>(n);
  std::vector<T> h((size_t)(n * ld)), hb((size_t)(n * nrhs));
  gather(n, n, lu, lrs, lcs, h.data(), ld);
  gather(n, nrhs, b, brs, bcs, hb.data(), nrhs);
  LAIR_HIP(ensure(ctx->ws_tmp, sizeof(T) * n * ld));
  LAIR_HIP(ensure(ctx->ws_rhs, sizeof(T) * n * nrhs));
  LAIR_HIP(ensure(ctx->ws_ipiv, sizeof(int64_t) * n));
  T *da = static_cast<T *>(ctx->ws_tmp.ptr);
  T *db = static_cast<T *>(ctx->ws_rhs.ptr);
  int64_t *dp = static_cast<int64_t *>(ctx->ws_ipiv.ptr);
  LAIR_HIP(hipMemcpyAsync(da, h.data(), sizeof(T) * n * ld, hipMemcpyHostToDevice, ctx->stream));
  LAIR_HIP(hipMemcpyAsync(db, hb.data(), sizeof(T) * n * nrhs, hipMemcpyHostToDevice, ctx->stream));
  LAIR_HIP(hipMemcpyAsync(dp, ipiv, sizeof(int64_t) * n, hipMemcpyHostToDevice, ctx->stream));
  lair_status st;
  if constexpr (cplx_traits<T>::is)
    st = trans ? LAIR_ERR_UNSUPPORTED : cgetrs_impl<T>(ctx, n, nrhs, da, ld, dp, db, nrhs);
  else
    st = getrs_impl<T>(ctx, n, nrhs, da, ld, dp, db, nrhs, trans);
  if (st != LAIR_OK) return st;
  LAIR_HIP(hipMemcpyAsync(hb.data(), db, sizeof(T) * n * nrhs, hipMemcpyDeviceToHost, ctx->stream));
  st = sync_status(ctx);
  if (st != LAIR_OK) return st;
  scatter(n, nrhs, hb.data(), nrhs, b, brs, bcs);
  return LAIR_OK;
}


// equation::solve (src/equation.rs:32-60) for Complex<R>: the same checks
// and errors as lair_dsolve.
template <typename T>
static lair_status host_csolve(int64_t m, int64_t n, const T *a, int64_t rs, int64_t cs, int64_t blen,
                               const T *b, T *x) {
  if (m < 0 || n < 0 || blen < 0) return -1;
  if (m != n) {
    set_error("input matrix is not square");
    return LAIR_ERR_SHAPE;
  }
  if (blen != m) {
    set_error("The number of elements in `b` must be the same as the number of rows in `a`");
    return LAIR_ERR_SHAPE;
  }
  if (n == 0) return LAIR_OK;
  std::vector<T> w((size_t)(n * n));
  gather(n, n, a, rs, cs, w.data(), n);
  std::vector<int64_t> piv((size_t)n);
  int64_t sing = -1;
  lair_status st = host_getrf<T>(n, n, w.data(), n, 1, piv.data(), &sing);
  if (st != LAIR_OK) return st;
  if (sing >= 0) {
    set_error("`a` is a singular matrix");
    return LAIR_ERR_SINGULAR;
  }
  std::vector<T> xb((size_t)n);
  gather(n, 1, b, 1, 1, xb.data(), 1);
  st = host_getrs<T>(n, 1, w.data(), n, 1, piv.data(), xb.data(), 1, 1, 0);
  if (st != LAIR_OK) return st;
  scatter(n, 1, xb.data(), 1, x, 1, 1);
  return LAIR_OK;
}

// Complex device entry points: interleaved (re, im) storage, leading
// dimensions in complex elements, pointers aligned to the complex element.
template <typename T>
static lair_status cgetrf_device(lair_context *ctx, int64_t m, int64_t n, void *a, int64_t lda, int64_t *ipiv,
                                 int64_t nb, int64_t *singular) {
  if (!ctx) return -1;
  if (a && reinterpret_cast<uintptr_t>(a) % sizeof(T)) return -4;
  std::lock_guard<std::mutex> lk(ctx->mu);
  return getrf_impl<T>(ctx, m, n, static_cast<T *>(a), lda, ipiv, nb < 0 ? 0 : nb, gemm_mode_of<T>(), singular);
}
template <typename T>
static lair_status cgetrs_device(lair_context *ctx, int64_t n, int64_t nrhs, const void *lu, int64_t lda,
                                 const int64_t *ipiv, void *b, int64_t ldb) {
  if (!ctx) return -1;
  if ((lu && reinterpret_cast<uintptr_t>(lu) % sizeof(T)) || (b && reinterpret_cast<uintptr_t>(b) % sizeof(T)))
    return -4;
  std::lock_guard<std::mutex> lk(ctx->mu);
  lair_status st = cgetrs_impl<T>(ctx, n, nrhs, static_cast<const T *>(lu), lda, ipiv, static_cast<T *>(b), ldb);
  return st != LAIR_OK ? st : sync_status(ctx);
}

extern "C" {

const char *lair_last_error(void) { return g_last_error.c_str(); }

lair_status lair_context_create(lair_context **out) {
  if (!out) return -1;
  std::unique_ptr<lair_context> ctx(new lair_context());
  LAIR_HIP(hipGetDevice(&ctx->device));
  {
    const lair_status st = noalloc_scratch_address(ctx->device, &ctx->noalloc_scratch);
    if (st != LAIR_OK) return st;
  }
  LAIR_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  LAIR_HIP(hipMalloc(&ctx->rec, sizeof(unsigned long long) * 2 * kMaxPanelWG * 64));
  LAIR_HIP(hipMalloc(&ctx->drec, sizeof(unsigned long long) * 2 * 64));
  LAIR_HIP(hipMalloc(&ctx->crec, sizeof(unsigned long long) * kCaluWords));
  LAIR_HIP(hipMalloc(&ctx->err, sizeof(int32_t)));
  LAIR_HIP(hipMalloc(&ctx->rp_hdr, sizeof(unsigned long long) * kRpHdrTotal));
  LAIR_HIP(hipMalloc(&ctx->rp_rowd, sizeof(unsigned long long) * 2 * kMaxPanelWG * kRpRowWords));
  LAIR_HIP(hipMalloc(&ctx->sing, sizeof(int64_t)));
  LAIR_HIP(hipMalloc(&ctx->swap_rows, sizeof(int) * kMaxSwapRows));
  LAIR_HIP(hipMalloc(&ctx->swap_src, sizeof(int) * kMaxSwapRows));
  LAIR_HIP(hipMalloc(&ctx->swap_count, sizeof(int)));
  LAIR_HIP(hipMalloc(&ctx->swap_rows_alt, sizeof(int) * kMaxSwapRows));
  LAIR_HIP(hipMalloc(&ctx->swap_src_alt, sizeof(int) * kMaxSwapRows));
  LAIR_HIP(hipMalloc(&ctx->swap_count_alt, sizeof(int)));
  for (int q = 0; q < 2; ++q) {
    LAIR_HIP(hipMalloc(&ctx->pmap_rows[q], sizeof(int) * kMaxSwapRows));
    LAIR_HIP(hipMalloc(&ctx->pmap_src[q], sizeof(int) * kMaxSwapRows));
    LAIR_HIP(hipMalloc(&ctx->pmap_count[q], sizeof(int)));
  }
  // The look-ahead stream carries the bulk trailing updates; their grids are
  // capped (see launch_gemm_minus's max_wg) so that kReservedCUs CUs stay free
  // for the critical path on `stream`.
  LAIR_HIP(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
  {
    hipDeviceProp_t prop;
    LAIR_HIP(hipGetDeviceProperties(&prop, ctx->device));
    ctx->num_cu = prop.multiProcessorCount;
  }
  LAIR_HIP(hipEventCreateWithFlags(&ctx->ev_panel, hipEventDisableTiming));
  LAIR_HIP(hipEventCreateWithFlags(&ctx->ev_bulk, hipEventDisableTiming));
  LAIR_HIP(hipMalloc(&ctx->norms, sizeof(double) * 4096));
  // Pinned, coherent host words for the refinement's published norms; without
  // them the norms are read back with a stream synchronization.
  if (hipHostMalloc(reinterpret_cast<void **>(&ctx->pub_host), sizeof(unsigned long long) * (2 + 4096),
                    hipHostMallocCoherent | hipHostMallocMapped) == hipSuccess) {
    std::memset(ctx->pub_host, 0, sizeof(unsigned long long) * (2 + 4096));
    if (hipHostGetDevicePointer(reinterpret_cast<void **>(&ctx->pub_dev), ctx->pub_host, 0) != hipSuccess) {
      (void)hipHostFree(ctx->pub_host);
      ctx->pub_host = ctx->pub_dev = nullptr;
    }
  }
  (void)hipGetLastError();
  LAIR_HIP(launch_fill_u32(ctx.get(), reinterpret_cast<uint32_t *>(ctx->rec), 4 * kMaxPanelWG * 64, 0u));
  LAIR_HIP(launch_fill_u32(ctx.get(), reinterpret_cast<uint32_t *>(ctx->drec), 4 * 64, 0u));
  LAIR_HIP(launch_fill_u32(ctx.get(), reinterpret_cast<uint32_t *>(ctx->crec), 2 * kCaluWords, 0u));
  LAIR_HIP(launch_fill_u32(ctx.get(), reinterpret_cast<uint32_t *>(ctx->rp_hdr), 2 * kRpHdrTotal, 0u));
  LAIR_HIP(launch_fill_u32(ctx.get(), reinterpret_cast<uint32_t *>(ctx->rp_rowd),
                           2 * 2 * kMaxPanelWG * kRpRowWords, 0u));
  ctx->seq = 1;
  ctx->rp_seq = 1;
  if (const char *ev = std::getenv("LAIR_PANEL_MAX_WG")) ctx->rp_max_wg = std::max(1, std::atoi(ev));
  LAIR_HIP(launch_fill_u32(ctx.get(), reinterpret_cast<uint32_t *>(ctx->err), 1, 0u));
  LAIR_HIP(hipStreamSynchronize(ctx->stream));
  *out = ctx.release();
  return LAIR_OK;
}

lair_status lair_context_destroy(lair_context *ctx) {
  if (!ctx) return -1;
  hipStreamSynchronize(ctx->stream);
  for (auto e : ctx->event_pool) hipEventDestroy(e);
  hipFree(ctx->rec);
  hipFree(ctx->drec);
  hipFree(ctx->crec);
  hipFree(ctx->err);
  hipFree(ctx->rp_hdr);
  hipFree(ctx->rp_rowd);
  hipFree(ctx->sing);
  hipFree(ctx->norms);
  if (ctx->pub_host) hipHostFree(ctx->pub_host);
  hipFree(ctx->swap_rows);
  hipFree(ctx->swap_src);
  hipFree(ctx->swap_count);
  hipFree(ctx->swap_rows_alt);
  hipFree(ctx->swap_src_alt);
  hipFree(ctx->swap_count_alt);
  for (int q = 0; q < 2; ++q) {
    hipFree(ctx->pmap_rows[q]);
    hipFree(ctx->pmap_src[q]);
    hipFree(ctx->pmap_count[q]);
  }
  if (ctx->stream2) {
    hipStreamSynchronize(ctx->stream2);
    hipStreamDestroy(ctx->stream2);
  }
  if (ctx->ev_panel) hipEventDestroy(ctx->ev_panel);
  if (ctx->ev_bulk) hipEventDestroy(ctx->ev_bulk);
  for (Buffer *b : {&ctx->ws_a32, &ctx->ws_lu64, &ctx->ws_ipiv, &ctx->ws_x32, &ctx->ws_r,
                    &ctx->ws_tmp, &ctx->ws_rhs, &ctx->ws_perm, &ctx->ws_linv, &ctx->ws_uinv,
                    &ctx->ws_flags, &ctx->ws_tmp2, &ctx->ws_cexp, &ctx->ws_split})
    if (b->ptr) hipFree(b->ptr);
  hipStreamDestroy(ctx->stream);
  if (ctx == g_default) g_default = nullptr;
  delete ctx;
  return LAIR_OK;
}

lair_status lair_context_synchronize(lair_context *ctx) {
  if (!ctx) ctx = default_ctx();
  if (!ctx) return LAIR_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(ctx->mu);
  return sync_status(ctx);
}

lair_status lair_context_set_wait_limit(lair_context *ctx, int64_t polls) {
  if (!ctx) return -1;
  if (polls < 0 || polls > INT32_MAX) return -2;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->wait_limit = (int32_t)polls;
  return LAIR_OK;
}

void *lair_context_stream(lair_context *ctx) {
  if (!ctx) ctx = default_ctx();
  return ctx ? (void *)ctx->stream : nullptr;
}

lair_status lair_malloc(void **dptr, uint64_t bytes) {
  if (!dptr) return -1;
  LAIR_HIP(hipMalloc(dptr, bytes ? bytes : 1));
  return LAIR_OK;
}
lair_status lair_free(void *dptr) {
  LAIR_HIP(hipFree(dptr));
  return LAIR_OK;
}
lair_status lair_memcpy_h2d(void *dst, const void *src, uint64_t bytes) {
  LAIR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return LAIR_OK;
}
lair_status lair_memcpy_d2h(void *dst, const void *src, uint64_t bytes) {
  LAIR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return LAIR_OK;
}
lair_status lair_memcpy_d2d(void *dst, const void *src, uint64_t bytes) {
  // A device-to-device hipMemcpy may return before the copy completes; the
  // copies here are blocking like the other two, so work a caller then queues
  // on a context's (non-blocking) stream sees the copied data.
  LAIR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  LAIR_HIP(hipDeviceSynchronize());
  return LAIR_OK;
}

lair_status lair_dgetrf_device(lair_context *ctx, int64_t m, int64_t n, double *a, int64_t lda,
                               int64_t *ipiv, int64_t nb, int64_t *singular) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (nb < 0) {
    lair_status st = tune_nb<double>(ctx, m, n, a, lda, kGemmF64, &nb);
    if (st != LAIR_OK) return st;
  }
  return getrf_impl<double>(ctx, m, n, a, lda, ipiv, nb, kGemmF64, singular);
}

lair_status lair_sgetrf_device(lair_context *ctx, int64_t m, int64_t n, float *a, int64_t lda,
                               int64_t *ipiv, int64_t nb, int64_t *singular) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (nb < 0) {
    lair_status st = tune_nb<float>(ctx, m, n, a, lda, kGemmF32, &nb);
    if (st != LAIR_OK) return st;
  }
  return getrf_impl<float>(ctx, m, n, a, lda, ipiv, nb, kGemmF32, singular);
}

lair_status lair_sgetrf_pivot_device(lair_context *ctx, int64_t m, int64_t n, float *a, int64_t lda,
                                     int64_t *ipiv, int64_t nb, int32_t pivoting, int32_t prec,
                                     int64_t *singular) {
  if (!ctx) return -1;
  if (pivoting != 0 && pivoting != LAIR_PIVOT_TOURNAMENT) return -8;
  if (prec != LAIR_PREC_FP32 && prec != LAIR_PREC_BF16X6 && prec != LAIR_PREC_BF16) return -9;
  const int mode = prec == LAIR_PREC_FP32 ? kGemmF32 : (prec == LAIR_PREC_BF16X6 ? kGemmBF16x6 : kGemmBF16);
  std::lock_guard<std::mutex> lk(ctx->mu);
  return getrf_impl<float>(ctx, m, n, a, lda, ipiv, nb, mode, singular, nullptr,
                           pivoting == LAIR_PIVOT_TOURNAMENT);
}

lair_status lair_dgetrs_device(lair_context *ctx, int64_t n, int64_t nrhs, const double *lu,
                               int64_t lda, const int64_t *ipiv, double *b, int64_t ldb,
                               int32_t trans) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  lair_status st = getrs_impl<double>(ctx, n, nrhs, lu, lda, ipiv, b, ldb, trans);
  return st != LAIR_OK ? st : sync_status(ctx);
}

lair_status lair_sgetrs_device(lair_context *ctx, int64_t n, int64_t nrhs, const float *lu,
                               int64_t lda, const int64_t *ipiv, float *b, int64_t ldb,
                               int32_t trans) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  lair_status st = getrs_impl<float>(ctx, n, nrhs, lu, lda, ipiv, b, ldb, trans);
  return st != LAIR_OK ? st : sync_status(ctx);
}

}  // extern "C"

// C -= A B on device data (src/blas/gemm.rs:6-32 with alpha = -1 and no
// conjugation), on the trailing-update kernels.
template <typename T>
static lair_status gemm_minus_impl(lair_context *ctx, int mode, int64_t m, int64_t n, int64_t k, const T *a,
                                   int64_t lda, const T *b, int64_t ldb, T *c, int64_t ldc, int arg0) {
  const int64_t lim = (int64_t)1 << 30;
  if (m < 0 || m > lim) return -arg0;
  if (n < 0 || n > lim) return -(arg0 + 1);
  if (k < 0 || k > lim) return -(arg0 + 2);
  if (m == 0 || n == 0 || k == 0) return LAIR_OK;
  if (!a) return -(arg0 + 3);
  if (lda < k) return -(arg0 + 4);
  if (!b) return -(arg0 + 5);
  if (ldb < n) return -(arg0 + 6);
  if (!c) return -(arg0 + 7);
  if (ldc < n) return -(arg0 + 8);
  // The tile kernels address 128 rows of each operand with 32-bit offsets.
  const int64_t ldmax = std::max(std::max(lda, ldb), ldc);
  if (ldmax > ((int64_t)1 << 31) / (128 * (int64_t)sizeof(T))) {
    set_error("gemm: leading dimension too large");
    return LAIR_ERR_UNSUPPORTED;
  }
  hipError_t e = launch_gemm_minus<T>(ctx, mode, (int)m, (int)n, (int)k, a, lda, false, b, ldb, c, ldc,
                                      LAIR_KCLASS_GEMM);
  if (e != hipSuccess) return hip_status(e, "gemm");
  return sync_status(ctx);
}

extern "C" {

lair_status lair_dgemm_minus_device(lair_context *ctx, int64_t m, int64_t n, int64_t k, const double *a,
                                    int64_t lda, const double *b, int64_t ldb, double *c, int64_t ldc) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  return gemm_minus_impl<double>(ctx, kGemmF32, m, n, k, a, lda, b, ldb, c, ldc, 2);
}

lair_status lair_sgemm_minus_device(lair_context *ctx, int32_t prec, int64_t m, int64_t n, int64_t k,
                                    const float *a, int64_t lda, const float *b, int64_t ldb, float *c,
                                    int64_t ldc) {
  if (!ctx) return -1;
  if (prec != LAIR_PREC_FP32 && prec != LAIR_PREC_BF16X6 && prec != LAIR_PREC_BF16) return -2;
  const int mode = prec == LAIR_PREC_FP32 ? kGemmF32 : (prec == LAIR_PREC_BF16X6 ? kGemmBF16x6 : kGemmBF16);
  std::lock_guard<std::mutex> lk(ctx->mu);
  return gemm_minus_impl<float>(ctx, mode, m, n, k, a, lda, b, ldb, c, ldc, 3);
}

lair_status lair_zgetrf_device(lair_context *ctx, int64_t m, int64_t n, double *a, int64_t lda, int64_t *ipiv,
                               int64_t nb, int64_t *singular) {
  return cgetrf_device<zcplx>(ctx, m, n, a, lda, ipiv, nb, singular);
}
lair_status lair_cgetrf_device(lair_context *ctx, int64_t m, int64_t n, float *a, int64_t lda, int64_t *ipiv,
                               int64_t nb, int64_t *singular) {
  return cgetrf_device<ccplx>(ctx, m, n, a, lda, ipiv, nb, singular);
}
lair_status lair_zgetrs_device(lair_context *ctx, int64_t n, int64_t nrhs, const double *lu, int64_t lda,
                               const int64_t *ipiv, double *b, int64_t ldb) {
  return cgetrs_device<zcplx>(ctx, n, nrhs, lu, lda, ipiv, b, ldb);
}
lair_status lair_cgetrs_device(lair_context *ctx, int64_t n, int64_t nrhs, const float *lu, int64_t lda,
                               const int64_t *ipiv, float *b, int64_t ldb) {
  return cgetrs_device<ccplx>(ctx, n, nrhs, lu, lda, ipiv, b, ldb);
}

lair_status lair_dgetrs_device_noalloc(lair_context *ctx, int64_t n, int64_t nrhs, const double *lu,
                                       int64_t lda, const int64_t *ipiv, double *b, int64_t ldb,
                                       int32_t trans) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (n > kNoAllocMaxN || nrhs > kNoAllocMaxRhs) {
    set_error("lair_dgetrs_device_noalloc: n <= 16384 and nrhs <= 16 (the size of its scratch)");
    return LAIR_ERR_UNSUPPORTED;
  }
  if (!ctx->noalloc_scratch) {
    set_error("lair_dgetrs_device_noalloc: no scratch address on this device");
    return LAIR_ERR_DEVICE;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (n == 0 || nrhs == 0) return LAIR_OK;
  char *base = static_cast<char *>(ctx->noalloc_scratch);
  GetrsScratch<double> sc;
  sc.perm = reinterpret_cast<int *>(base);
  sc.linv = reinterpret_cast<double *>(base + kNoAllocPermBytes);
  sc.uinv = reinterpret_cast<double *>(base + kNoAllocPermBytes + kNoAllocInvBytes);
  sc.tmp = reinterpret_cast<double *>(base + kNoAllocPermBytes + 2 * kNoAllocInvBytes);
  unsigned long long *xt =
      reinterpret_cast<unsigned long long *>(base + kNoAllocPermBytes + 2 * kNoAllocInvBytes + kNoAllocTmpBytes);
  // The tagged values start from zero on every call: the array keeps the
  // last call's contents, possibly from another context on this GPU.
  hipError_t e = launch_fill_u32(ctx, reinterpret_cast<uint32_t *>(xt), 4 * n * nrhs, 0u);
  if (e != hipSuccess) return hip_status(e, "fill");
  ctx->xt_override = xt;
  const lair_status st = getrs_impl<double>(ctx, n, nrhs, lu, lda, ipiv, b, ldb, trans, &sc);
  ctx->xt_override = nullptr;
  return st != LAIR_OK ? st : sync_status(ctx);
}

lair_status lair_dgetrf(int64_t m, int64_t n, double *a, int64_t rs, int64_t cs, int64_t *ipiv,
                        int64_t *singular) {
  return host_getrf<double>(m, n, a, rs, cs, ipiv, singular);
}
lair_status lair_sgetrf(int64_t m, int64_t n, float *a, int64_t rs, int64_t cs, int64_t *ipiv,
                        int64_t *singular) {
  return host_getrf<float>(m, n, a, rs, cs, ipiv, singular);
}

lair_status lair_dgetrs(int64_t n, int64_t nrhs, const double *lu, int64_t lrs, int64_t lcs,
                        const int64_t *ipiv, double *b, int64_t brs, int64_t bcs, int32_t trans) {
  return host_getrs<double>(n, nrhs, lu, lrs, lcs, ipiv, b, brs, bcs, trans);
}
lair_status lair_sgetrs(int64_t n, int64_t nrhs, const float *lu, int64_t lrs, int64_t lcs,
                        const int64_t *ipiv, float *b, int64_t brs, int64_t bcs, int32_t trans) {
  return host_getrs<float>(n, nrhs, lu, lrs, lcs, ipiv, b, brs, bcs, trans);
}

lair_status lair_zgetrf(int64_t m, int64_t n, double *a, int64_t rs, int64_t cs, int64_t *ipiv,
                        int64_t *singular) {
  return host_getrf<zcplx>(m, n, reinterpret_cast<zcplx *>(a), rs, cs, ipiv, singular);
}
lair_status lair_cgetrf(int64_t m, int64_t n, float *a, int64_t rs, int64_t cs, int64_t *ipiv,
                        int64_t *singular) {
  return host_getrf<ccplx>(m, n, reinterpret_cast<ccplx *>(a), rs, cs, ipiv, singular);
}
lair_status lair_zgetrs(int64_t n, int64_t nrhs, const double *lu, int64_t lrs, int64_t lcs,
                        const int64_t *ipiv, double *b, int64_t brs, int64_t bcs) {
  return host_getrs<zcplx>(n, nrhs, reinterpret_cast<const zcplx *>(lu), lrs, lcs, ipiv,
                           reinterpret_cast<zcplx *>(b), brs, bcs, 0);
}
lair_status lair_cgetrs(int64_t n, int64_t nrhs, const float *lu, int64_t lrs, int64_t lcs,
                        const int64_t *ipiv, float *b, int64_t brs, int64_t bcs) {
  return host_getrs<ccplx>(n, nrhs, reinterpret_cast<const ccplx *>(lu), lrs, lcs, ipiv,
                           reinterpret_cast<ccplx *>(b), brs, bcs, 0);
}

lair_status lair_dlaswp(int64_t ncols, double *a, int64_t rs, int64_t cs, int64_t begin,
                        int64_t npiv, const int64_t *ipiv) {
  if (ncols < 0) return -1;
  if (npiv < 0) return -6;
  if (begin < 0) return -5;
  if (npiv == 0 || begin >= npiv) return LAIR_OK;
  if (!a) return -2;
  if (!ipiv) return -7;
  int64_t rows = npiv;
  for (int64_t i = begin; i < npiv; ++i) {
    if (ipiv[i] < 0) return -7;
    rows = std::max(rows, ipiv[i] + 1);
  }
  // The reference swaps the first column pair even when ncols == 0
  // (laswp.rs:28-38); keep that by treating 0 as 1.
  const int64_t nc = ncols == 0 ? 1 : ncols;
  lair_context *ctx = default_ctx();
  if (!ctx) return LAIR_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(ctx->mu);
  std::vector<double> h((size_t)(rows * nc));
  gather(rows, nc, a, rs, cs, h.data(), nc);
  LAIR_HIP(ensure(ctx->ws_tmp, sizeof(double) * rows * nc));
  LAIR_HIP(ensure(ctx->ws_ipiv, sizeof(int64_t) * npiv));
  double *da = static_cast<double *>(ctx->ws_tmp.ptr);
  int64_t *dp = static_cast<int64_t *>(ctx->ws_ipiv.ptr);
  LAIR_HIP(hipMemcpyAsync(da, h.data(), sizeof(double) * rows * nc, hipMemcpyHostToDevice, ctx->stream));
  LAIR_HIP(hipMemcpyAsync(dp, ipiv, sizeof(int64_t) * npiv, hipMemcpyHostToDevice, ctx->stream));
  hipError_t e = launch_laswp<double>(ctx, da, nc, 0, (int)nc, 0, 0, dp, (int)begin, (int)npiv, false);
  if (e != hipSuccess) return hip_status(e, "laswp");
  LAIR_HIP(hipMemcpyAsync(h.data(), da, sizeof(double) * rows * nc, hipMemcpyDeviceToHost, ctx->stream));
  lair_status st = sync_status(ctx);
  if (st != LAIR_OK) return st;
  scatter(rows, nc, h.data(), nc, a, rs, cs);
  return LAIR_OK;
}

lair_status lair_dsolve(int64_t m, int64_t n, const double *a, int64_t rs, int64_t cs,
                        int64_t blen, const double *b, double *x) {
  if (m < 0 || n < 0 || blen < 0) return -1;
  if (m != n) {
    set_error("input matrix is not square");
    return LAIR_ERR_SHAPE;
  }
  if (blen != m) {
    set_error("The number of elements in `b` must be the same as the number of rows in `a`");
    return LAIR_ERR_SHAPE;
  }
  if (n == 0) return LAIR_OK;
  std::vector<double> w((size_t)(n * n));
  gather(n, n, a, rs, cs, w.data(), n);
  std::vector<int64_t> piv((size_t)n);
  int64_t sing = -1;
  lair_status st = lair_dgetrf(n, n, w.data(), n, 1, piv.data(), &sing);
  if (st != LAIR_OK) return st;
  if (sing >= 0) {
    set_error("`a` is a singular matrix");
    return LAIR_ERR_SINGULAR;
  }
  std::memcpy(x, b, sizeof(double) * n);
  return lair_dgetrs(n, 1, w.data(), n, 1, piv.data(), x, 1, 1, 0);
}

lair_status lair_zsolve(int64_t m, int64_t n, const double *a, int64_t rs, int64_t cs, int64_t blen,
                        const double *b, double *x) {
  return host_csolve<zcplx>(m, n, reinterpret_cast<const zcplx *>(a), rs, cs, blen,
                            reinterpret_cast<const zcplx *>(b), reinterpret_cast<zcplx *>(x));
}
lair_status lair_csolve(int64_t m, int64_t n, const float *a, int64_t rs, int64_t cs, int64_t blen,
                        const float *b, float *x) {
  return host_csolve<ccplx>(m, n, reinterpret_cast<const ccplx *>(a), rs, cs, blen,
                            reinterpret_cast<const ccplx *>(b), reinterpret_cast<ccplx *>(x));
}

// ---- device solvers ------------------------------------------------------

static lair_status dgesv_device_locked(lair_context *ctx, int64_t n, int64_t nrhs, const double *a,
                                       int64_t lda, const double *b, int64_t ldb, double *x,
                                       int64_t ldx) {
  const int64_t ld = pad_ld<double>(n);
  LAIR_HIP(ensure(ctx->ws_lu64, sizeof(double) * n * ld));
  LAIR_HIP(ensure(ctx->ws_ipiv, sizeof(int64_t) * n));
  LAIR_HIP(ensure(ctx->ws_perm, sizeof(int) * n));
  double *lu = static_cast<double *>(ctx->ws_lu64.ptr);
  int64_t *piv = static_cast<int64_t *>(ctx->ws_ipiv.ptr);
  int *perm = static_cast<int *>(ctx->ws_perm.ptr);
  hipError_t e = launch_copy_d2d(ctx, (int)n, (int)n, a, lda, lu, ld);
  if (e != hipSuccess) return hip_status(e, "copy");
  int64_t sing = -1;
  lair_status st = getrf_impl<double>(ctx, n, n, lu, ld, piv, 0, kGemmF64, &sing, perm);
  if (st != LAIR_OK) return st;
  if (sing >= 0) {
    set_error("`a` is a singular matrix");
    return LAIR_ERR_SINGULAR;
  }
  e = launch_copy_d2d(ctx, (int)n, (int)nrhs, b, ldb, x, ldx);
  if (e != hipSuccess) return hip_status(e, "copy");
  if (n <= kExactOrderMax) return getrs_impl<double>(ctx, n, nrhs, lu, ld, piv, x, ldx, 0);
  SolveAux<double> aux;
  st = make_aux<double>(ctx, (int)n, lu, ld, piv, perm, &aux);
  if (st != LAIR_OK) return st;
  return getrs_aux<double>(ctx, (int)n, (int)nrhs, lu, ld, aux, x, ldx, 0);
}

lair_status lair_dgesv_device(lair_context *ctx, int64_t n, int64_t nrhs, const double *a,
                              int64_t lda, const double *b, int64_t ldb, double *x, int64_t ldx) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (n == 0 || nrhs == 0) return LAIR_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  lair_status st = dgesv_device_locked(ctx, n, nrhs, a, lda, b, ldb, x, ldx);
  if (st != LAIR_OK) return st;
  return sync_status(ctx);
}

static lair_status berr_locked(lair_context *ctx, int n, int nrhs, const double *a, int64_t lda,
                               const double *b, int64_t ldb, const double *x, int64_t ldx,
                               double *berr, bool norm_a_ready) {
  LAIR_HIP(ensure(ctx->ws_r, sizeof(double) * n * nrhs));
  double *r = static_cast<double *>(ctx->ws_r.ptr);
  unsigned long long *norms = reinterpret_cast<unsigned long long *>(ctx->norms);
  hipError_t e;
  if (!norm_a_ready) {
    e = launch_fill_f64(ctx, ctx->norms, 1, 0.0);
    if (e != hipSuccess) return hip_status(e, "fill");
    e = launch_rows_norm(ctx, n, n, a, lda, norms);
    if (e != hipSuccess) return hip_status(e, "norm");
  }
  e = launch_residual(ctx, n, nrhs, a, lda, x, ldx, b, ldb, r, nrhs);
  if (e != hipSuccess) return hip_status(e, "residual");
  return backward_error_from(ctx, n, nrhs, r, nrhs, x, ldx, norms, berr);
}

lair_status lair_dbackward_error(lair_context *ctx, int64_t n, int64_t nrhs, const double *a,
                                 int64_t lda, const double *b, int64_t ldb, const double *x,
                                 int64_t ldx, double *berr) {
  if (!ctx) return -1;
  if (!berr) return -10;
  if (n <= 0 || nrhs <= 0) {
    *berr = 0.0;
    return LAIR_OK;
  }
  if (nrhs > 2000) return -3;
  std::lock_guard<std::mutex> lk(ctx->mu);
  return berr_locked(ctx, (int)n, (int)nrhs, a, lda, b, ldb, x, ldx, berr, false);
}

lair_status lair_dgesv_mixed_device(lair_context *ctx, int64_t n64, int64_t nrhs64, const double *a,
                                    int64_t lda, const double *b, int64_t ldb, double *x,
                                    int64_t ldx, int32_t prec, int32_t max_iter, double tol,
                                    lair_ir_info *info) {
  if (!ctx) return -1;
  if (n64 < 0) return -2;
  if (nrhs64 < 0 || nrhs64 > 2000) return -3;
  const bool tournament = (prec & LAIR_PIVOT_TOURNAMENT) != 0;
  prec &= ~LAIR_PIVOT_TOURNAMENT;
  if (prec < LAIR_PREC_FP64 || prec > LAIR_PREC_BF16) return -11;
  if (max_iter < 0 || max_iter > 30) return -12;
  lair_ir_info local;
  if (!info) info = &local;
  std::memset(info, 0, sizeof(*info));
  info->precision = prec;
  if (n64 == 0 || nrhs64 == 0) {
    info->converged = 1;
    return LAIR_OK;
  }
  const int n = (int)n64, nrhs = (int)nrhs64;
  if (tol <= 0.0) tol = (double)n * DBL_EPSILON;
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipError_t e;
  unsigned long long *norms = reinterpret_cast<unsigned long long *>(ctx->norms);
  if (prec == LAIR_PREC_FP64) {
    lair_status st = dgesv_device_locked(ctx, n, nrhs, a, lda, b, ldb, x, ldx);
    if (st != LAIR_OK) return st;
    double be;
    st = berr_locked(ctx, n, nrhs, a, lda, b, ldb, x, ldx, &be, false);
    if (st != LAIR_OK) return st;
    info->backward_error = be;
    info->history[0] = be;
    info->converged = be <= tol;
    return LAIR_OK;
  }
  const int mode = prec == LAIR_PREC_FP32 ? kGemmF32 : (prec == LAIR_PREC_BF16X6 ? kGemmBF16x6 : kGemmBF16);
  const int64_t ld32 = pad_ld<float>(n);
  LAIR_HIP(ensure(ctx->ws_a32, sizeof(float) * n * ld32));
  LAIR_HIP(ensure(ctx->ws_ipiv, sizeof(int64_t) * n));
  LAIR_HIP(ensure(ctx->ws_perm, sizeof(int) * n));
  LAIR_HIP(ensure(ctx->ws_x32, sizeof(float) * n * nrhs));
  LAIR_HIP(ensure(ctx->ws_r, sizeof(double) * n * nrhs));
  float *a32 = static_cast<float *>(ctx->ws_a32.ptr);
  int64_t *piv = static_cast<int64_t *>(ctx->ws_ipiv.ptr);
  int *perm = static_cast<int *>(ctx->ws_perm.ptr);
  float *d32 = static_cast<float *>(ctx->ws_x32.ptr);
  double *r = static_cast<double *>(ctx->ws_r.ptr);
  e = launch_fill_f64(ctx, ctx->norms, 1, 0.0);
  if (e != hipSuccess) return hip_status(e, "fill");
  e = launch_convert_d2s_norm(ctx, n, n, a, lda, a32, ld32, norms);
  if (e != hipSuccess) return hip_status(e, "convert");
  lair_status st = getrf_impl<float>(ctx, n, n, a32, ld32, piv, 0, mode, nullptr, perm, tournament);
  if (st != LAIR_OK) return st;
  SolveAux<float> aux;
  if (n > kExactOrderMax) {
    st = make_aux<float>(ctx, n, a32, ld32, piv, perm, &aux);
    if (st != LAIR_OK) return st;
  }
  // d = (LU)^-1 P rhs in FP32: the row gather and the FP64->FP32 conversion
  // are one pass.
  auto lowprec_solve = [&](const double *rhs, int64_t ldr) -> lair_status {
    if (n <= kExactOrderMax) {
      hipError_t e2 = launch_copy_d2s(ctx, n, nrhs, rhs, ldr, d32, nrhs);
      if (e2 != hipSuccess) return hip_status(e2, "copy");
      return getrs_impl<float>(ctx, n, nrhs, a32, ld32, piv, d32, nrhs, 0);
    }
    LAIR_HIP(ensure(ctx->ws_tmp2, sizeof(float) * (size_t)n * nrhs));
    float *t = static_cast<float *>(ctx->ws_tmp2.ptr);
    hipError_t e2 = launch_permute_rows<double, float>(ctx, n, nrhs, aux.perm, rhs, ldr, t, nrhs, false);
    if (e2 != hipSuccess) return hip_status(e2, "permute");
    e2 = launch_trsv_dataflow<float>(ctx, a32, ld32, aux.linv, false, t, t, nrhs, n, nrhs, true, false);
    if (e2 != hipSuccess) return hip_status(e2, "trsv");
    e2 = launch_trsv_dataflow<float>(ctx, a32, ld32, aux.uinv, false, t, d32, nrhs, n, nrhs, false,
                                     false);
    return hip_status(e2, "trsv");
  };
  bool converged = false;
  int it = 0;
  double be = INFINITY;
  // The fused step's workgroups (128 rows of the solve each, and 48 rows of
  // the residual for one right-hand side, 24 for up to four) must all be
  // resident at two per CU, or the residual's later workgroups would start
  // only after the solve.
  const int64_t fused_wgs = (n + 127) / 128 + (nrhs == 1 ? (n + 47) / 48 : (n + 23) / 24);
  if (n > kExactOrderMax && nrhs <= 4 && fused_wgs <= 2 * (int64_t)ctx->num_cu) {
    // Each step is the gather of the right-hand side (which also clears the
    // norm words), the forward solve, and one launch for the backward solve,
    // x += d and the residual with its norms.  Step 0 solves with b from
    // x = 0, so x = d exactly as a copy would give.
    LAIR_HIP(ensure(ctx->ws_tmp2, sizeof(float) * (size_t)n * nrhs));
    float *t = static_cast<float *>(ctx->ws_tmp2.ptr);
    for (;;) {
      const double *rhs = it == 0 ? b : r;
      const int64_t ldr = it == 0 ? ldb : nrhs;
      e = launch_permute_rows<double, float>(ctx, n, nrhs, aux.perm, rhs, ldr, t, nrhs, false, norms + 1,
                                             2 * nrhs);
      if (e != hipSuccess) return hip_status(e, "permute");
      e = launch_trsv_dataflow<float>(ctx, a32, ld32, aux.linv, false, t, t, nrhs, n, nrhs, true, false);
      if (e != hipSuccess) return hip_status(e, "trsv");
      e = launch_trsv_ir(ctx, a32, ld32, aux.uinv, t, d32, n, nrhs, x, ldx, it == 0, a, lda, b, ldb, r, nrhs,
                         norms + 1 + nrhs, norms + 1);
      if (e != hipSuccess) return hip_status(e, "refinement step");
      st = backward_error_published(ctx, nrhs, norms, &be);
      if (st != LAIR_OK) return st;
      if (it < 32) info->history[it] = be;
      if (be <= tol) {
        converged = true;
        break;
      }
      if (it >= max_iter || !std::isfinite(be)) break;
      ++it;
    }
  } else {
    st = lowprec_solve(b, ldb);
    if (st != LAIR_OK) return st;
    e = launch_copy_s2d(ctx, n, nrhs, d32, nrhs, x, ldx);
    if (e != hipSuccess) return hip_status(e, "copy");
    // The residual kernel folds in the norms of r and x; their words are
    // cleared here once and then by each correction's x += d.
    e = launch_fill_f64(ctx, reinterpret_cast<double *>(norms + 1), 2 * nrhs, 0.0);
    if (e != hipSuccess) return hip_status(e, "fill");
    for (;;) {
      e = launch_residual(ctx, n, nrhs, a, lda, x, ldx, b, ldb, r, nrhs, norms + 1, norms + 1 + nrhs);
      if (e != hipSuccess) return hip_status(e, "residual");
      st = backward_error_published(ctx, nrhs, norms, &be);
      if (st != LAIR_OK) return st;
      if (it < 32) info->history[it] = be;
      if (be <= tol) {
        converged = true;
        break;
      }
      if (it >= max_iter || !std::isfinite(be)) break;
      // Correction d = LU^-1 r in the factorization precision, x += d in FP64.
      st = lowprec_solve(r, nrhs);
      if (st != LAIR_OK) return st;
      e = launch_axpy_s2d(ctx, n, nrhs, d32, nrhs, x, ldx, norms + 1, 2 * nrhs);
      if (e != hipSuccess) return hip_status(e, "axpy");
      ++it;
    }
  }
  info->iterations = it;
  info->backward_error = be;
  info->converged = converged ? 1 : 0;
  if (converged) return LAIR_OK;
  // Fallback: FP64 factorization.
  st = dgesv_device_locked(ctx, n, nrhs, a, lda, b, ldb, x, ldx);
  if (st != LAIR_OK) return st;
  st = berr_locked(ctx, n, nrhs, a, lda, b, ldb, x, ldx, &be, true);
  if (st != LAIR_OK) return st;
  info->fell_back = 1;
  info->precision = LAIR_PREC_FP64;
  info->backward_error = be;
  return LAIR_ERR_NOCONV;
}

// ---- batched -------------------------------------------------------------

struct lair_batched_plan {
  lair_context *ctx = nullptr;
  int prec = 0;
  int count = 0;
  void *a = nullptr, *b = nullptr;
  int64_t *ipiv = nullptr;
  int *info = nullptr;
  void *meta = nullptr;  // one device allocation for the metadata arrays
  // One persistent launch per size class, largest systems first; each class
  // has its own launch shape.
  std::vector<lair::BatchLaunch> launches;
  size_t a_elems = 0, b_elems = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static hipError_t batched_launch(lair_batched_plan *pl, const lair::BatchLaunch &bl, void *a, void *b) {
  if (pl->prec == 0)
    return launch_batched_gesv<double>(pl->ctx, bl, static_cast<double *>(a), static_cast<double *>(b),
                                       pl->ipiv, pl->info);
  return launch_batched_gesv<float>(pl->ctx, bl, static_cast<float *>(a), static_cast<float *>(b),
                                    pl->ipiv, pl->info);
}

// Candidate launch shapes: panel width, LDS budget for whole-system-in-LDS,
// workgroups per CU.  The LDS budget trades systems solved on chip against
// workgroups resident per CU; 3 and 4 workgroups per CU use kernels whose
// registers are capped to fit that many.
static std::vector<lair::BatchLaunch> batched_candidates(const lair::BatchLaunch &base, int esize,
                                                        int ncu) {
  std::vector<lair::BatchLaunch> out;
  const size_t lds_cap = 160 * 1024;
  for (int nb : {16, 8})
    for (int per_cu : {1, 2, 3, 4})
      for (int frac : {0, 1, 2}) {
        lair::BatchLaunch c = base;
        c.nb = nb;
        c.occ = per_cu >= 3 ? per_cu : 1;
        const size_t fixed = batched_lds_bytes_for(esize, nb, 0);
        const size_t per_wg = lds_cap / per_cu;
        if (per_wg <= fixed + 1024) continue;
        const size_t avail = (per_wg - fixed - 1024) / esize;  // elements
        c.lds_max_elems = frac == 0 ? 0 : (frac == 1 ? (int)(avail / 2) : (int)avail);
        c.grid = ncu * per_cu;
        out.push_back(c);
      }
  return out;
}

extern "C" {

lair_status lair_gesv_batched_plan(lair_context *ctx, int32_t precision, int32_t count,
                                   const int32_t *n, const int32_t *nrhs, void *a, void *b,
                                   int64_t *ipiv, int32_t *info, int32_t flags,
                                   lair_batched_plan **plan) {
  if (!ctx) return -1;
  if (precision != 0 && precision != 1) return -2;
  if (count < 0) return -3;
  if (!plan) return -11;
  if (count > 0 && (!n || !nrhs)) return -4;
  for (int i = 0; i < count; ++i) {
    if (n[i] < 1 || n[i] > 256) return -4;
    if (nrhs[i] < 1 || nrhs[i] > 16) return -5;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  std::unique_ptr<lair_batched_plan> pl(new lair_batched_plan());
  pl->ctx = ctx;
  pl->prec = precision;
  pl->count = count;
  pl->a = a;
  pl->b = b;
  pl->ipiv = ipiv;
  pl->info = info;
  const int esize = precision == 0 ? 8 : 4;
  std::vector<int64_t> aoff(count), boff(count), poff(count);
  std::vector<int> order(count);
  int64_t sa = 0, sb = 0, sp = 0;
  for (int i = 0; i < count; ++i) {
    aoff[i] = sa;
    boff[i] = sb;
    poff[i] = sp;
    sa += (int64_t)n[i] * n[i];
    sb += (int64_t)n[i] * nrhs[i];
    sp += n[i];
  }
  pl->a_elems = (size_t)sa;
  pl->b_elems = (size_t)sb;
  // Largest systems first (by order, then right-hand sides), so that the
  // queue hands out the long systems early and a size class is a prefix.
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return n[x] != n[y] ? n[x] > n[y] : nrhs[x] > nrhs[y];
  });
  const size_t bytes = sizeof(int) * 2 * count + sizeof(int64_t) * 3 * count + sizeof(int) * count + 64;
  LAIR_HIP(hipMalloc(&pl->meta, bytes));
  char *m = static_cast<char *>(pl->meta);
  int64_t *d_aoff = reinterpret_cast<int64_t *>(m);
  int64_t *d_boff = d_aoff + count;
  int64_t *d_poff = d_boff + count;
  int *d_n = reinterpret_cast<int *>(d_poff + count);
  int *d_r = d_n + count;
  int *d_order = d_r + count;
  int *d_queue = d_order + count;
  if (count > 0) {
    LAIR_HIP(hipMemcpy(d_aoff, aoff.data(), sizeof(int64_t) * count, hipMemcpyHostToDevice));
    LAIR_HIP(hipMemcpy(d_boff, boff.data(), sizeof(int64_t) * count, hipMemcpyHostToDevice));
    LAIR_HIP(hipMemcpy(d_poff, poff.data(), sizeof(int64_t) * count, hipMemcpyHostToDevice));
    LAIR_HIP(hipMemcpy(d_n, n, sizeof(int) * count, hipMemcpyHostToDevice));
    LAIR_HIP(hipMemcpy(d_r, nrhs, sizeof(int) * count, hipMemcpyHostToDevice));
    LAIR_HIP(hipMemcpy(d_order, order.data(), sizeof(int) * count, hipMemcpyHostToDevice));
  }
  lair::BatchLaunch bl;
  bl.count = count;
  bl.n = d_n;
  bl.nrhs = d_r;
  bl.a_off = d_aoff;
  bl.b_off = d_boff;
  bl.p_off = d_poff;
  bl.order = d_order;
  bl.queue = d_queue;
  int ncu = 256;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) ncu = prop.multiProcessorCount;
  }
  batched_prepare();
  // Size classes: without the autotuner, one class with the first candidate
  // shape.  With it, every threshold T splits the sorted systems into n > T
  // and n <= T; each class gets its fastest candidate, timed on scratch copies
  // of the inputs (the solve is in place), and the split with the smallest
  // total wins (T = 256: no split).
  std::vector<lair::BatchLaunch> best_plan{batched_candidates(bl, esize, ncu).front()};
  if ((flags & LAIR_BATCHED_AUTOTUNE) && count > 0) {
    void *sa_d = nullptr, *sb_d = nullptr;
    LAIR_HIP(hipMalloc(&sa_d, esize * pl->a_elems));
    LAIR_HIP(hipMalloc(&sb_d, esize * pl->b_elems));
    hipEvent_t e0, e1;
    LAIR_HIP(hipEventCreate(&e0));
    LAIR_HIP(hipEventCreate(&e1));
    lair_status st = LAIR_OK;
    auto time_launch = [&](const lair::BatchLaunch &c, float *best_ms) -> lair_status {
      for (int rep = 0; rep < 2; ++rep) {
        LAIR_HIP(hipMemcpyAsync(sa_d, a, esize * pl->a_elems, hipMemcpyDeviceToDevice, ctx->stream));
        LAIR_HIP(hipMemcpyAsync(sb_d, b, esize * pl->b_elems, hipMemcpyDeviceToDevice, ctx->stream));
        LAIR_HIP(hipEventRecord(e0, ctx->stream));
        LAIR_HIP(batched_launch(pl.get(), c, sa_d, sb_d));
        LAIR_HIP(hipEventRecord(e1, ctx->stream));
        LAIR_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        LAIR_HIP(hipEventElapsedTime(&ms, e0, e1));
        *best_ms = std::min(*best_ms, ms);
      }
      return LAIR_OK;
    };
    // The fastest candidate for the systems order[first, first + cnt).
    auto tune_class = [&](int first, int cnt, int queue_slot, lair::BatchLaunch *out,
                          float *out_ms) -> lair_status {
      lair::BatchLaunch base = bl;
      base.count = cnt;
      base.order = d_order + first;
      base.queue = d_queue + queue_slot;
      *out_ms = 1e30f;
      for (const auto &c : batched_candidates(base, esize, ncu)) {
        float ms = 1e30f;
        lair_status s2 = time_launch(c, &ms);
        if (s2 != LAIR_OK) return s2;
        if (ms < *out_ms) {
          *out_ms = ms;
          *out = c;
        }
      }
      return LAIR_OK;
    };
    float best_total = 1e30f;
    for (int T : {256, 192, 144}) {
      int split = 0;  // systems with n > T (a prefix of the order)
      while (split < count && n[order[split]] > T) ++split;
      if (T != 256 && (split == 0 || split == count)) continue;
      std::vector<lair::BatchLaunch> cand_plan;
      float total = 0.f;
      if (split > 0) {
        lair::BatchLaunch la;
        float ms;
        st = tune_class(0, split, 0, &la, &ms);
        if (st != LAIR_OK) break;
        cand_plan.push_back(la);
        total += ms;
      }
      if (split < count) {
        lair::BatchLaunch lb;
        float ms;
        st = tune_class(split, count - split, 1, &lb, &ms);
        if (st != LAIR_OK) break;
        cand_plan.push_back(lb);
        total += ms;
      }
      if (total < best_total) {
        best_total = total;
        best_plan = cand_plan;
      }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(sa_d);
    hipFree(sb_d);
    if (st != LAIR_OK) return st;
  }
  pl->launches = best_plan;
  // Capture the (queue reset, persistent kernel) pair of every class.
  LAIR_HIP(hipStreamSynchronize(ctx->stream));
  LAIR_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  for (const auto &l : pl->launches)
    if (le == hipSuccess) le = batched_launch(pl.get(), l, a, b);
  hipGraph_t g = nullptr;
  hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
  if (le != hipSuccess) return hip_status(le, "batched capture");
  if (ce != hipSuccess) return hip_status(ce, "end capture");
  pl->graph = g;
  LAIR_HIP(hipGraphInstantiate(&pl->exec, g, nullptr, nullptr, 0));
  *plan = pl.release();
  return LAIR_OK;
}

lair_status lair_gesv_batched_run(lair_batched_plan *plan) {
  if (!plan) return -1;
  if (plan->count == 0) return LAIR_OK;
  std::lock_guard<std::mutex> lk(plan->ctx->mu);
  LAIR_HIP(hipGraphLaunch(plan->exec, plan->ctx->stream));
  return LAIR_OK;
}

lair_status lair_gesv_batched_config(lair_batched_plan *plan, int32_t *nb, int32_t *grid,
                                     int32_t *lds_elems) {
  if (!plan) return -1;
  // The launch shape of the largest systems' class.
  if (nb) *nb = plan->launches.front().nb;
  if (grid) *grid = plan->launches.front().grid;
  if (lds_elems) *lds_elems = plan->launches.front().lds_max_elems;
  return LAIR_OK;
}

lair_status lair_gesv_batched_destroy(lair_batched_plan *plan) {
  if (!plan) return -1;
  hipStreamSynchronize(plan->ctx->stream);
  if (plan->exec) hipGraphExecDestroy(plan->exec);
  if (plan->graph) hipGraphDestroy(plan->graph);
  if (plan->meta) hipFree(plan->meta);
  delete plan;
  return LAIR_OK;
}

}  // extern "C"

// ---- adjoint -------------------------------------------------------------

// BBAR = A^-T XBAR by a transposed getrs (on the handle's permutation and
// diagonal inverses when aux is given, else composed from ipiv), then
// ABAR = -BBAR X^T.
static lair_status solve_backward_locked(lair_context *ctx, int n, int nrhs, const double *lu,
                                         int64_t lda, const int64_t *ipiv, const double *x,
                                         int64_t ldx, const double *xbar, int64_t ldxb, double *bbar,
                                         int64_t ldbb, double *abar, int64_t ldab,
                                         const SolveAux<double> *aux = nullptr) {
  hipError_t e = launch_copy_d2d(ctx, n, nrhs, xbar, ldxb, bbar, ldbb);
  if (e != hipSuccess) return hip_status(e, "copy");
  lair_status st = aux ? getrs_aux<double>(ctx, n, nrhs, lu, lda, *aux, bbar, ldbb, 1)
                       : getrs_impl<double>(ctx, n, nrhs, lu, lda, ipiv, bbar, ldbb, 1);
  if (st != LAIR_OK) return st;
  // X^T (nrhs x n) so that ABAR = -BBAR * X^T is a plain row-major product.
  LAIR_HIP(ensure(ctx->ws_rhs, sizeof(double) * (size_t)n * nrhs));
  double *xt = static_cast<double *>(ctx->ws_rhs.ptr);
  e = launch_transpose_d(ctx, n, nrhs, x, ldx, xt, n);
  if (e != hipSuccess) return hip_status(e, "transpose");
  const int slot = prof_begin(ctx, LAIR_KCLASS_GEMM);
  e = launch_dgemm_tile(ctx, n, n, nrhs, bbar, ldbb, false, xt, n, abar, ldab, true);
  prof_end(ctx, slot, LAIR_KCLASS_GEMM, 2.0 * n * (double)n * nrhs,
           sizeof(double) * ((double)n * n + 2.0 * n * nrhs));
  if (e != hipSuccess) return hip_status(e, "gemm");
  return sync_status(ctx);
}

lair_status lair_dsolve_backward_device(lair_context *ctx, int64_t n, int64_t nrhs,
                                        const double *lu, int64_t lda, const int64_t *ipiv,
                                        const double *x, int64_t ldx, const double *xbar,
                                        int64_t ldxb, double *bbar, int64_t ldbb, double *abar,
                                        int64_t ldab) {
  if (!ctx) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (n == 0 || nrhs == 0) return LAIR_OK;
  if (n > INT32_MAX || nrhs > INT32_MAX) return LAIR_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(ctx->mu);
  return solve_backward_locked(ctx, (int)n, (int)nrhs, lu, lda, ipiv, x, ldx, xbar, ldxb, bbar,
                               ldbb, abar, ldab);
}

lair_status lair_dsolve_backward(int64_t n, int64_t nrhs, const double *lu, const int64_t *ipiv,
                                 const double *x, const double *xbar, double *bbar, double *abar) {
  if (n < 0) return -1;
  if (nrhs < 0) return -2;
  if (n == 0 || nrhs == 0) return LAIR_OK;
  if (!lu || !ipiv || !x || !xbar || !bbar || !abar) return -3;
  lair_context *ctx = default_ctx();
  if (!ctx) return LAIR_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(ctx->mu);
  const size_t nn = (size_t)n * n, nr = (size_t)n * nrhs;
  // One device allocation: lu, abar (n x n each), x, xbar, bbar (n x nrhs), ipiv.
  const size_t bytes = sizeof(double) * (2 * nn + 3 * nr) + sizeof(int64_t) * n;
  LAIR_HIP(ensure(ctx->ws_tmp, bytes));
  double *d_lu = static_cast<double *>(ctx->ws_tmp.ptr);
  double *d_ab = d_lu + nn, *d_x = d_ab + nn, *d_xb = d_x + nr, *d_bb = d_xb + nr;
  int64_t *d_piv = reinterpret_cast<int64_t *>(d_bb + nr);
  LAIR_HIP(hipMemcpyAsync(d_lu, lu, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream));
  LAIR_HIP(hipMemcpyAsync(d_x, x, sizeof(double) * nr, hipMemcpyHostToDevice, ctx->stream));
  LAIR_HIP(hipMemcpyAsync(d_xb, xbar, sizeof(double) * nr, hipMemcpyHostToDevice, ctx->stream));
  LAIR_HIP(hipMemcpyAsync(d_piv, ipiv, sizeof(int64_t) * n, hipMemcpyHostToDevice, ctx->stream));
  lair_status st = solve_backward_locked(ctx, (int)n, (int)nrhs, d_lu, n, d_piv, d_x, nrhs, d_xb,
                                         nrhs, d_bb, nrhs, d_ab, n);
  if (st != LAIR_OK) return st;
  LAIR_HIP(hipMemcpyAsync(bbar, d_bb, sizeof(double) * nr, hipMemcpyDeviceToHost, ctx->stream));
  LAIR_HIP(hipMemcpyAsync(abar, d_ab, sizeof(double) * nn, hipMemcpyDeviceToHost, ctx->stream));
  return sync_status(ctx);
}

// ---- device-resident factors ----------------------------------------------

// lu::Factorized (src/decomposition/lu.rs:12-20) kept on the device: the
// packed factors, the pivots, the row permutation the factorization built and
// (square systems, built at the first solve) the inverses of the diagonal
// blocks of L and U, so that every later solve and adjoint reuses them.
struct lair_lu {
  int device = 0;
  int64_t m = 0, n = 0, ld = 0;
  double *lu = nullptr;
  int64_t *ipiv = nullptr;
  int *perm = nullptr;
  double *linv = nullptr, *uinv = nullptr;
  int64_t singular = -1;
};

static void lu_free(lair_lu *h) {
  for (void *p : {(void *)h->lu, (void *)h->ipiv, (void *)h->perm, (void *)h->linv, (void *)h->uinv})
    if (p) hipFree(p);
  delete h;
}

// The handle's SolveAux; the diagonal inverses are built on first use.
static lair_status lu_aux(lair_context *ctx, lair_lu *h, SolveAux<double> *aux) {
  constexpr int BS = TrsvBlock<double>::BS;
  if (!h->linv) {
    const size_t bytes = sizeof(double) * (size_t)((h->n + BS - 1) / BS) * BS * BS;
    LAIR_HIP(hipMalloc(&h->linv, bytes));
    LAIR_HIP(hipMalloc(&h->uinv, bytes));
    hipError_t e = launch_trtri_diag<double>(ctx, h->lu, h->ld, (int)h->n, h->linv, h->uinv);
    if (e != hipSuccess) return hip_status(e, "trtri");
  }
  aux->perm = h->perm;
  aux->linv = h->linv;
  aux->uinv = h->uinv;
  return LAIR_OK;
}

extern "C" {

lair_status lair_dlu_factor(lair_context *ctx, int64_t m, int64_t n, const double *a, int64_t lda, int64_t nb,
                            lair_lu **out) {
  if (!ctx) return -1;
  if (m < 1) return -2;
  if (n < 1) return -3;
  if (!a) return -4;
  if (lda < n) return -5;
  if (!out) return -7;
  if (m > INT32_MAX || n > INT32_MAX) return LAIR_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(ctx->mu);
  std::unique_ptr<lair_lu, void (*)(lair_lu *)> h(new lair_lu(), lu_free);
  h->device = ctx->device;
  h->m = m;
  h->n = n;
  h->ld = pad_ld<double>(n);
  LAIR_HIP(hipMalloc(&h->lu, sizeof(double) * (size_t)m * h->ld));
  LAIR_HIP(hipMalloc(&h->ipiv, sizeof(int64_t) * (size_t)std::min(m, n)));
  LAIR_HIP(hipMalloc(&h->perm, sizeof(int) * (size_t)m));
  hipError_t e = launch_copy_d2d(ctx, (int)m, (int)n, a, lda, h->lu, h->ld);
  if (e != hipSuccess) return hip_status(e, "copy");
  if (nb < 0) {
    lair_status st = tune_nb<double>(ctx, m, n, h->lu, h->ld, kGemmF64, &nb);
    if (st != LAIR_OK) return st;
  }
  lair_status st = getrf_impl<double>(ctx, m, n, h->lu, h->ld, h->ipiv, nb, kGemmF64, &h->singular, h->perm);
  if (st != LAIR_OK) return st;
  *out = h.release();
  return LAIR_OK;
}

lair_status lair_dlu_destroy(lair_lu *h) {
  if (!h) return -1;
  hipDeviceSynchronize();
  lu_free(h);
  return LAIR_OK;
}

lair_status lair_dlu_info(const lair_lu *h, int64_t *m, int64_t *n, int64_t *singular) {
  if (!h) return -1;
  if (m) *m = h->m;
  if (n) *n = h->n;
  if (singular) *singular = h->singular;
  return LAIR_OK;
}

lair_status lair_dlu_factors(const lair_lu *h, const double **lu, int64_t *ld, const int64_t **ipiv,
                             const int32_t **perm) {
  if (!h) return -1;
  if (lu) *lu = h->lu;
  if (ld) *ld = h->ld;
  if (ipiv) *ipiv = h->ipiv;
  if (perm) *perm = h->perm;
  return LAIR_OK;
}

lair_status lair_dlu_solve(lair_context *ctx, lair_lu *h, int64_t nrhs, double *b, int64_t ldb,
                           int32_t trans) {
  if (!ctx) return -1;
  if (!h) return -2;
  if (nrhs < 0) return -3;
  if (nrhs > 0 && (!b || ldb < nrhs)) return -4;
  if (h->m != h->n) {
    set_error("solve needs a square factorization");
    return LAIR_ERR_SHAPE;
  }
  if (nrhs == 0) return LAIR_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  const int n = (int)h->n;
  lair_status st;
  if (!trans && n <= kExactOrderMax) {
    st = hip_status(launch_getrs_exact<double>(ctx, n, (int)nrhs, h->lu, h->ld, h->ipiv, b, ldb), "getrs");
  } else {
    SolveAux<double> aux;
    st = lu_aux(ctx, h, &aux);
    if (st == LAIR_OK) st = getrs_aux<double>(ctx, n, (int)nrhs, h->lu, h->ld, aux, b, ldb, trans);
  }
  return st != LAIR_OK ? st : sync_status(ctx);
}

lair_status lair_dlu_solve_backward(lair_context *ctx, lair_lu *h, int64_t nrhs, const double *x, int64_t ldx,
                                    const double *xbar, int64_t ldxb, double *bbar, int64_t ldbb,
                                    double *abar, int64_t ldab) {
  if (!ctx) return -1;
  if (!h) return -2;
  if (nrhs < 0) return -3;
  if (h->m != h->n) {
    set_error("solve_backward needs a square factorization");
    return LAIR_ERR_SHAPE;
  }
  if (nrhs == 0) return LAIR_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  SolveAux<double> aux;
  lair_status st = lu_aux(ctx, h, &aux);
  if (st != LAIR_OK) return st;
  return solve_backward_locked(ctx, (int)h->n, (int)nrhs, h->lu, h->ld, h->ipiv, x, ldx, xbar, ldxb, bbar,
                               ldbb, abar, ldab, &aux);
}

lair_status lair_dlu_extract(lair_context *ctx, const lair_lu *h, int32_t which, double *out, int64_t ldo) {
  if (!ctx) return -1;
  if (!h) return -2;
  if (which < LAIR_LU_P || which > LAIR_LU_PL) return -3;
  if (!out) return -4;
  const int64_t k = std::min(h->m, h->n);
  const int64_t cols = which == LAIR_LU_P ? h->m : (which == LAIR_LU_U ? h->n : k);
  if (ldo < cols) return -5;
  std::lock_guard<std::mutex> lk(ctx->mu);
  hipError_t e = launch_lu_extract(ctx, which, (int)h->m, (int)h->n, h->lu, h->ld, h->perm, out, ldo);
  if (e != hipSuccess) return hip_status(e, "extract");
  return sync_status(ctx);
}

}  // extern "C"

// ---- instrumentation -----------------------------------------------------

lair_status lair_profile_enable(lair_context *ctx, int32_t class_mask) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->prof_mask = class_mask;
  return LAIR_OK;
}

lair_status lair_profile_read(lair_context *ctx, int32_t kclass, double *ms, int64_t *launches,
                              double *flops, double *bytes) {
  if (!ctx) return -1;
  if (kclass < 0 || kclass >= LAIR_KCLASS_COUNT) return -2;
  std::lock_guard<std::mutex> lk(ctx->mu);
  prof_collect(ctx);
  if (ms) *ms = ctx->prof_ms[kclass];
  if (launches) *launches = ctx->prof_n[kclass];
  if (flops) *flops = ctx->prof_flops[kclass];
  if (bytes) *bytes = ctx->prof_bytes[kclass];
  return LAIR_OK;
}

lair_status lair_profile_reset(lair_context *ctx) {
  if (!ctx) return -1;
  std::lock_guard<std::mutex> lk(ctx->mu);
  prof_collect(ctx);
  for (int i = 0; i < LAIR_KCLASS_COUNT; ++i) {
    ctx->prof_ms[i] = 0;
    ctx->prof_n[i] = 0;
    ctx->prof_flops[i] = 0;
    ctx->prof_bytes[i] = 0;
  }
  return LAIR_OK;
}

}  // extern "C"
