// Resident-panel LU leaf with exact partial pivoting: the column loop of
// getrf_row_major (src/lapack/getrf.rs:55-117) with iamax's choice rule
// (src/blas/iamax.rs:6-21), for a tall leaf whose rows are spread over P
// cooperating workgroups.
//
//  * Workgroup g owns R consecutive rows of the leaf.  They are loaded into
//    LDS once and written back once, each to its final position.  Rows do not
//    move while the leaf is factored: every row carries its current position
//    (its row index after the interchanges so far, as the reference would
//    see it), which is also iamax's tie-break key (lowest index wins).
//  * One column per iteration of a runtime loop, so the code stays small
//    enough for the instruction cache.
//  * Per column, each workgroup publishes a header (|value|, position) for
//    its best candidate and that candidate's row as tagged 64-bit words;
//    every workgroup polls the P headers, picks the winner by the iamax rule
//    and reads the winner's row.  A word is valid when its tag equals the
//    column's tag, so no fences are needed.
//  * The rank-1 update of pivot j (getrf.rs:76-88: l = a * (1/pivot), then
//    a -= l * u with separate multiply and subtract) is split: column j + 1
//    is updated at once because it decides the next pivot; the other columns
//    are updated at the start of the next iteration while the exchange is in
//    flight.  A published candidate row therefore still lacks the previous
//    pivot's update beyond its pivot column.  The readers apply it with the
//    same operands in the same order, so every copy of the pivot row is
//    bit-identical to the owner's.
#include <climits>

#include "device_util.h"
#include "lair_internal.h"

namespace lair {

constexpr int kRpNT = 256;              // threads per workgroup
constexpr int kRpNW = kRpNT / 64;       // waves per workgroup
constexpr uint32_t kRpNone = 0x1FFFFu;  // header position meaning "no candidate"

template <typename T>
struct RpArgs {
  T *a;
  int64_t lda;
  int row0, m, col0, wcols, npiv, P;
  int64_t *ipiv;
  int64_t *sing;
  unsigned long long *hdr;   // [2][kMaxPanelWG][2] header words, by column parity
  unsigned long long *rowd;  // [2][kMaxPanelWG][kRpRowWords] row words, by column parity
  uint32_t seq;              // column j's tag is seq + j + 1
  int32_t *err;
  int wait_limit;            // polls before a wait gives up (0: every wait gives up at once)
};

// Header word `word` of a candidate (v = |value| >= 0, pos = its position or
// INT_MAX for none).  Word 0: tag (16 bits) | position (17 bits) | the high
// 31 bits of v (the sign bit of a non-negative value is zero).  FP64 adds
// word 1: tag (32 bits) | the low 32 bits of v.
template <typename T>
__device__ __forceinline__ unsigned long long rp_header(int word, uint32_t tag, double v, int pos) {
  uint32_t hi, lo = 0;
  if constexpr (sizeof(T) == 8) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    hi = (uint32_t)(b >> 32);
    lo = (uint32_t)b;
  } else {
    hi = __float_as_uint((float)v);
  }
  if (word == 1) return tagged(tag, lo);
  const uint32_t p = pos == INT_MAX ? kRpNone : (uint32_t)pos;
  return ((unsigned long long)(tag & 0xFFFFu) << 48) | ((unsigned long long)p << 31) |
         (hi & 0x7FFFFFFFu);
}

template <typename T, int NB, int R>
__global__ void __launch_bounds__(kRpNT) rpanel_kernel(RpArgs<T> pa) {
#pragma clang fp contract(off)
  constexpr int RPT = R / kRpNT;
  constexpr int LD = NB + 1;
  constexpr int WPE = sizeof(T) / 4;
  constexpr int NI = (NB + 63) / 64;
  extern __shared__ unsigned char smem_raw[];
  T *s_a = reinterpret_cast<T *>(smem_raw);  // [R][LD] the workgroup's rows
  __shared__ T s_wrow[2][kRpNW][NB];         // wave winners' rows (columns j-1..)
  __shared__ T s_u[2][NB];                   // pivot rows
  __shared__ double s_wv[2][kRpNW];
  __shared__ int s_wp[2][kRpNW];
  __shared__ int s_at[NB];  // original row now at position row0 + e
  __shared__ int s_pw[2], s_occ[2], s_none[2];
  __shared__ int s_fin[R];  // final positions for the write-back
  __shared__ int s_bail;

  const int g = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int base = pa.row0 + g * R;
  const int ncols = pa.wcols;

  // Load: wave w copies rows w, w + 4, ... (coalesced along the row).
  for (int i = wave; i < R; i += kRpNW) {
    const int r = base + i;
    const bool ok = r < pa.m;
    const T *src = pa.a + (int64_t)(ok ? r : 0) * pa.lda + pa.col0;
#pragma unroll
    for (int q = 0; q < NI; ++q) {
      const int c = lane + 64 * q;
      if (c < NB) s_a[i * LD + c] = (ok && c < ncols) ? src[c] : T(0);
    }
  }
  int cpos[RPT];
  bool act[RPT];
  T lr[RPT];
#pragma unroll
  for (int s = 0; s < RPT; ++s) {
    cpos[s] = base + t + s * kRpNT;
    act[s] = cpos[s] < pa.m;
    lr[s] = T(0);
  }
  for (int e = t; e < NB; e += kRpNT) s_at[e] = pa.row0 + e;
  if (t == 0) s_bail = 0;
  __syncthreads();

  bool pend = false;  // the previous pivot's update of columns > j is pending
  for (int j = 0; j < pa.npiv; ++j) {
    const int par = j & 1;
    const int d = pa.row0 + j;
    const uint32_t tg = pa.seq + (uint32_t)j + 1u;

    // 1. Candidates of column j: |a| > 0, largest first, lowest position on ties.
    double bv = 0.0;
    int bp = INT_MAX, bs = -1;
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      if (act[s]) {
        const double x = fabs((double)s_a[(t + s * kRpNT) * LD + j]);
        if (x > bv || (x == bv && x > 0.0 && cpos[s] < bp)) {
          bv = x;
          bp = cpos[s];
          bs = s;
        }
      }
    }
    double wv;
    int wp;
    wave_argmax(bv, bp, wv, wp);
    if (wp != INT_MAX) {
      // The wave copies its winner's row from column j - 1 (the multiplier of
      // the pending update) before the pending update changes it.
      const unsigned long long hit = __ballot(bs >= 0 && bp == wp);
      const int wl = __ffsll((long long)hit) - 1;
      const int wsel = __builtin_amdgcn_readlane(bs, wl);
      const int li = wave * 64 + wl + wsel * kRpNT;
      for (int c = (j > 0 ? j - 1 : 0) + lane; c < NB; c += 64) s_wrow[par][wave][c] = s_a[li * LD + c];
    }
    if (lane == 0) {
      s_wv[par][wave] = wv;
      s_wp[par][wave] = wp;
    }
    __syncthreads();

    // 2. Wave 0 publishes the workgroup's candidate: header, then the row.
    if (wave == 0) {
      double cv = s_wv[par][0];
      int cp = s_wp[par][0], cw = 0;
#pragma unroll
      for (int q = 1; q < kRpNW; ++q) {
        const double ov = s_wv[par][q];
        const int op = s_wp[par][q];
        if (ov > cv || (ov == cv && op < cp)) {
          cv = ov;
          cp = op;
          cw = q;
        }
      }
      unsigned long long *hs = pa.hdr + ((size_t)par * kMaxPanelWG + g) * 2;
      if (lane < WPE) st_agent(hs + lane, rp_header<T>(lane, tg, cv, cp));
      if (cp != INT_MAX) {
        unsigned long long *rs = pa.rowd + ((size_t)par * kMaxPanelWG + g) * kRpRowWords;
        for (int idx = (j > 0 ? j - 1 : 0) * WPE + lane; idx < NB * WPE; idx += 64)
          st_agent(rs + idx, tagged(tg, word_of<T>(s_wrow[par][cw][idx / WPE], idx % WPE)));
      }
    }

    // 3. The pending update of columns > j with the previous pivot row.
    if (pend) {
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        if (act[s]) {
          T *row = s_a + (t + s * kRpNT) * LD;
          const T l = lr[s];
          for (int c = j + 1; c < ncols; ++c) {
            const T e = l * s_u[par ^ 1][c];
            row[c] = row[c] - e;
          }
        }
      }
    }

    // 4. Wave 0: all headers, the winner, the winner's row.
    if (wave == 0) {
      bool bail = false;
      double hv = 0.0;
      int hp = INT_MAX;
      {
        const unsigned long long *hb = pa.hdr + (size_t)par * kMaxPanelWG * 2;
        int spins = 0;
        for (;;) {
          if (spins >= pa.wait_limit) {
            bail = true;
            break;
          }
          bool ok = true;
          if (lane < pa.P) {
            const unsigned long long w0 = ld_agent(hb + lane * 2);
            uint32_t lo = 0;
            ok = (uint32_t)(w0 >> 48) == (tg & 0xFFFFu);
            if constexpr (WPE == 2) {
              const unsigned long long w1 = ld_agent(hb + lane * 2 + 1);
              ok = ok && (uint32_t)(w1 >> 32) == tg;
              lo = (uint32_t)w1;
            }
            const uint32_t p = (uint32_t)(w0 >> 31) & 0x1FFFFu;
            const uint32_t hi = (uint32_t)w0 & 0x7FFFFFFFu;
            if constexpr (WPE == 2)
              hv = __hiloint2double((int)hi, (int)lo);
            else
              hv = (double)__uint_as_float(hi);
            hp = p == kRpNone ? INT_MAX : (int)p;
            if (hp == INT_MAX) hv = 0.0;
          }
          if (__all(ok)) break;
          if ((++spins & 63) == 0 && ld_agent(pa.err) != 0) {
            bail = true;
            break;
          }
        }
      }
      double gv;
      int gp;
      wave_argmax(hv, hp, gv, gp);
      const bool none = gp == INT_MAX;
      const int pw = none ? d : gp;
      if (!bail && !none) {
        const int q = __ffsll((long long)__ballot(lane < pa.P && hp == gp)) - 1;
        const unsigned long long *rs = pa.rowd + ((size_t)par * kMaxPanelWG + q) * kRpRowWords;
        T rv[NI];
        T mult = T(0);
        int spins = 0;
        for (;;) {
          if (spins >= pa.wait_limit) {
            bail = true;
            break;
          }
          bool ok = true;
#pragma unroll
          for (int i = 0; i < NI; ++i) {
            const int c = j + lane + 64 * i;
            rv[i] = T(0);
            if (c < NB) {
              uint32_t w[WPE];
#pragma unroll
              for (int p = 0; p < WPE; ++p) {
                const unsigned long long x = ld_agent(rs + c * WPE + p);
                ok = ok && (uint32_t)(x >> 32) == tg;
                w[p] = (uint32_t)x;
              }
              rv[i] = from_words<T>(w);
            }
          }
          if (pend) {
            uint32_t w[WPE];
#pragma unroll
            for (int p = 0; p < WPE; ++p) {
              const unsigned long long x = ld_agent(rs + (j - 1) * WPE + p);
              ok = ok && (uint32_t)(x >> 32) == tg;
              w[p] = (uint32_t)x;
            }
            mult = from_words<T>(w);
          }
          if (__all(ok)) break;
          if ((++spins & 63) == 0 && ld_agent(pa.err) != 0) {
            bail = true;
            break;
          }
        }
        // The winner's row as its owner holds it after the pending update.
#pragma unroll
        for (int i = 0; i < NI; ++i) {
          const int c = j + lane + 64 * i;
          if (c < NB) {
            T u = rv[i];
            if (pend && c > j) {
              const T e = mult * s_u[par ^ 1][c];
              u = u - e;
            }
            s_u[par][c] = u;
          }
        }
      }
      if (lane == 0) {
        const int occ = s_at[j];
        s_pw[par] = pw;
        s_occ[par] = occ;
        s_none[par] = none ? 1 : 0;
        if (pw != d && pw - pa.row0 < NB) s_at[pw - pa.row0] = occ;
        if (g == 0 && !bail) {
          pa.ipiv[d] = pw;
          if (none) *pa.sing = d;
        }
        if (bail) {
          s_bail = 1;
          atomic_or_agent(pa.err, kErrWait);
        }
      }
    }
    __syncthreads();
    if (s_bail) {
      // Given up: identity pivots for the columns not decided keep the
      // interchanges queued after this launch in bounds (thread 0 of
      // workgroup 0 writes every pivot of the leaf).
      if (g == 0 && t == 0)
        for (int jj = j; jj < pa.npiv; ++jj) pa.ipiv[pa.row0 + jj] = pa.row0 + jj;
      return;
    }

    // 5. Interchange bookkeeping: the pivot row takes position d, the row at
    //    position d takes the pivot's old position.
    const int pw = s_pw[par], occ = s_occ[par];
    const bool none = s_none[par] != 0;
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      if (act[s] && cpos[s] == pw) {
        act[s] = false;
        cpos[s] = d;
      } else if (base + t + s * kRpNT == occ) {
        cpos[s] = pw;
      }
    }

    // 6. Multipliers of column j and the update of column j + 1.  A column
    //    without a nonzero candidate is left unscaled (getrf.rs:72-73).
    pend = !none;
    if (!none) {
      const T recip = T(1) / s_u[par][j];
      const T u1 = j + 1 < NB ? s_u[par][j + 1] : T(0);
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        if (act[s]) {
          T *row = s_a + (t + s * kRpNT) * LD;
          const T l = row[j] * recip;
          row[j] = l;
          lr[s] = l;
          if (j + 1 < ncols) {
            const T e = l * u1;
            row[j + 1] = row[j + 1] - e;
          }
        }
      }
    }
  }

  // The last pivot's update of the columns beyond the pivots (wide leaves).
  if (pend) {
    const int jl = pa.npiv - 1;
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      if (act[s]) {
        T *row = s_a + (t + s * kRpNT) * LD;
        for (int c = jl + 2; c < ncols; ++c) {
          const T e = lr[s] * s_u[jl & 1][c];
          row[c] = row[c] - e;
        }
      }
    }
  }

  // Write-back: every row to its final position.
#pragma unroll
  for (int s = 0; s < RPT; ++s) s_fin[t + s * kRpNT] = base + t + s * kRpNT < pa.m ? cpos[s] : -1;
  __syncthreads();
  for (int i = wave; i < R; i += kRpNW) {
    const int dst = s_fin[i];
    if (dst < 0) continue;
    T *out = pa.a + (int64_t)dst * pa.lda + pa.col0;
#pragma unroll
    for (int q = 0; q < NI; ++q) {
      const int c = lane + 64 * q;
      if (c < ncols) out[c] = s_a[i * LD + c];
    }
  }
}

template <typename T, int NB, int R>
static hipError_t rp_launch(lair_context *ctx, const RpArgs<T> &pa) {
  auto *k = &rpanel_kernel<T, NB, R>;
  const size_t shm = (size_t)R * (NB + 1) * sizeof(T);
  static bool once = ((void)hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                                (int)shm),
                      (void)hipGetLastError(), true);
  (void)once;
  k<<<pa.P, kRpNT, shm, ctx->stream>>>(pa);
  return hipGetLastError();
}

// Rows per workgroup for a leaf of width nbw and `rows` rows: the smallest R
// with an instance whose grid fits ctx->rp_max_wg; 0 if none.
template <typename T>
int rpanel_rows_per_wg(lair_context *ctx, int nbw, int rows) {
  const int maxp = std::min(ctx->rp_max_wg, kMaxPanelWG);
  static const int kF[][2] = {{128, 256}, {64, 256}, {64, 512}, {32, 512}, {32, 1024}};
  static const int kD[][2] = {{64, 256}, {32, 256}, {32, 512}, {16, 512}, {16, 1024}};
  const auto &tab = sizeof(T) == 4 ? kF : kD;
  for (const auto &e : tab)
    if (e[0] == nbw && (rows + e[1] - 1) / e[1] <= maxp) return e[1];
  return 0;
}

// Widest leaf (<= want) that the resident kernel can factor for `rows` rows.
template <typename T>
int rpanel_leaf_width(lair_context *ctx, int want, int rows) {
  for (int w : {128, 64, 32, 16})
    if (w <= want && rpanel_rows_per_wg<T>(ctx, w, rows) > 0) return w;
  return 0;
}

template <typename T>
hipError_t launch_rpanel(lair_context *ctx, T *a, int64_t lda, int row0, int m, int col0, int wcols,
                         int nbw, int64_t *ipiv) {
  const int rows = m - row0;
  if (rows <= 0 || wcols <= 0) return hipSuccess;
  if (wcols > nbw || m >= (int)kRpNone) return hipErrorInvalidValue;
  const int R = rpanel_rows_per_wg<T>(ctx, nbw, rows);
  if (R == 0) return hipErrorInvalidValue;
  // Header tags carry 16 bits: start a new tag epoch (cleared slots) before
  // they could wrap inside this launch.
  if ((ctx->rp_seq & 0xFFFFu) + (uint32_t)nbw + 2u > 0xFFFFu || ctx->rp_seq > 0xF0000000u) {
    hipError_t e = launch_fill_u32(ctx, reinterpret_cast<uint32_t *>(ctx->rp_hdr),
                                   2 * 2 * kMaxPanelWG * 2, 0u);
    if (e == hipSuccess)
      e = launch_fill_u32(ctx, reinterpret_cast<uint32_t *>(ctx->rp_rowd),
                          2 * 2 * kMaxPanelWG * kRpRowWords, 0u);
    if (e != hipSuccess) return e;
    ctx->rp_seq = ctx->rp_seq > 0xF0000000u ? 1u : ((ctx->rp_seq | 0xFFFFu) + 2u);
  }
  RpArgs<T> pa;
  pa.a = a;
  pa.lda = lda;
  pa.row0 = row0;
  pa.m = m;
  pa.col0 = col0;
  pa.wcols = wcols;
  pa.npiv = wcols < rows ? wcols : rows;
  pa.P = (rows + R - 1) / R;
  pa.ipiv = ipiv;
  pa.sing = ctx->sing;
  pa.hdr = ctx->rp_hdr;
  pa.rowd = ctx->rp_rowd;
  pa.seq = ctx->rp_seq;
  pa.err = ctx->err;
  pa.wait_limit = ctx->wait_limit;
  const int slot = prof_begin(ctx, LAIR_KCLASS_PANEL);
  hipError_t e;
  if constexpr (sizeof(T) == 4) {
    if (nbw == 128) e = rp_launch<T, 128, 256>(ctx, pa);
    else if (nbw == 64 && R == 256) e = rp_launch<T, 64, 256>(ctx, pa);
    else if (nbw == 64) e = rp_launch<T, 64, 512>(ctx, pa);
    else if (R == 512) e = rp_launch<T, 32, 512>(ctx, pa);
    else e = rp_launch<T, 32, 1024>(ctx, pa);
  } else {
    if (nbw == 64) e = rp_launch<T, 64, 256>(ctx, pa);
    else if (nbw == 32 && R == 256) e = rp_launch<T, 32, 256>(ctx, pa);
    else if (nbw == 32) e = rp_launch<T, 32, 512>(ctx, pa);
    else if (R == 512) e = rp_launch<T, 16, 512>(ctx, pa);
    else e = rp_launch<T, 16, 1024>(ctx, pa);
  }
  prof_end(ctx, slot, LAIR_KCLASS_PANEL, (double)pa.npiv * rows * wcols,
           2.0 * sizeof(T) * (double)rows * wcols);
  ctx->rp_seq += (uint32_t)nbw + 1u;
  return e;
}

template int rpanel_leaf_width<float>(lair_context *, int, int);
template int rpanel_leaf_width<double>(lair_context *, int, int);
template hipError_t launch_rpanel<float>(lair_context *, float *, int64_t, int, int, int, int, int,
                                         int64_t *);
template hipError_t launch_rpanel<double>(lair_context *, double *, int64_t, int, int, int, int, int,
                                          int64_t *);

}  // namespace lair
