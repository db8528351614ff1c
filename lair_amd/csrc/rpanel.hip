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

constexpr uint32_t kRpNone = 0x1FFFFu;  // header position meaning "no candidate"

template <typename T>
struct RpArgs {
  T *a;
  int64_t lda;
  int row0, m, col0, wcols, npiv, P;
  int64_t *ipiv;
  int64_t *sing;
  unsigned long long *hdr;   // [2][kMaxPanelWG][RpHeader<T>::S] header words, by column parity
  unsigned long long *rowd;  // [2][kMaxPanelWG][kRpRowWords] row words, by column parity
  int *map_rows, *map_src, *map_count;  // row map of the leaf's interchanges (optional)
  uint32_t seq;              // column j's tag is seq + j + 1
  int32_t *err;
  int wait_limit;            // polls before a wait gives up (0: every wait gives up at once)
};

// A candidate's header: its position, its signed pivot value v and u1, its
// entry in the next column after every earlier pivot's update (what the
// other workgroups need to continue before the rest of the row arrives).
//   word 0: tag (16 bits) | position (17 bits) | high 31 bits of |v|
//   FP32 word 1: tag (31 bits) | sign of v | bits of u1
//   FP64 word 1: tag (32 bits) | low 32 bits of |v|
//   FP64 word 2: tag (31 bits) | sign of v | high 32 bits of u1
//   FP64 word 3: tag (32 bits) | low 32 bits of u1
// S is the stride between two workgroups' headers, in words.
template <typename T>
struct RpHeader {
  static constexpr int W = sizeof(T) == 8 ? 4 : 2;
  static constexpr int S = kRpHdrWords;
  static __device__ __forceinline__ unsigned long long word(int w, uint32_t tag, T v, T u1, int pos) {
    const uint32_t p = pos == INT_MAX ? kRpNone : (uint32_t)pos;
    const uint32_t sgn = v < T(0) ? 1u : 0u;
    const T av = v < T(0) ? -v : v;
    if constexpr (sizeof(T) == 8) {
      const unsigned long long b = (unsigned long long)__double_as_longlong(av);
      const unsigned long long ub = (unsigned long long)__double_as_longlong(u1);
      if (w == 0)
        return ((unsigned long long)(tag & 0xFFFFu) << 48) | ((unsigned long long)p << 31) |
               ((uint32_t)(b >> 32) & 0x7FFFFFFFu);
      if (w == 1) return tagged(tag, (uint32_t)b);
      if (w == 2) return ((unsigned long long)(tag & 0x7FFFFFFFu) << 33) | ((unsigned long long)sgn << 32) |
                         (uint32_t)(ub >> 32);
      return tagged(tag, (uint32_t)ub);
    } else {
      if (w == 0)
        return ((unsigned long long)(tag & 0xFFFFu) << 48) | ((unsigned long long)p << 31) |
               (__float_as_uint(av) & 0x7FFFFFFFu);
      return ((unsigned long long)(tag & 0x7FFFFFFFu) << 33) | ((unsigned long long)sgn << 32) |
             __float_as_uint(u1);
    }
  }
  // Tag check of word w.
  static __device__ __forceinline__ bool ok(int w, unsigned long long x, uint32_t tag) {
    if (w == 0) return (uint32_t)(x >> 48) == (tag & 0xFFFFu);
    if (sizeof(T) == 4 || w == 2) return (uint32_t)(x >> 33) == (tag & 0x7FFFFFFFu);
    return (uint32_t)(x >> 32) == tag;
  }
  // Position, |v| (the iamax key), v and u1 of a header whose words all
  // carry the tag.
  static __device__ __forceinline__ uint32_t decode(const unsigned long long *w, double &hv, T &hvs, T &hu1) {
    const uint32_t p = (uint32_t)(w[0] >> 31) & 0x1FFFFu;
    const uint32_t hi = (uint32_t)w[0] & 0x7FFFFFFFu;
    uint32_t sgn;
    if constexpr (sizeof(T) == 8) {
      hv = __hiloint2double((int)hi, (int)(uint32_t)w[1]);
      sgn = (uint32_t)(w[2] >> 32) & 1u;
      hu1 = __hiloint2double((int)(uint32_t)w[2], (int)(uint32_t)w[3]);
    } else {
      hv = (double)__uint_as_float(hi);
      sgn = (uint32_t)(w[1] >> 32) & 1u;
      hu1 = __uint_as_float((uint32_t)w[1]);
    }
    hvs = sgn ? -(T)hv : (T)hv;
    return p;
  }
};

// iamax's key (src/blas/iamax.rs:14: |re| + |im| in the real type) and the
// pivot reciprocal (getrf.rs:76) of a real or complex element.
template <typename T>
__device__ __forceinline__ double rp_key(T x) {
  if constexpr (cplx_traits<T>::is) {
    const typename cplx_traits<T>::real k = fabs(x.re) + fabs(x.im);
    return (double)k;
  } else {
    return fabs((double)x);
  }
}
template <typename T>
__device__ __forceinline__ T rp_recip(T v) {
  if constexpr (cplx_traits<T>::is)
    return cplx_recip(v);
  else
    return T(1) / v;
}

// Complex header: the key is recomputed from v by every reader, so the
// header carries the position and the full values of v and u1, each 32-bit
// part under a 32-bit tag.
//   word 0: tag | position; words 1..H: v; words H+1..2H: u1
template <typename R>
struct RpHeader<cplx<R>> {
  typedef cplx<R> T;
  static constexpr int H = sizeof(T) / 4;
  static constexpr int W = 1 + 2 * H;
  static constexpr int S = kRpHdrWordsMax;
  static __device__ __forceinline__ unsigned long long word(int w, uint32_t tag, T v, T u1, int pos) {
    const uint32_t p = pos == INT_MAX ? kRpNone : (uint32_t)pos;
    if (w == 0) return tagged(tag, p);
    return tagged(tag, w <= H ? word_of<T>(v, w - 1) : word_of<T>(u1, w - 1 - H));
  }
  static __device__ __forceinline__ bool ok(int, unsigned long long x, uint32_t tag) {
    return (uint32_t)(x >> 32) == tag;
  }
  static __device__ __forceinline__ uint32_t decode(const unsigned long long *w, double &hv, T &hvs, T &hu1) {
    uint32_t a[H], b[H];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      a[k] = (uint32_t)w[1 + k];
      b[k] = (uint32_t)w[1 + H + k];
    }
    hvs = from_words<T>(a);
    hu1 = from_words<T>(b);
    hv = rp_key(hvs);
    return (uint32_t)w[0];
  }
};

// Waves: wave 0 communicates (publishes, polls, rebuilds pivot rows) and
// owns no rows; waves 1..NWK own RPT rows per lane, R = 64 * NWK * RPT rows.
template <typename T, int NB, int NWK, int RPT>
__global__ void __launch_bounds__(64 * (NWK + 1)) rpanel_kernel(RpArgs<T> pa) {
#pragma clang fp contract(off)
  constexpr int NT = 64 * (NWK + 1);
  constexpr int NW = 64 * NWK;  // worker threads
  constexpr int R = NW * RPT;
  // Column-major tile with a padded column: a worker's sweep along its row
  // and a wave's copy of one row (lanes along the columns) are both free of
  // bank conflicts.
  constexpr int LDC = R + 1;
  constexpr int WPE = sizeof(T) / 4;
  constexpr int NI = (NB + 63) / 64;
  constexpr int HW = RpHeader<T>::W;
  static_assert(NB * WPE <= kRpRowWords && HW <= RpHeader<T>::S, "exchange record sizes");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *s_a = reinterpret_cast<T *>(smem_raw);  // [NB][LDC] the workgroup's rows, by column
  __shared__ T s_wrow[2][NWK][NB];           // worker waves' best rows (columns j-1..)
  __shared__ T s_u[2][NB];                   // pivot rows
  __shared__ double s_wv[2][NWK];
  __shared__ int s_wp[2][NWK];
  __shared__ T s_recip[2], s_u1[2];
  __shared__ int s_at[NB];  // original row now at position row0 + e
  __shared__ int s_pw[2], s_occ[2], s_none[2], s_q[2];
  __shared__ int s_fin[R];  // final positions for the write-back
  __shared__ int s_bail;

  const int g = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int tw = t - 64;  // worker index (rows tw + s * NW)
  const int base = pa.row0 + g * R;
  const int ncols = pa.wcols;

  // Load: wave w copies rows w, w + NWK + 1, ... (coalesced along the row).
  for (int i = wave; i < R; i += NWK + 1) {
    const int r = base + i;
    const bool okr = r < pa.m;
    const T *src = pa.a + (int64_t)(okr ? r : 0) * pa.lda + pa.col0;
#pragma unroll
    for (int q = 0; q < NI; ++q) {
      const int c = lane + 64 * q;
      if (c < NB) s_a[c * LDC + i] = (okr && c < ncols) ? src[c] : T(0);
    }
  }
  int cpos[RPT];
  bool act[RPT];
  T lr[RPT];
  // Entry j of each row (all updates through pivot j - 1) and entry j + 1
  // after the pending update, kept in registers between the phases that
  // produce and use them.
  T cj[RPT], c1[RPT];
#pragma unroll
  for (int s = 0; s < RPT; ++s) {
    cpos[s] = base + tw + s * NW;
    act[s] = wave > 0 && cpos[s] < pa.m;
    lr[s] = T(0);
  }
  for (int e = t; e < NB; e += NT) s_at[e] = pa.row0 + e;
  if (t == 0) s_bail = 0;
  __syncthreads();
#pragma unroll
  for (int s = 0; s < RPT; ++s) {
    cj[s] = wave > 0 ? s_a[tw + s * NW] : T(0);
    c1[s] = T(0);
  }

  bool pend = false;  // the previous pivot's update of columns > j is pending
  // Wave 0: the winning row's entries j + 1 .. NB - 1 and its multiplier
  // entry j - 1 (tagged words; columns past the row read as valid zeros).
  unsigned long long rw[NI][WPE], mw[WPE];
  auto load_row = [&](int q, int jj, uint32_t tag, bool pnd) {
    const unsigned long long *rs = pa.rowd + ((size_t)(jj & 1) * kMaxPanelWG + q) * kRpRowWords;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int c = jj + 1 + lane + 64 * i;
#pragma unroll
      for (int p = 0; p < WPE; ++p) rw[i][p] = c < NB ? ld_agent(rs + c * WPE + p) : tagged(tag, 0);
    }
#pragma unroll
    for (int p = 0; p < WPE; ++p) mw[p] = pnd ? ld_agent(rs + (jj - 1) * WPE + p) : tagged(tag, 0);
  };
  int j = 0;
  for (; j < pa.npiv; ++j) {
    const int par = j & 1;
    const int d = pa.row0 + j;
    const uint32_t tg = pa.seq + (uint32_t)j + 1u;

    if (wave > 0) {
      // W1. Candidates of column j: |a| > 0, largest first, lowest position
      //     on ties; the wave's best row is copied (from column j - 1, the
      //     multiplier of the pending update) before that update changes it.
      double bv = 0.0;
      int bp = INT_MAX, bs = -1;
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        if (act[s]) {
          const double x = rp_key(cj[s]);
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
        const unsigned long long hit = __ballot(bs >= 0 && bp == wp);
        const int wl = __ffsll((long long)hit) - 1;
        const int wsel = __builtin_amdgcn_readlane(bs, wl);
        const int li = (wave - 1) * 64 + wl + wsel * NW;
        for (int c = (j > 0 ? j - 1 : 0) + lane; c < NB; c += 64) s_wrow[par][wave - 1][c] = s_a[c * LDC + li];
      }
      if (lane == 0) {
        s_wv[par][wave - 1] = wv;
        s_wp[par][wave - 1] = wp;
      }
    }
    __syncthreads();  // B1

    if (wave == 0) {
      // C2. The workgroup's candidate, its header and its row.
      double cv = s_wv[par][0];
      int cp = s_wp[par][0], cw = 0;
#pragma unroll
      for (int q = 1; q < NWK; ++q) {
        const double ov = s_wv[par][q];
        const int op = s_wp[par][q];
        if (ov > cv || (ov == cv && op < cp)) {
          cv = ov;
          cp = op;
          cw = q;
        }
      }
      T cvs = T(0), cu1 = T(0);
      if (cp != INT_MAX) {
        const T *wr = s_wrow[par][cw];
        cvs = wr[j];
        if (j + 1 < NB) {
          cu1 = wr[j + 1];
          if (pend) {
            const T e = wr[j - 1] * s_u[par ^ 1][j + 1];
            cu1 = cu1 - e;
          }
        }
      }
      unsigned long long *hs = pa.hdr + ((size_t)par * kMaxPanelWG + g) * RpHeader<T>::S;
      if (lane < HW) st_agent(hs + lane, RpHeader<T>::word(lane, tg, cvs, cu1, cp));
      if (cp != INT_MAX) {
        unsigned long long *rs = pa.rowd + ((size_t)par * kMaxPanelWG + g) * kRpRowWords;
        for (int idx = (j > 0 ? j - 1 : 0) * WPE + lane; idx < NB * WPE; idx += 64)
          st_agent(rs + idx, tagged(tg, word_of<T>(s_wrow[par][cw][idx / WPE], idx % WPE)));
      }

      // C3. All headers; the winner by the iamax rule.
      bool bail = false;
      double hv = 0.0;
      int hp = INT_MAX;
      T hvs = T(0), hu1 = T(0);
      {
        const unsigned long long *hb = pa.hdr + (size_t)par * kMaxPanelWG * RpHeader<T>::S;
        int spins = 0;
        for (;;) {
          if (spins >= pa.wait_limit) {
            bail = true;
            break;
          }
          bool ok = true;
          if (lane < pa.P) {
            unsigned long long w[HW];
#pragma unroll
            for (int k = 0; k < HW; ++k) {
              w[k] = ld_agent(hb + lane * RpHeader<T>::S + k);
              ok = ok && RpHeader<T>::ok(k, w[k], tg);
            }
            const uint32_t p = RpHeader<T>::decode(w, hv, hvs, hu1);
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
      // C4. The winner's row loads are issued now and stay in flight
      //     through the bookkeeping and the barrier (C5 completes them).
      int q = 0;
      if (!none && !bail) {
        q = __ffsll((long long)__ballot(lane < pa.P && hp == gp)) - 1;
        load_row(q, j, tg, pend);
      }
      if (lane == 0) {
        const int occ = s_at[j];
        const int pw = none ? d : gp;
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
      if (!none) {
        const T v = readlane_t(hvs, q), u1 = readlane_t(hu1, q);
        if (lane == 0) {
          s_recip[par] = rp_recip(v);
          s_u1[par] = u1;
          s_q[par] = q;
          s_u[par][j] = v;
        }
      }
    } else if (pend) {
      // W3. The pending update of columns > j with the previous pivot row,
      //     while wave 0 exchanges; eight columns in flight at a time.
      const T *u = s_u[par ^ 1];
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        if (act[s]) {
          T *row = s_a + tw + s * NW;
          const T l = lr[s];
          int c = j + 1;
          if (c < ncols) {
            const T e = l * u[c];
            c1[s] = row[c * LDC] - e;
            row[c * LDC] = c1[s];
            ++c;
          }
          for (; c + 8 <= ncols; c += 8) {
            T x[8], y[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
              x[q] = row[(c + q) * LDC];
              y[q] = u[c + q];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
              const T e = l * y[q];
              row[(c + q) * LDC] = x[q] - e;
            }
          }
          for (; c < ncols; ++c) {
            const T e = l * u[c];
            row[c * LDC] = row[c * LDC] - e;
          }
        }
      }
    }
    __syncthreads();  // B2
    if (s_bail) break;
    const bool none = s_none[par] != 0;

    if (wave == 0) {
      // C5. The winner's full row, rebuilt as its owner holds it after the
      //     pending update (same operands, same order), for the next
      //     iteration's update; read while the workers go on.
      if (!none) {
        bool bail = false;
        int spins = 0;
        for (;;) {
          bool ok = true;
#pragma unroll
          for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int p = 0; p < WPE; ++p) ok = ok && (uint32_t)(rw[i][p] >> 32) == tg;
#pragma unroll
          for (int p = 0; p < WPE; ++p) ok = ok && (uint32_t)(mw[p] >> 32) == tg;
          if (__all(ok)) break;
          if (spins >= pa.wait_limit || ((++spins & 63) == 0 && ld_agent(pa.err) != 0)) {
            bail = true;
            break;
          }
          load_row(s_q[par], j, tg, pend);
        }
        T rv[NI];
        T mult = T(0);
#pragma unroll
        for (int i = 0; i < NI; ++i) {
          uint32_t w[WPE];
#pragma unroll
          for (int p = 0; p < WPE; ++p) w[p] = (uint32_t)rw[i][p];
          rv[i] = from_words<T>(w);
        }
        if (pend) {
          uint32_t w[WPE];
#pragma unroll
          for (int p = 0; p < WPE; ++p) w[p] = (uint32_t)mw[p];
          mult = from_words<T>(w);
        }
        if (bail) {
          // The status word is set, so the next exchange gives up at once
          // and the loop ends at the next barrier.
          if (lane == 0) {
            s_bail = 1;
            atomic_or_agent(pa.err, kErrWait);
          }
        } else {
#pragma unroll
          for (int i = 0; i < NI; ++i) {
            const int c = j + 1 + lane + 64 * i;
            if (c < NB) {
              T u = rv[i];
              if (pend) {
                const T e = mult * s_u[par ^ 1][c];
                u = u - e;
              }
              s_u[par][c] = u;
            }
          }
        }
      }
    } else {
      // W5. Interchange bookkeeping (the pivot row takes position d, the row
      //     at d takes the pivot's old position), then the multipliers of
      //     column j and the update of column j + 1.  A column without a
      //     nonzero candidate is left unscaled (getrf.rs:72-73).
      const int pw = s_pw[par], occ = s_occ[par];
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        if (act[s] && cpos[s] == pw) {
          act[s] = false;
          cpos[s] = d;
        } else if (base + tw + s * NW == occ) {
          cpos[s] = pw;
        }
      }
      const T recip = s_recip[par], u1 = s_u1[par];
#pragma unroll
      for (int s = 0; s < RPT; ++s) {
        if (act[s]) {
          T *row = s_a + tw + s * NW;
          const T v1 = (j + 1 < ncols) ? (pend ? c1[s] : row[(j + 1) * LDC]) : T(0);
          if (!none) {
            const T l = cj[s] * recip;
            row[j * LDC] = l;
            lr[s] = l;
            if (j + 1 < ncols) {
              const T e = l * u1;
              cj[s] = v1 - e;
              row[(j + 1) * LDC] = cj[s];
            }
          } else {
            cj[s] = v1;
          }
        }
      }
    }
    pend = !none;
  }
  __syncthreads();
  if (s_bail) {
    // Given up: identity pivots for the columns not decided keep the
    // interchanges queued after this launch in bounds (thread 0 of workgroup
    // 0 writes every pivot of the leaf).
    if (g == 0 && t == 0) {
      for (int jj = j; jj < pa.npiv; ++jj) pa.ipiv[pa.row0 + jj] = pa.row0 + jj;
      if (pa.map_count) *pa.map_count = 0;
    }
    return;
  }

  // The last pivot's update of the columns beyond the pivots (wide leaves).
  if (pend && wave > 0) {
    const int jl = pa.npiv - 1;
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      if (act[s]) {
        T *row = s_a + tw + s * NW;
        for (int c = jl + 2; c < ncols; ++c) {
          const T e = lr[s] * s_u[jl & 1][c];
          row[c * LDC] = row[c * LDC] - e;
        }
      }
    }
  }

  // Row map of the leaf's interchanges in gather form (rows[e] receives the
  // old contents of src[e]; rows[e] == src[e] is a no-op).  Slot e < npiv:
  // the final occupant of position row0 + e.  Slot npiv + e: where the row
  // that started at row0 + e went, when it left [row0, row0 + npiv).  Each
  // slot is written once, by the owner of the row it describes.
  if (wave > 0) {
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
      const int orig = base + tw + s * NW;
      const bool valid = orig < pa.m;
      s_fin[tw + s * NW] = valid ? cpos[s] : -1;
      if (pa.map_rows && valid) {
        const int top1 = pa.row0 + pa.npiv;
        const int fin = cpos[s];
        if (fin < top1) {
          pa.map_rows[fin - pa.row0] = fin;
          pa.map_src[fin - pa.row0] = orig;
        }
        if (orig < top1) {
          pa.map_rows[pa.npiv + orig - pa.row0] = fin < top1 ? orig : fin;
          pa.map_src[pa.npiv + orig - pa.row0] = orig;
        }
      }
    }
  }
  if (pa.map_count && g == 0 && t == 0) *pa.map_count = 2 * pa.npiv;
  __syncthreads();

  // Write-back: every row to its final position.
  for (int i = wave; i < R; i += NWK + 1) {
    const int dst = s_fin[i];
    if (dst < 0) continue;
    T *out = pa.a + (int64_t)dst * pa.lda + pa.col0;
#pragma unroll
    for (int q = 0; q < NI; ++q) {
      const int c = lane + 64 * q;
      if (c < ncols) out[c] = s_a[c * LDC + i];
    }
  }
}

template <typename T, int NB, int NWK, int RPT>
static hipError_t rp_launch(lair_context *ctx, const RpArgs<T> &pa) {
  auto *k = &rpanel_kernel<T, NB, NWK, RPT>;
  constexpr int R = 64 * NWK * RPT;
  const size_t shm = (size_t)NB * (R + 1) * sizeof(T);
  static bool once = ((void)hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                                (int)shm),
                      (void)hipGetLastError(), true);
  (void)once;
  k<<<pa.P, 64 * (NWK + 1), shm, ctx->stream>>>(pa);
  return hipGetLastError();
}

template <typename T>
int rpanel_rows_per_wg(lair_context *ctx, int nbw, int rows) {
  const int maxp = std::min(ctx->rp_max_wg, kMaxPanelWG);
  // {leaf width, rows per workgroup}: the rows of a workgroup fill about
  // 130 KB of LDS, and a header or row record fits its exchange slot.
  static const int kF[][2] = {{128, 256}, {64, 512}, {32, 1024}};
  static const int kD[][2] = {{64, 256}, {32, 512}, {16, 1024}};
  static const int kC[][2] = {{64, 256}, {32, 512}, {16, 1024}};  // complex<float>
  static const int kZ[][2] = {{32, 256}, {16, 512}, {0, 0}};      // complex<double>
  const auto &tab = sizeof(T) == 4 ? kF : sizeof(T) == 16 ? kZ : cplx_traits<T>::is ? kC : kD;
  for (const auto &e : tab)
    if (e[0] > 0 && e[0] == nbw && (rows + e[1] - 1) / e[1] <= maxp) return e[1];
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
                         int nbw, int64_t *ipiv, int *map_rows, int *map_src, int *map_count) {
  const int rows = m - row0;
  if (rows <= 0 || wcols <= 0) return hipSuccess;
  if (wcols > nbw || m >= (int)kRpNone) return hipErrorInvalidValue;
  const int R = rpanel_rows_per_wg<T>(ctx, nbw, rows);
  if (R == 0) return hipErrorInvalidValue;
  // Header tags carry 16 bits: start a new tag epoch (cleared slots) before
  // they could wrap inside this launch.
  if ((ctx->rp_seq & 0xFFFFu) + (uint32_t)nbw + 2u > 0xFFFFu || ctx->rp_seq > 0xF0000000u) {
    hipError_t e = launch_fill_u32(ctx, reinterpret_cast<uint32_t *>(ctx->rp_hdr),
                                   2 * kRpHdrTotal, 0u);
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
  pa.map_rows = map_rows;
  pa.map_src = map_src;
  pa.map_count = map_count;
  pa.hdr = ctx->rp_hdr + (cplx_traits<T>::is ? kRpHdrComplexOff : 0);
  pa.rowd = ctx->rp_rowd;
  pa.seq = ctx->rp_seq;
  pa.err = ctx->err;
  pa.wait_limit = ctx->wait_limit;
  const int slot = prof_begin(ctx, LAIR_KCLASS_PANEL);
  hipError_t e;
  if constexpr (sizeof(T) == 16) {
    if (nbw == 32) e = rp_launch<T, 32, 4, 1>(ctx, pa);
    else e = rp_launch<T, 16, 4, 2>(ctx, pa);
  } else if constexpr (cplx_traits<T>::is) {
    if (nbw == 64) e = rp_launch<T, 64, 4, 1>(ctx, pa);
    else if (nbw == 32) e = rp_launch<T, 32, 4, 2>(ctx, pa);
    else e = rp_launch<T, 16, 4, 4>(ctx, pa);
  } else if constexpr (sizeof(T) == 4) {
    if (nbw == 128) e = rp_launch<T, 128, 4, 1>(ctx, pa);
    else if (nbw == 64) e = rp_launch<T, 64, 4, 2>(ctx, pa);
    else e = rp_launch<T, 32, 4, 4>(ctx, pa);
  } else {
    if (nbw == 64) e = rp_launch<T, 64, 4, 1>(ctx, pa);
    else if (nbw == 32) e = rp_launch<T, 32, 4, 2>(ctx, pa);
    else e = rp_launch<T, 16, 4, 4>(ctx, pa);
  }
  prof_end(ctx, slot, LAIR_KCLASS_PANEL, (double)pa.npiv * rows * wcols,
           2.0 * sizeof(T) * (double)rows * wcols);
  ctx->rp_seq += (uint32_t)nbw + 1u;
  return e;
}

template <typename T>
int rpanel_grid(lair_context *ctx, int nbw, int rows) {
  const int R = rpanel_rows_per_wg<T>(ctx, nbw, rows);
  return R > 0 ? (rows + R - 1) / R : 0;
}

template int rpanel_grid<float>(lair_context *, int, int);
template int rpanel_grid<double>(lair_context *, int, int);
template int rpanel_leaf_width<float>(lair_context *, int, int);
template int rpanel_leaf_width<double>(lair_context *, int, int);
template hipError_t launch_rpanel<float>(lair_context *, float *, int64_t, int, int, int, int, int,
                                         int64_t *, int *, int *, int *);
template hipError_t launch_rpanel<double>(lair_context *, double *, int64_t, int, int, int, int, int,
                                          int64_t *, int *, int *, int *);
template int rpanel_grid<zcplx>(lair_context *, int, int);
template int rpanel_grid<ccplx>(lair_context *, int, int);
template int rpanel_leaf_width<zcplx>(lair_context *, int, int);
template int rpanel_leaf_width<ccplx>(lair_context *, int, int);
template hipError_t launch_rpanel<zcplx>(lair_context *, zcplx *, int64_t, int, int, int, int, int, int64_t *,
                                         int *, int *, int *);
template hipError_t launch_rpanel<ccplx>(lair_context *, ccplx *, int64_t, int, int, int, int, int, int64_t *,
                                         int *, int *, int *);

}  // namespace lair
