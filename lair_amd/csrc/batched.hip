// Batched solve of many independent small systems A_i X_i = B_i (n_i <= 256,
// nrhs_i <= 16) in persistent launches, one per size class of the plan.
//
// Workgroups pull system indices from an atomic queue (an order sorted by
// decreasing n, so the largest systems start first) until the queue is
// empty; every workgroup reaches that exit.  Each system is factored as the
// augmented matrix [A | B] with a blocked right-looking LU with partial
// pivoting (src/lapack/getrf.rs:46-120, blocked; the interchanges are applied
// to B as getrs.rs:23 does), so the forward substitution comes with the
// factorization; a blocked back substitution with U gives X (getrs.rs:30-36).
//
// Per panel of NB columns:
//  * thread t keeps panel row kb + t in registers; per column, a DPP argmax
//    per wave and a fixed-order reduction of the four wave winners (ties ->
//    lowest row; NaN and zero never chosen, src/blas/iamax.rs:6-21), the
//    interchange, the reciprocal scaling and the rank-1 update
//    (getrf.rs:76-88); once the remaining rows fit one wave, wave 0 alone;
//  * the panel's interchanges form one row map, applied to the other columns
//    of [A | B] as a gather;
//  * U12 = L11^-1 A12 column by column, then A22 -= L21 U12 with L21 and U12
//    staged in LDS and 4 x 4 register tiles.
// The augmented matrix lives in LDS when it fits (copied in and out), else it
// is updated in place in global memory (A and B buffers).
#include <algorithm>
#include <climits>
#include <numeric>
#include <vector>

#include "lair_internal.h"

namespace lair {

constexpr int kBatchMaxN = 256;
constexpr int kBatchMaxRhs = 16;
constexpr int kBatchLDU = kBatchMaxN + kBatchMaxRhs + 1;

template <typename T>
struct BatchArgs {
  int count;
  const int *n;          // [count]
  const int *nrhs;       // [count]
  const int64_t *a_off;  // element offsets of A_i (row-major, ld = n_i)
  const int64_t *b_off;  // element offsets of B_i (row-major, ld = nrhs_i)
  const int64_t *p_off;  // offsets of ipiv_i (n_i entries)
  const int *order;      // processing order
  T *a;
  T *b;
  int64_t *ipiv;
  int *info;             // -1, or the last zero-pivot column (getrf.rs:72-73)
  int *queue;            // work counter, zero at launch
  int lds_max_elems;     // largest augmented matrix (n * (n + nrhs)) kept in LDS
};

// [A | B] with A and B in separate row-major buffers.
template <typename T>
struct SplitMat {
  T *a;
  T *b;
  int n, r;
  __device__ __forceinline__ T &operator()(int i, int c) const {
    return c < n ? a[i * n + c] : b[i * r + (c - n)];
  }
};

// [A | B] contiguous (leading dimension n + r), e.g. in LDS.
template <typename T>
struct DenseMat {
  T *m;
  int ld;
  __device__ __forceinline__ T &operator()(int i, int c) const { return m[i * ld + c]; }
};

template <int CTRL, int ROWM = 0xF>
__device__ __forceinline__ double bdpp(double x) {
  const int lo = __double2loint(x), hi = __double2hiint(x);
  const int l2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROWM, 0xF, false);
  const int h2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROWM, 0xF, false);
  return __hiloint2double(h2, l2);
}

// Maximum of non-negative doubles over the wave (result uniform).
__device__ __forceinline__ double bwave_max(double v) {
  v = fmax(v, bdpp<0xB1>(v));        // quad_perm [1,0,3,2]
  v = fmax(v, bdpp<0x4E>(v));        // quad_perm [2,3,0,1]
  v = fmax(v, bdpp<0x141>(v));       // row_half_mirror
  v = fmax(v, bdpp<0x140>(v));       // row_mirror
  v = fmax(v, bdpp<0x142, 0xA>(v));  // row_bcast:15
  v = fmax(v, bdpp<0x143, 0xC>(v));  // row_bcast:31
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, int NB>
struct BatchSolver {
  int t, lane, wave;
  T *sL;      // [kBatchMaxN][NB]: L11 and L21 of the current panel
  T *sU;      // [NB][kBatchLDU]: U12 (and the rows of B) of the current panel
  T *scand;   // [2][4][NB]: each wave's candidate pivot row, by column parity
  T *sdiag;   // [2][NB]: the diagonal row, by column parity
  int *spiv;  // [NB]
  int *srows, *ssrc, *scnt;  // row map of the panel's interchanges
  int *s_srcx;                // [2][5]: origins of the candidate rows and the diagonal row, by column parity
  int *sinfo;
  double *s_wv;  // [2][4] wave maxima (workgroup panel), by column parity
  int *s_wr;     // [2][4] their rows

  // The panel's columns kb..kb+w factored by the whole workgroup: thread t
  // keeps row kb + t (n <= kBatchMaxN = 256 threads).  Per column, one
  // barrier: every wave finds its argmax with DPP (lowest row among equal
  // maxima; NaN and zero never chosen, iamax.rs:6-21) and its winning lane
  // publishes the whole row, while the thread holding the diagonal row
  // publishes that row; after the barrier every thread reduces the four wave
  // winners in a fixed order, and the interchange, reciprocal scaling and
  // rank-1 update (separate multiply and subtract, getrf.rs:76-88) follow
  // from LDS.  The buffers alternate by column parity, so the next column's
  // writes cannot overtake this column's reads.  Every thread also tracks
  // which original row its contents came from; the rows whose origin changed
  // form the panel's row map (rows[e] receives row src[e]), listed in any
  // order since it is a permutation of distinct rows.
  template <class Mat>
  __device__ __forceinline__ void panel_wg(const Mat &M, int n, int kb, int w, int64_t *ipiv_out) {
#pragma clang fp contract(off)
    const int row = kb + t;
    const bool has = row < n;
    int src = row;
    if (t == 0) *scnt = 0;
    T v[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) v[c] = (has && c < w) ? M(row, kb + c) : T(0);
    if (n - kb <= 64) {
      // Every remaining row belongs to wave 0, which runs the column loop
      // alone: the same argmax, tie rule and arithmetic, with a wavefront
      // fence in place of the workgroup barrier and no cross-wave reduction.
      if (wave == 0) {
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          if (j < w) {
            const int d = kb + j;
            const int par = j & 1;
            T *prow = scand + par * 4 * NB;
            T *diag = sdiag + par * NB;
            int *srcs = s_srcx + par * 5;
            const double x = (has && row >= d) ? fabs((double)v[j]) : 0.0;
            const int xr = (has && row >= d && x > 0.0) ? row : INT_MAX;
            const double vmax = bwave_max(xr != INT_MAX ? x : 0.0);
            unsigned long long mm = __ballot(xr != INT_MAX && x == vmax);
            int wr = INT_MAX;
            while (mm) {
              const int l = __ffsll((long long)mm) - 1;
              mm &= mm - 1;
              const int rl = __builtin_amdgcn_readlane(xr, l);
              wr = rl < wr ? rl : wr;
            }
            const bool singular = (wr == INT_MAX);
            const int pr = singular ? d : wr;
            if (row == pr) {
#pragma unroll
              for (int c = 0; c < NB; ++c) prow[c] = v[c];
              srcs[0] = src;
            }
            if (row == d) {
#pragma unroll
              for (int c = 0; c < NB; ++c) diag[c] = v[c];
              srcs[4] = src;
            }
            wave_sync();
            if (pr != d) {
              if (row == d) {
#pragma unroll
                for (int c = 0; c < NB; ++c) v[c] = prow[c];
                src = srcs[0];
              } else if (row == pr) {
#pragma unroll
                for (int c = 0; c < NB; ++c) v[c] = diag[c];
                src = srcs[4];
              }
            }
            if (lane == 0) {
              spiv[j] = pr;
              if (singular) *sinfo = d;
            }
            if (!singular && has && row > d) {
              const T recip = T(1) / prow[j];
              const T l = v[j] * recip;
              v[j] = l;
#pragma unroll
              for (int c = j + 1; c < NB; ++c) {
                const T e = l * prow[c];
                v[c] = v[c] - e;
              }
            }
            wave_sync();
          }
        }
      }
    } else {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < w) {
        const int d = kb + j;
        const int par = j & 1;
        T *cand = scand + par * 4 * NB;
        T *diag = sdiag + par * NB;
        int *srcs = s_srcx + par * 5;  // [4 candidates | diagonal]
        const double x = (has && row >= d) ? fabs((double)v[j]) : 0.0;
        const int xr = (has && row >= d && x > 0.0) ? row : INT_MAX;
        const double vmax = bwave_max(xr != INT_MAX ? x : 0.0);
        const unsigned long long m = __ballot(xr != INT_MAX && x == vmax);
        int wr = INT_MAX;
        unsigned long long mm = m;
        while (mm) {
          const int l = __ffsll((long long)mm) - 1;
          mm &= mm - 1;
          const int rl = __builtin_amdgcn_readlane(xr, l);
          wr = rl < wr ? rl : wr;
        }
        if (lane == 0) {
          s_wv[par * 4 + wave] = m ? vmax : 0.0;
          s_wr[par * 4 + wave] = wr;
        }
        if (row == wr) {
#pragma unroll
          for (int c = 0; c < NB; ++c) cand[wave * NB + c] = v[c];
          srcs[wave] = src;
        }
        if (row == d) {
#pragma unroll
          for (int c = 0; c < NB; ++c) diag[c] = v[c];
          srcs[4] = src;
        }
        __syncthreads();
        double bv = s_wv[par * 4];
        int pr = s_wr[par * 4], pw = 0;
#pragma unroll
        for (int q = 1; q < 4; ++q) {
          const double ov = s_wv[par * 4 + q];
          const int orr = s_wr[par * 4 + q];
          if (ov > bv || (ov == bv && orr < pr)) {
            bv = ov;
            pr = orr;
            pw = q;
          }
        }
        const bool singular = (pr == INT_MAX);
        if (singular) pr = d;
        const T *prow = cand + pw * NB;  // the pivot row (valid unless singular)
        if (pr != d) {
          if (row == d) {
#pragma unroll
            for (int c = 0; c < NB; ++c) v[c] = prow[c];
            src = srcs[pw];
          } else if (row == pr) {
#pragma unroll
            for (int c = 0; c < NB; ++c) v[c] = diag[c];
            src = srcs[4];
          }
        }
        if (t == 0) {
          spiv[j] = pr;
          if (singular) *sinfo = d;
        }
        if (!singular && has && row > d) {
          const T recip = T(1) / prow[j];
          const T l = v[j] * recip;
          v[j] = l;
#pragma unroll
          for (int c = j + 1; c < NB; ++c) {
            const T e = l * prow[c];
            v[c] = v[c] - e;
          }
        }
      }
    }
    }
    if (has) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < w) {
          M(row, kb + c) = v[c];
          sL[t * NB + c] = v[c];
        }
      if (src != row) {
        const int e = atomicAdd(scnt, 1);
        srows[e] = row;
        ssrc[e] = src;
      }
    }
    __syncthreads();
    if (t < w) ipiv_out[kb + t] = spiv[t];
  }

  template <class Mat>
  __device__ __forceinline__ void solve(const Mat &M, int n, int r, int64_t *ipiv_out) {
    const int nc = n + r;
    if (t == 0) *sinfo = -1;
    for (int kb = 0; kb < n; kb += NB) {
      const int w = n - kb < NB ? n - kb : NB;
      panel_wg(M, n, kb, w, ipiv_out);
      __syncthreads();
      // Interchanges on the columns of [A | B] outside the panel.
      const int cnt = *scnt;
      for (int ci = t; ci < nc - w; ci += 256) {
        __asm__ volatile("" ::: "memory");  // the row map is re-read per column, not held in registers
        const int c = ci < kb ? ci : ci + w;
        T vals[2 * NB];
#pragma unroll
        for (int e = 0; e < 2 * NB; ++e)
          if (e < cnt) vals[e] = M(ssrc[e], c);
#pragma unroll
        for (int e = 0; e < 2 * NB; ++e)
          if (e < cnt) M(srows[e], c) = vals[e];
      }
      __syncthreads();
      // U12 = L11^-1 A12 for the columns right of the panel (and of B).
      const int c0 = kb + w;
      for (int c = c0 + t; c < nc; c += 256) {
        T x[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) x[i] = i < w ? M(kb + i, c) : T(0);
#pragma unroll
        for (int i = 1; i < NB; ++i) {
          // Row i's loads of L11 are issued after row i - 1 is solved (an
          // empty asm that uses x[i - 1] and orders memory): scheduled
          // freely, all NB (NB - 1) / 2 entries would be held in registers.
          __asm__ volatile("" : "+v"(x[i - 1]) : : "memory");
          T acc = x[i];
#pragma unroll
          for (int q = 0; q < i; ++q) acc -= sL[i * NB + q] * x[q];
          x[i] = acc;
        }
#pragma unroll
        for (int i = 0; i < NB; ++i)
          if (i < w) {
            M(kb + i, c) = x[i];
            sU[i * kBatchLDU + (c - c0)] = x[i];
          }
      }
      __syncthreads();
      // A22 -= L21 U12 over rows c0..n-1 and columns c0..nc-1.
      const int rows = n - c0, cols = nc - c0;
      if (rows > 0) {
        const int tr = (rows + 3) / 4, tc = (cols + 3) / 4;
        for (int tile = t; tile < tr * tc; tile += 256) {
          const int i0 = (tile / tc) * 4, j0 = (tile % tc) * 4;
          // The tile of A22 is loaded first so that its latency overlaps the
          // products from LDS.
          T cv[4][4], acc[4][4];
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
              cv[a][b] = (i0 + a < rows && j0 + b < cols) ? M(c0 + i0 + a, c0 + j0 + b) : T(0);
              acc[a][b] = T(0);
            }
#pragma unroll
          for (int q = 0; q < NB; ++q) {
            if (q < w) {
              T l[4], u[4];
#pragma unroll
              for (int a = 0; a < 4; ++a) l[a] = (i0 + a < rows) ? sL[(w + i0 + a) * NB + q] : T(0);
#pragma unroll
              for (int b = 0; b < 4; ++b) u[b] = (j0 + b < cols) ? sU[q * kBatchLDU + j0 + b] : T(0);
#pragma unroll
              for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] += l[a] * u[b];
            }
          }
#pragma unroll
          for (int a = 0; a < 4; ++a)
            if (i0 + a < rows) {
#pragma unroll
              for (int b = 0; b < 4; ++b)
                if (j0 + b < cols) M(c0 + i0 + a, c0 + j0 + b) = cv[a][b] - acc[a][b];
            }
        }
      }
      __syncthreads();
    }
    // Back substitution U X = Y on the right-hand-side columns, blocked from
    // the bottom (getrs.rs:30-36, true division).  Y stays in LDS from start
    // to end.  Per block of NB rows, one thread per right-hand side solves the
    // diagonal block in registers, then one thread per row above subtracts
    // the block's contribution.  The next block's diagonal entries and row
    // segments of U are loaded into registers while the current block is
    // solved, so the chain of blocks carries no memory round trip.
    static_assert(kBatchMaxN * kBatchMaxRhs + NB * NB <= kBatchMaxN * NB + NB * kBatchLDU,
                  "Y and the diagonal block fit the panel buffers");
    T *sY = sL;                               // [n][r]: Y, then X
    T *sD = sL + kBatchMaxN * kBatchMaxRhs;   // [NB][NB]: diagonal block of U
    for (int e = t; e < n * r; e += 256) sY[e] = M(e / r, n + e % r);
    auto load_d = [&](int kb) -> T {
      const int w = n - kb < NB ? n - kb : NB, i = t / NB, q = t % NB;
      return (t < NB * NB && i < w && q < w && q >= i) ? M(kb + i, kb + q) : T(0);
    };
    auto load_u = [&](int kb, T (&u)[NB]) {
      const int w = n - kb < NB ? n - kb : NB;
#pragma unroll
      for (int q = 0; q < NB; ++q) u[q] = (t < kb && q < w) ? M(t, kb + q) : T(0);
    };
    int kb = ((n - 1) / NB) * NB;
    T dcur = load_d(kb), ucur[NB];
    load_u(kb, ucur);
    for (;;) {
      const int w = n - kb < NB ? n - kb : NB;
      const int kn = kb - NB;
      if (t < NB * NB) sD[t] = dcur;
      T dnext = T(0), unext[NB];
#pragma unroll
      for (int q = 0; q < NB; ++q) unext[q] = T(0);
      if (kn >= 0) {
        dnext = load_d(kn);
        load_u(kn, unext);
      }
      __syncthreads();
      if (t < r) {
        T x[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) x[i] = i < w ? sY[(kb + i) * r + t] : T(0);
#pragma unroll
        for (int i = NB - 1; i >= 0; --i) {
          if (i < w) {
            T acc = x[i];
#pragma unroll
            for (int q = i + 1; q < NB; ++q)
              if (q < w) acc -= sD[i * NB + q] * x[q];
            x[i] = acc / sD[i * NB + i];
          }
        }
#pragma unroll
        for (int i = 0; i < NB; ++i)
          if (i < w) sY[(kb + i) * r + t] = x[i];
      }
      __syncthreads();
      if (t < kb) {
        for (int c = 0; c < r; ++c) {
          T acc = sY[t * r + c];
#pragma unroll
          for (int q = 0; q < NB; ++q)
            if (q < w) acc -= ucur[q] * sY[(kb + q) * r + c];
          sY[t * r + c] = acc;
        }
      }
      if (kn < 0) break;
      __syncthreads();
      dcur = dnext;
#pragma unroll
      for (int q = 0; q < NB; ++q) ucur[q] = unext[q];
      kb = kn;
    }
    __syncthreads();
    for (int e = t; e < n * r; e += 256) M(e / r, n + e % r) = sY[e];
    // The caller reads M with another thread-to-element mapping (the copy out
    // of LDS), and the next system reuses sY.
    __syncthreads();
  }
};

// OCC > 1 caps the registers so that OCC workgroups fit on a CU (each wave
// of a 256-thread workgroup runs on its own SIMD); values beyond the cap
// spill to scratch.  More resident workgroups hide more memory latency.
template <typename T, int NB, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC)))
batched_gesv_kernel(BatchArgs<T> p) {
  extern __shared__ unsigned char smem_raw[];
  T *sL = reinterpret_cast<T *>(smem_raw);
  T *sU = sL + kBatchMaxN * NB;
  T *scand = sU + NB * kBatchLDU;
  T *sdiag = scand + 8 * NB;
  T *sM = sdiag + 2 * NB;  // augmented matrix when it fits
  __shared__ int spiv[NB];
  __shared__ int srows[2 * NB], ssrc[2 * NB];
  __shared__ int scnt, sinfo, ssys;
  __shared__ int s_srcx[10];
  __shared__ double s_wv[8];
  __shared__ int s_wr[8];
  BatchSolver<T, NB> S;
  S.t = threadIdx.x;
  S.lane = threadIdx.x & 63;
  S.wave = threadIdx.x >> 6;
  S.sL = sL;
  S.sU = sU;
  S.scand = scand;
  S.sdiag = sdiag;
  S.spiv = spiv;
  S.srows = srows;
  S.ssrc = ssrc;
  S.scnt = &scnt;
  S.s_srcx = s_srcx;
  S.sinfo = &sinfo;
  S.s_wv = s_wv;
  S.s_wr = s_wr;
  for (;;) {
    if (threadIdx.x == 0) ssys = atomicAdd(p.queue, 1);
    __syncthreads();
    const int idx = ssys;
    if (idx >= p.count) break;
    const int sys = p.order[idx];
    const int n = p.n[sys], r = p.nrhs[sys];
    T *A = p.a + p.a_off[sys];
    T *B = p.b + p.b_off[sys];
    int64_t *piv = p.ipiv + p.p_off[sys];
    const int nc = n + r;
    if (n * nc <= p.lds_max_elems) {
      for (int e = threadIdx.x; e < n * nc; e += 256) {
        const int i = e / nc, c = e % nc;
        sM[e] = c < n ? A[i * n + c] : B[i * r + (c - n)];
      }
      __syncthreads();
      S.solve(DenseMat<T>{sM, nc}, n, r, piv);
      for (int e = threadIdx.x; e < n * nc; e += 256) {
        const int i = e / nc, c = e % nc;
        if (c < n)
          A[i * n + c] = sM[e];
        else
          B[i * r + (c - n)] = sM[e];
      }
    } else {
      S.solve(SplitMat<T>{A, B, n, r}, n, r, piv);
    }
    if (threadIdx.x == 0) p.info[sys] = sinfo;
    __syncthreads();
  }
}

template <typename T, int NB>
static size_t batched_lds_bytes(int lds_max_elems) {
  return sizeof(T) * ((size_t)kBatchMaxN * NB + (size_t)NB * kBatchLDU + 10 * NB + lds_max_elems);
}

__global__ void zero_int_kernel(int *p) { *p = 0; }

template <typename T>
hipError_t launch_batched_gesv(lair_context *ctx, const BatchLaunch &bl, T *a, T *b, int64_t *ipiv,
                               int *info) {
  BatchArgs<T> p;
  p.count = bl.count;
  p.n = bl.n;
  p.nrhs = bl.nrhs;
  p.a_off = bl.a_off;
  p.b_off = bl.b_off;
  p.p_off = bl.p_off;
  p.order = bl.order;
  p.a = a;
  p.b = b;
  p.ipiv = ipiv;
  p.info = info;
  p.queue = bl.queue;
  p.lds_max_elems = bl.lds_max_elems;
  // The queue counter is cleared by a kernel (not a memset), so the pair can
  // be captured in a graph and replayed.
  zero_int_kernel<<<1, 1, 0, ctx->stream>>>(bl.queue);
  const unsigned grid = (unsigned)bl.grid;
  const size_t shm = bl.nb == 8 ? batched_lds_bytes<T, 8>(bl.lds_max_elems)
                                : batched_lds_bytes<T, 16>(bl.lds_max_elems);
  auto go = [&](auto kernel) { kernel<<<grid, 256, shm, ctx->stream>>>(p); };
  if (bl.nb == 8) {
    if (bl.occ == 4)
      go(batched_gesv_kernel<T, 8, 4>);
    else if (bl.occ == 3)
      go(batched_gesv_kernel<T, 8, 3>);
    else
      go(batched_gesv_kernel<T, 8, 1>);
  } else {
    if (bl.occ == 4)
      go(batched_gesv_kernel<T, 16, 4>);
    else if (bl.occ == 3)
      go(batched_gesv_kernel<T, 16, 3>);
    else
      go(batched_gesv_kernel<T, 16, 1>);
  }
  return hipGetLastError();
}

// Lifts the dynamic-LDS limit of the batched kernels to the full 160 KB; done
// before any launch is captured into a graph.
void batched_prepare() {
  static bool done = false;
  if (done) return;
  for (const void *f : {reinterpret_cast<const void *>(batched_gesv_kernel<double, 8, 1>),
                        reinterpret_cast<const void *>(batched_gesv_kernel<double, 8, 3>),
                        reinterpret_cast<const void *>(batched_gesv_kernel<double, 8, 4>),
                        reinterpret_cast<const void *>(batched_gesv_kernel<double, 16, 1>),
                        reinterpret_cast<const void *>(batched_gesv_kernel<double, 16, 3>),
                        reinterpret_cast<const void *>(batched_gesv_kernel<double, 16, 4>),
                        reinterpret_cast<const void *>(batched_gesv_kernel<float, 8, 1>),
                        reinterpret_cast<const void *>(batched_gesv_kernel<float, 8, 3>),
                        reinterpret_cast<const void *>(batched_gesv_kernel<float, 8, 4>),
                        reinterpret_cast<const void *>(batched_gesv_kernel<float, 16, 1>),
                        reinterpret_cast<const void *>(batched_gesv_kernel<float, 16, 3>),
                        reinterpret_cast<const void *>(batched_gesv_kernel<float, 16, 4>)})
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  (void)hipGetLastError();
  done = true;
}

template hipError_t launch_batched_gesv<double>(lair_context *, const BatchLaunch &, double *,
                                                double *, int64_t *, int *);
template hipError_t launch_batched_gesv<float>(lair_context *, const BatchLaunch &, float *, float *,
                                               int64_t *, int *);

size_t batched_lds_bytes_for(int elem_size, int nb, int lds_max_elems) {
  return (size_t)elem_size *
         ((size_t)kBatchMaxN * nb + (size_t)nb * kBatchLDU + 10 * nb + lds_max_elems);
}

}  // namespace lair
