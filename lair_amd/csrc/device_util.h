// Device helpers shared by the panel kernels: agent-scope loads and stores,
// tagged 64-bit exchange words and the wave-level iamax reduction.
#pragma once

#include <climits>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace lair {

template <typename T>
__device__ __forceinline__ void st_agent(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_agent(const T *p) {
  return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bits of the context's device status word.
constexpr int32_t kErrWait = 1;   // a bounded wait gave up
constexpr int32_t kErrPivot = 2;  // a caller-supplied pivot was outside [0, n)

__device__ __forceinline__ void atomic_or_agent(int32_t *p, int32_t bits) {
  __hip_atomic_fetch_or(p, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long tagged(uint32_t tag, uint32_t data) {
  return ((unsigned long long)tag << 32) | data;
}

// Complex element (re, im), stored interleaved like num-complex's
// Complex<R>.  The arithmetic follows num-complex 0.4 with every product and
// sum rounded on its own (no fused multiply-add):
//   a * b  = (a.re b.re - a.im b.im, a.re b.im + a.im b.re)
//   a - b  = (a.re - b.re, a.im - b.im)
// The reciprocal and quotient are in cplx_recip / cplx_div below.
template <typename R>
struct alignas(2 * sizeof(R)) cplx {
  R re, im;
  cplx() = default;
  __host__ __device__ constexpr cplx(R r) : re(r), im(R(0)) {}
  __host__ __device__ constexpr cplx(R r, R i) : re(r), im(i) {}
};

template <typename T>
struct cplx_traits {
  static constexpr bool is = false;
  typedef T real;
};
template <typename R>
struct cplx_traits<cplx<R>> {
  static constexpr bool is = true;
  typedef R real;
};

template <typename R>
__device__ __forceinline__ cplx<R> operator*(cplx<R> a, cplx<R> b) {
#pragma clang fp contract(off)
  const R rr = a.re * b.re, ii = a.im * b.im, ri = a.re * b.im, ir = a.im * b.re;
  return cplx<R>(rr - ii, ri + ir);
}
template <typename R>
__device__ __forceinline__ cplx<R> operator-(cplx<R> a, cplx<R> b) {
  return cplx<R>(a.re - b.re, a.im - b.im);
}
template <typename R>
__device__ __forceinline__ cplx<R> &operator-=(cplx<R> &a, cplx<R> b) {
  a = a - b;
  return a;
}

// One / z as num-complex's Div of Complex::one() by z (src/lapack/getrf.rs:76
// computes A::one() / pivot): ((1 re + 0 im) / |z|^2, (0 re - 1 im) / |z|^2)
// with |z|^2 = re re + im im.
template <typename R>
__device__ __forceinline__ cplx<R> cplx_recip(cplx<R> z) {
#pragma clang fp contract(off)
  const R rr = z.re * z.re, ii = z.im * z.im;
  const R ns = rr + ii;
  const R one_re = R(1) * z.re, zero_im = R(0) * z.im, zero_re = R(0) * z.re, one_im = R(1) * z.im;
  return cplx<R>((one_re + zero_im) / ns, (zero_re - one_im) / ns);
}

// x /= d as num-complex's DivAssign (src/lapack/getrs.rs:34):
// re = (x.re d.re + x.im d.im) / |d|^2, im = (x.im d.re - x.re d.im) / |d|^2.
template <typename R>
__device__ __forceinline__ cplx<R> cplx_div(cplx<R> x, cplx<R> d) {
#pragma clang fp contract(off)
  const R rr = d.re * d.re, ii = d.im * d.im;
  const R ns = rr + ii;
  const R a = x.re * d.re, b = x.im * d.im, c = x.im * d.re, e = x.re * d.im;
  return cplx<R>((a + b) / ns, (c - e) / ns);
}

template <typename T>
__device__ __forceinline__ uint32_t word_of(T x, int part) {
  if constexpr (cplx_traits<T>::is) {
    typedef typename cplx_traits<T>::real R;
    constexpr int H = sizeof(R) / 4;
    return part < H ? word_of<R>(x.re, part) : word_of<R>(x.im, part - H);
  } else if constexpr (sizeof(T) == 8) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return part == 0 ? (uint32_t)(b >> 32) : (uint32_t)b;
  } else {
    return (uint32_t)__float_as_uint(x);
  }
}

template <typename T>
__device__ __forceinline__ T from_words(const uint32_t *w) {
  if constexpr (cplx_traits<T>::is) {
    typedef typename cplx_traits<T>::real R;
    constexpr int H = sizeof(R) / 4;
    return T(from_words<R>(w), from_words<R>(w + H));
  } else if constexpr (sizeof(T) == 8) {
    return __longlong_as_double((long long)(((unsigned long long)w[0] << 32) | w[1]));
  } else {
    return __uint_as_float(w[0]);
  }
}

// Lane shuffle and read-lane of a real or complex value.
template <typename T>
__device__ __forceinline__ T shfl_t(T x, int src) {
  if constexpr (cplx_traits<T>::is)
    return T(__shfl(x.re, src), __shfl(x.im, src));
  else
    return __shfl(x, src);
}
template <typename R>
__device__ __forceinline__ R readlane_real(R x, int lane) {
  if constexpr (sizeof(R) == 8)
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), lane),
                            __builtin_amdgcn_readlane(__double2loint(x), lane));
  else
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}
template <typename T>
__device__ __forceinline__ T readlane_t(T x, int lane) {
  if constexpr (cplx_traits<T>::is)
    return T(readlane_real(x.re, lane), readlane_real(x.im, lane));
  else
    return readlane_real(x, lane);
}

// DPP move of a double (two 32-bit halves); lanes without a source read 0.
template <int CTRL, int ROWM = 0xF>
__device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __double2loint(x), hi = __double2hiint(x);
  const int l2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROWM, 0xF, false);
  const int h2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROWM, 0xF, false);
  return __hiloint2double(h2, l2);
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
  return __hiloint2double(hi, lo);
}

// Maximum of non-negative doubles over the 64 lanes: quad swaps, half-row and
// row mirrors, then row broadcasts 15 and 31 leave the maximum in lane 63.
__device__ __forceinline__ double wave_max_nonneg(double v) {
  v = fmax(v, dpp_f64<0xB1>(v));        // quad_perm [1,0,3,2]
  v = fmax(v, dpp_f64<0x4E>(v));        // quad_perm [2,3,0,1]
  v = fmax(v, dpp_f64<0x141>(v));       // row_half_mirror
  v = fmax(v, dpp_f64<0x140>(v));       // row_mirror
  v = fmax(v, dpp_f64<0x142, 0xA>(v));  // row_bcast:15 -> rows 1, 3
  v = fmax(v, dpp_f64<0x143, 0xC>(v));  // row_bcast:31 -> rows 2, 3
  return readlane_f64(v, 63);
}

// iamax over one wave: each lane holds (value, row) with value > 0 for a
// candidate and (0, INT_MAX) otherwise.  Returns the largest value and, among
// the lanes holding it, the lowest row (uniform across the wave).
__device__ __forceinline__ void wave_argmax(double v, int row, double &bv, int &br) {
  const double vmax = wave_max_nonneg(v);
  const unsigned long long m = __ballot(row != INT_MAX && v == vmax);
  int best = INT_MAX;
  unsigned long long mm = m;
  while (mm) {
    const int l = __ffsll((long long)mm) - 1;
    mm &= mm - 1;
    const int rl = __builtin_amdgcn_readlane(row, l);
    best = rl < best ? rl : best;
  }
  bv = m ? vmax : 0.0;
  br = best;
}

}  // namespace lair
