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

template <typename T>
__device__ __forceinline__ uint32_t word_of(T x, int part) {
  if constexpr (sizeof(T) == 8) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return part == 0 ? (uint32_t)(b >> 32) : (uint32_t)b;
  } else {
    return (uint32_t)__float_as_uint(x);
  }
}

template <typename T>
__device__ __forceinline__ T from_words(const uint32_t *w) {
  if constexpr (sizeof(T) == 8)
    return __longlong_as_double((long long)(((unsigned long long)w[0] << 32) | w[1]));
  else
    return __uint_as_float(w[0]);
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
