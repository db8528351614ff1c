// Synthetic input generator for the benchmark and the large tests (not part of
// the solver library; built as a separate liblair_gen.so).
//
//  * lair_gen_gaussian: i.i.d. N(0,1) entries on the device (SplitMix64 ->
//    53-bit uniforms -> Box-Muller), row-major with leading dimension.
//  * lair_gen_kappa: A = U diag(s) V^T with U, V Haar-distributed orthogonal
//    matrices (QR of a Gaussian matrix, columns of Q multiplied by the signs of
//    diag(R) -- Mezzadri 2007) and s_i = kappa^(i/(n-1)), so ||A||_2 = kappa
//    and cond_2(A) = kappa.  The QR factorizations and the product use
//    rocSOLVER/rocBLAS; they only build inputs and are never on the solver path.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <cstdint>

namespace {

__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void gaussian_kernel(double *a, int64_t rows, int64_t cols, int64_t ld, uint64_t seed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = rows * cols;
  const int64_t pair = 2 * i;
  if (pair >= total) return;
  const uint64_t h1 = splitmix(seed ^ splitmix((uint64_t)pair));
  const uint64_t h2 = splitmix(h1 + 0x632BE59BD9B4E019ull);
  const double u1 = ((h1 >> 11) + 1) * 0x1.0p-53;  // (0, 1]
  const double u2 = (h2 >> 11) * 0x1.0p-53;        // [0, 1)
  const double r = sqrt(-2.0 * log(u1));
  const double z0 = r * cospi(2.0 * u2), z1 = r * sinpi(2.0 * u2);
  a[(pair / cols) * ld + pair % cols] = z0;
  if (pair + 1 < total) a[((pair + 1) / cols) * ld + (pair + 1) % cols] = z1;
}

// Column-major n x n Q: column j scaled by sign(R_jj) * s_j.
__global__ void scale_columns(double *q, int64_t n, int64_t ld, const double *rdiag, double log_kappa,
                              int use_s) {
  const int64_t j = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || j >= n) return;
  double f = rdiag[j] < 0.0 ? -1.0 : 1.0;
  if (use_s) f *= exp(log_kappa * (double)j / (double)(n > 1 ? n - 1 : 1));
  q[j * ld + i] *= f;
}

__global__ void take_diag(const double *r, int64_t n, int64_t ld, double *d) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) d[j] = r[j * ld + j];
}

int haar_in_place(rocblas_handle h, double *g, int64_t n, int64_t ld, double *tau, double *rd,
                  double log_kappa, int use_s, hipStream_t st) {
  if (rocsolver_dgeqrf(h, (rocblas_int)n, (rocblas_int)n, g, (rocblas_int)ld, tau) !=
      rocblas_status_success)
    return 3;
  take_diag<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(g, n, ld, rd);
  if (rocsolver_dorgqr(h, (rocblas_int)n, (rocblas_int)n, (rocblas_int)n, g, (rocblas_int)ld,
                       tau) != rocblas_status_success)
    return 3;
  dim3 grid((unsigned)((n + 255) / 256), (unsigned)n);
  scale_columns<<<grid, 256, 0, st>>>(g, n, ld, rd, log_kappa, use_s);
  return hipGetLastError() == hipSuccess ? 0 : 3;
}

}  // namespace

// Plain 16-byte copy of `n16` 16-byte words: the known-size reference that
// calibrates HBM byte counters (FETCH_SIZE / WRITE_SIZE) against the bytes a
// kernel must move.
__global__ void copy16_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// Adds scale * n_i to the diagonal of every packed system (order ns[i] at
// element offset offs[i], row-major): one workgroup per system.
__global__ void batch_shift_kernel(double *a, const int *ns, const int64_t *offs, int count) {
  const int s = blockIdx.x;
  if (s >= count) return;
  const int n = ns[s];
  for (int i = threadIdx.x; i < n; i += blockDim.x) a[offs[s] + (int64_t)i * n + i] += (double)n;
}

__global__ void to_f32_kernel(const double *src, float *dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = (float)src[i];
}

__global__ void shift_diag_kernel(double *a, int64_t n, int64_t ld, double s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i * ld + i] += s;
}

extern "C" {

// Packed batch of `count` systems: a already holds N(0,1) entries; adds n_i on
// each system's diagonal (ns, offs: device arrays).
int lair_gen_batch_shift(double *a, const int *ns, const int64_t *offs, int count) {
  if (count <= 0) return 0;
  batch_shift_kernel<<<count, 256>>>(a, ns, offs, count);
  return hipDeviceSynchronize() == hipSuccess ? 0 : 3;
}

// dst[i] = (float) src[i], i < n (device pointers).
int lair_gen_to_f32(const double *src, float *dst, int64_t n) {
  if (n <= 0) return 0;
  to_f32_kernel<<<2048, 256>>>(src, dst, n);
  return hipDeviceSynchronize() == hipSuccess ? 0 : 3;
}

// a[i][i] += s for i < n (row-major, leading dimension ld).
int lair_gen_shift_diag(double *a, int64_t n, int64_t ld, double s) {
  if (n <= 0) return 0;
  shift_diag_kernel<<<(unsigned)((n + 255) / 256), 256>>>(a, n, ld, s);
  return hipDeviceSynchronize() == hipSuccess ? 0 : 3;
}

// dst <- src, `bytes` a multiple of 16 (device pointers, 16-byte aligned).
int lair_gen_copy(void *dst, const void *src, int64_t bytes) {
  if (bytes <= 0 || bytes % 16 != 0) return 1;
  copy16_kernel<<<2048, 256>>>(static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), bytes / 16);
  if (hipDeviceSynchronize() != hipSuccess) return 3;
  return 0;
}

int lair_gen_gaussian(double *a, int64_t rows, int64_t cols, int64_t ld, uint64_t seed) {
  const int64_t total = rows * cols;
  if (total <= 0) return 0;
  const int64_t pairs = (total + 1) / 2;
  gaussian_kernel<<<(unsigned)((pairs + 255) / 256), 256>>>(a, rows, cols, ld, seed);
  if (hipGetLastError() != hipSuccess) return 3;
  return hipDeviceSynchronize() == hipSuccess ? 0 : 3;
}

// a: n x n output (row-major view of the column-major product, leading
// dimension ld >= n).  Needs about 2 n^2 doubles of scratch beyond a.
int lair_gen_kappa(double *a, int64_t n, int64_t ld, double kappa, uint64_t seed) {
  if (n <= 0) return 0;
  double *u = nullptr, *v = nullptr, *tau = nullptr, *rd = nullptr;
  const size_t bytes = sizeof(double) * (size_t)n * (size_t)n;
  if (hipMalloc(&u, bytes) != hipSuccess) return 4;
  if (hipMalloc(&v, bytes) != hipSuccess) {
    hipFree(u);
    return 4;
  }
  hipMalloc(&tau, sizeof(double) * n);
  hipMalloc(&rd, sizeof(double) * n);
  rocblas_handle h;
  rocblas_create_handle(&h);
  hipStream_t st = 0;
  rocblas_set_stream(h, st);
  int rc = lair_gen_gaussian(u, n, n, n, seed);
  if (!rc) rc = lair_gen_gaussian(v, n, n, n, seed ^ 0x5DEECE66Dull);
  if (!rc) rc = haar_in_place(h, u, n, n, tau, rd, std::log(kappa), 1, st);  // U diag(s)
  if (!rc) rc = haar_in_place(h, v, n, n, tau, rd, 0.0, 0, st);              // V
  if (!rc) {
    const double one = 1.0, zero = 0.0;
    // column-major C = (U diag(s)) V^T written with leading dimension ld
    if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)n,
                      (rocblas_int)n, (rocblas_int)n, &one, u, (rocblas_int)n, v, (rocblas_int)n,
                      &zero, a, (rocblas_int)ld) != rocblas_status_success)
      rc = 3;
  }
  if (hipDeviceSynchronize() != hipSuccess) rc = 3;
  rocblas_destroy_handle(h);
  hipFree(u);
  hipFree(v);
  hipFree(tau);
  hipFree(rd);
  return rc;
}

}  // extern "C"
