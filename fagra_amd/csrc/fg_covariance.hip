// fg_covariance.hip -- selected covariance from undamped information on the device
// (SelectedCovariance, src/covariance.rs:163-262; DenseNormalCholesky::information_factor,
// src/normal.rs:341-392).
//
// The information H (n x n, column-major, lower triangle significant) is factored in place by
// a right-looking blocked Cholesky in the reference's coordinate order (64-column panels: the
// diagonal tile factored in LDS by one workgroup, the panel below solved row by row against
// that tile, the trailing lower triangle updated by 64 x 64 LDS-tiled f64 products). The
// reference's rank screen follows (pivot L_ii against sqrt(H_ii)), then Y = L^-1 E for the
// selected unit columns E by blocked forward substitution, and the covariance Y^T Y.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <mutex>
#include <vector>

#include "fg_internal.h"
#include "fg_kernels.h"

namespace {

using namespace fg;

constexpr int kNB = 64;  // panel width

#define CV_TRY(x)                                  \
  do {                                             \
    if ((x) != hipSuccess) return FG_DEVICE_ERROR; \
  } while (0)

// diagonal copy and finiteness of the lower triangle (normal.rs:350-358)
__global__ void k_cov_check(const double* __restrict__ h, size_t n, double* __restrict__ diag, int* flags) {
  const size_t total = n * n;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t col = idx / n, row = idx % n;
    if (row < col) continue;
    const double v = h[idx];
    if (!isfinite(v)) atomicOr(flags, 1);
    if (row == col) diag[col] = v;
  }
}

// Cholesky of the diagonal tile [k0, k0 + kb) by one wavefront, then the inverse of its factor
// into linv (64 x 64, row-major, identity-padded beyond kb) for the panel and selected-column
// solves. A nonpositive or nonfinite pivot sets flags[1] (faer's LLT error,
// SingularInformation). Lane i keeps row i of the tile in registers (both loops are unrolled,
// so every index is a compile-time one); the column being eliminated is broadcast through a
// 64-entry LDS vector. The inverse is formed column by column in the same way: lane c keeps
// column c of L^-1 and reads the rows of L from LDS, which nothing writes meanwhile, so those
// loads are in flight together. (Rows kept in LDS and updated in place there took 161 us per
// tile: each read-modify-write waited for the previous one.) Entries above the diagonal of a
// row are never read back: a lane at or above the column being eliminated only changes them.
__global__ void __launch_bounds__(kWave) k_cov_potrf(double* __restrict__ a, int n, int k0, int kb,
                                                     double* __restrict__ linv, int* flags) {
  __shared__ double t[kNB][kNB + 1];
  __shared__ double col[kNB];
  __shared__ int fail;
  if (flags[1]) return;
  const int i = threadIdx.x;
  double row[kNB];
#pragma unroll
  for (int c = 0; c < kNB; ++c) row[c] = (i < kb && c <= i) ? a[(size_t)(k0 + c) * n + k0 + i] : (i == c ? 1.0 : 0.0);
  if (i == 0) fail = 0;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kNB; ++j) {
    if (i == j) {
      const double d = row[j];
      if (!(d > 0.0) || !isfinite(d)) {
        fail = 1;
      } else {
        row[j] = sqrt(d);
        col[j] = row[j];
      }
    }
    __syncthreads();
    if (fail) break;
    if (i > j) {
      row[j] /= col[j];
      col[i] = row[j];
    }
    __syncthreads();
    const double f = row[j];  // L[i][j] below the diagonal; zero above it
#pragma unroll
    for (int l = j + 1; l < kNB; ++l) row[l] -= f * col[l];
    __syncthreads();
  }
  if (fail) {
    if (i == 0) flags[1] = 1;
    return;
  }
#pragma unroll
  for (int c = 0; c < kNB; ++c) {
    if (i < kb && c <= i) a[(size_t)(k0 + c) * n + k0 + i] = row[c];
    t[i][c] = c <= i ? row[c] : 0.0;
  }
  __syncthreads();
  // column i of L^-1 by forward substitution on e_i; rows above i come out zero
  double v[kNB];
#pragma unroll
  for (int r = 0; r < kNB; ++r) {
    double s[4] = {r == i ? 1.0 : 0.0, 0.0, 0.0, 0.0};  // four chains, added in a fixed order
#pragma unroll
    for (int q = 0; q < r; ++q) s[q & 3] -= t[r][q] * v[q];
    v[r] = ((s[0] + s[1]) + (s[2] + s[3])) / t[r][r];
  }
#pragma unroll
  for (int r = 0; r < kNB; ++r) linv[r * kNB + i] = v[r];
}

// Panel rows below the tile, X = A_panel L^-T as a product with the tile inverse: 64 rows per
// workgroup, each thread a 4 x 4 strided block (out[r][j] = sum_t A[r][t] Linv[j][t]).
__global__ void __launch_bounds__(kBlock) k_cov_trsm(double* __restrict__ a, int n, int k0, int kb,
                                                     const double* __restrict__ linv, const int* flags) {
  __shared__ double p[kNB][kNB + 1];
  __shared__ double l[kNB][kNB + 1];
  if (flags[1]) return;
  const int r0 = k0 + kb + kNB * blockIdx.x;
  for (int e = threadIdx.x; e < kNB * kNB; e += blockDim.x) {
    const int r = e % kNB, c = e / kNB;
    p[r][c] = (c < kb && r0 + r < n) ? a[(size_t)(k0 + c) * n + r0 + r] : 0.0;
    l[e / kNB][e % kNB] = linv[e];
  }
  __syncthreads();
  const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
  double acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int w = 0; w < 4; ++w) acc[u][w] = 0.0;
  for (int t = 0; t < kb; ++t) {
    double pv[4], lv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) pv[u] = p[ty + 16 * u][t];
#pragma unroll
    for (int w = 0; w < 4; ++w) lv[w] = l[tx + 16 * w][t];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int w = 0; w < 4; ++w) acc[u][w] += pv[u] * lv[w];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const int row = r0 + ty + 16 * u, col = tx + 16 * w;
      if (row < n && col < kb) a[(size_t)(k0 + col) * n + row] = acc[u][w];
    }
}

// Trailing update A[i, j] -= sum_t P[i, t] P[j, t] for m0 <= j <= i < n over 64 x 64 lower tiles;
// each thread owns a 4 x 4 strided block of its tile.
__global__ void __launch_bounds__(kBlock) k_cov_syrk(double* __restrict__ a, int n, int k0, int kb, int m0,
                                                     const int* flags) {
  __shared__ double p[kNB][kNB + 1];
  __shared__ double q[kNB][kNB + 1];
  if (flags[1]) return;
  const long long qi = blockIdx.x;
  int ti = (int)((sqrt(8.0 * (double)qi + 1.0) - 1.0) / 2.0);
  while ((long long)(ti + 1) * (ti + 2) / 2 <= qi) ++ti;
  while ((long long)ti * (ti + 1) / 2 > qi) --ti;
  const int tj = (int)(qi - (long long)ti * (ti + 1) / 2);
  const int r0 = m0 + kNB * ti, c0 = m0 + kNB * tj;
  for (int e = threadIdx.x; e < kNB * kNB; e += blockDim.x) {
    const int r = e % kNB, t = e / kNB;
    p[r][t] = (t < kb && r0 + r < n) ? a[(size_t)(k0 + t) * n + r0 + r] : 0.0;
    q[r][t] = (t < kb && c0 + r < n) ? a[(size_t)(k0 + t) * n + c0 + r] : 0.0;
  }
  __syncthreads();
  const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
  double acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
  for (int t = 0; t < kb; ++t) {
    double pv[4], qv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) pv[u] = p[ty + 16 * u][t];
#pragma unroll
    for (int v = 0; v < 4; ++v) qv[v] = q[tx + 16 * v][t];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] += pv[u] * qv[v];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = r0 + ty + 16 * u, col = c0 + tx + 16 * v;
      if (row < n && col < n && row >= col) a[(size_t)col * n + row] -= acc[u][v];
    }
}

// rank screen (normal.rs:380-390): the first failing coordinate decides, as in the reference's
// sequential loop; code 2 = nonfinite (InvalidInformation), 1 = SingularInformation
__global__ void k_cov_screen(const double* __restrict__ l, int n, const double* __restrict__ diag, double tol,
                             int* first) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double pivot = l[(size_t)i * n + i], d = diag[i];
  int code = 0;
  if (!isfinite(pivot) || !isfinite(d)) {
    code = 2;
  } else {
    const double ratio = pivot / sqrt(d);
    if (d <= 0.0 || pivot <= 0.0 || ratio * ratio <= tol) code = 1;
  }
  if (code) atomicMin(first, 4 * i + code);
}

// E: unit columns at the selected coordinates
__global__ void k_cov_unit(double* __restrict__ y, int n, int k, const int* __restrict__ cols) {
  const size_t total = (size_t)n * k;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
    const int col = (int)(idx / n), row = (int)(idx % n);
    y[idx] = row == cols[col] ? 1.0 : 0.0;
  }
}

// Y[k0 : k0 + kb, a] = L_kk^-1 Y[k0 : k0 + kb, a]: one workgroup of 64 per selected column
__global__ void __launch_bounds__(kWave) k_cov_fwd_tile(const double* __restrict__ linv, int n, int k0, int kb,
                                                        double* __restrict__ y) {
  __shared__ double yt[kNB];
  double* yc = y + (size_t)blockIdx.x * n;
  const int j = threadIdx.x;
  yt[j] = j < kb ? yc[k0 + j] : 0.0;
  __syncthreads();
  double s = 0.0;
  for (int t = 0; t <= j; ++t) s += linv[j * kNB + t] * yt[t];
  if (j < kb) yc[k0 + j] = s;
}

// Y[i, a] -= sum_t L[i, k0 + t] Y[k0 + t, a] for rows below the tile
__global__ void __launch_bounds__(kBlock) k_cov_fwd_update(const double* __restrict__ a, int n, int k0, int kb,
                                                           double* __restrict__ y, int k) {
  __shared__ double yt[kNB];
  const int c = blockIdx.y;
  double* yc = y + (size_t)c * n;
  if (threadIdx.x < kNB) yt[threadIdx.x] = threadIdx.x < kb ? yc[k0 + threadIdx.x] : 0.0;
  __syncthreads();
  const int i = k0 + kb + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = yc[i];
  for (int t = 0; t < kb; ++t) v -= a[(size_t)(k0 + t) * n + i] * yt[t];
  yc[i] = v;
}

// C[a, b] = Y[:, a] . Y[:, b] for b >= a (one workgroup each), mirrored; nonfinite -> flag
__global__ void __launch_bounds__(kBlock) k_cov_gram(const double* __restrict__ y, int n, int k,
                                                     double* __restrict__ c, int* flags) {
  __shared__ double smem[kBlock / kWave];
  const int qi = blockIdx.x;
  int b = (int)((sqrt(8.0 * (double)qi + 1.0) - 1.0) / 2.0);
  while ((b + 1) * (b + 2) / 2 <= qi) ++b;
  while (b * (b + 1) / 2 > qi) --b;
  const int a2 = qi - b * (b + 1) / 2;  // a2 <= b
  const double* ya = y + (size_t)a2 * n;
  const double* yb = y + (size_t)b * n;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += ya[i] * yb[i];
  s = block_sum(s, smem);
  if (threadIdx.x == 0) {
    if (!isfinite(s)) atomicOr(flags, 4);
    c[(size_t)b * k + a2] = s;
    c[(size_t)a2 * k + b] = s;
  }
}

int grid_of(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + kBlock - 1) / kBlock, 65535)); }

}  // namespace

namespace {

// Device workspace of one covariance: the diagonal of H, the flags, Y (n x k), the k x k
// covariance, the selected columns and the inverse of every panel's diagonal factor tile. Each
// part starts on a 256-byte boundary.
struct CovWork {
  double* diag;
  int* flags;
  double* y;
  double* c;
  int* cols;
  double* linv;
};

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t cov_work_bytes(size_t n, size_t k) {
  return align256(n * sizeof(double)) + align256(4 * sizeof(int)) + align256((n * k > 0 ? n * k : 1) * sizeof(double)) +
         align256((k * k > 0 ? k * k : 1) * sizeof(double)) + align256((k > 0 ? k : 1) * sizeof(int)) +
         align256(((n + kNB - 1) / kNB) * kNB * kNB * sizeof(double));
}
CovWork cov_work_carve(unsigned char* base, size_t n, size_t k) {
  CovWork w;
  w.diag = (double*)base;
  base += align256(n * sizeof(double));
  w.flags = (int*)base;
  base += align256(4 * sizeof(int));
  w.y = (double*)base;
  base += align256((n * k > 0 ? n * k : 1) * sizeof(double));
  w.c = (double*)base;
  base += align256((k * k > 0 ? k * k : 1) * sizeof(double));
  w.cols = (int*)base;
  base += align256((k > 0 ? k : 1) * sizeof(int));
  w.linv = (double*)base;
  return w;
}

// The factorization, rank screen and selected solves of H (device, n x n, overwritten by its
// factor) in the workspace w, on stream s; host_cols are the k selected coordinates (checked by
// the caller). Allocates nothing.
fg_status dense_covariance_ws(hipStream_t s, double* h, size_t n, const int* host_cols, size_t k, double tolerance,
                              double* out, const CovWork& w) {
  double *diag = w.diag, *y = w.y, *c = w.c, *linv = w.linv;
  int *flags = w.flags, *cols = w.cols;
  auto fail = [&](fg_status st) {
    (void)hipStreamSynchronize(s);
    return st;
  };
  const int ni = (int)n, ki = (int)k;
  int init[4] = {0, 0, INT_MAX, 0};
  if (hipMemcpyAsync(flags, init, sizeof(init), hipMemcpyHostToDevice, s) != hipSuccess ||
      (k && hipMemcpyAsync(cols, host_cols, k * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess))
    return fail(FG_DEVICE_ERROR);
  hipLaunchKernelGGL(k_cov_check, dim3(grid_of(n * n)), dim3(kBlock), 0, s, (const double*)h, n, diag, flags);
  int hf[4];
  if (hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(FG_DEVICE_ERROR);
  if (hf[0]) return fail(FG_INVALID_INFORMATION);  // nonfinite information
  for (int k0 = 0; k0 < ni; k0 += kNB) {
    const int kb = std::min(kNB, ni - k0), m0 = k0 + kb, rest = ni - m0;
    double* li = linv + (size_t)(k0 / kNB) * kNB * kNB;
    hipLaunchKernelGGL(k_cov_potrf, dim3(1), dim3(kWave), 0, s, h, ni, k0, kb, li, flags);
    if (rest > 0) {
      hipLaunchKernelGGL(k_cov_trsm, dim3((rest + kNB - 1) / kNB), dim3(kBlock), 0, s, h, ni, k0, kb,
                         (const double*)li, (const int*)flags);
      const long long m = (rest + kNB - 1) / kNB;
      hipLaunchKernelGGL(k_cov_syrk, dim3((unsigned)(m * (m + 1) / 2)), dim3(kBlock), 0, s, h, ni, k0, kb, m0,
                         (const int*)flags);
    }
  }
  hipLaunchKernelGGL(k_cov_screen, dim3(grid_of(n)), dim3(kBlock), 0, s, (const double*)h, ni, (const double*)diag,
                     tolerance, flags + 2);
  if (hipGetLastError() != hipSuccess) return fail(FG_DEVICE_ERROR);
  if (hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(FG_DEVICE_ERROR);
  if (hf[1]) return FG_SINGULAR_INFORMATION;  // the factorization failed
  if (hf[2] != INT_MAX) return (hf[2] & 3) == 2 ? FG_INVALID_INFORMATION : FG_SINGULAR_INFORMATION;
  if (k == 0) return FG_OK;
  hipLaunchKernelGGL(k_cov_unit, dim3(grid_of(n * k)), dim3(kBlock), 0, s, y, ni, ki, (const int*)cols);
  for (int k0 = 0; k0 < ni; k0 += kNB) {
    const int kb = std::min(kNB, ni - k0), rest = ni - k0 - kb;
    hipLaunchKernelGGL(k_cov_fwd_tile, dim3(ki), dim3(kWave), 0, s,
                       (const double*)(linv + (size_t)(k0 / kNB) * kNB * kNB), ni, k0, kb, y);
    if (rest > 0)
      hipLaunchKernelGGL(k_cov_fwd_update, dim3(grid_of(rest), ki), dim3(kBlock), 0, s, (const double*)h, ni, k0,
                         kb, y, ki);
  }
  hipLaunchKernelGGL(k_cov_gram, dim3((unsigned)(k * (k + 1) / 2)), dim3(kBlock), 0, s, (const double*)y, ni, ki, c,
                     flags);
  if (hipGetLastError() != hipSuccess) return fail(FG_DEVICE_ERROR);
  if (hipMemcpyAsync(out, c, k * k * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(FG_DEVICE_ERROR);
  return (hf[0] & 4) ? FG_INVALID_INFORMATION : FG_OK;
}

// Checks shared by the covariance entry points (src/covariance.rs:136-160, 214-230).
fg_status check_selection(size_t n, const size_t* columns, size_t k, double tolerance) {
  if (!std::isfinite(tolerance) || tolerance < 0.0 || tolerance >= 1.0) return FG_INVALID_RANK_TOLERANCE;
  for (size_t a = 0; a < k; ++a)
    if (columns[a] >= n) return FG_UNKNOWN_KEY;
  if (n > (size_t)INT_MAX / 2 || k > (size_t)INT_MAX / 2 || (k && n > SIZE_MAX / 8 / k)) return FG_INVALID_INFORMATION;
  return FG_OK;
}

// Largest problem of fg_dense_covariance_noalloc, and its scratch: the n x n staging copy of the
// information followed by the workspace of an (n, k) covariance.
constexpr size_t kNoAllocMaxN = 2048, kNoAllocMaxK = 256;
constexpr size_t kScratchBytes = align256(kNoAllocMaxN * kNoAllocMaxN * sizeof(double)) +
                                 cov_work_bytes(kNoAllocMaxN, kNoAllocMaxK);
constexpr int kMaxDevices = 64;

}  // namespace

// Device memory of fg_dense_covariance_noalloc, one copy per GPU for the life of the process.
// Host code never uses this name as an address: it reads each device's address once with
// hipGetSymbolAddress and passes that pointer.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of scratch, looked up on the first call on that device.
fg_status scratch_of_current_device(unsigned char** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FG_DEVICE_ERROR;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FG_DEVICE_ERROR;
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return FG_OK;
}

}  // namespace

namespace fg {

fg_status dense_covariance(hipStream_t s, double* h, size_t n, const size_t* columns, size_t k, double tolerance,
                           double* out) {
  fg_status st = check_selection(n, columns, k, tolerance);
  if (st != FG_OK) return st;
  if (n == 0) return FG_OK;  // k == 0 here: the empty selection of an empty model
  void* base = nullptr;
  if (hipMalloc(&base, cov_work_bytes(n, k)) != hipSuccess) return FG_DEVICE_ERROR;
  std::vector<int> hc(columns, columns + k);
  st = dense_covariance_ws(s, h, n, hc.data(), k, tolerance, out, cov_work_carve((unsigned char*)base, n, k));
  (void)hipStreamSynchronize(s);
  (void)hipFree(base);
  return st;
}

}  // namespace fg

extern "C" {

fg_status fg_covariance_options_default(fg_covariance_options* o) {
  o->relative_pivot_tolerance = 64.0 * 2.220446049250313e-16;  // 64 epsilon (covariance.rs:145-151)
  return FG_OK;
}

fg_status fg_dense_covariance_noalloc(const double* information, size_t n, const size_t* columns, size_t k,
                                      const fg_covariance_options* options, double* out) {
  // the checks of fg_dense_covariance, in its order
  if (!std::isfinite(options->relative_pivot_tolerance) || options->relative_pivot_tolerance < 0.0 ||
      options->relative_pivot_tolerance >= 1.0)
    return FG_INVALID_RANK_TOLERANCE;
  for (size_t a = 0; a < k; ++a) {
    if (columns[a] >= n) return FG_UNKNOWN_KEY;
    for (size_t b = 0; b < a; ++b)
      if (columns[b] == columns[a]) return FG_DUPLICATE_COVARIANCE_BLOCK;
  }
  fg_status st = check_selection(n, columns, k, options->relative_pivot_tolerance);
  if (st != FG_OK) return st;
  if (n > kNoAllocMaxN || k > kNoAllocMaxK) return FG_INVALID_OPTIONS;  // beyond the scratch
  if (n == 0) return FG_OK;
  unsigned char* base = nullptr;
  st = scratch_of_current_device(&base);
  if (st != FG_OK) return st;
  hipStream_t s = nullptr;  // the null stream: no stream is created either
  double* h = (double*)base;
  const CovWork w = cov_work_carve(base + align256(kNoAllocMaxN * kNoAllocMaxN * sizeof(double)), n, k);
  int hc[kNoAllocMaxK];
  for (size_t a = 0; a < k; ++a) hc[a] = (int)columns[a];
  if (hipMemcpyAsync(h, information, n * n * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)
    return FG_DEVICE_ERROR;
  st = dense_covariance_ws(s, h, n, hc, k, options->relative_pivot_tolerance, out, w);
  if (hipStreamSynchronize(s) != hipSuccess) return FG_DEVICE_ERROR;
  return st;
}

fg_status fg_dense_covariance(const double* information, size_t n, const size_t* columns, size_t k,
                              const fg_covariance_options* options, double* out) {
  if (!std::isfinite(options->relative_pivot_tolerance) || options->relative_pivot_tolerance < 0.0 ||
      options->relative_pivot_tolerance >= 1.0)
    return FG_INVALID_RANK_TOLERANCE;
  for (size_t a = 0; a < k; ++a) {
    if (columns[a] >= n) return FG_UNKNOWN_KEY;
    for (size_t b = 0; b < a; ++b)
      if (columns[b] == columns[a]) return FG_DUPLICATE_COVARIANCE_BLOCK;
  }
  if (n == 0) return FG_OK;
  hipStream_t s = nullptr;
  double* h = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return FG_DEVICE_ERROR;
  fg_status st = FG_DEVICE_ERROR;
  if (hipMalloc(&h, n * n * sizeof(double)) == hipSuccess &&
      hipMemcpyAsync(h, information, n * n * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess)
    st = fg::dense_covariance(s, h, n, columns, k, options->relative_pivot_tolerance, out);
  (void)hipStreamSynchronize(s);
  if (h) (void)hipFree(h);
  (void)hipStreamDestroy(s);
  return st;
}

}  // extern "C"
