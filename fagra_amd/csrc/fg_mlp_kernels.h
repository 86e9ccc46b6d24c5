// fg_mlp_kernels.h -- forward, backward and Adam kernels of the preconditioner network, plus
// the per-camera feature and training-target kernels. Batches are padded to a multiple of 64
// rows (one 16-row tile per wave, four waves per workgroup); padded rows are zero.
#pragma once
#include "fg_kernels.h"
#include "fg_mlp.h"

namespace fg {

// Upper-triangle (row-major, i <= j) position k of a 9x9 block -> (i, j).
__host__ __device__ inline void upper_index(int k, int& i, int& j) {
  int r = 0, start = 0;
  while (k >= start + (9 - r)) {
    start += 9 - r;
    ++r;
  }
  i = r;
  j = r + (k - start);
}

// Feature k of camera c's block: for the 45 upper-triangle entries of the unnormalized Gram
// block, log(H_ii) on the diagonal and the correlation H_ij / sqrt(H_ii H_jj) off it; then
// log(observations + 1) and log10(lambda), all scaled to O(1); 0 for the padding k >= 47.
// The pieces, shared by camera_feature and the fused preconditioner block:
__device__ __forceinline__ float diagonal_log_feature(double hii) { return float(log(fmax(hii, 1e-30)) / 8.0); }
__device__ __forceinline__ float correlation_feature(double hii, double hjj, double hij) {
  const double d = sqrt(hii * hjj);
  return d > 0.0 ? float(hij / d) : 0.f;
}
__device__ __forceinline__ float count_feature(int n) { return float(log(double(n) + 1.0) / 8.0); }
__device__ __forceinline__ float damping_feature(double lambda) { return float(log10(fmax(lambda, 1e-30)) / 8.0); }

template <typename T>
__device__ __forceinline__ float camera_feature(int c, int k, const int* __restrict__ cam_off,
                                                const T* __restrict__ hcc, double lambda) {
  const T* h = hcc + 45 * c;
  float f = 0.f;
  if (k < 45) {
    int i, j;
    upper_index(k, i, j);
    const double hii = double(h[sym9(i, i)]), hjj = double(h[sym9(j, j)]);
    if (i == j) {
      f = diagonal_log_feature(hii);
    } else {
      f = correlation_feature(hii, hjj, double(h[sym9(i, j)]));
    }
  } else if (k == 45) {
    f = count_feature(cam_off[c + 1] - cam_off[c]);
  } else if (k == 46) {
    f = damping_feature(lambda);
  }
  return f;
}

// The features of every camera as padded bf16 rows (and f32 rows in xf when non-null). One lane
// per (camera, feature).
template <typename T>
__global__ void k_mlp_features(int nc, const int* __restrict__ cam_off, const T* __restrict__ hcc, double lambda,
                               uint16_t* __restrict__ x, float* __restrict__ xf) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = gid / kInPad, k = gid % kInPad;
  if (c >= nc) return;
  const float f = camera_feature<T>(c, k, cam_off, hcc, lambda);
  x[(size_t)kInPad * c + k] = f2bf(f);
  if (xf && k < kIn) xf[(size_t)kIn * c + k] = f;
}

// Fused three-layer forward for 16 rows per wave, one to four waves per workgroup (small
// inference batches launch one-wave workgroups to spread over more CUs). Z1/H1/Z2/H2 are
// stored when non-null (training); Y receives the f32 output rows (kOutPad wide).
static __global__ void __launch_bounds__(256) k_mlp_forward(const uint16_t* __restrict__ x, const uint16_t* __restrict__ w1,
                                                     const float* __restrict__ b1, const uint16_t* __restrict__ w2,
                                                     const float* __restrict__ b2, const uint16_t* __restrict__ w3,
                                                     const float* __restrict__ b3, float* __restrict__ z1,
                                                     uint16_t* __restrict__ h1, float* __restrict__ z2,
                                                     uint16_t* __restrict__ h2, float* __restrict__ y) {
  __shared__ uint16_t stage[2][4][16][kHidden];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  const size_t m0 = ((size_t)blockIdx.x * (blockDim.x >> 6) + wave) * 16;
  // layer 1
  for (int nt = 0; nt < kHidden / 16; ++nt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(x + m0 * kInPad, kInPad, w1 + (size_t)nt * 16 * kInPad, kInPad, kInPad, acc);
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * q + i, col = nt * 16 + r;
      const float z = acc[i] + b1[col];
      const uint16_t hb = f2bf(gelu(z));
      stage[0][wave][row][col] = hb;
      if (z1) z1[(m0 + row) * kHidden + col] = z;
      if (h1) h1[(m0 + row) * kHidden + col] = hb;
    }
  }
  __syncthreads();
  // layer 2
  for (int nt = 0; nt < kHidden / 16; ++nt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(&stage[0][wave][0][0], kHidden, w2 + (size_t)nt * 16 * kHidden, kHidden, kHidden, acc);
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * q + i, col = nt * 16 + r;
      const float z = acc[i] + b2[col];
      const uint16_t hb = f2bf(gelu(z));
      stage[1][wave][row][col] = hb;
      if (z2) z2[(m0 + row) * kHidden + col] = z;
      if (h2) h2[(m0 + row) * kHidden + col] = hb;
    }
  }
  __syncthreads();
  // layer 3 (linear)
  for (int nt = 0; nt < kOutPad / 16; ++nt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(&stage[1][wave][0][0], kHidden, w3 + (size_t)nt * 16 * kHidden, kHidden, kHidden, acc);
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * q + i, col = nt * 16 + r;
      y[(m0 + row) * kOutPad + col] = acc[i] + b3[col];
    }
  }
}

// Inference forward (no stored activations): one 16-row block per four-wave workgroup, the
// waves splitting each layer's 16-column output tiles and exchanging the activations through
// LDS. Every tile is the same MFMA chain over the same operands as in k_mlp_forward, so the
// output rows are the same; the chain per wave is a quarter as long.
static __global__ void __launch_bounds__(256) k_mlp_infer(const uint16_t* __restrict__ x, const uint16_t* __restrict__ w1,
                                                   const float* __restrict__ b1, const uint16_t* __restrict__ w2,
                                                   const float* __restrict__ b2, const uint16_t* __restrict__ w3,
                                                   const float* __restrict__ b3, float* __restrict__ y) {
  __shared__ uint16_t h1[16][kHidden];
  __shared__ uint16_t h2[16][kHidden];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  const int waves = blockDim.x >> 6;
  const size_t m0 = (size_t)blockIdx.x * 16;
  for (int nt = wave; nt < kHidden / 16; nt += waves) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(x + m0 * kInPad, kInPad, w1 + (size_t)nt * 16 * kInPad, kInPad, kInPad, acc);
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * q + i, col = nt * 16 + r;
      h1[row][col] = f2bf(gelu(acc[i] + b1[col]));
    }
  }
  __syncthreads();
  for (int nt = wave; nt < kHidden / 16; nt += waves) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(&h1[0][0], kHidden, w2 + (size_t)nt * 16 * kHidden, kHidden, kHidden, acc);
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * q + i, col = nt * 16 + r;
      h2[row][col] = f2bf(gelu(acc[i] + b2[col]));
    }
  }
  __syncthreads();
  for (int nt = wave; nt < kOutPad / 16; nt += waves) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(&h2[0][0], kHidden, w3 + (size_t)nt * 16 * kHidden, kHidden, kHidden, acc);
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * q + i, col = nt * 16 + r;
      y[(m0 + row) * kOutPad + col] = acc[i] + b3[col];
    }
  }
}

// G = I + s tril(y): packed lower column-major 9x9 per camera.
__device__ __forceinline__ void build_g(const float* yrow, float g[45]) {
  for (int k = 0; k < 45; ++k) g[k] = kOutScale * yrow[k];
  for (int j = 0; j < 9; ++j) g[sym_index(9, j, j)] += 1.f;
}

// The learned inverse per observed camera, M = L^-T G G^T L^-1 (packed symmetric, f64), from
// the network output y and the block-Jacobi factor L of B + lambda I, in two steps: column j of
// W = L^-T G (a triangular solve), then entry e of M = W W^T.
// With range (device [2][kIn], lower then upper bounds) a camera whose features xf (f32, kIn)
// leave the range the network was trained on keeps G = I, i.e. block-Jacobi.
__device__ __forceinline__ void inverse_column(int j, const float* yr, const double* __restrict__ l45,
                                               const float* xf, const float* __restrict__ range, double col[9]) {
  bool inside = true;
  if (range)
    for (int k = 0; k < kIn; ++k) {
      const float f = xf[k];
      inside = inside && f >= range[k] && f <= range[kIn + k];
    }
  double l[45];
  for (int i = 0; i < 45; ++i) l[i] = l45[i];
  for (int i = 0; i < 9; ++i)  // column j of G = I + s tril(y)
    col[i] = i < j ? 0.0 : (inside ? double(kOutScale * yr[sym_index(9, i, j)]) : 0.0) + (i == j ? 1.0 : 0.0);
  solve_upper_packed<double, 9>(l, col);
}
// Entry e (packed lower, column-major) of M = W W^T from W (9x9, w[i][t]).
__device__ __forceinline__ void inverse_entry(int e, const double (*w)[9], double* __restrict__ m45) {
  int j = 0, k = e;
  while (k >= 9 - j) {  // packed lower, column-major: (i, j) with i >= j
    k -= 9 - j;
    ++j;
  }
  const int i = j + k;
  double s = 0.0;
  for (int t = 0; t < 9; ++t) s += w[i][t] * w[j][t];
  m45[sym9(i, j)] = s;
}

// One 64-lane group per camera: lanes 0-8 form the columns of W, lanes 0-44 one entry of M each.
static __global__ void __launch_bounds__(256) k_mlp_build_g(int nc, const int* __restrict__ cam_off,
                                                            const float* __restrict__ y,
                                                            const double* __restrict__ lc,
                                                            const float* __restrict__ xf,
                                                            const float* __restrict__ range,
                                                            double* __restrict__ mout) {
  __shared__ double w[4][9][9];
  const int grp = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + grp;
  const bool active = c < nc && cam_off[c + 1] > cam_off[c];
  if (active && lane < 9) {
    double col[9];
    inverse_column(lane, y + (size_t)kOutPad * c, lc + 45 * c, range ? xf + (size_t)kIn * c : nullptr, range, col);
    for (int i = 0; i < 9; ++i) w[grp][i][lane] = col[i];
  }
  __syncthreads();
  if (active && lane < 45) inverse_entry(lane, w[grp], mout + 45 * c);
}

// The whole learned preconditioner of 16 cameras in one workgroup (features, the network as
// k_mlp_infer computes it, and M as k_mlp_build_g), every intermediate in LDS: no round trips
// through memory between them. The same arithmetic on the same operands as the three kernels,
// so M is the same. Every thread of the workgroup must call it (it has workgroup barriers).
struct PrecondArgs {
  int nc;
  const int* cam_off;
  const void* hcc;  // T [nc][45]
  double lambda;
  const uint16_t *w1, *w2, *w3;
  const float *b1, *b2, *b3;
  const void* bs;  // T [nc][45]: the scaled camera blocks B, factored here as k_factor_cameras does
  const float* range;
  double* mout;
};
template <typename T>
__device__ __forceinline__ void mlp_precond_block(int blk, const PrecondArgs& pa) {
  const int nc = pa.nc;
  const int* __restrict__ cam_off = pa.cam_off;
  const T* __restrict__ hcc = (const T*)pa.hcc;
  const double lambda = pa.lambda;
  const uint16_t *__restrict__ w1 = pa.w1, *__restrict__ w2 = pa.w2, *__restrict__ w3 = pa.w3;
  const float *__restrict__ b1 = pa.b1, *__restrict__ b2 = pa.b2, *__restrict__ b3 = pa.b3;
  const T* __restrict__ bs = (const T*)pa.bs;
  const float* __restrict__ range = pa.range;
  double* __restrict__ mout = pa.mout;
  // The stages' LDS arrays alias where their lifetimes do not overlap (xs: features and layer 1,
  // h1: layers 1-2, h2: layers 2-3, w: the inverse stage after layer 3; the features' Gram
  // blocks and logarithms share the space of h1 and h2): 22.6 KB.
  struct Pool {
    union {
      struct {
        uint16_t xs[16][kInPad];
        union {
          struct {
            uint16_t h1[16][kHidden];
            uint16_t h2[16][kHidden];
          } l;  // layers 1-3
          struct {
            T hs[16][45];        // the cameras' Gram blocks
            float logs[16][11];  // per camera: the 9 diagonal features, the count and the damping
            int offs[17];        // camera offsets c0 .. c0 + 16
          } f;  // features
        } v;
      } a;
      double w[16][9][9];  // inverse stage
    } u;
    float xfs[16][kIn];
    float ys[16][kOutPad];
    double lcs[16][45];
    float rng[2 * kIn];
  };
  __shared__ Pool pool;
  auto& xs = pool.u.a.xs;
  auto& h1 = pool.u.a.v.l.h1;
  auto& h2 = pool.u.a.v.l.h2;
  auto& hs = pool.u.a.v.f.hs;
  auto& logs = pool.u.a.v.f.logs;
  auto& offs = pool.u.a.v.f.offs;
  auto& w = pool.u.w;
  auto& xfs = pool.xfs;
  auto& ys = pool.ys;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  const int c0 = blk * 16;
  // Everything the block reads from memory is loaded here, in one round trip: the 16 cameras'
  // Gram blocks (contiguous, 45 each: at most 3 per thread), their offsets, the feature range
  // and the biases of this lane's output columns; the first 16 threads factor one camera block
  // B + lambda I each (camera_chol, as k_factor_cameras does, so the block needs no other
  // kernel's factors); the layers' weights are touched below.
  const int ncb = max(0, min(16, nc - c0));
  const int nl = 45 * ncb;
  T hreg[3];
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const int e = (int)threadIdx.x + n * (int)blockDim.x;
    hreg[n] = e < nl ? hcc[(size_t)45 * c0 + e] : T(0);
  }
  const int oreg = (int)threadIdx.x <= ncb ? cam_off[c0 + threadIdx.x] : 0;
  const float rreg = range && threadIdx.x < 2 * kIn ? range[threadIdx.x] : 0.f;
  float bias1[2], bias2[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    bias1[u] = b1[(wave + 4 * u) * 16 + r];
    bias2[u] = b2[(wave + 4 * u) * 16 + r];
  }
  const float bias3 = wave < kOutPad / 16 ? b3[wave * 16 + r] : 0.f;
  double fac[45];
  bool fok = false;
  if ((int)threadIdx.x < ncb) fok = camera_chol<T>(bs + 45 * (size_t)(c0 + threadIdx.x), lambda, fac);
  // The three layers' weights (60 KB, read by every workgroup) are touched once per 128-byte
  // line now, so the layers find them in L2 instead of waiting on memory behind the point
  // factorization's traffic; the loaded words meet in pf, which an empty asm consumes at the end.
  constexpr int kW1 = kHidden * kInPad, kW2 = kHidden * kHidden, kW3 = kOutPad * kHidden;  // bf16
  constexpr int kLines = (kW1 + kW2 + kW3) / 64;
  const __attribute__((address_space(1))) uint32_t* wl[3] = {
      (const __attribute__((address_space(1))) uint32_t*)w1, (const __attribute__((address_space(1))) uint32_t*)w2,
      (const __attribute__((address_space(1))) uint32_t*)w3};
  __asm__ volatile("" : "+s"(wl[0]), "+s"(wl[1]), "+s"(wl[2]));
  uint32_t pf = 0;
#pragma unroll
  for (int n = 0; n < (kLines + 255) / 256; ++n) {
    const int line = (int)threadIdx.x + 256 * n;
    if (line < kLines) {
      const int b16 = 64 * line;  // bf16 index of the line's first word
      pf ^= b16 < kW1 ? wl[0][b16 / 2] : b16 < kW1 + kW2 ? wl[1][(b16 - kW1) / 2] : wl[2][(b16 - kW1 - kW2) / 2];
    }
  }
  __asm__ volatile("" ::: "memory");
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const int e = (int)threadIdx.x + n * (int)blockDim.x;
    if (e < 16 * 45) (&hs[0][0])[e] = hreg[n];
  }
  if ((int)threadIdx.x < ncb) {
#pragma unroll
    for (int i = 0; i < 45; ++i) pool.lcs[threadIdx.x][i] = fok ? fac[i] : 0.0;
  }
  if ((int)threadIdx.x <= ncb) offs[threadIdx.x] = oreg;
  if (range && threadIdx.x < 2 * kIn) pool.rng[threadIdx.x] = rreg;
  __syncthreads();
  // the logarithms once per camera (camera_feature's pieces), then every feature from them and
  // the off-diagonal correlations
  if (threadIdx.x < 16 * 11) {
    const int cl = threadIdx.x / 11, m = threadIdx.x % 11;
    float f = 0.f;
    if (cl < ncb)
      f = m < 9 ? diagonal_log_feature(double(hs[cl][sym9(m, m)]))
                : m == 9 ? count_feature(offs[cl + 1] - offs[cl]) : damping_feature(lambda);
    logs[cl][m] = f;
  }
  __syncthreads();
  static_assert(16 * kInPad % 256 == 0, "features per thread");
#pragma unroll
  for (int n = 0; n < 16 * kInPad / 256; ++n) {
    const int e = (int)threadIdx.x + 256 * n;
    const int cl = e / kInPad, k = e % kInPad, c = c0 + cl;
    float f = 0.f;
    if (c < nc && k < kIn) {
      if (k >= 45) {
        f = logs[cl][9 + (k - 45)];
      } else {
        int i, j;
        upper_index(k, i, j);
        f = i == j ? logs[cl][i]
                   : correlation_feature(double(hs[cl][sym9(i, i)]), double(hs[cl][sym9(j, j)]),
                                         double(hs[cl][sym9(i, j)]));
      }
    }
    xs[cl][k] = f2bf(f);
    if (k < kIn) xfs[cl][k] = f;
  }
  __syncthreads();
  // each wave's two output tiles of a hidden layer unrolled, so both tiles' weight loads issue
  // together
  static_assert(kHidden / 16 == 8, "two hidden tiles per wave");
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int nt = wave + 4 * u;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(&xs[0][0], kInPad, w1 + (size_t)nt * 16 * kInPad, kInPad, kInPad, acc);
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * q + i, col = nt * 16 + r;
      h1[row][col] = f2bf(gelu(acc[i] + bias1[u]));
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int nt = wave + 4 * u;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(&h1[0][0], kHidden, w2 + (size_t)nt * 16 * kHidden, kHidden, kHidden, acc);
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * q + i, col = nt * 16 + r;
      h2[row][col] = f2bf(gelu(acc[i] + bias2[u]));
    }
  }
  __syncthreads();
  static_assert(kOutPad / 16 <= 4, "one output tile per wave");
  if (wave < kOutPad / 16) {
    const int nt = wave;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(&h2[0][0], kHidden, w3 + (size_t)nt * 16 * kHidden, kHidden, kHidden, acc);
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * q + i, col = nt * 16 + r;
      ys[row][col] = acc[i] + bias3;
    }
  }
  __syncthreads();
  // 16 threads per camera: 9 of them form the columns of W (as inverse_column does, the factor
  // and the range read from LDS), then all 16 the 45 entries of M
  const int cl = threadIdx.x >> 4, t = threadIdx.x & 15, c = c0 + cl;
  const bool active = c < nc && cam_off[c + 1] > cam_off[c];
  // the camera's 16 lanes test three features each against the training range, then combine
  int inside = 1;
  if (range)
    for (int k = t; k < kIn; k += 16) {
      const float f = xfs[cl][k];
      inside &= (f >= pool.rng[k]) & (f <= pool.rng[kIn + k]);
    }
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) inside &= __shfl_xor(inside, o, 16);
  if (active && t < 9) {
    double col[9];
    for (int i = 0; i < 9; ++i)  // column t of G = I + s tril(y)
      col[i] = i < t ? 0.0 : (inside ? double(kOutScale * ys[cl][sym_index(9, i, t)]) : 0.0) + (i == t ? 1.0 : 0.0);
    solve_upper_packed<double, 9>(&pool.lcs[cl][0], col);
    for (int i = 0; i < 9; ++i) w[cl][i][t] = col[i];
  }
  __syncthreads();
  if (active)
    for (int e = t; e < 45; e += 16) inverse_entry(e, w[cl], mout + 45 * c);
  __asm__ volatile("" ::"v"(pf));
}

template <typename T>
static __global__ void __launch_bounds__(256) k_mlp_precond(PrecondArgs pa) {
  mlp_precond_block<T>(blockIdx.x, pa);
}

// The attempt's point factorization (k_factor_points) and the learned preconditioner in one
// launch: workgroups 0 .. learned_blocks - 1 form M for 16 cameras each (mlp_precond_block, which
// factors its cameras' blocks itself, so it needs nothing the camera factorization forms), the
// rest factor 256 points each. The network's few latency-bound workgroups run beside the point
// factorization's instead of after it; their one memory round trip waits behind its traffic.
template <typename T>
static __global__ void __launch_bounds__(256) k_factor_points_precond(
    int learned_blocks, PrecondArgs pa, int np, const int* __restrict__ pt_len, const double* __restrict__ cs,
    double lambda, double* __restrict__ lp, int* fail_count, const T* __restrict__ gp, const T* __restrict__ dpv,
    double* __restrict__ t_out, const double* __restrict__ hpp, double* __restrict__ lpp) {
  if ((int)blockIdx.x < learned_blocks) {  // uniform over the workgroup
    mlp_precond_block<T>(blockIdx.x, pa);
    return;
  }
  factor_point<T>(((int)blockIdx.x - learned_blocks) * blockDim.x + threadIdx.x, np, pt_len, cs, lambda, lp,
                  fail_count, gp, dpv, t_out, hpp, lpp);
}

// Loss ||G^T A G - I||_F^2 / n per row and its gradient w.r.t. y (bf16 for the backward GEMM).
// A is the packed (lower, column-major) 9x9 target; rows >= n get zero gradient.
static __global__ void k_mlp_loss(int n, int npad, const float* __restrict__ y, const float* __restrict__ a,
                           uint16_t* __restrict__ dy, double* __restrict__ parts) {
  __shared__ double smem[kBlock / kWave];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  double loss = 0.0;
  if (b < npad) {
    float dyv[kOutPad];
    for (int k = 0; k < kOutPad; ++k) dyv[k] = 0.f;
    if (b < n) {
      float g[45], A[9][9], G[9][9], AG[9][9], E[9][9];
      build_g(y + (size_t)kOutPad * b, g);
      for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) {
          A[i][j] = a[(size_t)45 * b + sym9(i, j)];
          G[i][j] = i >= j ? g[sym_index(9, i, j)] : 0.f;
        }
      for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) {
          float s = 0.f;
          for (int k = 0; k < 9; ++k) s += A[i][k] * G[k][j];
          AG[i][j] = s;
        }
      for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) {
          float s = 0.f;
          for (int k = 0; k < 9; ++k) s += G[k][i] * AG[k][j];
          E[i][j] = s - (i == j ? 1.f : 0.f);
          loss += double(E[i][j]) * double(E[i][j]);
        }
      // dL/dG = 4 A G E, restricted to the lower triangle; dy = s dG / n
      const float scale = 4.f * kOutScale / float(n);
      for (int j = 0; j < 9; ++j)
        for (int i = j; i < 9; ++i) {
          float s = 0.f;
          for (int k = 0; k < 9; ++k) s += AG[i][k] * E[k][j];
          dyv[sym_index(9, i, j)] = scale * s;
        }
      loss /= double(n);
    }
    for (int k = 0; k < kOutPad; ++k) dy[(size_t)kOutPad * b + k] = f2bf(dyv[k]);
  }
  const double s = block_sum(loss, smem);
  if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

// dH = dZ W (per 16-row tile; wt = W transposed, [in][out]); then dZ_prev = dH * gelu'(Z_prev).
static __global__ void __launch_bounds__(256) k_mlp_backward_data(int nout, int nin, const uint16_t* __restrict__ dz,
                                                           const uint16_t* __restrict__ wt, const float* __restrict__ zprev,
                                                           uint16_t* __restrict__ dzprev) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  const size_t m0 = ((size_t)blockIdx.x * 4 + wave) * 16;
  for (int nt = 0; nt < nin / 16; ++nt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma_tile(dz + m0 * nout, nout, wt + (size_t)nt * 16 * nout, nout, nout, acc);
    for (int i = 0; i < 4; ++i) {
      const size_t row = m0 + 4 * q + i;
      const int col = nt * 16 + r;
      dzprev[row * nin + col] = f2bf(acc[i] * gelu_grad(zprev[row * nin + col]));
    }
  }
}

// dW[o][k] += sum_b dZ[b][o] X[b][k] and db[o] += sum_b dZ[b][o] over a 256-row chunk per
// workgroup (MFMA over the batch dimension, f32 atomics into the gradient).
static __global__ void __launch_bounds__(256) k_mlp_wgrad(int npad, int nout, int nin, const uint16_t* __restrict__ dz,
                                                   const uint16_t* __restrict__ xin, float* __restrict__ dw,
                                                   float* __restrict__ db) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  const int b0 = blockIdx.x * 256;
  const int b1 = min(b0 + 256, npad);
  const int tiles_o = nout / 16, tiles_k = nin / 16;
  for (int t = wave; t < tiles_o * tiles_k; t += 4) {
    const int o0 = (t / tiles_k) * 16, k0 = (t % tiles_k) * 16;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int bb = b0; bb < b1; bb += 16) {
      bf16x4 av, bv;
      for (int i = 0; i < 4; ++i) {
        av[i] = (short)dz[(size_t)(bb + 4 * q + i) * nout + o0 + r];
        bv[i] = (short)xin[(size_t)(bb + 4 * q + i) * nin + k0 + r];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    for (int i = 0; i < 4; ++i) atomicAdd(dw + (size_t)(o0 + 4 * q + i) * nin + k0 + r, acc[i]);
  }
  for (int o = threadIdx.x; o < nout; o += blockDim.x) {
    float s = 0.f;
    for (int bb = b0; bb < b1; ++bb) s += bf2f(dz[(size_t)bb * nout + o]);
    atomicAdd(db + o, s);
  }
}

// Adam on the f32 master parameters (bias-corrected), then bf16 copies W [out][in] and
// WT [in][out] of every weight matrix.
static __global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float beta1, float beta2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = beta1 * m[i] + (1.f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

static __global__ void k_mlp_convert(const float* __restrict__ p, uint16_t* __restrict__ w1, uint16_t* __restrict__ w1t,
                              uint16_t* __restrict__ w2, uint16_t* __restrict__ w2t, uint16_t* __restrict__ w3,
                              uint16_t* __restrict__ w3t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kHidden * kInPad) {
    const int o = i / kInPad, k = i % kInPad;
    const uint16_t h = f2bf(p[kW1 + i]);
    w1[i] = h;
    w1t[k * kHidden + o] = h;
  }
  if (i < kHidden * kHidden) {
    const int o = i / kHidden, k = i % kHidden;
    const uint16_t h = f2bf(p[kW2 + i]);
    w2[i] = h;
    w2t[k * kHidden + o] = h;
  }
  if (i < kOutPad * kHidden) {
    const int o = i / kHidden, k = i % kHidden;
    const uint16_t h = f2bf(p[kW3 + i]);
    w3[i] = h;
    w3t[k * kOutPad + o] = h;
  }
}

// Training target of one camera: A = L^-1 S_cc L^-T with S_cc = B + lambda I - D^-1 (sum_k
// Jc^T Jp M_p Jp^T Jc) D^-1 the camera's diagonal block of the reduced system and L the
// Cholesky factor of B + lambda I (one workgroup per camera, packed 45 output).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_schur_diagonal(int nc, const int* __restrict__ cam_off,
                                                           const int* __restrict__ k_pslot,
                                                           const Vec4<T>* __restrict__ rec_cm, const T* __restrict__ tab,
                                                           const double* __restrict__ mp, const T* __restrict__ dpv,
                                                           const T* __restrict__ bs,
                                                           const T* __restrict__ dc, const double* __restrict__ lc,
                                                           T lambda, float* __restrict__ out) {
  __shared__ T ct[kCamTab];
  __shared__ T red[45][kBlock / kWave];
  const int c = blockIdx.x;
  if (threadIdx.x < kCamTab) ct[threadIdx.x] = tab[kCamTab * c + threadIdx.x];
  __syncthreads();
  T acc[45];
  for (int i = 0; i < 45; ++i) acc[i] = T(0);
  const int k0 = cam_off[c], k1 = cam_off[c + 1];
  for (int k = k0 + threadIdx.x; k < k1; k += kBlock) {
    ObsGeom<T> geo;
    obs_geom(rec_cm[k], ct, geo);
    T jc[18];
    jc_block(geo, ct, jc);
    const int ps = k_pslot[k];
    // Jp (2x3), K2 = Jp M Jp^T (2x2) with M = (Hpp + lambda D^2)^-1 from the point's factor
    T jp[2][3];
    for (int a = 0; a < 3; ++a) {
      jp[0][a] = geo.sw * (geo.dp0[0] * ct[tR(0, a)] + geo.dp0[1] * ct[tR(1, a)] + geo.dp0[2] * ct[tR(2, a)]);
      jp[1][a] = geo.sw * (geo.dp1[0] * ct[tR(0, a)] + geo.dp1[1] * ct[tR(1, a)] + geo.dp1[2] * ct[tR(2, a)]);
    }
    T mj[2][3];
    for (int row = 0; row < 2; ++row) {
      const double v[3] = {double(jp[row][0]), double(jp[row][1]), double(jp[row][2])};
      double t[3];
      point_solve(mp, dpv, ps, v, t);
      for (int a = 0; a < 3; ++a) mj[row][a] = T(t[a]);
    }
    T k2[2][2];
    for (int u = 0; u < 2; ++u)
      for (int w = 0; w < 2; ++w) k2[u][w] = jp[u][0] * mj[w][0] + jp[u][1] * mj[w][1] + jp[u][2] * mj[w][2];
    for (int a = 0; a < 9; ++a) {
      const T t0 = k2[0][0] * jc[a] + k2[0][1] * jc[9 + a];
      const T t1 = k2[1][0] * jc[a] + k2[1][1] * jc[9 + a];
      for (int b = a; b < 9; ++b) acc[sym9(b, a)] += t0 * jc[b] + t1 * jc[9 + b];
    }
  }
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (int i = 0; i < 45; ++i) {
    const T v = wave_sum(acc[i]);
    if (lane == 0) red[i][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    T s[45], l[45];
    const T* d = dc + 9 * c;
    for (int j = 0; j < 9; ++j)
      for (int i = j; i < 9; ++i) {
        T e = T(0);
        for (int w = 0; w < kBlock / kWave; ++w) e += red[sym9(i, j)][w];
        s[sym9(i, j)] = bs[45 * c + sym9(i, j)] + (i == j ? lambda : T(0)) - (e / d[i]) / d[j];
      }
    for (int i = 0; i < 45; ++i) l[i] = T(lc[45 * c + i]);
    // A = L^-1 S L^-T: columns of L^-1 S, then rows
    T X[9][9];
    for (int j = 0; j < 9; ++j) {
      T col[9];
      for (int i = 0; i < 9; ++i) col[i] = s[sym9(i, j)];
      solve_lower_packed<T, 9>(l, col);
      for (int i = 0; i < 9; ++i) X[i][j] = col[i];
    }
    for (int i = 0; i < 9; ++i) {
      T row[9];
      for (int j = 0; j < 9; ++j) row[j] = X[i][j];
      solve_lower_packed<T, 9>(l, row);  // (L^-1 (L^-1 S)^T)^T = L^-1 S L^-T
      for (int j = 0; j <= i; ++j) out[(size_t)45 * c + sym9(i, j)] = float(row[j]);
    }
  }
}

}  // namespace fg
