// fg_kernels.h -- HIP kernels of the bundle-adjustment path (gfx950, wave64).
//
// Data layouts (all device memory):
//   * Points are processed in "point slots": points sorted by track length (longest first).
//     A chunk of 64 consecutive point slots is one wavefront; lane = point. The chunk's
//     observations are stored SELL-style: entry (j, lane) at chunk_base + 64 j + lane, so
//     every per-observation load of a point-major kernel is a coalesced 64-lane access.
//     Padding entries have camera -1.
//   * Camera-major kernels walk camera c's rows k in [off[c], off[c+1]) and keep a
//     camera-major copy of the records. Within a camera the rows are grouped by the point-slot
//     range of their point (point_ranges<T>() ranges of equal observation count; the reference's row
//     order within a group), so the Schur product's camera pass can give each range to its own
//     wave (k_schur_cameras); cm_ref maps a camera-major row to the reference row.
//   * Camera tables: kCamTab scalars per camera (see fg_math.h).
#pragma once
#include "fg_math.h"

namespace fg {

constexpr int kBlock = 256;          // threads per workgroup
constexpr int kParts = 6;            // partial sums per workgroup / camera
constexpr int kChunksPerBlock = kBlock / kWave;
// Point-slot ranges of the Schur product's camera pass (k_schur_cameras). Workgroup b takes
// range b % R, and workgroup index mod 8 is the XCD that runs it; each of the 8 XCDs has its own
// L2, so with R ranges an XCD gathers the per-point solves of 1/R of the points instead of all
// of them. FP64 (32-byte solves, 16 MB at 500,000 points): 8 ranges, 2 MB per XCD (unsplit,
// every XCD touches every line of the solves). FP32 (16-byte solves): 2 ranges; on the
// 1,500-camera workload 2, 4 and 8 ranges fetch 110, 100 and 95 MB per launch in 28.7, 29.6
// and 31.9 us: the pass waits on L1 misses in flight, which finer splits do not reduce, while
// they launch more, shorter waves with the same fixed costs (scalar loads, pipeline fill, the
// 9-sum wave reduction).
template <typename T>
constexpr int point_ranges() { return sizeof(T) == 4 ? 2 : 8; }

// PCG state. Kernels of launch i read slot i % 2 (the direction kernel) or (i + 1) % 2 (the
// product and update kernels); the direction kernel of launch i writes slot (i + 1) % 2 from
// block 0 only. No workgroup ever reads a slot that another workgroup of the same kernel writes,
// so no cross-workgroup synchronization is needed inside a kernel.
struct CgState {
  double rz, tol2, bb, target, rr, lambda;
  int done, iters, term, use_learned;
  int fallback_taken, pad0, pad1, pad2;
};
// LSMR state (the scalar recurrences of Fong & Saunders 2011, norms of the unnormalized
// vectors U = beta u and V = alpha v); declared here so the product kernels can read the
// update coefficients and the finished flag.
struct LsmrState {
  double alpha, beta, alphabar, zetabar, rho, rhobar, cbar, sbar;
  double f1, f2, f3;        // hbar, y and h update factors of the current iteration
  double norm_atb, abs_res;
  int done, converged, iters, max_iters;
  double rel_tol, root;             // sqrt(lambda)
  double beta_new, alpha_new;        // norms summed this iteration, adopted by k_lsmr_scalar
};
struct DevScalars {
  CgState st[2];
  int max_iters, fallback_at, pad0, pad1;
  double rel_tol, fallback_factor;
};

// Camera table rows and records of storage type T widened to f64 registers.
template <typename T>
__device__ __forceinline__ void load_cam16_d(const T* __restrict__ tab, int c, double out[16]) {
  T t[16];
  load_cam16(tab, c, t);
#pragma unroll
  for (int i = 0; i < 16; ++i) out[i] = double(t[i]);
}
// Product row of a camera (kPRow scalars): the table's (R | t) rows, (f, k1, k2, fw0), fw1-4,
// fw5-8 -- everything an observation of a point-major product needs from its camera in six
// 4-vector loads (a table row and a forward table take seven).
constexpr int kPRow = 24;
template <typename T>
__device__ __forceinline__ void load_prow_d(const T* __restrict__ prow, int c, double ct[16], double f[kFwd]) {
  const Vec4<T>* r4 = reinterpret_cast<const Vec4<T>*>(prow) + (kPRow / 4) * c;
  Vec4<T> v[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) v[i] = r4[i];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ct[4 * i] = double(v[i].x); ct[4 * i + 1] = double(v[i].y); ct[4 * i + 2] = double(v[i].z);
    ct[4 * i + 3] = double(v[i].w);
  }
  ct[12] = double(v[3].x); ct[13] = double(v[3].y); ct[14] = double(v[3].z); ct[15] = 0.0;
  f[0] = double(v[3].w);
  f[1] = double(v[4].x); f[2] = double(v[4].y); f[3] = double(v[4].z); f[4] = double(v[4].w);
  f[5] = double(v[5].x); f[6] = double(v[5].y); f[7] = double(v[5].z); f[8] = double(v[5].w);
  f[9] = f[10] = f[11] = 0.0;
}
// The product row from a table row and the forward table values v (T).
template <typename T>
__device__ __forceinline__ void store_prow(const T* __restrict__ ct, const T v[kFwd], T* __restrict__ prow) {
  Vec4<T>* r4 = reinterpret_cast<Vec4<T>*>(prow);
#pragma unroll
  for (int i = 0; i < 3; ++i) r4[i] = Vec4<T>{ct[4 * i], ct[4 * i + 1], ct[4 * i + 2], ct[4 * i + 3]};
  r4[3] = Vec4<T>{ct[tF], ct[tK1], ct[tK2], v[0]};
  r4[4] = Vec4<T>{v[1], v[2], v[3], v[4]};
  r4[5] = Vec4<T>{v[5], v[6], v[7], v[8]};
}

template <typename T>
__device__ __forceinline__ void load_fwd_d(const T* __restrict__ fw, int c, double out[kFwd]) {
  T t[kFwd];
  load_fwd(fw, c, t);
#pragma unroll
  for (int i = 0; i < kFwd; ++i) out[i] = double(t[i]);
}
template <typename T>
__device__ __forceinline__ Vec4<double> to_d(const Vec4<T>& r) {
  return Vec4<double>{double(r.x), double(r.y), double(r.z), double(r.w)};
}

// ---------------------------------------------------------------------------------------
// Small reduction helpers
// ---------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* smem /* kBlock/kWave */) {
  v = wave_sum(v);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  __syncthreads();
  if (lane == 0) smem[wave] = v;
  __syncthreads();
  T r = T(0);
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i) r += smem[i];
  return r;  // valid in thread 0
}

// Three block sums in one round (each the same tree as block_sum); valid in thread 0.
__device__ __forceinline__ void block_sum3(double v[3]) {
  __shared__ double red[3][kBlock / kWave];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int c = 0; c < 3; ++c) v[c] = wave_sum(v[c]);
  __syncthreads();
  if (lane == 0)
    for (int c = 0; c < 3; ++c) red[c][wave] = v[c];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) {
      double r = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) r += red[c][w];
      v[c] = r;
    }
}

// Deterministic sums of columns 0-2 of a (n x kParts) partial array by one workgroup (fixed
// assignment, fixed tree, the same per column as block_sum_parts_col); results broadcast to
// all threads, with two barriers for the three columns.
__device__ __forceinline__ void block_sum_parts_cols3(const double* parts, int n, double out[3]) {
  __shared__ double red[3][kBlock / kWave];
  __shared__ double bc[3];
  double s[3] = {0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
#pragma unroll
    for (int c = 0; c < 3; ++c) s[c] += parts[kParts * i + c];
  }
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int c = 0; c < 3; ++c) s[c] = wave_sum(s[c]);
  if (lane == 0)
    for (int c = 0; c < 3; ++c) red[c][wave] = s[c];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) {
      double r = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) r += red[c][w];
      bc[c] = r;
    }
  __syncthreads();
  out[0] = bc[0];
  out[1] = bc[1];
  out[2] = bc[2];
}

// Deterministic sum of column col of a (n x kParts) partial array by one workgroup (fixed
// assignment, fixed tree); result broadcast to all threads.
__device__ __forceinline__ double block_sum_parts_col(const double* parts, int n, int col, double* smem) {
  __shared__ double bcast;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += parts[kParts * i + col];
  s = block_sum(s, smem);
  if (threadIdx.x == 0) bcast = s;
  __syncthreads();
  return bcast;
}

// ---------------------------------------------------------------------------------------
// Camera tables: Projection::new (examples/bal/model.rs:126-146) in f64, stored in T.
// ---------------------------------------------------------------------------------------
// Projection::new for camera c from its 9 parameters cam (f64): the table row in tab64 and, with
// lo, its rounding to T and that rounding widened back to f64 (lo_wide).
template <typename T>
__device__ __forceinline__ void camera_table(int c, const double* cam, double* __restrict__ tab64,
                                             T* __restrict__ lo, double* __restrict__ lo_wide) {
  const double w0 = cam[0], w1 = cam[1], w2 = cam[2];
  const double t2 = w0 * w0 + w1 * w1 + w2 * w2;
  double a, b;
  if (t2 < 1e-8) {
    a = 0.5 - t2 / 24. + t2 * t2 / 720.;
    b = 1. / 6. - t2 / 120. + t2 * t2 / 5040.;
  } else {
    const double t = sqrt(t2);
    a = (1. - cos(t)) / t2;
    b = (t - sin(t)) / (t2 * t);
  }
  // K = [w]x ; Jr = I - a K + b K^2
  const double K[9] = {0, -w2, w1, w2, 0, -w0, -w1, w0, 0};
  double Jr[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double kk = 0;
      for (int k = 0; k < 3; ++k) kk += K[3 * i + k] * K[3 * k + j];
      Jr[3 * i + j] = (i == j ? 1.0 : 0.0) - a * K[3 * i + j] + b * kk;
    }
  // Rotation of the unit quaternion exp(w/2) (textbook quaternion-to-matrix formula).
  const double v0 = w0 / 2, v1 = w1 / 2, v2 = w2 / 2;
  const double nn = v0 * v0 + v1 * v1 + v2 * v2;
  const double eps = 2.220446049250313e-16;
  double qi = 0, qj = 0, qk = 0, qw = 1;
  if (!(nn <= eps * eps)) {
    const double n = sqrt(nn);
    const double s = sin(n) / n;
    qi = v0 * s; qj = v1 * s; qk = v2 * s; qw = cos(n);
  }
  const double ww = qw * qw, ii = qi * qi, jj = qj * qj, kk2 = qk * qk;
  const double ij = qi * qj * 2, wk = qw * qk * 2, wj = qw * qj * 2, ik = qi * qk * 2, jk = qj * qk * 2,
               wi = qw * qi * 2;
  const double R[9] = {ww + ii - jj - kk2, ij - wk, wj + ik, wk + ij, ww - ii + jj - kk2,
                       jk - wi, ik - wj, wi + jk, ww - ii - jj + kk2};
  double o[kCamTab];
  for (int i = 0; i < kCamTab; ++i) o[i] = 0.0;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      o[tR(i, j)] = R[3 * i + j];
      double s = 0;
      for (int k = 0; k < 3; ++k) s += R[3 * i + k] * Jr[3 * k + j];
      o[tRJ(i, j)] = s;
    }
    o[tT(i)] = cam[3 + i];
  }
  o[tF] = cam[6]; o[tK1] = cam[7]; o[tK2] = cam[8];
#pragma unroll
  for (int i = 0; i < kCamTab; ++i) tab64[kCamTab * c + i] = o[i];
  if (lo != nullptr) {
#pragma unroll
    for (int i = 0; i < kCamTab; ++i) {
      const T r = T(o[i]);
      lo[kCamTab * c + i] = r;
      lo_wide[kCamTab * c + i] = double(r);
    }
  }
}
template <typename T>
__global__ void k_camera_tables(int nc, const double* __restrict__ cams, double* __restrict__ tab64,
                                T* __restrict__ lo, double* __restrict__ lo_wide) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  camera_table<T>(c, cams + 9 * c, tab64, lo, lo_wide);
}

// Blockwise Huber (src/factors.rs:42-52): cost and sqrt(weight).
template <typename T>
__device__ __forceinline__ void huber_weight(T rx, T ry, T delta, T& cost, T& sw) {
  const T s2 = rx * rx + ry * ry;
  if (!(delta > T(0)) || s2 <= delta * delta) {
    cost = T(0.5) * s2;
    sw = T(1);
  } else {
    const T s = sqrt(s2);
    cost = delta * (s - T(0.5) * delta);
    sw = sqrt(delta / s);
  }
}

// Projection::value (model.rs:147-156). Returns false on z == 0 or a nonfinite pixel.
template <typename T>
__device__ __forceinline__ bool project(const T* __restrict__ tab, T x0, T x1, T x2, T& px, T& py, T& qx,
                                        T& qy, T& z) {
  const T p0 = tab[tR(0, 0)] * x0 + tab[tR(0, 1)] * x1 + tab[tR(0, 2)] * x2 + tab[tT(0)];
  const T p1 = tab[tR(1, 0)] * x0 + tab[tR(1, 1)] * x1 + tab[tR(1, 2)] * x2 + tab[tT(1)];
  const T p2 = tab[tR(2, 0)] * x0 + tab[tR(2, 1)] * x1 + tab[tR(2, 2)] * x2 + tab[tT(2)];
  qx = -p0 / p2;
  qy = -p1 / p2;
  z = p2;
  const T r2 = qx * qx + qy * qy;
  const T s = tab[tF] * (T(1) + tab[tK1] * r2 + tab[tK2] * r2 * r2);
  px = s * qx;
  py = s * qy;
  return p2 != T(0) && isfinite(px) && isfinite(py);
}

// Cooperative row gather. A lane-private gather of a per-camera row makes every 16-byte load
// instruction touch up to 64 cache lines, one per lane, and those per-line lookups bound the
// point-major passes. Here a wave fetches the 64 rows it needs (R 16-byte pieces each, row c
// starting at piece stride * c) as 64 R consecutive pieces: in round k (0 .. R-1) lane m fetches
// piece P = 64k + m, i.e. piece P % R of the row of lane P / R (whose camera a lane shuffle
// provides), so one instruction covers the lines of about 64 / R rows. The rows meet in the
// wave's LDS slice (row stride R + 1 pieces, so the 16-byte reads of consecutive rows fall on
// distinct banks in every group of 8 lanes) and each lane reads its own. Every lane of the
// wave must call both halves (lane shuffles); c < 0 rows are not fetched. Pure data movement:
// the values are bit-identical to a lane-private load.
template <int R>
__device__ __forceinline__ void coop_fetch(const Vec4<float>* __restrict__ rows, int stride, int c, int lane,
                                           Vec4<float> v[R]) {
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int piece = kWave * k + lane;
    const int r = piece / R;
    const int cr = __shfl(c, r, kWave);
    v[k] = Vec4<float>{0.f, 0.f, 0.f, 0.f};
    if (cr >= 0) v[k] = rows[stride * cr + (piece - R * r)];
  }
}
template <int R>
__device__ __forceinline__ void coop_exchange(const Vec4<float> v[R], int lane, Vec4<float>* __restrict__ lds,
                                              Vec4<float> u[R]) {
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int piece = kWave * k + lane;
    const int r = piece / R;
    lds[r * (R + 1) + (piece - R * r)] = v[k];
  }
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int k = 0; k < R; ++k) u[k] = lds[lane * (R + 1) + k];
  __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------------------
// Cost: sum over observations of 0.5||r||^2 or Huber, accumulated in f64 (Frame::cost,
// model.rs:201-213). Deterministic: per-lane sequential, fixed block tree; per-workgroup
// partials are summed by k_reduce_parts.
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cost(int nchunks, int np, const int* __restrict__ chunk_base,
                                                 const int* __restrict__ chunk_len, const int* __restrict__ slot_cam,
                                                 const Vec2<T>* __restrict__ slot_pix, const T* __restrict__ tab,
                                                 const T* __restrict__ pts, T huber, double* parts, int* flags) {
  __shared__ double smem[kBlock / kWave];
  // the first 16 table scalars of each observation's camera (R pieces of 16 bytes) are gathered
  // cooperatively by the wave, so every lane of a live chunk sweeps it (lanes past np read their
  // empty slots, whose camera is -1)
  constexpr int R = 16 * sizeof(T) / 16;
  __shared__ Vec4<float> rows[kChunksPerBlock * kWave * (R + 1)];
  const int wave = threadIdx.x / kWave;
  const int chunk = blockIdx.x * kChunksPerBlock + wave;
  const int lane = threadIdx.x % kWave;
  double acc = 0.0;
  bool bad = false;
  if (chunk < nchunks) {
    const int i = chunk * kWave + lane;
    const bool live = i < np;
    const T x0 = live ? pts[3 * i] : T(0), x1 = live ? pts[3 * i + 1] : T(0), x2 = live ? pts[3 * i + 2] : T(0);
    const int base = chunk_base[chunk], len = chunk_len[chunk];
    Vec4<float>* lds = rows + wave * kWave * (R + 1);
    for (int j = 0; j < len; ++j) {
      const int s = base + j * kWave + lane;
      const int c = slot_cam[s];
      Vec4<float> v[R], u[R];
      coop_fetch<R>(reinterpret_cast<const Vec4<float>*>(tab), kCamTab * (int)sizeof(T) / 16, c, lane, v);
      coop_exchange<R>(v, lane, lds, u);
      if (c >= 0) {
        T ct[16];
        __builtin_memcpy(ct, u, sizeof(ct));
        T px, py, qx, qy, z;
        bool ok = project(ct, x0, x1, x2, px, py, qx, qy, z);
        const Vec2<T> ob = slot_pix[s];
        T cost, sw;
        huber_weight(px - ob.x, py - ob.y, huber, cost, sw);
        bad |= !ok;
        acc += double(cost);
      }
    }
  }
  if (bad) atomicOr(flags, 1);
  const double b = block_sum(acc, smem);
  if (threadIdx.x == 0) parts[blockIdx.x] = b;
}

// The gradient stage's end in one workgroup: ||g||_inf from k_norm_inf's per-workgroup partials
// (the maximum k_reduce_parts forms, exact in any order) and the first three flag words
// (norm, linearization, damping scales), into host-visible memory: one kernel instead of a
// reduction and two readback copies.
struct GradReadback {
  double norm;
  int flags[4];
};
static __global__ void __launch_bounds__(kBlock) k_gradient_tail(const double* __restrict__ parts, int n,
                                                                 const int* __restrict__ flags, GradReadback* out) {
  __shared__ double smem[kBlock / kWave];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v = fmax(v, parts[i]);
  v = wave_max(v);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (lane == 0) smem[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i) v = fmax(v, smem[i]);
    out->norm = v;
    for (int i = 0; i < 3; ++i) out->flags[i] = flags[i];
  }
}

// Final reduction of per-workgroup partials by one workgroup: sum (op 0) or maximum (op 1).
static __global__ void k_reduce_parts(const double* __restrict__ parts, int n, int op, double* __restrict__ out) {
  __shared__ double smem[kBlock / kWave];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v = op == 0 ? v + parts[i] : fmax(v, parts[i]);
  if (op == 0) {
    v = block_sum(v, smem);
  } else {
    v = wave_max(v);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) smem[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < (int)(blockDim.x / kWave); ++i) v = fmax(v, smem[i]);
  }
  if (threadIdx.x == 0) *out = v;
}

// ---------------------------------------------------------------------------------------
// Linearization, point-major pass: projection, residual, IRLS weight, the compact record,
// weighted residual, the point Gram block Hpp = Jp^T Jp and gradient gp = Jp^T r.
// (Frame::linearize, model.rs:214-236; LinearizedModel::accumulate, linear_model.rs:180-210)
// The residual is the small difference of two pixel positions of a few hundred pixels, so the
// projection and the subtraction run in f64 from the f64 state, tables and observations in
// every precision. The record is stored in T; Jp is re-formed from the stored record and
// table in f64, the same way the Schur products form it, and Hpp and the gradient accumulate
// in f64, so the point blocks are the exact Gram blocks of the stored Jacobian.
// ---------------------------------------------------------------------------------------
// The compact record and weighted residual of one observation from the f64 camera table, the
// f64 point and the f64 pixel: projection, residual, IRLS weight, rounded to T. Both
// linearization passes form their copies with this function, with contraction off, so the
// point-major and camera-major records are the same bits. Returns false on a zero depth or a
// nonfinite pixel.
template <typename T>
__device__ __forceinline__ bool record_of(const double* __restrict__ cd, double x0, double x1, double x2,
                                          const Vec2<double>& ob, double huber, Vec4<T>& r4, Vec2<T>& rw,
                                          double& qx, double& qy, double& z, double& sw, double& wrx,
                                          double& wry) {
#pragma clang fp contract(off)
  double px, py;
  const double p0 = cd[tR(0, 0)] * x0 + cd[tR(0, 1)] * x1 + cd[tR(0, 2)] * x2 + cd[tT(0)];
  const double p1 = cd[tR(1, 0)] * x0 + cd[tR(1, 1)] * x1 + cd[tR(1, 2)] * x2 + cd[tT(1)];
  const double p2 = cd[tR(2, 0)] * x0 + cd[tR(2, 1)] * x1 + cd[tR(2, 2)] * x2 + cd[tT(2)];
  qx = -p0 / p2;
  qy = -p1 / p2;
  z = p2;
  const double r2 = qx * qx + qy * qy;
  const double sc = cd[tF] * (1.0 + cd[tK1] * r2 + cd[tK2] * r2 * r2);
  px = sc * qx;
  py = sc * qy;
  const bool ok = p2 != 0.0 && isfinite(px) && isfinite(py);
  const double rx = px - ob.x, ry = py - ob.y;
  const double s2 = rx * rx + ry * ry;
  if (!(huber > 0.0) || s2 <= huber * huber) {
    sw = 1.0;
  } else {
    sw = sqrt(huber / sqrt(s2));
  }
  r4 = Vec4<T>{T(qx), T(qy), T(z), T(sw)};
  wrx = sw * rx;
  wry = sw * ry;
  rw = Vec2<T>{T(wrx), T(wry)};
  return ok;
}

// One observation of the point-major linearization. cd is the f64 camera table; the table in
// T is its rounding (k_camera_tables stores both from the same f64 values), formed here
// instead of gathered a second time.
template <typename T>
__device__ __forceinline__ bool linearize_obs(const double cd[16], double x0, double x1, double x2,
                                              const Vec2<double>& ob, double huber, Vec4<T>& r4, Vec2<T>& rw,
                                              double h[6], double g[3]) {
  double qx, qy, z, sw, wrx, wry;
  bool ok = record_of<T>(cd, x0, x1, x2, ob, huber, r4, rw, qx, qy, z, sw, wrx, wry);
  double ct[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) ct[i] = double(T(cd[i]));
  ObsGeom<double> geo, g64;
  obs_geom(to_d(r4), ct, geo);
  obs_geom(Vec4<double>{qx, qy, z, sw}, cd, g64);
  // Jp columns: Jp[:, k] = sw * dp * R[:, k], from the stored record (Hpp) and from the f64
  // geometry (gradient)
  double jp0[3], jp1[3], e0[3], e1[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    jp0[k] = geo.sw * (geo.dp0[0] * ct[tR(0, k)] + geo.dp0[1] * ct[tR(1, k)] + geo.dp0[2] * ct[tR(2, k)]);
    jp1[k] = geo.sw * (geo.dp1[0] * ct[tR(0, k)] + geo.dp1[1] * ct[tR(1, k)] + geo.dp1[2] * ct[tR(2, k)]);
    e0[k] = g64.sw * (g64.dp0[0] * cd[tR(0, k)] + g64.dp0[1] * cd[tR(1, k)] + g64.dp0[2] * cd[tR(2, k)]);
    e1[k] = g64.sw * (g64.dp1[0] * cd[tR(0, k)] + g64.dp1[1] * cd[tR(1, k)] + g64.dp1[2] * cd[tR(2, k)]);
    ok = ok && isfinite(jp0[k]) && isfinite(jp1[k]);
  }
  ok = ok && isfinite(wrx) && isfinite(wry);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = a; b < 3; ++b) h[sym3(b, a)] += jp0[a] * jp0[b] + jp1[a] * jp1[b];
    g[a] += e0[a] * wrx + e1[a] * wry;
  }
  return ok;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_linearize_points(
    int nchunks, int np, const int* __restrict__ chunk_base, const int* __restrict__ chunk_len,
    const int* __restrict__ slot_cam, const Vec2<double>* __restrict__ slot_pix, const double* __restrict__ tab64,
    const double* __restrict__ pts, double huber, Vec4<T>* __restrict__ rec, Vec2<T>* __restrict__ res,
    double* __restrict__ hpp, T* __restrict__ gp, int* flags) {
  // the f64 camera tables (first 16 of 32 scalars: 8 pieces of a 16-piece row) are gathered
  // cooperatively by the wave, so every lane of a live chunk sweeps it (lanes past np read
  // their empty slots, whose camera is -1)
  __shared__ Vec4<float> rows[kChunksPerBlock * kWave * 9];
  const int wave = threadIdx.x / kWave;
  const int chunk = blockIdx.x * kChunksPerBlock + wave;
  const int lane = threadIdx.x % kWave;
  if (chunk >= nchunks) return;  // uniform over the wave
  const int i = chunk * kWave + lane;
  const bool live = i < np;
  const double x0 = live ? pts[3 * i] : 0.0, x1 = live ? pts[3 * i + 1] : 0.0, x2 = live ? pts[3 * i + 2] : 0.0;
  const int base = chunk_base[chunk], len = chunk_len[chunk];
  Vec4<float>* lds = rows + wave * kWave * 9;
  double h[6] = {0, 0, 0, 0, 0, 0};
  double g[3] = {0, 0, 0};
  bool bad = false;
  for (int j = 0; j < len; ++j) {
    const int s0 = base + j * kWave + lane;
    const int c0 = slot_cam[s0];
    Vec4<float> v[8], u[8];
    coop_fetch<8>(reinterpret_cast<const Vec4<float>*>(tab64), kCamTab / 2, c0, lane, v);
    coop_exchange<8>(v, lane, lds, u);
    if (c0 >= 0) {
      double cd0[16];
      __builtin_memcpy(cd0, u, sizeof(cd0));
      Vec4<T> r4;
      Vec2<T> rw;
      bad |= !linearize_obs<T>(cd0, x0, x1, x2, slot_pix[s0], huber, r4, rw, h, g);
      rec[s0] = r4;
      res[s0] = rw;
    }
  }
  if (!live) return;
  if (bad) atomicOr(flags, 1);
#pragma unroll
  for (int k = 0; k < 6; ++k) hpp[6 * i + k] = h[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) gp[3 * i + k] = T(g[k]);
}

// ---------------------------------------------------------------------------------------
// Linearization, camera-major pass split like k_schur_cameras: one wave per (camera, point
// range), four cameras of one range per workgroup, so the XCD running the workgroup gathers the
// f64 points of one range only. Per row the record and weighted residual are formed from the
// point, the f64 camera table and the pixel with record_of (as k_linearize_cameras does) and
// written to the camera-major copies; the wave's share of the 45 Gram entries (accumulated in
// T) and the 9 gradient entries (f64) goes to shares[54 sh ..], sh = kR c + g, and
// k_linearize_finish adds a camera's shares in range order. Rows are pipelined per lane as in
// k_schur_cameras; a lane past its last row re-forms the wave's last row (the same record, so
// its stores write the same bits) with weight 0 in the sums. FP32 storage only: in f64 the 45
// Gram sums take 90 registers, and the split pass ran at one or two waves per SIMD (171-236
// against 132 us for k_linearize_cameras), which FP64 keeps.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_linearize_cameras_split(
    int nc, const int* __restrict__ cam_off_r, const int* __restrict__ k_pslot, const double* __restrict__ pts,
    const Vec2<double>* __restrict__ cam_pix, const double* __restrict__ tab64, const T* __restrict__ tab,
    double huber, Vec4<T>* __restrict__ rec_cm, Vec2<T>* __restrict__ res_cm, double* __restrict__ shares,
    int* flags) {
  constexpr int kR = point_ranges<T>();
  constexpr int kWavesPerBlock = kBlock / kWave;
  const int lane = threadIdx.x % kWave;
  const int g = (int)blockIdx.x % kR;
  const int c = __builtin_amdgcn_readfirstlane(((int)blockIdx.x / kR) * kWavesPerBlock + (int)threadIdx.x / kWave);
  if (c >= nc) return;  // whole waves; no workgroup barrier follows
  const int sh = kR * c + g;
  const int k0 = cam_off_r[sh], k1 = cam_off_r[sh + 1];
  double cd[16];
  T ct[kCamTab];
#pragma unroll
  for (int a = 0; a < 16; ++a) cd[a] = tab64[kCamTab * c + a];
#pragma unroll
  for (int a = 0; a < kCamTab; ++a) ct[a] = tab[kCamTab * c + a];
  T acc[45];
  double gacc[9];
#pragma unroll
  for (int i = 0; i < 45; ++i) acc[i] = T(0);
#pragma unroll
  for (int i = 0; i < 9; ++i) gacc[i] = 0.0;
  bool bad = false;
  using D2 = double __attribute__((ext_vector_type(2)));
  using GD2 = const __attribute__((address_space(1))) D2;
  using GD = const __attribute__((address_space(1))) double;
  using GInt = const __attribute__((address_space(1))) int;
  GD* ptv = (GD*)(pts);
  GD2* pixv = (GD2*)(cam_pix);
  GInt* psv = (GInt*)(k_pslot);
  __asm__ volatile("" : "+s"(ptv), "+s"(pixv), "+s"(psv));
  struct Row {
    double x0, x1, x2;
    D2 ob;
  };
  const int n = (k1 - k0 + kWave - 1) / kWave;  // uniform over the wave
  if (n > 0) {
    const int last = k1 - 1;
    int k = k0 + lane;
    auto fetch = [&](int kk, int ps, Row& r) {
      r.x0 = ptv[3 * ps];
      r.x1 = ptv[3 * ps + 1];
      r.x2 = ptv[3 * ps + 2];
      r.ob = pixv[kk];
    };
    auto row = [&](int kk, const Row& r, bool valid) {
      Vec4<T> r4;
      Vec2<T> rw;
      double qx, qy, z, sw, wrx, wry;
      bad |= !record_of<T>(cd, r.x0, r.x1, r.x2, Vec2<double>{r.ob.x, r.ob.y}, huber, r4, rw, qx, qy, z, sw, wrx,
                           wry);
      rec_cm[kk] = r4;
      res_cm[kk] = rw;
      ObsGeom<T> geo;
      obs_geom(r4, ct, geo);
      T jc[18];
      jc_block(geo, ct, jc);
#pragma unroll
      for (int a = 0; a < 18; ++a) bad |= !isfinite(jc[a]);
      const T wt = valid ? T(1) : T(0);
#pragma unroll
      for (int a = 0; a < 18; ++a) jc[a] *= wt;
#pragma unroll
      for (int a = 0; a < 9; ++a) {
#pragma unroll
        for (int b = a; b < 9; ++b) acc[sym9(b, a)] += jc[a] * jc[b] + jc[9 + a] * jc[9 + b];
      }
#pragma unroll
      for (int a = 0; a < 9; ++a) gacc[a] += double(jc[a]) * double(rw.x) + double(jc[9 + a]) * double(rw.y);
    };
    {
      Row ra, rb;
      fetch(min(k, last), psv[min(k, last)], ra);
      int ps_b = psv[min(k + kWave, last)];
      for (int m = 0;; m += 2) {
        const int ps_a = psv[min(k + 2 * kWave, last)];
        fetch(min(k + kWave, last), ps_b, rb);
        __asm__ volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        row(min(k, last), ra, k < k1);
        if (m + 1 >= n) break;
        ps_b = psv[min(k + 3 * kWave, last)];
        fetch(min(k + 2 * kWave, last), ps_a, ra);
        __asm__ volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        row(min(k + kWave, last), rb, k + kWave < k1);
        if (m + 2 >= n) break;
        k += 2 * kWave;
      }
    }
  }
  if (__builtin_amdgcn_ballot_w64(bad) != 0 && lane == 0) atomicOr(flags, 1);
#pragma unroll
  for (int i = 0; i < 45; ++i) {
    const T v = wave_sum(acc[i]);
    if (lane == 0) shares[54 * sh + i] = double(v);
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const double v = wave_sum(gacc[i]);
    if (lane == 0) shares[54 * sh + 45 + i] = v;
  }
}

// One thread per (camera, entry): the camera's shares of k_linearize_cameras_split in range
// order into Hcc (45 packed) and gc (9).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_linearize_finish(int nc, const double* __restrict__ shares,
                                                             T* __restrict__ hcc, T* __restrict__ gc) {
  constexpr int kR = point_ranges<T>();
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = id / 54, e = id % 54;
  if (c >= nc) return;
  double s = 0.0;
#pragma unroll
  for (int g = 0; g < kR; ++g) s += shares[54 * (kR * c + g) + e];
  if (e < 45) hcc[45 * c + e] = T(s);
  else gc[9 * c + e - 45] = T(s);
}

// ---------------------------------------------------------------------------------------
// Linearization, camera-major pass (one workgroup per camera): the camera-major copies of the
// records, formed again from the f64 points (gathered by point slot from a 12 MB array in the
// 1,500-camera case, instead of gathering the 78 MB of slot-order records and residuals) with
// the same record_of as the point-major pass, then the camera Gram block Hcc = Jc^T Jc (45
// packed) and gradient gc = Jc^T r.
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_linearize_cameras(
    const int* __restrict__ cam_off, const int* __restrict__ k_pslot, const double* __restrict__ pts,
    const Vec2<double>* __restrict__ cam_pix, const double* __restrict__ tab64, const T* __restrict__ tab,
    double huber, Vec4<T>* __restrict__ rec_cm, Vec2<T>* __restrict__ res_cm, T* __restrict__ hcc,
    T* __restrict__ gc, int* flags) {
  // The weighted residual was formed in f64 and rounded once to T (relative error eps of the
  // residual itself), so the gradient needs only f64 accumulation here.
  __shared__ T ct[kCamTab];
  __shared__ double red[54][kBlock / kWave];
  const int c = blockIdx.x;
  if (threadIdx.x < kCamTab) ct[threadIdx.x] = tab[kCamTab * c + threadIdx.x];
  __syncthreads();
  T acc[45];
  double gacc[9];
#pragma unroll
  for (int i = 0; i < 45; ++i) acc[i] = T(0);
#pragma unroll
  for (int i = 0; i < 9; ++i) gacc[i] = 0.0;
  bool bad = false;
  const int k0 = cam_off[c], k1 = cam_off[c + 1];
  const double* __restrict__ cd = tab64 + kCamTab * c;  // uniform: scalar loads
  auto one = [&](int k, const double x[3], const Vec2<double>& ob) {
    Vec4<T> r4;
    Vec2<T> rw;
    double qx, qy, z, sw, wrx, wry;
    bad |= !record_of<T>(cd, x[0], x[1], x[2], ob, huber, r4, rw, qx, qy, z, sw, wrx, wry);
    rec_cm[k] = r4;
    res_cm[k] = rw;
    ObsGeom<T> geo;
    obs_geom(r4, ct, geo);
    T jc[18];
    jc_block(geo, ct, jc);
#pragma unroll
    for (int a = 0; a < 18; ++a) bad |= !isfinite(jc[a]);
#pragma unroll
    for (int a = 0; a < 9; ++a) {
#pragma unroll
      for (int b = a; b < 9; ++b) acc[sym9(b, a)] += jc[a] * jc[b] + jc[9 + a] * jc[9 + b];
    }
#pragma unroll
    for (int a = 0; a < 9; ++a) gacc[a] += double(jc[a]) * double(rw.x) + double(jc[9 + a]) * double(rw.y);
  };
  // one row per step: two in flight held 171 (FP32) / 256 (FP64) VGPRs, 2 / 1 waves per SIMD
  for (int k = k0 + threadIdx.x; k < k1; k += kBlock) {
    const int pa = k_pslot[k];
    const double xa[3] = {pts[3 * pa], pts[3 * pa + 1], pts[3 * pa + 2]};
    one(k, xa, cam_pix[k]);
  }
  if (bad) atomicOr(flags, 1);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int i = 0; i < 45; ++i) {
    const T v = wave_sum(acc[i]);
    if (lane == 0) red[i][wave] = double(v);
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const double v = wave_sum(gacc[i]);
    if (lane == 0) red[45 + i][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < 54) {
    double s = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) s += red[threadIdx.x][w];
    if (threadIdx.x < 45) hcc[45 * c + threadIdx.x] = T(s);
    else gc[9 * c + threadIdx.x - 45] = T(s);
  }
}

__device__ __forceinline__ bool memcmp_bits(float a, float b) { return __float_as_uint(a) == __float_as_uint(b); }
__device__ __forceinline__ bool memcmp_bits(double a, double b) {
  return __double_as_longlong(a) == __double_as_longlong(b);
}

// Consistency check of the two record copies: counts camera-major rows whose record or
// weighted residual differs in any bit from the point-major copy of the same observation.
template <typename T>
__global__ void k_record_mismatch(int no, const int* __restrict__ k_slot, const Vec4<T>* __restrict__ rec,
                                  const Vec2<T>* __restrict__ res, const Vec4<T>* __restrict__ rec_cm,
                                  const Vec2<T>* __restrict__ res_cm, int* count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= no) return;
  const int s = k_slot[k];
  const Vec4<T> a = rec[s], b = rec_cm[k];
  const Vec2<T> u = res[s], v = res_cm[k];
  const bool same = memcmp_bits(a.x, b.x) && memcmp_bits(a.y, b.y) && memcmp_bits(a.z, b.z) &&
                    memcmp_bits(a.w, b.w) && memcmp_bits(u.x, v.x) && memcmp_bits(u.y, v.y);
  if (!same) atomicAdd(count, 1);
}

// ---------------------------------------------------------------------------------------
// Infinity norm with nonfinite detection (norm_inf, src/optimization.rs:653-662).
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_norm_inf(int n, const T* __restrict__ a, int m, const T* __restrict__ b,
                                                     double* parts, int* flags) {
  __shared__ double smem[kBlock / kWave];
  double v = 0.0;
  bool bad = false;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n + m; i += gridDim.x * blockDim.x) {
    const double x = double(i < n ? a[i] : b[i - n]);
    bad |= !isfinite(x);
    v = fmax(v, fabs(x));
  }
  if (bad) atomicOr(flags, 2);
  v = wave_max(v);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (lane == 0) smem[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0;
    for (int i = 0; i < kBlock / kWave; ++i) r = fmax(r, smem[i]);
    parts[blockIdx.x] = r;
  }
}

// ---------------------------------------------------------------------------------------
// prepare_scales (linear_model.rs:211-230) and the normalized Gram blocks
// (block_factors.rs:47-111): D = max(column norm, floor); B = D^-1 Hcc D^-1; C = D^-1 Hpp D^-1.
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void k_scales(int nc, int np, const T* __restrict__ hcc, const double* __restrict__ hpp, T floor,
                         T* __restrict__ dc, T* __restrict__ dp, T* __restrict__ bs, double* __restrict__ cs,
                         int* flags) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < nc) {
    const T* h = hcc + 45 * idx;
    T d[9];
    bool bad = false;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const T n = sqrt(h[sym9(j, j)]);
      bad |= !isfinite(n);
      d[j] = max(n, floor);
      dc[9 * idx + j] = d[j];
    }
#pragma unroll
    for (int j = 0; j < 9; ++j)
#pragma unroll
      for (int i = j; i < 9; ++i) bs[45 * idx + sym9(i, j)] = (h[sym9(i, j)] / d[i]) / d[j];
    if (bad) atomicOr(flags, 4);
  } else if (idx < nc + np) {
    const int p = idx - nc;
    const double* h = hpp + 6 * p;
    double d[3];
    bool bad = false;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double n = sqrt(h[sym3(j, j)]);
      bad |= !isfinite(n);
      // the stored (T) scale is the one every later use of Dp reads
      dp[3 * p + j] = T(max(n, double(floor)));
      d[j] = double(dp[3 * p + j]);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int i = j; i < 3; ++i) cs[6 * p + sym3(i, j)] = (h[sym3(i, j)] / d[i]) / d[j];
    if (bad) atomicOr(flags, 4);
  }
}

// ---------------------------------------------------------------------------------------
// Precision split of the solve. T is the storage type of the bulk per-observation data
// (records, weighted residuals, camera tables) and of the Gram blocks; every product of the
// reduced system is evaluated in f64 registers from those values, and the camera-sized
// vectors (PCG state, forward tables) and the per-point solves are f64. The reduced matrix is
// then the exact Schur complement of the Jacobian the stored records define, which keeps it
// positive semidefinite down to the smallest damping: with f32 arithmetic, the rounding of
// the products along the similarity gauge (where S has eigenvalue lambda) exceeds lambda once
// lambda falls below ~1e-7 and PCG breaks down on its first direction. Two rounding points
// remain, chosen so that their error in p.Sp is first order in the eliminated residual
// e = u - Jp t (whose square p.Sp is) rather than in u: the forward table of p is stored in T
// (both passes then see the same rounded direction), while the per-point solve t stays f64
// (a rounded t would leave an error of order eps |u|^2 in p.Sp).
// ---------------------------------------------------------------------------------------

// ---------------------------------------------------------------------------------------
// BlockFactors::factor (block_factors.rs:113-138), point blocks: the f64 Cholesky factor of
// C + lambda I (packed 6), applied by substitution (point_solve). Solving with the factor is
// backward stable, so the eliminated projection I - Jp (Hpp + lambda D^2)^-1 Jp^T stays
// positive semidefinite to rounding; an explicit inverse is off by eps * cond(Hpp), which at
// lambda = 1e-12 is enough to stall a 1e-10 PCG on the gauge directions. Unobserved blocks
// are skipped (no Gram block exists for them in the reference): a zero factor, zero solve.
// ---------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void point_solve(const double* __restrict__ lp, const T* __restrict__ dpv, int i,
                                            const double v[3], double t[3]);
template <typename T>
__device__ __forceinline__ void store_tp(double* tp, int i, const double t[3]);

// With t_out: also the right-hand side's point solves t_p = (Hpp + lambda Dp^2)^-1 gp of
// factor_and_reduce_rhs (what k_points<kRhs> computes), from the factor just stored.
// With lpp: also the unscaled factor of Hpp + lambda Dp^2 from the linearization's f64 Hpp
// (the Gram of the Jp the products re-form), which k_points<kSchur> solves with; a zero
// factor where it fails (an unobserved point at lambda = 0), where kSchur's t is zero.
template <typename T>
__device__ __forceinline__ void factor_point(int p, int np, const int* __restrict__ pt_len,
                                             const double* __restrict__ cs, double lambda, double* __restrict__ lp,
                                             int* fail_count, const T* __restrict__ gp, const T* __restrict__ dpv,
                                             double* __restrict__ t_out, const double* __restrict__ hpp,
                                             double* __restrict__ lpp) {
  if (p >= np) return;
  double* l = lp + 6 * p;
  if (lpp != nullptr) {
    double h[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) h[k] = pt_len[p] == 0 ? 0.0 : hpp[6 * p + k];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double d = double(dpv[3 * p + a]);
      h[sym3(a, a)] += lambda * d * d;
    }
    const bool ok = pt_len[p] != 0 && chol_packed<double, 3>(h);
#pragma unroll
    for (int k = 0; k < 6; ++k) lpp[6 * p + k] = ok ? h[k] : 0.0;
  }
  if (pt_len[p] == 0) {
#pragma unroll
    for (int i = 0; i < 6; ++i) l[i] = 0.0;
    if (t_out != nullptr) {
      const double z[3] = {0.0, 0.0, 0.0};
      store_tp<T>(t_out, p, z);
    }
    return;
  }
  double a[6];
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    a[i] = cs[6 * p + i];
    ok = ok && isfinite(a[i]);
  }
  a[0] += lambda; a[3] += lambda; a[5] += lambda;
  ok = ok && chol_packed<double, 3>(a);
#pragma unroll
  for (int i = 0; i < 6; ++i) l[i] = ok ? a[i] : 0.0;  // zero factor: failed (see cameras)
  if (!ok) atomicAdd(fail_count, 1);
  if (t_out != nullptr) {
    const double v[3] = {double(gp[3 * p]), double(gp[3 * p + 1]), double(gp[3 * p + 2])};
    double t[3];
    point_solve(lp, dpv, p, v, t);
    store_tp<T>(t_out, p, t);
  }
}
template <typename T>
__global__ void k_factor_points(int np, const int* __restrict__ pt_len, const double* __restrict__ cs,
                                double lambda, double* __restrict__ lp, int* fail_count,
                                const T* __restrict__ gp, const T* __restrict__ dpv, double* __restrict__ t_out,
                                const double* __restrict__ hpp, double* __restrict__ lpp) {
  factor_point<T>(blockIdx.x * blockDim.x + threadIdx.x, np, pt_len, cs, lambda, lp, fail_count, gp, dpv, t_out, hpp,
                  lpp);
}

// t = (Hpp + lambda D^2)^-1 v = D^-1 (C + lambda I)^-1 D^-1 v with the stored factor of point i.
template <typename T>
__device__ __forceinline__ void point_solve(const double* __restrict__ lp, const T* __restrict__ dpv, int i,
                                            const double v[3], double t[3]) {
  double l[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) l[k] = lp[6 * i + k];
  if (l[0] == 0.0) {  // unobserved point
    t[0] = t[1] = t[2] = 0.0;
    return;
  }
  double d[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    d[a] = double(dpv[3 * i + a]);
    t[a] = v[a] / d[a];
  }
  solve_lower_packed<double, 3>(l, t);
  solve_upper_packed<double, 3>(l, t);
#pragma unroll
  for (int a = 0; a < 3; ++a) t[a] /= d[a];
}

// The factor of camera block B + lambda I (B: the scaled Gram, packed lower, in T) in registers;
// false where an entry is nonfinite or the factorization fails.
template <typename T>
__device__ __forceinline__ bool camera_chol(const T* __restrict__ b, double lambda, double a[45]) {
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 45; ++i) {
    a[i] = double(b[i]);
    ok = ok && isfinite(a[i]);
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) a[sym9(k, k)] += lambda;
  return ok && chol_packed<double, 9>(a);
}

// Camera blocks of BlockFactors::factor: nine lanes per camera each factor B + lambda I in
// registers (the same arithmetic, so the same factor) and solve one column of its explicit
// inverse; lane 0 stores the factor and counts a failure. Thread gid: camera gid / 9, inverse
// column gid % 9.
template <typename T>
__device__ __forceinline__ void factor_camera(int gid, int nc, const int* __restrict__ cam_off,
                                              const T* __restrict__ bs, double lambda, double* __restrict__ lc,
                                              double* __restrict__ mbj, int* fail_count) {
  const int c = gid / 9, j = gid % 9;
  if (c >= nc) return;
  if (cam_off[c + 1] == cam_off[c]) {  // unobserved: no block; a zero factor marks it
    if (j == 0)
      for (int i = 0; i < 45; ++i) lc[45 * c + i] = 0.0;
    return;
  }
  double a[45];
  const bool ok = camera_chol<T>(bs + 45 * c, lambda, a);
  if (j == 0) {
    // a failed block stores a zero factor: Schur reports it, LSMR keeps diagonal treatment
#pragma unroll
    for (int i = 0; i < 45; ++i) lc[45 * c + i] = ok ? a[i] : 0.0;
    if (!ok) atomicAdd(fail_count, 1);
  }
  // column j of (B + lambda I)^-1 = L^-T L^-1 e_j, lower part only
  double e[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) e[i] = i == j ? 1.0 : 0.0;
  solve_lower_packed<double, 9>(a, e);
  solve_upper_packed<double, 9>(a, e);
  for (int i = j; i < 9; ++i) mbj[45 * c + sym9(i, j)] = e[i];
}
template <typename T>
__global__ void k_factor_cameras(int nc, const int* __restrict__ cam_off, const T* __restrict__ bs, double lambda,
                                 double* __restrict__ lc, double* __restrict__ mbj, int* fail_count) {
  factor_camera<T>(blockIdx.x * blockDim.x + threadIdx.x, nc, cam_off, bs, lambda, lc, mbj, fail_count);
}

// ---------------------------------------------------------------------------------------
// Forward camera tables from a retained-coordinate vector. scaled != 0: z = x / Dc (the
// reduced system's scaled coordinates); otherwise z = x. fw = (RJ z_rot, z_t, z_intr).
// ---------------------------------------------------------------------------------------
// fww (f32 storage only): the same table widened to f64, which the camera-major kernels read
// with uniform loads.
template <typename T>
__device__ __forceinline__ void forward_table(const T* __restrict__ ct, const double* x, const T* __restrict__ d,
                                              bool scaled, T* __restrict__ fw, double* __restrict__ fww,
                                              T* __restrict__ prow, double z[9]) {
#pragma unroll
  for (int j = 0; j < 9; ++j) z[j] = scaled ? x[j] / double(d[j]) : x[j];
  T v[kFwd];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    v[i] = T(double(ct[tRJ(i, 0)]) * z[0] + double(ct[tRJ(i, 1)]) * z[1] + double(ct[tRJ(i, 2)]) * z[2]);
#pragma unroll
  for (int j = 3; j < 9; ++j) v[j] = T(z[j]);
  v[9] = v[10] = v[11] = T(0);
#pragma unroll
  for (int i = 0; i < kFwd; ++i) fw[i] = v[i];
  if (fww != nullptr) {
#pragma unroll
    for (int i = 0; i < kFwd; ++i) fww[i] = double(v[i]);
  }
  store_prow(ct, v, prow);
}

// forward_table (scaled) from registers: rj = the table's R Jr rows (T values widened), d = Dc.
template <typename T>
__device__ __forceinline__ void forward_table_rows(const double rj[9], const double* x, const double d[9],
                                                   const T trow[16], T* __restrict__ fw, double* __restrict__ fww,
                                                   T* __restrict__ prow) {
  double z[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) z[j] = x[j] / d[j];
  T v[kFwd];
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] = T(rj[3 * i] * z[0] + rj[3 * i + 1] * z[1] + rj[3 * i + 2] * z[2]);
#pragma unroll
  for (int j = 3; j < 9; ++j) v[j] = T(z[j]);
  v[9] = v[10] = v[11] = T(0);
#pragma unroll
  for (int i = 0; i < kFwd; ++i) fw[i] = v[i];
  if (fww != nullptr) {
#pragma unroll
    for (int i = 0; i < kFwd; ++i) fww[i] = double(v[i]);
  }
  store_prow(trow, v, prow);
}

template <typename T>
__global__ void k_forward_tables(int nc, const T* __restrict__ tab, const double* __restrict__ x,
                                 const T* __restrict__ dc, int scaled, T* __restrict__ fw, double* __restrict__ fww,
                                 T* __restrict__ prow) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double z[9];
  forward_table(tab + kCamTab * c, x + 9 * c, dc + 9 * c, scaled != 0, fw + kFwd * c,
                fww != nullptr ? fww + kFwd * c : nullptr, prow + kPRow * c, z);
}

// ---------------------------------------------------------------------------------------
// Point-major half of the Schur products (one lane per point, SELL chunks).
//   kSchur:   t_p = (Hpp + lambda Dp^2)^-1 sum_k Jp^T Jc z (the eliminated-block solve of
//             E^T z), with Hpp re-formed in f64 from the same records and factored in registers
//   kRecover: delta_p = -(Hpp + lambda Dp^2)^-1 (gp + sum_k Jp^T Jc z)   (Schur::recover_step,
//             schur.rs:312-344)
//   kRhs:     t_p = (Hpp + lambda Dp^2)^-1 gp   (factor_and_reduce_rhs, schur.rs:204-247)
//   kVerify:  normal residual Jp^T (J delta + r) + lambda Dp^2 delta_p and ||J delta||^2
//   kJv:      rows 2k,2k+1 of J x               (Jacobian::apply, linear_model.rs:39-55)
//   kJtu:     x_p = sum_k Jp^T u_k              (Jacobian::transpose_apply, :64-87)
//   kLsmr:    one LSMR step over rows kept in slot order: U_s = J ws - (alpha/beta) U_s for
//             each observation, then x_p = sum_k Jp^T U_k, the point damping rows
//             U_p = sqrt(lambda) Dp ws_p - (alpha/beta) U_p, and ||U||^2 partials; with no
//             forward table it only forms x_p and ||U||^2 of the current U (the start vector)
// ---------------------------------------------------------------------------------------
//   kCheck:   the whole point side of a damped attempt's recovery and verification in one
//             sweep over the point's observations: kSchur at z (t_p for the reduced residual
//             ||S z - rhs||), kRecover (delta_p from the same sum), and kVerify's terms at
//             delta = (z / D, delta_p) expanded in sums of that sweep
enum PointMode { kSchur = 0, kRecover = 1, kRhs = 2, kVerify = 3, kJv = 4, kJtu = 5, kLsmr = 6, kCheck = 7 };
// Sweep mode of kCheck only: kSchur's sums plus sum ||Jc z||^2 (uu) and sum Jp^T r (gr).
constexpr int kSchurSums = 8;

// Chunks at least this long (the first ones: points are ordered longest track first) are
// swept by all four waves of one workgroup, each taking a quarter of the observation steps,
// so that a long-tailed track distribution does not leave a few waves looping for the rest.
constexpr int kSplitLen = 32;

struct PointArgs {
  int nchunks, np;
  int nsplit;  // leading chunks that each take a whole workgroup
  const int* chunk_base;
  const int* chunk_len;
  const int* slot_cam;
  const int* slot_row;
};

// The eliminated-block solves t_p handed from the point half of a product to the camera half,
// one 4-vector row per point in the storage type (16 bytes in f32, gathered once per row).
// A rounded t makes the f32 product S~p = lambda p + Dc^-1 Jc^T (u - Jp t~), whose curvature
// p.S~p is off by (Jp^T u).(t~ - t), first order in the rounding: along the similarity gauge at
// lambda ~ 1e-12 it can come out nonpositive (NotPositiveDefinite, a rejected LM trial). f64
// rows remove that but slowed the f32 camera pass from 34.4 to 47-56 us in every layout tried.
template <typename T>
__device__ __forceinline__ void store_tp(double* tp, int i, const double t[3]) {
  reinterpret_cast<Vec4<T>*>(tp)[i] = Vec4<T>{T(t[0]), T(t[1]), T(t[2]), T(0)};
}
template <typename T>
__device__ __forceinline__ Vec4<double> load_tp(const double* tp, int i) {
  const Vec4<T> v = reinterpret_cast<const Vec4<T>*>(tp)[i];
  return Vec4<double>{double(v.x), double(v.y), double(v.z), 0.0};
}

// The point-major modes that apply a camera vector read the cameras' product rows.
template <int MODE>
constexpr bool point_fwd_mode() {
  return MODE == kSchur || MODE == kSchurSums || MODE == kRecover || MODE == kVerify || MODE == kJv ||
         MODE == kLsmr;
}

// One observation of a point-major pass with its camera's table (and, with fwd, product row)
// already in registers.
template <typename T, int MODE>
__device__ __forceinline__ void point_obs_row(const PointArgs& pa, int s, const Vec4<T>& rec,
                                              const Vec2<T>* __restrict__ res, const double ct[16],
                                              const double f[kFwd], bool fwd, const double dpx[3],
                                              const double* __restrict__ rows_in, double* __restrict__ rows_out,
                                              double w[3], double h[6], double& uu, double coef, double gr[3]) {
  ObsGeom<double> geo;
  obs_geom(to_d(rec), ct, geo);
  if (MODE == kSchur) {
    // w += Jp^T Jc z; the point block comes factored (k_factor_points)
    double u0, u1;
    jc_apply(geo, f, u0, u1);
    jpt_accumulate(geo, ct, u0, u1, w);
  } else if (MODE == kSchurSums) {
    double u0, u1;
    jc_apply(geo, f, u0, u1);
    double j0[3], j1[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      j0[k] = geo.sw * (geo.dp0[0] * ct[tR(0, k)] + geo.dp0[1] * ct[tR(1, k)] + geo.dp0[2] * ct[tR(2, k)]);
      j1[k] = geo.sw * (geo.dp1[0] * ct[tR(0, k)] + geo.dp1[1] * ct[tR(1, k)] + geo.dp1[2] * ct[tR(2, k)]);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int b = a; b < 3; ++b) h[sym3(b, a)] += j0[a] * j0[b] + j1[a] * j1[b];
      w[a] += j0[a] * u0 + j1[a] * u1;
    }
    if (MODE == kSchurSums) {
      uu += u0 * u0 + u1 * u1;
      const Vec2<T> r = res[s];
#pragma unroll
      for (int a = 0; a < 3; ++a) gr[a] += j0[a] * double(r.x) + j1[a] * double(r.y);
    }
  } else if (MODE == kRecover) {
    double u0, u1;
    jc_apply(geo, f, u0, u1);
    jpt_accumulate(geo, ct, u0, u1, w);
  } else if (MODE == kVerify || MODE == kJv) {
    double u0, u1, v0, v1;
    jc_apply(geo, f, u0, u1);
    jp_apply(geo, ct, dpx, v0, v1);
    u0 += v0;
    u1 += v1;
    if (MODE == kJv) {
      const int row = pa.slot_row[s];
      rows_out[2 * row] = u0;
      rows_out[2 * row + 1] = u1;
    } else {
      uu += u0 * u0 + u1 * u1;
      const Vec2<T> r = res[s];
      jpt_accumulate(geo, ct, u0 + double(r.x), u1 + double(r.y), w);
    }
  } else if (MODE == kJtu) {
    const int row = pa.slot_row[s];
    jpt_accumulate(geo, ct, rows_in[2 * row], rows_in[2 * row + 1], w);
  } else if (MODE == kLsmr) {
    const Vec2<double> old = reinterpret_cast<const Vec2<double>*>(rows_in)[s];
    double u0 = old.x, u1 = old.y;
    if (fwd) {
      double v0, v1, x0, x1;
      jc_apply(geo, f, v0, v1);
      jp_apply(geo, ct, dpx, x0, x1);
      u0 = (v0 + x0) - coef * u0;
      u1 = (v1 + x1) - coef * u1;
      reinterpret_cast<Vec2<double>*>(rows_out)[s] = Vec2<double>{u0, u1};
    }
    uu += u0 * u0 + u1 * u1;
    jpt_accumulate(geo, ct, u0, u1, w);
  }
}

// One observation of a point-major pass: record, camera table and forward table in registers.
template <typename T, int MODE>
__device__ __forceinline__ void point_obs(const PointArgs& pa, int s, int c, const Vec4<T>& rec,
                                          const Vec2<T>* __restrict__ res, const T* __restrict__ tab,
                                          const T* __restrict__ fw, const double dpx[3],
                                          const double* __restrict__ rows_in, double* __restrict__ rows_out,
                                          double w[3], double h[6], double& uu, double coef, double gr[3]) {
  double ct[16], f[kFwd];
  if (point_fwd_mode<MODE>() && fw != nullptr) load_prow_d(fw, c, ct, f);
  else load_cam16_d(tab, c, ct);
  point_obs_row<T, MODE>(pa, s, rec, res, ct, f, fw != nullptr, dpx, rows_in, rows_out, w, h, uu, coef, gr);
}

// The FP32 product row (24 floats = 6 pieces) of lane's camera c, through coop_fetch /
// coop_exchange, widened to f64.
constexpr int kCoopStride = 7;  // LDS Vec4 per row for R = 6
__device__ __forceinline__ void coop_prow(const float* __restrict__ prow, int c, int lane,
                                          Vec4<float>* __restrict__ lds, double ct[16], double f[kFwd]) {
  Vec4<float> v[6], u[6];
  coop_fetch<6>(reinterpret_cast<const Vec4<float>*>(prow), kPRow / 4, c, lane, v);
  coop_exchange<6>(v, lane, lds, u);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ct[4 * k] = double(u[k].x); ct[4 * k + 1] = double(u[k].y); ct[4 * k + 2] = double(u[k].z);
    ct[4 * k + 3] = double(u[k].w);
  }
  ct[12] = double(u[3].x); ct[13] = double(u[3].y); ct[14] = double(u[3].z); ct[15] = 0.0;
  f[0] = double(u[3].w);
  f[1] = double(u[4].x); f[2] = double(u[4].y); f[3] = double(u[4].z); f[4] = double(u[4].w);
  f[5] = double(u[5].x); f[6] = double(u[5].y); f[7] = double(u[5].z); f[8] = double(u[5].w);
  f[9] = f[10] = f[11] = 0.0;
}

// The wave's 64 FP32 product rows (6 pieces of 16 bytes each) straight from L2 into its 6 KB
// LDS slice with global_load_lds: round k moves pieces 64k .. 64k + 63, lane m fetching piece
// P = 64k + m, i.e. piece P % 6 of the row of lane P / 6 (whose camera a lane shuffle
// provides), and the hardware stores it at slice + 16 P, so row r lands at slice + 96 r. No
// registers stage the data and no LDS store instructions are issued. Rows of empty slots
// (camera -1) fetch camera 0's row, which no lane reads. Every lane of the wave must call this.
// The waits: the previous row's LDS reads complete before the new data may land (lgkmcnt),
// and the new data has landed before it is read (vmcnt).
constexpr int kRowPieces = kPRow / 4;  // 6
// The camera of the row each of lane's pieces belongs to (lane shuffles), camera -1 as 0.
__device__ __forceinline__ void piece_cameras(int c, int lane, int cr[kRowPieces]) {
#pragma unroll
  for (int k = 0; k < kRowPieces; ++k) cr[k] = max(__shfl(c, (kWave * k + lane) / kRowPieces, kWave), 0);
}
__device__ __forceinline__ void rows_to_lds_dma(const Vec4<float>* __restrict__ prow, const int cr[kRowPieces],
                                                int lane, Vec4<float>* __restrict__ lds) {
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
#pragma unroll
  for (int k = 0; k < kRowPieces; ++k) {
    const int piece = kWave * k + lane;
    const int r = piece / kRowPieces;
    __builtin_amdgcn_global_load_lds(
        reinterpret_cast<const void*>(prow + __umul24(cr[k], kRowPieces) + (piece - kRowPieces * r)),
        (__attribute__((address_space(3))) void*)(lds + kWave * k), 16, 0, 0);
  }
}
// The wait for rows_to_lds_dma (and every other vector-memory load issued so far).
__device__ __forceinline__ void rows_landed() {
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  __builtin_amdgcn_wave_barrier();
}

// -1/z from the hardware reciprocal refined by two Newton steps (within an ulp of the division,
// which the compiler expands to eleven instructions). z > 0 for every linearized row.
__device__ __forceinline__ double neg_inv_depth(double z) {
  double r = __builtin_amdgcn_rcp(z);
  double e = __builtin_fma(-z, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-z, r, 1.0);
  return -__builtin_fma(r, e, r);
}

// w += Jp^T Jc z for one observation of the f32 Schur product, its camera's product row read
// from LDS piece by piece as the computation needs it (pieces 0-2: rows of R with t_i; 3: f,
// k1, k2, fw0; 4, 5: fw1-8), which keeps fewer values live than the whole row in f64 registers.
// With the projection derivative dp = iz A P of schur_camera_row (iz = -1/z, A = f (radial I +
// c2 q q^T), P = [I | q]) and v = w' x (P - t) + z_t: u = Jc z = sw v', v' = iz A P v + q s, and
// Jp^T u = R^T e with e = sw^2 iz P^T A v'; f64 arithmetic.
__device__ __forceinline__ void schur_obs_lds(const Vec4<float>* __restrict__ row, const Vec4<float>& rec,
                                              double w[3]) {
  const double qx = rec.x, qy = rec.y, z = rec.z, sw = rec.w;
  const Vec4<float> p3 = row[3];
  const double f = p3.x, k1 = p3.y, k2 = p3.z;
  const double r2 = __builtin_fma(qx, qx, qy * qy);
  const double radial = __builtin_fma(r2, __builtin_fma(k2, r2, k1), 1.0);
  const double fc2 = f * __builtin_fma(4.0 * k2, r2, 2.0 * k1);
  const double frad = f * radial;
  const double cx = fc2 * qx, cy = fc2 * qy;
  const double a = __builtin_fma(cx, qx, frad), b = cx * qy, d = __builtin_fma(cy, qy, frad);
  const double iz = neg_inv_depth(z);
  const Vec4<float> p0 = row[0], p1 = row[1], p2 = row[2];
  const double pt0 = __builtin_fma(-qx, z, -double(p0.w));
  const double pt1 = __builtin_fma(-qy, z, -double(p1.w));
  const double pt2 = z - double(p2.w);
  const Vec4<float> p4 = row[4], p5 = row[5];
  const double f0 = p3.w, f1 = p4.x, f2 = p4.y, f3 = p4.z, f4 = p4.w, f5 = p5.x, f6 = p5.y, f7 = p5.z,
               f8 = p5.w;
  const double x0 = __builtin_fma(f1, pt2, __builtin_fma(-f2, pt1, f3));
  const double x1 = __builtin_fma(f2, pt0, __builtin_fma(-f0, pt2, f4));
  const double x2 = __builtin_fma(f0, pt1, __builtin_fma(-f1, pt0, f5));
  const double fr2 = f * r2, fr4 = fr2 * r2;
  const double sv = __builtin_fma(radial, f6, __builtin_fma(fr2, f7, fr4 * f8));
  const double y0 = __builtin_fma(qx, x2, x0), y1 = __builtin_fma(qy, x2, x1);
  const double v0 = __builtin_fma(iz, __builtin_fma(a, y0, b * y1), qx * sv);
  const double v1 = __builtin_fma(iz, __builtin_fma(b, y0, d * y1), qy * sv);
  const double h0 = __builtin_fma(a, v0, b * v1), h1 = __builtin_fma(b, v0, d * v1);
  const double wz = sw * sw * iz;
  const double e0 = wz * h0, e1 = wz * h1, e2 = wz * __builtin_fma(qx, h0, qy * h1);
  w[0] = __builtin_fma(double(p0.x), e0, __builtin_fma(double(p1.x), e1, __builtin_fma(double(p2.x), e2, w[0])));
  w[1] = __builtin_fma(double(p0.y), e0, __builtin_fma(double(p1.y), e1, __builtin_fma(double(p2.y), e2, w[1])));
  w[2] = __builtin_fma(double(p0.z), e0, __builtin_fma(double(p1.z), e1, __builtin_fma(double(p2.z), e2, w[2])));
}

// The attempt check's sums for one observation of the f32 pass (kCheck), its camera's product
// row (the forward table of the solution z) read from LDS as in schur_obs_lds: with u = Jc z =
// sw v' and the weighted residual r of the row,
//   w += Jp^T u = R^T e,  e = sw^2 iz P^T A v';  uu += ||u||^2;  gr += Jp^T r = R^T e',
//   e' = sw iz P^T A r;
// f64 arithmetic.
__device__ __forceinline__ void check_obs_lds(const Vec4<float>* __restrict__ row, const Vec4<float>& rec,
                                              const Vec2<float>& rr, double w[3], double& uu, double gr[3]) {
  const double qx = rec.x, qy = rec.y, z = rec.z, sw = rec.w;
  const Vec4<float> p3 = row[3];
  const double f = p3.x, k1 = p3.y, k2 = p3.z;
  const double r2 = __builtin_fma(qx, qx, qy * qy);
  const double radial = __builtin_fma(r2, __builtin_fma(k2, r2, k1), 1.0);
  const double fc2 = f * __builtin_fma(4.0 * k2, r2, 2.0 * k1);
  const double frad = f * radial;
  const double cx = fc2 * qx, cy = fc2 * qy;
  const double a = __builtin_fma(cx, qx, frad), b = cx * qy, d = __builtin_fma(cy, qy, frad);
  const double iz = neg_inv_depth(z);
  const Vec4<float> p0 = row[0], p1 = row[1], p2 = row[2];
  const double pt0 = __builtin_fma(-qx, z, -double(p0.w));
  const double pt1 = __builtin_fma(-qy, z, -double(p1.w));
  const double pt2 = z - double(p2.w);
  const Vec4<float> p4 = row[4], p5 = row[5];
  const double f0 = p3.w, f1 = p4.x, f2 = p4.y, f3 = p4.z, f4 = p4.w, f5 = p5.x, f6 = p5.y, f7 = p5.z,
               f8 = p5.w;
  const double x0 = __builtin_fma(f1, pt2, __builtin_fma(-f2, pt1, f3));
  const double x1 = __builtin_fma(f2, pt0, __builtin_fma(-f0, pt2, f4));
  const double x2 = __builtin_fma(f0, pt1, __builtin_fma(-f1, pt0, f5));
  const double fr2 = f * r2, fr4 = fr2 * r2;
  const double sv = __builtin_fma(radial, f6, __builtin_fma(fr2, f7, fr4 * f8));
  const double y0 = __builtin_fma(qx, x2, x0), y1 = __builtin_fma(qy, x2, x1);
  const double v0 = __builtin_fma(iz, __builtin_fma(a, y0, b * y1), qx * sv);
  const double v1 = __builtin_fma(iz, __builtin_fma(b, y0, d * y1), qy * sv);
  const double h0 = __builtin_fma(a, v0, b * v1), h1 = __builtin_fma(b, v0, d * v1);
  const double w2 = sw * sw;
  const double wz = w2 * iz;
  const double e0 = wz * h0, e1 = wz * h1, e2 = wz * __builtin_fma(qx, h0, qy * h1);
  uu = __builtin_fma(w2, __builtin_fma(v0, v0, v1 * v1), uu);
  const double rx = rr.x, ry = rr.y;
  const double g0 = __builtin_fma(a, rx, b * ry), g1 = __builtin_fma(b, rx, d * ry);
  const double sz = sw * iz;
  const double q0 = sz * g0, q1 = sz * g1, q2 = sz * __builtin_fma(qx, g0, qy * g1);
  w[0] = __builtin_fma(double(p0.x), e0, __builtin_fma(double(p1.x), e1, __builtin_fma(double(p2.x), e2, w[0])));
  w[1] = __builtin_fma(double(p0.y), e0, __builtin_fma(double(p1.y), e1, __builtin_fma(double(p2.y), e2, w[1])));
  w[2] = __builtin_fma(double(p0.z), e0, __builtin_fma(double(p1.z), e1, __builtin_fma(double(p2.z), e2, w[2])));
  gr[0] = __builtin_fma(double(p0.x), q0, __builtin_fma(double(p1.x), q1, __builtin_fma(double(p2.x), q2, gr[0])));
  gr[1] = __builtin_fma(double(p0.y), q0, __builtin_fma(double(p1.y), q1, __builtin_fma(double(p2.y), q2, gr[1])));
  gr[2] = __builtin_fma(double(p0.z), q0, __builtin_fma(double(p1.z), q1, __builtin_fma(double(p2.z), q2, gr[2])));
}

// The f32 attempt check's sweep (kCheck): the rows of the solution's forward table move into
// LDS with global_load_lds and the next step's camera index, record and weighted residual are
// fetched behind them, as in the Schur product's sweep (register sets A and B; loads past the
// chunk's last step re-read it, so exactly three loads follow the rows and vmcnt(3) is "rows
// landed"). Accumulates w = sum Jp^T Jc z, uu = sum ||Jc z||^2 and gr = sum Jp^T r; the point's
// Hpp and its damped factor come from the linearization and the attempt's factorization
// instead of being re-formed here. Every lane of the wave must call this.
__device__ __forceinline__ void check_sweep_dma(const PointArgs& pa, int chunk, int lane, int j0, int j1,
                                                const Vec4<float>* __restrict__ rec,
                                                const Vec2<float>* __restrict__ res, const float* __restrict__ fw,
                                                double w[3], double& uu, double gr[3],
                                                Vec4<float>* __restrict__ coop) {
  using V4f = float __attribute__((ext_vector_type(4)));
  using V2f = float __attribute__((ext_vector_type(2)));
  using GV = const __attribute__((address_space(1))) V4f;
  using GV2 = const __attribute__((address_space(1))) V2f;
  using GI = const __attribute__((address_space(1))) int;
  GV* rcv = (GV*)(rec);
  GV2* rsv = (GV2*)(res);
  GI* scv = (GI*)(pa.slot_cam);
  __asm__ volatile("" : "+s"(rcv), "+s"(rsv), "+s"(scv));
  const Vec4<float>* pr = reinterpret_cast<const Vec4<float>*>(fw);
  if (j0 >= j1) return;  // uniform over the wave
  const int base = pa.chunk_base[chunk];
  const int last = base + (j1 - 1) * kWave + lane;
  auto slot = [&](int j) { return min(base + j * kWave + lane, last); };
  int cr[kRowPieces];
  int c_a = scv[slot(j0)], c_b;
  V4f r_a = rcv[slot(j0)], r_b;
  V2f q_a = rsv[slot(j0)], q_b;
  piece_cameras(c_a, lane, cr);
  auto step = [&](int j, int c, const V4f& r, const V2f& q, int& c_next, V4f& r_next, V2f& q_next) {
    rows_to_lds_dma(pr, cr, lane, coop);
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    c_next = scv[slot(j + 1)];
    r_next = rcv[slot(j + 1)];
    q_next = rsv[slot(j + 1)];
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(0x0F73);  // vmcnt(3): the rows and every load before them
    __builtin_amdgcn_wave_barrier();
    if (c >= 0)
      check_obs_lds(coop + lane * kRowPieces, Vec4<float>{r.x, r.y, r.z, r.w}, Vec2<float>{q.x, q.y}, w, uu, gr);
    piece_cameras(c_next, lane, cr);
  };
  for (int j = j0;; j += 2) {
    step(j, c_a, r_a, q_a, c_b, r_b, q_b);
    if (j + 1 >= j1) break;
    step(j + 1, c_b, r_b, q_b, c_a, r_a, q_a);
    if (j + 2 >= j1) break;
  }
}

// One lane's observations of a point-major pass (SELL chunk, two observations per step so
// that both loads issue before either is used). With COOP (FP32 product rows) the rows come
// through coop_prow, and every lane of the wave must take part.
template <typename T, int MODE, bool COOP = false>
__device__ __forceinline__ void point_sweep(const PointArgs& pa, int chunk, int lane, int j0, int j1,
                                            const Vec4<T>* __restrict__ rec, const Vec2<T>* __restrict__ res,
                                            const T* __restrict__ tab, const T* __restrict__ fw, const double dpx[3],
                                            const double* __restrict__ rows_in, double* __restrict__ rows_out,
                                            double w[3], double h[6], double& uu, double coef, double gr[3],
                                            Vec4<float>* __restrict__ coop = nullptr) {
  const int base = pa.chunk_base[chunk];
  if constexpr (COOP && MODE == kSchur && sizeof(T) == 4) {
    // Step j's rows travel from L2 into LDS while the camera index and record of step j + 1 are
    // fetched; the wait before step j's arithmetic covers the rows only (the two loads issued
    // after them stay in flight through the arithmetic), and step j + 1's camera shuffles come
    // after step j's arithmetic. Two register sets A (even steps) and B (odd steps) carry the
    // index and record, so no copy waits on a load in flight. Loads past the chunk's last step
    // re-read its last step (in bounds, unused), so exactly two loads follow the rows and
    // vmcnt(2) is "rows landed". The index and record loads go through pointer copies the
    // compiler cannot trace to the restrict-qualified arguments, so the compiler barriers keep
    // them after the row fetch.
    using V4f = float __attribute__((ext_vector_type(4)));
    using GV = const __attribute__((address_space(1))) V4f;
    using GI = const __attribute__((address_space(1))) int;
    GV* rcv = (GV*)(rec);
    GI* scv = (GI*)(pa.slot_cam);
    __asm__ volatile("" : "+s"(rcv), "+s"(scv));
    const Vec4<float>* pr = reinterpret_cast<const Vec4<float>*>(fw);
    if (j0 >= j1) return;  // uniform over the wave
    const int last = base + (j1 - 1) * kWave + lane;
    auto slot = [&](int j) { return min(base + j * kWave + lane, last); };
    int cr[kRowPieces];
    int c_a = scv[slot(j0)], c_b;
    V4f r_a = rcv[slot(j0)], r_b;
    piece_cameras(c_a, lane, cr);
    auto step = [&](int j, int c, const V4f& r, int& c_next, V4f& r_next) {
      rows_to_lds_dma(pr, cr, lane, coop);
      __asm__ volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
      c_next = scv[slot(j + 1)];
      r_next = rcv[slot(j + 1)];
      __asm__ volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
      __builtin_amdgcn_s_waitcnt(0x0F72);  // vmcnt(2): the rows and every load before them
      __builtin_amdgcn_wave_barrier();
      if (c >= 0) schur_obs_lds(coop + lane * kRowPieces, Vec4<float>{r.x, r.y, r.z, r.w}, w);
      piece_cameras(c_next, lane, cr);
    };
    for (int j = j0;; j += 2) {
      step(j, c_a, r_a, c_b, r_b);
      if (j + 1 >= j1) break;
      step(j + 1, c_b, r_b, c_a, r_a);
      if (j + 2 >= j1) break;
    }
    return;
  } else if constexpr (COOP) {
    const float* pr = reinterpret_cast<const float*>(fw);
    // one observation per step: two in flight cost 40 more VGPRs and a wave per SIMD, and
    // measured slower
    for (int j = j0; j < j1; ++j) {
      const int s0 = base + j * kWave + lane;
      const int c0 = pa.slot_cam[s0];
      const Vec4<T> r0 = rec[s0];
      double ct[16], f[kFwd];
      coop_prow(pr, c0, lane, coop, ct, f);
      if (c0 >= 0) point_obs_row<T, MODE>(pa, s0, r0, res, ct, f, true, dpx, rows_in, rows_out, w, h, uu, coef, gr);
    }
    return;
  }
  int j = j0;
  for (; j + 1 < j1; j += 2) {
    const int s0 = base + j * kWave + lane, s1 = s0 + kWave;
    const int c0 = pa.slot_cam[s0], c1 = pa.slot_cam[s1];
    const Vec4<T> r0 = rec[s0], r1 = rec[s1];
    if (c0 >= 0) point_obs<T, MODE>(pa, s0, c0, r0, res, tab, fw, dpx, rows_in, rows_out, w, h, uu, coef, gr);
    if (c1 >= 0) point_obs<T, MODE>(pa, s1, c1, r1, res, tab, fw, dpx, rows_in, rows_out, w, h, uu, coef, gr);
  }
  if (j < j1) {
    const int s0 = base + j * kWave + lane;
    const int c0 = pa.slot_cam[s0];
    if (c0 >= 0) point_obs<T, MODE>(pa, s0, c0, rec[s0], res, tab, fw, dpx, rows_in, rows_out, w, h, uu, coef, gr);
  }
}

// Waves per SIMD the point-major kernel is compiled for: the f32 Schur product is held to 96
// VGPRs for five (its LDS row exchange, 30 KB per workgroup, also admits five); the other
// modes keep the compiler's default.
template <typename T, int MODE>
constexpr int point_waves() { return 1; }

template <typename T, int MODE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(point_waves<T, MODE>()))) k_points(
    PointArgs pa, const Vec4<T>* __restrict__ rec, const Vec2<T>* __restrict__ res, const T* __restrict__ tab,
    const T* __restrict__ fw, const double* __restrict__ mp, const T* __restrict__ gp,
    const T* __restrict__ dpv, const double* __restrict__ delta_p_in, const double* __restrict__ rows_in,
    double lambda, double* __restrict__ t_out, double* __restrict__ delta_out, double* __restrict__ rows_out,
    double* parts, const CgState* st, const LsmrState* ls, const double* __restrict__ lpp,
    const double* __restrict__ hpp) {
  __shared__ double smem[kBlock / kWave];
  if (MODE == kSchur && st != nullptr) {
    if (st->done) return;
    lambda = st->lambda;  // inside PCG the damping travels with the state
  }
  if (ls != nullptr && ls->done) return;  // a finished LSMR solve's queued tail
  // kLsmr: U update coefficient alpha/beta and sqrt(lambda) of the damping rows
  const double coef = MODE == kLsmr ? ls->alpha / ls->beta : 0.0;
  const double root = MODE == kLsmr ? ls->root : 0.0;
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  // The first pa.nsplit chunks (the longest tracks) take a whole workgroup each: its four waves
  // sweep consecutive quarters of the chunk's observation steps, add their lanes' sums in LDS
  // in wave order, and wave 0 finishes the points.
  const bool split = (int)blockIdx.x < pa.nsplit;
  const int chunk = split ? (int)blockIdx.x : pa.nsplit + ((int)blockIdx.x - pa.nsplit) * kChunksPerBlock + wave;
  const int i = chunk * kWave + lane;
  const bool live = chunk < pa.nchunks && i < pa.np;
  double uu = 0.0, n2 = 0.0, n2s = 0.0, gd = 0.0, g2 = 0.0, g2s = 0.0;  // verify partials
  double w[3] = {0.0, 0.0, 0.0};
  double h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double gr[3] = {0.0, 0.0, 0.0};  // kCheck: sum Jp^T r
  double dpx[3] = {0.0, 0.0, 0.0};
  if (live && (MODE == kVerify || MODE == kJv || (MODE == kLsmr && fw != nullptr))) {
    dpx[0] = delta_p_in[3 * i];
    dpx[1] = delta_p_in[3 * i + 1];
    dpx[2] = delta_p_in[3 * i + 2];
  }
  // kCheck's one sweep: w = sum Jp^T Jc z and Hpp re-formed as kSchur, plus sum ||Jc z||^2
  // (into uu) and sum Jp^T r
  constexpr int kSweep = MODE == kCheck ? kSchurSums : MODE;
  // FP32 product rows are gathered cooperatively by the whole wave (lanes past np sweep their
  // chunk's empty slots, whose camera is -1); not in kCheck, whose larger sweep went from 96 to
  // 132 VGPRs with it and ran slower (63.8 against 56.2 us per launch)
  constexpr bool kCoop = sizeof(T) == 4 && point_fwd_mode<kSweep>() && MODE != kCheck;
  // the f32 check gathers its rows with global_load_lds like kSchur (check_sweep_dma), and takes
  // the point's Hpp and damped factor from memory
  constexpr bool kDmaCheck = sizeof(T) == 4 && MODE == kCheck;
  // kSchur's and the check's DMA rows need 6 pieces per row, the register exchange of the other
  // modes 7
  constexpr int kCoopRow = (MODE == kSchur || kDmaCheck) ? kRowPieces : kCoopStride;
  __shared__ Vec4<float> coop[(kCoop || kDmaCheck) ? kChunksPerBlock * kWave * kCoopRow : 1];
  if (MODE != kRhs && chunk < pa.nchunks) {
    const int len = pa.chunk_len[chunk];
    const int j0 = split ? len * wave / kChunksPerBlock : 0;
    const int j1 = split ? len * (wave + 1) / kChunksPerBlock : len;
    if constexpr (kCoop && MODE == kSchur) {
      // every Schur product names its forward table; only this sweep is compiled, so the
      // register budget is the sweep's own
      point_sweep<T, kSweep, kCoop>(pa, chunk, lane, j0, j1, rec, res, tab, fw, dpx, rows_in, rows_out, w, h, uu,
                                    coef, gr, coop + wave * kWave * kCoopRow);
    } else if constexpr (kDmaCheck) {
      check_sweep_dma(pa, chunk, lane, j0, j1, reinterpret_cast<const Vec4<float>*>(rec),
                      reinterpret_cast<const Vec2<float>*>(res), reinterpret_cast<const float*>(fw), w, uu, gr,
                      coop + wave * kWave * kCoopRow);
    } else if (kCoop && fw != nullptr) {
      point_sweep<T, kSweep, kCoop>(pa, chunk, lane, j0, j1, rec, res, tab, fw, dpx, rows_in, rows_out, w, h, uu,
                                    coef, gr, coop + wave * kWave * kCoopRow);
    } else if (live) {
      point_sweep<T, kSweep>(pa, chunk, lane, j0, j1, rec, res, tab, fw, dpx, rows_in, rows_out, w, h, uu, coef,
                             gr);
    }
  }
  // sums a split chunk's waves must combine before the points are finished (uu is additive
  // over the block and goes to the block sum as it is)
  constexpr int kMeet = MODE == kCheck ? 12 : (MODE == kRhs || MODE == kJv) ? 0 : 3;
  if (kMeet > 0 && split) {
    __shared__ double meet[kMeet > 0 ? kMeet * kWave : 1];
    double v[12];
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = w[a];
#pragma unroll
    for (int a = 0; a < 6; ++a) v[3 + a] = h[a];
#pragma unroll
    for (int a = 0; a < 3; ++a) v[9 + a] = gr[a];
    for (int src = 1; src < kChunksPerBlock; ++src) {
      __syncthreads();
      if (wave == src) {
#pragma unroll
        for (int a = 0; a < kMeet; ++a) meet[a * kWave + lane] = v[a];
      }
      __syncthreads();
      if (wave == 0) {
#pragma unroll
        for (int a = 0; a < kMeet; ++a) v[a] += meet[a * kWave + lane];
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) w[a] = v[a];
#pragma unroll
    for (int a = 0; a < 6; ++a) h[a] = v[3 + a];
#pragma unroll
    for (int a = 0; a < 3; ++a) gr[a] = v[9 + a];
  }
  if (live && (!split || wave == 0)) {
    if (MODE == kCheck) {
      // The verification of delta = (z / D, delta_p) follows from the sweep's sums without
      // a second sweep:
      //   ||J delta||^2 over the point's rows = sum ||Jc z||^2 + 2 delta_p.w + delta_p^T Hpp delta_p
      //   Jp^T (J delta + r)                  = w + Hpp delta_p + sum Jp^T r
      // (f32: Hpp and the factor of Hpp + lambda Dp^2 as the linearization and the attempt's
      // factorization stored them, formed from the same rows in the same order)
      double hu[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) hu[k] = kDmaCheck ? hpp[6 * i + k] : h[k];  // undamped Hpp
      double t[3] = {w[0], w[1], w[2]};
      if constexpr (kDmaCheck) {
#pragma unroll
        for (int k = 0; k < 6; ++k) h[k] = lpp[6 * i + k];
      } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const double d = double(dpv[3 * i + a]);
          h[sym3(a, a)] += lambda * d * d;
        }
      }
      if (kDmaCheck ? h[0] != 0.0 : chol_packed<double, 3>(h)) {
        solve_lower_packed<double, 3>(h, t);
        solve_upper_packed<double, 3>(h, t);
      } else {
        t[0] = t[1] = t[2] = 0.0;
      }
      // f32: t and delta_p side by side in one 32-byte row per point (t_out holds 2 np Vec4),
      // so the check's camera pass gathers one line per row
      const Vec4<float> t4{float(t[0]), float(t[1]), float(t[2]), 0.f};
      if constexpr (!kDmaCheck) store_tp<T>(t_out, i, t);
      const double v[3] = {double(gp[3 * i]) + w[0], double(gp[3 * i + 1]) + w[1], double(gp[3 * i + 2]) + w[2]};
      point_solve(mp, dpv, i, v, t);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        dpx[a] = -t[a];
        delta_out[3 * i + a] = dpx[a];
      }
      if constexpr (kDmaCheck) {
        Vec4<float>* row = reinterpret_cast<Vec4<float>*>(t_out) + 2 * i;
        row[0] = t4;
        row[1] = Vec4<float>{float(dpx[0]), float(dpx[1]), float(dpx[2]), 0.f};
      }
      double hd[3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
        hd[a] = hu[sym3(a, 0)] * dpx[0] + hu[sym3(a, 1)] * dpx[1] + hu[sym3(a, 2)] * dpx[2];
      uu += 2.0 * (dpx[0] * w[0] + dpx[1] * w[1] + dpx[2] * w[2]) +
            (dpx[0] * hd[0] + dpx[1] * hd[1] + dpx[2] * hd[2]);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double d = double(dpv[3 * i + a]);
        const double nr = (w[a] + hd[a] + gr[a]) + (lambda * (d * dpx[a])) * d;
        n2 += nr * nr;
        n2s += (nr / d) * (nr / d);
        const double g = double(gp[3 * i + a]);
        gd += g * dpx[a];
        g2 += g * g;
        g2s += (g / d) * (g / d);
      }
    } else if (MODE == kSchur) {
      // t = (Hpp + lambda Dp^2)^-1 w with the unscaled factor of this attempt: Hpp is the f64
      // Gram of the Jp this pass re-forms (accumulated at linearization in the same order), so
      // the eliminated projection stays exact for it (a factor of the scaled block C + lambda
      // I, rounded differently, slowed tight PCG solves 2x)
      double t[3] = {w[0], w[1], w[2]};
      double l[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) l[k] = lpp[6 * i + k];
      if (l[0] != 0.0) {
        solve_lower_packed<double, 3>(l, t);
        solve_upper_packed<double, 3>(l, t);
      } else {
        t[0] = t[1] = t[2] = 0.0;  // unobserved point at lambda = 0: no elimination term
      }
      store_tp<T>(t_out, i, t);
    } else if (MODE == kRhs || MODE == kRecover) {
      double v[3];
      if (MODE == kRhs) {
        v[0] = double(gp[3 * i]); v[1] = double(gp[3 * i + 1]); v[2] = double(gp[3 * i + 2]);
      } else {
        v[0] = double(gp[3 * i]) + w[0]; v[1] = double(gp[3 * i + 1]) + w[1]; v[2] = double(gp[3 * i + 2]) + w[2];
      }
      double t[3];
      point_solve(mp, dpv, i, v, t);
      if (MODE == kRecover) {
#pragma unroll
        for (int a = 0; a < 3; ++a) delta_out[3 * i + a] = -t[a];
      } else {
        store_tp<T>(t_out, i, t);
      }
    } else if (MODE == kVerify) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double d = double(dpv[3 * i + a]);
        const double nr = w[a] + (lambda * (d * dpx[a])) * d;
        if (delta_out) delta_out[3 * i + a] = nr;  // the normal residual itself, on request
        n2 += nr * nr;
        n2s += (nr / d) * (nr / d);
        const double g = double(gp[3 * i + a]);
        gd += g * delta_p_in[3 * i + a];
        g2 += g * g;
        g2s += (g / d) * (g / d);
      }
    } else if (MODE == kJtu) {
#pragma unroll
      for (int a = 0; a < 3; ++a) delta_out[3 * i + a] = w[a];
    } else if (MODE == kLsmr) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        delta_out[3 * i + a] = w[a];
        if (t_out != nullptr && fw != nullptr) {  // damping rows of this point
          const double ub = root * double(dpv[3 * i + a]) * dpx[a] - coef * t_out[3 * i + a];
          t_out[3 * i + a] = ub;
          uu += ub * ub;
        }
      }
    }
  }
  if (MODE == kLsmr) {
    const double b = block_sum(uu, smem);
    if (threadIdx.x == 0) parts[blockIdx.x] = b;
  }
  if (MODE == kVerify || MODE == kCheck) {
    const double v[6] = {n2, n2s, gd, g2, g2s, uu};
    for (int a = 0; a < 6; ++a) {
      const double b = block_sum(v[a], smem);
      if (threadIdx.x == 0) parts[kParts * blockIdx.x + a] = b;
    }
  }
}

// ---------------------------------------------------------------------------------------
// Camera half of the Schur product inside PCG: S p = (B + lambda I) p - E (C + lambda I)^-1 E^T p
// (Reduced::apply, schur.rs:606-652) in its factored form lambda p + Dc^-1 T(acc) with
// acc = sum over the camera's rows of the factors of Jc^T (u - Jp t), u = Jc Dc^-1 p (forward
// table of p) and t the point's solve from k_points<kSchur>; no B is assembled and the
// cancellation happens per row on 2-vectors. T(acc) maps the rotation slots by (R Jr)^T.
//
// One wave per (camera, point range): a camera's rows are grouped by the point-slot range of
// their point (point_ranges ranges), and workgroup b (four waves) takes range g = b % kR of the
// four cameras 4 (b / kR) .. + 3, so the XCD that runs it gathers the solves of that range only.
// The wave writes its share y = Dc^-1 T(acc) of camera c to out[9 sh ..] with sh = kR c + g,
// and p_c.y (plus lambda ||p_c||^2 in range 0) to parts[sh]; k_cg_update forms
// q_c = lambda p_c + sum_g y in range order.
//
// Rows are software-pipelined per lane (rows k0 + lane + 64 m), with two register sets A (even
// m) and B (odd m): while row m is evaluated, the record and solve of row m + 1 and the slot
// index of row m + 2 are in flight, so the chain index -> solve -> arithmetic is exposed once
// per wave instead of once per row. The trip count is uniform over the wave; a lane past its
// last row re-reads the wave's last row (in bounds) with weight 0, which adds exact zeros. The
// loop body has no divergent branch, so each wait covers only loads issued one row earlier.
// acc += the factors of Jc^T (Jc p~ - Jp t) for one row of the Schur product (the camera half's
// row: jc_apply, jp_apply and jct_factors of obs_geom in one pass). The row's projection
// derivative is dp = iz A P with iz = -1/z, A = f (radial I + c2 q q^T) (2x2 symmetric, entries
// a, b, d) and P = [I | q] (2x3), so with v = w' x (P - t) + z_t (the forward table of p~) and
// x = v - R t:
//   v' = iz A P x + q s          (s = radial z_i0 + f r2 z_i1 + f r2^2 z_i2; the row is sw v')
//   e = sw^2 iz P^T A v'          (= sw dp^T (sw v'))
//   acc_rot += (P - t) x e, acc_t += e, acc_i += sw^2 (q.v') (radial, f r2, f r2^2).
// ct: the camera table (R rows with t, f, k1, k2), fw: the forward table; f64 arithmetic.
__device__ __forceinline__ void schur_camera_row(double qx, double qy, double z, double sw, const double tp[3],
                                                 const double* ct, const double* fw, double acc[9]) {
  const double f = ct[tF];
  const double r2 = __builtin_fma(qx, qx, qy * qy);
  const double radial = __builtin_fma(r2, __builtin_fma(ct[tK2], r2, ct[tK1]), 1.0);
  const double fc2 = f * __builtin_fma(4.0 * ct[tK2], r2, 2.0 * ct[tK1]);
  const double frad = f * radial;
  const double cx = fc2 * qx, cy = fc2 * qy;
  const double a = __builtin_fma(cx, qx, frad), b = cx * qy, d = __builtin_fma(cy, qy, frad);
  const double iz = neg_inv_depth(z);
  const double pt0 = __builtin_fma(-qx, z, -ct[tT(0)]);
  const double pt1 = __builtin_fma(-qy, z, -ct[tT(1)]);
  const double pt2 = z - ct[tT(2)];
  // x = w' x pt + z_t - R t
  double x0 = __builtin_fma(fw[1], pt2, __builtin_fma(-fw[2], pt1, fw[3]));
  double x1 = __builtin_fma(fw[2], pt0, __builtin_fma(-fw[0], pt2, fw[4]));
  double x2 = __builtin_fma(fw[0], pt1, __builtin_fma(-fw[1], pt0, fw[5]));
  x0 = __builtin_fma(-ct[tR(0, 0)], tp[0], __builtin_fma(-ct[tR(0, 1)], tp[1], __builtin_fma(-ct[tR(0, 2)], tp[2], x0)));
  x1 = __builtin_fma(-ct[tR(1, 0)], tp[0], __builtin_fma(-ct[tR(1, 1)], tp[1], __builtin_fma(-ct[tR(1, 2)], tp[2], x1)));
  x2 = __builtin_fma(-ct[tR(2, 0)], tp[0], __builtin_fma(-ct[tR(2, 1)], tp[1], __builtin_fma(-ct[tR(2, 2)], tp[2], x2)));
  const double fr2 = f * r2, fr4 = fr2 * r2;
  const double sv = __builtin_fma(radial, fw[6], __builtin_fma(fr2, fw[7], fr4 * fw[8]));
  const double y0 = __builtin_fma(qx, x2, x0), y1 = __builtin_fma(qy, x2, x1);
  const double v0 = __builtin_fma(iz, __builtin_fma(a, y0, b * y1), qx * sv);
  const double v1 = __builtin_fma(iz, __builtin_fma(b, y0, d * y1), qy * sv);
  const double h0 = __builtin_fma(a, v0, b * v1), h1 = __builtin_fma(b, v0, d * v1);
  const double w2 = sw * sw;
  const double wz = w2 * iz;
  const double e0 = wz * h0, e1 = wz * h1, e2 = wz * __builtin_fma(qx, h0, qy * h1);
  acc[0] = __builtin_fma(pt1, e2, __builtin_fma(-pt2, e1, acc[0]));
  acc[1] = __builtin_fma(pt2, e0, __builtin_fma(-pt0, e2, acc[1]));
  acc[2] = __builtin_fma(pt0, e1, __builtin_fma(-pt1, e0, acc[2]));
  acc[3] += e0;
  acc[4] += e1;
  acc[5] += e2;
  const double si = w2 * __builtin_fma(qx, v0, qy * v1);
  acc[6] = __builtin_fma(radial, si, acc[6]);
  acc[7] = __builtin_fma(fr2, si, acc[7]);
  acc[8] = __builtin_fma(fr4, si, acc[8]);
}

template <typename T>
constexpr int schur_camera_waves() { return sizeof(T) == 4 ? 6 : 5; }
// Rows per register set: two per set measured no faster than one (30.4 against 30.1 us in FP32).
template <typename T>
constexpr int schur_rows_per_step() { return 1; }
template <typename T>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(schur_camera_waves<T>())))
k_schur_cameras(int nc, const int* __restrict__ cam_off_r, const int* __restrict__ k_pslot,
                const Vec4<T>* __restrict__ rec_cm, const double* __restrict__ tabw, const double* __restrict__ tp,
                const double* __restrict__ fww, const T* __restrict__ dc, const double* __restrict__ pin,
                double lambda, double* __restrict__ out, double* __restrict__ parts, const CgState* st) {
  constexpr int kR = point_ranges<T>();
  constexpr int kWavesPerBlock = kBlock / kWave;
  if (st != nullptr) {
    if (st->done) return;
    lambda = st->lambda;
  }
  const int lane = threadIdx.x % kWave;
  const int g = (int)blockIdx.x % kR;
  const int c = __builtin_amdgcn_readfirstlane(((int)blockIdx.x / kR) * kWavesPerBlock + (int)threadIdx.x / kWave);
  if (c >= nc) return;  // whole waves; no workgroup barrier follows
  const int sh = kR * c + g;
  const int k0 = cam_off_r[sh], k1 = cam_off_r[sh + 1];
  const double* __restrict__ ct = tabw + kCamTab * c;
  double ctl[16], fwl[kFwd];
#pragma unroll
  for (int a = 0; a < 16; ++a) ctl[a] = ct[a];
#pragma unroll
  for (int a = 0; a < kFwd; ++a) fwl[a] = fww[kFwd * c + a];
  double acc[9];
#pragma unroll
  for (int a = 0; a < 9; ++a) acc[a] = 0.0;
  // The pipelined loads go through copies of the row pointers that the compiler cannot trace
  // to the restrict-qualified arguments, so the compiler barriers in the loop also order them
  // (otherwise each load is moved next to its use a row later).
  using V4 = T __attribute__((ext_vector_type(4)));
  using GVec = const __attribute__((address_space(1))) V4;
  using GInt = const __attribute__((address_space(1))) int;
  GVec* tpv = (GVec*)(tp);
  GVec* recv = (GVec*)(rec_cm);
  GInt* psv = (GInt*)(k_pslot);
  __asm__ volatile("" : "+s"(tpv), "+s"(recv), "+s"(psv));
  // rows k0 + lane + 64 (P G + i), i < P, form group G; the lane's groups alternate between the
  // register sets A and B
  constexpr int P = schur_rows_per_step<T>();
  const int ng = ((k1 - k0 + kWave - 1) / kWave + P - 1) / P;  // uniform over the wave
  if (ng > 0) {
    const int last = k1 - 1;
    const int kk = k0 + lane;
    auto at = [&](int G, int i) { return kk + kWave * (P * G + i); };
    auto row = [&](const V4& r, const V4& t, bool valid) {
      const double x[3] = {double(t.x), double(t.y), double(t.z)};
      schur_camera_row(double(r.x), double(r.y), double(r.z), valid ? double(r.w) : 0.0, x, ctl, fwl, acc);
    };
    V4 rec_a[P], t_a[P];
    int ps_b[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
      rec_a[i] = recv[min(at(0, i), last)];
      t_a[i] = tpv[psv[min(at(0, i), last)]];
    }
#pragma unroll
    for (int i = 0; i < P; ++i) ps_b[i] = psv[min(at(1, i), last)];
    for (int G = 0;; G += 2) {
      int ps_a[P];
      V4 rec_b[P], t_b[P];
#pragma unroll
      for (int i = 0; i < P; ++i) ps_a[i] = psv[min(at(G + 2, i), last)];
#pragma unroll
      for (int i = 0; i < P; ++i) rec_b[i] = recv[min(at(G + 1, i), last)];
#pragma unroll
      for (int i = 0; i < P; ++i) t_b[i] = tpv[ps_b[i]];
      __asm__ volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < P; ++i) row(rec_a[i], t_a[i], at(G, i) < k1);
      if (G + 1 >= ng) break;
#pragma unroll
      for (int i = 0; i < P; ++i) ps_b[i] = psv[min(at(G + 3, i), last)];
#pragma unroll
      for (int i = 0; i < P; ++i) rec_a[i] = recv[min(at(G + 2, i), last)];
#pragma unroll
      for (int i = 0; i < P; ++i) t_a[i] = tpv[ps_a[i]];
      __asm__ volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < P; ++i) row(rec_b[i], t_b[i], at(G + 1, i) < k1);
      if (G + 2 >= ng) break;
    }
  }
#pragma unroll
  for (int a = 0; a < 9; ++a) acc[a] = wave_sum(acc[a]);
  if (lane == 0) {
    double y[9];
#pragma unroll
    for (int j = 0; j < 3; ++j) y[j] = ct[tRJ(0, j)] * acc[0] + ct[tRJ(1, j)] * acc[1] + ct[tRJ(2, j)] * acc[2];
#pragma unroll
    for (int j = 3; j < 9; ++j) y[j] = acc[j];
    const double* p = pin + 9 * c;
    double part = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const double yi = y[i] / double(dc[9 * c + i]);
      out[9 * sh + i] = yi;
      part += p[i] * yi;
      if (g == 0) part += lambda * (p[i] * p[i]);
    }
    parts[sh] = part;
  }
}

// ---------------------------------------------------------------------------------------
// Camera-major half (one workgroup per camera). Sums the factors of Jc^T v over the camera's
// rows and finishes the camera block (the PCG product's mode cSchur is k_schur_cameras):
//   kResid:   q = lambda p + Dc^-1 T(acc) with v = Jc Dc^-1 p - Jp t (S p in factored form),
//             minus rhs, with ||q - rhs||^2 and ||rhs||^2 partials
//   kRhs:     rhs = Dc^-1 (-gc + T(acc)),  v = Jp t
//   kVerify:  nr = T(acc) + lambda Dc^2 delta_c, v = J delta + r; norms and ||J delta||^2
//   kJtu:     x_c = T(acc), v = u of the row's reference index cm_ref[k]
//   kLsmr:    x_c = T(acc) with v = U gathered from slot order (k_pslot holds the
//             observation slots), then the camera damping rows
//             U_c = sqrt(lambda) Dc ws_c - (alpha/beta) U_c and their ||.||^2 (no ws: x_c only)
// T(acc) maps the rotation slots by (R Jr)^T.
// ---------------------------------------------------------------------------------------
//   kCheck:   cResid and cVerify of one attempt in one pass: per row u = Jc z (forward table
//             of z), v1 = u - Jp t_p for q = S z, v2 = u + Jp delta_p + r for the normal
//             residual, delta_c = z / D written; residual partials in parts[c], verification
//             partials in parts[nc + c]
// (cCheck runs as k_camera_split and k_camera_finish; finish_camera holds the camera's end of
// every mode)
enum CameraMode { cSchur = 0, cResid = 1, cRhs = 2, cVerify = 3, cJtu = 4, cLsmr = 5, cCheck = 6 };

// Waves per SIMD the camera-major kernel is compiled for: the f32 residual mode fits six (80
// VGPRs), so all of a 1,500-camera launch is resident at once; the other modes keep the
// compiler's default.
template <typename T, int MODE>
constexpr int camera_waves() { return sizeof(T) == 4 && MODE == cResid ? 6 : 1; }

// The end of a camera-major pass for camera c, from the camera's column sums s of the factors
// of Jc^T v (and, in cCheck, s2 of the second row vector): T(acc), the camera block's outputs
// and its partial sums, as the modes above describe.
template <typename T, int MODE>
__device__ __forceinline__ void finish_camera(int nc, int c, const double s[9], const double s2[9],
                                              const double* __restrict__ ct, const T* __restrict__ dc,
                                              const T* __restrict__ gc, const double* __restrict__ vin,
                                              const double* __restrict__ delta_c, const double* __restrict__ rhs,
                                              double lambda, double* __restrict__ out, double* __restrict__ out_d,
                                              double* parts, const LsmrState* ls) {
  double part0 = 0.0, part1 = 0.0, part2 = 0.0, part3 = 0.0, part4 = 0.0;
  // T(acc): rotation slots through (R Jr)^T
  double y[9];
#pragma unroll
  for (int j = 0; j < 3; ++j) y[j] = ct[tRJ(0, j)] * s[0] + ct[tRJ(1, j)] * s[1] + ct[tRJ(2, j)] * s[2];
#pragma unroll
  for (int j = 3; j < 9; ++j) y[j] = s[j];
  double d[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) d[i] = double(dc[9 * c + i]);
  if (MODE == cCheck) {
    // verification of delta = (z / D, delta_p) in original coordinates
    double y2[9];
#pragma unroll
    for (int j = 0; j < 3; ++j) y2[j] = ct[tRJ(0, j)] * s2[0] + ct[tRJ(1, j)] * s2[1] + ct[tRJ(2, j)] * s2[2];
#pragma unroll
    for (int j = 3; j < 9; ++j) y2[j] = s2[j];
    double v0s = 0.0, v1s = 0.0, v2s = 0.0, v3s = 0.0, v4s = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const double dd = d[i];
      const double dc_i = vin[9 * c + i] / dd;
      out_d[9 * c + i] = dc_i;
      const double nr = y2[i] + (lambda * (dd * dc_i)) * dd;
      v0s += nr * nr;
      v1s += (nr / dd) * (nr / dd);
      const double g = double(gc[9 * c + i]);
      v2s += g * dc_i;
      v3s += g * g;
      v4s += (g / dd) * (g / dd);
    }
    double* pv = parts + kParts * (nc + c);
    pv[0] = v0s;
    pv[1] = v1s;
    pv[2] = v2s;
    pv[3] = v3s;
    pv[4] = v4s;
    pv[5] = 0.0;
  }
  if (MODE == cResid || MODE == cCheck) {
    const double* p = vin + 9 * c;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const double q = lambda * p[i] + y[i] / d[i];
      out[9 * c + i] = q;
      const double e = q - rhs[9 * c + i];
      part0 += e * e;
      part1 += rhs[9 * c + i] * rhs[9 * c + i];
    }
  } else if (MODE == cRhs) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const double r = (-double(gc[9 * c + i]) + y[i]) / d[i];
      out[9 * c + i] = r;
      part0 += r * r;
    }
  } else if (MODE == cVerify) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const double dd = d[i];
      const double nr = y[i] + (lambda * (dd * delta_c[9 * c + i])) * dd;
      if (out_d) out_d[9 * c + i] = nr;  // the normal residual itself, on request
      part0 += nr * nr;
      part1 += (nr / dd) * (nr / dd);
      const double g = double(gc[9 * c + i]);
      part2 += g * delta_c[9 * c + i];
      part3 += g * g;
      part4 += (g / dd) * (g / dd);
    }
  } else if (MODE == cJtu) {
#pragma unroll
    for (int i = 0; i < 9; ++i) out_d[9 * c + i] = y[i];
  } else if (MODE == cLsmr) {
    const double coef = ls->alpha / ls->beta, root = ls->root;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      out_d[9 * c + i] = y[i];
      if (vin != nullptr && out != nullptr) {  // damping rows of this camera
        const double ub = root * d[i] * vin[9 * c + i] - coef * out[9 * c + i];
        out[9 * c + i] = ub;
        part0 += ub * ub;
      }
    }
    parts[c] = part0;
  }
  if (MODE != cJtu && MODE != cLsmr) {
    double* pc = parts + kParts * c;
    pc[0] = part0;
    pc[1] = part1;
    pc[2] = part2;
    pc[3] = part3;
    pc[4] = part4;
  }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(camera_waves<T, MODE>()))) k_cameras(
    int nc, const int* __restrict__ cam_off, const int* __restrict__ k_pslot, const Vec4<T>* __restrict__ rec_cm,
    const Vec2<T>* __restrict__ res_cm, const double* __restrict__ tabw, const double* __restrict__ tp,
    const double* __restrict__ delta_p, const double* __restrict__ rows_in, const double* __restrict__ fww,
    const T* __restrict__ dc, const T* __restrict__ gc, const double* __restrict__ vin,
    const double* __restrict__ delta_c, const double* __restrict__ rhs, double lambda, double* __restrict__ out,
    double* __restrict__ out_d, double* parts, const CgState* st, const LsmrState* ls,
    const int* __restrict__ cm_ref) {
  static_assert(MODE != cSchur && MODE != cCheck, "cSchur is k_schur_cameras, cCheck k_camera_split");
  constexpr int kB = kBlock;
  __shared__ double red[9][kB / kWave];
  (void)st;
  if (ls != nullptr && ls->done) return;
  const int c = (int)blockIdx.x;
  // the camera's table and forward table as f64 rows (the T values widened): uniform over the
  // workgroup, so they are read with scalar loads and stay out of the vector registers
  const double* __restrict__ ct = tabw + kCamTab * c;
  const double* __restrict__ cfw = fww + kFwd * c;
  double acc[9];
#pragma unroll
  for (int a = 0; a < 9; ++a) acc[a] = 0.0;
  const int k0 = cam_off[c];
  const int k1 = cam_off[c + 1];
  auto one = [&](int k, const Vec4<T>& rec, const Vec4<double>& tv) {
    ObsGeom<double> geo;
    obs_geom(to_d(rec), ct, geo);
    double v0, v1;
    if (MODE == cJtu) {
      const int row = cm_ref[k];
      v0 = rows_in[2 * row];
      v1 = rows_in[2 * row + 1];
    } else if (MODE == cLsmr) {
      v0 = tv.x;
      v1 = tv.y;
    } else if (MODE == cVerify) {
      const double x[3] = {tv.x, tv.y, tv.z};
      double u0, u1, w0, w1;
      jc_apply(geo, cfw, u0, u1);
      jp_apply(geo, ct, x, w0, w1);
      const Vec2<T> r = res_cm[k];
      v0 = u0 + w0 + double(r.x);
      v1 = u1 + w1 + double(r.y);
    } else if (MODE == cResid) {
      const double x[3] = {tv.x, tv.y, tv.z};
      double u0, u1, w0, w1;
      jc_apply(geo, cfw, u0, u1);
      jp_apply(geo, ct, x, w0, w1);
      v0 = u0 - w0;
      v1 = u1 - w1;
    } else {
      const double x[3] = {tv.x, tv.y, tv.z};
      jp_apply(geo, ct, x, v0, v1);
    }
    jct_factors(geo, v0, v1, acc);
  };
  auto gather = [&](int k) -> Vec4<double> {
    if (MODE == cJtu) return Vec4<double>{0.0, 0.0, 0.0, 0.0};
    const int ps = k_pslot[k];
    if (MODE == cLsmr) {
      const Vec2<double> u = reinterpret_cast<const Vec2<double>*>(rows_in)[ps];
      return Vec4<double>{u.x, u.y, 0.0, 0.0};
    }
    if (MODE == cVerify) return Vec4<double>{delta_p[3 * ps], delta_p[3 * ps + 1], delta_p[3 * ps + 2], 0.0};
    return load_tp<T>(tp, ps);
  };
  int k = k0 + threadIdx.x;
  // two rows per step: records and gathers of both issue before either is used
  for (; k + kB < k1; k += 2 * kB) {
    const Vec4<T> ra = rec_cm[k], rb = rec_cm[k + kB];
    const Vec4<double> ta = gather(k), tb = gather(k + kB);
    one(k, ra, ta);
    one(k + kB, rb, tb);
  }
  if (k < k1) one(k, rec_cm[k], gather(k));
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    const double v = wave_sum(acc[a]);
    if (lane == 0) red[a][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s[9];
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      s[a] = 0.0;
      for (int w = 0; w < kB / kWave; ++w) s[a] += red[a][w];
    }
    finish_camera<T, MODE>(nc, c, s, nullptr, ct, dc, gc, vin, delta_c, rhs, lambda, out, out_d, parts, ls);
  }
}

// ---------------------------------------------------------------------------------------
// The camera side of the attempt check (mode cCheck) split like k_schur_cameras: one wave per
// (camera, point range), four cameras of one range per workgroup, so the XCD running the
// workgroup gathers the per-point vectors of one range of points only. Per row u = Jc z
// (forward table of z), v1 = u - Jp t_p and v2 = u + Jp delta_p + r. The wave writes its share
// of the column sums of the factors of Jc^T v1 and Jc^T v2 to shares[W sh ..], W = share_width,
// sh = kR c + g, and k_camera_finish adds a camera's shares in range order and finishes the
// camera (finish_camera). Rows are pipelined per lane as in k_schur_cameras (register sets A
// and B; a lane past its last row re-reads the wave's last row with weight 0). The same split
// of the reduced right-hand side (cRhs, v = Jp t_p) measured no faster in FP32 (28.6 + 5.0
// against 32.8 us) and changed the rounding of the right-hand side enough to move a borderline
// FP32 solve at lambda = 1e-10 from convergence to the iteration limit, so cRhs stays in
// k_cameras.
template <int MODE>
constexpr int share_width() { return MODE == cCheck ? 18 : 10; }
template <typename T, int MODE>
__global__ void __launch_bounds__(kBlock) k_camera_split(
    int nc, const int* __restrict__ cam_off_r, const int* __restrict__ k_pslot, const Vec4<T>* __restrict__ rec_cm,
    const Vec2<T>* __restrict__ res_cm, const double* __restrict__ tabw, const double* __restrict__ tp,
    const double* __restrict__ delta_p, const double* __restrict__ fww, double* __restrict__ shares) {
  static_assert(MODE == cCheck, "split camera pass: cCheck");
  constexpr bool kCheck = MODE == cCheck;
  constexpr int kR = point_ranges<T>();
  constexpr int kW = share_width<MODE>();
  constexpr int kWavesPerBlock = kBlock / kWave;
  const int lane = threadIdx.x % kWave;
  const int g = (int)blockIdx.x % kR;
  const int c = __builtin_amdgcn_readfirstlane(((int)blockIdx.x / kR) * kWavesPerBlock + (int)threadIdx.x / kWave);
  if (c >= nc) return;  // whole waves; no workgroup barrier follows
  const int sh = kR * c + g;
  const int k0 = cam_off_r[sh], k1 = cam_off_r[sh + 1];
  double ctl[16], fwl[kFwd];
#pragma unroll
  for (int a = 0; a < 16; ++a) ctl[a] = tabw[kCamTab * c + a];
#pragma unroll
  for (int a = 0; a < kFwd; ++a) fwl[a] = kCheck ? fww[kFwd * c + a] : 0.0;
  double acc[9], acc2[9];
#pragma unroll
  for (int a = 0; a < 9; ++a) acc[a] = acc2[a] = 0.0;
  // pipelined loads through pointer copies the compiler cannot trace to the restrict-qualified
  // arguments, so the compiler barriers in the loop also order them
  using V4 = T __attribute__((ext_vector_type(4)));
  using V2 = T __attribute__((ext_vector_type(2)));
  using GV4 = const __attribute__((address_space(1))) V4;
  using GV2 = const __attribute__((address_space(1))) V2;
  using GD = const __attribute__((address_space(1))) double;
  using GInt = const __attribute__((address_space(1))) int;
  GV4* tpv = (GV4*)(tp);
  GV4* recv = (GV4*)(rec_cm);
  GV2* resv = (GV2*)(res_cm);
  GD* dpv = (GD*)(delta_p);
  GInt* psv = (GInt*)(k_pslot);
  __asm__ volatile("" : "+s"(tpv), "+s"(recv), "+s"(resv), "+s"(dpv), "+s"(psv));
  // f32: tp holds t and delta_p side by side, two Vec4 per point (one 32-byte line piece);
  // f64: t rows in tp, delta_p (f64) in delta_p
  constexpr bool kPair = sizeof(T) == 4;
  struct Row {
    V4 rec, t, td;
    V2 res;
    double d0, d1, d2;
  };
  const int n = (k1 - k0 + kWave - 1) / kWave;  // uniform over the wave
  if (n > 0) {
    const int last = k1 - 1;
    int k = k0 + lane;
    auto fetch = [&](int kk, int ps, Row& r) {
      r.rec = recv[kk];
      if constexpr (kPair) {
        r.t = tpv[2 * ps];
        r.td = tpv[2 * ps + 1];
      } else {
        r.t = tpv[ps];
      }
      if constexpr (kCheck) {
        r.res = resv[kk];
        if constexpr (!kPair) {
          r.d0 = dpv[3 * ps];
          r.d1 = dpv[3 * ps + 1];
          r.d2 = dpv[3 * ps + 2];
        }
      }
    };
    auto row = [&](const Row& r, bool valid) {
      ObsGeom<double> geo;
      obs_geom(Vec4<double>{double(r.rec.x), double(r.rec.y), double(r.rec.z), double(r.rec.w)}, ctl, geo);
      if (!valid) geo.sw = 0.0;
      const double xt[3] = {double(r.t.x), double(r.t.y), double(r.t.z)};
      double a0, a1;
      jp_apply(geo, ctl, xt, a0, a1);
      if constexpr (kCheck) {
        const double xd[3] = {kPair ? double(r.td.x) : r.d0, kPair ? double(r.td.y) : r.d1,
                              kPair ? double(r.td.z) : r.d2};
        double u0, u1, b0, b1;
        jc_apply(geo, fwl, u0, u1);
        jp_apply(geo, ctl, xd, b0, b1);
        jct_factors(geo, u0 - a0, u1 - a1, acc);
        jct_factors(geo, u0 + b0 + double(r.res.x), u1 + b1 + double(r.res.y), acc2);
      } else {
        jct_factors(geo, a0, a1, acc);
      }
    };
    Row ra, rb;
    fetch(min(k, last), psv[min(k, last)], ra);
    int ps_b = psv[min(k + kWave, last)];
    for (int m = 0;; m += 2) {
      const int ps_a = psv[min(k + 2 * kWave, last)];
      fetch(min(k + kWave, last), ps_b, rb);
      __asm__ volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
      row(ra, k < k1);
      if (m + 1 >= n) break;
      ps_b = psv[min(k + 3 * kWave, last)];
      fetch(min(k + 2 * kWave, last), ps_a, ra);
      __asm__ volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
      row(rb, k + kWave < k1);
      if (m + 2 >= n) break;
      k += 2 * kWave;
    }
  }
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    acc[a] = wave_sum(acc[a]);
    if (kCheck) acc2[a] = wave_sum(acc2[a]);
  }
  if (lane == 0) {
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      shares[kW * sh + a] = acc[a];
      if (kCheck) shares[kW * sh + 9 + a] = acc2[a];
    }
  }
}

// One thread per camera (single-wave workgroups, spread over more CUs): the camera's shares of
// k_camera_split in range order, then the end of the camera-major pass (finish_camera).
template <typename T, int MODE>
__global__ void __launch_bounds__(kWave) k_camera_finish(int nc, const double* __restrict__ shares,
                                                         const double* __restrict__ tabw, const T* __restrict__ dc,
                                                         const T* __restrict__ gc, const double* __restrict__ vin,
                                                         const double* __restrict__ rhs, double lambda,
                                                         double* __restrict__ out, double* __restrict__ out_d,
                                                         double* parts) {
  constexpr int kR = point_ranges<T>();
  constexpr int kW = share_width<MODE>();
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  // all of the camera's shares are loaded (16-byte pieces) before the first sum
  Vec2<double> v[kW / 2 * kR];
  const Vec2<double>* sh = reinterpret_cast<const Vec2<double>*>(shares + kW * kR * c);
#pragma unroll
  for (int i = 0; i < kW / 2 * kR; ++i) v[i] = sh[i];
  double s[9], s2[9];
#pragma unroll
  for (int a = 0; a < 9; ++a) s[a] = s2[a] = 0.0;
#pragma unroll
  for (int g = 0; g < kR; ++g) {
    double e[kW];
#pragma unroll
    for (int i = 0; i < kW / 2; ++i) {
      e[2 * i] = v[kW / 2 * g + i].x;
      e[2 * i + 1] = v[kW / 2 * g + i].y;
    }
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      s[a] += e[a];
      if (MODE == cCheck) s2[a] += e[9 + a];
    }
  }
  finish_camera<T, MODE>(nc, c, s, s2, tabw + kCamTab * c, dc, gc, vin, nullptr, rhs, lambda, out, out_d, parts,
                         nullptr);
}

// ---------------------------------------------------------------------------------------
// Column sums of a (rows x kParts) partial array in fixed order (deterministic).
// ---------------------------------------------------------------------------------------
// Up to four column-sum jobs: job j sums rows[j] rows of the partial array starting at row
// first[j] into out[slot[j] * kParts ...].
struct SumJobs {
  int rows[4], first[4], slot[4];
};
// Column sums of `rows` rows of kParts partials by one workgroup (fixed assignment and tree:
// per-thread strided sums, wave sums, waves in order); valid in threads 0 .. kParts - 1.
__device__ __forceinline__ double column_sums(const double* __restrict__ p, int rows, double (*smem)[kBlock / kWave]) {
  double s[kParts];
#pragma unroll
  for (int col = 0; col < kParts; ++col) s[col] = 0.0;
  for (int i = threadIdx.x; i < rows; i += blockDim.x)
#pragma unroll
    for (int col = 0; col < kParts; ++col) s[col] += p[kParts * i + col];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int col = 0; col < kParts; ++col) {
    const double v = wave_sum(s[col]);
    if (lane == 0) smem[col][wave] = v;
  }
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x < kParts)
    for (int w = 0; w < kBlock / kWave; ++w) t += smem[threadIdx.x][w];
  __syncthreads();
  return t;
}

// The end of an attempt in one launch (no separate reduction launches): the check's column
// sums (jobs, the tree of column_sums, one workgroup each), the trial cost from k_cost's per-workgroup partials
// (the tree of k_reduce_parts' sum) when cost_parts is set, ||delta||_inf and its nonfinite flag
// from k_stage's partials when norm_parts is set (else from norm / norm_flag), and the PCG
// state and factor failures, gathered into the host-visible readback.
struct AttemptReadback {
  DevScalars sc;
  int fail, cost_flag, norm_flag, pad;
  double cost, norm;
  double sums[3 * kParts];
};
struct TailArgs {
  const DevScalars* sc;
  int* fail;  // read, then cleared for the next attempt's factorization
  const double* parts;  // the check's partial rows
  SumJobs jobs;
  int njobs;
  const double* cost_parts;  // k_cost's partials (nullptr: no trial cost)
  int cost_blocks;
  const int* cost_flag;
  const double* norm_parts;  // k_stage's [max, bad] pairs (nullptr: norm, norm_flag)
  int norm_blocks;
  const double* norm;
  const int* norm_flag;
  double* sums;  // the column sums, also kept on the device
  AttemptReadback* out;
};
static __global__ void __launch_bounds__(kBlock) k_attempt_tail(TailArgs a) {
  // workgroup j < njobs: column-sum job j; the last workgroup: cost, norm, state (launched with
  // njobs + 1 workgroups, which share nothing)
  __shared__ double smem[kParts][kBlock / kWave];
  const int t = threadIdx.x;
  const int j = blockIdx.x;
  if (j < a.njobs) {
    const double v = column_sums(a.parts + (size_t)kParts * a.jobs.first[j], a.jobs.rows[j], smem);
    if (t < kParts) {
      a.sums[kParts * a.jobs.slot[j] + t] = v;
      a.out->sums[kParts * a.jobs.slot[j] + t] = v;
    }
    return;
  }
  if (a.cost_parts != nullptr) {  // uniform
    double v = 0.0;
    for (int i = t; i < a.cost_blocks; i += blockDim.x) v += a.cost_parts[i];
    v = block_sum(v, &smem[0][0]);
    if (t == 0) {
      a.out->cost = v;
      a.out->cost_flag = *a.cost_flag;
    }
    __syncthreads();
  } else if (t == 0) {
    a.out->cost = 0.0;
    a.out->cost_flag = 0;
  }
  if (a.norm_parts != nullptr) {  // uniform
    double m = 0.0;
    int bad = 0;
    for (int i = t; i < a.norm_blocks; i += blockDim.x) {
      m = fmax(m, a.norm_parts[2 * i]);
      bad |= a.norm_parts[2 * i + 1] != 0.0;
    }
    m = wave_max(m);
    bad = __any(bad) ? 1 : 0;
    const int wave = t / kWave, lane = t % kWave;
    if (lane == 0) {
      smem[0][wave] = m;
      smem[1][wave] = bad;
    }
    __syncthreads();
    if (t == 0) {
      double r = 0.0, b = 0.0;
      for (int w = 0; w < kBlock / kWave; ++w) {
        r = fmax(r, smem[0][w]);
        b = fmax(b, smem[1][w]);
      }
      a.out->norm = r;
      a.out->norm_flag = b != 0.0 ? 2 : 0;
    }
  } else if (t == 0) {
    a.out->norm = *a.norm;
    a.out->norm_flag = *a.norm_flag;
  }
  if (t == 0) {
    a.out->sc = *a.sc;
    a.out->fail = *a.fail;
    *a.fail = 0;
  }
}


static __global__ void k_sum_parts(const double* __restrict__ parts, int rows, double* __restrict__ out) {
  __shared__ double smem[kParts][kBlock / kWave];
  double s[kParts];
#pragma unroll
  for (int col = 0; col < kParts; ++col) s[col] = 0.0;
  for (int i = threadIdx.x; i < rows; i += blockDim.x)
#pragma unroll
    for (int col = 0; col < kParts; ++col) s[col] += parts[kParts * i + col];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int col = 0; col < kParts; ++col) {
    const double v = wave_sum(s[col]);
    if (lane == 0) smem[col][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < kParts) {
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) t += smem[threadIdx.x][w];
    out[threadIdx.x] = t;
  }
}

// ---------------------------------------------------------------------------------------
// Reduced-system PCG (textbook; replaces faer's conjugate_gradient at schur.rs:249-310).
// One thread per camera block for the vector work; dot products and norms in f64.
// ---------------------------------------------------------------------------------------
// Block-Jacobi z = (B + lambda I)^-1 r per camera (RetainedPreconditioner, schur.rs:670-677;
// unobserved blocks divide by lambda, block_factors.rs:193-203), or with the learned
// correction z = L^-T G G^T L^-1 r. Both inverses are formed explicitly per camera (packed
// symmetric 45) once per solve, so one application is a symmetric 9x9 product.
// precondition() with the camera's block already in registers.
__device__ __forceinline__ void precondition_block(bool observed, const double m[45], double lambda, const double r[9],
                                                   double z[9]) {
  if (observed) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 9; ++j) s += m[sym9(i, j)] * r[j];
      z[i] = s;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) z[i] = lambda != 0.0 ? r[i] / lambda : r[i];
  }
}

__device__ __forceinline__ void precondition(int c, const int* __restrict__ cam_off, const double* __restrict__ minv,
                                             double lambda, const double r[9], double z[9]) {
  if (cam_off[c + 1] > cam_off[c]) {
    double m[45];
#pragma unroll
    for (int i = 0; i < 45; ++i) m[i] = minv[45 * c + i];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 9; ++j) s += m[sym9(i, j)] * r[j];
      z[i] = s;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) z[i] = lambda != 0.0 ? r[i] / lambda : r[i];
  }
}

// x = 0, r = b, z = M r, p = z; per-workgroup partials r.r (col 0) and r.z (col 1). The PCG's
// scalar state arrives as an argument and workgroup 0 publishes it (no host copy before the
// solve).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_init(int nc, const int* __restrict__ cam_off,
                                                    const double* __restrict__ mbj,
                                                    const double* __restrict__ mlearn,
                                                    const double* __restrict__ rhs, double* __restrict__ x,
                                                    double* __restrict__ r, double* __restrict__ z,
                                                    double* __restrict__ p, double* parts, DevScalars init,
                                                    DevScalars* sc) {
  __shared__ double smem[kBlock / kWave];
  if (blockIdx.x == 0 && threadIdx.x == 0) *sc = init;
  const bool learned = init.st[0].use_learned != 0;
  const double lambda = init.st[0].lambda;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0, rz = 0.0;
  if (c < nc) {
    double rv[9], zv[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      rv[i] = rhs[9 * c + i];
      x[9 * c + i] = 0.0;
      r[9 * c + i] = rv[i];
    }
    precondition(c, cam_off, learned ? mlearn : mbj, lambda, rv, zv);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      z[9 * c + i] = zv[i];
      p[9 * c + i] = zv[i];
      rr += rv[i] * rv[i];
      rz += rv[i] * zv[i];
    }
  }
  const double b0 = block_sum(rr, smem);
  const double b1 = block_sum(rz, smem);
  if (threadIdx.x == 0) {
    parts[kParts * blockIdx.x] = b0;
    parts[kParts * blockIdx.x + 1] = b1;
    parts[kParts * blockIdx.x + 2] = 0.0;
  }
}

// Start of PCG iteration i: every workgroup sums the previous update's partials (the same way,
// so all agree), decides convergence / exhaustion / fallback, and forms p = z + beta p with the
// forward table of p; workgroup 0 publishes the new state into slot out.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_direction(int nc, const int* __restrict__ cam_off,
                                                         const T* __restrict__ tab, const T* __restrict__ dc,
                                                         const double* __restrict__ zb, const double* __restrict__ z,
                                                         double* __restrict__ p, T* __restrict__ fw,
                                                         double* __restrict__ fww, T* __restrict__ prow,
                                                         const double* __restrict__ parts, int nparts,
                                                         const CgState* __restrict__ in, CgState* out,
                                                         const DevScalars* __restrict__ sc,
                                                         const int* __restrict__ factor_fails, CgState* mirror) {
  const int max_iters = sc->max_iters, fallback_at = sc->fallback_at;
  const double rel_tol = sc->rel_tol, fallback_factor = sc->fallback_factor;
  CgState s = *in;
  if (s.done) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      *out = s;
      if (mirror) *mirror = s;
    }
    return;
  }
  // this camera's vectors and table row load before the partial sums, overlapping them
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double zc[9], pc[9], rj[9], dv[9];
  T trow[16];
  if (c < nc) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      zc[i] = z[9 * c + i];
      pc[i] = p[9 * c + i];
      rj[i] = double(tab[kCamTab * c + tRJ(i / 3, i % 3)]);
      dv[i] = double(dc[9 * c + i]);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) trow[i] = tab[kCamTab * c + i];
  }
  double sums[3];
  block_sum_parts_cols3(parts, nparts, sums);
  const double rr = sums[0], rz = sums[1], rzb = sums[2];
  double beta = 0.0;
  bool restart = false;
  s.rr = rr;
  if (s.iters < 0) {  // first direction after k_cg_init: p = z already
    s.bb = rr;
    s.tol2 = rel_tol * rel_tol * rr;
    s.target = fallback_factor * fallback_factor * rr;
    s.rz = rz;
    s.iters = 0;
    if (*factor_fails != 0) {  // a block of B + lambda I or C + lambda I did not factor
      s.done = 1;
      s.term = 2;  // NotPositiveDefinite
    } else if (rr <= s.tol2) {
      s.done = 1;
      s.term = 0;  // Converged
    }
  } else {
    s.iters += 1;
    if (rr <= s.tol2) {
      s.done = 1;
      s.term = 0;  // Converged
    } else if (s.iters >= max_iters) {
      s.done = 1;
      s.term = 1;  // IterationLimit
    } else if (s.use_learned && s.iters == fallback_at && rr > s.target) {
      s.use_learned = 0;  // fall back to block-Jacobi and restart the direction
      s.fallback_taken = 1;
      restart = true;
      s.rz = rzb;
    } else {
      beta = rz / s.rz;
      s.rz = rz;
    }
  }
  // workgroup 0 publishes the state, and a copy to host memory the host polls (mirror)
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *out = s;
    if (mirror) *mirror = s;
  }
  if (s.done) return;
  if (c >= nc) return;
  const bool first = in->iters < 0;
  double pv[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    if (restart) pv[i] = zb[9 * c + i];
    else if (first) pv[i] = pc[i];
    else pv[i] = zc[i] + beta * pc[i];
    p[9 * c + i] = pv[i];
  }
  forward_table_rows<T>(rj, pv, dv, trow, fw + kFwd * c, fww != nullptr ? fww + kFwd * c : nullptr, prow + kPRow * c);
}

// x += alpha p; r -= alpha q; z = M r. Every workgroup forms p.q from the camera kernel's
// npq per-(camera, range) partials (alpha = r.z / p.q; p.q <= 0 ends the solve
// NotPositiveDefinite), and q_c = lambda p_c + the camera's point_ranges<T>() shares y in order.
// Partials: r.r (col 0), r.z (col 1) and, before a possible fallback, r.z_blockJacobi (col 2).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_update(int nc, const int* __restrict__ cam_off,
                                                      const double* __restrict__ mbj,
                                                      const double* __restrict__ mlearn,
                                                      const double* __restrict__ p,
                                                      const double* __restrict__ qy, double* __restrict__ x,
                                                      double* __restrict__ r, double* __restrict__ z,
                                                      double* __restrict__ zb, const double* __restrict__ pq_parts,
                                                      int npq, double* parts, CgState* st,
                                                      const DevScalars* __restrict__ sc) {
  __shared__ double smem[kBlock / kWave];
  __shared__ double bcast;
  const CgState s = *st;
  if (s.done) return;
  const double lambda = s.lambda;
  const int fallback_at = sc->fallback_at;
  const bool learned = s.use_learned != 0;
  // this camera's vectors and preconditioner block load before the p.q sum, overlapping it
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double pc[9], qc[9], xc[9], rc[9], m[45];
  bool observed = false;
  if (c < nc) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      pc[i] = p[9 * c + i];
      xc[i] = x[9 * c + i];
      rc[i] = r[9 * c + i];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      double y = 0.0;
      for (int g = 0; g < point_ranges<T>(); ++g) y += qy[9 * (point_ranges<T>() * c + g) + i];
      qc[i] = lambda * pc[i] + y;
    }
    observed = cam_off[c + 1] > cam_off[c];
    const double* mc = (learned ? mlearn : mbj) + 45 * c;
#pragma unroll
    for (int i = 0; i < 45; ++i) m[i] = mc[i];
  }
  // p.q from its partials: each thread sums a contiguous run (an even count, so its start is
  // 16-byte aligned) two values a load with four independent sums, so the loads of its run
  // are in flight together (one dependent sum over 188 partials per thread took 44 us with
  // 12,000 partials); every workgroup uses the same order
  double pq = 0.0;
  {
    const int per = ((npq + (int)blockDim.x - 1) / (int)blockDim.x + 1) & ~1;
    const int i0 = min(npq, (int)threadIdx.x * per), i1 = min(npq, i0 + per);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int i = i0;
    for (; i + 8 <= i1; i += 8) {
      const Vec2<double> v0 = *reinterpret_cast<const Vec2<double>*>(pq_parts + i);
      const Vec2<double> v1 = *reinterpret_cast<const Vec2<double>*>(pq_parts + i + 2);
      const Vec2<double> v2 = *reinterpret_cast<const Vec2<double>*>(pq_parts + i + 4);
      const Vec2<double> v3 = *reinterpret_cast<const Vec2<double>*>(pq_parts + i + 6);
      a0 += v0.x + v0.y;
      a1 += v1.x + v1.y;
      a2 += v2.x + v2.y;
      a3 += v3.x + v3.y;
    }
    for (; i < i1; ++i) a0 += pq_parts[i];
    pq = (a0 + a1) + (a2 + a3);
  }
  pq = block_sum(pq, smem);
  if (threadIdx.x == 0) bcast = pq;
  __syncthreads();
  pq = bcast;
  if (!(pq > 0.0) || !isfinite(pq)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->term = 2;  // NotPositiveDefinite
    }
    return;
  }
  const double alpha = s.rz / pq;
  const bool both = learned && s.iters + 1 == fallback_at;
  double rr = 0.0, rz = 0.0, rzb = 0.0;
  if (c < nc) {
    double rv[9], zv[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      x[9 * c + i] = xc[i] + alpha * pc[i];
      rv[i] = rc[i] - alpha * qc[i];
      r[9 * c + i] = rv[i];
    }
    precondition_block(observed, m, lambda, rv, zv);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      z[9 * c + i] = zv[i];
      rr += rv[i] * rv[i];
      rz += rv[i] * zv[i];
    }
    if (both) {
      precondition(c, cam_off, mbj, lambda, rv, zv);
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        zb[9 * c + i] = zv[i];
        rzb += rv[i] * zv[i];
      }
    }
  }
  double b[3] = {rr, rz, rzb};
  block_sum3(b);
  if (threadIdx.x == 0) {
    parts[kParts * blockIdx.x] = b[0];
    parts[kParts * blockIdx.x + 1] = b[1];
    parts[kParts * blockIdx.x + 2] = b[2];
  }
}

// delta_c = x / Dc (original coordinates, f64)
template <typename T>
__global__ void k_recover_cameras(int n, const double* __restrict__ x, const T* __restrict__ dc,
                                  double* __restrict__ delta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) delta[i] = x[i] / double(dc[i]);
}

// State staging: x_trial = x + delta (additive BAL retraction, model.rs:20-28), plus the
// compute-precision point copy.
// The trial state cur + delta into the other buffer: one thread per camera stages its 9
// parameters and, with tab64, forms the trial camera tables from them (camera_table, as
// k_camera_tables would: f64 table, and with lo its rounding to T and the widening); the points
// are staged element by element. With nparts: also each workgroup's max |delta| and whether a
// delta entry is nonfinite (nparts[2 b], nparts[2 b + 1]), which k_attempt_tail reduces to
// ||delta||_inf and its flag (the same maximum k_norm_inf forms, max being exact in any order);
// with zero_flag: also clears that flag word for the kernels that follow in stream order.
// Grid-stride.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_stage(int nc, int np, const double* __restrict__ cams,
                                                  const double* __restrict__ pts, const double* __restrict__ dcam,
                                                  const double* __restrict__ dpt, double* __restrict__ tcams,
                                                  double* __restrict__ tpts, T* __restrict__ tpts_t,
                                                  double* __restrict__ tab64, T* __restrict__ lo,
                                                  double* __restrict__ lo_wide, double* __restrict__ nparts,
                                                  int* zero_flag) {
  __shared__ double smem[kBlock / kWave];
  __shared__ int sbad;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x, gsz = gridDim.x * blockDim.x;
  double v = 0.0;
  bool bad = false;
  for (int c = gid; c < nc; c += gsz) {
    double tc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double d = dcam[9 * c + k];
      tc[k] = cams[9 * c + k] + d;
      tcams[9 * c + k] = tc[k];
      bad |= !isfinite(d);
      v = fmax(v, fabs(d));
    }
    if (tab64 != nullptr) camera_table<T>(c, tc, tab64, lo, lo_wide);
  }
  for (int j = gid; j < 3 * np; j += gsz) {
    const double d = dpt[j];
    const double x = pts[j] + d;
    tpts[j] = x;
    tpts_t[j] = T(x);
    bad |= !isfinite(d);
    v = fmax(v, fabs(d));
  }
  if (zero_flag != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *zero_flag = 0;
  if (nparts == nullptr) return;  // uniform over the grid
  if (threadIdx.x == 0) sbad = 0;
  v = wave_max(v);
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  __syncthreads();
  if (bad) sbad = 1;
  if (lane == 0) smem[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) r = fmax(r, smem[w]);
    nparts[2 * blockIdx.x] = r;
    nparts[2 * blockIdx.x + 1] = sbad ? 1.0 : 0.0;
  }
}

template <typename T>
__global__ void k_convert(int n, const double* __restrict__ a, T* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = T(a[i]);
}

}  // namespace fg
