// fg_math.h -- device math shared by the bundle-adjustment kernels.
//
// The BAL reprojection Jacobian (examples/bal/model.rs:157-184) is never stored. Each
// observation keeps a 4-scalar record (qx, qy, z, sw): the normalized image point
// q = -P.xy / P.z, the camera-frame depth z = P.z and the square root of its IRLS weight.
// Together with the per-camera table (R, R*Jr, t, f, k1, k2) and nothing else, the record
// reproduces the weighted blocks
//   Jp = sw * dp * R                                   (2x3)
//   Jc = sw * [ ((P - t) x dp_i)^T (R Jr) | dp | radial q, f r2 q, f r2^2 q ]   (2x9)
// with P = (-qx z, -qy z, z), dp = -(1/z) [[a, b, a qx + b qy], [b, d, b qx + d qy]] and
// [[a, b], [b, d]] = f (radial I + (2 k1 + 4 k2 r2) q q^T). The rotation columns use
// R [X]x = [R X]x R, so the point itself is not needed either.
#pragma once
#include <hip/hip_runtime.h>

namespace fg {

constexpr int kWave = 64;
// Per-camera table, 32 scalars read as eight 4-vectors: rows of R with t_i in the fourth
// lane, then (f, k1, k2, 0), then rows of R Jr (fourth lane 0), then padding.
constexpr int kCamTab = 32;
__host__ __device__ constexpr int tR(int i, int j) { return 4 * i + j; }
__host__ __device__ constexpr int tT(int i) { return 4 * i + 3; }
constexpr int tF = 12, tK1 = 13, tK2 = 14;
__host__ __device__ constexpr int tRJ(int i, int j) { return 16 + 4 * i + j; }
// Per-camera input table of a retained vector z: w' = RJ z_rot (3), z_t (3), z_intr (3), pad.
constexpr int kFwd = 12;

// Packed lower-triangular column-major index for a symmetric n x n block (i >= j).
__host__ __device__ constexpr int sym_index(int n, int i, int j) {
  return j * n - (j * (j - 1)) / 2 + (i - j);
}
__host__ __device__ constexpr int sym9(int i, int j) { return i >= j ? sym_index(9, i, j) : sym_index(9, j, i); }
__host__ __device__ constexpr int sym3(int i, int j) { return i >= j ? sym_index(3, i, j) : sym_index(3, j, i); }

template <typename T> struct Vec2 { T x, y; };
template <typename T> struct alignas(4 * sizeof(T)) Vec4 { T x, y, z, w; };

// First 16 scalars of a camera table (R, t, f, k1, k2) with four 4-vector loads.
template <typename T>
__device__ __forceinline__ void load_cam16(const T* __restrict__ tab, int c, T out[16]) {
  const Vec4<T>* t4 = reinterpret_cast<const Vec4<T>*>(tab) + (kCamTab / 4) * c;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const Vec4<T> v = t4[i];
    out[4 * i] = v.x; out[4 * i + 1] = v.y; out[4 * i + 2] = v.z; out[4 * i + 3] = v.w;
  }
}
// Forward table of a camera (w', z_t, z_intr) with three 4-vector loads.
template <typename T>
__device__ __forceinline__ void load_fwd(const T* __restrict__ fw, int c, T out[kFwd]) {
  const Vec4<T>* f4 = reinterpret_cast<const Vec4<T>*>(fw) + (kFwd / 4) * c;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const Vec4<T> v = f4[i];
    out[4 * i] = v.x; out[4 * i + 1] = v.y; out[4 * i + 2] = v.z; out[4 * i + 3] = v.w;
  }
}

// Quantities of one observation re-formed from its record and its camera's table.
template <typename T>
struct ObsGeom {
  T qx, qy, z, sw;
  T r2, radial;
  T dp0[3], dp1[3];  // rows of dp (unweighted)
  T pt[3];           // P - t = R X
  T f, fr2, fr4;     // f, f r2, f r2^2
};

template <typename T>
__device__ __forceinline__ void obs_geom(const Vec4<T>& rec, const T* __restrict__ tab, ObsGeom<T>& g) {
  g.qx = rec.x; g.qy = rec.y; g.z = rec.z; g.sw = rec.w;
  const T f = tab[tF], k1 = tab[tK1], k2 = tab[tK2];
  g.r2 = g.qx * g.qx + g.qy * g.qy;
  g.radial = T(1) + k1 * g.r2 + k2 * g.r2 * g.r2;
  const T c2 = T(2) * k1 + T(4) * k2 * g.r2;
  const T a = f * (g.radial + c2 * g.qx * g.qx);
  const T b = f * (c2 * g.qx * g.qy);
  const T d = f * (g.radial + c2 * g.qy * g.qy);
  const T iz = T(-1) / g.z;
  g.dp0[0] = iz * a; g.dp0[1] = iz * b; g.dp0[2] = iz * (a * g.qx + b * g.qy);
  g.dp1[0] = iz * b; g.dp1[1] = iz * d; g.dp1[2] = iz * (b * g.qx + d * g.qy);
  g.pt[0] = -g.qx * g.z - tab[tT(0)];
  g.pt[1] = -g.qy * g.z - tab[tT(1)];
  g.pt[2] = g.z - tab[tT(2)];
  g.f = f;
  g.fr2 = f * g.r2;
  g.fr4 = g.fr2 * g.r2;
}

// u = Jc z for a camera vector z summarized by its forward table (w', z_t, z_intr).
template <typename T>
__device__ __forceinline__ void jc_apply(const ObsGeom<T>& g, const T* __restrict__ fw, T& u0, T& u1) {
  // w' x (P - t) + z_t
  const T v0 = fw[1] * g.pt[2] - fw[2] * g.pt[1] + fw[3];
  const T v1 = fw[2] * g.pt[0] - fw[0] * g.pt[2] + fw[4];
  const T v2 = fw[0] * g.pt[1] - fw[1] * g.pt[0] + fw[5];
  const T s = g.radial * fw[6] + g.fr2 * fw[7] + g.fr4 * fw[8];
  u0 = g.sw * (g.dp0[0] * v0 + g.dp0[1] * v1 + g.dp0[2] * v2 + g.qx * s);
  u1 = g.sw * (g.dp1[0] * v0 + g.dp1[1] * v1 + g.dp1[2] * v2 + g.qy * s);
}

// u = Jp x for a point increment x (3).
template <typename T>
__device__ __forceinline__ void jp_apply(const ObsGeom<T>& g, const T* __restrict__ tab, const T x[3], T& u0, T& u1) {
  const T y0 = tab[tR(0, 0)] * x[0] + tab[tR(0, 1)] * x[1] + tab[tR(0, 2)] * x[2];
  const T y1 = tab[tR(1, 0)] * x[0] + tab[tR(1, 1)] * x[1] + tab[tR(1, 2)] * x[2];
  const T y2 = tab[tR(2, 0)] * x[0] + tab[tR(2, 1)] * x[1] + tab[tR(2, 2)] * x[2];
  u0 = g.sw * (g.dp0[0] * y0 + g.dp0[1] * y1 + g.dp0[2] * y2);
  u1 = g.sw * (g.dp1[0] * y0 + g.dp1[1] * y1 + g.dp1[2] * y2);
}

// w += Jp^T v (3).
template <typename T>
__device__ __forceinline__ void jpt_accumulate(const ObsGeom<T>& g, const T* __restrict__ tab, T v0, T v1, T w[3]) {
  const T e0 = g.sw * (v0 * g.dp0[0] + v1 * g.dp1[0]);
  const T e1 = g.sw * (v0 * g.dp0[1] + v1 * g.dp1[1]);
  const T e2 = g.sw * (v0 * g.dp0[2] + v1 * g.dp1[2]);
  w[0] += tab[tR(0, 0)] * e0 + tab[tR(1, 0)] * e1 + tab[tR(2, 0)] * e2;
  w[1] += tab[tR(0, 1)] * e0 + tab[tR(1, 1)] * e1 + tab[tR(2, 1)] * e2;
  w[2] += tab[tR(0, 2)] * e0 + tab[tR(1, 2)] * e1 + tab[tR(2, 2)] * e2;
}

// acc += the camera-frame factors of Jc^T v: rotation slots hold (P - t) x e and are mapped by
// (R Jr)^T once per camera (camera_finalize); translation slots hold e; intrinsics hold
// (radial, f r2, f r2^2) * (q . v). All weighted by sw.
template <typename T>
__device__ __forceinline__ void jct_factors(const ObsGeom<T>& g, T v0, T v1, T acc[9]) {
  const T e0 = g.sw * (v0 * g.dp0[0] + v1 * g.dp1[0]);
  const T e1 = g.sw * (v0 * g.dp0[1] + v1 * g.dp1[1]);
  const T e2 = g.sw * (v0 * g.dp0[2] + v1 * g.dp1[2]);
  acc[0] += g.pt[1] * e2 - g.pt[2] * e1;
  acc[1] += g.pt[2] * e0 - g.pt[0] * e2;
  acc[2] += g.pt[0] * e1 - g.pt[1] * e0;
  acc[3] += e0;
  acc[4] += e1;
  acc[5] += e2;
  const T s = g.sw * (g.qx * v0 + g.qy * v1);
  acc[6] += g.radial * s;
  acc[7] += g.fr2 * s;
  acc[8] += g.fr4 * s;
}

// Full weighted camera block (2x9, row-major jc[row*9+col]) for Gram assembly.
template <typename T>
__device__ __forceinline__ void jc_block(const ObsGeom<T>& g, const T* __restrict__ tab, T jc[18]) {
  const T* dps[2] = {g.dp0, g.dp1};
  const T qs[2] = {g.qx, g.qy};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const T* d = dps[i];
    const T c0 = g.pt[1] * d[2] - g.pt[2] * d[1];
    const T c1 = g.pt[2] * d[0] - g.pt[0] * d[2];
    const T c2 = g.pt[0] * d[1] - g.pt[1] * d[0];
#pragma unroll
    for (int j = 0; j < 3; ++j) jc[9 * i + j] = g.sw * (c0 * tab[tRJ(0, j)] + c1 * tab[tRJ(1, j)] + c2 * tab[tRJ(2, j)]);
    jc[9 * i + 3] = g.sw * d[0];
    jc[9 * i + 4] = g.sw * d[1];
    jc[9 * i + 5] = g.sw * d[2];
    jc[9 * i + 6] = g.sw * g.radial * qs[i];
    jc[9 * i + 7] = g.sw * g.fr2 * qs[i];
    jc[9 * i + 8] = g.sw * g.fr4 * qs[i];
  }
}

// In-register Cholesky of a packed symmetric n x n block (lower, column-major packed).
// Returns false on a nonpositive or nonfinite pivot, the failure classes of the reference's
// BlockFactors::factor (block_factors.rs:113-138).
template <typename T, int N>
__device__ __forceinline__ bool chol_packed(T* a) {
  bool ok = true;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    T d = a[sym_index(N, j, j)];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= a[sym_index(N, j, k)] * a[sym_index(N, j, k)];
    ok = ok && (d > T(0)) && isfinite(d);
    d = sqrt(d);
    a[sym_index(N, j, j)] = d;
    const T inv = T(1) / d;
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      T s = a[sym_index(N, i, j)];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= a[sym_index(N, i, k)] * a[sym_index(N, j, k)];
      a[sym_index(N, i, j)] = s * inv;
    }
  }
  return ok;
}

// x <- L^-1 x, then x <- L^-T x (packed lower L).
template <typename T, int N>
__device__ __forceinline__ void solve_lower_packed(const T* l, T* x) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    T s = x[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= l[sym_index(N, i, k)] * x[k];
    x[i] = s / l[sym_index(N, i, i)];
  }
}
template <typename T, int N>
__device__ __forceinline__ void solve_upper_packed(const T* l, T* x) {
#pragma unroll
  for (int ii = N - 1; ii >= 0; --ii) {
    T s = x[ii];
#pragma unroll
    for (int k = ii + 1; k < N; ++k) s -= l[sym_index(N, k, ii)] * x[k];
    x[ii] = s / l[sym_index(N, ii, ii)];
  }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

}  // namespace fg
