// fg_generic.hip -- the Schur backend for emissions of any factor family: Schur::new(retained,
// rw, ew) (src/schur.rs:116-656) over explicit Jacobian blocks accumulated in bulk from the
// host, so models other than the BAL evaluator (their own cost and linearization on the host)
// solve their damped systems on the GPU.
//
// Model (device, f64): each emission keeps its rows' residuals and at most one retained
// (width rw, at a multiple of rw below `retained`) and one eliminated (width ew) Jacobian
// block, column-major, in emission order (LinearizedModel::accumulate, linear_model.rs:180-210,
// with Schur::accumulate's partition rules, schur.rs:438-482). Blocks form a regular grid, so a
// variable block is its grid index; CSR lists give each block its emissions in emission order.
// Per block: column scales D, the scaled Gram (lower triangle, column-major), its damped
// Cholesky factor; per edge (emission with both blocks): the scaled cross block
// E = D_r^-1 J_r^T J_e D_e^-1 (layout of schur.rs:516, cross[(k ew + p) rw + c]).
//
// The reduced system S = lambda I + B - E (C + lambda I)^-1 E^T is solved by block-Jacobi PCG
// with device-resident scalars (one thread per variable block in every vector kernel; widths up
// to kGW). Unobserved retained coordinates carry only lambda (Reduced::apply, schur.rs:606-652)
// and their preconditioner divides by lambda (block_factors.rs:193-203).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "fg_kernels.h"
#include "fg_internal.h"

namespace {

using namespace fg;

constexpr int kGW = 16;  // widest variable block of the generic backend
const double kNaN = std::numeric_limits<double>::quiet_NaN();

#define GS_TRY(x)                                       \
  do {                                                  \
    if ((x) != hipSuccess) return FG_DEVICE_ERROR;      \
  } while (0)
#define GS_CHECK(x)                  \
  do {                               \
    const fg_status st_ = (x);       \
    if (st_ != FG_OK) return st_;    \
  } while (0)

inline int blocks_for(size_t n) { return (int)std::max<size_t>(1, (n + kBlock - 1) / kBlock); }

// Per-emission descriptor (device and host).
struct Em {
  int row0, rows;   // first row, row count
  int rblk, roff;   // retained block index and value offset (-1: none)
  int eblk, eoff;   // eliminated block index and value offset (-1: none)
  int edge, pad;    // edge index (-1: not an edge)
};

// Generic PCG state (one struct, device). Kernels of an iteration return at once when done.
struct GcgState {
  double rz, rr, bb, tol2, alpha, beta;
  int done, iters, term, max_iters;
};

// ---------------------------------------------------------------------------------------
// Small dense helpers on column-major w x w arrays (lower triangle significant).
// ---------------------------------------------------------------------------------------
__device__ bool chol_cm(double* a, int w) {
  for (int j = 0; j < w; ++j) {
    double d = a[j * w + j];
    for (int k = 0; k < j; ++k) d -= a[k * w + j] * a[k * w + j];
    if (!(d > 0.0) || !isfinite(d)) return false;
    d = sqrt(d);
    a[j * w + j] = d;
    for (int i = j + 1; i < w; ++i) {
      double v = a[j * w + i];
      for (int k = 0; k < j; ++k) v -= a[k * w + i] * a[k * w + j];
      a[j * w + i] = v / d;
    }
  }
  return true;
}
// b <- L^-1 b
__device__ void lower_solve_cm(const double* l, int w, double* b) {
  for (int i = 0; i < w; ++i) {
    double v = b[i];
    for (int k = 0; k < i; ++k) v -= l[k * w + i] * b[k];
    b[i] = v / l[i * w + i];
  }
}
// b <- L^-T b
__device__ void upper_solve_cm(const double* l, int w, double* b) {
  for (int i = w - 1; i >= 0; --i) {
    double v = b[i];
    for (int k = i + 1; k < w; ++k) v -= l[i * w + k] * b[k];
    b[i] = v / l[i * w + i];
  }
}
// BlockFactors::solve_normal for one block (block_factors.rs:178-204): (G + lambda I)^-1 b
// through the stored factor, or b / lambda (b at lambda = 0) for a coordinate without a block.
__device__ void block_solve(const double* fac, int w, bool present, double lambda, double* b) {
  if (present) {
    lower_solve_cm(fac, w, b);
    upper_solve_cm(fac, w, b);
  } else if (lambda != 0.0) {
    for (int i = 0; i < w; ++i) b[i] /= lambda;
  }
}

// ---------------------------------------------------------------------------------------
// Model kernels
// ---------------------------------------------------------------------------------------
// Per block column: g = J^T r (emission order) and, when scales != null, D = max(||J col||, floor)
// with a flag for nonfinite norms (LinearizedModel::prepare_scales, linear_model.rs:211-230).
__global__ void k_gs_columns(int nblk, int w, int col0, bool retained_side, const int* __restrict__ ptr,
                             const int* __restrict__ idx, const Em* __restrict__ em,
                             const double* __restrict__ vals, const double* __restrict__ res, double floor,
                             double* __restrict__ g, double* __restrict__ scales, int* flags) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblk * w) return;
  const int b = t / w, j = t % w;
  double gj = 0.0, n2 = 0.0;
  for (int q = ptr[b]; q < ptr[b + 1]; ++q) {
    const Em m = em[idx[q]];
    const int off = retained_side ? m.roff : m.eoff;
    for (int i = 0; i < m.rows; ++i) {
      const double v = vals[off + j * m.rows + i];
      gj += v * res[m.row0 + i];
      n2 += v * v;
    }
  }
  const int col = col0 + b * w + j;
  if (g) g[col] = gj;
  if (scales) {
    const double nrm = sqrt(n2);
    if (!isfinite(nrm)) atomicOr(flags, 1);
    scales[col] = fmax(nrm, floor);
  }
}

// Scaled Gram of each block, lower triangle column-major: one thread per (block, entry).
__global__ void k_gs_gram(int nblk, int w, int col0, bool retained_side, const int* __restrict__ ptr,
                          const int* __restrict__ idx, const Em* __restrict__ em, const double* __restrict__ vals,
                          const double* __restrict__ scales, double* __restrict__ gram) {
  const int tri = w * (w + 1) / 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblk * tri) return;
  const int b = t / tri;
  int k = t % tri, col = 0;
  while (k >= w - col) {  // entry k of the packed lower columns -> (row, col), row >= col
    k -= w - col;
    ++col;
  }
  const int row = col + k;
  const double dc = scales[col0 + b * w + col], dr = scales[col0 + b * w + row];
  double s = 0.0;
  for (int q = ptr[b]; q < ptr[b + 1]; ++q) {
    const Em m = em[idx[q]];
    const int off = retained_side ? m.roff : m.eoff;
    double e = 0.0;
    for (int i = 0; i < m.rows; ++i) e += (vals[off + col * m.rows + i] / dc) * (vals[off + row * m.rows + i] / dr);
    s += e;
  }
  gram[(size_t)b * w * w + col * w + row] = s;
  if (row != col) gram[(size_t)b * w * w + row * w + col] = 0.0;  // upper triangle stays zero
}

// Scaled cross blocks: one thread per (edge, p, c).
__global__ void k_gs_cross(int nedges, int rw, int ew, int retained, const int* __restrict__ edge_em,
                           const Em* __restrict__ em, const double* __restrict__ vals,
                           const double* __restrict__ scales, double* __restrict__ cross, int* flags) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nedges * rw * ew) return;
  const int k = t / (rw * ew), pc = t % (rw * ew), p = pc / rw, c = pc % rw;
  const Em m = em[edge_em[k]];
  const double dc = scales[m.rblk * rw + c], dp = scales[retained + m.eblk * ew + p];
  double s = 0.0;
  for (int i = 0; i < m.rows; ++i) s += (vals[m.roff + c * m.rows + i] / dc) * (vals[m.eoff + p * m.rows + i] / dp);
  if (!isfinite(s)) atomicOr(flags, 2);
  cross[((size_t)k * ew + p) * rw + c] = s;
}

// BlockFactors::factor (block_factors.rs:113-138): Gram + lambda I, in-place Cholesky, fallback
// on nonfinite entries or a nonpositive pivot. Blocks without emissions are absent.
__global__ void k_gs_factor(int nblk, int w, const int* __restrict__ ptr, const double* __restrict__ gram,
                            double lambda, double* __restrict__ fac, int* fail_count) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk || ptr[b + 1] == ptr[b]) return;
  double a[kGW * kGW];
  const double* gsrc = gram + (size_t)b * w * w;
  bool finite = true;
  for (int i = 0; i < w * w; ++i) {
    a[i] = gsrc[i];
    if (!isfinite(a[i])) finite = false;
  }
  for (int j = 0; j < w; ++j) a[j * w + j] += lambda;
  for (int j = 0; j < w && finite; ++j)
    if (!isfinite(a[j * w + j])) finite = false;
  const bool ok = finite && chol_cm(a, w);
  if (!ok) atomicAdd(fail_count, 1);
  double* dst = fac + (size_t)b * w * w;
  for (int i = 0; i < w * w; ++i) dst[i] = a[i];
}

enum ElimMode { eSchur = 0, eRhs = 1, eRecover = 2 };
// One thread per eliminated block:
//   eSchur:   t_b = (C_b + lambda I)^-1 sum_e E_e^T x_c(e)
//   eRhs:     t_b = (C_b + lambda I)^-1 (-g_b / D_b)                     (schur.rs:204-247)
//   eRecover: z_b = (C_b + lambda I)^-1 (-g_b / D_b - sum_e E_e^T x_c(e)), delta_b = z_b / D_b
template <int MODE>
__global__ void k_gs_elim(int neb, int rw, int ew, int retained, const int* __restrict__ eptr,
                          const int* __restrict__ eidx, const Em* __restrict__ em, const double* __restrict__ cross,
                          const double* __restrict__ fac_e, const double* __restrict__ g,
                          const double* __restrict__ scales, const double* __restrict__ x, double lambda,
                          double* __restrict__ t_out, double* __restrict__ delta, const GcgState* st) {
  if (st != nullptr && st->done) return;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= neb) return;
  double v[kGW];
  for (int p = 0; p < ew; ++p) v[p] = 0.0;
  if (MODE != eRhs) {
    for (int q = eptr[b]; q < eptr[b + 1]; ++q) {
      const Em m = em[eidx[q]];
      if (m.edge < 0) continue;
      const double* E = cross + (size_t)m.edge * rw * ew;
      const double* xc = x + (size_t)m.rblk * rw;
      for (int p = 0; p < ew; ++p) {
        double s = 0.0;
        for (int c = 0; c < rw; ++c) s += E[p * rw + c] * xc[c];
        v[p] += s;
      }
    }
  }
  const int col0 = retained + b * ew;
  if (MODE != eSchur)
    for (int p = 0; p < ew; ++p) v[p] = -g[col0 + p] / scales[col0 + p] - v[p];
  block_solve(fac_e + (size_t)b * ew * ew, ew, eptr[b + 1] > eptr[b], lambda, v);
  if (MODE == eRecover) {
    for (int p = 0; p < ew; ++p) delta[col0 + p] = v[p] / scales[col0 + p];
  } else {
    for (int p = 0; p < ew; ++p) t_out[(size_t)b * ew + p] = v[p];
  }
}

enum RetMode { rSchur = 0, rResid = 1, rRhs = 2 };
// One thread per retained block, u = sum_e E_e t_b(e):
//   rSchur: q = lambda x + B x - u, partial x.q            (Reduced::apply, schur.rs:606-652)
//   rResid: q as rSchur; partials ||q - rhs||^2, ||rhs||^2  (verify_step, schur.rs:346-360)
//   rRhs:   rhs = -g / D - u; partial ||rhs||^2             (factor_and_reduce_rhs, :204-247)
template <int MODE>
__global__ void __launch_bounds__(kBlock) k_gs_ret(int nrb, int rw, int ew, const int* __restrict__ rptr,
                                                   const int* __restrict__ ridx, const Em* __restrict__ em,
                                                   const double* __restrict__ cross, const double* __restrict__ gram_r,
                                                   const double* __restrict__ g, const double* __restrict__ scales,
                                                   const double* __restrict__ t, const double* __restrict__ x,
                                                   double lambda, double* __restrict__ out,
                                                   const double* __restrict__ rhs, double* __restrict__ parts,
                                                   const GcgState* st) {
  __shared__ double smem[kBlock / kWave];
  if (st != nullptr && st->done) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double s0 = 0.0, s1 = 0.0;
  if (c < nrb) {
    double u[kGW];
    for (int i = 0; i < rw; ++i) u[i] = 0.0;
    for (int q = rptr[c]; q < rptr[c + 1]; ++q) {
      const Em m = em[ridx[q]];
      if (m.edge < 0) continue;
      const double* E = cross + (size_t)m.edge * rw * ew;
      const double* tb = t + (size_t)m.eblk * ew;
      for (int p = 0; p < ew; ++p)
        for (int i = 0; i < rw; ++i) u[i] += E[p * rw + i] * tb[p];
    }
    const int col0 = c * rw;
    if (MODE == rRhs) {
      for (int i = 0; i < rw; ++i) {
        const double v = -g[col0 + i] / scales[col0 + i] - u[i];
        out[col0 + i] = v;
        s0 += v * v;
      }
    } else {
      const double* G = gram_r + (size_t)c * rw * rw;
      const bool present = rptr[c + 1] > rptr[c];
      for (int i = 0; i < rw; ++i) {
        double q = lambda * x[col0 + i] - u[i];
        if (present)
          for (int j = 0; j < rw; ++j) q += G[min(i, j) * rw + max(i, j)] * x[col0 + j];
        out[col0 + i] = q;
        if (MODE == rSchur) {
          s0 += x[col0 + i] * q;
        } else {
          const double e = q - rhs[col0 + i];
          s0 += e * e;
          s1 += rhs[col0 + i] * rhs[col0 + i];
        }
      }
    }
  }
  const double b0 = block_sum(s0, smem);
  const double b1 = block_sum(s1, smem);
  if (threadIdx.x == 0) {
    parts[kParts * blockIdx.x] = b0;
    parts[kParts * blockIdx.x + 1] = b1;
    for (int k = 2; k < kParts; ++k) parts[kParts * blockIdx.x + k] = 0.0;
  }
}

// ---------------------------------------------------------------------------------------
// PCG (textbook, x0 = 0, ||r|| <= tol ||b||; replaces faer's CG, schur.rs:249-310)
// ---------------------------------------------------------------------------------------
__device__ double sum_col(const double* parts, int n, int col, double* smem) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += parts[kParts * i + col];
  return block_sum(s, smem);
}

// x = 0, r = rhs, z = M r, p = z; partials r.r, r.z
__global__ void __launch_bounds__(kBlock) k_gcg_init(int nrb, int rw, const int* __restrict__ rptr,
                                                     const double* __restrict__ fac_r, double lambda,
                                                     const double* __restrict__ rhs, double* __restrict__ x,
                                                     double* __restrict__ r, double* __restrict__ z,
                                                     double* __restrict__ p, double* __restrict__ parts) {
  __shared__ double smem[kBlock / kWave];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0, rz = 0.0;
  if (c < nrb) {
    double v[kGW];
    for (int i = 0; i < rw; ++i) v[i] = rhs[c * rw + i];
    block_solve(fac_r + (size_t)c * rw * rw, rw, rptr[c + 1] > rptr[c], lambda, v);
    for (int i = 0; i < rw; ++i) {
      const double ri = rhs[c * rw + i];
      x[c * rw + i] = 0.0;
      r[c * rw + i] = ri;
      z[c * rw + i] = v[i];
      p[c * rw + i] = v[i];
      rr += ri * ri;
      rz += ri * v[i];
    }
  }
  const double b0 = block_sum(rr, smem);
  const double b1 = block_sum(rz, smem);
  if (threadIdx.x == 0) {
    parts[kParts * blockIdx.x] = b0;
    parts[kParts * blockIdx.x + 1] = b1;
  }
}

__global__ void k_gcg_start(const double* __restrict__ parts, int n, double rel_tol, int max_iters, GcgState* st) {
  __shared__ double smem[kBlock / kWave];
  const double rr = sum_col(parts, n, 0, smem);
  const double rz = sum_col(parts, n, 1, smem);
  if (threadIdx.x == 0) {
    GcgState s{};
    s.rr = s.bb = rr;
    s.rz = rz;
    s.tol2 = rel_tol * rel_tol * rr;
    s.max_iters = max_iters;
    s.term = -1;
    if (rr <= s.tol2) {  // includes b = 0
      s.done = 1;
      s.term = 0;
    }
    *st = s;
  }
}

// alpha = rz / p.q; a nonpositive or nonfinite curvature ends the solve (NotPositiveDefinite)
__global__ void k_gcg_alpha(const double* __restrict__ parts, int n, GcgState* st) {
  __shared__ double smem[kBlock / kWave];
  if (st->done) return;
  const double pq = sum_col(parts, n, 0, smem);
  if (threadIdx.x == 0) {
    if (!(pq > 0.0) || !isfinite(pq)) {
      st->done = 1;
      st->term = 2;
    } else {
      st->alpha = st->rz / pq;
    }
  }
}

// x += alpha p, r -= alpha q, z = M r; partials r.r, r.z
__global__ void __launch_bounds__(kBlock) k_gcg_update(int nrb, int rw, const int* __restrict__ rptr,
                                                       const double* __restrict__ fac_r, double lambda,
                                                       const double* __restrict__ p, const double* __restrict__ q,
                                                       double* __restrict__ x, double* __restrict__ r,
                                                       double* __restrict__ z, double* __restrict__ parts,
                                                       const GcgState* st) {
  __shared__ double smem[kBlock / kWave];
  if (st->done) return;
  const double alpha = st->alpha;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0, rz = 0.0;
  if (c < nrb) {
    double v[kGW];
    for (int i = 0; i < rw; ++i) {
      x[c * rw + i] += alpha * p[c * rw + i];
      v[i] = r[c * rw + i] - alpha * q[c * rw + i];
      r[c * rw + i] = v[i];
      rr += v[i] * v[i];
    }
    double w[kGW];
    for (int i = 0; i < rw; ++i) w[i] = v[i];
    block_solve(fac_r + (size_t)c * rw * rw, rw, rptr[c + 1] > rptr[c], lambda, w);
    for (int i = 0; i < rw; ++i) {
      z[c * rw + i] = w[i];
      rz += v[i] * w[i];
    }
  }
  const double b0 = block_sum(rr, smem);
  const double b1 = block_sum(rz, smem);
  if (threadIdx.x == 0) {
    parts[kParts * blockIdx.x] = b0;
    parts[kParts * blockIdx.x + 1] = b1;
  }
}

// iteration count, convergence / limit, beta = rz_new / rz
__global__ void k_gcg_beta(const double* __restrict__ parts, int n, GcgState* st) {
  __shared__ double smem[kBlock / kWave];
  if (st->done) return;
  const double rr = sum_col(parts, n, 0, smem);
  const double rz = sum_col(parts, n, 1, smem);
  if (threadIdx.x == 0) {
    GcgState s = *st;
    s.iters += 1;
    s.rr = rr;
    if (rr <= s.tol2) {
      s.done = 1;
      s.term = 0;
    } else if (s.iters >= s.max_iters) {
      s.done = 1;
      s.term = 1;
    }
    s.beta = rz / s.rz;
    s.rz = rz;
    *st = s;
  }
}

__global__ void k_gcg_direction(int n, const double* __restrict__ z, double* __restrict__ p, const GcgState* st) {
  if (st->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + st->beta * p[i];
}

// ---------------------------------------------------------------------------------------
// Verification (LinearizedModel::verify_step, linear_model.rs:233-282)
// ---------------------------------------------------------------------------------------
// delta_c = x / D on the retained coordinates
__global__ void k_gs_unscale(int n, const double* __restrict__ x, const double* __restrict__ scales,
                             double* __restrict__ delta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) delta[i] = x[i] / scales[i];
}

// rows of J delta (one thread per emission); partial ||J delta||^2 in column 5
__global__ void __launch_bounds__(kBlock) k_gs_jv(int ne, int rw, int ew, int retained, const Em* __restrict__ em,
                                                  const double* __restrict__ vals, const double* __restrict__ delta,
                                                  double* __restrict__ jd, double* __restrict__ parts) {
  __shared__ double smem[kBlock / kWave];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (e < ne) {
    const Em m = em[e];
    for (int i = 0; i < m.rows; ++i) {
      double v = 0.0;
      if (m.rblk >= 0)
        for (int c = 0; c < rw; ++c) v += vals[m.roff + c * m.rows + i] * delta[m.rblk * rw + c];
      if (m.eblk >= 0)
        for (int p = 0; p < ew; ++p) v += vals[m.eoff + p * m.rows + i] * delta[retained + m.eblk * ew + p];
      jd[m.row0 + i] = v;
      s += v * v;
    }
  }
  const double b = block_sum(s, smem);
  if (threadIdx.x == 0) {
    for (int k = 0; k < kParts; ++k) parts[kParts * blockIdx.x + k] = 0.0;
    parts[kParts * blockIdx.x + 5] = b;
  }
}

// Per block column: nr = J^T (J delta + r) + lambda D^2 delta; partials ||nr||^2, ||nr/D||^2,
// g.delta, ||g||^2, ||g/D||^2 (columns 0-4)
__global__ void __launch_bounds__(kBlock) k_gs_normal(int nblk, int w, int col0, bool retained_side,
                                                      const int* __restrict__ ptr, const int* __restrict__ idx,
                                                      const Em* __restrict__ em, const double* __restrict__ vals,
                                                      const double* __restrict__ res, const double* __restrict__ jd,
                                                      const double* __restrict__ g, const double* __restrict__ scales,
                                                      const double* __restrict__ delta, double lambda,
                                                      double* __restrict__ nr_out, double* __restrict__ parts) {
  __shared__ double smem[kBlock / kWave];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (t < nblk * w) {
    const int b = t / w, j = t % w;
    double acc = 0.0;
    for (int q = ptr[b]; q < ptr[b + 1]; ++q) {
      const Em m = em[idx[q]];
      const int off = retained_side ? m.roff : m.eoff;
      for (int i = 0; i < m.rows; ++i) acc += vals[off + j * m.rows + i] * (jd[m.row0 + i] + res[m.row0 + i]);
    }
    const int col = col0 + b * w + j;
    const double s = scales[col], d = delta[col], gj = g[col];
    const double nr = acc + (lambda * (s * d)) * s;
    if (nr_out) nr_out[col] = nr;
    v[0] = nr * nr;
    v[1] = (nr / s) * (nr / s);
    v[2] = gj * d;
    v[3] = gj * gj;
    v[4] = (gj / s) * (gj / s);
  }
  for (int k = 0; k < 5; ++k) {
    const double b = block_sum(v[k], smem);
    if (threadIdx.x == 0) parts[kParts * blockIdx.x + k] = b;
  }
  if (threadIdx.x == 0) parts[kParts * blockIdx.x + 5] = 0.0;
}

// max |x_i| with a nonfinite flag (norm_inf, src/optimization.rs:653-662); partial per workgroup
__global__ void __launch_bounds__(kBlock) k_gs_absmax(int n, const double* __restrict__ x, double* __restrict__ parts,
                                                      int* flags) {
  __shared__ double smem[kBlock / kWave];
  double m = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double a = fabs(x[i]);
    if (!isfinite(a)) atomicOr(flags, 4);
    m = fmax(m, a);
  }
  for (int o = kWave / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, kWave));
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (lane == 0) smem[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock / kWave; ++k) m = fmax(m, smem[k]);
    parts[blockIdx.x] = m;
  }
}

// Dense undamped information J^T J (LinearizedModel::covariance, linear_model.rs:122-167), lower
// triangle, deterministic: every entry has one owner, which sums its contributions in emission
// order. Thread r of k_gs_info_ret owns retained block r's diagonal block (its emissions through
// the retained CSR list); thread e of k_gs_info_elim owns eliminated block e's diagonal block
// and its cross blocks with the retained blocks (all through e's CSR list, since every emission
// that couples r and e is one of e's). n: dimension, h: column-major n x n, zeroed.
__device__ __forceinline__ void gs_block_product(const Em& m, const double* __restrict__ vals, int offx, int wx,
                                                 int colx, int offy, int wy, int coly, bool diag, size_t n,
                                                 double* __restrict__ h) {
  for (int ca = 0; ca < wx; ++ca)
    for (int cb = 0; cb < (diag ? ca + 1 : wy); ++cb) {
      double v = 0.0;
      for (int i = 0; i < m.rows; ++i) v += vals[offx + ca * m.rows + i] * vals[offy + cb * m.rows + i];
      h[(size_t)(coly + cb) * n + (size_t)(colx + ca)] += v;
    }
}
__global__ void k_gs_info_ret(int nrb, int rw, size_t n, const int* __restrict__ rptr, const int* __restrict__ ridx,
                              const Em* __restrict__ em, const double* __restrict__ vals, double* __restrict__ h) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrb) return;
  for (int k = rptr[r]; k < rptr[r + 1]; ++k) {
    const Em m = em[ridx[k]];
    gs_block_product(m, vals, m.roff, rw, r * rw, m.roff, rw, r * rw, true, n, h);
  }
}
__global__ void k_gs_info_elim(int neb, int rw, int ew, int retained, size_t n, const int* __restrict__ eptr,
                               const int* __restrict__ eidx, const Em* __restrict__ em,
                               const double* __restrict__ vals, double* __restrict__ h) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= neb) return;
  const int col = retained + e * ew;
  for (int k = eptr[e]; k < eptr[e + 1]; ++k) {
    const Em m = em[eidx[k]];
    gs_block_product(m, vals, m.eoff, ew, col, m.eoff, ew, col, true, n, h);
    // eliminated columns follow the retained ones, so the cross block lies below the diagonal
    if (m.rblk >= 0) gs_block_product(m, vals, m.eoff, ew, col, m.roff, rw, m.rblk * rw, false, n, h);
  }
}

}  // namespace

// ------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------
struct fg_gschur {
  fg_schur_config cfg;
  size_t retained = 0, rw = 0, ew = 0, dim = 0;
  bool prepared = false, ready = false, dirty = true;
  // host staging, emission order
  std::vector<double> h_res, h_vals;
  std::vector<Em> h_em;
  std::vector<double> h_g, h_delta;
  size_t rows = 0;
  // device
  hipStream_t stream = nullptr;
  Em* d_em = nullptr;
  double *d_res = nullptr, *d_vals = nullptr;
  int *d_rptr = nullptr, *d_ridx = nullptr, *d_eptr = nullptr, *d_eidx = nullptr, *d_edge_em = nullptr;
  size_t cap_em = 0, cap_rows = 0, cap_vals = 0, cap_dim = 0, cap_edges = 0, cap_rptr = 0, cap_eptr = 0,
         cap_ridx = 0, cap_eidx = 0, cap_edge_em = 0, cap_jd = 0, cap_parts = 0;
  size_t nedges = 0;
  double *d_g = nullptr, *d_D = nullptr, *d_gram_r = nullptr, *d_gram_e = nullptr, *d_fac_r = nullptr,
         *d_fac_e = nullptr, *d_cross = nullptr, *d_t = nullptr, *d_rhs = nullptr, *d_x = nullptr, *d_r = nullptr,
         *d_z = nullptr, *d_p = nullptr, *d_q = nullptr, *d_delta = nullptr, *d_jd = nullptr, *d_nr = nullptr,
         *d_parts = nullptr, *d_sums = nullptr;
  int* d_flags = nullptr;
  GcgState* d_st = nullptr;
  GcgState* h_st = nullptr;   // pinned, 3 slots
  double* h_sums = nullptr;   // pinned, 4 x kParts
  int* h_int = nullptr;       // pinned
  hipEvent_t poll[2] = {nullptr, nullptr};
  size_t solves = 0, iterations = 0;
  fg_schur_diagnostics diag;
  fg_schur_timings tim;
  double last_step_norm = kNaN;
  bool step_valid = false;
};

namespace {

void gs_reset_diag(fg_gschur* g, double damping) {
  g->diag = fg_schur_diagnostics{};
  g->diag.termination = FG_CG_NONE;
  g->diag.iterations = 0;
  g->diag.damping = damping;
  g->diag.estimated_relative_residual = kNaN;
  g->diag.reduced_relative_residual = kNaN;
  g->diag.scaled_relative_residual = kNaN;
  g->diag.original_relative_residual = kNaN;
  g->diag.used_fallback = 0;
}

template <typename P>
hipError_t grow(P** p, size_t* cap, size_t need) {
  if (*p != nullptr && *cap >= need) return hipSuccess;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  const size_t n = std::max<size_t>(need, 1);
  hipError_t e = hipMalloc((void**)p, n * sizeof(P));
  *cap = e == hipSuccess ? n : 0;
  return e;
}

void gs_free(fg_gschur* g) {
  auto f = [](auto*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
  };
  f(g->d_em);
  f(g->d_res);
  f(g->d_vals);
  f(g->d_rptr);
  f(g->d_ridx);
  f(g->d_eptr);
  f(g->d_eidx);
  f(g->d_edge_em);
  for (double** v : {&g->d_g, &g->d_D, &g->d_gram_r, &g->d_gram_e, &g->d_fac_r, &g->d_fac_e, &g->d_cross, &g->d_t,
                     &g->d_rhs, &g->d_x, &g->d_r, &g->d_z, &g->d_p, &g->d_q, &g->d_delta, &g->d_jd, &g->d_nr,
                     &g->d_parts, &g->d_sums})
    f(*v);
  f(g->d_flags);
  f(g->d_st);
  g->cap_em = g->cap_rows = g->cap_vals = g->cap_dim = g->cap_edges = g->cap_rptr = g->cap_eptr = g->cap_ridx =
      g->cap_eidx = g->cap_edge_em = g->cap_jd = g->cap_parts = 0;
}

size_t nrb(const fg_gschur* g) { return g->retained / g->rw; }
size_t neb(const fg_gschur* g) { return (g->dim - g->retained) / g->ew; }

// Per-dimension arrays (grow only).
fg_status gs_alloc_dim(fg_gschur* g) {
  if (g->cap_dim >= g->dim && g->d_g) return FG_OK;
  size_t cap = 0;
  const size_t n = std::max<size_t>(g->dim, 1), nr = std::max<size_t>(g->retained, 1);
  const size_t ne = std::max<size_t>(g->dim - g->retained, 1);
  GS_TRY(grow(&g->d_g, &cap, n));
  GS_TRY(grow(&g->d_D, &cap, n));
  GS_TRY(grow(&g->d_delta, &cap, n));
  GS_TRY(grow(&g->d_nr, &cap, n));
  GS_TRY(grow(&g->d_gram_r, &cap, nrb(g) * g->rw * g->rw + 1));
  GS_TRY(grow(&g->d_fac_r, &cap, nrb(g) * g->rw * g->rw + 1));
  GS_TRY(grow(&g->d_gram_e, &cap, neb(g) * g->ew * g->ew + 1));
  GS_TRY(grow(&g->d_fac_e, &cap, neb(g) * g->ew * g->ew + 1));
  GS_TRY(grow(&g->d_t, &cap, ne));
  for (double** v : {&g->d_rhs, &g->d_x, &g->d_r, &g->d_z, &g->d_p, &g->d_q}) GS_TRY(grow(v, &cap, nr));
  g->cap_dim = g->dim;
  return FG_OK;
}

// Upload the accumulated model and compute the gradient (once per accumulate phase).
fg_status gs_finalize(fg_gschur* g) {
  if (!g->dirty) return FG_OK;
  const size_t ne = g->h_em.size();
  GS_TRY(grow(&g->d_em, &g->cap_em, ne));
  GS_TRY(grow(&g->d_res, &g->cap_rows, g->rows));
  GS_TRY(grow(&g->d_vals, &g->cap_vals, g->h_vals.size()));
  GS_CHECK(gs_alloc_dim(g));
  GS_TRY(grow(&g->d_jd, &g->cap_jd, g->rows));
  // partial rows: the widest per-workgroup pass (emissions, block columns, blocks)
  const size_t prow = (size_t)std::max({blocks_for(ne), blocks_for(nrb(g) * g->rw), blocks_for(neb(g) * g->ew),
                                        1024});
  GS_TRY(grow(&g->d_parts, &g->cap_parts, kParts * prow));
  // CSR of emissions per retained / eliminated block, emission order
  const size_t R = nrb(g), E = neb(g);
  std::vector<int> rptr(R + 1, 0), eptr(E + 1, 0), ridx, eidx, edge_em;
  for (const Em& m : g->h_em) {
    if (m.rblk >= 0) ++rptr[m.rblk + 1];
    if (m.eblk >= 0) ++eptr[m.eblk + 1];
  }
  for (size_t i = 0; i < R; ++i) rptr[i + 1] += rptr[i];
  for (size_t i = 0; i < E; ++i) eptr[i + 1] += eptr[i];
  ridx.resize(rptr[R]);
  eidx.resize(eptr[E]);
  {
    std::vector<int> rf(rptr.begin(), rptr.end() - 1), ef(eptr.begin(), eptr.end() - 1);
    for (size_t e = 0; e < ne; ++e) {
      Em& m = g->h_em[e];
      if (m.rblk >= 0) ridx[rf[m.rblk]++] = (int)e;
      if (m.eblk >= 0) eidx[ef[m.eblk]++] = (int)e;
      if (m.rblk >= 0 && m.eblk >= 0) {
        m.edge = (int)edge_em.size();
        edge_em.push_back((int)e);
      } else {
        m.edge = -1;
      }
    }
  }
  g->nedges = edge_em.size();
  GS_TRY(grow(&g->d_rptr, &g->cap_rptr, R + 1));
  GS_TRY(grow(&g->d_eptr, &g->cap_eptr, E + 1));
  GS_TRY(grow(&g->d_ridx, &g->cap_ridx, ridx.size()));
  GS_TRY(grow(&g->d_eidx, &g->cap_eidx, eidx.size()));
  GS_TRY(grow(&g->d_edge_em, &g->cap_edge_em, edge_em.size()));
  GS_TRY(grow(&g->d_cross, &g->cap_edges, g->nedges * g->rw * g->ew));
  hipStream_t s = g->stream;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
  };
  GS_TRY(up(g->d_em, g->h_em.data(), ne * sizeof(Em)));
  GS_TRY(up(g->d_res, g->h_res.data(), g->rows * sizeof(double)));
  GS_TRY(up(g->d_vals, g->h_vals.data(), g->h_vals.size() * sizeof(double)));
  GS_TRY(up(g->d_rptr, rptr.data(), rptr.size() * sizeof(int)));
  GS_TRY(up(g->d_eptr, eptr.data(), eptr.size() * sizeof(int)));
  GS_TRY(up(g->d_ridx, ridx.data(), ridx.size() * sizeof(int)));
  GS_TRY(up(g->d_eidx, eidx.data(), eidx.size() * sizeof(int)));
  GS_TRY(up(g->d_edge_em, edge_em.data(), edge_em.size() * sizeof(int)));
  // gradient
  if (R)
    hipLaunchKernelGGL(k_gs_columns, dim3(blocks_for(R * g->rw)), dim3(kBlock), 0, s, (int)R, (int)g->rw, 0, true,
                       g->d_rptr, g->d_ridx, g->d_em, g->d_vals, g->d_res, 0.0, g->d_g, (double*)nullptr,
                       g->d_flags);
  if (E)
    hipLaunchKernelGGL(k_gs_columns, dim3(blocks_for(E * g->ew)), dim3(kBlock), 0, s, (int)E, (int)g->ew,
                       (int)g->retained, false, g->d_eptr, g->d_eidx, g->d_em, g->d_vals, g->d_res, 0.0, g->d_g,
                       (double*)nullptr, g->d_flags);
  GS_TRY(hipGetLastError());
  g->h_g.resize(g->dim);
  if (g->dim) GS_TRY(hipMemcpyAsync(g->h_g.data(), g->d_g, g->dim * sizeof(double), hipMemcpyDeviceToHost, s));
  GS_TRY(hipStreamSynchronize(s));
  g->dirty = false;
  return FG_OK;
}

// Sum the kParts columns of `rows` partial rows into h_sums[slot * kParts ...] (device side;
// read back by the caller's synchronization).
void gs_sum(fg_gschur* g, int rows, int slot) {
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, g->stream, (const double*)g->d_parts, rows,
                     g->d_sums + slot * kParts);
}

// S x into out (x, out: scaled retained coordinates), partial x.out summed into slot.
void gs_product(fg_gschur* g, const double* x, double lambda, double* out, const GcgState* st) {
  const int R = (int)nrb(g), E = (int)neb(g);
  if (E)
    hipLaunchKernelGGL(k_gs_elim<eSchur>, dim3(blocks_for(E)), dim3(kBlock), 0, g->stream, E, (int)g->rw,
                       (int)g->ew, (int)g->retained, g->d_eptr, g->d_eidx, g->d_em, g->d_cross, g->d_fac_e, g->d_g,
                       g->d_D, x, lambda, g->d_t, (double*)nullptr, st);
  hipLaunchKernelGGL(k_gs_ret<rSchur>, dim3(blocks_for(R)), dim3(kBlock), 0, g->stream, R, (int)g->rw, (int)g->ew,
                     g->d_rptr, g->d_ridx, g->d_em, g->d_cross, g->d_gram_r, g->d_g, g->d_D, g->d_t, x, lambda, out,
                     (const double*)nullptr, g->d_parts, st);
}

fg_status gs_solve_damped(fg_gschur* g, double lambda, double* predicted) {
  hipStream_t s = g->stream;
  const int R = (int)nrb(g), E = (int)neb(g);
  const int rblocks = blocks_for(R);
  // factor (any failed block: NotPositiveDefinite before any iteration)
  GS_TRY(hipMemsetAsync(g->d_flags, 0, 2 * sizeof(int), s));
  if (R)
    hipLaunchKernelGGL(k_gs_factor, dim3(blocks_for(R)), dim3(kBlock), 0, s, R, (int)g->rw, g->d_rptr, g->d_gram_r,
                       lambda, g->d_fac_r, g->d_flags + 1);
  if (E)
    hipLaunchKernelGGL(k_gs_factor, dim3(blocks_for(E)), dim3(kBlock), 0, s, E, (int)g->ew, g->d_eptr, g->d_gram_e,
                       lambda, g->d_fac_e, g->d_flags + 1);
  GS_TRY(hipGetLastError());
  GS_TRY(hipMemcpyAsync(g->h_int, g->d_flags + 1, sizeof(int), hipMemcpyDeviceToHost, s));
  GS_TRY(hipStreamSynchronize(s));
  if (g->h_int[0] != 0) {
    g->diag.termination = FG_CG_NOT_POSITIVE_DEFINITE;
    return FG_LINEAR_SOLVE_FAILED;
  }
  // reduced right-hand side
  if (E)
    hipLaunchKernelGGL(k_gs_elim<eRhs>, dim3(blocks_for(E)), dim3(kBlock), 0, s, E, (int)g->rw, (int)g->ew,
                       (int)g->retained, g->d_eptr, g->d_eidx, g->d_em, g->d_cross, g->d_fac_e, g->d_g, g->d_D,
                       (const double*)nullptr, lambda, g->d_t, (double*)nullptr, (const GcgState*)nullptr);
  hipLaunchKernelGGL(k_gs_ret<rRhs>, dim3(rblocks), dim3(kBlock), 0, s, R, (int)g->rw, (int)g->ew, g->d_rptr,
                     g->d_ridx, g->d_em, g->d_cross, g->d_gram_r, g->d_g, g->d_D, g->d_t, (const double*)nullptr,
                     lambda, g->d_rhs, (const double*)nullptr, g->d_parts, (const GcgState*)nullptr);
  // PCG
  hipLaunchKernelGGL(k_gcg_init, dim3(rblocks), dim3(kBlock), 0, s, R, (int)g->rw, g->d_rptr, g->d_fac_r, lambda,
                     g->d_rhs, g->d_x, g->d_r, g->d_z, g->d_p, g->d_parts);
  const int max_it = (int)std::min<size_t>(g->cfg.max_iterations, (size_t)std::numeric_limits<int>::max());
  hipLaunchKernelGGL(k_gcg_start, dim3(1), dim3(kBlock), 0, s, (const double*)g->d_parts, rblocks,
                     g->cfg.relative_tolerance, max_it, g->d_st);
  GS_TRY(hipGetLastError());
  const int kBatch = 4;
  int launched = 0, batch = 0;
  bool stop = false;
  while (!stop && launched < max_it) {
    for (int i = 0; i < kBatch; ++i) {
      gs_product(g, g->d_p, lambda, g->d_q, g->d_st);
      hipLaunchKernelGGL(k_gcg_alpha, dim3(1), dim3(kBlock), 0, s, (const double*)g->d_parts, rblocks, g->d_st);
      hipLaunchKernelGGL(k_gcg_update, dim3(rblocks), dim3(kBlock), 0, s, R, (int)g->rw, g->d_rptr, g->d_fac_r,
                         lambda, (const double*)g->d_p, (const double*)g->d_q, g->d_x, g->d_r, g->d_z, g->d_parts,
                         (const GcgState*)g->d_st);
      hipLaunchKernelGGL(k_gcg_beta, dim3(1), dim3(kBlock), 0, s, (const double*)g->d_parts, rblocks, g->d_st);
      hipLaunchKernelGGL(k_gcg_direction, dim3(blocks_for(g->retained)), dim3(kBlock), 0, s, (int)g->retained,
                         (const double*)g->d_z, g->d_p, (const GcgState*)g->d_st);
    }
    GS_TRY(hipGetLastError());
    launched += kBatch;
    const int slot = batch & 1;
    GS_TRY(hipMemcpyAsync(&g->h_st[1 + slot], g->d_st, sizeof(GcgState), hipMemcpyDeviceToHost, s));
    GS_TRY(hipEventRecord(g->poll[slot], s));
    if (batch > 0) {
      GS_TRY(hipEventSynchronize(g->poll[slot ^ 1]));
      if (g->h_st[1 + (slot ^ 1)].done) stop = true;
    }
    ++batch;
  }
  GS_TRY(hipMemcpyAsync(&g->h_st[0], g->d_st, sizeof(GcgState), hipMemcpyDeviceToHost, s));
  // recover_step (schur.rs:312-344): delta_c = x / D, z_e = (C + lambda)^-1 (-g/D - E^T x)
  hipLaunchKernelGGL(k_gs_unscale, dim3(blocks_for(g->retained)), dim3(kBlock), 0, s, (int)g->retained,
                     (const double*)g->d_x, (const double*)g->d_D, g->d_delta);
  if (E)
    hipLaunchKernelGGL(k_gs_elim<eRecover>, dim3(blocks_for(E)), dim3(kBlock), 0, s, E, (int)g->rw, (int)g->ew,
                       (int)g->retained, g->d_eptr, g->d_eidx, g->d_em, g->d_cross, g->d_fac_e, g->d_g, g->d_D,
                       (const double*)g->d_x, lambda, g->d_t, g->d_delta, (const GcgState*)nullptr);
  // verify_step (schur.rs:346-397): reduced residual, then the full model
  if (E)
    hipLaunchKernelGGL(k_gs_elim<eSchur>, dim3(blocks_for(E)), dim3(kBlock), 0, s, E, (int)g->rw, (int)g->ew,
                       (int)g->retained, g->d_eptr, g->d_eidx, g->d_em, g->d_cross, g->d_fac_e, g->d_g, g->d_D,
                       (const double*)g->d_x, lambda, g->d_t, (double*)nullptr, (const GcgState*)nullptr);
  hipLaunchKernelGGL(k_gs_ret<rResid>, dim3(rblocks), dim3(kBlock), 0, s, R, (int)g->rw, (int)g->ew, g->d_rptr,
                     g->d_ridx, g->d_em, g->d_cross, g->d_gram_r, g->d_g, g->d_D, g->d_t, (const double*)g->d_x,
                     lambda, g->d_q, (const double*)g->d_rhs, g->d_parts, (const GcgState*)nullptr);
  gs_sum(g, rblocks, 0);
  const int ne = (int)g->h_em.size();
  hipLaunchKernelGGL(k_gs_jv, dim3(blocks_for(ne)), dim3(kBlock), 0, s, ne, (int)g->rw, (int)g->ew,
                     (int)g->retained, g->d_em, g->d_vals, (const double*)g->d_delta, g->d_jd, g->d_parts);
  gs_sum(g, blocks_for(ne), 1);
  if (R) {
    hipLaunchKernelGGL(k_gs_normal, dim3(blocks_for(R * g->rw)), dim3(kBlock), 0, s, R, (int)g->rw, 0, true,
                       g->d_rptr, g->d_ridx, g->d_em, g->d_vals, g->d_res, (const double*)g->d_jd, g->d_g, g->d_D,
                       (const double*)g->d_delta, lambda, g->d_nr, g->d_parts);
    gs_sum(g, blocks_for(R * g->rw), 2);
  }
  if (E) {
    hipLaunchKernelGGL(k_gs_normal, dim3(blocks_for(E * g->ew)), dim3(kBlock), 0, s, E, (int)g->ew,
                       (int)g->retained, false, g->d_eptr, g->d_eidx, g->d_em, g->d_vals, g->d_res,
                       (const double*)g->d_jd, g->d_g, g->d_D, (const double*)g->d_delta, lambda, g->d_nr,
                       g->d_parts);
    gs_sum(g, blocks_for(E * g->ew), 3);
  }
  // ||delta||_inf and finiteness
  GS_TRY(hipMemsetAsync(g->d_flags, 0, sizeof(int), s));
  const int mb = std::min(1024, blocks_for(g->dim));
  double* absparts = g->d_parts;  // the sums above were enqueued first
  hipLaunchKernelGGL(k_gs_absmax, dim3(mb), dim3(kBlock), 0, s, (int)g->dim, (const double*)g->d_delta, absparts,
                     g->d_flags);
  GS_TRY(hipGetLastError());
  std::vector<double> am(mb);
  GS_TRY(hipMemcpyAsync(am.data(), absparts, mb * sizeof(double), hipMemcpyDeviceToHost, s));
  GS_TRY(hipMemcpyAsync(g->h_sums, g->d_sums, 4 * kParts * sizeof(double), hipMemcpyDeviceToHost, s));
  GS_TRY(hipMemcpyAsync(g->h_int, g->d_flags, sizeof(int), hipMemcpyDeviceToHost, s));
  GS_TRY(hipStreamSynchronize(s));
  const GcgState& st = g->h_st[0];
  g->diag.iterations = (size_t)st.iters;
  g->iterations += (size_t)st.iters;
  g->diag.termination = st.term == 0 ? FG_CG_CONVERGED
                                     : st.term == 1 ? FG_CG_ITERATION_LIMIT
                                                    : st.term == 2 ? FG_CG_NOT_POSITIVE_DEFINITE : FG_CG_NONE;
  auto relative = [](double a, double r) { return r == 0.0 ? a : a / r; };  // linear_model.rs relative()
  if (st.term == 0) g->diag.estimated_relative_residual = relative(std::sqrt(st.rr), std::sqrt(st.bb));
  const double* rs = g->h_sums;              // ||Sx - rhs||^2, ||rhs||^2
  const double jd2 = g->h_sums[kParts + 5];  // ||J delta||^2
  double n2 = 0, n2s = 0, gd = 0, g2 = 0, g2s = 0;
  for (int slot = 2; slot <= 3; ++slot) {
    if ((slot == 2 && !R) || (slot == 3 && !E)) continue;
    const double* v = g->h_sums + slot * kParts;
    n2 += v[0];
    n2s += v[1];
    gd += v[2];
    g2 += v[3];
    g2s += v[4];
  }
  g->diag.reduced_relative_residual = relative(std::sqrt(rs[0]), std::sqrt(rs[1]));
  g->diag.original_relative_residual = relative(std::sqrt(n2), std::sqrt(g2));
  g->diag.scaled_relative_residual = relative(std::sqrt(n2s), std::sqrt(g2s));
  const bool finite_step = g->h_int[0] == 0;
  const double pred = -gd - 0.5 * jd2;
  bool bad = !finite_step || !std::isfinite(pred);
  for (double r : {g->diag.reduced_relative_residual, g->diag.scaled_relative_residual,
                   g->diag.original_relative_residual})
    if (!std::isfinite(r)) bad = true;
  double mx = 0.0;
  for (double v : am) mx = std::max(mx, v);
  g->last_step_norm = finite_step ? mx : kNaN;
  if (bad) {
    g->diag.termination = FG_CG_NON_FINITE;
    return FG_LINEAR_SOLVE_FAILED;
  }
  if (g->diag.termination != FG_CG_CONVERGED) return FG_LINEAR_SOLVE_FAILED;
  *predicted = pred;
  return FG_OK;
}

}  // namespace

extern "C" {

fg_status fg_gschur_create(const fg_schur_config* cfg, size_t retained, size_t retained_width,
                           size_t eliminated_width, fg_gschur** out) {
  *out = nullptr;
  if (cfg->precision != FG_FP64 || cfg->preconditioner != FG_PRECOND_BLOCK_JACOBI) return FG_INVALID_OPTIONS;
  if (retained_width > (size_t)kGW || eliminated_width > (size_t)kGW) return FG_UNSUPPORTED_STRUCTURE;
  fg_gschur* g = new fg_gschur();
  g->cfg = *cfg;
  g->retained = retained;
  g->rw = retained_width;
  g->ew = eliminated_width;
  gs_reset_diag(g, 0.0);
  std::memset(&g->tim, 0, sizeof(g->tim));
  bool ok = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc((void**)&g->d_flags, 4 * sizeof(int)) == hipSuccess &&
            hipMalloc((void**)&g->d_st, sizeof(GcgState)) == hipSuccess &&
            hipMalloc((void**)&g->d_sums, 4 * kParts * sizeof(double)) == hipSuccess &&
            hipHostMalloc((void**)&g->h_st, 3 * sizeof(GcgState), 0) == hipSuccess &&
            hipHostMalloc((void**)&g->h_sums, 4 * kParts * sizeof(double), 0) == hipSuccess &&
            hipHostMalloc((void**)&g->h_int, 4 * sizeof(int), 0) == hipSuccess &&
            hipEventCreateWithFlags(&g->poll[0], hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&g->poll[1], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    fg_gschur_destroy(g);
    return FG_DEVICE_ERROR;
  }
  *out = g;
  return FG_OK;
}

void fg_gschur_destroy(fg_gschur* g) {
  if (!g) return;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  gs_free(g);
  if (g->h_st) (void)hipHostFree(g->h_st);
  if (g->h_sums) (void)hipHostFree(g->h_sums);
  if (g->h_int) (void)hipHostFree(g->h_int);
  for (hipEvent_t e : g->poll)
    if (e) (void)hipEventDestroy(e);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

fg_status fg_gschur_set_config(fg_gschur* g, const fg_schur_config* cfg) {
  if (cfg->precision != FG_FP64 || cfg->preconditioner != FG_PRECOND_BLOCK_JACOBI) return FG_INVALID_OPTIONS;
  g->cfg = *cfg;
  return FG_OK;
}

// prepare (schur.rs:409-431)
fg_status fg_gschur_prepare(fg_gschur* g, size_t dimension) {
  g->prepared = g->ready = false;
  g->solves = g->iterations = 0;
  gs_reset_diag(g, 0.0);
  std::memset(&g->tim, 0, sizeof(g->tim));
  if (g->rw == 0 || g->ew == 0 || g->retained > dimension || g->retained % g->rw != 0 ||
      (dimension - g->retained) % g->ew != 0 || g->cfg.max_iterations == 0 ||
      !(g->cfg.relative_tolerance >= 0.0 && g->cfg.relative_tolerance < 1.0))
    return FG_INVALID_OPTIONS;
  if (dimension > (size_t)std::numeric_limits<int>::max() / 2) return FG_UNSUPPORTED_STRUCTURE;
  g->dim = dimension;
  g->h_delta.assign(dimension, 0.0);
  g->h_g.assign(dimension, 0.0);
  fg_gschur_clear(g);
  g->prepared = true;
  return FG_OK;
}

// clear (schur.rs:432-437)
void fg_gschur_clear(fg_gschur* g) {
  g->ready = false;
  g->step_valid = false;
  g->h_em.clear();
  g->h_res.clear();
  g->h_vals.clear();
  g->rows = 0;
  g->dirty = true;
}

// Schur::accumulate (schur.rs:438-482) for n emissions of the same shape; the whole batch is
// validated before any of it is appended.
fg_status fg_gschur_accumulate(fg_gschur* g, size_t n, size_t rows, const double* residual, size_t n_blocks,
                               const size_t* cols, const size_t* widths, const double* values) {
  g->ready = false;
  if (!g->prepared) return FG_INVALID_OPTIONS;
  size_t per = 0;
  for (size_t k = 0; k < n_blocks; ++k) per += rows * widths[k];
  std::vector<Em> add(n);
  for (size_t e = 0; e < n; ++e) {
    int left = -1, right = -1, loff = 0, roff = 0;
    size_t voff = 0;
    for (size_t k = 0; k < n_blocks; ++k) {
      const size_t col = cols[e * n_blocks + k], w = widths[k];
      if (col > g->dim || w > g->dim - col) return FG_DIMENSION_MISMATCH;
      if (col < g->retained) {
        if (col % g->rw != 0 || w != g->rw || col + w > g->retained || left >= 0) return FG_UNSUPPORTED_STRUCTURE;
        left = (int)(col / g->rw);
        loff = (int)voff;
      } else {
        if ((col - g->retained) % g->ew != 0 || w != g->ew || right >= 0) return FG_UNSUPPORTED_STRUCTURE;
        right = (int)((col - g->retained) / g->ew);
        roff = (int)voff;
      }
      voff += rows * w;
    }
    add[e] = Em{0, (int)rows, left, loff, right, roff, -1, 0};
  }
  if (g->rows + n * rows > (size_t)std::numeric_limits<int>::max() ||
      g->h_vals.size() + n * per > (size_t)std::numeric_limits<int>::max())
    return FG_UNSUPPORTED_STRUCTURE;
  for (size_t e = 0; e < n; ++e) {
    Em m = add[e];
    m.row0 = (int)g->rows;
    const size_t base = g->h_vals.size();
    if (m.rblk >= 0) m.roff += (int)base;
    if (m.eblk >= 0) m.eoff += (int)base;
    g->h_vals.insert(g->h_vals.end(), values + e * per, values + (e + 1) * per);
    g->h_res.insert(g->h_res.end(), residual + e * rows, residual + (e + 1) * rows);
    g->rows += rows;
    g->h_em.push_back(m);
  }
  g->dirty = true;
  return FG_OK;
}

fg_status fg_gschur_gradient_norm(fg_gschur* g, double* out) {
  if (!g->prepared) return FG_INVALID_OPTIONS;
  GS_CHECK(gs_finalize(g));
  double m = 0.0;
  for (double v : g->h_g) {
    if (!std::isfinite(v)) return FG_INVALID_EVALUATION;
    m = std::max(m, std::fabs(v));
  }
  *out = m;
  return FG_OK;
}

// prepare_damping (schur.rs:490-526): scales, block Grams, cross blocks
fg_status fg_gschur_prepare_damping(fg_gschur* g, double floor) {
  g->ready = false;
  if (!g->prepared) return FG_INVALID_OPTIONS;
  if (!std::isfinite(floor) || floor <= 0.0) return FG_INVALID_OPTIONS;
  GS_CHECK(gs_finalize(g));
  hipStream_t s = g->stream;
  const int R = (int)nrb(g), E = (int)neb(g);
  GS_TRY(hipMemsetAsync(g->d_flags, 0, sizeof(int), s));
  if (R)
    hipLaunchKernelGGL(k_gs_columns, dim3(blocks_for(R * g->rw)), dim3(kBlock), 0, s, R, (int)g->rw, 0, true,
                       g->d_rptr, g->d_ridx, g->d_em, g->d_vals, g->d_res, floor, (double*)nullptr, g->d_D,
                       g->d_flags);
  if (E)
    hipLaunchKernelGGL(k_gs_columns, dim3(blocks_for(E * g->ew)), dim3(kBlock), 0, s, E, (int)g->ew, (int)g->retained,
                       false, g->d_eptr, g->d_eidx, g->d_em, g->d_vals, g->d_res, floor, (double*)nullptr, g->d_D,
                       g->d_flags);
  const int tr = (int)(g->rw * (g->rw + 1) / 2), te = (int)(g->ew * (g->ew + 1) / 2);
  if (R)
    hipLaunchKernelGGL(k_gs_gram, dim3(blocks_for((size_t)R * tr)), dim3(kBlock), 0, s, R, (int)g->rw, 0, true,
                       g->d_rptr, g->d_ridx, g->d_em, g->d_vals, (const double*)g->d_D, g->d_gram_r);
  if (E)
    hipLaunchKernelGGL(k_gs_gram, dim3(blocks_for((size_t)E * te)), dim3(kBlock), 0, s, E, (int)g->ew,
                       (int)g->retained, false, g->d_eptr, g->d_eidx, g->d_em, g->d_vals, (const double*)g->d_D,
                       g->d_gram_e);
  if (g->nedges)
    hipLaunchKernelGGL(k_gs_cross, dim3(blocks_for(g->nedges * g->rw * g->ew)), dim3(kBlock), 0, s, (int)g->nedges,
                       (int)g->rw, (int)g->ew, (int)g->retained, g->d_edge_em, g->d_em, g->d_vals,
                       (const double*)g->d_D, g->d_cross, g->d_flags);
  GS_TRY(hipGetLastError());
  GS_TRY(hipMemcpyAsync(g->h_int, g->d_flags, sizeof(int), hipMemcpyDeviceToHost, s));
  GS_TRY(hipStreamSynchronize(s));
  if (g->h_int[0] != 0) return FG_INVALID_EVALUATION;
  g->ready = true;
  return FG_OK;
}

fg_status fg_gschur_solve_damped(fg_gschur* g, double lambda, const double** delta, double* predicted) {
  gs_reset_diag(g, lambda);
  g->step_valid = false;
  if (!g->ready || !std::isfinite(lambda) || lambda < 0.0) return FG_INVALID_OPTIONS;
  g->solves += 1;
  double pred = 0.0;
  const fg_status st = gs_solve_damped(g, lambda, &pred);
  if (g->cfg.on_solve) g->cfg.on_solve(&g->diag, g->cfg.on_solve_user);
  if (st != FG_OK) return st;
  if (g->dim) GS_TRY(hipMemcpy(g->h_delta.data(), g->d_delta, g->dim * sizeof(double), hipMemcpyDeviceToHost));
  g->step_valid = true;
  if (delta) *delta = g->h_delta.data();
  if (predicted) *predicted = pred;
  return FG_OK;
}

// solve (schur.rs:486-489): prepare_damping(1e-6) then solve_damped(0)
fg_status fg_gschur_solve(fg_gschur* g, const double** delta) {
  GS_CHECK(fg_gschur_prepare_damping(g, 1e-6));
  return fg_gschur_solve_damped(g, 0.0, delta, nullptr);
}

void fg_gschur_diagnostics_get(const fg_gschur* g, fg_schur_diagnostics* d) { *d = g->diag; }
void fg_gschur_statistics(const fg_gschur* g, size_t* solves, size_t* iterations) {
  if (solves) *solves = g->solves;
  if (iterations) *iterations = g->iterations;
}
double fg_gschur_last_step_norm(const fg_gschur* g) { return g->last_step_norm; }

// CovarianceBackend for Schur (src/schur.rs:548-556) over the accumulated emissions.
fg_status fg_gschur_covariance(fg_gschur* g, const size_t* offsets, const size_t* widths, size_t n_blocks,
                               const fg_covariance_options* options, double* out) {
  const double tol = options->relative_pivot_tolerance;
  if (!std::isfinite(tol) || tol < 0.0 || tol >= 1.0) return FG_INVALID_RANK_TOLERANCE;
  if (!g->prepared) return FG_INVALID_OPTIONS;
  const size_t n = g->dim;
  std::vector<size_t> cols;
  for (size_t i = 0; i < n_blocks; ++i) {  // SelectedCovariance::prepare (src/covariance.rs:180-207)
    const size_t o = offsets[i], w = widths[i];
    for (size_t j = 0; j < i; ++j)
      if (offsets[j] == o) return FG_DUPLICATE_COVARIANCE_BLOCK;
    const bool ret = o < g->retained && o % g->rw == 0 && w == g->rw;
    const bool elim = o >= g->retained && o < n && (o - g->retained) % g->ew == 0 && w == g->ew;
    if (!ret && !elim) return FG_UNKNOWN_KEY;
    for (size_t c = o; c < o + w; ++c) cols.push_back(c);
  }
  GS_CHECK(gs_finalize(g));
  if (n > 0 && n > SIZE_MAX / 8 / n) return FG_INVALID_INFORMATION;
  double* h = nullptr;
  GS_TRY(hipMalloc(&h, std::max<size_t>(1, n * n) * sizeof(double)));
  fg_status st = FG_DEVICE_ERROR;
  const int ne = (int)g->h_em.size();
  if (hipMemsetAsync(h, 0, n * n * sizeof(double), g->stream) == hipSuccess) {
    if (ne) {
      const int R = (int)nrb(g), E = (int)neb(g);
      if (R)
        hipLaunchKernelGGL(k_gs_info_ret, dim3(blocks_for(R)), dim3(kBlock), 0, g->stream, R, (int)g->rw, n,
                           (const int*)g->d_rptr, (const int*)g->d_ridx, (const Em*)g->d_em,
                           (const double*)g->d_vals, h);
      if (E)
        hipLaunchKernelGGL(k_gs_info_elim, dim3(blocks_for(E)), dim3(kBlock), 0, g->stream, E, (int)g->rw,
                           (int)g->ew, (int)g->retained, n, (const int*)g->d_eptr, (const int*)g->d_eidx,
                           (const Em*)g->d_em, (const double*)g->d_vals, h);
    }
    if (hipGetLastError() == hipSuccess) st = fg::dense_covariance(g->stream, h, n, cols.data(), cols.size(), tol, out);
  }
  (void)hipStreamSynchronize(g->stream);
  (void)hipFree(h);
  return st;
}

}  // extern "C"
