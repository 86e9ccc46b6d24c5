// fg_lsmr_kernels.h -- device side of the LSMR backend (src/lsmr.rs): the right block
// preconditioner, the vector updates and the scalar recurrences of LSMR (faer 0.24.4's
// matrix_free::lsmr, i.e. Fong & Saunders 2011 without reorthogonalization). The operator
// products are the fused point-major and camera-major passes k_points<kLsmr> and
// k_cameras<cLsmr>.
//
// Vectors live in the library's internal order: cameras (9 nc, reference order) then points
// (3 np, slot order). The Jacobian rows of U are kept in the point-major slot order (the
// order of the records, so the point pass reads and writes them coalesced and the camera
// pass gathers them), the n damping rows in coordinate order. Row order does not change
// LSMR: only norms and J / J^T products of U enter. LSMR's u and v are kept unnormalized
// (U, V) with their norms beta, alpha in device state, so no kernel ever waits for a norm it
// has not itself summed.
#pragma once
#include "fg_kernels.h"

namespace fg {


// One block of P = L^-T (transpose = 0) or P^T = L^-1 (transpose = 1) in place on x: camera
// blocks (9, factor lc) for c < nc, point blocks (3, factor lp) after. A zero factor
// (unobserved or failed block) leaves the block as it is (diagonal treatment).
__device__ __forceinline__ void lsmr_block_precond(int blk, int nc, const double* __restrict__ lc,
                                                   const double* __restrict__ lp, double* x, int width,
                                                   bool transpose) {
  if (width == 9) {
    double l[45];
#pragma unroll
    for (int i = 0; i < 45; ++i) l[i] = lc[45 * blk + i];
    if (l[0] == 0.0) return;
    if (transpose) solve_lower_packed<double, 9>(l, x);
    else solve_upper_packed<double, 9>(l, x);
  } else {
    double l[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) l[i] = lp[6 * (blk - nc) + i];
    if (l[0] == 0.0) return;
    if (transpose) solve_lower_packed<double, 3>(l, x);
    else solve_upper_packed<double, 3>(l, x);
  }
}

// ws = P V / (alpha D) (scaled modes) or P V / alpha (identity), one thread per block; with
// final != 0 it instead writes x = P y / D (or P y), the recovered step in original
// coordinates.
template <typename T>
__global__ void k_lsmr_pv(int nc, int np, const double* __restrict__ v, const double* __restrict__ lc,
                          const double* __restrict__ lp, const T* __restrict__ dc, const T* __restrict__ dp,
                          int block, int scaled, int final_step, const LsmrState* __restrict__ st,
                          double* __restrict__ ws) {
  const int blk = blockIdx.x * blockDim.x + threadIdx.x;
  if (blk >= nc + np) return;
  if (!final_step && st->done) return;
  const int width = blk < nc ? 9 : 3;
  const int off = blk < nc ? 9 * blk : 9 * nc + 3 * (blk - nc);
  double x[9];
  for (int i = 0; i < width; ++i) x[i] = v[off + i];
  if (block) lsmr_block_precond(blk, nc, lc, lp, x, width, false);
  const double inv_alpha = final_step ? 1.0 : 1.0 / st->alpha;
  for (int i = 0; i < width; ++i) {
    const double d = blk < nc ? double(dc[off + i]) : double(dp[3 * (blk - nc) + i]);
    ws[off + i] = scaled ? x[i] * inv_alpha / d : x[i] * inv_alpha;
  }
}

// beta (or alpha) = sqrt(sum of partials) into the state; which = 0 beta, 1 alpha.
static __global__ void k_lsmr_norm(const double* __restrict__ parts, int n, int which, LsmrState* st) {
  __shared__ double smem[kBlock / kWave];
  if (st->done) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += parts[i];
  s = block_sum(s, smem);
  if (threadIdx.x == 0) {
    const double r = sqrt(s);
    if (which == 0) st->beta_new = r;
    else st->alpha_new = r;
  }
}

// V = P^T z - (beta_new / alpha) V with z = (zt / D + sqrt(lambda) U_bottom) / beta_new
// (scaled) or (zt + sqrt(lambda) D U_bottom) / beta_new; one thread per block, partials of
// ||V||^2 per workgroup. init != 0: the first v (V_old = 0).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_lsmr_v(int nc, int np, const double* __restrict__ zt,
                                                   const double* __restrict__ u_bottom, const double* __restrict__ lc,
                                                   const double* __restrict__ lp, const T* __restrict__ dc,
                                                   const T* __restrict__ dp, int block, int scaled, int damped,
                                                   int init, const LsmrState* __restrict__ st, double* __restrict__ v,
                                                   double* __restrict__ parts) {
  __shared__ double smem[kBlock / kWave];
  if (st->done) return;
  const double beta_new = st->beta_new, root = st->root;
  const double keep = init ? 0.0 : beta_new / st->alpha;
  const int blk = blockIdx.x * blockDim.x + threadIdx.x;
  double ss = 0.0;
  if (blk < nc + np) {
    const int width = blk < nc ? 9 : 3;
    const int off = blk < nc ? 9 * blk : 9 * nc + 3 * (blk - nc);
    double z[9];
    for (int i = 0; i < width; ++i) {
      const double d = blk < nc ? double(dc[off + i]) : double(dp[3 * (blk - nc) + i]);
      double t = scaled ? zt[off + i] / d : zt[off + i];
      if (damped) t += (scaled ? root : root * d) * u_bottom[off + i];
      z[i] = beta_new > 0.0 ? t / beta_new : t;
    }
    if (block) lsmr_block_precond(blk, nc, lc, lp, z, width, true);
    for (int i = 0; i < width; ++i) {
      // the first v reads no old V: the buffer is uninitialized then, and 0 * NaN is NaN
      const double vn = init ? z[i] : z[i] - keep * v[off + i];
      v[off + i] = vn;
      ss += vn * vn;
    }
  }
  const double b = block_sum(ss, smem);
  if (threadIdx.x == 0) parts[blockIdx.x] = b;
}

// The scalar recurrences of one LSMR iteration (one thread). init != 0: the start values
// after beta_1 and alpha_1.
static __global__ void k_lsmr_scalar(LsmrState* st, int init) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  LsmrState s = *st;
  if (s.done) return;
  const double beta = s.beta_new, alpha = s.alpha_new;
  if (init) {
    s.beta = beta;
    s.alpha = alpha;
    s.alphabar = alpha;
    s.zetabar = alpha * beta;
    s.rho = 1.0;
    s.rhobar = 1.0;
    s.cbar = 1.0;
    s.sbar = 0.0;
    s.norm_atb = alpha * beta;
    s.abs_res = fabs(s.zetabar);
    s.iters = 0;
    if (!(s.norm_atb > 0.0)) {  // A^T b = 0: the zero step is exact
      s.done = 1;
      s.converged = 1;
      s.abs_res = 0.0;
    }
    *st = s;
    return;
  }
  const double rho_old = s.rho, rhobar_old = s.rhobar;
  s.rho = hypot(s.alphabar, beta);
  const double c = s.alphabar / s.rho, sn = beta / s.rho;
  const double theta_new = sn * alpha;
  s.alphabar = c * alpha;
  const double thetabar = s.sbar * s.rho;
  s.rhobar = hypot(s.cbar * s.rho, theta_new);
  s.cbar = s.cbar * s.rho / s.rhobar;
  s.sbar = theta_new / s.rhobar;
  const double zeta = s.cbar * s.zetabar;
  s.zetabar = -s.sbar * s.zetabar;
  s.f1 = thetabar * s.rho / (rho_old * rhobar_old);
  s.f2 = zeta / (s.rho * s.rhobar);
  s.f3 = theta_new / s.rho;
  s.beta = beta;
  s.alpha = alpha;
  s.iters += 1;
  s.abs_res = fabs(s.zetabar);
  if (s.abs_res <= s.rel_tol * s.norm_atb) {
    s.converged = 1;
    s.done = 2;  // the vector update of this iteration still runs
  } else if (s.iters >= s.max_iters) {
    s.done = 2;
  }
  *st = s;
}

// hbar = h - f1 hbar; y += f2 hbar; h = V/alpha - f3 h (or, init != 0, h = V/alpha, hbar = 0,
// y = 0). Marks the state finished after the last update.
static __global__ void k_lsmr_h(int n, const double* __restrict__ v, LsmrState* st, int init, double* __restrict__ h,
                                double* __restrict__ hbar, double* __restrict__ y) {
  const LsmrState s = *st;
  if (!init && s.done == 1) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double vi = s.alpha > 0.0 ? v[i] / s.alpha : 0.0;
    if (init) {
      h[i] = vi;
      hbar[i] = 0.0;
      y[i] = 0.0;
    } else {
      const double hb = h[i] - s.f1 * hbar[i];
      hbar[i] = hb;
      y[i] += s.f2 * hb;
      h[i] = vi - s.f3 * h[i];
    }
  }
}

// Ends an iteration: a state marked "last update pending" becomes finished.
static __global__ void k_lsmr_close(LsmrState* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && st->done == 2) st->done = 1;
}

// U = b: the negative weighted residual in slot order (zero on padding slots) and zero
// damping rows. Its norm comes from the first transpose pass.
template <typename T>
__global__ void k_lsmr_init_u(int nslots, int n_bottom, const int* __restrict__ slot_cam,
                              const Vec2<T>* __restrict__ res, Vec2<double>* __restrict__ u,
                              double* __restrict__ u_bottom) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nslots) {
    const Vec2<T> r = res[i];
    u[i] = slot_cam[i] >= 0 ? Vec2<double>{-double(r.x), -double(r.y)} : Vec2<double>{0.0, 0.0};
  }
  if (i < n_bottom) u_bottom[i] = 0.0;
}

// The scaled normal residual in block coordinates: per block, L^-1 (nr / D) and L^-1 (g / D)
// (the reference applies L^-1 D^-1 to both, src/lsmr.rs:440-452); partials of both squared
// norms in columns 0 and 1 (grid-stride over blocks).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_lsmr_scaled_residual(int nc, int np, const double* __restrict__ nr,
                                                                 const T* __restrict__ gc, const T* __restrict__ gp,
                                                                 const double* __restrict__ lc,
                                                                 const double* __restrict__ lp,
                                                                 const T* __restrict__ dc, const T* __restrict__ dp,
                                                                 double* __restrict__ parts) {
  __shared__ double smem[kBlock / kWave];
  double s0 = 0.0, s1 = 0.0;
  for (int blk = blockIdx.x * blockDim.x + threadIdx.x; blk < nc + np; blk += gridDim.x * blockDim.x) {
    const int width = blk < nc ? 9 : 3;
    const int off = blk < nc ? 9 * blk : 9 * nc + 3 * (blk - nc);
    double a[9], g[9];
    for (int i = 0; i < width; ++i) {
      const double d = blk < nc ? double(dc[off + i]) : double(dp[3 * (blk - nc) + i]);
      a[i] = nr[off + i] / d;
      g[i] = (blk < nc ? double(gc[off + i]) : double(gp[3 * (blk - nc) + i])) / d;
    }
    lsmr_block_precond(blk, nc, lc, lp, a, width, true);
    lsmr_block_precond(blk, nc, lc, lp, g, width, true);
    for (int i = 0; i < width; ++i) {
      s0 += a[i] * a[i];
      s1 += g[i] * g[i];
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

// Minimum and maximum of the column scales (partials: col 0 min, col 1 max).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_scale_range(int n_cam, int n_pt, const T* __restrict__ dc,
                                                        const T* __restrict__ dp, double* __restrict__ parts) {
  __shared__ double smin[kBlock / kWave], smax[kBlock / kWave];
  double lo = INFINITY, hi = -INFINITY;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_cam + n_pt; i += gridDim.x * blockDim.x) {
    const double d = i < n_cam ? double(dc[i]) : double(dp[i - n_cam]);
    lo = fmin(lo, d);
    hi = fmax(hi, d);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
  }
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (lane == 0) {
    smin[wave] = lo;
    smax[wave] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x / kWave); ++w) {
      lo = fmin(lo, smin[w]);
      hi = fmax(hi, smax[w]);
    }
    parts[kParts * blockIdx.x] = lo;
    parts[kParts * blockIdx.x + 1] = hi;
  }
}

}  // namespace fg
