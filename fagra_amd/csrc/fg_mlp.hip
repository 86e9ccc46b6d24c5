// fg_mlp.hip -- host side of the preconditioner network: parameters, inference, gradients and
// training with the hand-written BF16-MFMA kernels of fg_mlp_kernels.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "fg_internal.h"
#include "fg_mlp_kernels.h"

using namespace fg;

#define FG_TRY(expr)                              \
  do {                                            \
    hipError_t e_ = (expr);                       \
    if (e_ != hipSuccess) return FG_DEVICE_ERROR; \
  } while (0)

namespace fg {

static int blocks_for(size_t n, int block = 256) { return (int)((n + block - 1) / block); }

fg_status mlp_reserve(fg_mlp* m, size_t n) {
  const size_t np = pad64(std::max<size_t>(n, 1));
  if (np <= m->cap) return FG_OK;
  (void)hipStreamSynchronize(m->stream);
  void** bufs[] = {(void**)&m->x, (void**)&m->z1, (void**)&m->z2, (void**)&m->h1, (void**)&m->h2, (void**)&m->y,
                   (void**)&m->dy, (void**)&m->dz2, (void**)&m->dz1, (void**)&m->target, (void**)&m->parts};
  for (void** b : bufs)
    if (*b) {
      (void)hipFree(*b);
      *b = nullptr;
    }
  FG_TRY(hipMalloc((void**)&m->x, np * kInPad * 2));
  FG_TRY(hipMalloc((void**)&m->z1, np * kHidden * 4));
  FG_TRY(hipMalloc((void**)&m->z2, np * kHidden * 4));
  FG_TRY(hipMalloc((void**)&m->h1, np * kHidden * 2));
  FG_TRY(hipMalloc((void**)&m->h2, np * kHidden * 2));
  FG_TRY(hipMalloc((void**)&m->y, np * kOutPad * 4));
  FG_TRY(hipMalloc((void**)&m->dy, np * kOutPad * 2));
  FG_TRY(hipMalloc((void**)&m->dz2, np * kHidden * 2));
  FG_TRY(hipMalloc((void**)&m->dz1, np * kHidden * 2));
  FG_TRY(hipMalloc((void**)&m->target, np * 45 * 4));
  FG_TRY(hipMalloc((void**)&m->parts, (np / 256 + 1) * sizeof(double)));
  FG_TRY(hipMemset(m->x, 0, np * kInPad * 2));
  // hipMemset runs on the null stream, which does not order against the non-blocking work
  // streams: finish it before any later upload can be overtaken.
  FG_TRY(hipDeviceSynchronize());
  m->cap = np;
  return FG_OK;
}

void mlp_forward(fg_mlp* m, hipStream_t stream, size_t npad, const uint16_t* x, float* y) {
  // 16 rows per four-wave workgroup: a 1,536-row batch spreads over 96 workgroups
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)(npad / 16)), dim3(256), 0, stream, x, m->w1, m->params + kB1,
                     m->w2, m->params + kB2, m->w3, m->params + kB3, y);
}

}  // namespace fg

namespace {

fg_status convert(fg_mlp* m) {
  hipLaunchKernelGGL(k_mlp_convert, dim3(blocks_for(kHidden * kHidden)), dim3(256), 0, m->stream, m->params, m->w1,
                     m->w1t, m->w2, m->w2t, m->w3, m->w3t);
  FG_TRY(hipGetLastError());
  return FG_OK;
}

// Upload n feature rows (47 f32 each, host) as padded bf16 rows and the optional targets.
fg_status upload_batch(fg_mlp* m, size_t n, const float* features, const float* targets) {
  FG_TRY((hipError_t)(mlp_reserve(m, n) == FG_OK ? hipSuccess : hipErrorOutOfMemory));
  const size_t np = pad64(n);
  std::vector<uint16_t> xb(np * kInPad, 0);
  for (size_t b = 0; b < n; ++b)
    for (int k = 0; k < kIn; ++k) xb[b * kInPad + k] = f2bf(features[b * kIn + k]);
  FG_TRY(hipMemcpyAsync(m->x, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, m->stream));
  if (targets) FG_TRY(hipMemcpyAsync(m->target, targets, n * 45 * 4, hipMemcpyHostToDevice, m->stream));
  FG_TRY(hipStreamSynchronize(m->stream));
  return FG_OK;
}

// Forward with stored activations, loss, backward into m->grads (zeroed first).
fg_status loss_and_grad(fg_mlp* m, size_t n, double* loss) {
  const size_t np = pad64(n);
  const unsigned nb = (unsigned)(np / 64);
  FG_TRY(hipMemsetAsync(m->grads, 0, kParams * 4, m->stream));
  hipLaunchKernelGGL(k_mlp_forward, dim3(nb), dim3(256), 0, m->stream, m->x, m->w1, m->params + kB1, m->w2,
                     m->params + kB2, m->w3, m->params + kB3, m->z1, m->h1, m->z2, m->h2, m->y);
  const int lb = blocks_for(np);
  hipLaunchKernelGGL(k_mlp_loss, dim3(lb), dim3(256), 0, m->stream, (int)n, (int)np, m->y, m->target, m->dy,
                     m->parts);
  // layer 3: dW3 = dY^T H2, dH2 = dY W3 -> dZ2
  const unsigned wb = (unsigned)((np + 255) / 256);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(wb), dim3(256), 0, m->stream, (int)np, kOutPad, kHidden, m->dy, m->h2,
                     m->grads + kW3, m->grads + kB3);
  hipLaunchKernelGGL(k_mlp_backward_data, dim3(nb), dim3(256), 0, m->stream, kOutPad, kHidden, m->dy, m->w3t, m->z2,
                     m->dz2);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(wb), dim3(256), 0, m->stream, (int)np, kHidden, kHidden, m->dz2, m->h1,
                     m->grads + kW2, m->grads + kB2);
  hipLaunchKernelGGL(k_mlp_backward_data, dim3(nb), dim3(256), 0, m->stream, kHidden, kHidden, m->dz2, m->w2t,
                     m->z1, m->dz1);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(wb), dim3(256), 0, m->stream, (int)np, kHidden, kInPad, m->dz1, m->x,
                     m->grads + kW1, m->grads + kB1);
  FG_TRY(hipGetLastError());
  if (loss) {
    std::vector<double> parts(lb);
    FG_TRY(hipMemcpyAsync(parts.data(), m->parts, lb * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    FG_TRY(hipStreamSynchronize(m->stream));
    *loss = std::accumulate(parts.begin(), parts.end(), 0.0);
  }
  return FG_OK;
}

}  // namespace

extern "C" {

size_t fg_mlp_param_count(void) { return kParams; }

fg_status fg_mlp_create(uint64_t seed, fg_mlp** out) {
  *out = nullptr;
  fg_mlp* m = new fg_mlp();
  auto fail = [&]() {
    fg_mlp_destroy(m);
    return FG_DEVICE_ERROR;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) return fail();
  float** f32[] = {&m->params, &m->grads, &m->adam_m, &m->adam_v};
  for (float** p : f32)
    if (hipMalloc((void**)p, kParams * 4) != hipSuccess) return fail();
  uint16_t** w[] = {&m->w1, &m->w1t, &m->w2, &m->w2t, &m->w3, &m->w3t};
  const size_t wsz[] = {kHidden * kInPad, kHidden * kInPad, kHidden * kHidden, kHidden * kHidden,
                        kOutPad * kHidden, kOutPad * kHidden};
  for (int i = 0; i < 6; ++i)
    if (hipMalloc((void**)w[i], wsz[i] * 2) != hipSuccess) return fail();
  if (hipHostMalloc((void**)&m->h_parts, 64 * sizeof(double), 0) != hipSuccess) return fail();
  // Initialization: uniform(+-sqrt(6/(in+out))) hidden layers, zero output layer, so the initial
  // correction is G = I (plain block-Jacobi).
  std::vector<float> p(kParams, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int rows, int cols, int used_cols) {
    std::uniform_real_distribution<float> u(-std::sqrt(6.f / (used_cols + rows)), std::sqrt(6.f / (used_cols + rows)));
    for (int o = 0; o < rows; ++o)
      for (int k = 0; k < cols; ++k) p[off + o * cols + k] = k < used_cols ? u(rng) : 0.f;
  };
  fill(kW1, kHidden, kInPad, kIn);
  fill(kW2, kHidden, kHidden, kHidden);
  if (fg_mlp_set_params(m, p.data()) != FG_OK) return fail();
  *out = m;
  return FG_OK;
}

void fg_mlp_destroy(fg_mlp* m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  void* bufs[] = {m->params, m->grads, m->adam_m, m->adam_v, m->w1, m->w1t, m->w2, m->w2t, m->w3, m->w3t,
                  m->x, m->z1, m->z2, m->h1, m->h2, m->y, m->dy, m->dz2, m->dz1, m->target, m->parts};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (m->range) (void)hipFree(m->range);
  if (m->h_parts) (void)hipHostFree(m->h_parts);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

fg_status fg_mlp_set_params(fg_mlp* m, const float* params) {
  std::vector<float> p(params, params + kParams);
  // padded input columns and padded output rows stay zero
  for (int o = 0; o < kHidden; ++o)
    for (int k = kIn; k < kInPad; ++k) p[kW1 + o * kInPad + k] = 0.f;
  for (int o = kOut; o < kOutPad; ++o) {
    for (int k = 0; k < kHidden; ++k) p[kW3 + o * kHidden + k] = 0.f;
    p[kB3 + o] = 0.f;
  }
  FG_TRY(hipMemcpyAsync(m->params, p.data(), kParams * 4, hipMemcpyHostToDevice, m->stream));
  FG_TRY(hipMemsetAsync(m->adam_m, 0, kParams * 4, m->stream));
  FG_TRY(hipMemsetAsync(m->adam_v, 0, kParams * 4, m->stream));
  m->step = 0;
  if (convert(m) != FG_OK) return FG_DEVICE_ERROR;
  FG_TRY(hipStreamSynchronize(m->stream));
  return FG_OK;
}

fg_status fg_mlp_set_input_range(fg_mlp* m, const float* lo, const float* hi) {
  if (!lo || !hi) {  // clear: the network applies to every camera
    if (m->range) {
      (void)hipStreamSynchronize(m->stream);
      (void)hipFree(m->range);
      m->range = nullptr;
    }
    return FG_OK;
  }
  float r[2 * kIn];
  for (int k = 0; k < kIn; ++k) {
    if (!(lo[k] <= hi[k])) return FG_INVALID_OPTIONS;  // also rejects NaN bounds
    r[k] = lo[k];
    r[kIn + k] = hi[k];
  }
  if (!m->range) FG_TRY(hipMalloc((void**)&m->range, sizeof(r)));
  FG_TRY(hipMemcpyAsync(m->range, r, sizeof(r), hipMemcpyHostToDevice, m->stream));
  FG_TRY(hipStreamSynchronize(m->stream));
  return FG_OK;
}

fg_status fg_mlp_get_params(fg_mlp* m, float* params) {
  FG_TRY(hipMemcpyAsync(params, m->params, kParams * 4, hipMemcpyDeviceToHost, m->stream));
  FG_TRY(hipStreamSynchronize(m->stream));
  return FG_OK;
}

fg_status fg_mlp_forward(fg_mlp* m, size_t n, const float* features, float* out) {
  if (n == 0) return FG_OK;
  fg_status st = upload_batch(m, n, features, nullptr);
  if (st != FG_OK) return st;
  const size_t np = pad64(n);
  mlp_forward(m, m->stream, np, m->x, m->y);
  FG_TRY(hipGetLastError());
  std::vector<float> y(np * kOutPad);
  FG_TRY(hipMemcpyAsync(y.data(), m->y, y.size() * 4, hipMemcpyDeviceToHost, m->stream));
  FG_TRY(hipStreamSynchronize(m->stream));
  for (size_t b = 0; b < n; ++b)
    for (int k = 0; k < kOut; ++k) out[b * kOut + k] = y[b * kOutPad + k];
  return FG_OK;
}

fg_status fg_mlp_loss_grad(fg_mlp* m, size_t n, const float* features, const float* targets, double* loss,
                           float* grads) {
  if (n == 0) return FG_INVALID_OPTIONS;
  fg_status st = upload_batch(m, n, features, targets);
  if (st != FG_OK) return st;
  st = loss_and_grad(m, n, loss);
  if (st != FG_OK) return st;
  if (grads) {
    FG_TRY(hipMemcpyAsync(grads, m->grads, kParams * 4, hipMemcpyDeviceToHost, m->stream));
    FG_TRY(hipStreamSynchronize(m->stream));
  }
  return FG_OK;
}

// Adam epochs over host data (features n x 47, targets n x 45), shuffled per epoch with seed.
fg_status fg_mlp_train(fg_mlp* m, size_t n, const float* features, const float* targets, size_t epochs,
                       size_t batch, double lr, uint64_t seed, double* epoch_loss) {
  if (n == 0 || batch == 0 || !(lr > 0.0)) return FG_INVALID_OPTIONS;
  batch = std::min(batch, n);
  std::mt19937_64 rng(seed);
  std::vector<size_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::vector<float> fb(batch * kIn), tb(batch * 45);
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  for (size_t e = 0; e < epochs; ++e) {
    std::shuffle(order.begin(), order.end(), rng);
    double total = 0.0;
    size_t count = 0;
    for (size_t s = 0; s + batch <= n; s += batch) {
      for (size_t b = 0; b < batch; ++b) {
        std::memcpy(&fb[b * kIn], features + order[s + b] * kIn, kIn * 4);
        std::memcpy(&tb[b * 45], targets + order[s + b] * 45, 45 * 4);
      }
      fg_status st = upload_batch(m, batch, fb.data(), tb.data());
      if (st != FG_OK) return st;
      double loss = 0.0;
      st = loss_and_grad(m, batch, &loss);
      if (st != FG_OK) return st;
      total += loss;
      ++count;
      m->step += 1;
      const float c1 = 1.f - std::pow(beta1, (float)m->step), c2 = 1.f - std::pow(beta2, (float)m->step);
      hipLaunchKernelGGL(k_adam, dim3(blocks_for(kParams)), dim3(256), 0, m->stream, kParams, m->params, m->grads,
                         m->adam_m, m->adam_v, (float)lr, beta1, beta2, eps, c1, c2);
      if (convert(m) != FG_OK) return FG_DEVICE_ERROR;
    }
    FG_TRY(hipStreamSynchronize(m->stream));
    if (epoch_loss) epoch_loss[e] = count ? total / count : 0.0;
  }
  return FG_OK;
}

}  // extern "C"
