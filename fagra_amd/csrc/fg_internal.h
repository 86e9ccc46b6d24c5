// fg_internal.h -- shared host-side declarations between the backend and the network code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/fagra_gpu.h"

// Preconditioner network: f32 master parameters, bf16 operand copies, Adam state and
// batch scratch (grown on demand outside the solve loop).
struct fg_mlp {
  float* params = nullptr;   // kParams
  float* grads = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  uint16_t *w1 = nullptr, *w1t = nullptr, *w2 = nullptr, *w2t = nullptr, *w3 = nullptr, *w3t = nullptr;
  size_t cap = 0;            // padded rows of the scratch below
  uint16_t* x = nullptr;     // [cap][64]
  float *z1 = nullptr, *z2 = nullptr;
  uint16_t *h1 = nullptr, *h2 = nullptr;
  float* y = nullptr;        // [cap][48]
  uint16_t *dy = nullptr, *dz2 = nullptr, *dz1 = nullptr;
  float* target = nullptr;   // [cap][45]
  double* parts = nullptr;
  double* h_parts = nullptr;
  hipStream_t stream = nullptr;
  int step = 0;
  // Input range of the trained network (device [2][kIn]: lower bounds, then upper bounds);
  // cameras whose features leave it get G = I. Null: the network is applied everywhere.
  float* range = nullptr;
};

namespace fg {
// Ensure scratch for n rows (padded to 64); allocates only when growing.
fg_status mlp_reserve(fg_mlp* m, size_t n);
// y[n][48] = network(x[n][64]) on the given stream; x and y are device buffers of >= npad rows.
void mlp_forward(fg_mlp* m, hipStream_t stream, size_t npad, const uint16_t* x, float* y);
inline size_t pad64(size_t n) { return (n + 63) / 64 * 64; }
// Selected covariance of the dense information h (device, n x n column-major, lower triangle
// significant; overwritten by its Cholesky factor) at the given coordinates, k x k into the
// host array out, on stream s. Status codes of SelectedCovariance::compute.
fg_status dense_covariance(hipStream_t s, double* h, size_t n, const size_t* columns, size_t k, double tolerance,
                           double* out);
// fg_lm_optimize with a damping hint before each linearization (fg_lm.cpp).
void lm_optimize_hinted(const fg_problem* pb, const fg_optimize_options* options, const fg_lm_options* lm,
                        fg_on_accept_fn on_accept, void* user, fg_lm_outcome* out,
                        void (*hint)(void* ctx, double damping));
}  // namespace fg
