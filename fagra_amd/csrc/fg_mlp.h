// fg_mlp.h -- the per-camera preconditioner network (47 -> 128 -> 128 -> 45, GELU), BF16
// operands on the gfx950 matrix cores with FP32 accumulation.
//
// The network maps features of a camera's Gram block to a lower-triangular correction G =
// I + s tril(y) of the block-Jacobi inverse: M = L^-T G G^T L^-1 with L L^T = B + lambda I.
// M is symmetric positive definite whenever G is nonsingular. Training fits G G^T to
// (L^-1 S_cc L^-T)^-1, S_cc the camera's diagonal block of the reduced (Schur) system, by
// minimizing || G^T A G - I ||_F^2 with A = L^-1 S_cc L^-T.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fg {

constexpr int kIn = 47, kInPad = 64, kHidden = 128, kOut = 45, kOutPad = 48;
constexpr float kOutScale = 0.125f;  // s: y -> G = I + s tril(y)

// Parameter block layout (f32 master copy, also the serialized form):
//   W1 [kHidden][kInPad], b1 [kHidden], W2 [kHidden][kHidden], b2 [kHidden],
//   W3 [kOutPad][kHidden], b3 [kOutPad]  (W stored as [out][in]).
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kHidden * kInPad;
constexpr int kW2 = kB1 + kHidden;
constexpr int kB2 = kW2 + kHidden * kHidden;
constexpr int kW3 = kB2 + kHidden;
constexpr int kB3 = kW3 + kOutPad * kHidden;
constexpr int kParams = kB3 + kOutPad;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// tanh-form GELU and its derivative
__device__ __forceinline__ float gelu(float x) {
  const float c = 0.7978845608028654f;
  const float t = tanhf(c * (x + 0.044715f * x * x * x));
  return 0.5f * x * (1.f + t);
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float c = 0.7978845608028654f;
  const float u = c * (x + 0.044715f * x * x * x);
  const float t = tanhf(u);
  return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * c * (1.f + 3.f * 0.044715f * x * x);
}

// One 16x16 output tile += A(16 x K) * B(K x 16) where A rows are contiguous in K (row-major,
// leading dimension lda) and B is given transposed (Bt[n][k], leading dimension ldb), both bf16.
// Operand layout of v_mfma_f32_16x16x16_bf16: lane l supplies A[l%16][4(l/16)+i] and
// B[4(l/16)+i][l%16]; the result lane l holds D[4(l/16)+i][l%16], i = 0..3.
__device__ __forceinline__ f32x4 mfma_tile(const uint16_t* a, int lda, const uint16_t* bt, int ldb, int kdim,
                                           f32x4 acc) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  for (int k0 = 0; k0 < kdim; k0 += 16) {
    const bf16x4 av = *(const bf16x4*)(a + (size_t)r * lda + k0 + 4 * q);
    const bf16x4 bv = *(const bf16x4*)(bt + (size_t)r * ldb + k0 + 4 * q);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
  return acc;
}

}  // namespace fg
