// BF16 matrix-core building blocks of the learned warm start (gfx950) and
// the per-element MLP they serve.  Device functions only (no kernels), so
// both translation units that run the network -- the solver's fused
// inference and the training kernels -- include it.
//
// v_mfma_f32_16x16x32_bf16: D(16x16, fp32) += A(16x32, bf16) B(32x16, bf16)
// across one 64-lane wavefront.  Lane l holds
//   A[l % 16][8 (l / 16) + i], i = 0..7          (8 bf16)
//   B[8 (l / 16) + i][l % 16], i = 0..7          (8 bf16)
//   D[4 (l / 16) + j][l % 16], j = 0..3          (4 fp32)
// (checked on the device by an in-bounds probe of the builtin).  Both
// operand fragments are therefore 8 consecutive values of one row of a
// row-major matrix: A as stored, B from its transpose (n-major).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mlx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Network shape: 21 features (padded to one K step of 32), two hidden
// layers of 64, 2 outputs (padded to one 16-column tile).
constexpr int kMlpIn = 21, kMlpInPad = 32, kMlpHidden = 64, kMlpOut = 2, kMlpOutPad = 16;

__device__ __forceinline__ int lane64() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ f32x4 mfma_16x16x32(const bf16x8 &a, const bf16x8 &b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Operand fragment: rows r0 .. r0 + 15 of a row-major bf16 matrix with row
// stride ld, columns k0 + 8 (l / 16) .. + 7 (ld and k0 multiples of 8).
__device__ __forceinline__ bf16x8 frag_rows(const __bf16 *M, int ld, int r0, int k0) {
  const int l = lane64();
  return *reinterpret_cast<const bf16x8 *>(M + (size_t)(r0 + (l & 15)) * ld + k0 + 8 * (l >> 4));
}

__device__ __forceinline__ float sigmoidf(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float siluf(float z) { return z * sigmoidf(z); }
// d silu / dz = s (1 + z (1 - s)), s = sigmoid(z)
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoidf(z);
  return s * (1.0f + z * (1.0f - s));
}

// Device view of a network.  W1 64 x 32, W2 64 x 64, W3 16 x 64: row-major
// (out x in, zero padding), bf16.  Biases, feature normalisation (x - mean)
// * istd, and the output map y = raw * oscale + oshift are fp32.
struct MlpDev {
  const __bf16 *W1, *W2, *W3;
  const float *b1, *b2, *b3;
  const float *fmean, *fistd;  // kMlpIn each
  const float *oscale, *oshift;  // kMlpOut each
};

// Weight fragments of one wave, held in registers across the tiles it runs.
struct MlpFrags {
  bf16x8 w1[4];     // n-tile 0..3 of layer 1 (K = 32)
  bf16x8 w2[4][2];  // n-tile x k-step of layer 2
  bf16x8 w3[2];     // k-step of layer 3 (one n-tile)
  float b1[4], b2[4], b3;
};

__device__ __forceinline__ void mlp_load_frags(const MlpDev &N, MlpFrags &F) {
  const int c = lane64() & 15;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    F.w1[nt] = frag_rows(N.W1, kMlpInPad, 16 * nt, 0);
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) F.w2[nt][kt] = frag_rows(N.W2, kMlpHidden, 16 * nt, 32 * kt);
    F.b1[nt] = N.b1[16 * nt + c];
    F.b2[nt] = N.b2[16 * nt + c];
  }
#pragma unroll
  for (int kt = 0; kt < 2; ++kt) F.w3[kt] = frag_rows(N.W3, kMlpHidden, 0, 32 * kt);
  F.b3 = N.b3[c];
}

// D fragment (16 x 16, this lane's 4 rows of one column) to rows
// r0 + 4 (l / 16) + j, column c0 + l % 16 of a row-major bf16 tile.
__device__ __forceinline__ void store_d_bf16(__bf16 *M, int ld, int c0, const f32x4 &d) {
  const int l = lane64();
#pragma unroll
  for (int j = 0; j < 4; ++j) M[(size_t)(4 * (l >> 4) + j) * ld + c0 + (l & 15)] = (__bf16)d[j];
}

// The network on one 16-row tile: X (16 x 32 bf16, normalised features,
// row-major in LDS), H (16 x 64 bf16 LDS scratch).  Returns this lane's
// D fragment of the (unmapped) output layer: rows 4 (l / 16) + j, output
// column l % 16 (columns >= kMlpOut are padding).  sync() separates the
// LDS writes of one layer from the next layer's fragment reads (the whole
// workgroup runs the same tiles).  Z1/Z2 (optional, 16 x 64 fp32 each, row
// stride 64) receive the pre-activations for the backward pass.
template <typename Sync>
__device__ __forceinline__ f32x4 mlp_tile(const MlpFrags &F, const __bf16 *X, __bf16 *H,
                                          Sync &&sync, float *Z1 = nullptr, float *Z2 = nullptr,
                                          __bf16 *H1out = nullptr, __bf16 *H2out = nullptr) {
  const int l = lane64(), c = l & 15;
  const bf16x8 xa = frag_rows(X, kMlpInPad, 0, 0);
  f32x4 z[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    z[nt] = mfma_16x16x32(xa, F.w1[nt], f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int j = 0; j < 4; ++j) z[nt][j] += F.b1[nt];
  }
  sync();  // previous users of H are done
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    f32x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (Z1) Z1[(4 * (l >> 4) + j) * kMlpHidden + 16 * nt + c] = z[nt][j];
      a[j] = siluf(z[nt][j]);
    }
    store_d_bf16(H, kMlpHidden, 16 * nt, a);
    if (H1out) store_d_bf16(H1out, kMlpHidden, 16 * nt, a);
  }
  sync();
  const bf16x8 h0 = frag_rows(H, kMlpHidden, 0, 0), h1 = frag_rows(H, kMlpHidden, 0, 32);
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    z[nt] = mfma_16x16x32(h0, F.w2[nt][0], f32x4{0.f, 0.f, 0.f, 0.f});
    z[nt] = mfma_16x16x32(h1, F.w2[nt][1], z[nt]);
#pragma unroll
    for (int j = 0; j < 4; ++j) z[nt][j] += F.b2[nt];
  }
  sync();
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    f32x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (Z2) Z2[(4 * (l >> 4) + j) * kMlpHidden + 16 * nt + c] = z[nt][j];
      a[j] = siluf(z[nt][j]);
    }
    store_d_bf16(H, kMlpHidden, 16 * nt, a);
    if (H2out) store_d_bf16(H2out, kMlpHidden, 16 * nt, a);
  }
  sync();
  const bf16x8 g0 = frag_rows(H, kMlpHidden, 0, 0), g1 = frag_rows(H, kMlpHidden, 0, 32);
  f32x4 y = mfma_16x16x32(g0, F.w3[0], f32x4{0.f, 0.f, 0.f, 0.f});
  y = mfma_16x16x32(g1, F.w3[1], y);
#pragma unroll
  for (int j = 0; j < 4; ++j) y[j] += F.b3;
  return y;
}

}  // namespace mlx
