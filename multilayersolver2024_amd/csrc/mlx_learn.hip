// The per-element MLP of the learned warm start: model storage, inference on
// host-given features and training (forward, backward and Adam) on the
// device, every matrix product on the BF16 matrix cores
// (v_mfma_f32_16x16x32_bf16, fp32 accumulate).  The solver's fused
// inference (features from the substep state) is in mlx_learn_kernels.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "mlx_internal.h"
#include "mlx_mfma.h"

namespace mlx {

// Flat parameter vector (fp32 master copy, gradients and Adam moments share
// the layout): W1 64x32 | W2 64x64 | W3 16x64 | b1 64 | b2 64 | b3 16.
constexpr int kOW1 = 0, kOW2 = kOW1 + kMlpHidden * kMlpInPad,
              kOW3 = kOW2 + kMlpHidden * kMlpHidden, kOB1 = kOW3 + kMlpOutPad * kMlpHidden,
              kOB2 = kOB1 + kMlpHidden, kOB3 = kOB2 + kMlpHidden, kNParam = kOB3 + kMlpOutPad;
constexpr int kTrainWaves = 4, kTrainBlock = 64 * kTrainWaves;
constexpr int kGradChunk = 32;  // samples per K step of the weight-gradient products

}  // namespace mlx

using namespace mlx;

struct mlx_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  DevBuf d_master, d_m, d_v;          // fp32 kNParam each
  DevBuf d_w;                         // bf16: W1 | W2 | W3 | W2T | W3T
  DevBuf d_norm;                      // fp32: fmean 21 | fistd 21 | oscale 2 | oshift 2
  int64_t step = 0;                   // Adam steps taken
};

namespace mlx {

// bf16 copies laid out in d_w
constexpr int kBW1 = 0, kBW2 = kBW1 + kMlpHidden * kMlpInPad,
              kBW3 = kBW2 + kMlpHidden * kMlpHidden, kBW2T = kBW3 + kMlpOutPad * kMlpHidden,
              kBW3T = kBW2T + kMlpHidden * kMlpHidden, kBWTotal = kBW3T + kMlpHidden * kMlpInPad;

MlpDev mlp_dev(const mlx_mlp &M) {
  MlpDev N;
  const __bf16 *w = M.d_w.as<__bf16>();
  const float *p = M.d_master.as<float>(), *nrm = M.d_norm.as<float>();
  N.W1 = w + kBW1;
  N.W2 = w + kBW2;
  N.W3 = w + kBW3;
  N.b1 = p + kOB1;
  N.b2 = p + kOB2;
  N.b3 = p + kOB3;
  N.fmean = nrm;
  N.fistd = nrm + kMlpIn;
  N.oscale = nrm + 2 * kMlpIn;
  N.oshift = nrm + 2 * kMlpIn + kMlpOut;
  return N;
}

// bf16 weight copies (and the transposes the backward pass reads) from the
// fp32 master vector.
__global__ void k_mlp_cast(const float *P, __bf16 *w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kMlpHidden * kMlpInPad) w[kBW1 + i] = (__bf16)P[kOW1 + i];
  if (i < kMlpHidden * kMlpHidden) {
    w[kBW2 + i] = (__bf16)P[kOW2 + i];
    const int o = i / kMlpHidden, in = i % kMlpHidden;  // W2T[in][o] = W2[o][in]
    w[kBW2T + in * kMlpHidden + o] = (__bf16)P[kOW2 + i];
  }
  if (i < kMlpOutPad * kMlpHidden) {
    w[kBW3 + i] = (__bf16)P[kOW3 + i];
    const int o = i / kMlpHidden, in = i % kMlpHidden;  // W3T[in][o] (o padded to 32)
    w[kBW3T + in * kMlpInPad + o] = (__bf16)P[kOW3 + i];
  }
  if (i < kMlpHidden * (kMlpInPad - kMlpOutPad)) {
    const int in = i / (kMlpInPad - kMlpOutPad), o = kMlpOutPad + i % (kMlpInPad - kMlpOutPad);
    w[kBW3T + in * kMlpInPad + o] = (__bf16)0.0f;
  }
}

// One tile of 16 samples per wave: normalised features -> X (LDS, bf16).
__device__ __forceinline__ void load_features(const MlpDev &N, const float *X, const int32_t *idx,
                                              int64_t b0, int64_t nb, __bf16 *Xs,
                                              __bf16 *H0 /* optional global 16 x 32 */) {
  const int l = lane64();
  for (int q = l; q < 16 * kMlpInPad; q += 64) {
    const int r = q / kMlpInPad, k = q % kMlpInPad;
    float v = 0.0f;
    if (b0 + r < nb && k < kMlpIn) {
      const int64_t smp = idx ? idx[b0 + r] : b0 + r;
      v = (X[smp * kMlpIn + k] - N.fmean[k]) * N.fistd[k];
    }
    Xs[q] = (__bf16)v;
    if (H0) H0[q] = (__bf16)v;
  }
}

// Inference on host-given raw features X (n x 21): Y (n x 2).
__global__ __launch_bounds__(kTrainBlock) void k_mlp_infer_x(MlpDev N, const float *X, int64_t n,
                                                             float *Y) {
  __shared__ __align__(16) __bf16 sX[kTrainWaves][16 * kMlpInPad];
  __shared__ __align__(16) __bf16 sH[kTrainWaves][16 * kMlpHidden];
  const int w = threadIdx.x >> 6, l = lane64();
  MlpFrags F;
  mlp_load_frags(N, F);
  const int64_t tiles = (n + 15) / 16;
  const int64_t per = (int64_t)gridDim.x * kTrainWaves;
  const int64_t rounds = (tiles + per - 1) / per;
  for (int64_t r = 0; r < rounds; ++r) {
    const int64_t t = r * per + (int64_t)blockIdx.x * kTrainWaves + w;
    load_features(N, X, nullptr, 16 * t, n, sX[w], nullptr);
    __syncthreads();
    const f32x4 y = mlp_tile(F, sX[w], sH[w], [] { __syncthreads(); });
    const int c = l & 15;
    if (c < kMlpOut)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t row = 16 * t + 4 * (l >> 4) + j;
        if (row < n) Y[row * kMlpOut + c] = y[j] * N.oscale[c] + N.oshift[c];
      }
    __syncthreads();
  }
}

// Training forward over one minibatch (samples idx[0 .. nb)): the network
// with its activations kept (Z1, Z2 fp32; H0, H1, H2 bf16; sample-major),
// the weighted squared error and its gradient dY (nb x 16, fp32, columns
// >= 2 zero, already divided by the weight total).  lossPart: per block.
__global__ __launch_bounds__(kTrainBlock) void k_mlp_fwd_train(
    MlpDev N, const float *X, const float *T, const float *Wt, const int32_t *idx, int64_t nb,
    double wsum, float *Z1, float *Z2, __bf16 *H0, __bf16 *H1, __bf16 *H2, float *dY,
    double *lossPart) {
  __shared__ __align__(16) __bf16 sX[kTrainWaves][16 * kMlpInPad];
  __shared__ __align__(16) __bf16 sH[kTrainWaves][16 * kMlpHidden];
  __shared__ double sLoss[kTrainBlock];
  const int w = threadIdx.x >> 6, l = lane64();
  MlpFrags F;
  mlp_load_frags(N, F);
  double loss = 0;
  const int64_t tiles = (nb + 15) / 16;
  const int64_t per = (int64_t)gridDim.x * kTrainWaves;
  const int64_t rounds = (tiles + per - 1) / per;
  for (int64_t r = 0; r < rounds; ++r) {
    const int64_t t = r * per + (int64_t)blockIdx.x * kTrainWaves + w;
    const bool live = t < tiles;
    const int64_t b0 = 16 * t;
    load_features(N, X, idx, b0, nb, sX[w], live ? H0 + b0 * kMlpInPad : nullptr);
    __syncthreads();
    const f32x4 y = mlp_tile(F, sX[w], sH[w], [] { __syncthreads(); },
                             live ? Z1 + b0 * kMlpHidden : nullptr,
                             live ? Z2 + b0 * kMlpHidden : nullptr,
                             live ? H1 + b0 * kMlpHidden : nullptr,
                             live ? H2 + b0 * kMlpHidden : nullptr);
    const int c = l & 15;
    if (live)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t row = b0 + 4 * (l >> 4) + j;
        float g = 0.0f;
        if (row < nb && c < kMlpOut) {
          const int64_t smp = idx[row];
          const float wt = Wt ? Wt[smp] : 1.0f;
          const float e = y[j] * N.oscale[c] + N.oshift[c] - T[smp * kMlpOut + c];
          loss += (double)wt * e * e;
          // d loss / d raw output, loss = sum_s w_s |y_s - t_s|^2 / wsum
          g = (float)(2.0 * wt * e * N.oscale[c] / wsum);
        }
        dY[row * kMlpOutPad + c] = g;
      }
    __syncthreads();
  }
  sLoss[threadIdx.x] = loss;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int i = 0; i < kTrainBlock; ++i) s += sLoss[i];
    lossPart[blockIdx.x] = s / wsum;
  }
}

// Backward through the layers: dH2 = dY W3, dZ2 = dH2 * silu'(Z2),
// dH1 = dZ2 W2, dZ1 = dH1 * silu'(Z1) (dZ1, dZ2 bf16 sample-major).
__global__ __launch_bounds__(kTrainBlock) void k_mlp_bwd_data(const __bf16 *w, const float *dY,
                                                              const float *Z1, const float *Z2,
                                                              int64_t nb, __bf16 *dZ1,
                                                              __bf16 *dZ2) {
  __shared__ __align__(16) __bf16 sA[kTrainWaves][16 * kMlpHidden];
  const int wv = threadIdx.x >> 6, l = lane64(), c = l & 15;
  const __bf16 *W2T = w + kBW2T, *W3T = w + kBW3T;
  const int64_t tiles = (nb + 15) / 16;
  const int64_t per = (int64_t)gridDim.x * kTrainWaves;
  const int64_t rounds = (tiles + per - 1) / per;
  for (int64_t r = 0; r < rounds; ++r) {
    const int64_t t = r * per + (int64_t)blockIdx.x * kTrainWaves + wv;
    const bool live = t < tiles;
    const int64_t b0 = 16 * t;
    // dY tile (16 x 32, columns >= 16 zero) as bf16 rows
    for (int q = l; q < 16 * kMlpInPad; q += 64) {
      const int rr = q / kMlpInPad, k = q % kMlpInPad;
      sA[wv][q] = (__bf16)(live && k < kMlpOutPad ? dY[(b0 + rr) * kMlpOutPad + k] : 0.0f);
    }
    __syncthreads();
    const bf16x8 ya = frag_rows(sA[wv], kMlpInPad, 0, 0);
    f32x4 g[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
      g[nt] = mfma_16x16x32(ya, frag_rows(W3T, kMlpInPad, 16 * nt, 0), f32x4{0.f, 0.f, 0.f, 0.f});
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 d;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int rr = 4 * (l >> 4) + j, col = 16 * nt + c;
        d[j] = live ? g[nt][j] * silu_grad(Z2[(b0 + rr) * kMlpHidden + col]) : 0.0f;
      }
      store_d_bf16(sA[wv], kMlpHidden, 16 * nt, d);
      if (live) store_d_bf16(dZ2 + b0 * kMlpHidden, kMlpHidden, 16 * nt, d);
    }
    __syncthreads();
    const bf16x8 a0 = frag_rows(sA[wv], kMlpHidden, 0, 0), a1 = frag_rows(sA[wv], kMlpHidden, 0, 32);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      g[nt] = mfma_16x16x32(a0, frag_rows(W2T, kMlpHidden, 16 * nt, 0), f32x4{0.f, 0.f, 0.f, 0.f});
      g[nt] = mfma_16x16x32(a1, frag_rows(W2T, kMlpHidden, 16 * nt, 32), g[nt]);
    }
    if (live)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        f32x4 d;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int rr = 4 * (l >> 4) + j, col = 16 * nt + c;
          d[j] = g[nt][j] * silu_grad(Z1[(b0 + rr) * kMlpHidden + col]);
        }
        store_d_bf16(dZ1 + b0 * kMlpHidden, kMlpHidden, 16 * nt, d);
      }
    __syncthreads();
  }
}

// Weight gradients: dW1 = dZ1^T H0, dW2 = dZ2^T H1, dW3 = dY^T H2 and the
// bias gradients (column sums), each block over its share of the samples
// in K steps of 32 staged (transposed) in LDS; 28 output tiles dealt over
// the block's waves.  part: one kNParam vector per block.
__global__ __launch_bounds__(kTrainBlock) void k_mlp_wgrad(const float *dY, const __bf16 *dZ1,
                                                           const __bf16 *dZ2, const __bf16 *H0,
                                                           const __bf16 *H1, const __bf16 *H2,
                                                           int64_t nb, int64_t chunksPerBlock,
                                                           float *part) {
  __shared__ __align__(16) __bf16 tZ1[kMlpHidden * kGradChunk], tZ2[kMlpHidden * kGradChunk],
      tH1[kMlpHidden * kGradChunk], tH2[kMlpHidden * kGradChunk],
      tH0[kMlpInPad * kGradChunk], tY[kMlpOutPad * kGradChunk];
  const int wv = threadIdx.x >> 6, l = lane64();
  f32x4 acc[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bacc = 0.0f;  // thread i < 144: bias gradient entry i
  const int64_t c0 = (int64_t)blockIdx.x * chunksPerBlock;
  for (int64_t ch = c0; ch < c0 + chunksPerBlock; ++ch) {
    const int64_t s0 = ch * kGradChunk;
    if (s0 >= nb) break;  // uniform over the block
    // stage the chunk transposed: t[feature][sample]
    for (int q = threadIdx.x; q < kGradChunk * kMlpHidden; q += kTrainBlock) {
      const int smp = q / kMlpHidden, f = q % kMlpHidden;
      const bool ok = s0 + smp < nb;
      const int64_t g = (s0 + smp) * kMlpHidden + f;
      tZ1[f * kGradChunk + smp] = ok ? dZ1[g] : (__bf16)0.0f;
      tZ2[f * kGradChunk + smp] = ok ? dZ2[g] : (__bf16)0.0f;
      tH1[f * kGradChunk + smp] = ok ? H1[g] : (__bf16)0.0f;
      tH2[f * kGradChunk + smp] = ok ? H2[g] : (__bf16)0.0f;
    }
    for (int q = threadIdx.x; q < kGradChunk * kMlpInPad; q += kTrainBlock) {
      const int smp = q / kMlpInPad, f = q % kMlpInPad;
      tH0[f * kGradChunk + smp] = s0 + smp < nb ? H0[(s0 + smp) * kMlpInPad + f] : (__bf16)0.0f;
    }
    for (int q = threadIdx.x; q < kGradChunk * kMlpOutPad; q += kTrainBlock) {
      const int smp = q / kMlpOutPad, f = q % kMlpOutPad;
      tY[f * kGradChunk + smp] =
          (__bf16)(s0 + smp < nb ? dY[(s0 + smp) * kMlpOutPad + f] : 0.0f);
    }
    __syncthreads();
    // 28 tiles: 0..15 dW2 (mt, nt), 16..23 dW1 (mt, nt < 2), 24..27 dW3 (nt)
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      const int tile = wv + kTrainWaves * q;
      if (tile < 16) {
        const int mt = tile / 4, nt = tile % 4;
        acc[q] = mfma_16x16x32(frag_rows(tZ2, kGradChunk, 16 * mt, 0),
                               frag_rows(tH1, kGradChunk, 16 * nt, 0), acc[q]);
      } else if (tile < 24) {
        const int mt = (tile - 16) / 2, nt = (tile - 16) % 2;
        acc[q] = mfma_16x16x32(frag_rows(tZ1, kGradChunk, 16 * mt, 0),
                               frag_rows(tH0, kGradChunk, 16 * nt, 0), acc[q]);
      } else {
        const int nt = tile - 24;
        acc[q] = mfma_16x16x32(frag_rows(tY, kGradChunk, 0, 0),
                               frag_rows(tH2, kGradChunk, 16 * nt, 0), acc[q]);
      }
    }
    // bias gradients: column sums of dZ1 | dZ2 | dY in sample order
    if (threadIdx.x < kMlpHidden) {
      for (int smp = 0; smp < kGradChunk; ++smp) bacc += (float)tZ1[threadIdx.x * kGradChunk + smp];
    } else if (threadIdx.x < 2 * kMlpHidden) {
      const int f = threadIdx.x - kMlpHidden;
      for (int smp = 0; smp < kGradChunk; ++smp) bacc += (float)tZ2[f * kGradChunk + smp];
    } else if (threadIdx.x < 2 * kMlpHidden + kMlpOutPad) {
      const int f = threadIdx.x - 2 * kMlpHidden;
      for (int smp = 0; smp < kGradChunk; ++smp)
        if (s0 + smp < nb) bacc += dY[(s0 + smp) * kMlpOutPad + f];
    }
    __syncthreads();
  }
  float *out = part + (size_t)blockIdx.x * kNParam;
  const int c = l & 15;
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const int tile = wv + kTrainWaves * q;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int rr = 4 * (l >> 4) + j;  // row within the 16 x 16 tile
      if (tile < 16) {
        const int mt = tile / 4, nt = tile % 4;
        out[kOW2 + (16 * mt + rr) * kMlpHidden + 16 * nt + c] = acc[q][j];
      } else if (tile < 24) {
        const int mt = (tile - 16) / 2, nt = (tile - 16) % 2;
        out[kOW1 + (16 * mt + rr) * kMlpInPad + 16 * nt + c] = acc[q][j];
      } else {
        const int nt = tile - 24;
        out[kOW3 + rr * kMlpHidden + 16 * nt + c] = acc[q][j];
      }
    }
  }
  if (threadIdx.x < 2 * kMlpHidden + kMlpOutPad) out[kOB1 + threadIdx.x] = bacc;
}

// Sum of the blocks' partial gradients in block order (deterministic).
__global__ void k_mlp_grad_sum(const float *part, int nblocks, float *grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNParam) return;
  float s = 0.0f;
  for (int b = 0; b < nblocks; ++b) s += part[(size_t)b * kNParam + i];
  grad[i] = s;
}

// Adam on the fp32 master vector (bias-corrected), then the bf16 copies.
// Padding entries (inputs >= 21, outputs >= 2) receive zero gradients and
// stay zero.
__global__ void k_mlp_adam(float *P, float *m, float *v, const float *grad, float lr, float b1,
                           float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNParam) return;
  const float g = grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * g, vi = b2 * v[i] + (1.0f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace mlx

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

namespace {

int mlp_cast(mlx_mlp &M) {
  const int n = kMlpHidden * kMlpHidden;
  hipLaunchKernelGGL(k_mlp_cast, dim3((n + 255) / 256), dim3(256), 0, M.stream,
                     M.d_master.as<float>(), M.d_w.as<__bf16>());
  MLX_CHECK_HIP(hipGetLastError());
  return MLX_OK;
}

// Training buffers for a minibatch of up to nb samples.
struct TrainBufs {
  DevBuf buf;
  float *Z1, *Z2, *dY, *grad, *part;
  __bf16 *H0, *H1, *H2, *dZ1, *dZ2;
  double *lossPart;
  int32_t *idx;
  int nblk, gblk;
  int64_t chunksPerBlock, cap;
};

int train_bufs(TrainBufs &B, int64_t nb) {
  const int64_t nbp = (nb + 31) / 32 * 32;
  B.cap = nbp;
  B.nblk = (int)std::min<int64_t>(256, std::max<int64_t>(1, (nbp / 16 + kTrainWaves - 1) / kTrainWaves));
  const int64_t chunks = nbp / kGradChunk;
  B.gblk = (int)std::min<int64_t>(128, std::max<int64_t>(1, chunks));
  B.chunksPerBlock = (chunks + B.gblk - 1) / B.gblk;
  const size_t f4 = sizeof(float), b2 = sizeof(__bf16);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t oZ1 = take(f4 * nbp * kMlpHidden), oZ2 = take(f4 * nbp * kMlpHidden),
               odY = take(f4 * nbp * kMlpOutPad), ogr = take(f4 * kNParam),
               opa = take(f4 * (size_t)B.gblk * kNParam), oH0 = take(b2 * nbp * kMlpInPad),
               oH1 = take(b2 * nbp * kMlpHidden), oH2 = take(b2 * nbp * kMlpHidden),
               odZ1 = take(b2 * nbp * kMlpHidden), odZ2 = take(b2 * nbp * kMlpHidden),
               olp = take(sizeof(double) * B.nblk), oid = take(sizeof(int32_t) * nbp);
  MLX_TRY(B.buf.reserve(off));
  char *p = B.buf.as<char>();
  B.Z1 = reinterpret_cast<float *>(p + oZ1);
  B.Z2 = reinterpret_cast<float *>(p + oZ2);
  B.dY = reinterpret_cast<float *>(p + odY);
  B.grad = reinterpret_cast<float *>(p + ogr);
  B.part = reinterpret_cast<float *>(p + opa);
  B.H0 = reinterpret_cast<__bf16 *>(p + oH0);
  B.H1 = reinterpret_cast<__bf16 *>(p + oH1);
  B.H2 = reinterpret_cast<__bf16 *>(p + oH2);
  B.dZ1 = reinterpret_cast<__bf16 *>(p + odZ1);
  B.dZ2 = reinterpret_cast<__bf16 *>(p + odZ2);
  B.lossPart = reinterpret_cast<double *>(p + olp);
  B.idx = reinterpret_cast<int32_t *>(p + oid);
  return MLX_OK;
}

// Forward, backward and the summed gradient of one minibatch (idx already
// on the device).  Returns the weighted mean squared error.
int minibatch_grad(mlx_mlp &M, TrainBufs &B, const float *X, const float *T, const float *W,
                   int64_t nb, double wsum, double *loss) {
  const MlpDev N = mlp_dev(M);
  hipStream_t s = M.stream;
  hipLaunchKernelGGL(k_mlp_fwd_train, dim3(B.nblk), dim3(kTrainBlock), 0, s, N, X, T, W, B.idx,
                     nb, wsum, B.Z1, B.Z2, B.H0, B.H1, B.H2, B.dY, B.lossPart);
  hipLaunchKernelGGL(k_mlp_bwd_data, dim3(B.nblk), dim3(kTrainBlock), 0, s, M.d_w.as<__bf16>(),
                     B.dY, B.Z1, B.Z2, nb, B.dZ1, B.dZ2);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(B.gblk), dim3(kTrainBlock), 0, s, B.dY, B.dZ1, B.dZ2, B.H0,
                     B.H1, B.H2, nb, B.chunksPerBlock, B.part);
  hipLaunchKernelGGL(k_mlp_grad_sum, dim3((kNParam + 255) / 256), dim3(256), 0, s, B.part, B.gblk,
                     B.grad);
  MLX_CHECK_HIP(hipGetLastError());
  if (loss) {
    std::vector<double> lp(B.nblk);
    MLX_CHECK_HIP(hipMemcpyAsync(lp.data(), B.lossPart, sizeof(double) * B.nblk,
                                 hipMemcpyDeviceToHost, s));
    MLX_CHECK_HIP(hipStreamSynchronize(s));
    double t = 0;
    for (double x : lp) t += x;
    *loss = t;
  }
  return MLX_OK;
}

// Device copies of a dataset (features, targets, weights).
struct DataBufs {
  DevBuf buf;
  float *X = nullptr, *T = nullptr, *W = nullptr;
};

int upload_data(mlx_mlp &M, DataBufs &D, const float *X, const float *T, const float *W,
                int64_t n) {
  const size_t sx = sizeof(float) * n * kMlpIn, st = sizeof(float) * n * kMlpOut,
               sw = W ? sizeof(float) * n : 0;
  MLX_TRY(D.buf.reserve(sx + st + sw + 1024));
  char *p = D.buf.as<char>();
  D.X = reinterpret_cast<float *>(p);
  D.T = reinterpret_cast<float *>(p + (sx + 255) / 256 * 256);
  D.W = W ? reinterpret_cast<float *>(p + (sx + 255) / 256 * 256 + (st + 255) / 256 * 256) : nullptr;
  MLX_CHECK_HIP(hipMemcpyAsync(D.X, X, sx, hipMemcpyHostToDevice, M.stream));
  if (T) MLX_CHECK_HIP(hipMemcpyAsync(D.T, T, st, hipMemcpyHostToDevice, M.stream));
  if (W) MLX_CHECK_HIP(hipMemcpyAsync(D.W, W, sw, hipMemcpyHostToDevice, M.stream));
  return MLX_OK;
}

bool finite_all(const float *a, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(a[i])) return false;
  return true;
}

}  // namespace

namespace mlx {
MlpDev mlp_device_view(const mlx_mlp *M) { return mlp_dev(*M); }
int mlp_device_of(const mlx_mlp *M) { return M->device; }
}  // namespace mlx

extern "C" {

int mlx_mlp_create(mlx_mlp **out) {
  if (!out) {
    set_error("mlx_mlp_create: null output");
    return MLX_ERR_ARG;
  }
  int dev = 0;
  MLX_CHECK_HIP(hipGetDevice(&dev));
  mlx_mlp *M = new mlx_mlp();
  M->device = dev;
  auto fail = [&](int r) {
    mlx_mlp_destroy(M);
    return r;
  };
  if (hipStreamCreate(&M->stream) != hipSuccess) {
    set_error("mlx_mlp_create: hipStreamCreate failed");
    delete M;
    return MLX_ERR_HIP;
  }
  int r;
  if ((r = M->d_master.reserve(sizeof(float) * kNParam)) != MLX_OK ||
      (r = M->d_m.reserve(sizeof(float) * kNParam)) != MLX_OK ||
      (r = M->d_v.reserve(sizeof(float) * kNParam)) != MLX_OK ||
      (r = M->d_w.reserve(sizeof(__bf16) * kBWTotal)) != MLX_OK ||
      (r = M->d_norm.reserve(sizeof(float) * (2 * kMlpIn + 2 * kMlpOut))) != MLX_OK)
    return fail(r);
  if (hipMemsetAsync(M->d_master.ptr, 0, sizeof(float) * kNParam, M->stream) != hipSuccess ||
      hipMemsetAsync(M->d_m.ptr, 0, sizeof(float) * kNParam, M->stream) != hipSuccess ||
      hipMemsetAsync(M->d_v.ptr, 0, sizeof(float) * kNParam, M->stream) != hipSuccess) {
    set_error("mlx_mlp_create: memset failed");
    return fail(MLX_ERR_HIP);
  }
  std::vector<float> nrm(2 * kMlpIn + 2 * kMlpOut, 0.0f);
  for (int k = 0; k < kMlpIn; ++k) nrm[kMlpIn + k] = 1.0f;
  for (int k = 0; k < kMlpOut; ++k) nrm[2 * kMlpIn + k] = 1.0f;
  if (hipMemcpy(M->d_norm.ptr, nrm.data(), sizeof(float) * nrm.size(), hipMemcpyHostToDevice) !=
          hipSuccess ||
      mlp_cast(*M) != MLX_OK || hipStreamSynchronize(M->stream) != hipSuccess) {
    set_error("mlx_mlp_create: upload failed");
    return fail(MLX_ERR_HIP);
  }
  *out = M;
  return MLX_OK;
}

int mlx_mlp_destroy(mlx_mlp *M) {
  if (!M) return MLX_OK;
  if (M->stream) (void)hipStreamSynchronize(M->stream);
  M->d_master.release();
  M->d_m.release();
  M->d_v.release();
  M->d_w.release();
  M->d_norm.release();
  if (M->stream) (void)hipStreamDestroy(M->stream);
  delete M;
  return MLX_OK;
}

int mlx_mlp_set(mlx_mlp *M, const float *W1, const float *b1, const float *W2, const float *b2,
                const float *W3, const float *b3, const float *featMean,
                const float *featInvStd, const float *outScale, const float *outShift) {
  if (!M || !W1 || !b1 || !W2 || !b2 || !W3 || !b3 || !featMean || !featInvStd || !outScale ||
      !outShift) {
    set_error("mlx_mlp_set: null argument");
    return MLX_ERR_ARG;
  }
  std::vector<float> P(kNParam, 0.0f);
  for (int o = 0; o < kMlpHidden; ++o)
    for (int i = 0; i < kMlpIn; ++i) P[kOW1 + o * kMlpInPad + i] = W1[o * kMlpIn + i];
  for (int i = 0; i < kMlpHidden * kMlpHidden; ++i) P[kOW2 + i] = W2[i];
  for (int o = 0; o < kMlpOut; ++o)
    for (int i = 0; i < kMlpHidden; ++i) P[kOW3 + o * kMlpHidden + i] = W3[o * kMlpHidden + i];
  for (int o = 0; o < kMlpHidden; ++o) P[kOB1 + o] = b1[o];
  for (int o = 0; o < kMlpHidden; ++o) P[kOB2 + o] = b2[o];
  for (int o = 0; o < kMlpOut; ++o) P[kOB3 + o] = b3[o];
  std::vector<float> nrm(2 * kMlpIn + 2 * kMlpOut);
  for (int k = 0; k < kMlpIn; ++k) {
    nrm[k] = featMean[k];
    nrm[kMlpIn + k] = featInvStd[k];
  }
  for (int k = 0; k < kMlpOut; ++k) {
    nrm[2 * kMlpIn + k] = outScale[k];
    nrm[2 * kMlpIn + kMlpOut + k] = outShift[k];
  }
  if (!finite_all(P.data(), P.size()) || !finite_all(nrm.data(), nrm.size())) {
    set_error("mlx_mlp_set: non-finite parameter");
    return MLX_ERR_ARG;
  }
  MLX_CHECK_HIP(hipMemcpyAsync(M->d_master.ptr, P.data(), sizeof(float) * kNParam,
                               hipMemcpyHostToDevice, M->stream));
  MLX_CHECK_HIP(hipMemcpyAsync(M->d_norm.ptr, nrm.data(), sizeof(float) * nrm.size(),
                               hipMemcpyHostToDevice, M->stream));
  MLX_CHECK_HIP(hipMemsetAsync(M->d_m.ptr, 0, sizeof(float) * kNParam, M->stream));
  MLX_CHECK_HIP(hipMemsetAsync(M->d_v.ptr, 0, sizeof(float) * kNParam, M->stream));
  M->step = 0;
  MLX_TRY(mlp_cast(*M));
  MLX_CHECK_HIP(hipStreamSynchronize(M->stream));
  return MLX_OK;
}

int mlx_mlp_get(mlx_mlp *M, float *W1, float *b1, float *W2, float *b2, float *W3, float *b3) {
  if (!M) {
    set_error("mlx_mlp_get: null model");
    return MLX_ERR_ARG;
  }
  std::vector<float> P(kNParam);
  MLX_CHECK_HIP(hipMemcpyAsync(P.data(), M->d_master.ptr, sizeof(float) * kNParam,
                               hipMemcpyDeviceToHost, M->stream));
  MLX_CHECK_HIP(hipStreamSynchronize(M->stream));
  if (W1)
    for (int o = 0; o < kMlpHidden; ++o)
      for (int i = 0; i < kMlpIn; ++i) W1[o * kMlpIn + i] = P[kOW1 + o * kMlpInPad + i];
  if (W2) std::memcpy(W2, P.data() + kOW2, sizeof(float) * kMlpHidden * kMlpHidden);
  if (W3)
    for (int o = 0; o < kMlpOut; ++o)
      for (int i = 0; i < kMlpHidden; ++i) W3[o * kMlpHidden + i] = P[kOW3 + o * kMlpHidden + i];
  if (b1) std::memcpy(b1, P.data() + kOB1, sizeof(float) * kMlpHidden);
  if (b2) std::memcpy(b2, P.data() + kOB2, sizeof(float) * kMlpHidden);
  if (b3) std::memcpy(b3, P.data() + kOB3, sizeof(float) * kMlpOut);
  return MLX_OK;
}

int mlx_mlp_infer(mlx_mlp *M, const float *X, int64_t n, float *Y) {
  if (!M || !X || !Y || n <= 0) {
    set_error("mlx_mlp_infer: bad arguments");
    return MLX_ERR_ARG;
  }
  DataBufs D;
  MLX_TRY(upload_data(*M, D, X, nullptr, nullptr, n));
  DevBuf yb;
  MLX_TRY(yb.reserve(sizeof(float) * n * kMlpOut));
  const int64_t tiles = (n + 15) / 16;
  const int nblk = (int)std::min<int64_t>(1024, (tiles + kTrainWaves - 1) / kTrainWaves);
  hipLaunchKernelGGL(k_mlp_infer_x, dim3(nblk), dim3(kTrainBlock), 0, M->stream, mlp_dev(*M),
                     D.X, n, yb.as<float>());
  MLX_CHECK_HIP(hipGetLastError());
  MLX_CHECK_HIP(hipMemcpyAsync(Y, yb.ptr, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToHost,
                               M->stream));
  MLX_CHECK_HIP(hipStreamSynchronize(M->stream));
  return MLX_OK;
}

int mlx_mlp_grad(mlx_mlp *M, const float *X, const float *T, const float *W, int64_t n,
                 float *grad, double *loss) {
  if (!M || !X || !T || !grad || n <= 0 || n > (1 << 24)) {
    set_error("mlx_mlp_grad: bad arguments");
    return MLX_ERR_ARG;
  }
  DataBufs D;
  MLX_TRY(upload_data(*M, D, X, T, W, n));
  TrainBufs B;
  MLX_TRY(train_bufs(B, n));
  std::vector<int32_t> idx(n);
  double wsum = 0;
  for (int64_t i = 0; i < n; ++i) {
    idx[i] = (int32_t)i;
    wsum += W ? W[i] : 1.0;
  }
  if (!(wsum > 0)) {
    set_error("mlx_mlp_grad: weights sum to zero");
    return MLX_ERR_ARG;
  }
  MLX_CHECK_HIP(hipMemcpyAsync(B.idx, idx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice,
                               M->stream));
  MLX_TRY(minibatch_grad(*M, B, D.X, D.T, D.W, n, wsum, loss));
  std::vector<float> g(kNParam);
  MLX_CHECK_HIP(hipMemcpyAsync(g.data(), B.grad, sizeof(float) * kNParam, hipMemcpyDeviceToHost,
                               M->stream));
  MLX_CHECK_HIP(hipStreamSynchronize(M->stream));
  // in the layout of mlx_mlp_set: dW1 64x21 | db1 | dW2 | db2 | dW3 2x64 | db3
  float *o = grad;
  for (int r = 0; r < kMlpHidden; ++r)
    for (int i = 0; i < kMlpIn; ++i) *o++ = g[kOW1 + r * kMlpInPad + i];
  for (int r = 0; r < kMlpHidden; ++r) *o++ = g[kOB1 + r];
  for (int i = 0; i < kMlpHidden * kMlpHidden; ++i) *o++ = g[kOW2 + i];
  for (int r = 0; r < kMlpHidden; ++r) *o++ = g[kOB2 + r];
  for (int r = 0; r < kMlpOut; ++r)
    for (int i = 0; i < kMlpHidden; ++i) *o++ = g[kOW3 + r * kMlpHidden + i];
  for (int r = 0; r < kMlpOut; ++r) *o++ = g[kOB3 + r];
  return MLX_OK;
}

int mlx_mlp_train(mlx_mlp *M, const float *X, const float *T, const float *W, int64_t n,
                  int32_t epochs, int32_t batch, float lr, uint64_t seed, float *epochLoss) {
  if (!M || !X || !T || n <= 0 || n > INT32_MAX || epochs < 0 || batch < 16 || batch > (1 << 20) ||
      !(lr > 0)) {
    set_error("mlx_mlp_train: bad arguments");
    return MLX_ERR_ARG;
  }
  DataBufs D;
  MLX_TRY(upload_data(*M, D, X, T, W, n));
  TrainBufs B;
  MLX_TRY(train_bufs(B, std::min<int64_t>(batch, n)));
  std::mt19937_64 rng(seed);
  std::vector<int32_t> perm(n);
  DevBuf dperm;
  MLX_TRY(dperm.reserve(sizeof(int32_t) * n));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    std::shuffle(perm.begin(), perm.end(), rng);
    MLX_CHECK_HIP(hipMemcpyAsync(dperm.ptr, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice,
                                 M->stream));
    double epLoss = 0, epW = 0;
    for (int64_t b0 = 0; b0 < n; b0 += batch) {
      const int64_t nb = std::min<int64_t>(batch, n - b0);
      double wsum = 0;
      for (int64_t i = b0; i < b0 + nb; ++i) wsum += W ? W[perm[i]] : 1.0;
      if (!(wsum > 0)) continue;
      B.idx = dperm.as<int32_t>() + b0;
      double loss = 0;
      const bool last = b0 + batch >= n;
      MLX_TRY(minibatch_grad(*M, B, D.X, D.T, D.W, nb, wsum, last || epochLoss ? &loss : nullptr));
      epLoss += loss * wsum;
      epW += wsum;
      M->step += 1;
      const float c1 = 1.0f - std::pow(b1, (float)M->step), c2 = 1.0f - std::pow(b2, (float)M->step);
      hipLaunchKernelGGL(k_mlp_adam, dim3((kNParam + 255) / 256), dim3(256), 0, M->stream,
                         M->d_master.as<float>(), M->d_m.as<float>(), M->d_v.as<float>(), B.grad,
                         lr, b1, b2, eps, c1, c2);
      MLX_TRY(mlp_cast(*M));
    }
    if (epochLoss) epochLoss[ep] = (float)(epW > 0 ? epLoss / epW : 0.0);
  }
  MLX_CHECK_HIP(hipStreamSynchronize(M->stream));
  return MLX_OK;
}

}  // extern "C"
