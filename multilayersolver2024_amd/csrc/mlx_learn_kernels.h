// Kernels of the learned warm start that run on the solver's state (3D):
// the fused per-element features + network inference, the strain-rate key
// correction of the coarse layers, and the warm start of the finest layer
// with its fallback to the plain (cold) start.  Included once, by
// mlx_solver.hip, after mlx_kernels.h.
//
// Features of element e at the start of a substep (21, in this order):
//   F (9, row-major) of the current positions, dF/dt = (F - F_prev) / h
//   (9) with F_prev from the previous substep's positions, the previous
//   substep's multipliers as log10(|lambda_e| / (vol h^2 E) + 1e-12), log10 E
//   and nu of the element (from its stored compliances).
// Outputs: kappa -- the warm start lambda_0 = kappa lambda_prev (the
// reference restarts lambda at zero every substep, xpbdLayer3D.m:666) -- and
// c, the log10 factor by which the element's strain-rate key is corrected
// before the coarse layers are built (the rate predicted for the coming
// substep instead of the last one).
#pragma once
#include "mlx_mfma.h"

namespace mlx {

constexpr int kLearnWaves = 4, kLearnBlock = 64 * kLearnWaves;

// Young's modulus and Poisson ratio of a slot from its material scalars.
// StVK-Voigt: alpha_s = 1/(2 mu), the normal block of K = inv(alpha) has
// diagonal lambda + 4 mu and off-diagonal lambda, i.e. with p = alpha(1,1),
// q = alpha(1,2): a - b = 1/(p - q), a + 2b = 1/(p + 2q).  NeoHookean:
// -1/mu, 1/lambda.
template <typename T>
__device__ __forceinline__ void slot_moduli(const SweepParams<T> &P, int64_t s, float &E,
                                            float &nu) {
  const int64_t ld = P.ld;
  const double m0 = (double)P.mat[s], m1 = (double)P.mat[ld + s], m2 = (double)P.mat[2 * ld + s];
  double mu, la;
  if (m0 < 0) {
    mu = -1.0 / m0;
    la = 1.0 / m1;
  } else {
    mu = 0.5 / m2;
    const double amb = 1.0 / (m0 - m1), ap2b = 1.0 / (m0 + 2.0 * m1);
    la = (ap2b - amb) / 3.0;
  }
  E = (float)(mu * (3.0 * la + 2.0 * mu) / (la + mu));
  nu = (float)(la / (2.0 * (la + mu)));
}

template <typename T>
__device__ __forceinline__ void slot_features(const SweepParams<T> &P, const vec4_t<T> *xprev,
                                              int64_t s, float ih, float ih2, float (&f)[kMlpIn]) {
  const int64_t ld = P.ld;
  T D[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) D[k] = P.dminv[k * ld + s];
  vec4_t<T> X[4], Q[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int v = P.vid[a * ld + s];
    X[a] = P.x[v];
    Q[a] = xprev[v];
  }
  T F[3][3], G[3][3];
  tet_F<T>(X, D, F);
  tet_F<T>(Q, D, G);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      f[3 * r + c] = (float)F[r][c];
      f[9 + 3 * r + c] = (float)((F[r][c] - G[r][c]) * (T)ih);
    }
  float E, nu;
  slot_moduli<T>(P, s, E, nu);
  double l2 = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const double li = (double)P.lam[i * ld + s];
    l2 += li * li;
  }
  const double scale = (double)P.vol[s] * (double)E / (double)ih2;  // vol h^2 E
  f[18] = (float)log10(sqrt(l2) / scale + 1e-12);
  f[19] = log10f(E);
  f[20] = nu;
}

// One wave per 64 slots: lane l forms the features of slot 64 g + l into
// LDS, then the wave runs the network on its four 16-row tiles and lane l
// maps its slot's outputs.  lam0 (6 per slot, slot-major like lambda):
// kappa times the slot's current (previous substep's) multipliers;
// keyScale (per slot, optional): 10^c, c clamped to [-3, 3] (1 for a
// non-finite c); feat (21 per slot, optional): the raw features.
template <typename T>
__global__ __launch_bounds__(kLearnBlock) void k_learn_infer(SweepParams<T> P, MlpDev N,
                                                             const vec4_t<T> *xprev, float ih,
                                                             float ih2, T *lam0, double *keyScale,
                                                             float kappaMax, float *feat) {
  __shared__ __align__(16) __bf16 sX[kLearnWaves][64 * kMlpInPad];
  __shared__ __align__(16) __bf16 sH[kLearnWaves][16 * kMlpHidden];
  __shared__ float sY[kLearnWaves][64 * kMlpOut];
  const int w = threadIdx.x >> 6, l = lane64();
  MlpFrags F;
  mlp_load_frags(N, F);
  const int64_t g = (int64_t)blockIdx.x * kLearnWaves + w;
  const int64_t s = 64 * g + l;
  const bool live = s < P.ld && P.vid[s] >= 0;
  float f[kMlpIn];
  if (live) {
    slot_features<T>(P, xprev, s, ih, ih2, f);
    if (feat)
#pragma unroll
      for (int k = 0; k < kMlpIn; ++k) feat[(size_t)s * kMlpIn + k] = f[k];
  } else {
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) f[k] = N.fmean[k];
  }
#pragma unroll
  for (int k = 0; k < kMlpInPad; ++k)
    sX[w][l * kMlpInPad + k] = (__bf16)(k < kMlpIn ? (f[k] - N.fmean[k]) * N.fistd[k] : 0.0f);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const f32x4 y = mlp_tile(F, sX[w] + 16 * q * kMlpInPad, sH[w], [] { __syncthreads(); });
    const int c = l & 15;
    if (c < kMlpOut)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        sY[w][(16 * q + 4 * (l >> 4) + j) * kMlpOut + c] = y[j] * N.oscale[c] + N.oshift[c];
  }
  __syncthreads();
  if (!live) return;
  const int64_t ld = P.ld;
  float kappa = sY[w][l * kMlpOut];
  kappa = kappa == kappa ? fminf(fmaxf(kappa, 0.0f), kappaMax) : 0.0f;
#pragma unroll
  for (int i = 0; i < 6; ++i) lam0[i * ld + s] = (T)kappa * P.lam[i * ld + s];
  if (keyScale) {
    const float c = sY[w][l * kMlpOut + 1];
    keyScale[s] = c == c ? exp10((double)fminf(fmaxf(c, -3.0f), 3.0f)) : 1.0;
  }
}

// Strain-rate keys (element order) times the factor of their slot.
__global__ void k_learn_keys(int64_t nT, const int32_t *slotOf, const double *keyScale,
                             double *keys) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= nT) return;
  keys[e] = keys[e] * keyScale[slotOf[e]];
}

// Residual of the XPBD equations the projection drives to zero, squared,
// of one elastic slot: per constraint the numerator of its update,
// C + alpha~ lambda + gamma grad C (x - x_old) (StVK-Voigt, the normal and
// shear rows as the projection forms them) or C + alpha~ lambda (the
// NeoHookean pair).  Unlike the reference's monitor measure
// max(|C| - |alpha~ lambda|, 0), it grows when lambda overshoots, so it can
// judge a warm start.
template <typename T>
__device__ __forceinline__ double xpbd_eq_residual_x(const TetIn<T> &in, const vec4_t<T> (&X)[4]) {
  T D[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) D[k] = in.D[k];
  T F[3][3];
  tet_F<T>(X, D, F);
  if (in.tad < T(0)) {
    T PH[3][3], n, J;
    nh_terms<T>(F, n, PH, J);
    const T iv = T(1) / in.vol;
    const double rD = (double)(n + iv * (-in.tad) * in.lam[0]);
    const double rH = (double)((J - in.tas) + iv * in.tao * in.lam[1]);
    return rD * rD + rH * rH;
  }
  vec4_t<T> Dx[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    Dx[a] = X[a];
    Dx[a].x = X[a].x - in.O[a].x;
    Dx[a].y = X[a].y - in.O[a].y;
    Dx[a].z = X[a].z - in.O[a].z;
  }
  T Fd[3][3];
  tet_F<T>(Dx, D, Fd);
  const T vol = in.vol, half = T(0.5);
  T Cg[3][3], H[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Cg[j][k] = F[0][j] * F[0][k] + F[1][j] * F[1][k] + F[2][j] * F[2][k];
      H[j][k] = F[0][j] * Fd[0][k] + F[1][j] * Fd[1][k] + F[2][j] * Fd[2][k];
    }
  double acc = 0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int k = (j + 1) % 3, l = (j + 2) % 3;
    const double r = (double)(vol * (half * (Cg[j][j] - T(1))) +
                              (in.tad * in.lam[j] + in.tao * in.lam[k] + in.tao * in.lam[l]) +
                              in.g * (vol * H[j][j]));
    acc += r * r;
  }
  const int sj[3] = {0, 0, 1}, sk[3] = {1, 2, 2};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int j = sj[q], k = sk[q];
    const double r = (double)(vol * (half * Cg[j][k]) + in.tas * in.lam[3 + q] +
                              in.g * (vol * (half * (H[j][k] + H[k][j]))));
    acc += r * r;
  }
  return acc;
}

template <typename T>
__device__ __forceinline__ double xpbd_eq_residual(const SweepParams<T> &P, int64_t s) {
  TetIn<T> in;
  tet_fetch<T, kMatMixed>(P, s, in);
  vec4_t<T> X[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) X[a] = P.x[in.v[a]];
  return xpbd_eq_residual_x<T>(in, X);
}

// The warm-start judge in two launches: block partials of the squared
// equation residuals of the elastic slots (P.flags, null = all) and of the
// contacts (r phi)^2 (their multipliers are zero when a layer starts), then
// their sum in block order into *out (sqrt taken).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_learn_res_part(SweepParams<T> P, int nE,
                                                           ContactArrays c, double resistance,
                                                           double *part) {
  __shared__ double sh[kBlock];
  const int t = threadIdx.x;
  double acc = 0;
  if ((int)blockIdx.x < nE) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + t;
    if (s < P.ld && P.vid[s] >= 0 && (!P.flags || P.flags[s])) acc = xpbd_eq_residual<T>(P, s);
  } else {
    const int i = ((int)blockIdx.x - nE) * kBlock + t;
    if (i < c.counts[0]) {
      const vec4_t<T> p = P.x[c.vert[i]];
      const double *n = c.normal + 3 * i, *pc = c.point + 3 * i;
      double phi = n[0] * ((double)p.x - pc[0]) + n[1] * ((double)p.y - pc[1]) +
                   n[2] * ((double)p.z - pc[2]);
      phi = phi < 0 ? phi : 0;
      acc = (resistance * phi) * (resistance * phi);
    }
  }
  double o[1];
  double v1[1] = {acc};
  block_sum<1>(v1, sh, o);
  if (t == 0) part[blockIdx.x] = o[0];
}

__global__ __launch_bounds__(kBlock) void k_learn_res_final(const double *part, int n,
                                                            double *out) {
  __shared__ double sh[kBlock];
  double a = 0;
  for (int q = threadIdx.x; q < n; q += kBlock) a += part[q];
  double o[1];
  double v1[1] = {a};
  block_sum<1>(v1, sh, o);
  if (threadIdx.x == 0) *out = sqrt(o[0]);
}

// The warm start of one fresh elastic slot with its vertices' current
// positions X: lambda takes lambda_0, dx the displacement of each vertex.
template <typename T>
__device__ __forceinline__ void learn_disp_slot(const SweepParams<T> &P, const int64_t s,
                                                const T *lam0, T *dx, const vec4_t<T> (&X)[4]) {
  const int64_t ld = P.ld;
  T D[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) D[k] = P.dminv[k * ld + s];
  T Fm[3][3];
  tet_F<T>(X, D, Fm);
  T G[4][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    G[1][c] = D[3 * c];
    G[2][c] = D[1 + 3 * c];
    G[3][c] = D[2 + 3 * c];
    G[0][c] = -(G[1][c] + G[2][c] + G[3][c]);
  }
  T dl[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    dl[i] = lam0[i * ld + s] - P.lam[i * ld + s];
    P.lam[i * ld + s] = lam0[i * ld + s];
  }
  const T vol = P.vol[s];
  if (P.mat[s] < T(0)) {
    // NeoHookean: dlambda_D grad ||F|| + dlambda_H grad det F
    T PH[3][3], n, J;
    nh_terms<T>(Fm, n, PH, J);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const T gD = (Fm[r][0] * G[a][0] + Fm[r][1] * G[a][1] + Fm[r][2] * G[a][2]) / n;
        const T gH = PH[r][0] * G[a][0] + PH[r][1] * G[a][1] + PH[r][2] * G[a][2];
        dx[(3 * a + r) * ld + s] = X[a].w * (gD * dl[0] + gH * dl[1]);
      }
    return;
  }
  // StVK-Voigt: sum_i dl_i grad_a C_i = vol F S g_a, S = sum_i dl_i P_i
  const T half = T(0.5);
  const T S[3][3] = {{dl[0], half * dl[3], half * dl[4]},
                     {half * dl[3], dl[1], half * dl[5]},
                     {half * dl[4], half * dl[5], dl[2]}};
  T FS[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      FS[r][c] = Fm[r][0] * S[0][c] + Fm[r][1] * S[1][c] + Fm[r][2] * S[2][c];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      dx[(3 * a + r) * ld + s] =
          X[a].w * vol * (FS[r][0] * G[a][0] + FS[r][1] * G[a][1] + FS[r][2] * G[a][2]);
}

// Warm start of the elastic slots whose multipliers are still zero:
// lambda takes lambda_0 and each vertex the matching displacement
// (1/m_a) grad_a C(x)^T lambda_0 (XPBD keeps x = x~ + M^-1 grad C^T
// lambda), written per (vertex a, slot) into dx (12 ld: component 3a + d of
// slot s at (3a + d) ld + s).  Other slots write zeros.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_learn_disp(SweepParams<T> P, const T *lam0, T *dx) {
  const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (s >= P.ld) return;
  const int64_t ld = P.ld;
  bool el = P.vid[s] >= 0 && (!P.flags || P.flags[s]);
  if (el)
#pragma unroll
    for (int i = 0; i < 6; ++i) el = el && P.lam[i * ld + s] == T(0);
  if (!el) {
#pragma unroll
    for (int k = 0; k < 12; ++k) dx[k * ld + s] = T(0);
    return;
  }
  vec4_t<T> X[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) X[a] = P.x[P.vid[a * ld + s]];
  learn_disp_slot<T>(P, s, lam0, dx, X);
}

// Each unpinned vertex adds its incident (vertex a, slot s) displacements in
// ascending slot order (inc: CSR of a * ld + s per vertex).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_learn_gather(int64_t nV, int64_t ld,
                                                         const int32_t *incOff,
                                                         const int32_t *inc, const T *dx,
                                                         const uint8_t *pinMask, vec4_t<T> *x) {
  const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (v >= nV || pinMask[3 * v]) return;
  T d[3] = {T(0), T(0), T(0)};
  for (int q = incOff[v]; q < incOff[v + 1]; ++q) {
    const int32_t e = inc[q];
    const int64_t a = e / ld, s = e % ld;
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] += dx[(3 * a + c) * ld + s];
  }
  vec4_t<T> p = x[v];
  p.x += d[0];
  p.y += d[1];
  p.z += d[2];
  x[v] = p;
}

// Keep the warm start only if it does not raise the equation residual:
// res[0] cold, res[1] warm.  Otherwise restore the saved positions and multipliers.
// counts[0] / counts[1]: warm starts kept / rejected.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_learn_select(int64_t nV, int64_t nLam,
                                                         const double *res, vec4_t<T> *x,
                                                         const vec4_t<T> *xs, T *lam,
                                                         const T *lams, unsigned long long *counts) {
  const double cold = res[0], warm = res[1];
  const bool keep = warm <= cold;  // false for a NaN warm residual
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) counts[keep ? 0 : 1] += 1ull;
  if (keep) return;
  if (i < nV) x[i] = xs[i];
  if (i < nLam) lam[i] = lams[i];
}

}  // namespace mlx
