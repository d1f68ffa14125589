// Device kernels of libmlxpbd (gfx950).  Included once, by mlx_solver.hip.
//
// Layout: per-element data lives in color-major "slots" as structure-of-
// arrays (field f of slot s at f*ld + s), so the elements of one graph color
// are contiguous and every field load of a wavefront is one coalesced
// transaction.  Vertex positions are vec4 (x, y, z, 1/mass) so one 16- or
// 32-byte load per vertex brings the position and the inverse mass used by
// the projection.  Rigid-body, boundary and contact arithmetic is always
// double; the element projection runs in T (float or double).
#pragma once
#include <hip/hip_runtime.h>

#include "mlx_internal.h"


namespace mlx {

constexpr int kBlock = 256;

// Per-body state (double), one record per rigid body of the current layer.
struct Body {
  double M, invM;
  double COM[3], COMDOT[3], omega[3];
  double R[9], I0[9], I[9], Rstep[9];  // column-major
  double angle, Iz;                    // 2D pose angle and scalar inertia
};

__device__ __forceinline__ bool iteration_stopped(const Ctl *ctl) {
  return (ctl->stopSolve | ctl->endLayer) != 0;
}

// Load of data that other workgroups of the same launch write (COH): device-
// scope relaxed atomic loads in 8- or 4-byte pieces, which bypass the CU's
// vector L1 and read the L2, so no L1 invalidation is needed after a
// barrier among workgroups of one XCD.  Otherwise an ordinary load.
template <bool COH, typename V>
__device__ __forceinline__ V ldc(const V *p) {
  if constexpr (!COH) {
    return *p;
  } else if constexpr (sizeof(V) % 8 == 0) {
    uint64_t w[sizeof(V) / 8];
    const uint64_t *q = reinterpret_cast<const uint64_t *>(p);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(V) / 8); ++k)
      w[k] = __hip_atomic_load(q + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(V, w);
  } else {
    static_assert(sizeof(V) % 4 == 0, "ldc: size must be a multiple of 4");
    uint32_t w[sizeof(V) / 4];
    const uint32_t *q = reinterpret_cast<const uint32_t *>(p);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(V) / 4); ++k)
      w[k] = __hip_atomic_load(q + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(V, w);
  }
}

// ---------------------------------------------------------------------------
// small dense helpers (double)
// ---------------------------------------------------------------------------

// A x = b, 3x3 column-major, Gaussian elimination with partial pivoting
// (the method of Matlab's mldivide on a general square page).
__device__ __forceinline__ void solve3_pivot(const double *Ain, const double *bin,
                                             double *x) {
  double A[9], b[3];
#pragma unroll
  for (int k = 0; k < 9; ++k) A[k] = Ain[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) b[k] = bin[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int piv = k;
    double best = fabs(A[k + 3 * k]);
    for (int i = k + 1; i < 3; ++i) {
      double v = fabs(A[i + 3 * k]);
      if (v > best) { best = v; piv = i; }
    }
    if (piv != k) {
      for (int j = 0; j < 3; ++j) {
        double t = A[k + 3 * j]; A[k + 3 * j] = A[piv + 3 * j]; A[piv + 3 * j] = t;
      }
      double t = b[k]; b[k] = b[piv]; b[piv] = t;
    }
    double d = A[k + 3 * k];
    for (int i = k + 1; i < 3; ++i) {
      double l = A[i + 3 * k] / d;
      for (int j = k + 1; j < 3; ++j) A[i + 3 * j] = A[i + 3 * j] - l * A[k + 3 * j];
      b[i] = b[i] - l * b[k];
    }
  }
  for (int i = 2; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < 3; ++j) s = s - A[i + 3 * j] * x[j];
    x[i] = s / A[i + 3 * i];
  }
}

__device__ __forceinline__ void matmul3(const double *A, const double *B, double *C) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      C[i + 3 * j] = A[i] * B[3 * j] + A[i + 3] * B[1 + 3 * j] + A[i + 6] * B[2 + 3 * j];
}

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// R = exp(h [omega]x) by Rodrigues' formula; omega = 0 gives I exactly.
__device__ __forceinline__ void exp_rodrigues(const double *om, double h, double *R) {
  double mag = sqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
  double w[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    w[i] = om[i] / mag;
    if (w[i] != w[i]) w[i] = 0;
  }
  double th = h * mag;
  double K[9] = {0, w[2], -w[1], -w[2], 0, w[0], w[1], -w[0], 0};
  double K2[9];
  matmul3(K, K, K2);
  double s = sin(th), c1 = 1 - cos(th);
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k] = ((k % 4) == 0 ? 1.0 : 0.0) + s * K[k] + c1 * K2[k];
}

// pos = R r + c
__device__ __forceinline__ void rigid_pos(const double *R, const double *r,
                                          const double *c, double *out) {
#pragma unroll
  for (int i = 0; i < 3; ++i) out[i] = (R[i] * r[0] + R[i + 3] * r[1] + R[i + 6] * r[2]) + c[i];
}

// I = R (I0 R^T)
__device__ __forceinline__ void inertia_update(const double *R, const double *I0, double *I) {
  double Rt[9], tmp[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Rt[i + 3 * j] = R[j + 3 * i];
  matmul3(I0, Rt, tmp);
  matmul3(R, tmp, I);
}

// w = minv + (r x n) . (I \ (r x n))
__device__ __forceinline__ double gen_inv_mass(double minv, const double *r,
                                               const double *I, const double *n) {
  double ax[3], ic[3];
  cross3(r, n, ax);
  solve3_pivot(I, ax, ic);
  return minv + (ax[0] * ic[0] + ax[1] * ic[1] + ax[2] * ic[2]);
}

// Block-wide sum of NV doubles per thread into out[0..NV) (valid in all
// threads after the call).  Fixed tree: deterministic.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *sh, double *out) {
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < NV; ++k) sh[k * kBlock + t] = v[k];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (t < s) {
#pragma unroll
      for (int k = 0; k < NV; ++k) sh[k * kBlock + t] += sh[k * kBlock + t + s];
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = sh[k * kBlock];
  __syncthreads();
}

// Workgroup sum of one double per thread in a fixed order: a butterfly
// within each wavefront (lane 0's value), then the wavefront sums in
// wavefront order.  Needs NT / 64 doubles of LDS; the result is valid in
// every thread.
template <int NT>
__device__ __forceinline__ double wg_sum(double v, double *shw) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) shw[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) t += shw[w];
  __syncthreads();
  return t;
}

// M x = n for a 3x3 matrix by cofactors (x_i = sum_j cof_ji n_j / det).
template <typename T>
__device__ __forceinline__ void solve3_cof(const T (&M)[3][3], const T (&n)[3], T (&x)[3]) {
  const T c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
  const T c01 = M[1][2] * M[2][0] - M[1][0] * M[2][2];
  const T c02 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
  const T c10 = M[0][2] * M[2][1] - M[0][1] * M[2][2];
  const T c11 = M[0][0] * M[2][2] - M[0][2] * M[2][0];
  const T c12 = M[0][1] * M[2][0] - M[0][0] * M[2][1];
  const T c20 = M[0][1] * M[1][2] - M[0][2] * M[1][1];
  const T c21 = M[0][2] * M[1][0] - M[0][0] * M[1][2];
  const T c22 = M[0][0] * M[1][1] - M[0][1] * M[1][0];
  const T idet = T(1) / (M[0][0] * c00 + M[0][1] * c01 + M[0][2] * c02);
  x[0] = (c00 * n[0] + c10 * n[1] + c20 * n[2]) * idet;
  x[1] = (c01 * n[0] + c11 * n[1] + c21 * n[2]) * idet;
  x[2] = (c02 * n[0] + c12 * n[1] + c22 * n[2]) * idet;
}

// ---------------------------------------------------------------------------
// element projection: one Gauss-Seidel pass over one color
// ---------------------------------------------------------------------------

template <typename T>
struct SweepParams {
  const int32_t *__restrict__ vid;
  const T *__restrict__ dminv;
  const T *__restrict__ vol;
  const T *__restrict__ mat;  // alpha_d, alpha_o, alpha_s, alpha1
  T *__restrict__ lam;
  const uint8_t *__restrict__ flags;  // elastic flags (slot order) or null
  vec4_t<T> *__restrict__ x;
  const vec4_t<T> *__restrict__ xold;
  const Ctl *ctl;
  int64_t ld, begin, end;
  T ih2, ih;  // 1/h^2, 1/h
  int nh;     // some slot holds a NeoHookean tetrahedron (uniform per launch)
};

// StVK-Voigt tetrahedron (matStvkVoigtConstraint3D.m:1-45).  With F the
// deformation gradient, Cg = F^T F, E = (Cg - I)/2, constraint i = vol*E_jk
// and its gradient w.r.t. vertex a is vol * F P_i g_a, g_a the rows of
// [-1^T DmInv; DmInv] and P_i the symmetric selector of (j,k).  Hence
//   gradC_i . gradC_j weighted by 1/m  = vol^2 tr(P_i Cg P_j W),
//   W = sum_a (1/m_a) g_a g_a^T,  gradC_i . (x - x_old) = vol tr(P_i F^T Fd),
// and the displacement is (1/m_a) vol F S g_a with S = sum_i dlambda_i P_i.
// gamma = alpha~ beta~ / h is alpha1/h times I for the isotropic material.
// The normal block is solved as a 3x3 system, the three shear rows as
// scalars with the (4,4) compliance, as decomposeSolve does.
// What a tet projection reads besides the current positions: slot data and
// the positions at the start of the substep (neither changes during a
// layer solve), so it can be fetched ahead of the barrier that publishes
// the positions.  lam is the slot's own lambda, written only by the thread
// that projects the slot.
template <typename T>
struct TetIn {
  int v[4];
  T D[9];
  T vol, tad, tao, tas, g;
  T m0, m1, m2, m3;  // the material scalars as stored (the solves scale them)
  T lam[6];
  vec4_t<T> O[4];
};

// Material mode of a specialised projection / residual: StVK-Voigt slots
// only, NeoHookean slots only, or either (chosen per slot from the sign of
// the first material scalar).  The batched kernel runs a scene with the
// single-material body, the per-color launches of a mesh with NeoHookean
// elements the mixed one.
constexpr int kMatStVK = 0, kMatMixed = 1, kMatNH = 2;

// Third material scalar as the projection uses it: alpha_s / h^2
// (StVK-Voigt) or gamma as stored (NeoHookean).
template <typename T, int MAT>
__device__ __forceinline__ T mat_third(T m0, T m2, T ih2) {
  if constexpr (MAT == kMatStVK) return m2 * ih2;
  else if constexpr (MAT == kMatNH) return m2;
  else return m0 < T(0) ? m2 : m2 * ih2;
}

// Everything but the vertex ids, which are in in.v already.
template <typename T, int MAT = 0>
__device__ __forceinline__ void tet_fetch_rest(const SweepParams<T> &P, const int64_t s,
                                               TetIn<T> &in) {
  const int64_t ld = P.ld;
#pragma unroll
  for (int a = 0; a < 4; ++a) in.O[a] = P.xold[in.v[a]];
#pragma unroll
  for (int k = 0; k < 9; ++k) in.D[k] = P.dminv[k * ld + s];
  in.vol = P.vol[s];
  const T m0 = P.mat[s], m2 = P.mat[2 * ld + s];
  const T m1 = P.mat[ld + s], m3 = P.mat[3 * ld + s];
  in.m0 = m0;
  in.m1 = m1;
  in.m2 = m2;
  in.m3 = m3;
  in.tad = m0 * P.ih2;
  in.tao = m1 * P.ih2;
  in.tas = mat_third<T, MAT>(m0, m2, P.ih2);
  in.g = m3 * P.ih;
#pragma unroll
  for (int i = 0; i < 6; ++i) in.lam[i] = P.lam[i * ld + s];
}

template <typename T, int MAT = 0>
__device__ __forceinline__ void tet_fetch(const SweepParams<T> &P, const int64_t s, TetIn<T> &in) {
#pragma unroll
  for (int a = 0; a < 4; ++a) in.v[a] = P.vid[a * P.ld + s];
  tet_fetch_rest<T, MAT>(P, s, in);
}

// Material scalars of a slot.  StVK-Voigt: alpha_d = alpha(1,1),
// alpha_o = alpha(1,2), alpha_s = alpha(4,4), alpha1 (all > 0 but alpha_o).
// NeoHookean tetrahedron (the deviatoric / volumetric pair of
// mexNeoHookeanDConstraint.cpp:46-138 and mexNeoHookeanHConstraint.cpp:47-138
// lifted to three dimensions): -1/mu, 1/lambda, gamma = 1 + mu/lambda, 0.
// The negative first scalar tags the slot (alpha(1,1) of an isotropic
// stiffness is positive), so one slot layout serves both materials.

// F of a tetrahedron from its current vertices and DmInv (column-major).
template <typename T>
__device__ __forceinline__ void tet_F(const vec4_t<T> (&X)[4], const T (&D)[9], T (&F)[3][3]) {
  const T e[3][3] = {{X[1].x - X[0].x, X[1].y - X[0].y, X[1].z - X[0].z},
                     {X[2].x - X[0].x, X[2].y - X[0].y, X[2].z - X[0].z},
                     {X[3].x - X[0].x, X[3].y - X[0].y, X[3].z - X[0].z}};
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      F[r][c] = e[0][r] * D[3 * c] + e[1][r] * D[1 + 3 * c] + e[2][r] * D[2 + 3 * c];
}

// ||F||_F and cof(F) (the gradient of det F with respect to F); det F.
template <typename T>
__device__ __forceinline__ void nh_terms(const T (&F)[3][3], T &n, T (&PH)[3][3], T &J) {
  T n2 = T(0);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) n2 += F[r][c] * F[r][c];
  n = sqrt(n2);
  PH[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
  PH[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
  PH[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
  PH[1][0] = F[0][2] * F[2][1] - F[0][1] * F[2][2];
  PH[1][1] = F[0][0] * F[2][2] - F[0][2] * F[2][0];
  PH[1][2] = F[0][1] * F[2][0] - F[0][0] * F[2][1];
  PH[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
  PH[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
  PH[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
  J = F[0][0] * PH[0][0] + F[0][1] * PH[0][1] + F[0][2] * PH[0][2];
}

// The NeoHookean pair on one tetrahedron: C_D = ||F||_F with compliance
// 1/(vol h^2 mu), C_H = det F - gamma with compliance 1/(vol h^2 lambda),
// both evaluated at the positions the projection starts from; each is a
// scalar XPBD step dl = -(C + a~ l) / (sum_a w_a |g_a|^2 + a~) with vertex
// gradients g_a = (dC/dF) grad_a, grad_a the rows of [-1^T DmInv; DmInv];
// the displacements are summed, D first.  lambda_D, lambda_H are the slot's
// first two multipliers.
template <typename T, typename Store>
__device__ __forceinline__ void tet_solve_nh_st(const SweepParams<T> &P, const int64_t s,
                                                const TetIn<T> &in, const vec4_t<T> (&X)[4],
                                                Store &&st) {
  const int64_t ld = P.ld;
  T D[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) D[k] = in.D[k];
  T F[3][3], PH[3][3], n, J;
  tet_F<T>(X, D, F);
  nh_terms<T>(F, n, PH, J);
  T G[4][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    G[1][c] = D[3 * c];
    G[2][c] = D[1 + 3 * c];
    G[3][c] = D[2 + 3 * c];
    G[0][c] = -(G[1][c] + G[2][c] + G[3][c]);
  }
  const T w[4] = {X[0].w, X[1].w, X[2].w, X[3].w};
  // the scaled material scalars are formed here, next to their uses, so
  // every kernel that runs this solve rounds them the same way
  const T iv = T(1) / in.vol;
  const T aD = iv * (-(in.m0 * P.ih2)), aH = iv * (in.m1 * P.ih2), gam = in.m2;
  // vertex gradients: gD_a = F g_a / n, gH_a = cof(F) g_a; formed twice (for
  // the denominators, then for the displacements) rather than kept live.
  // One reciprocal of n serves the 24 components: a correctly rounded
  // division costs about ten instructions, and the scenes that run this
  // solve to the iteration cap are bound by the CU's vector ALU.
  const T rn = T(1) / n;
  auto grads = [&](int a, T (&gD)[3], T (&gH)[3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      gD[r] = (F[r][0] * G[a][0] + F[r][1] * G[a][1] + F[r][2] * G[a][2]) * rn;
      gH[r] = PH[r][0] * G[a][0] + PH[r][1] * G[a][1] + PH[r][2] * G[a][2];
    }
  };
  T sD = aD, sH = aH;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    T gD[3], gH[3];
    grads(a, gD, gH);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      sD += w[a] * (gD[r] * gD[r]);
      sH += w[a] * (gH[r] * gH[r]);
    }
  }
  const T dlD = -(n + aD * in.lam[0]) / sD;
  const T dlH = -((J - gam) + aH * in.lam[1]) / sH;
  P.lam[s] = in.lam[0] + dlD;
  P.lam[ld + s] = in.lam[1] + dlH;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    T gD[3], gH[3];
    grads(a, gD, gH);
    vec4_t<T> o = X[a];
    o.x = (o.x + w[a] * gD[0] * dlD) + w[a] * gH[0] * dlH;
    o.y = (o.y + w[a] * gD[1] * dlD) + w[a] * gH[1] * dlH;
    o.z = (o.z + w[a] * gD[2] * dlD) + w[a] * gH[2] * dlH;
    st(a, o);
  }
}

// The projection proper, on fetched slot data and the current positions X
// of the tet's four vertices; the moved vertex a is handed to st(a, x).
template <typename T, int MAT = 0, typename Store>
__device__ __forceinline__ void tet_solve_st(const SweepParams<T> &P, const int64_t s,
                                             const TetIn<T> &in, const vec4_t<T> (&X)[4],
                                             Store &&st) {
  if constexpr (MAT == kMatNH) {
    tet_solve_nh_st<T>(P, s, in, X, st);
    return;
  } else if constexpr (MAT == kMatMixed) {
    if (in.m0 < T(0)) {
      tet_solve_nh_st<T>(P, s, in, X, st);
      return;
    }
  }
  const int64_t ld = P.ld;
  const vec4_t<T> X0 = X[0], X1 = X[1], X2 = X[2], X3 = X[3];
  const vec4_t<T> O0 = in.O[0], O1 = in.O[1], O2 = in.O[2], O3 = in.O[3];
  T D[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) D[k] = in.D[k];
  const T vol = in.vol;
  // the scaled material scalars are formed here, next to their uses, so
  // every kernel that runs this solve rounds (and fuses) them the same way
  const T tad = in.m0 * P.ih2, tao = in.m1 * P.ih2, tas = in.m2 * P.ih2;
  const T g = in.m3 * P.ih;

  // edge vectors (current and displacement since x_old)
  const T e[3][3] = {{X1.x - X0.x, X1.y - X0.y, X1.z - X0.z},
                     {X2.x - X0.x, X2.y - X0.y, X2.z - X0.z},
                     {X3.x - X0.x, X3.y - X0.y, X3.z - X0.z}};
  const T d0[3] = {X0.x - O0.x, X0.y - O0.y, X0.z - O0.z};
  const T f[3][3] = {{(X1.x - O1.x) - d0[0], (X1.y - O1.y) - d0[1], (X1.z - O1.z) - d0[2]},
                     {(X2.x - O2.x) - d0[0], (X2.y - O2.y) - d0[1], (X2.z - O2.z) - d0[2]},
                     {(X3.x - O3.x) - d0[0], (X3.y - O3.y) - d0[1], (X3.z - O3.z) - d0[2]}};
  T F[3][3], Fd[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      F[r][c] = e[0][r] * D[3 * c] + e[1][r] * D[1 + 3 * c] + e[2][r] * D[2 + 3 * c];
      Fd[r][c] = f[0][r] * D[3 * c] + f[1][r] * D[1 + 3 * c] + f[2][r] * D[2 + 3 * c];
    }
  // g_a: g1..g3 = rows of DmInv, g0 = -(g1+g2+g3)
  T G[4][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    G[1][c] = D[3 * c];
    G[2][c] = D[1 + 3 * c];
    G[3][c] = D[2 + 3 * c];
    G[0][c] = -(G[1][c] + G[2][c] + G[3][c]);
  }
  const T w[4] = {X0.w, X1.w, X2.w, X3.w};
  T Cg[3][3], H[3][3], W[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Cg[j][k] = F[0][j] * F[0][k] + F[1][j] * F[1][k] + F[2][j] * F[2][k];
      H[j][k] = F[0][j] * Fd[0][k] + F[1][j] * Fd[1][k] + F[2][j] * Fd[2][k];
      W[j][k] = w[0] * G[0][j] * G[0][k] + w[1] * G[1][j] * G[1][k] +
                w[2] * G[2][j] * G[2][k] + w[3] * G[3][j] * G[3][k];
    }
  const T half = T(0.5), vol2 = vol * vol;
  T lam[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) lam[i] = in.lam[i];
  // normal block
  T C[3], num[3], M[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    C[j] = vol * (half * (Cg[j][j] - T(1)));
    const int k = (j + 1) % 3, l = (j + 2) % 3;
    num[j] = -C[j] - (tad * lam[j] + tao * lam[k] + tao * lam[l]) - g * (vol * H[j][j]);
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      M[j][k] = (T(1) + g) * (vol2 * Cg[j][k] * W[j][k]) + (j == k ? tad : tao);
  T dl[6];
  {
    T xs[3];
    solve3_cof(M, num, xs);
    dl[0] = xs[0];
    dl[1] = xs[1];
    dl[2] = xs[2];
  }
  // shear rows (0,1), (0,2), (1,2)
  const int sj[3] = {0, 0, 1}, sk[3] = {1, 2, 2};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int j = sj[q], k = sk[q];
    const T Cs = vol * (half * Cg[j][k]);
    const T gd = vol * (half * (H[j][k] + H[k][j]));
    const T A = vol2 * (T(0.25) * (T(2) * Cg[j][k] * W[j][k] + Cg[k][k] * W[j][j] + Cg[j][j] * W[k][k]));
    dl[3 + q] = (-Cs - tas * lam[3 + q] - g * gd) / ((T(1) + g) * A + tas);
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) P.lam[i * ld + s] = lam[i] + dl[i];
  // S = sum_i dl_i P_i, FS = F S
  const T S[3][3] = {{dl[0], half * dl[3], half * dl[4]},
                     {half * dl[3], dl[1], half * dl[5]},
                     {half * dl[4], half * dl[5], dl[2]}};
  T FS[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) FS[r][c] = F[r][0] * S[0][c] + F[r][1] * S[1][c] + F[r][2] * S[2][c];
  const vec4_t<T> Xs[4] = {X0, X1, X2, X3};
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const T sc = w[a] * vol;
    vec4_t<T> o = Xs[a];
    o.x += sc * (FS[0][0] * G[a][0] + FS[0][1] * G[a][1] + FS[0][2] * G[a][2]);
    o.y += sc * (FS[1][0] * G[a][0] + FS[1][1] * G[a][1] + FS[1][2] * G[a][2]);
    o.z += sc * (FS[2][0] * G[a][0] + FS[2][1] * G[a][1] + FS[2][2] * G[a][2]);
    st(a, o);
  }
}

template <typename T, int MAT = 0>
__device__ __forceinline__ void tet_solve_x(const SweepParams<T> &P, const int64_t s,
                                            const TetIn<T> &in, const vec4_t<T> (&X)[4]) {
  tet_solve_st<T, MAT>(P, s, in, X, [&](int a, const vec4_t<T> &o) { P.x[in.v[a]] = o; });
}

// Current positions of a fetched tet's vertices (ldc<COH>).
template <typename T, bool COH>
__device__ __forceinline__ void tet_gather(const SweepParams<T> &P, const TetIn<T> &in,
                                           vec4_t<T> (&X)[4]) {
#pragma unroll
  for (int a = 0; a < 4; ++a) X[a] = ldc<COH>(P.x + in.v[a]);
}

template <typename T, bool COH, int MAT = 0>
__device__ __forceinline__ void tet_solve(const SweepParams<T> &P, const int64_t s,
                                          const TetIn<T> &in) {
  vec4_t<T> X[4];
  tet_gather<T, COH>(P, in, X);
  tet_solve_x<T, MAT>(P, s, in, X);
}

template <typename T, int MAT = 0>
__device__ __forceinline__ void project_tet(const SweepParams<T> &P, const int64_t s) {
  TetIn<T> in;
  tet_fetch<T, MAT>(P, s, in);
  tet_solve<T, false, MAT>(P, s, in);
}

// The slots s, s + step, ... < e of one color for one thread, two at a time:
// both tets' slot data and positions are loaded before either is solved.
// The tets of a color share no vertex, so no other thread writes these
// positions while the color runs, and this thread's first solve does not
// touch the second tet's vertices.
template <typename T, bool COH>
__device__ __forceinline__ void color_slots(const SweepParams<T> &P, int64_t s, const int64_t e,
                                            const int64_t step) {
  for (; s + step < e; s += 2 * step) {
    TetIn<T> a, b;
    vec4_t<T> Xa[4], Xb[4];
    tet_fetch<T>(P, s, a);
    tet_fetch<T>(P, s + step, b);
    tet_gather<T, COH>(P, a, Xa);
    tet_gather<T, COH>(P, b, Xb);
    tet_solve_x<T>(P, s, a, Xa);
    tet_solve_x<T>(P, s + step, b, Xb);
  }
  if (s < e) {
    TetIn<T> a;
    tet_fetch<T>(P, s, a);
    tet_solve<T, COH>(P, s, a);
  }
}

// Keeps a loaded value live up to this point, so the load cannot be moved
// below a later branch (the empty asm emits no instruction).
template <typename T>
__device__ __forceinline__ void hold(const T &v) {
  asm volatile("" ::"v"(v));
}

// Stand-in read when a launch has no control block (never written: global
// memory, so the selected address stays global).
__device__ int kNoStop[2] = {0, 0};

// One slot of one color.  In a layer with rigid elements (P.flags set) a
// rigid slot returns after its flag, so its data is never read.  The
// control word (has the monitor stopped the layer?) and the slot's vertex
// ids are loaded together and tested together, so the ids' latency, the
// first link of the id -> position chain, overlaps the control word's; a
// launch after the stop then returns having read only the ids.  The rest
// of the slot data and the positions follow.
template <typename T, int MAT>
__device__ __forceinline__ void sweep_slot(const SweepParams<T> &P, const int64_t s) {
  if (P.flags && !P.flags[s]) return;
  // address chosen without a branch, so the loads issue together
  const int *cw = P.ctl ? &P.ctl->stopSolve : kNoStop;
  static_assert(offsetof(Ctl, endLayer) == offsetof(Ctl, stopSolve) + sizeof(int),
                "stopSolve and endLayer are adjacent");
  const int stopWord = cw[0] | cw[1];
  TetIn<T> in;
#pragma unroll
  for (int a = 0; a < 4; ++a) in.v[a] = P.vid[a * P.ld + s];
#pragma unroll
  for (int a = 0; a < 4; ++a) hold(in.v[a]);
  if (stopWord != 0) return;
  tet_fetch_rest<T, MAT>(P, s, in);
  vec4_t<T> X[4];
  tet_gather<T, false>(P, in, X);
  tet_solve_x<T, MAT>(P, s, in, X);
}

// One color of a layer, one thread per slot.  MAT is kMatMixed for a mesh
// with NeoHookean elements (P.nh), else kMatStVK: chosen at launch, so the
// kernel body has no branch between its loads.
template <typename T, int MAT>
__global__ __launch_bounds__(kBlock) void k_sweep3d(SweepParams<T> P) {
  const int64_t s = P.begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.end) return;
  sweep_slot<T, MAT>(P, s);
}

// Elastic slots of each color of a layer, compacted in ascending slot order
// into list[colorOff[c] ..) (a color's list fits in its own slot range),
// their number into cnt[c].  One workgroup per color.
__global__ __launch_bounds__(kBlock) void k_color_lists(const int64_t *colorOff,
                                                        const int64_t *colorCnt,
                                                        const uint8_t *flags, int32_t *list,
                                                        int32_t *cnt) {
  __shared__ int32_t sh[kBlock];
  __shared__ int32_t base;
  const int c = blockIdx.x, t = threadIdx.x;
  const int64_t a = colorOff[c], n = colorCnt[c];
  if (t == 0) base = 0;
  __syncthreads();
  for (int64_t s0 = 0; s0 < n; s0 += kBlock) {
    const int64_t s = a + s0 + t;
    const int f = (s0 + t < n && flags[s]) ? 1 : 0;
    sh[t] = f;
    __syncthreads();
    // inclusive scan (Hillis-Steele) over the workgroup
    for (int o = 1; o < kBlock; o <<= 1) {
      const int32_t v = t >= o ? sh[t - o] : 0;
      __syncthreads();
      sh[t] += v;
      __syncthreads();
    }
    if (f) list[a + base + sh[t] - 1] = (int32_t)s;
    __syncthreads();
    if (t == kBlock - 1) base += sh[t];
    __syncthreads();
  }
  if (t == 0) cnt[c] = base;
}

// One projectGS sweep over all colors of a layer with few elastic elements
// per color, in one workgroup: the compacted elastic slots of each color
// (k_color_lists), colors separated by workgroup barriers (which wait for
// the workgroup's stores; its waves share one L1).  Same projections in the
// same color order as one k_sweep3d launch per color, through the same
// slot routine (sweep_slot), so the arithmetic is the same instruction for
// instruction and the states are bit-identical.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_sweep_wg(SweepParams<T> P, const int64_t *colorOff,
                                                 const int32_t *list, const int32_t *cnt,
                                                 int numColors) {
  if (P.ctl && iteration_stopped(P.ctl)) return;
  for (int c = 0; c < numColors; ++c) {
    const int n = cnt[c];
    const int32_t *l = list + colorOff[c];
    for (int i = threadIdx.x; i < n; i += NT) {
      if (P.nh)
        sweep_slot<T, kMatMixed>(P, l[i]);
      else
        sweep_slot<T, kMatStVK>(P, l[i]);
    }
    __syncthreads();
  }
}

// StVK-Voigt triangle (matStvkVoigtConstraint.m:1-34): same algebra in 2D
// with constraints (E11, E22, E12) and the full 3x3 solve of justSolve.
template <typename T>
__device__ __forceinline__ void project_tri(const SweepParams<T> &P, const int64_t s) {
  const int64_t ld = P.ld;
  const int v0 = P.vid[s], v1 = P.vid[ld + s], v2 = P.vid[2 * ld + s];
  const vec4_t<T> X0 = P.x[v0], X1 = P.x[v1], X2 = P.x[v2];
  const vec4_t<T> O0 = P.xold[v0], O1 = P.xold[v1], O2 = P.xold[v2];
  const T D[4] = {P.dminv[s], P.dminv[ld + s], P.dminv[2 * ld + s], P.dminv[3 * ld + s]};
  const T vol = P.vol[s];
  const T tad = P.mat[s] * P.ih2, tao = P.mat[ld + s] * P.ih2, tas = P.mat[2 * ld + s] * P.ih2;
  const T g = P.mat[3 * ld + s] * P.ih;
  // Ds = [p0 - p2, p1 - p2], F = Ds DmInv
  const T e[2][2] = {{X0.x - X2.x, X0.y - X2.y}, {X1.x - X2.x, X1.y - X2.y}};
  const T d2[2] = {X2.x - O2.x, X2.y - O2.y};
  const T f[2][2] = {{(X0.x - O0.x) - d2[0], (X0.y - O0.y) - d2[1]},
                     {(X1.x - O1.x) - d2[0], (X1.y - O1.y) - d2[1]}};
  T F[2][2], Fd[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      F[r][c] = e[0][r] * D[2 * c] + e[1][r] * D[1 + 2 * c];
      Fd[r][c] = f[0][r] * D[2 * c] + f[1][r] * D[1 + 2 * c];
    }
  T G[3][2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    G[0][c] = D[2 * c];
    G[1][c] = D[1 + 2 * c];
    G[2][c] = -(G[0][c] + G[1][c]);
  }
  const T w[3] = {X0.w, X1.w, X2.w};
  T Cg[2][2], H[2][2], W[2][2];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      Cg[j][k] = F[0][j] * F[0][k] + F[1][j] * F[1][k];
      H[j][k] = F[0][j] * Fd[0][k] + F[1][j] * Fd[1][k];
      W[j][k] = w[0] * G[0][j] * G[0][k] + w[1] * G[1][j] * G[1][k] + w[2] * G[2][j] * G[2][k];
    }
  const T half = T(0.5), vol2 = vol * vol;
  const int pa[3] = {0, 1, 0}, pb[3] = {0, 1, 1};
  T lam[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) lam[i] = P.lam[i * ld + s];
  T C[3], gd[3];
  C[0] = vol * (half * (Cg[0][0] - T(1)));
  C[1] = vol * (half * (Cg[1][1] - T(1)));
  C[2] = vol * (half * Cg[0][1]);
  gd[0] = vol * H[0][0];
  gd[1] = vol * H[1][1];
  gd[2] = vol * (half * (H[0][1] + H[1][0]));
  const T at[3][3] = {{tad, tao, T(0)}, {tao, tad, T(0)}, {T(0), T(0), tas}};
  T M[3][3], num[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    num[i] = -C[i] - (at[i][0] * lam[0] + at[i][1] * lam[1] + at[i][2] * lam[2]) - g * gd[i];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int a = pa[i], b = pb[i], c = pa[j], d = pb[j];
      const T A = vol2 * (T(0.25) * (Cg[b][c] * W[d][a] + Cg[b][d] * W[c][a] +
                                     Cg[a][c] * W[d][b] + Cg[a][d] * W[c][b]));
      M[i][j] = (T(1) + g) * A + at[i][j];
    }
  }
  T dl[3];
  solve3_cof(M, num, dl);
#pragma unroll
  for (int i = 0; i < 3; ++i) P.lam[i * ld + s] = lam[i] + dl[i];
  const T S[2][2] = {{dl[0], half * dl[2]}, {half * dl[2], dl[1]}};
  T FS[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) FS[r][c] = F[r][0] * S[0][c] + F[r][1] * S[1][c];
  const vec4_t<T> Xs[3] = {X0, X1, X2};
  const int vs[3] = {v0, v1, v2};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const T sc = w[a] * vol;
    vec4_t<T> o = Xs[a];
    o.x += sc * (FS[0][0] * G[a][0] + FS[0][1] * G[a][1]);
    o.y += sc * (FS[1][0] * G[a][0] + FS[1][1] * G[a][1]);
    P.x[vs[a]] = o;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sweep2d(SweepParams<T> P) {
  if (P.ctl && iteration_stopped(P.ctl)) return;
  const int64_t s = P.begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.end) return;
  if (P.flags && !P.flags[s]) return;
  project_tri<T>(P, s);
}

}  // namespace mlx

namespace mlx {

// ---------------------------------------------------------------------------
// rigid layer: restriction (per body), step, prolongation
// ---------------------------------------------------------------------------

// Per-layer lists, CSR by body (0-based body k owns items off[k]..off[k+1]).
struct LayerLists {
  const int32_t *rig_off, *rig_vert;  // rigid vertices of each body
  const int32_t *bnd_off, *bnd_vert;  // boundary vertices of each body
  const int32_t *rc_off, *rc_idx;     // rigid contacts of each body
  const int32_t *ec_idx;              // elastic contacts
  const int32_t *chk_off;             // first restriction chunk of each body (nb + 1), or null
  const int32_t *bchk_off;            // first boundary chunk of each body (nb + 1), or null
  int nb, nbnd, nrc, nec, nchunks, nbchunks;
};

// Rigid vertices per workgroup of the chunked body restriction (one per thread).
constexpr int kRestrictChunk = 256;
// Boundary vertices per workgroup of the chunked boundary accumulation.
constexpr int kBoundaryChunk = 256;

// makeRigidsFromConnectivity / matComputeRigidBodyProperties3D
// (AdaptiveMesh3D.m:32-43, RigidBody3D.m:198-265): mass, COM, COM velocity,
// inertia sum(-m rhat rhat), angular momentum sum(m r x (v - vbar)),
// omega = I \ L, R = I, r_i = x - COM.  One workgroup per body, fixed
// reduction tree.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_restrict(LayerLists L, const vec4_t<T> *x,
                                                     const T *rv, const double *mass,
                                                     Body *bodies, double *ri) {
  __shared__ double sh[12 * kBlock];
  const int k = blockIdx.x;
  const int off = L.rig_off[k], n = L.rig_off[k + 1] - off;
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int q = threadIdx.x; q < n; q += kBlock) {
    const int v = L.rig_vert[off + q];
    const double m = mass[v];
    const vec4_t<T> p = x[v];
    a[0] += m;
    a[1] += (double)p.x * m;
    a[2] += (double)p.y * m;
    a[3] += (double)p.z * m;
    a[4] += (double)rv[3 * v] * m;
    a[5] += (double)rv[3 * v + 1] * m;
    a[6] += (double)rv[3 * v + 2] * m;
  }
  double s7[7];
  block_sum<7>(a, sh, s7);
  const double M = s7[0];
  const double com[3] = {s7[1] / M, s7[2] / M, s7[3] / M};
  const double cdot[3] = {s7[4] / M, s7[5] / M, s7[6] / M};
  double b[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = threadIdx.x; q < n; q += kBlock) {
    const int v = L.rig_vert[off + q];
    const double m = mass[v];
    const vec4_t<T> p = x[v];
    const double r[3] = {(double)p.x - com[0], (double)p.y - com[1], (double)p.z - com[2]};
    const double vd[3] = {(double)rv[3 * v] - cdot[0], (double)rv[3 * v + 1] - cdot[1],
                          (double)rv[3 * v + 2] - cdot[2]};
    // -m * (rhat rhat) = m (|r|^2 I - r r^T)
    const double rh[9] = {0, r[2], -r[1], -r[2], 0, r[0], r[1], -r[0], 0};
    double rr[9];
    matmul3(rh, rh, rr);
#pragma unroll
    for (int i = 0; i < 9; ++i) b[i] += (-m) * rr[i];
    double c[3];
    cross3(r, vd, c);
    b[9] += m * c[0];
    b[10] += m * c[1];
    b[11] += m * c[2];
    ri[3 * v] = r[0];
    ri[3 * v + 1] = r[1];
    ri[3 * v + 2] = r[2];
  }
  double s12[12];
  block_sum<12>(b, sh, s12);
  if (threadIdx.x == 0) {
    Body &B = bodies[k];
    B.M = M;
    B.invM = 1.0 / M;
    for (int d = 0; d < 3; ++d) {
      B.COM[d] = com[d];
      B.COMDOT[d] = cdot[d];
    }
    for (int i = 0; i < 9; ++i) {
      B.I0[i] = s12[i];
      B.I[i] = s12[i];
      B.R[i] = (i % 4 == 0) ? 1.0 : 0.0;
      B.Rstep[i] = B.R[i];
    }
    solve3_pivot(s12, s12 + 9, B.omega);
  }
}

// The same restriction with each body's rigid vertices split in chunks of
// kRestrictChunk, one workgroup per chunk, for bodies that span most of the
// mesh.  Two launches: k_restrict_mass (mass, COM, COM velocity) then
// k_restrict_inertia (r_i, inertia, angular momentum, omega).  Each chunk
// stores its sums past the caches and counts itself on its body's counter;
// the first wave of the body's last chunk adds the chunks' sums in a fixed
// order (chunk_total), finishes the body and re-arms the counter: the result
// does not depend on which chunk arrives last.
__device__ __forceinline__ int body_of_chunk(const int32_t *chk, int nb, int i) {
  int lo = 0, hi = nb;  // chk[lo] <= i < chk[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (chk[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Returns, in every thread of the workgroup, whether this chunk arrived
// last on its body's counter.
template <int NV>
__device__ __forceinline__ bool chunk_publish(const double (&v)[NV], double *part, unsigned *cnt,
                                              int i, int k, int nch, int *sLast) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j)
      __hip_atomic_store(part + 12 * (size_t)i + j, v[j], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    // release: the sums above are visible before the count; acquire: the
    // last chunk sees every other chunk's sums.
    const unsigned prev = __hip_atomic_fetch_add(cnt + k, 1u, __ATOMIC_ACQ_REL,
                                                 __HIP_MEMORY_SCOPE_AGENT);
    *sLast = prev == (unsigned)(nch - 1);
  }
  __syncthreads();
  return *sLast != 0;
}

// Sum of the body's chunk sums by the first wave of its last chunk (call
// with threadIdx.x < 64): lane l adds chunks l, l + 64, ... in order, then a
// fixed butterfly over the lanes; every lane ends with the same totals.
// Lane 0 re-arms the body's counter.
template <int NV>
__device__ __forceinline__ void chunk_total(const double *part, unsigned *cnt, int c0, int nch,
                                            int k, double (&t)[NV]) {
  const int lane = threadIdx.x;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#pragma unroll
  for (int j = 0; j < NV; ++j) t[j] = 0;
  for (int c = lane; c < nch; c += 64)
#pragma unroll
    for (int j = 0; j < NV; ++j)
      t[j] += __hip_atomic_load(part + 12 * (size_t)(c0 + c) + j, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int j = 0; j < NV; ++j) t[j] += __shfl_xor(t[j], o, 64);
  if (lane == 0) __hip_atomic_store(cnt + k, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_restrict_mass(LayerLists L, const vec4_t<T> *x,
                                                          const T *rv, const double *mass,
                                                          Body *bodies, double *part,
                                                          unsigned *cnt) {
  __shared__ double sh[7 * kBlock];
  __shared__ int sLast;
  const int i = blockIdx.x;
  const int k = body_of_chunk(L.chk_off, L.nb, i);
  const int c0 = L.chk_off[k], nch = L.chk_off[k + 1] - c0;
  const int off = L.rig_off[k], n = L.rig_off[k + 1] - off;
  const int q0 = (i - c0) * kRestrictChunk, q1 = min(n, q0 + kRestrictChunk);
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int q = q0 + threadIdx.x; q < q1; q += kBlock) {
    const int v = L.rig_vert[off + q];
    const double m = mass[v];
    const vec4_t<T> p = x[v];
    a[0] += m;
    a[1] += (double)p.x * m;
    a[2] += (double)p.y * m;
    a[3] += (double)p.z * m;
    a[4] += (double)rv[3 * v] * m;
    a[5] += (double)rv[3 * v + 1] * m;
    a[6] += (double)rv[3 * v + 2] * m;
  }
  double s7[7];
  block_sum<7>(a, sh, s7);
  if (!chunk_publish<7>(s7, part, cnt, i, k, nch, &sLast) || threadIdx.x >= 64) return;
  double t[7];
  chunk_total<7>(part, cnt, c0, nch, k, t);
  if (threadIdx.x != 0) return;
  Body &B = bodies[k];
  B.M = t[0];
  B.invM = 1.0 / t[0];
  for (int d = 0; d < 3; ++d) {
    B.COM[d] = t[1 + d] / t[0];
    B.COMDOT[d] = t[4 + d] / t[0];
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_restrict_inertia(LayerLists L, const vec4_t<T> *x,
                                                             const T *rv, const double *mass,
                                                             Body *bodies, double *ri,
                                                             double *part, unsigned *cnt) {
  __shared__ double sh[12 * kBlock];
  __shared__ int sLast;
  const int i = blockIdx.x;
  const int k = body_of_chunk(L.chk_off, L.nb, i);
  const int c0 = L.chk_off[k], nch = L.chk_off[k + 1] - c0;
  const int off = L.rig_off[k], n = L.rig_off[k + 1] - off;
  const int q0 = (i - c0) * kRestrictChunk, q1 = min(n, q0 + kRestrictChunk);
  const Body &Bk = bodies[k];
  const double com[3] = {Bk.COM[0], Bk.COM[1], Bk.COM[2]};
  const double cdot[3] = {Bk.COMDOT[0], Bk.COMDOT[1], Bk.COMDOT[2]};
  double b[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = q0 + threadIdx.x; q < q1; q += kBlock) {
    const int v = L.rig_vert[off + q];
    const double m = mass[v];
    const vec4_t<T> p = x[v];
    const double r[3] = {(double)p.x - com[0], (double)p.y - com[1], (double)p.z - com[2]};
    const double vd[3] = {(double)rv[3 * v] - cdot[0], (double)rv[3 * v + 1] - cdot[1],
                          (double)rv[3 * v + 2] - cdot[2]};
    // -m * (rhat rhat) = m (|r|^2 I - r r^T)
    const double rh[9] = {0, r[2], -r[1], -r[2], 0, r[0], r[1], -r[0], 0};
    double rr[9];
    matmul3(rh, rh, rr);
#pragma unroll
    for (int j = 0; j < 9; ++j) b[j] += (-m) * rr[j];
    double c[3];
    cross3(r, vd, c);
    b[9] += m * c[0];
    b[10] += m * c[1];
    b[11] += m * c[2];
    ri[3 * v] = r[0];
    ri[3 * v + 1] = r[1];
    ri[3 * v + 2] = r[2];
  }
  double s12[12];
  block_sum<12>(b, sh, s12);
  if (!chunk_publish<12>(s12, part, cnt, i, k, nch, &sLast) || threadIdx.x >= 64) return;
  double t[12];
  chunk_total<12>(part, cnt, c0, nch, k, t);
  if (threadIdx.x != 0) return;
  Body &B = bodies[k];
  for (int j = 0; j < 9; ++j) {
    B.I0[j] = t[j];
    B.I[j] = t[j];
    B.R[j] = (j % 4 == 0) ? 1.0 : 0.0;
    B.Rstep[j] = B.R[j];
  }
  solve3_pivot(t, t + 9, B.omega);
}

// rigidSymplecticImperativeStep (RigidBody3D.m:174-179).
__global__ void k_body_step(Body *bodies, int nb, double h) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  Body &B = bodies[k];
  for (int d = 0; d < 3; ++d) B.COM[d] = B.COM[d] + h * B.COMDOT[d];
  double dR[9], Rn[9], tmp[9], Rt[9];
  exp_rodrigues(B.omega, h, dR);
  matmul3(dR, B.R, Rn);
  for (int i = 0; i < 9; ++i) B.R[i] = Rn[i];
  matmul3(B.R, B.I0, tmp);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Rt[i + 3 * j] = B.R[j + 3 * i];
  matmul3(tmp, Rt, B.I);
  for (int i = 0; i < 9; ++i) B.Rstep[i] = B.R[i];
}

// Layer start on the vertices (xpbdLayer3D.m:138-158): elastic vertices are
// released (x += h v_res, v_res = 0); when the layer has rigid elements,
// rigid vertices move with their stepped body and their residual velocity
// becomes R v_res - (x_new - x_prev)/h.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_layer_vertices(int64_t nV, const int32_t *vbody,
                                                           vec4_t<T> *x, T *rv,
                                                           const Body *bodies,
                                                           const double *ri, double h,
                                                           int stepRigid) {
  const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (v >= nV) return;
  const int b = vbody[v];
  vec4_t<T> p = x[v];
  if (b == 0) {
    p.x = p.x + T(h) * rv[3 * v];
    p.y = p.y + T(h) * rv[3 * v + 1];
    p.z = p.z + T(h) * rv[3 * v + 2];
    rv[3 * v] = T(0);
    rv[3 * v + 1] = T(0);
    rv[3 * v + 2] = T(0);
    x[v] = p;
    return;
  }
  if (!stepRigid) return;
  const Body &B = bodies[b - 1];
  double xn[3];
  rigid_pos(B.R, ri + 3 * v, B.COM, xn);
  const double prev[3] = {(double)p.x, (double)p.y, (double)p.z};
  p.x = (T)xn[0];
  p.y = (T)xn[1];
  p.z = (T)xn[2];
  x[v] = p;
  const double vr[3] = {(double)rv[3 * v], (double)rv[3 * v + 1], (double)rv[3 * v + 2]};
  const double zero[3] = {0, 0, 0};
  double rot[3];
  rigid_pos(B.R, vr, zero, rot);
  rv[3 * v] = (T)(rot[0] - ((double)p.x - prev[0]) / h);
  rv[3 * v + 1] = (T)(rot[1] - ((double)p.y - prev[1]) / h);
  rv[3 * v + 2] = (T)(rot[2] - ((double)p.z - prev[2]) / h);
}

// rotateAllVerts (xpbdLayer3D.m:508-513): rigid vertices to R r_i + COM.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rotate_all(int64_t nV, const int32_t *vbody,
                                                       vec4_t<T> *x, const Body *bodies,
                                                       const double *ri) {
  const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (v >= nV) return;
  const int b = vbody[v];
  if (b == 0) return;
  const Body &B = bodies[b - 1];
  double xn[3];
  rigid_pos(B.R, ri + 3 * v, B.COM, xn);
  vec4_t<T> p = x[v];
  p.x = (T)xn[0];
  p.y = (T)xn[1];
  p.z = (T)xn[2];
  x[v] = p;
}

// rotateResidual with deltaR = R_projected R_stepped^T (xpbdLayer3D.m:217-241).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rotate_residual(int64_t nV, const int32_t *vbody,
                                                            T *rv, const Body *bodies) {
  const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (v >= nV) return;
  const int b = vbody[v];
  if (b == 0) return;
  const Body &B = bodies[b - 1];
  double Rt[9], dR[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Rt[i + 3 * j] = B.Rstep[j + 3 * i];
  matmul3(B.R, Rt, dR);
  const double vr[3] = {(double)rv[3 * v], (double)rv[3 * v + 1], (double)rv[3 * v + 2]};
  const double zero[3] = {0, 0, 0};
  double o[3];
  rigid_pos(dR, vr, zero, o);
  rv[3 * v] = (T)o[0];
  rv[3 * v + 1] = (T)o[1];
  rv[3 * v + 2] = (T)o[2];
}

// Apply accumulated dCOM / dOmega to body k (xpbdLayer3D.m:400-414):
// COM += dCOM; w = I \ dOmega; R = exp(h w) R; I = R I0 R^T.
__device__ __forceinline__ void body_apply(Body &B, const double *dCOM, const double *dOm,
                                           double h) {
  for (int d = 0; d < 3; ++d) B.COM[d] = B.COM[d] + dCOM[d];
  double w[3], dR[9], Rn[9];
  solve3_pivot(B.I, dOm, w);
  exp_rodrigues(w, h, dR);
  matmul3(dR, B.R, Rn);
  for (int i = 0; i < 9; ++i) B.R[i] = Rn[i];
  inertia_update(B.R, B.I0, B.I);
}

// rigidBoundaryConstraint.project (rigidBoundaryConstraint.m:4-35) for
// boundary entry q (vertex v) of body B: updates its multiplier and adds the
// impulse's dCOM and dOmega terms to acc.
// boundary_project on the vertex's position p, loaded by the caller.
template <typename T>
__device__ __forceinline__ void boundary_project_at(const Body &B, int v, const vec4_t<T> &p,
                                                    const double *minv, const double *ri,
                                                    double *lamq, double compliance, double h,
                                                    double (&acc)[6]) {
  double rx[3], r[3], dir[3], nn[3];
  rigid_pos(B.R, ri + 3 * v, B.COM, rx);
  const double xe[3] = {(double)p.x, (double)p.y, (double)p.z};
  for (int d = 0; d < 3; ++d) {
    r[d] = rx[d] - B.COM[d];
    dir[d] = rx[d] - xe[d];
  }
  const double C = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  for (int d = 0; d < 3; ++d) {
    nn[d] = dir[d] / C;
    if (nn[d] != nn[d]) nn[d] = 0;
  }
  const double w = gen_inv_mass(B.invM, r, B.I, nn);
  const double dl = (-C - compliance * *lamq) / ((minv[v] + w) + compliance);
  *lamq += dl;
  const double pp[3] = {dl * nn[0], dl * nn[1], dl * nn[2]};
  const double ph[3] = {pp[0] / h, pp[1] / h, pp[2] / h};
  double dom[3];
  cross3(r, ph, dom);
  for (int d = 0; d < 3; ++d) {
    acc[d] += pp[d] * B.invM;
    acc[3 + d] += dom[d];
  }
}

template <typename T>
__device__ __forceinline__ void boundary_project(const Body &B, int v, const vec4_t<T> *x,
                                                 const double *minv, const double *ri,
                                                 double *lamq, double compliance, double h,
                                                 double (&acc)[6]) {
  boundary_project_at<T>(B, v, x[v], minv, ri, lamq, compliance, h, acc);
}

template <typename T>
__device__ __forceinline__ void boundary_snap(const double *R, const double *COM, int v,
                                              const double *ri, vec4_t<T> *x) {
  double xn[3];
  rigid_pos(R, ri + 3 * v, COM, xn);
  vec4_t<T> p = x[v];
  p.x = (T)xn[0];
  p.y = (T)xn[1];
  p.z = (T)xn[2];
  x[v] = p;
}

// rigidElasticBoundaryAccumulationConstraints (xpbdLayer3D.m:381-420) with
// rigidBoundaryConstraint.project (rigidBoundaryConstraint.m:4-35).  One
// workgroup per body: project its boundary vertices, reduce, apply, snap.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_boundary(LayerLists L, vec4_t<T> *x,
                                                     const double *minv, Body *bodies,
                                                     const double *ri, double *lamB,
                                                     double compliance, double h,
                                                     const Ctl *ctl) {
  if (iteration_stopped(ctl)) return;
  __shared__ double sh[6 * kBlock];
  const int k = blockIdx.x;
  Body &B = bodies[k];
  const int off = L.bnd_off[k], n = L.bnd_off[k + 1] - off;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int q = threadIdx.x; q < n; q += kBlock)
    boundary_project<T>(B, L.bnd_vert[off + q], x, minv, ri, lamB + off + q, compliance, h, acc);
  double s6[6];
  block_sum<6>(acc, sh, s6);
  __shared__ double sR[9], sC[3];
  if (threadIdx.x == 0) {
    body_apply(B, s6, s6 + 3, h);
    for (int i = 0; i < 9; ++i) sR[i] = B.R[i];
    for (int i = 0; i < 3; ++i) sC[i] = B.COM[i];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += kBlock)
    boundary_snap<T>(sR, sC, L.bnd_vert[off + q], ri, x);
}

// The same accumulation with each body's boundary vertices split in chunks
// of kBoundaryChunk, one workgroup per chunk (every body has at least one
// chunk, so a body without boundary vertices is still applied, with zero
// sums, as above).  Each chunk projects its vertices against the body's
// pose at the start and publishes its sums; the body's last chunk adds them
// in chunk order and applies them.  k_boundary_snap then moves every
// boundary vertex to its body's new pose.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_boundary_chunk(LayerLists L, const vec4_t<T> *x,
                                                           const double *minv, Body *bodies,
                                                           const double *ri, double *lamB,
                                                           double compliance, double h,
                                                           const Ctl *ctl, double *part,
                                                           unsigned *cnt) {
  if (iteration_stopped(ctl)) return;
  __shared__ double sh[6 * kBlock];
  __shared__ int sLast;
  const int i = blockIdx.x;
  const int k = body_of_chunk(L.bchk_off, L.nb, i);
  const int c0 = L.bchk_off[k], nch = L.bchk_off[k + 1] - c0;
  const int off = L.bnd_off[k], n = L.bnd_off[k + 1] - off;
  const int q0 = (i - c0) * kBoundaryChunk, q1 = min(n, q0 + kBoundaryChunk);
  const Body &Bk = bodies[k];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int q = q0 + threadIdx.x; q < q1; q += kBlock)
    boundary_project<T>(Bk, L.bnd_vert[off + q], x, minv, ri, lamB + off + q, compliance, h, acc);
  double s6[6];
  block_sum<6>(acc, sh, s6);
  if (!chunk_publish<6>(s6, part, cnt, i, k, nch, &sLast) || threadIdx.x >= 64) return;
  double t[6];
  chunk_total<6>(part, cnt, c0, nch, k, t);
  if (threadIdx.x != 0) return;
  body_apply(bodies[k], t, t + 3, h);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_boundary_snap(LayerLists L, vec4_t<T> *x,
                                                          const Body *bodies, const double *ri,
                                                          const Ctl *ctl) {
  if (iteration_stopped(ctl)) return;
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= L.nbnd) return;
  const int k = body_of_chunk(L.bnd_off, L.nb, q);
  const Body &B = bodies[k];
  boundary_snap<T>(B.R, B.COM, L.bnd_vert[q], ri, x);
}

// ---------------------------------------------------------------------------
// rigid layer, 2D
// ---------------------------------------------------------------------------

// RigidBody.pageRelativeVertexPosition with angle2rotm2D (RigidBody.m:166-169,
// 2d/tools/math/angle2rotm2D.m:1-5): pos = [c -s; s c] r_i + COM.
__device__ __forceinline__ void rigid_pos2d(double angle, const double *ri, const double *COM,
                                            double *out) {
  const double c = cos(angle), s = sin(angle);
  out[0] = (c * ri[0] + (-s) * ri[1]) + COM[0];
  out[1] = (s * ri[0] + c * ri[1]) + COM[1];
}

// makeRigidsFromConnectivity / matComputeRigidBodyProperties2D
// (AdaptiveMesh.m:27-37, RigidBody.m:171-215): mass, COM = sum(m x)/M,
// scalar inertia sum(m |x - COM|^2), angle = 0, r_i = x - COM (stride 3).
// The 2D layerSolve does not step bodies, so velocities are not formed.
// One workgroup per body, fixed reduction tree.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_restrict2d(LayerLists L, const vec4_t<T> *x,
                                                       const double *mass, Body *bodies,
                                                       double *ri) {
  __shared__ double sh[3 * kBlock];
  const int k = blockIdx.x;
  const int off = L.rig_off[k], n = L.rig_off[k + 1] - off;
  double a[3] = {0, 0, 0};
  for (int q = threadIdx.x; q < n; q += kBlock) {
    const int v = L.rig_vert[off + q];
    const double m = mass[v];
    const vec4_t<T> p = x[v];
    a[0] += m;
    a[1] += (double)p.x * m;
    a[2] += (double)p.y * m;
  }
  double s3[3];
  block_sum<3>(a, sh, s3);
  const double M = s3[0];
  const double com[2] = {s3[1] / M, s3[2] / M};
  double b[1] = {0};
  for (int q = threadIdx.x; q < n; q += kBlock) {
    const int v = L.rig_vert[off + q];
    const vec4_t<T> p = x[v];
    const double r0 = (double)p.x - com[0], r1 = (double)p.y - com[1];
    b[0] += mass[v] * (r0 * r0 + r1 * r1);
    ri[3 * v] = r0;
    ri[3 * v + 1] = r1;
    ri[3 * v + 2] = 0;
  }
  double s1[1];
  block_sum<1>(b, sh, s1);
  if (threadIdx.x == 0) {
    Body &B = bodies[k];
    B.M = M;
    B.invM = 1.0 / M;
    B.COM[0] = com[0];
    B.COM[1] = com[1];
    B.COM[2] = 0;
    B.angle = 0;
    B.Iz = s1[0];
  }
}

// rigidElasticBoundaryAccumulationConstraints (xpbdLayer2D.m:296-330) with
// rigidBoundaryConstraint2D.projectVectorized (rigidBoundaryConstraint2D.m:4-34).
// One workgroup per body: each boundary vertex against the body's pose at
// the start, w = (1/M)/nb + Iinv (r x n)^2 with nb the body's boundary-vertex
// count, alpha~ = alpha/h^2; sums of p/M and (Iinv/2) r x p are applied and
// the boundary vertices snap to the rigid positions.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_boundary2d(LayerLists L, vec4_t<T> *x,
                                                       const double *minv, Body *bodies,
                                                       const double *ri, double *lamB,
                                                       double aT, const Ctl *ctl) {
  if (iteration_stopped(ctl)) return;
  __shared__ double sh[3 * kBlock];
  const int k = blockIdx.x;
  Body &B = bodies[k];
  const int off = L.bnd_off[k], n = L.bnd_off[k + 1] - off;
  const double Iinv = 1.0 / B.Iz, invMs = B.invM / n;
  double acc[3] = {0, 0, 0};
  for (int q = threadIdx.x; q < n; q += kBlock) {
    const int v = L.bnd_vert[off + q];
    double rx[2];
    rigid_pos2d(B.angle, ri + 3 * v, B.COM, rx);
    const vec4_t<T> p = x[v];
    const double r[2] = {rx[0] - B.COM[0], rx[1] - B.COM[1]};
    const double dir[2] = {rx[0] - (double)p.x, rx[1] - (double)p.y};
    const double C = sqrt(dir[0] * dir[0] + dir[1] * dir[1]);
    double nn[2] = {dir[0] / C, dir[1] / C};
    if (C == 0) nn[0] = nn[1] = 0;
    const double crn = r[0] * nn[1] - r[1] * nn[0];
    const double w = invMs + Iinv * (crn * crn);
    const double dl = (-C - aT * lamB[off + q]) / ((minv[v] + w) + aT);
    lamB[off + q] += dl;
    const double pp[2] = {dl * nn[0], dl * nn[1]};
    acc[0] += pp[0] * B.invM;
    acc[1] += pp[1] * B.invM;
    acc[2] += (0.5 * Iinv) * (r[0] * pp[1] - r[1] * pp[0]);
  }
  double s3[3];
  block_sum<3>(acc, sh, s3);
  __shared__ double sA, sC[3];
  if (threadIdx.x == 0) {
    B.COM[0] = B.COM[0] + s3[0];
    B.COM[1] = B.COM[1] + s3[1];
    B.angle = B.angle + s3[2];
    sA = B.angle;
    sC[0] = B.COM[0];
    sC[1] = B.COM[1];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += kBlock) {
    const int v = L.bnd_vert[off + q];
    double xn[2];
    rigid_pos2d(sA, ri + 3 * v, sC, xn);
    vec4_t<T> p = x[v];
    p.x = (T)xn[0];
    p.y = (T)xn[1];
    x[v] = p;
  }
}

// rotateAllVerts2D.m:1-7: rigid vertices to R(angle) r_i + COM.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rotate_all2d(int64_t nV, const int32_t *vbody,
                                                         vec4_t<T> *x, const Body *bodies,
                                                         const double *ri) {
  const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (v >= nV) return;
  const int b = vbody[v];
  if (b == 0) return;
  const Body &B = bodies[b - 1];
  double xn[2];
  rigid_pos2d(B.angle, ri + 3 * v, B.COM, xn);
  vec4_t<T> p = x[v];
  p.x = (T)xn[0];
  p.y = (T)xn[1];
  x[v] = p;
}

// ---------------------------------------------------------------------------
// contacts
// ---------------------------------------------------------------------------

// Contact arrays at fixed offsets (capacity = vertex count: one contact per
// vertex at most), so kernels captured in a graph stay valid when the
// number of contacts changes.  counts = {K, elastic contacts of the layer,
// rigid contacts of the layer, distinct compliance values}, on the device.
struct ContactArrays {
  const int32_t *vert;      // K, 0-based
  const double *normal;     // 3K (xyz per contact)
  const double *point;      // 3K
  const double *compliance; // K
  double *lambda;           // K
  const int32_t *counts;    // device {K, nec, nrc, nUniq}
  int K;                    // host copy of counts[0] (kernels outside graphs)
};

// penaltyConstraint.projectElastic (penaltyConstraint.m:4-28), 3D;
// penaltyConstraint2D.projectElastic (penaltyConstraint2D.m:4-26) in 2D:
// contact i on an elastic vertex.
template <typename T, int DIM, bool COH = false>
__device__ __forceinline__ void contact_elastic_one(const ContactArrays &c, int i, vec4_t<T> *x,
                                                    const double *minv, double resistance,
                                                    double h) {
  const int v = c.vert[i];
  const double *n = c.normal + 3 * i, *pc = c.point + 3 * i;
  vec4_t<T> p = ldc<COH>(x + v);
  const double lam_i = ldc<COH>(c.lambda + i);
  const double xv[3] = {(double)p.x, (double)p.y, (double)p.z};
  double phi = n[0] * (xv[0] - pc[0]) + n[1] * (xv[1] - pc[1]);
  if (DIM == 3) phi = phi + n[2] * (xv[2] - pc[2]);
  if (phi > 0) phi = 0;
  const double C = resistance * phi;
  const double aT = c.compliance[i] / (h * h);
  const double mi = minv[v];
  double dl;
  if (DIM == 3) {
    dl = (-C - aT * lam_i) / (2 * mi + aT);
    p.x = (T)(xv[0] + mi * (dl * n[0]));
    p.y = (T)(xv[1] + mi * (dl * n[1]));
    p.z = (T)(xv[2] + mi * (dl * n[2]));
  } else {
    const double g0 = n[0] * 2 * mi, g1 = n[1] * 2 * mi;
    dl = -(C + aT * lam_i) / ((g0 * n[0] + g1 * n[1]) + aT);
    p.x = (T)(xv[0] + g0 * dl);
    p.y = (T)(xv[1] + g1 * dl);
  }
  c.lambda[i] = lam_i + dl;
  x[v] = p;
}

template <typename T, int DIM>
__global__ __launch_bounds__(kBlock) void k_contacts_elastic(LayerLists L, ContactArrays c,
                                                             vec4_t<T> *x, const double *minv,
                                                             double resistance, double h,
                                                             const Ctl *ctl) {
  if (iteration_stopped(ctl)) return;
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= c.counts[1]) return;
  contact_elastic_one<T, DIM>(c, L.ec_idx[q], x, minv, resistance, h);
}

// penaltyConstraint.projectRigid (penaltyConstraint.m:30-63) with the
// per-body accumulation of contactConstraints (xpbdLayer3D.m:445-484).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_contacts_rigid(LayerLists L, ContactArrays c,
                                                           vec4_t<T> *x, Body *bodies,
                                                           const double *ri,
                                                           double resistance, double h,
                                                           const Ctl *ctl) {
  if (iteration_stopped(ctl)) return;
  __shared__ double sh[6 * kBlock];
  const int k = blockIdx.x;
  Body &B = bodies[k];
  const int off = L.rc_off[k], n = L.rc_off[k + 1] - off;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int q = threadIdx.x; q < n; q += kBlock) {
    const int i = L.rc_idx[off + q];
    const int v = c.vert[i];
    const double *nn = c.normal + 3 * i, *pc = c.point + 3 * i;
    double rx[3], r[3];
    rigid_pos(B.R, ri + 3 * v, B.COM, rx);
    for (int d = 0; d < 3; ++d) r[d] = rx[d] - B.COM[d];
    const vec4_t<T> p = x[v];
    double phi = nn[0] * ((double)p.x - pc[0]) + nn[1] * ((double)p.y - pc[1]) +
                 nn[2] * ((double)p.z - pc[2]);
    if (phi > 0) phi = 0;
    const double C = resistance * phi;
    const double w = gen_inv_mass(B.invM, r, B.I, nn);
    const double aT = (2 * c.compliance[i]) / (h * h);
    const double dl = (-C - aT * c.lambda[i]) / (2 * w + aT);
    c.lambda[i] += dl;
    const double pp[3] = {dl * nn[0], dl * nn[1], dl * nn[2]};
    double cr[3];
    cross3(r, pp, cr);
    for (int d = 0; d < 3; ++d) {
      acc[d] += pp[d] * B.invM;
      acc[3 + d] += (1 / h) * cr[d];
    }
  }
  double s6[6];
  block_sum<6>(acc, sh, s6);
  __shared__ double sR[9], sC[3];
  if (threadIdx.x == 0) {
    body_apply(B, s6, s6 + 3, h);
    for (int i = 0; i < 9; ++i) sR[i] = B.R[i];
    for (int i = 0; i < 3; ++i) sC[i] = B.COM[i];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += kBlock) {
    const int v = c.vert[L.rc_idx[off + q]];
    double xn[3];
    rigid_pos(sR, ri + 3 * v, sC, xn);
    vec4_t<T> p = x[v];
    p.x = (T)xn[0];
    p.y = (T)xn[1];
    p.z = (T)xn[2];
    x[v] = p;
  }
}

// penaltyConstraint2D.projectRigid (penaltyConstraint2D.m:28-58) with the
// per-body accumulation of the 2D contactConstraints (xpbdLayer2D.m:418-444):
// nb = the body's rigid-contact count, w = (1/M)/nb + Iinv (r x n)^2,
// dlambda = (-C - alpha~ lambda) / (2 w + alpha~); sums of p/M and
// (Iinv/2) r x p applied; the contact vertices snap to the rigid positions.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_contacts_rigid2d(LayerLists L, ContactArrays c,
                                                             vec4_t<T> *x, Body *bodies,
                                                             const double *ri, double resistance,
                                                             double h, const Ctl *ctl) {
  if (iteration_stopped(ctl)) return;
  __shared__ double sh[3 * kBlock];
  const int k = blockIdx.x;
  Body &B = bodies[k];
  const int off = L.rc_off[k], n = L.rc_off[k + 1] - off;
  const double Iinv = 1.0 / B.Iz, invMs = B.invM / n;
  double acc[3] = {0, 0, 0};
  for (int q = threadIdx.x; q < n; q += kBlock) {
    const int i = L.rc_idx[off + q];
    const int v = c.vert[i];
    const double *nn = c.normal + 3 * i, *pc = c.point + 3 * i;
    double rx[2];
    rigid_pos2d(B.angle, ri + 3 * v, B.COM, rx);
    const double r[2] = {rx[0] - B.COM[0], rx[1] - B.COM[1]};
    const vec4_t<T> p = x[v];
    double phi = ((double)p.x - pc[0]) * nn[0] + ((double)p.y - pc[1]) * nn[1];
    if (phi > 0) phi = 0;
    const double C = resistance * phi;
    const double crn = r[0] * nn[1] - r[1] * nn[0];
    const double w = invMs + Iinv * (crn * crn);
    const double aT = c.compliance[i] / (h * h);
    const double dl = (-C - aT * c.lambda[i]) / (2 * w + aT);
    c.lambda[i] += dl;
    const double pp[2] = {dl * nn[0], dl * nn[1]};
    acc[0] += pp[0] * B.invM;
    acc[1] += pp[1] * B.invM;
    acc[2] += (0.5 * Iinv) * (r[0] * pp[1] - r[1] * pp[0]);
  }
  double s3[3];
  block_sum<3>(acc, sh, s3);
  __shared__ double sA, sC[3];
  if (threadIdx.x == 0) {
    B.COM[0] = B.COM[0] + s3[0];
    B.COM[1] = B.COM[1] + s3[1];
    B.angle = B.angle + s3[2];
    sA = B.angle;
    sC[0] = B.COM[0];
    sC[1] = B.COM[1];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += kBlock) {
    const int v = c.vert[L.rc_idx[off + q]];
    double xn[2];
    rigid_pos2d(sA, ri + 3 * v, sC, xn);
    vec4_t<T> p = x[v];
    p.x = (T)xn[0];
    p.y = (T)xn[1];
    x[v] = p;
  }
}

// xi(pinnedDOFs) = pinnedx (xpbdLayer3D.m:333; 2D pins to oldp, :187).
template <typename T, int DIM>
__device__ __forceinline__ void pin_one(const int32_t *dof, const T *val, vec4_t<T> *x,
                                        int64_t q) {
  const int d = dof[q];
  const int v = d / DIM, c = d % DIM;
  T *px = reinterpret_cast<T *>(&x[v]);
  px[c] = val[q];
}

template <typename T, int DIM>
__global__ __launch_bounds__(kBlock) void k_pins(int64_t n, const int32_t *dof, const T *val,
                                                 vec4_t<T> *x, const Ctl *ctl) {
  if (iteration_stopped(ctl)) return;
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= n) return;
  pin_one<T, DIM>(dof, val, x, q);
}

// ---------------------------------------------------------------------------
// residual and monitor
// ---------------------------------------------------------------------------

// Residual terms of one tetrahedron (xpbdLayer3D.m:555-568): sum over its six
// constraints of (|C_i| - min(|C_i|, |(alpha~ lambda)_i|))^2, in double.
// Inputs of one tet's residual terms.
template <typename T>
struct ResIn {
  vec4_t<T> X[4];
  T D[9];
  T tad, tao, tas, vol;
  T lam[6];
};

template <typename T, bool COH, int MAT = 0>
__device__ __forceinline__ void res_fetch(const SweepParams<T> &P, const int64_t s,
                                          ResIn<T> &in) {
  const int64_t ld = P.ld;
  const T m0 = P.mat[s], m2 = P.mat[2 * ld + s];
  in.tad = m0 * P.ih2;
  in.tao = P.mat[ld + s] * P.ih2;
  in.tas = mat_third<T, MAT>(m0, m2, P.ih2);
  in.vol = P.vol[s];
#pragma unroll
  for (int a = 0; a < 4; ++a) in.X[a] = ldc<COH>(P.x + P.vid[a * ld + s]);
#pragma unroll
  for (int k = 0; k < 9; ++k) in.D[k] = P.dminv[k * ld + s];
#pragma unroll
  for (int i = 0; i < 6; ++i) in.lam[i] = ldc<COH>(P.lam + i * ld + s);
}

// Residual terms of a NeoHookean tetrahedron: (|C| - min(|C|, |a~ l|))^2
// for C_D = ||F|| and C_H = det F - gamma, in double.
template <typename T>
__device__ __forceinline__ double res_compute_nh(const ResIn<T> &in) {
  T D[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) D[k] = in.D[k];
  T F[3][3], PH[3][3], n, J;
  tet_F<T>(in.X, D, F);
  nh_terms<T>(F, n, PH, J);
  const T iv = T(1) / in.vol;
  const double C[2] = {(double)n, (double)(J - in.tas)};
  const double al[2] = {(double)(iv * (-in.tad) * in.lam[0]), (double)(iv * in.tao * in.lam[1])};
  double acc = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double a = fabs(C[i]), cr = fabs(al[i]);
    const double t = a - (a < cr ? a : cr);
    acc += t * t;
  }
  return acc;
}

template <typename T, int MAT = 0>
__device__ __forceinline__ double res_compute(const ResIn<T> &in) {
  if constexpr (MAT == kMatNH) return res_compute_nh<T>(in);
  if constexpr (MAT == kMatMixed)
    if (in.tad < T(0)) return res_compute_nh<T>(in);
  const T tad = in.tad, tao = in.tao, tas = in.tas;
  const T vol = in.vol;
  const vec4_t<T> X0 = in.X[0], X1 = in.X[1], X2 = in.X[2], X3 = in.X[3];
  T D[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) D[k] = in.D[k];
  const T e[3][3] = {{X1.x - X0.x, X1.y - X0.y, X1.z - X0.z},
                     {X2.x - X0.x, X2.y - X0.y, X2.z - X0.z},
                     {X3.x - X0.x, X3.y - X0.y, X3.z - X0.z}};
  T F[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      F[r][c] = e[0][r] * D[3 * c] + e[1][r] * D[1 + 3 * c] + e[2][r] * D[2 + 3 * c];
  T Cg[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) Cg[j][k] = F[0][j] * F[0][k] + F[1][j] * F[1][k] + F[2][j] * F[2][k];
  T lam[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) lam[i] = in.lam[i];
  const T half = T(0.5);
  T C[6], al[6];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    C[j] = vol * (half * (Cg[j][j] - T(1)));
    const int k = (j + 1) % 3, l = (j + 2) % 3;
    al[j] = tad * lam[j] + tao * lam[k] + tao * lam[l];
  }
  C[3] = vol * (half * Cg[0][1]);
  C[4] = vol * (half * Cg[0][2]);
  C[5] = vol * (half * Cg[1][2]);
  al[3] = tas * lam[3];
  al[4] = tas * lam[4];
  al[5] = tas * lam[5];
  double acc = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const double a = fabs((double)C[i]), cr = fabs((double)al[i]);
    const double t = a - (a < cr ? a : cr);
    acc += t * t;
  }
  return acc;
}

// res_fetch with the positions from pos(vertex id) instead of P.x.
template <typename T, int MAT = 0, typename Pos>
__device__ __forceinline__ void res_fetch_pos(const SweepParams<T> &P, const int64_t s,
                                              ResIn<T> &in, Pos &&pos) {
  const int64_t ld = P.ld;
  const T m0 = P.mat[s], m2 = P.mat[2 * ld + s];
  in.tad = m0 * P.ih2;
  in.tao = P.mat[ld + s] * P.ih2;
  in.tas = mat_third<T, MAT>(m0, m2, P.ih2);
  in.vol = P.vol[s];
#pragma unroll
  for (int a = 0; a < 4; ++a) in.X[a] = pos(P.vid[a * ld + s]);
#pragma unroll
  for (int k = 0; k < 9; ++k) in.D[k] = P.dminv[k * ld + s];
#pragma unroll
  for (int i = 0; i < 6; ++i) in.lam[i] = P.lam[i * ld + s];
}

template <typename T, bool COH = false, int MAT = 0>
__device__ __forceinline__ double residual_tet(const SweepParams<T> &P, const int64_t s) {
  ResIn<T> in;
  res_fetch<T, COH, MAT>(P, s, in);
  return res_compute<T, MAT>(in);
}

// Sum of residual_tet over the real slots s, s + step, ... < e of one
// thread, U slots' inputs loaded before their terms are computed.
constexpr int kResPerThread = 2;  // slots per thread of k_residual_tail (a pair)
template <typename T, bool COH, int MAT = 0, int U = 4>
__device__ __forceinline__ double residual_slots(const SweepParams<T> &P, int64_t s,
                                                 const int64_t e, const int64_t step) {
  double acc = 0;
  for (; s < e; s += U * step) {
    ResIn<T> in[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t su = s + u * step;
      ok[u] = su < e && P.vid[su] >= 0;
      if (ok[u]) res_fetch<T, COH, MAT>(P, su, in[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (ok[u]) acc += res_compute<T, MAT>(in[u]);
  }
  return acc;
}

// Residual terms of slots s and s1 (> s) of one thread, both slots' inputs
// loaded before either term is computed; summed in slot order.
template <typename T, int MAT>
__device__ __forceinline__ double residual_pair(const SweepParams<T> &P, const int64_t s,
                                                const int64_t s1, const int64_t e) {
  const bool oa = s < e && P.vid[s] >= 0, ob = s1 < e && P.vid[s1] >= 0;
  ResIn<T> a, b;
  if (oa) res_fetch<T, false, MAT>(P, s, a);
  if (ob) res_fetch<T, false, MAT>(P, s1, b);
  double acc = 0;
  if (oa) acc += res_compute<T, MAT>(a);
  if (ob) acc += res_compute<T, MAT>(b);
  return acc;
}

// Element part of simpleEvaluateAllConstraintResiduals (xpbdLayer3D.m:546-568;
// 2D xpbdLayer2D.m:477-496): per element sum_i (|C_i| - min(|C_i|,
// |(alpha~ lambda)_i|))^2 in double.  lamSlot2d gives, for the 2D quirk, the
// slot whose lambda column feeds each page entry.
template <typename T, int DIM, bool COH = false>
__device__ __forceinline__ double residual_elem(const SweepParams<T> &P, int64_t s,
                                                const int32_t *lamSlot2d) {
  if (DIM == 3) {
    if (P.nh) return residual_tet<T, COH, kMatMixed>(P, s);
    return residual_tet<T, COH>(P, s);
  }
  static_assert(DIM == 3 || !COH, "coherent loads are for 3D only");
  const int64_t ld = P.ld;
  const T tad = P.mat[s] * P.ih2, tao = P.mat[ld + s] * P.ih2, tas = P.mat[2 * ld + s] * P.ih2;
  const T vol = P.vol[s];
  const vec4_t<T> X0 = P.x[P.vid[s]], X1 = P.x[P.vid[ld + s]], X2 = P.x[P.vid[2 * ld + s]];
  const T D[4] = {P.dminv[s], P.dminv[ld + s], P.dminv[2 * ld + s], P.dminv[3 * ld + s]};
  const T e[2][2] = {{X0.x - X2.x, X0.y - X2.y}, {X1.x - X2.x, X1.y - X2.y}};
  T F[2][2];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) F[r][c] = e[0][r] * D[2 * c] + e[1][r] * D[1 + 2 * c];
  const T c00 = F[0][0] * F[0][0] + F[1][0] * F[1][0];
  const T c11 = F[0][1] * F[0][1] + F[1][1] * F[1][1];
  const T c01 = F[0][0] * F[0][1] + F[1][0] * F[1][1];
  const T half = T(0.5);
  const T C[3] = {vol * (half * (c00 - T(1))), vol * (half * (c11 - T(1))), vol * (half * c01)};
  // lambda page e mixes elements: entry r is lambdai(L mod nT, L div nT)
  // with L = 3e + r (reshape of the nT x 3 array, xpbdLayer2D.m:489).
  T lp[3];
  for (int r = 0; r < 3; ++r) {
    const int32_t code = lamSlot2d[3 * s + r];  // slot*3 + component
    lp[r] = P.lam[(code % 3) * ld + code / 3];
  }
  const T al[3] = {tad * lp[0] + tao * lp[1], tao * lp[0] + tad * lp[1], tas * lp[2]};
  double acc = 0;
  for (int i = 0; i < 3; ++i) {
    const double a = fabs((double)C[i]), cr = fabs((double)al[i]);
    const double t = a - (a < cr ? a : cr);
    acc += t * t;
  }
  return acc;
}

// Contact part of the residual (xpbdLayer3D.m:573-583).  In 3D the
// compliance enters as a 1x1xK page array times lambda(:), which Matlab
// expands to K x 1 x K: every contact i contributes once per compliance
// value j.  uniqAlpha/uniqCount list the distinct compliance values with
// multiplicities (3D); in 2D each contact counts once with its own value.
template <typename T, int DIM, bool COH = false>
__device__ __forceinline__ double residual_contact_one(const ContactArrays &c, int i,
                                                       const vec4_t<T> *x, double resistance,
                                                       double h, const double *uniqAlpha,
                                                       const double *uniqCount, int nUniq) {
  const int v = c.vert[i];
  const double *n = c.normal + 3 * i, *pc = c.point + 3 * i;
  const vec4_t<T> p = ldc<COH>(x + v);
  const double lam_i = ldc<COH>(c.lambda + i);
  double phi = n[0] * ((double)p.x - pc[0]) + n[1] * ((double)p.y - pc[1]);
  if (DIM == 3) phi = phi + n[2] * ((double)p.z - pc[2]);
  if (phi > 0) phi = 0;
  const double a = fabs(resistance * phi);
  const double h2 = h * h;
  double acc = 0;
  if (DIM == 3) {
    for (int u = 0; u < nUniq; ++u) {
      const double cr = fabs((uniqAlpha[u] / h2) * lam_i);
      const double t = a - (a < cr ? a : cr);
      acc += uniqCount[u] * (t * t);
    }
  } else {
    const double cr = fabs((c.compliance[i] / h2) * lam_i);
    const double t = a - (a < cr ? a : cr);
    acc += t * t;
  }
  return acc;
}

// simpleMonitorError (xpbdLayer3D.m:515-544; 2D xpbdLayer2D.m:520-549) on
// one scene's control block, given the new residual.
__device__ __forceinline__ void monitor_update(Ctl &c, const MonitorParams &mp, double newRes,
                                               double *resid, int residCap) {
  const int idx = c.residualCount;
  if (resid && idx < residCap) resid[idx] = newRes;
  c.residualCount = idx + 1;
  const double last = fabs(c.lastResidual);
  c.lastResidual = newRes;
  const double nr = fabs(newRes);
  const double change = nr - last;
  if (c.setPercentStop > 0) {
    double ru = (1 - c.setPercentStop / 100.0) * nr;
    if (mp.is3d) ru = ru > 1e-8 ? ru : 1e-8;
    c.runUntil = ru;
    c.setPercentStop = 0;
  }
  if (change < 0) c.errorChangeDecreased = 1;
  if (mp.giveUpEnabled && change > mp.giveUpThreshold && c.errorChangeDecreased &&
      (!c.isLastLayer || mp.lastLayerGiveUpEnabled))
    c.endLayer = 1;
  const int stopHanging = c.residualCount > mp.hangingStop && mp.hangingStop >= 0;
  int stop;
  if (mp.is3d)
    stop = stopHanging || (c.runUntil >= 0 && c.runUntil >= nr &&
                           (c.isLastLayer || c.normresv < c.runUntil));
  else
    stop = (c.runUntil >= 0 && c.runUntil >= nr && (c.isLastLayer || c.normresv < mp.h * mp.h)) ||
           stopHanging;
  if (stop) {
    c.stopSolve = 1;
    c.stopLayer = c.layer;
  }
}

// The residual (simpleEvaluateAllConstraintResiduals, xpbdLayer3D.m:546-586;
// 2D xpbdLayer2D.m:477-518), the monitor and the iteration count of one
// projectGS iteration in one launch: blocks [0, nE) hold the element terms
// of their slots (kResPerThread slots per thread), the
// rest the contact terms; each block stores its partial past the caches
// (device-scope store), waits for it, and counts itself on *done.  The
// block that counts last reads every partial the same way (thread t takes
// partials t, t + 256, ...), sums them in a fixed order (wg_sum), runs the
// monitor and re-arms *done to 0 (so a captured graph can replay the
// launch).  Every block returns without writing when the layer has
// stopped, leaving *done at 0.
template <typename T, int DIM>
__global__ __launch_bounds__(kBlock) void k_residual_tail(
    SweepParams<T> P, int64_t nSlots, const int32_t *lamSlot2d, int nE, ContactArrays c,
    double resistance, double h, const double *uniqAlpha, const double *uniqCount,
    double *partials, unsigned *done, Ctl *ctl, MonitorParams mp, double *resid, int residCap) {
  __shared__ double sh[kBlock];
  __shared__ int sLast;
  const int t = threadIdx.x;
  // the control word is loaded beside the terms' inputs and tested once the
  // term is formed (hold), so its latency overlaps theirs; nothing is
  // written before the test
  const int stopWord = ctl->stopSolve | ctl->endLayer;
  double acc = 0;
  if ((int)blockIdx.x < nE) {
    // slots s, s + step (kResPerThread of them), in that order
    const int64_t s = (int64_t)blockIdx.x * kBlock + t, step = (int64_t)nE * kBlock;
    if constexpr (DIM == 3) {
      static_assert(kResPerThread == 2, "k_residual_tail takes slot pairs");
      acc = P.nh ? residual_pair<T, kMatMixed>(P, s, s + step, nSlots)
                 : residual_pair<T, kMatStVK>(P, s, s + step, nSlots);
    } else {
      for (int64_t q = s; q < nSlots; q += step)
        if (P.vid[q] >= 0) acc += residual_elem<T, DIM>(P, q, lamSlot2d);
    }
  } else {
    const int i = ((int)blockIdx.x - nE) * kBlock + t;
    if (i < c.counts[0])
      acc = residual_contact_one<T, DIM>(c, i, P.x, resistance, h, uniqAlpha, uniqCount,
                                         c.counts[3]);
  }
  hold(acc);
  if (stopWord) return;
  const double part = wg_sum<kBlock>(acc, sh);
  if (t == 0) {
    __hip_atomic_store(partials + blockIdx.x, part, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned prev =
        __hip_atomic_fetch_add(done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    sLast = prev == gridDim.x - 1;
  }
  __syncthreads();
  if (!sLast) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  double a2 = 0;
  for (int q = t; q < (int)gridDim.x; q += kBlock)
    a2 += __hip_atomic_load(partials + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const double total = wg_sum<kBlock>(a2, sh);
  if (t != 0) return;
  __hip_atomic_store(done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  ctl->iters += 1;
  monitor_update(*ctl, mp, sqrt(total), resid, residCap);
}

// End of one projectGS iteration: when residuals are monitored, sum the
// partials, append the residual and apply simpleMonitorError
// (xpbdLayer3D.m:515-544; 2D xpbdLayer2D.m:520-549); count the iteration.
__global__ __launch_bounds__(kBlock) void k_iter_end(Ctl *ctl, MonitorParams mp,
                                                     const double *partials, int nPartials,
                                                     double *resid, int residCap) {
  if (iteration_stopped(ctl)) return;
  __shared__ double sh[kBlock];
  double acc = 0;
  if (mp.computeResiduals)
    for (int q = threadIdx.x; q < nPartials; q += kBlock) acc += partials[q];
  double o[1];
  double v1[1] = {acc};
  block_sum<1>(v1, sh, o);
  if (threadIdx.x != 0) return;
  ctl->iters += 1;
  if (!mp.computeResiduals) return;
  monitor_update(*ctl, mp, sqrt(o[0]), resid, residCap);
}

// Sum of squares of n doubles-or-floats into partials (grid-stride, fixed grid).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sumsq(const T *a, int64_t n, double *partials) {
  __shared__ double sh[kBlock];
  double acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const double v = (double)a[i];
    acc += v * v;
  }
  double o[1];
  double v1[1] = {acc};
  block_sum<1>(v1, sh, o);
  if (threadIdx.x == 0) partials[blockIdx.x] = o[0];
}

// ctl->normresv = sqrt(sum(partials)); layer bookkeeping.
__global__ void k_layer_begin(Ctl *ctl, const double *partials, int nPartials, int layer,
                              int isLast, double setPercentStop) {
  __shared__ double sh[kBlock];
  double acc = 0;
  for (int q = threadIdx.x; q < nPartials; q += kBlock) acc += partials[q];
  double o[1];
  double v1[1] = {acc};
  block_sum<1>(v1, sh, o);
  if (threadIdx.x != 0) return;
  ctl->normresv = sqrt(o[0]);
  ctl->endLayer = 0;
  ctl->iters = 0;
  ctl->layer = layer;
  ctl->isLastLayer = isLast;
  ctl->setPercentStop = setPercentStop;
}

// ---------------------------------------------------------------------------
// layer building: strain-rate proxy
// ---------------------------------------------------------------------------

// Fa = B p and Fb = B p_old in the layout of computeB3D.m:31-41 (F row by
// row), summed over the element's vertices in ascending vertex index (the
// column order of the sparse product), then mexEdiffNorm3D.cpp:57-60.  No
// contraction into FMA, so the result is bit-identical to the same
// expression on the host.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_ediff3d(int64_t nT, const int32_t *T0,
                                                    const double *dminv, const vec4_t<T> *pa,
                                                    const vec4_t<T> *pb, double h,
                                                    double *out) {
#pragma clang fp contract(off)
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= nT) return;
  int t[4] = {T0[4 * e], T0[4 * e + 1], T0[4 * e + 2], T0[4 * e + 3]};
  const double *Mi = dminv + 9 * e;
  double Dc[4][3];
  for (int j = 0; j < 3; ++j) {
    Dc[0][j] = -((Mi[3 * j] + Mi[1 + 3 * j]) + Mi[2 + 3 * j]);
    for (int a = 1; a < 4; ++a) Dc[a][j] = Mi[(a - 1) + 3 * j];
  }
  int ord[4] = {0, 1, 2, 3};
  for (int i = 1; i < 4; ++i)
    for (int k = i; k > 0 && t[ord[k - 1]] > t[ord[k]]; --k) {
      int tmp = ord[k];
      ord[k] = ord[k - 1];
      ord[k - 1] = tmp;
    }
  double A[9], B[9];
  for (int d = 0; d < 3; ++d)
    for (int j = 0; j < 3; ++j) {
      double sa = 0.0, sb = 0.0;
      for (int q = 0; q < 4; ++q) {
        const int a = ord[q];
        const vec4_t<T> xa = pa[t[a]], xb = pb[t[a]];
        const double ca = d == 0 ? (double)xa.x : (d == 1 ? (double)xa.y : (double)xa.z);
        const double cb = d == 0 ? (double)xb.x : (d == 1 ? (double)xb.y : (double)xb.z);
        sa = sa + Dc[a][j] * ca;
        sb = sb + Dc[a][j] * cb;
      }
      A[3 * d + j] = sa;
      B[3 * d + j] = sb;
    }
  // mexEdiffNorm3D reads the 9 values column-major
  const double a11 = A[0], a21 = A[1], a31 = A[2], a12 = A[3], a22 = A[4], a32 = A[5],
               a13 = A[6], a23 = A[7], a33 = A[8];
  const double b11 = B[0], b21 = B[1], b31 = B[2], b12 = B[3], b22 = B[4], b32 = B[5],
               b13 = B[6], b23 = B[7], b33 = B[8];
  const double t2 = 1.0 / (h * h);
  double q1 = (a11 * a12) / 2.0 + (a21 * a22) / 2.0 + (a31 * a32) / 2.0 - (b11 * b12) / 2.0 - (b21 * b22) / 2.0 - (b31 * b32) / 2.0;
  double q2 = (a11 * a13) / 2.0 + (a21 * a23) / 2.0 + (a31 * a33) / 2.0 - (b11 * b13) / 2.0 - (b21 * b23) / 2.0 - (b31 * b33) / 2.0;
  double q3 = (a12 * a13) / 2.0 + (a22 * a23) / 2.0 + (a32 * a33) / 2.0 - (b12 * b13) / 2.0 - (b22 * b23) / 2.0 - (b32 * b33) / 2.0;
  double q4 = (a11 * a11) / 2.0 + (a21 * a21) / 2.0 + (a31 * a31) / 2.0 - (b11 * b11) / 2.0 - (b21 * b21) / 2.0 - (b31 * b31) / 2.0;
  double q5 = (a12 * a12) / 2.0 + (a22 * a22) / 2.0 + (a32 * a32) / 2.0 - (b12 * b12) / 2.0 - (b22 * b22) / 2.0 - (b32 * b32) / 2.0;
  double q6 = (a13 * a13) / 2.0 + (a23 * a23) / 2.0 + (a33 * a33) / 2.0 - (b13 * b13) / 2.0 - (b23 * b23) / 2.0 - (b33 * b33) / 2.0;
  out[e] = t2 * (q1 * q1) * 2.0 + t2 * (q2 * q2) * 2.0 + t2 * (q3 * q3) * 2.0 + t2 * (q4 * q4) +
           t2 * (q5 * q5) + t2 * (q6 * q6);
}

// 2D: F = Ds DmInv column-major (Mesh.m:458-467, mult2x2Stack.m), then
// mexEdiffNorm2D.cpp:47-51.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_ediff2d(int64_t nT, const int32_t *T0,
                                                    const double *dminv, const vec4_t<T> *pa,
                                                    const vec4_t<T> *pb, double h,
                                                    double *out) {
#pragma clang fp contract(off)
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= nT) return;
  int t[3] = {T0[3 * e], T0[3 * e + 1], T0[3 * e + 2]};
  const double *Mi = dminv + 4 * e;
  double Dc[3][2];
  for (int j = 0; j < 2; ++j) {
    Dc[0][j] = Mi[2 * j];
    Dc[1][j] = Mi[1 + 2 * j];
    Dc[2][j] = (-1 * Mi[2 * j]) + (-1 * Mi[1 + 2 * j]);
  }
  int ord[3] = {0, 1, 2};
  for (int i = 1; i < 3; ++i)
    for (int k = i; k > 0 && t[ord[k - 1]] > t[ord[k]]; --k) {
      int tmp = ord[k];
      ord[k] = ord[k - 1];
      ord[k - 1] = tmp;
    }
  double A[4], B[4];
  for (int j = 0; j < 2; ++j)
    for (int i = 0; i < 2; ++i) {
      double sa = 0.0, sb = 0.0;
      for (int q = 0; q < 3; ++q) {
        const int a = ord[q];
        const vec4_t<T> xa = pa[t[a]], xb = pb[t[a]];
        sa = sa + Dc[a][j] * (i == 0 ? (double)xa.x : (double)xa.y);
        sb = sb + Dc[a][j] * (i == 0 ? (double)xb.x : (double)xb.y);
      }
      A[i + 2 * j] = sa;
      B[i + 2 * j] = sb;
    }
  const double a = A[0] * A[0] + A[1] * A[1] - B[0] * B[0] - B[1] * B[1];
  const double b = A[0] * A[2] + A[1] * A[3] - B[0] * B[2] - B[1] * B[3];
  const double c = A[2] * A[2] + A[3] * A[3] - B[2] * B[2] - B[3] * B[3];
  out[e] = (0.25 * a * a + 0.5 * b * b + 0.25 * c * c) / h / h;
}

// ---------------------------------------------------------------------------
// integrate() prologue / epilogue
// ---------------------------------------------------------------------------

// frame-1 / NaN reset of cache.oldp and pinnedx (xpbdLayer3D.m:629-632).
template <typename T>
__global__ void k_any_nan(int64_t nV, const vec4_t<T> *xold, int *flag) {
  const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (v >= nV) return;
  const vec4_t<T> p = xold[v];
  if (p.x != p.x || p.y != p.y || p.z != p.z) *flag = 1;
}

// 3D: residualVelocity = v; += h f / m on unpinned DOFs with f = m g on z;
// *= (1 - airDrag) (xpbdLayer3D.m:638-664).
template <typename T>
__global__ void k_residual_velocity3d(int64_t nV, const T *v, T *rv, const double *mass,
                                      const uint8_t *pinnedDof, double h, double g,
                                      double airDrag) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= 3 * nV) return;
  double r = (double)v[i];
  if (!pinnedDof[i]) {
    const double m = mass[i / 3];
    const double f = (i % 3 == 2) ? 0.0 + m * g : 0.0;
    r = r + h * f / m;
  }
  rv[i] = (T)((1 - airDrag) * r);
}

// 2D explicit predictor v += h f/m; x += h v on unpinned DOFs with f = m g on
// y (xpbdLayer2D.m:812-835).
template <typename T>
__global__ void k_predict2d(int64_t nV, T *v, vec4_t<T> *x, const double *mass,
                            const uint8_t *pinnedDof, double h, double g) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nV) return;
  vec4_t<T> p = x[i];
  const double m = mass[i];
  for (int d = 0; d < 2; ++d) {
    if (pinnedDof[2 * i + d]) continue;
    const double f = (d == 1) ? 0.0 + m * g : 0.0;
    const double vn = (double)v[2 * i + d] + h * f / m;
    v[2 * i + d] = (T)vn;
    T *px = reinterpret_cast<T *>(&p);
    px[d] = (T)((double)px[d] + h * (double)(T)vn);
  }
  x[i] = p;
}

// v = (xi - oldp)/h (xpbdLayer3D.m:737; 2D BDF1 :907).
template <typename T, int DIM>
__global__ void k_velocity(int64_t nV, const vec4_t<T> *x, const vec4_t<T> *xold, T *v,
                           double h) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nV) return;
  const vec4_t<T> p = x[i], o = xold[i];
  v[DIM * i] = (T)(((double)p.x - (double)o.x) / h);
  v[DIM * i + 1] = (T)(((double)p.y - (double)o.y) / h);
  if (DIM == 3) v[DIM * i + 2] = (T)(((double)p.z - (double)o.z) / h);
}

// contactVelocityUpdate (xpbdLayer2D.m:577-603), contact i.
template <typename T, int DIM>
__device__ __forceinline__ void contact_velocity_one(const ContactArrays &c, int i, T *v, double h,
                                                     double g, double e) {
  const int id = c.vert[i];
  double vn[3] = {0, 0, 0}, nrm = 0;
  for (int d = 0; d < DIM; ++d) {
    vn[d] = (double)v[DIM * id + d];
    nrm = nrm + vn[d] * vn[d];
  }
  nrm = sqrt(nrm);
  double ei = e;
  if (nrm <= 2 * g * h) ei = 0;
  double vt = 0;
  for (int d = 0; d < DIM; ++d) vt = vt + c.normal[3 * i + d] * vn[d];
  double mn = -ei * vt;
  if (mn > 0) mn = 0;
  for (int d = 0; d < DIM; ++d) v[DIM * id + d] = (T)(vn[d] + c.normal[3 * i + d] * (-vn[d] + mn));
}

template <typename T, int DIM>
__global__ void k_contact_velocity(ContactArrays c, T *v, double h, double g, double e) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= c.K) return;
  contact_velocity_one<T, DIM>(c, i, v, h, g, e);
}

// The same with the contact count read from the device (c.counts[0]); the
// grid covers every possible contact.
template <typename T, int DIM>
__global__ void k_contact_velocity_dev(ContactArrays c, T *v, double h, double g, double e) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= c.counts[0]) return;
  contact_velocity_one<T, DIM>(c, i, v, h, g, e);
}

// Plane contact test (PlaneContactFinder3D.m:40-43; 2D strict).
template <typename T, int DIM>
__global__ void k_plane_flags(int64_t nV, const vec4_t<T> *x, double n0, double n1, double n2,
                              double D, int strict, int32_t *flag) {
  const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (v >= nV) return;
  const vec4_t<T> p = x[v];
  double dist = (double)p.x * n0 + (double)p.y * n1;
  if (DIM == 3) dist = dist + (double)p.z * n2;
  dist = dist - D;
  flag[v] = strict ? (dist < 0) : (dist <= 0);
}

}  // namespace mlx

namespace mlx {

// ---------------------------------------------------------------------------
// batched scenes: one workgroup runs one scene's whole layerSolve
// ---------------------------------------------------------------------------

template <typename T>
struct BatchParams {
  SweepParams<T> P;              // merged element/vertex arrays (ctl unused)
  const int32_t *color0;         // per scene: first entry in colorOff/colorCnt
  const int32_t *ncolor;         // per scene: number of colors
  const int64_t *colorOff;       // per (scene, color): first slot
  const int64_t *colorCnt;       // per (scene, color): elements
  const int64_t *slot0;          // nScenes+1 slot range per scene
  const int64_t *cont0;          // nScenes+1 contact range per scene
  const int64_t *pin0;           // nScenes+1 pinned-DOF range per scene
  const int64_t *vert0;          // nScenes+1 vertex range per scene
  int ldsVerts;                  // scenes with at most this many vertices keep positions in LDS
  const uint8_t *sceneNH;        // per scene: holds NeoHookean tetrahedra (null: none)
  const int32_t *pinDof;         // 0-based DOF
  const T *pinVal;
  ContactArrays c;
  const double *minv;
  MonitorParams mp;
  int iterations;
  double runUntil0, stopPct, contactResistance, contactCompliance, h;
  int hardCap;
  // per-scene outputs
  int32_t *iters;
  double *lastResidual;
  double *runUntilOut;
  int32_t *residualCount;
  const int32_t *sceneList;      // scene of each workgroup (null: workgroup i runs scene i)
};

// Positions of a scene's vertices: in HBM (the vec4 array) or in LDS (x, y,
// z of the scene's nv vertices as structure of arrays, from its first
// vertex vb; the inverse mass, constant, from x_old).  The batched scene
// body with rigid clusters (batch_scene_layered) is written once against
// load / load_w / store / store_comp.
template <typename T>
struct HbmPos {
  static constexpr bool kLds = false;
  vec4_t<T> *x;
  __device__ __forceinline__ vec4_t<T> load(int64_t v) const { return x[v]; }
  // with the inverse mass the caller already holds (ignored here)
  __device__ __forceinline__ vec4_t<T> load_w(int64_t v, T) const { return x[v]; }
  __device__ __forceinline__ void store(int64_t v, const vec4_t<T> &p) const { x[v] = p; }
  __device__ __forceinline__ void store_comp(int64_t v, int c, T val) const {
    reinterpret_cast<T *>(&x[v])[c] = val;
  }
};

template <typename T>
struct LdsPos {
  static constexpr bool kLds = true;
  T *pos;
  int64_t vb;
  int nv;
  const vec4_t<T> *xold;
  __device__ __forceinline__ vec4_t<T> load(int64_t v) const {
    const int l = (int)(v - vb);
    vec4_t<T> p;
    p.x = pos[l];
    p.y = pos[nv + l];
    p.z = pos[2 * nv + l];
    p.w = xold[v].w;
    return p;
  }
  __device__ __forceinline__ vec4_t<T> load_w(int64_t v, T w) const {
    const int l = (int)(v - vb);
    vec4_t<T> p;
    p.x = pos[l];
    p.y = pos[nv + l];
    p.z = pos[2 * nv + l];
    p.w = w;
    return p;
  }
  __device__ __forceinline__ void store(int64_t v, const vec4_t<T> &p) const {
    const int l = (int)(v - vb);
    pos[l] = p.x;
    pos[nv + l] = p.y;
    pos[2 * nv + l] = p.z;
  }
  __device__ __forceinline__ void store_comp(int64_t v, int c, T val) const {
    pos[c * nv + (int)(v - vb)] = val;
  }
};

// Tets one thread of a scene's workgroup has in flight together: a
// workgroup runs one scene on one CU, so with a single tet per thread its
// waves wait out the memory latency of every pass.  In FP32 up to 512
// threads (256 registers) the register file holds two tets' inputs per
// thread; at 768 and 1,024 threads (168 and 128 registers), and in FP64,
// one.
template <typename T, int NT>
constexpr bool kScenePairs = sizeof(T) == 4 && NT <= 512;

// One tet of a scene: slot data, then the current positions.
template <typename T, int MAT, typename Pos>
__device__ __forceinline__ void scene_fetch(const SweepParams<T> &P, const int64_t s,
                                            const Pos &x, TetIn<T> &in, vec4_t<T> (&X)[4]) {
  tet_fetch<T, MAT>(P, s, in);
#pragma unroll
  for (int q = 0; q < 4; ++q) X[q] = x.load_w(in.v[q], in.O[q].w);
}

template <typename T, int MAT, typename Pos>
__device__ __forceinline__ void scene_solve(const SweepParams<T> &P, const int64_t s,
                                            const Pos &x, const TetIn<T> &in,
                                            const vec4_t<T> (&X)[4]) {
  tet_solve_st<T, MAT>(P, s, in, X, [&](int q, const vec4_t<T> &o) { x.store(in.v[q], o); });
}

// One thread's slots s, s + step, ... < e of one color of a scene.  PAIRS:
// two at a time, both tets' slot data and current positions loaded before
// either is projected.  The tets of a color share no vertex, so no other
// projection of the color writes these positions and the two projections
// touch disjoint vertices: the same updates as one tet at a time.  With
// P.flags set, slots whose flag is clear are skipped.
template <typename T, int MAT, bool PAIRS, typename Pos>
__device__ __forceinline__ void scene_color(const SweepParams<T> &P, int64_t s, const int64_t e,
                                            const int64_t step, const Pos &x) {
  if constexpr (!PAIRS) {
    for (; s < e; s += step) {
      if (P.flags && !P.flags[s]) continue;
      TetIn<T> in;
      vec4_t<T> X[4];
      scene_fetch<T, MAT>(P, s, x, in, X);
      scene_solve<T, MAT>(P, s, x, in, X);
    }
  } else {
    for (; s < e; s += 2 * step) {
      const int64_t s1 = s + step;
      const bool oa = !P.flags || P.flags[s], ob = s1 < e && (!P.flags || P.flags[s1]);
      TetIn<T> a, b;
      vec4_t<T> Xa[4], Xb[4];
      if (oa) scene_fetch<T, MAT>(P, s, x, a, Xa);
      if (ob) scene_fetch<T, MAT>(P, s1, x, b, Xb);
      if (oa) scene_solve<T, MAT>(P, s, x, a, Xa);
      if (ob) scene_solve<T, MAT>(P, s1, x, b, Xb);
    }
  }
}

// Sum of the residual terms of one thread's real slots s, s + step, ... < e
// of a scene, in slot order; PAIRS: two slots' inputs loaded before their
// terms are computed.
template <typename T, int MAT, bool PAIRS, typename Pos>
__device__ __forceinline__ double scene_residual(const SweepParams<T> &P, int64_t s,
                                                 const int64_t e, const int64_t step,
                                                 const Pos &x) {
  const auto ld = [&](int v) { return x.load(v); };
  double acc = 0;
  if constexpr (!PAIRS) {
    for (; s < e; s += step)
      if (P.vid[s] >= 0) {
        ResIn<T> in;
        res_fetch_pos<T, MAT>(P, s, in, ld);
        acc += res_compute<T, MAT>(in);
      }
  } else {
    for (; s < e; s += 2 * step) {
      const int64_t s1 = s + step;
      const bool oa = P.vid[s] >= 0, ob = s1 < e && P.vid[s1] >= 0;
      ResIn<T> a, b;
      if (oa) res_fetch_pos<T, MAT>(P, s, a, ld);
      if (ob) res_fetch_pos<T, MAT>(P, s1, b, ld);
      if (oa) acc += res_compute<T, MAT>(a);
      if (ob) acc += res_compute<T, MAT>(b);
    }
  }
  return acc;
}

// Control block of a scene's stop rule, in LDS: only thread 0 updates it
// (monitor_update), so it occupies no registers of the other threads.
__device__ __forceinline__ void ctl_start(Ctl &c, double runUntil, double pct, int isLast) {
  c.stopSolve = 0;
  c.endLayer = 0;
  c.iters = 0;
  c.errorChangeDecreased = 0;
  c.residualCount = 1;
  c.stopLayer = -1;
  c.isLastLayer = isLast;
  c.layer = 0;
  c.runUntil = runUntil;
  c.setPercentStop = pct;
  c.normresv = 0;
  c.lastResidual = 0;
}

// k_batch_plain for a scene whose positions fit in LDS: x, y, z of its nv
// vertices live in LDS (structure of arrays, pos[d * nv + local vertex])
// from the first color to the last residual, and go back to HBM at the end.
// The inverse mass (w) comes from x_old, a copy of x taken at the start of
// the substep.  Same updates in the same order as the HBM path.
template <typename T, int NT, int MAT>
__device__ __forceinline__ void batch_scene_lds(const BatchParams<T> &B, const int sc, T *pos,
                                                double *sh) {
  __shared__ int sStop;
  __shared__ double sRunUntil;
  const int tid = threadIdx.x;
  const int64_t vb = B.vert0[sc];
  const int nv = (int)(B.vert0[sc + 1] - vb);
  for (int i = tid; i < nv; i += NT) {
    const vec4_t<T> p = B.P.x[vb + i];
    pos[i] = p.x;
    pos[nv + i] = p.y;
    pos[2 * nv + i] = p.z;
  }
  auto ld_pos = [&](int v) -> vec4_t<T> {
    const int l = v - (int)vb;
    vec4_t<T> p;
    p.x = pos[l];
    p.y = pos[nv + l];
    p.z = pos[2 * nv + l];
    p.w = B.P.xold[v].w;
    return p;
  };
  auto st_pos = [&](int v, const vec4_t<T> &p) {
    const int l = v - (int)vb;
    pos[l] = p.x;
    pos[nv + l] = p.y;
    pos[2 * nv + l] = p.z;
  };
  Ctl c;
  c.stopSolve = 0;
  c.endLayer = 0;
  c.iters = 0;
  c.errorChangeDecreased = 0;
  c.residualCount = 1;
  c.stopLayer = -1;
  c.isLastLayer = 1;
  c.layer = 0;
  c.runUntil = B.runUntil0;
  c.setPercentStop = B.stopPct;
  c.normresv = 0;
  c.lastResidual = 0;
  if (tid == 0) {
    sStop = 0;
    sRunUntil = B.runUntil0;
  }
  __syncthreads();
  const int nc = B.ncolor[sc], cb = B.color0[sc];
  const int64_t k0 = B.cont0[sc], k1 = B.cont0[sc + 1];
  const int64_t q0 = B.pin0[sc], q1 = B.pin0[sc + 1];
  const int64_t s0 = B.slot0[sc], s1 = B.slot0[sc + 1];
  const double h2 = B.h * B.h;
  int it = 0;
  while (it < B.hardCap) {
    for (int k = 0; k < nc; ++k) {
      const int64_t a = B.colorOff[cb + k], e = a + B.colorCnt[cb + k];
      for (int64_t s = a + tid; s < e; s += NT) {
        TetIn<T> in;
        tet_fetch<T, MAT>(B.P, s, in);
        vec4_t<T> X[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int l = in.v[q] - (int)vb;
          X[q].x = pos[l];
          X[q].y = pos[nv + l];
          X[q].z = pos[2 * nv + l];
          X[q].w = in.O[q].w;
        }
        tet_solve_st<T, MAT>(B.P, s, in, X,
                            [&](int q, const vec4_t<T> &o) { st_pos(in.v[q], o); });
      }
      __syncthreads();
    }
    for (int64_t i = k0 + tid; i < k1; i += NT) {
      const int v = B.c.vert[i];
      const double *n = B.c.normal + 3 * i, *pc = B.c.point + 3 * i;
      vec4_t<T> p = ld_pos(v);
      const double xv[3] = {(double)p.x, (double)p.y, (double)p.z};
      double phi = n[0] * (xv[0] - pc[0]) + n[1] * (xv[1] - pc[1]) + n[2] * (xv[2] - pc[2]);
      if (phi > 0) phi = 0;
      const double C = B.contactResistance * phi;
      const double aT = B.c.compliance[i] / h2;
      const double mi = B.minv[v];
      const double dl = (-C - aT * B.c.lambda[i]) / (2 * mi + aT);
      p.x = (T)(xv[0] + mi * (dl * n[0]));
      p.y = (T)(xv[1] + mi * (dl * n[1]));
      p.z = (T)(xv[2] + mi * (dl * n[2]));
      B.c.lambda[i] += dl;
      st_pos(v, p);
    }
    __syncthreads();
    for (int64_t q = q0 + tid; q < q1; q += NT) {
      const int d = B.pinDof[q];
      pos[(d % 3) * nv + (d / 3 - (int)vb)] = B.pinVal[q];
    }
    __syncthreads();
    ++it;
    if (B.mp.computeResiduals) {
      double acc = 0;
      for (int64_t s = s0 + tid; s < s1; s += NT)
        if (B.P.vid[s] >= 0) {
          ResIn<T> in;
          res_fetch_pos<T, MAT>(B.P, s, in, ld_pos);
          acc += res_compute<T, MAT>(in);
        }
      const double Ks = (double)(k1 - k0);
      for (int64_t i = k0 + tid; i < k1; i += NT) {
        const int v = B.c.vert[i];
        const double *n = B.c.normal + 3 * i, *pc = B.c.point + 3 * i;
        const vec4_t<T> p = ld_pos(v);
        double phi = n[0] * ((double)p.x - pc[0]) + n[1] * ((double)p.y - pc[1]) +
                     n[2] * ((double)p.z - pc[2]);
        if (phi > 0) phi = 0;
        const double a = fabs(B.contactResistance * phi);
        const double cr = fabs((B.contactCompliance / h2) * B.c.lambda[i]);
        const double t = a - (a < cr ? a : cr);
        acc += Ks * (t * t);
      }
      const double tot = wg_sum<NT>(acc, sh);
      if (tid == 0) {
        monitor_update(c, B.mp, sqrt(tot), nullptr, 0);
        sStop = c.stopSolve | c.endLayer;
        sRunUntil = c.runUntil;
      }
      __syncthreads();
    }
    const int stop = sStop;
    const double ru = sRunUntil;
    __syncthreads();
    if (stop) break;
    if (!(it < B.iterations || ru >= 0)) break;
  }
  for (int i = tid; i < nv; i += NT) {
    vec4_t<T> p = B.P.x[vb + i];
    p.x = pos[i];
    p.y = pos[nv + i];
    p.z = pos[2 * nv + i];
    B.P.x[vb + i] = p;
  }
  if (tid == 0) {
    B.iters[sc] = it;
    B.lastResidual[sc] = c.lastResidual;
    B.runUntilOut[sc] = c.runUntil;
    B.residualCount[sc] = c.residualCount;
  }
}

// One scene per workgroup (plain method: a single all-elastic layer): the
// projectGS loop of xpbdLayer3D.m:315-341 with colors separated by workgroup
// barriers, then contacts, pins and the scene's own residual/stop rule.
// Every loop is bounded (hardCap) so the kernel always returns.  Scenes
// with at most ldsVerts vertices keep their positions in the dynamic LDS
// (batch_scene_lds), the others in HBM (batch_scene_hbm).  Each scene runs
// the body specialised for its materials (StVK-Voigt only, or with
// NeoHookean tetrahedra), chosen once per workgroup, so the StVK body
// carries no material branch.
template <typename T, int NT, int MAT>
__device__ __forceinline__ void batch_scene_hbm(const BatchParams<T> &B, const int sc,
                                                double *sh) {
  __shared__ int sStop;
  __shared__ double sRunUntil;
  const int tid = threadIdx.x;
  Ctl c;
  c.stopSolve = 0;
  c.endLayer = 0;
  c.iters = 0;
  c.errorChangeDecreased = 0;
  c.residualCount = 1;
  c.stopLayer = -1;
  c.isLastLayer = 1;
  c.layer = 0;
  c.runUntil = B.runUntil0;
  c.setPercentStop = B.stopPct;
  c.normresv = 0;
  c.lastResidual = 0;
  if (tid == 0) {
    sStop = 0;
    sRunUntil = B.runUntil0;
  }
  __syncthreads();
  const int nc = B.ncolor[sc], cb = B.color0[sc];
  const int64_t k0 = B.cont0[sc], k1 = B.cont0[sc + 1];
  const int64_t q0 = B.pin0[sc], q1 = B.pin0[sc + 1];
  const int64_t s0 = B.slot0[sc], s1 = B.slot0[sc + 1];
  const double h2 = B.h * B.h;
  int it = 0;
  while (it < B.hardCap) {
    for (int k = 0; k < nc; ++k) {
      const int64_t a = B.colorOff[cb + k], e = a + B.colorCnt[cb + k];
      for (int64_t s = a + tid; s < e; s += NT) project_tet<T, MAT>(B.P, s);
      __syncthreads();
    }
    for (int64_t i = k0 + tid; i < k1; i += NT) {
      const int v = B.c.vert[i];
      const double *n = B.c.normal + 3 * i, *pc = B.c.point + 3 * i;
      vec4_t<T> p = B.P.x[v];
      const double xv[3] = {(double)p.x, (double)p.y, (double)p.z};
      double phi = n[0] * (xv[0] - pc[0]) + n[1] * (xv[1] - pc[1]) + n[2] * (xv[2] - pc[2]);
      if (phi > 0) phi = 0;
      const double C = B.contactResistance * phi;
      const double aT = B.c.compliance[i] / h2;
      const double mi = B.minv[v];
      const double dl = (-C - aT * B.c.lambda[i]) / (2 * mi + aT);
      p.x = (T)(xv[0] + mi * (dl * n[0]));
      p.y = (T)(xv[1] + mi * (dl * n[1]));
      p.z = (T)(xv[2] + mi * (dl * n[2]));
      B.c.lambda[i] += dl;
      B.P.x[v] = p;
    }
    __syncthreads();
    for (int64_t q = q0 + tid; q < q1; q += NT) {
      const int d = B.pinDof[q];
      T *px = reinterpret_cast<T *>(&B.P.x[d / 3]);
      px[d % 3] = B.pinVal[q];
    }
    __syncthreads();
    ++it;
    if (B.mp.computeResiduals) {
      double acc = 0;
      for (int64_t s = s0 + tid; s < s1; s += NT)
        if (B.P.vid[s] >= 0) acc += residual_tet<T, false, MAT>(B.P, s);
      const double Ks = (double)(k1 - k0);
      for (int64_t i = k0 + tid; i < k1; i += NT) {
        const int v = B.c.vert[i];
        const double *n = B.c.normal + 3 * i, *pc = B.c.point + 3 * i;
        const vec4_t<T> p = B.P.x[v];
        double phi = n[0] * ((double)p.x - pc[0]) + n[1] * ((double)p.y - pc[1]) +
                     n[2] * ((double)p.z - pc[2]);
        if (phi > 0) phi = 0;
        const double a = fabs(B.contactResistance * phi);
        const double cr = fabs((B.contactCompliance / h2) * B.c.lambda[i]);
        const double t = a - (a < cr ? a : cr);
        acc += Ks * (t * t);
      }
      const double tot = wg_sum<NT>(acc, sh);
      if (tid == 0) {
        monitor_update(c, B.mp, sqrt(tot), nullptr, 0);
        sStop = c.stopSolve | c.endLayer;
        sRunUntil = c.runUntil;
      }
      __syncthreads();
    }
    const int stop = sStop;
    const double ru = sRunUntil;
    __syncthreads();
    if (stop) break;
    if (!(it < B.iterations || ru >= 0)) break;
  }
  if (tid == 0) {
    B.iters[sc] = it;
    B.lastResidual[sc] = c.lastResidual;
    B.runUntilOut[sc] = c.runUntil;
    B.residualCount[sc] = c.residualCount;
  }
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_batch_plain(BatchParams<T> B) {
  __shared__ double sh[NT / 64];
  extern __shared__ __align__(16) unsigned char mlx_dyn_lds[];
  const int sc = B.sceneList ? B.sceneList[blockIdx.x] : (int)blockIdx.x;
  const bool nh = B.sceneNH && B.sceneNH[sc];  // the scene is all NeoHookean
  if (B.vert0[sc + 1] - B.vert0[sc] <= B.ldsVerts) {
    T *pos = reinterpret_cast<T *>(mlx_dyn_lds);
    if (nh)
      batch_scene_lds<T, NT, kMatNH>(B, sc, pos, sh);
    else
      batch_scene_lds<T, NT, kMatStVK>(B, sc, pos, sh);
    return;
  }
  if (nh)
    batch_scene_hbm<T, NT, kMatNH>(B, sc, sh);
  else
    batch_scene_hbm<T, NT, kMatStVK>(B, sc, sh);
}

}  // namespace mlx

namespace mlx {

// ---------------------------------------------------------------------------
// 2D NeoHookean constraint pair (xpbdLayer2D.neoHookeanConstraints)
// ---------------------------------------------------------------------------

// One color: per triangle the deviatoric constraint C_D = ||F||_F
// (compliance 1/(area h^2 mu)) and the volumetric C_H = det F - gamma
// (compliance 1/(area h^2 lambda), gamma = 1 + mu/lambda), both evaluated at
// the positions the color starts from, their displacements summed
// (mexNeoHookeanDConstraint.cpp / mexNeoHookeanHConstraint.cpp).  With
// F = sum_a x_a g_a^T: grad_a ||F|| = F g_a / ||F||, grad_a det F = cof(F) g_a.
// nh: mu | lambda_Lame | lambda_D | lambda_H, each ld long (slot order).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_neohookean2d(SweepParams<T> P, T *nh) {
  const int64_t s = P.begin + (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (s >= P.end) return;
  if (P.flags && !P.flags[s]) return;
  const int64_t ld = P.ld;
  const int v0 = P.vid[s], v1 = P.vid[ld + s], v2 = P.vid[2 * ld + s];
  const vec4_t<T> X0 = P.x[v0], X1 = P.x[v1], X2 = P.x[v2];
  const T D[4] = {P.dminv[s], P.dminv[ld + s], P.dminv[2 * ld + s], P.dminv[3 * ld + s]};
  const T area = P.vol[s];
  const T mu = nh[s], lamL = nh[ld + s];
  const T e[2][2] = {{X0.x - X2.x, X0.y - X2.y}, {X1.x - X2.x, X1.y - X2.y}};
  T F[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) F[r][c] = e[0][r] * D[2 * c] + e[1][r] * D[1 + 2 * c];
  const T G[3][2] = {{D[0], D[2]}, {D[1], D[3]}, {-(D[0] + D[1]), -(D[2] + D[3])}};
  const T w[3] = {X0.w, X1.w, X2.w};
  const T ih2 = P.ih2;
  // deviatoric
  const T n2 = F[0][0] * F[0][0] + F[0][1] * F[0][1] + F[1][0] * F[1][0] + F[1][1] * F[1][1];
  const T CD = sqrt(n2);
  const T atD = ((T(1) / area) * ih2) * (T(1) / mu);
  T gD[3][2], sD = atD;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int r = 0; r < 2; ++r) gD[a][r] = (F[r][0] * G[a][0] + F[r][1] * G[a][1]) / CD;
    sD += w[a] * (gD[a][0] * gD[a][0] + gD[a][1] * gD[a][1]);
  }
  const T lD = nh[2 * ld + s];
  const T dlD = -(CD + atD * lD) / sD;
  // volumetric
  const T detF = F[0][0] * F[1][1] - F[0][1] * F[1][0];
  const T gamma = T(1) + mu / lamL;
  const T CH = detF - gamma;
  const T atH = ((T(1) / area) * ih2) * (T(1) / lamL);
  const T cof[2][2] = {{F[1][1], -F[1][0]}, {-F[0][1], F[0][0]}};
  T gH[3][2], sH = atH;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int r = 0; r < 2; ++r) gH[a][r] = cof[r][0] * G[a][0] + cof[r][1] * G[a][1];
    sH += w[a] * (gH[a][0] * gH[a][0] + gH[a][1] * gH[a][1]);
  }
  const T lH = nh[3 * ld + s];
  const T dlH = -(CH + atH * lH) / sH;
  nh[2 * ld + s] = lD + dlD;
  nh[3 * ld + s] = lH + dlH;
  const vec4_t<T> Xs[3] = {X0, X1, X2};
  const int vs[3] = {v0, v1, v2};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    vec4_t<T> o = Xs[a];
    o.x = (o.x + w[a] * gD[a][0] * dlD) + w[a] * gH[a][0] * dlH;
    o.y = (o.y + w[a] * gD[a][1] * dlD) + w[a] * gH[a][1] * dlH;
    P.x[vs[a]] = o;
  }
}

}  // namespace mlx

namespace mlx {

// ---------------------------------------------------------------------------
// persistent iteration loop for layers without rigid bodies
// ---------------------------------------------------------------------------

// Grid-wide barrier of a cooperative launch: one arrival counter in device
// memory.  Every thread issues an agent-scope release fence (its own stores
// are written back past its XCD's L2), the workgroup meets, one thread
// arrives and waits for `target` arrivals, the workgroup meets again and
// every thread issues an agent-scope acquire fence (stale cache lines are
// dropped).  The k-th barrier of a launch waits for k * gridDim.x arrivals.
// The wait is bounded: a workgroup that never arrives sets *err and the
// barrier returns false instead of spinning forever.
constexpr long kBarrierSpinLimit = 1L << 26;

__device__ __forceinline__ bool grid_barrier(unsigned *count, unsigned target, int *err) {
  __shared__ int s_ok;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int ok = 1;
    long spins = 0;
    while (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > kBarrierSpinLimit) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
    }
    s_ok = ok;
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return s_ok != 0;
}

// XCD (accelerator complex die) that runs the calling wave: hardware
// register XCC_ID (gfx940+), low 4 bits.  Workgroups with the same id share
// one L2.
__device__ __forceinline__ unsigned xcd_id() {
  return __builtin_amdgcn_s_getreg(20 | (0 << 6) | (15 << 11)) & 15u;
}

// Barrier among workgroups that all run on one XCD, hence share its L2,
// and read each other's data only with ldc<true> (past the vector L1), in
// two halves so that loads of unchanging data can be issued in between.
// xcd_arrive: every thread waits until its own memory operations are
// acknowledged (stores have reached the L2, the vector L1 being
// write-through), the workgroup meets, one thread adds its arrival.
// xcd_wait: that thread waits for `target` arrivals, the workgroup meets.
// Nothing is written back from or dropped in the L2 (a device-scope
// acquire would invalidate it; a workgroup-scope buffer_inv does not drop
// the CU's L1 lines for ordinary workgroups, measured: every read of
// another workgroup's data through the L1 was stale with it, none with
// L1-bypassing loads).  The wait is bounded like grid_barrier's.
__device__ __forceinline__ void xcd_arrive(unsigned *count) {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool xcd_wait(unsigned *count, unsigned target, int *err) {
  __shared__ int s_ok;
  if (threadIdx.x == 0) {
    int ok = 1;
    long spins = 0;
    while (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > kBarrierSpinLimit) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
    }
    s_ok = ok;
  }
  __syncthreads();
  asm volatile("" ::: "memory");
  return s_ok != 0;
}

// Workgroup groups by XCD for a cooperative launch: each workgroup takes a
// rank on its XCD's counter reg[xcd], all workgroups meet once at a grid
// barrier (counter reg[16]), then the XCD with the most workgroups (lowest
// id on ties) is chosen.  Returns this workgroup's rank in the chosen group
// and sets *n to its size, or returns -1 for workgroups of other XCDs (they
// take no further part).  reg: 17 zeroed counters.
__device__ __forceinline__ int xcd_group(unsigned *reg, int *err, int *n) {
  __shared__ int s_rank, s_n;
  if (threadIdx.x == 0) {
    const unsigned x = xcd_id();
    s_rank = (int)__hip_atomic_fetch_add(&reg[x], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_n = (int)x;
  }
  if (!grid_barrier(reg + 16, gridDim.x, err)) return -1;
  if (threadIdx.x == 0) {
    unsigned best = 0, bn = 0;
    for (unsigned k = 0; k < 16; ++k) {
      const unsigned c = __hip_atomic_load(&reg[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (c > bn) {
        bn = c;
        best = k;
      }
    }
    s_n = (unsigned)s_n == best ? (int)bn : 0;
  }
  __syncthreads();
  *n = s_n;
  return s_n > 0 ? s_rank : -1;
}

template <typename T>
struct PersistParams {
  SweepParams<T> P;              // P.ctl is the layer's control block
  const int32_t *colorOff;       // first slot of each color
  const int32_t *colorCnt;       // elements of each color
  int numColors;
  const int32_t *lamSlot2d;      // 2D residual lambda pages (DIM == 2)
  ContactArrays ca;
  const int32_t *ecIdx;          // elastic contacts of the layer
  int nec;
  const double *minv;
  double resistance, h;
  const double *uniq, *ucnt;     // 3D contact-residual compliance values
  int nUniq;
  const int32_t *pinDof;
  const T *pinVal;
  int64_t nPin;
  MonitorParams mp;
  double *partials;              // gridDim.x entries
  double *resid;
  int residCap;
  unsigned *bar;                 // arrival counter, zero at launch
  int *err;                      // barrier failure flag
  unsigned *reg;                 // LOCAL: 17 XCD group counters, zero at launch
  int maxIters;
  int maxColorCnt;               // elements of the largest color
};

// Up to maxIters projectGS iterations (xpbdLayer3D.m:306-344; 2D
// xpbdLayer2D.m:174-198) of a layer whose elements are all elastic, in one
// cooperative launch: the colors in order, each split over the workgroups
// and closed by a barrier; the elastic contacts (one per vertex, all
// workgroups), a barrier, then the pins; the residual (each workgroup a
// fixed share of the elements and contacts) and the monitor.  Every
// workgroup re-reads the control block after the last barrier of an
// iteration, so all stop together when the monitor sets stopSolve/endLayer.
// LOCAL (3D): only the workgroups of one XCD take part (xcd_group), data
// written in the launch is read with ldc<true>, and they meet at
// xcd_arrive/xcd_wait; when every color fits one pass of the group, each
// thread fetches its slot of the next color (tet_fetch) between the two
// halves of the barrier, so only the position gathers follow it.
// Otherwise all workgroups take part and meet at grid_barrier.
template <typename T, int DIM, bool LOCAL>
__global__ __launch_bounds__(kBlock) void k_persist(PersistParams<T> Q) {
  static_assert(DIM == 3 || !LOCAL, "the XCD-local loop is 3D only");
  __shared__ double sh[kBlock];
  int G = gridDim.x, w = blockIdx.x;
  if (LOCAL) {
    w = xcd_group(Q.reg, Q.err, &G);
    if (w < 0) return;
  }
  const int t = threadIdx.x;
  const int64_t stride = (int64_t)G * kBlock;
  Ctl *ctl = const_cast<Ctl *>(Q.P.ctl);
  unsigned nbar = 0;
  auto arrive = [&]() {
    ++nbar;
    if (LOCAL) xcd_arrive(Q.bar);
  };
  auto wait = [&]() -> bool {
    return LOCAL ? xcd_wait(Q.bar, nbar * (unsigned)G, Q.err)
                 : grid_barrier(Q.bar, nbar * (unsigned)G, Q.err);
  };
  auto barrier = [&]() -> bool {
    arrive();
    return wait();
  };
  const bool onePass = LOCAL && Q.maxColorCnt <= stride;
  for (int it = 0; it < Q.maxIters; ++it) {
    if ((ldc<LOCAL>(&ctl->stopSolve) | ldc<LOCAL>(&ctl->endLayer)) != 0) break;
    bool swept = false;
    if constexpr (DIM == 3 && LOCAL) {
      if (onePass) {
        // slot of this thread in color c (colors fit one pass)
        TetIn<T> cur;
        int64_t sc = -1;
        int c = 0;
        while (c < Q.numColors && Q.colorCnt[c] == 0) ++c;
        if (c < Q.numColors) {
          const int64_t s0 = Q.colorOff[c] + (int64_t)w * kBlock + t;
          if (s0 < Q.colorOff[c] + Q.colorCnt[c]) {
            sc = s0;
            tet_fetch<T>(Q.P, sc, cur);
          }
        }
        while (c < Q.numColors) {
          if (sc >= 0) tet_solve<T, true>(Q.P, sc, cur);
          arrive();
          int cn = c + 1;
          while (cn < Q.numColors && Q.colorCnt[cn] == 0) ++cn;
          sc = -1;
          if (cn < Q.numColors) {
            const int64_t s0 = Q.colorOff[cn] + (int64_t)w * kBlock + t;
            if (s0 < Q.colorOff[cn] + Q.colorCnt[cn]) {
              sc = s0;
              tet_fetch<T>(Q.P, sc, cur);
            }
          }
          if (!wait()) return;
          c = cn;
        }
        swept = true;
      }
    }
    for (int c = 0; !swept && c < Q.numColors; ++c) {
      const int cnt = Q.colorCnt[c];
      if (cnt == 0) continue;
      const int64_t off = Q.colorOff[c], end = off + cnt;
      if constexpr (DIM == 3) {
        color_slots<T, LOCAL>(Q.P, off + (int64_t)w * kBlock + t, end, stride);
      } else {
        for (int64_t s = off + (int64_t)w * kBlock + t; s < end; s += stride)
          project_tri<T>(Q.P, s);
      }
      if (!barrier()) return;
    }
    // contacts (distinct vertices) over all workgroups, then the pins
    if (Q.nec > 0) {
      for (int64_t q = (int64_t)w * kBlock + t; q < Q.nec; q += stride)
        contact_elastic_one<T, DIM, LOCAL>(Q.ca, Q.ecIdx[q], Q.P.x, Q.minv, Q.resistance, Q.h);
      if (Q.nPin > 0 && !barrier()) return;
    }
    for (int64_t q = (int64_t)w * kBlock + t; q < Q.nPin; q += stride)
      pin_one<T, DIM>(Q.pinDof, Q.pinVal, Q.P.x, q);
    if (!barrier()) return;
    if (Q.mp.computeResiduals) {
      double acc = 0;
      if constexpr (DIM == 3) {
        acc = residual_slots<T, LOCAL>(Q.P, (int64_t)w * kBlock + t, Q.P.ld, stride);
      } else {
        for (int64_t s = (int64_t)w * kBlock + t; s < Q.P.ld; s += stride)
          if (Q.P.vid[s] >= 0) acc += residual_elem<T, DIM, LOCAL>(Q.P, s, Q.lamSlot2d);
      }
      for (int64_t i = (int64_t)w * kBlock + t; i < Q.ca.K; i += stride)
        acc += residual_contact_one<T, DIM, LOCAL>(Q.ca, (int)i, Q.P.x, Q.resistance, Q.h,
                                                   Q.uniq, Q.ucnt, Q.nUniq);
      double v1[1] = {acc}, o[1];
      block_sum<1>(v1, sh, o);
      if (t == 0) Q.partials[w] = o[0];
      if (!barrier()) return;
    }
    if (w == 0 && t == 0) {
      // this thread is the only writer of the control block in the launch
      ctl->iters += 1;
      if (Q.mp.computeResiduals) {
        double sum = 0;
        for (int q = 0; q < G; ++q) sum += ldc<LOCAL>(Q.partials + q);
        monitor_update(*ctl, Q.mp, sqrt(sum), Q.resid, Q.residCap);
      }
    }
    if (!barrier()) return;
  }
}

// ---------------------------------------------------------------------------
// dataflow color sweep (one iteration of a layer without rigid elements)
// ---------------------------------------------------------------------------

// Gauss-Seidel over graph colors fixes, for every vertex, the order in
// which its elements update it (ascending color; the elements of one color
// share no vertex) and nothing else.  Here the elements of a color are
// taken in chunks of 64 consecutive slots, one per wavefront, and a chunk
// may run as soon as the chunks holding the last lower-color updater of
// each of its vertices have run (its dependences, precomputed per mesh:
// waiting for the last updater of a vertex is enough, since that one waited
// for the vertex's earlier updaters).  Every vertex then sees its updates
// in the color order and the element arithmetic is sweep_slot's, so the
// results are bit-identical to per-color launches.  A chunk's flag holds
// the number of the layer's iterations it has finished.
template <typename T>
struct FlowParams {
  SweepParams<T> P;              // P.ctl the layer's control block, P.flags null
  const int32_t *colorOff;       // first slot of each color
  const int32_t *colorCnt;       // elements of each color
  int numColors;
  const int32_t *chunkBase;      // numColors + 1: first chunk of each color
  const int32_t *depOff;         // nChunks + 1
  const int32_t *dep;            // dependences of each chunk (chunk ids)
  unsigned *flags;               // nChunks: iterations finished (zero at the layer's start)
  unsigned base;                 // iterations of the layer run by earlier launches
  ContactArrays ca;
  const int32_t *ecIdx;          // elastic contacts of the layer
  int nec;
  const double *minv;
  double resistance, h;
  const int32_t *pinDof;
  const T *pinVal;
  int64_t nPin;
  unsigned *bar;                 // XCD barrier arrivals, zero at launch
  int *err;                      // wait or barrier failure flag
  unsigned *reg;                 // 17 XCD group counters, zero at launch
  long spinLimit;                // polls of one wait before it fails
};

// Whole wavefront: wait until every dependence of chunk g has finished
// this iteration (flag > base).
template <typename T>
__device__ __forceinline__ bool flow_wait(const FlowParams<T> &Q, int g, int lane) {
  const int d0 = Q.depOff[g], nd = Q.depOff[g + 1] - d0;
  long spins = 0;
  for (;;) {
    int ready = 1;
    for (int q = lane; q < nd; q += 64) ready &= ldc<true>(Q.flags + Q.dep[d0 + q]) > Q.base;
    if (__all(ready)) return true;
    ++spins;
    if (spins > Q.spinLimit || ((spins & 255) == 0 && ldc<true>(Q.err) != 0)) {
      if (lane == 0) __hip_atomic_store(Q.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// One iteration's sweep of a layer without rigid elements on the
// workgroups of one XCD (xcd_group; the others return): every wavefront
// takes its chunk of each color in color order and runs it when its
// dependences allow, with no barrier between colors; then, as projectGS
// does after its colors, the elastic contacts (distinct vertices) and the
// pins, each after a barrier of the group.  The residual and the monitor
// follow in k_residual_tail.  Launched cooperatively, so every workgroup
// is resident; every wait is bounded and a failed one sets *err (the
// kernel then returns without finishing the sweep).
template <typename T, int MAT>
__global__ __launch_bounds__(kBlock) void k_flow(FlowParams<T> Q) {
  int G = 0;
  const int w = xcd_group(Q.reg, Q.err, &G);
  if (w < 0) return;
  const Ctl *ctl = Q.P.ctl;
  if ((ctl->stopSolve | ctl->endLayer) != 0) return;  // every workgroup reads the same word
  const int t = threadIdx.x, lane = t & 63;
  const SweepParams<T> &P = Q.P;
  const int wfirst = w * (kBlock / 64) + (t >> 6), wstride = G * (kBlock / 64);
  bool ok = true;
  for (int c = 0; c < Q.numColors && ok; ++c) {
    const int64_t off = Q.colorOff[c], end = off + Q.colorCnt[c];
    const int nch = (int)((Q.colorCnt[c] + 63) / 64);
    for (int j = wfirst; j < nch && ok; j += wstride) {
      const int g = Q.chunkBase[c] + j;
      const int64_t s = off + 64 * (int64_t)j + lane;
      const bool mine = s < end;
      // the slot's data does not change in the launch: fetched before the wait
      TetIn<T> in;
      if (mine) tet_fetch<T, MAT>(P, s, in);
      ok = flow_wait(Q, g, lane);
      if (!ok) break;
      if (mine) {
        vec4_t<T> X[4];
        tet_gather<T, true>(P, in, X);
        tet_solve_x<T, MAT>(P, s, in, X);
      }
      // the chunk's positions reach the L2 before its flag says so;
      // readers load them past the vector L1 (ldc<true>)
      asm volatile("" ::: "memory");
      __builtin_amdgcn_s_waitcnt(0);
      asm volatile("" ::: "memory");
      if (lane == 0)
        __hip_atomic_store(Q.flags + g, Q.base + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  unsigned nbar = 0;
  xcd_arrive(Q.bar);
  if (!xcd_wait(Q.bar, ++nbar * (unsigned)G, Q.err)) return;
  if (ldc<true>(Q.err) != 0) return;
  const int64_t stride = (int64_t)G * kBlock, first = (int64_t)w * kBlock + t;
  for (int64_t q = first; q < Q.nec; q += stride)
    contact_elastic_one<T, 3, true>(Q.ca, Q.ecIdx[q], Q.P.x, Q.minv, Q.resistance, Q.h);
  if (Q.nPin > 0) {
    xcd_arrive(Q.bar);
    if (!xcd_wait(Q.bar, ++nbar * (unsigned)G, Q.err)) return;
    for (int64_t q = first; q < Q.nPin; q += stride) pin_one<T, 3>(Q.pinDof, Q.pinVal, Q.P.x, q);
  }
}

}  // namespace mlx

namespace mlx {

// ---------------------------------------------------------------------------
// batched scenes split over several workgroups
// ---------------------------------------------------------------------------

// Work split of k_batch_split.  The grid is 8 * P workgroups, dealt
// round-robin over the 8 XCDs; a workgroup takes a rank k on its XCD's
// counter and serves entry[xcd * P + k]: scene * 64 + rank of that scene's
// G workgroups, or -1 (idle).  All workgroups of a scene are on one XCD,
// so they share its L2 and meet at xcd_arrive/xcd_wait on bar[scene].
// After the ranks are taken, one grid barrier lets every workgroup check
// that each XCD received exactly P workgroups; otherwise all return at
// once and *mismatch is set (nothing has been written).
struct SplitParams {
  const int32_t *entry;   // 8 * P entries
  const int32_t *sceneG;  // workgroups per scene (1..64)
  const int32_t *pbase;   // first partial of each scene
  int P;                  // workgroups per XCD
  unsigned *reg;          // 17 counters (ranks per XCD, grid barrier), zero at launch
  unsigned *bar;          // per scene arrival counter, zero at launch
  int32_t *sceneStop;     // per scene stop flag published by rank 0
  double *sceneRunUntil;  // per scene runUntil published by rank 0
  double *partials;       // per (scene, rank) residual partial
  int *err;               // barrier failure flag
  int *mismatch;          // XCD placement differed from the plan
};

// One scene's plain-method layerSolve (the loop of k_batch_plain) on rank r
// of its G workgroups: each color split over the ranks, the contacts
// (distinct vertices) then the pins, the residual from all ranks, the
// monitor on rank 0, which publishes the stop decision.  COH (G > 1): data
// other ranks write is read with ldc<true> and the ranks meet at the
// scene's XCD barrier; otherwise a workgroup barrier.
template <typename T, int NT, bool COH>
__device__ __forceinline__ void batch_scene(const BatchParams<T> &B, const SplitParams &S,
                                            const int sc, const int r, const int G, double *sh) {
  __shared__ int sStop;
  __shared__ double sRunUntil;
  const int tid = threadIdx.x;
  unsigned nbar = 0;
  Ctl c;
  c.stopSolve = 0;
  c.endLayer = 0;
  c.iters = 0;
  c.errorChangeDecreased = 0;
  c.residualCount = 1;
  c.stopLayer = -1;
  c.isLastLayer = 1;
  c.layer = 0;
  c.runUntil = B.runUntil0;
  c.setPercentStop = B.stopPct;
  c.normresv = 0;
  c.lastResidual = 0;
  const int nc = B.ncolor[sc], cb = B.color0[sc];
  const int64_t k0 = B.cont0[sc], k1 = B.cont0[sc + 1];
  const int64_t q0 = B.pin0[sc], q1 = B.pin0[sc + 1];
  const int64_t s0 = B.slot0[sc], s1 = B.slot0[sc + 1];
  const int64_t stride = (int64_t)G * NT, first = (int64_t)r * NT + tid;
  double *part = S.partials + S.pbase[sc];
  auto sync = [&]() -> bool {
    if (!COH) {
      __syncthreads();
      return true;
    }
    ++nbar;
    xcd_arrive(S.bar + sc);
    return xcd_wait(S.bar + sc, nbar * (unsigned)G, S.err);
  };
  const double h2 = B.h * B.h;
  const double Ks = (double)(k1 - k0);
  int it = 0;
  double ru = B.runUntil0;
  while (it < B.hardCap) {
    for (int k = 0; k < nc; ++k) {
      const int64_t a = B.colorOff[cb + k], e = a + B.colorCnt[cb + k];
      for (int64_t s = a + first; s < e; s += stride) {
        TetIn<T> in;
        tet_fetch<T>(B.P, s, in);
        tet_solve<T, COH>(B.P, s, in);
      }
      if (!sync()) return;
    }
    for (int64_t i = k0 + first; i < k1; i += stride) {
      const int v = B.c.vert[i];
      const double *n = B.c.normal + 3 * i, *pc = B.c.point + 3 * i;
      vec4_t<T> p = ldc<COH>(B.P.x + v);
      const double xv[3] = {(double)p.x, (double)p.y, (double)p.z};
      double phi = n[0] * (xv[0] - pc[0]) + n[1] * (xv[1] - pc[1]) + n[2] * (xv[2] - pc[2]);
      if (phi > 0) phi = 0;
      const double C = B.contactResistance * phi;
      const double aT = B.c.compliance[i] / h2;
      const double mi = B.minv[v];
      const double li = B.c.lambda[i];  // written only by this thread
      const double dl = (-C - aT * li) / (2 * mi + aT);
      p.x = (T)(xv[0] + mi * (dl * n[0]));
      p.y = (T)(xv[1] + mi * (dl * n[1]));
      p.z = (T)(xv[2] + mi * (dl * n[2]));
      B.c.lambda[i] = li + dl;
      B.P.x[v] = p;
    }
    if (k1 > k0 && q1 > q0 && !sync()) return;
    for (int64_t q = q0 + first; q < q1; q += stride) {
      const int d = B.pinDof[q];
      T *px = reinterpret_cast<T *>(&B.P.x[d / 3]);
      px[d % 3] = B.pinVal[q];
    }
    if (!sync()) return;
    ++it;
    int stop = 0;
    if (B.mp.computeResiduals) {
      double acc = 0;
      for (int64_t s = s0 + first; s < s1; s += stride)
        if (B.P.vid[s] >= 0) acc += residual_tet<T, COH>(B.P, s);
      for (int64_t i = k0 + first; i < k1; i += stride) {
        const int v = B.c.vert[i];
        const double *n = B.c.normal + 3 * i, *pc = B.c.point + 3 * i;
        const vec4_t<T> p = ldc<COH>(B.P.x + v);
        double phi = n[0] * ((double)p.x - pc[0]) + n[1] * ((double)p.y - pc[1]) +
                     n[2] * ((double)p.z - pc[2]);
        if (phi > 0) phi = 0;
        const double a = fabs(B.contactResistance * phi);
        const double cr = fabs((B.contactCompliance / h2) * B.c.lambda[i]);
        const double t = a - (a < cr ? a : cr);
        acc += Ks * (t * t);
      }
      const double tot = wg_sum<NT>(acc, sh);
      if (tid == 0) part[r] = tot;
      if (!sync()) return;
      if (r == 0 && tid == 0) {
        double sum = 0;
        for (int q = 0; q < G; ++q) sum += ldc<COH>(part + q);
        monitor_update(c, B.mp, sqrt(sum), nullptr, 0);
        S.sceneStop[sc] = c.stopSolve | c.endLayer;
        S.sceneRunUntil[sc] = c.runUntil;
      }
      if (!sync()) return;
      if (tid == 0) {
        sStop = ldc<COH>(S.sceneStop + sc);
        sRunUntil = ldc<COH>(S.sceneRunUntil + sc);
      }
      __syncthreads();
      stop = sStop;
      ru = sRunUntil;
      __syncthreads();
    }
    if (stop) break;
    if (!(it < B.iterations || ru >= 0)) break;
  }
  if (r == 0 && tid == 0) {
    B.iters[sc] = it;
    B.lastResidual[sc] = c.lastResidual;
    B.runUntilOut[sc] = c.runUntil;
    B.residualCount[sc] = c.residualCount;
  }
}

// Scenes of uneven size split over workgroups of one XCD each (SplitParams).
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_batch_split(BatchParams<T> B, SplitParams S) {
  __shared__ double sh[NT / 64];
  __shared__ int sEntry, sOk;
  extern __shared__ __align__(16) unsigned char mlx_dyn_lds[];
  if (threadIdx.x == 0) {
    const unsigned x = xcd_id();
    const unsigned k =
        __hip_atomic_fetch_add(&S.reg[x], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sEntry = (x < 8u && k < (unsigned)S.P) ? S.entry[x * S.P + k] : -1;
  }
  if (!grid_barrier(S.reg + 16, gridDim.x, S.err)) return;
  if (threadIdx.x == 0) {
    int ok = 1;
    for (int x = 0; x < 8; ++x)
      ok &= __hip_atomic_load(&S.reg[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
            (unsigned)S.P;
    sOk = ok;
    if (!ok && blockIdx.x == 0) *S.mismatch = 1;
  }
  __syncthreads();
  if (!sOk || sEntry < 0) return;
  const int sc = sEntry >> 6, r = sEntry & 63, G = S.sceneG[sc];
  if (G > 1)
    batch_scene<T, NT, true>(B, S, sc, r, G, sh);
  else if (B.vert0[sc + 1] - B.vert0[sc] <= B.ldsVerts)
    batch_scene_lds<T, NT, kMatStVK>(B, sc, reinterpret_cast<T *>(mlx_dyn_lds), sh);
  else
    batch_scene<T, NT, false>(B, S, sc, 0, 1, sh);
}

}  // namespace mlx

