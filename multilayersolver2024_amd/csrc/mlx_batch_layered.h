// Batched scenes with rigid clusters: every scene runs the two-layer
// layerSolve of xpbdLayer3D.m:109-226 inside its own workgroup -- a coarse
// layer whose rigid elements (the scene's clusters, each connected group a
// body) move as bodies coupled to the elastic rest through the boundary
// constraints, then the all-elastic layer -- with the scene's own residual,
// stop rule and contacts.  One launch per substep for all scenes.  Included
// once, by mlx_solver.hip, after mlx_kernels.h.
#pragma once

namespace mlx {

constexpr int kMaxSceneBodies = 8;

// The learned warm start of every scene's all-elastic layer
// (mlx_batch_set_learned): lambda_0 per slot from the network at the start
// of the substep, and the scratch the scene's workgroup uses to apply and
// judge it.
template <typename T>
struct BatchLearn {
  const T *lam0;                // nc ld: lambda_0 (k_learn_infer)
  T *dx;                        // 12 ld: displacement of (vertex a, slot s)
  vec4_t<T> *xsave;             // nV: positions before the warm start
  T *lamsave;                   // nc ld: multipliers before the warm start
  const int32_t *incOff, *inc;  // incidences a ld + s of each vertex, ascending slot
  const uint8_t *pinMask;       // 3 nV
  unsigned long long *counts;   // warm starts kept, rejected
};

// The coarse layer's maps, fixed for the batch (mlx_batch_set_rigid).
template <typename T>
struct LayeredParams {
  const int32_t *body0;            // ns + 1: scene sc owns bodies body0[sc] .. body0[sc + 1]
  const int32_t *rigOff, *rigVert; // rigid vertices of each body (CSR)
  const int32_t *bndOff, *bndVert; // boundary vertices of each body (CSR)
  const int32_t *vbody;            // nV: 0 elastic, else global body + 1
  const uint8_t *flags0;           // ld: slot elastic in the coarse layer
  Body *bodies;
  double *ri;                      // 3 nV: r_i of the rigid vertices
  T *rv;                           // 3 nV: residual velocity
  double *lamB;                    // boundary multipliers (bndVert order)
  const double *mass;              // nV
  int32_t *coarseIters;            // per scene: iterations the coarse layer ran
  int it0, it1;                    // iterations of the coarse and the elastic layer
  double boundaryCompliance;
  BatchLearn<T> learn;             // used by the LEARN kernels only
};

// Workgroup sums of N doubles per thread in a fixed order (butterfly in each
// wavefront, then the wavefronts in order); shw holds N * NT / 64 doubles.
// Every thread gets the sums.
template <int NT, int N>
__device__ __forceinline__ void wg_sum_n(double (&v)[N], double *shw) {
  constexpr int NW = NT / 64;
#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) shw[k * NW + (threadIdx.x >> 6)] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double t = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) t += shw[k * NW + w];
    v[k] = t;
  }
  __syncthreads();
}

// The kernel's static LDS, declared once at kernel scope (function-scope
// __shared__ variables would be repeated for every instantiation of the
// scene body), so the dynamic LDS left for positions is as large as it can
// be: the largest sum is 7 doubles per wavefront.
template <int NT>
struct LayeredShared {
  double shw[7 * (NT / 64)];
  double sR[kMaxSceneBodies][9], sC[kMaxSceneBodies][3];
  Ctl c;  // the scene's stop rule, updated by thread 0 only
  double sRunUntil;
  int sStop, sStopSolve;
};

// Thread 0's part of wg_restrict: the body's mass, pose, inertia and
// angular velocity from the workgroup's sums.  Not inlined: its pivoted
// 3x3 solve in double would otherwise set the register budget of the whole
// scene kernel, whose hot loops then spill.
__device__ __noinline__ void restrict_finish(Body &B, double M, const double *com,
                                             const double *cdot, double *b) {
  B.M = M;
  B.invM = 1.0 / M;
  for (int d = 0; d < 3; ++d) {
    B.COM[d] = com[d];
    B.COMDOT[d] = cdot[d];
  }
  for (int i = 0; i < 9; ++i) {
    B.I0[i] = b[i];
    B.I[i] = b[i];
    B.R[i] = (i % 4 == 0) ? 1.0 : 0.0;
    B.Rstep[i] = B.R[i];
  }
  solve3_pivot(b, b + 9, B.omega);
}

// makeRigidsFromConnectivity / matComputeRigidBodyProperties3D
// (AdaptiveMesh3D.m:32-43, RigidBody3D.m:198-265) for body k of the scene,
// as k_restrict does it, with the workgroup's sums.
template <typename T, int NT, typename Pos>
__device__ __forceinline__ void wg_restrict(const LayeredParams<T> &Lp, int k, const Pos &x,
                                            const double *mass, double *shw) {
  const int off = Lp.rigOff[k], n = Lp.rigOff[k + 1] - off;
  const T *rv = Lp.rv;
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int q = threadIdx.x; q < n; q += NT) {
    const int v = Lp.rigVert[off + q];
    const double m = mass[v];
    const vec4_t<T> p = x.load(v);
    a[0] += m;
    a[1] += (double)p.x * m;
    a[2] += (double)p.y * m;
    a[3] += (double)p.z * m;
    a[4] += (double)rv[3 * v] * m;
    a[5] += (double)rv[3 * v + 1] * m;
    a[6] += (double)rv[3 * v + 2] * m;
  }
  wg_sum_n<NT, 7>(a, shw);
  const double M = a[0];
  const double com[3] = {a[1] / M, a[2] / M, a[3] / M};
  const double cdot[3] = {a[4] / M, a[5] / M, a[6] / M};
  double b[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = threadIdx.x; q < n; q += NT) {
    const int v = Lp.rigVert[off + q];
    const double m = mass[v];
    const vec4_t<T> p = x.load(v);
    const double r[3] = {(double)p.x - com[0], (double)p.y - com[1], (double)p.z - com[2]};
    const double vd[3] = {(double)rv[3 * v] - cdot[0], (double)rv[3 * v + 1] - cdot[1],
                          (double)rv[3 * v + 2] - cdot[2]};
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
    Lp.ri[3 * v] = r[0];
    Lp.ri[3 * v + 1] = r[1];
    Lp.ri[3 * v + 2] = r[2];
  }
  {
    // two sums of six (each component is reduced on its own, so this is the
    // 12-wide sum in half the LDS)
    double lo[6], hi[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) lo[i] = b[i], hi[i] = b[6 + i];
    wg_sum_n<NT, 6>(lo, shw);
    wg_sum_n<NT, 6>(hi, shw);
#pragma unroll
    for (int i = 0; i < 6; ++i) b[i] = lo[i], b[6 + i] = hi[i];
  }
  if (threadIdx.x == 0) restrict_finish(Lp.bodies[k], M, com, cdot, b);
}

// rigidSymplecticImperativeStep (RigidBody3D.m:174-179) of one body.
__device__ __noinline__ void body_step_one(Body &Bd, double h) {
  for (int d = 0; d < 3; ++d) Bd.COM[d] = Bd.COM[d] + h * Bd.COMDOT[d];
  double dR[9], Rn[9], tmp[9], Rt[9];
  exp_rodrigues(Bd.omega, h, dR);
  matmul3(dR, Bd.R, Rn);
  for (int i = 0; i < 9; ++i) Bd.R[i] = Rn[i];
  matmul3(Bd.R, Bd.I0, tmp);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Rt[i + 3 * j] = Bd.R[j + 3 * i];
  matmul3(tmp, Rt, Bd.I);
  for (int i = 0; i < 9; ++i) Bd.Rstep[i] = Bd.R[i];
}

// Boundary accumulation of body k (rigidElasticBoundaryAccumulation-
// Constraints, xpbdLayer3D.m:381-420): its boundary vertices projected
// against the body's pose, the sums applied by thread 0, the vertices
// snapped to the new pose.
template <typename T, int NT, typename Pos>
__device__ __forceinline__ void wg_boundary(const LayeredParams<T> &Lp, int kb, const Pos &x,
                                         const double *minv, double h, double *shw,
                                         double *sR, double *sC) {
  const Body &Bd = Lp.bodies[kb];
  const int off = Lp.bndOff[kb], n = Lp.bndOff[kb + 1] - off;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int q = threadIdx.x; q < n; q += NT) {
    const int v = Lp.bndVert[off + q];
    boundary_project_at<T>(Bd, v, x.load(v), minv, Lp.ri, Lp.lamB + off + q,
                           Lp.boundaryCompliance, h, acc);
  }
  wg_sum_n<NT, 6>(acc, shw);
  if (threadIdx.x == 0) {
    body_apply(Lp.bodies[kb], acc, acc + 3, h);
    for (int i = 0; i < 9; ++i) sR[i] = Lp.bodies[kb].R[i];
    for (int i = 0; i < 3; ++i) sC[i] = Lp.bodies[kb].COM[i];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += NT) {
    // boundary_snap through the accessor
    const int v = Lp.bndVert[off + q];
    double xn[3];
    rigid_pos(sR, Lp.ri + 3 * v, sC, xn);
    vec4_t<T> p = x.load(v);
    p.x = (T)xn[0];
    p.y = (T)xn[1];
    p.z = (T)xn[2];
    x.store(v, p);
  }
  __syncthreads();
}

// Rigid vertices of the scene: after the step, to R r_i + COM with residual
// R v - (x_new - x)/h (stepped = true); after the layer, to R r_i + COM with
// the residual rotated by R_projected R_stepped^T (xpbdLayer3D.m:217-241).
template <typename T, int NT, typename Pos>
__device__ __forceinline__ void wg_rigid_vertices(const LayeredParams<T> &Lp, int64_t vb,
                                               int64_t ve, const Pos &x, double h,
                                               bool stepped) {
  T *rv = Lp.rv;
  for (int64_t v = vb + threadIdx.x; v < ve; v += NT) {
    const int bid = Lp.vbody[v];
    if (bid == 0) continue;
    const Body &Bd = Lp.bodies[bid - 1];
    vec4_t<T> p = x.load(v);
    double xn[3];
    rigid_pos(Bd.R, Lp.ri + 3 * v, Bd.COM, xn);
    const double prev[3] = {(double)p.x, (double)p.y, (double)p.z};
    p.x = (T)xn[0];
    p.y = (T)xn[1];
    p.z = (T)xn[2];
    x.store(v, p);
    const double vr[3] = {(double)rv[3 * v], (double)rv[3 * v + 1], (double)rv[3 * v + 2]};
    const double zero[3] = {0, 0, 0};
    double o[3];
    if (stepped) {
      rigid_pos(Bd.R, vr, zero, o);
      const double now[3] = {(double)p.x, (double)p.y, (double)p.z};
      for (int d = 0; d < 3; ++d) o[d] = o[d] - (now[d] - prev[d]) / h;
    } else {
      double Rt[9], dR[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Rt[i + 3 * j] = Bd.Rstep[j + 3 * i];
      matmul3(Bd.R, Rt, dR);
      rigid_pos(dR, vr, zero, o);
    }
    rv[3 * v] = (T)o[0];
    rv[3 * v + 1] = (T)o[1];
    rv[3 * v + 2] = (T)o[2];
  }
  __syncthreads();
}

// One scene's two-layer layerSolve (positions in HBM).  The loop structure
// per layer is projectGS (xpbdLayer3D.m:315-341): colors (elastic slots
// only), boundary accumulation of every body (rigidElasticBoundary-
// AccumulationConstraints, :381-420), elastic contacts (:422-485; rigid
// contacts are off in this path), pins, residual and monitor.  Positions
// through x (HbmPos or LdsPos); the same arithmetic in the same order
// either way.
// The learned warm start of a scene's all-elastic layer, in its workgroup
// (the single-scene k_learn_* sequence on the scene's own slots, vertices
// and contacts): the equation residual of the cold start, copies of the
// positions and multipliers, lambda_0 and its displacement on every slot
// whose multipliers are still zero (each vertex adds its incident
// displacements in ascending slot order), the residual of the warm start,
// and the copies restored when the warm start is worse.  Ends with a
// workgroup barrier.
template <typename T, int NT, int MAT, typename Pos>
__device__ __forceinline__ void wg_warm_start(const BatchParams<T> &B, const LayeredParams<T> &Lp,
                                              const int sc, const Pos &x, double *shw) {
  const BatchLearn<T> &W = Lp.learn;
  const SweepParams<T> &P = B.P;
  const int tid = threadIdx.x;
  const int64_t ld = P.ld;
  const int64_t vb = B.vert0[sc], ve = B.vert0[sc + 1];
  const int64_t s0 = B.slot0[sc], s1 = B.slot0[sc + 1];
  const int64_t k0 = B.cont0[sc], k1 = B.cont0[sc + 1];
  auto eq_residual = [&]() -> double {
    double acc[1] = {0};
    for (int64_t s = s0 + tid; s < s1; s += NT) {
      if (P.vid[s] < 0) continue;
      TetIn<T> in;
      tet_fetch<T, MAT>(P, s, in);
      vec4_t<T> X[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) X[q] = x.load_w(in.v[q], in.O[q].w);
      acc[0] += xpbd_eq_residual_x<T>(in, X);
    }
    for (int64_t i = k0 + tid; i < k1; i += NT) {
      const vec4_t<T> p = x.load(B.c.vert[i]);
      const double *n = B.c.normal + 3 * i, *pc = B.c.point + 3 * i;
      double phi = n[0] * ((double)p.x - pc[0]) + n[1] * ((double)p.y - pc[1]) +
                   n[2] * ((double)p.z - pc[2]);
      phi = phi < 0 ? phi : 0;
      acc[0] += (B.contactResistance * phi) * (B.contactResistance * phi);
    }
    wg_sum_n<NT, 1>(acc, shw);
    return sqrt(acc[0]);
  };
  const double cold = eq_residual();
  for (int64_t v = vb + tid; v < ve; v += NT) W.xsave[v] = x.load(v);
  for (int64_t s = s0 + tid; s < s1; s += NT)
#pragma unroll
    for (int i = 0; i < 6; ++i) W.lamsave[i * ld + s] = P.lam[i * ld + s];
  for (int64_t s = s0 + tid; s < s1; s += NT) {
    bool fresh = P.vid[s] >= 0;
    if (fresh)
#pragma unroll
      for (int i = 0; i < 6; ++i) fresh = fresh && P.lam[i * ld + s] == T(0);
    if (!fresh) {
#pragma unroll
      for (int k = 0; k < 12; ++k) W.dx[k * ld + s] = T(0);
      continue;
    }
    vec4_t<T> X[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) X[q] = x.load(P.vid[q * ld + s]);
    learn_disp_slot<T>(P, s, W.lam0, W.dx, X);
  }
  __syncthreads();
  for (int64_t v = vb + tid; v < ve; v += NT) {
    if (W.pinMask[3 * v]) continue;
    T d[3] = {T(0), T(0), T(0)};
    for (int q = W.incOff[v]; q < W.incOff[v + 1]; ++q) {
      const int32_t e = W.inc[q];
      const int64_t a = e / ld, s = e % ld;
#pragma unroll
      for (int c = 0; c < 3; ++c) d[c] += W.dx[(3 * a + c) * ld + s];
    }
    vec4_t<T> p = x.load(v);
    p.x += d[0];
    p.y += d[1];
    p.z += d[2];
    x.store(v, p);
  }
  __syncthreads();
  const double warm = eq_residual();
  const bool keep = warm <= cold;  // false for a NaN warm residual
  if (tid == 0) atomicAdd(W.counts + (keep ? 0 : 1), 1ull);
  if (!keep) {
    for (int64_t v = vb + tid; v < ve; v += NT) x.store(v, W.xsave[v]);
    for (int64_t s = s0 + tid; s < s1; s += NT)
#pragma unroll
      for (int i = 0; i < 6; ++i) P.lam[i * ld + s] = W.lamsave[i * ld + s];
  }
  __syncthreads();
}

template <typename T, int NT, int MAT, typename Pos, bool LEARN>
__device__ __forceinline__ void batch_scene_layered(const BatchParams<T> &B,
                                                    const LayeredParams<T> &Lp, const int sc,
                                                    const Pos &x, LayeredShared<NT> &S) {
  double *shw = S.shw;
  int &sStop = S.sStop, &sStopSolve = S.sStopSolve;
  double &sRunUntil = S.sRunUntil;
  double(&sR)[kMaxSceneBodies][9] = S.sR;
  double(&sC)[kMaxSceneBodies][3] = S.sC;
  const int tid = threadIdx.x;
  const int64_t vb = B.vert0[sc], ve = B.vert0[sc + 1];
  const int b0 = Lp.body0[sc], nbS = Lp.body0[sc + 1] - b0;
  T *rv = Lp.rv;
  const double *mass = Lp.mass;
  Ctl &c = S.c;
  if (tid == 0) {
    ctl_start(c, B.runUntil0, 0.0, 0);
    sStop = 0;
    sStopSolve = 0;
    sRunUntil = B.runUntil0;
  }
  __syncthreads();
  const int nc = B.ncolor[sc], cb = B.color0[sc];
  const int64_t k0 = B.cont0[sc], k1 = B.cont0[sc + 1];
  const int64_t q0 = B.pin0[sc], q1 = B.pin0[sc + 1];
  const int64_t s0 = B.slot0[sc], s1 = B.slot0[sc + 1];
  const double h = B.h, h2 = h * h;
  int total = 0;
  for (int layer = 0; layer < 2; ++layer) {
    const bool coarse = layer == 0;
    const int nb = coarse ? nbS : 0;
    // restriction of the layer's bodies from the positions and residual
    // velocities at the layer's start
    for (int k = 0; k < nb; ++k) wg_restrict<T, NT>(Lp, b0 + k, x, mass, shw);
    // release of the elastic vertices (:138-139)
    for (int64_t v = vb + tid; v < ve; v += NT) {
      if (coarse && Lp.vbody[v] != 0) continue;
      vec4_t<T> p = x.load(v);
      p.x = p.x + T(h) * rv[3 * v];
      p.y = p.y + T(h) * rv[3 * v + 1];
      p.z = p.z + T(h) * rv[3 * v + 2];
      rv[3 * v] = T(0);
      rv[3 * v + 1] = T(0);
      rv[3 * v + 2] = T(0);
      x.store(v, p);
    }
    __syncthreads();
    if (nb > 0) {
      if (tid < nb) body_step_one(Lp.bodies[b0 + tid], h);
      __syncthreads();
      wg_rigid_vertices<T, NT>(Lp, vb, ve, x, h, true);
    }
    // the layer's normresv and start (k_layer_begin); contact and boundary
    // multipliers restart at zero
    double nr[1] = {0};
    for (int64_t v = vb + tid; v < ve; v += NT)
      for (int d = 0; d < 3; ++d) {
        const double r = (double)rv[3 * v + d];
        nr[0] += r * r;
      }
    wg_sum_n<NT, 1>(nr, shw);
    if (tid == 0) {
      c.normresv = sqrt(nr[0]);
      c.endLayer = 0;
      c.iters = 0;
      c.layer = layer;
      c.isLastLayer = !coarse;
      c.setPercentStop = coarse ? B.stopPct : 0.0;
    }
    for (int64_t i = k0 + tid; i < k1; i += NT) B.c.lambda[i] = 0;
    if (nb > 0)
      for (int q = Lp.bndOff[b0] + tid; q < Lp.bndOff[b0 + nb]; q += NT) Lp.lamB[q] = 0;
    if (tid == 0) sStop = 0;
    __syncthreads();
    if constexpr (LEARN)
      if (!coarse) wg_warm_start<T, NT, MAT>(B, Lp, sc, x, shw);
    SweepParams<T> P = B.P;
    P.flags = nb > 0 ? Lp.flags0 : nullptr;
    const int iters = coarse ? Lp.it0 : Lp.it1;
    int it = 0;
    while (it < B.hardCap) {
      for (int k = 0; k < nc; ++k) {
        const int64_t a = B.colorOff[cb + k], e = a + B.colorCnt[cb + k];
        scene_color<T, MAT, kScenePairs<T, NT>>(P, a + tid, e, NT, x);
        __syncthreads();
      }
      for (int k = 0; k < nb; ++k)
        wg_boundary<T, NT>(Lp, b0 + k, x, B.minv, h, shw, sR[k], sC[k]);
      for (int64_t i = k0 + tid; i < k1; i += NT) {
        const int v = B.c.vert[i];
        if (nb > 0 && Lp.vbody[v] != 0) continue;  // rigid contact: not projected
        const double *n = B.c.normal + 3 * i, *pc = B.c.point + 3 * i;
        vec4_t<T> p = x.load(v);
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
        x.store(v, p);
      }
      __syncthreads();
      for (int64_t q = q0 + tid; q < q1; q += NT) {
        const int d = B.pinDof[q];
        x.store_comp(d / 3, d % 3, B.pinVal[q]);
      }
      __syncthreads();
      ++it;
      if (B.mp.computeResiduals) {
        double acc[1] = {scene_residual<T, MAT, kScenePairs<T, NT>>(B.P, s0 + tid, s1, NT, x)};
        const double Ks = (double)(k1 - k0);
        for (int64_t i = k0 + tid; i < k1; i += NT) {
          const int v = B.c.vert[i];
          const double *n = B.c.normal + 3 * i, *pc = B.c.point + 3 * i;
          const vec4_t<T> p = x.load(v);
          double phi = n[0] * ((double)p.x - pc[0]) + n[1] * ((double)p.y - pc[1]) +
                       n[2] * ((double)p.z - pc[2]);
          if (phi > 0) phi = 0;
          const double a = fabs(B.contactResistance * phi);
          const double cr = fabs((B.contactCompliance / h2) * B.c.lambda[i]);
          const double t = a - (a < cr ? a : cr);
          acc[0] += Ks * (t * t);
        }
        wg_sum_n<NT, 1>(acc, shw);
        if (tid == 0) {
          monitor_update(c, B.mp, sqrt(acc[0]), nullptr, 0);
          sStop = c.stopSolve | c.endLayer;
          sStopSolve = c.stopSolve;
          sRunUntil = c.runUntil;
        }
        __syncthreads();
      }
      const int stop = sStop;
      const double ru = sRunUntil;
      __syncthreads();
      if (stop) break;
      // while it <= iterations || (~stopSolve && runUntil >= 0 && isLastLayer)
      if (!(it < iters || (!coarse && ru >= 0))) break;
    }
    total += it;
    if (coarse && tid == 0) Lp.coarseIters[sc] = it;
    if (nb > 0) wg_rigid_vertices<T, NT>(Lp, vb, ve, x, h, false);
    const int stopSolve = sStopSolve;
    __syncthreads();
    if (stopSolve) break;
  }
  if (tid == 0) {
    B.iters[sc] = total;
    B.lastResidual[sc] = c.lastResidual;
    B.runUntilOut[sc] = c.runUntil;
    B.residualCount[sc] = c.residualCount;
  }
}

template <typename T, int NT, int MAT, bool LEARN>
__device__ __forceinline__ void batch_scene_layered_at(const BatchParams<T> &B,
                                                       const LayeredParams<T> &Lp, const int sc,
                                                       T *lds, LayeredShared<NT> &shw) {
  const int64_t vb = B.vert0[sc];
  const int nv = (int)(B.vert0[sc + 1] - vb);
  if (nv > B.ldsVerts) {
    batch_scene_layered<T, NT, MAT, HbmPos<T>, LEARN>(B, Lp, sc, HbmPos<T>{B.P.x}, shw);
    return;
  }
  // the scene's positions in LDS from the start to the end of the substep
  for (int i = threadIdx.x; i < nv; i += NT) {
    const vec4_t<T> p = B.P.x[vb + i];
    lds[i] = p.x;
    lds[nv + i] = p.y;
    lds[2 * nv + i] = p.z;
  }
  __syncthreads();
  batch_scene_layered<T, NT, MAT, LdsPos<T>, LEARN>(B, Lp, sc, LdsPos<T>{lds, vb, nv, B.P.xold},
                                                   shw);
  __syncthreads();
  for (int i = threadIdx.x; i < nv; i += NT) {
    vec4_t<T> p = B.P.x[vb + i];
    p.x = lds[i];
    p.y = lds[nv + i];
    p.z = lds[2 * nv + i];
    B.P.x[vb + i] = p;
  }
}

// One scene per workgroup; a scene with at most B.ldsVerts vertices keeps
// its positions in the dynamic LDS (3 values of T per vertex).  LEARN: the
// all-elastic layer starts from the learned warm start (Lp.learn).
template <typename T, int NT, bool LEARN = false>
__global__ __launch_bounds__(NT) void k_batch_layered(BatchParams<T> B, LayeredParams<T> Lp) {
  extern __shared__ __align__(16) unsigned char mlx_dyn_lds[];
  __shared__ LayeredShared<NT> shw;
  const int sc = B.sceneList ? B.sceneList[blockIdx.x] : (int)blockIdx.x;
  T *lds = reinterpret_cast<T *>(mlx_dyn_lds);
  if (B.sceneNH && B.sceneNH[sc])
    batch_scene_layered_at<T, NT, kMatNH, LEARN>(B, Lp, sc, lds, shw);
  else
    batch_scene_layered_at<T, NT, kMatStVK, LEARN>(B, Lp, sc, lds, shw);
}

}  // namespace mlx
