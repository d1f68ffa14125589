/*
 * mlx_oracle.c -- CPU FP64 restatement of the multi-layer XPBD substep of
 * AlexandreMercierAubin/MultiLayerSolver2024 (a Matlab code base).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline.
 * The product (HIP) path never links or calls this file.
 *
 * Every function names the reference file:line it restates.  Where the
 * reference is vectorised Matlab over "pages" of elements, the restatement
 * loops over the same elements in the same (ascending) order; elements of
 * one graph color share no vertex, so the loop order inside a color does not
 * change any result.
 *
 * Parity pins (how each function is checked against something other than
 * itself) are listed per function; a function without one says
 * "parity unpinned".  Build with -O2 -ffp-contract=off.
 */
#include "mlx_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* small dense helpers                                                  */
/* ------------------------------------------------------------------ */

/* Solve A x = b for an n x n column-major A (n <= 6) by Gaussian elimination
 * with partial pivoting, which is what Matlab's mldivide/pagemldivide does
 * for a square non-triangular page (LAPACK getrf/getrs).  A and b are
 * copied.  Returns 0, or 1 for an exactly singular pivot (the result is then
 * Inf/NaN as in Matlab). */
/* Threads of the element loops (ora_set_threads; 1 = serial). */
static int g_threads = 1;

void ora_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

static int solve_n(int n, const double *A_in, const double *b_in, double *x) {
  double A[36], b[6];
  memcpy(A, A_in, sizeof(double) * n * n);
  memcpy(b, b_in, sizeof(double) * n);
  int sing = 0;
  for (int k = 0; k < n; ++k) {
    int piv = k;
    double best = fabs(A[k + n * k]);
    for (int i = k + 1; i < n; ++i) {
      double v = fabs(A[i + n * k]);
      if (v > best) { best = v; piv = i; }
    }
    if (piv != k) {
      for (int j = 0; j < n; ++j) {
        double t = A[k + n * j]; A[k + n * j] = A[piv + n * j]; A[piv + n * j] = t;
      }
      double t = b[k]; b[k] = b[piv]; b[piv] = t;
    }
    double d = A[k + n * k];
    if (d == 0.0) sing = 1;
    for (int i = k + 1; i < n; ++i) {
      double l = A[i + n * k] / d;
      A[i + n * k] = l;
      for (int j = k + 1; j < n; ++j) A[i + n * j] = A[i + n * j] - l * A[k + n * j];
      b[i] = b[i] - l * b[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s = s - A[i + n * j] * x[j];
    x[i] = s / A[i + n * i];
  }
  return sing;
}

/* Inverse of an n x n column-major matrix (Matlab inv/pageinv). */
static int inv_n(int n, const double *A, double *Ainv) {
  int sing = 0;
  double e[6], col[6];
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) e[i] = (i == j) ? 1.0 : 0.0;
    sing |= solve_n(n, A, e, col);
    for (int i = 0; i < n; ++i) Ainv[i + n * j] = col[i];
  }
  return sing;
}

/* C = A * B, n x n column-major, k summed in ascending order. */
static void matmul_n(int n, const double *A, const double *B, double *C) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = A[i] * B[n * j];
      for (int k = 1; k < n; ++k) s = s + A[i + n * k] * B[k + n * j];
      C[i + n * j] = s;
    }
}

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* ------------------------------------------------------------------ */
/* precompute                                                           */
/* ------------------------------------------------------------------ */

/* naiveGraphColoring.m:1-19 -- greedy, in element order: each element takes
 * the first color (1, 2, ...) that none of its vertices has used yet.
 * Pins: no two same-colored elements share a vertex; brute-force check of the
 * "first free color" definition on small meshes (tests). */
int ora_naive_graph_coloring(const int32_t *T, int nT, int vpe, int nV,
                             int32_t *color) {
  const int W = 16; /* up to 1024 colors */
  uint64_t *used = (uint64_t *)calloc((size_t)nV * W, sizeof(uint64_t));
  if (!used) return -1;
  for (int e = 0; e < nT; ++e) {
    int c = 1;
    for (;;) {
      int taken = 0;
      for (int k = 0; k < vpe; ++k) {
        int v = T[e * vpe + k];
        if (used[(size_t)v * W + (c - 1) / 64] >> ((c - 1) % 64) & 1ull) { taken = 1; break; }
      }
      if (!taken) break;
      ++c;
      if (c > 64 * W) { free(used); return -2; }
    }
    for (int k = 0; k < vpe; ++k) {
      int v = T[e * vpe + k];
      used[(size_t)v * W + (c - 1) / 64] |= 1ull << ((c - 1) % 64);
    }
    color[e] = c;
  }
  free(used);
  return 0;
}

/* toLame.m:1-6.  Pin: closed form (tests compare mu, lambda with E, nu). */
void ora_to_lame(double nu, double E, double *mu, double *lambda) {
  *mu = E / (2 * (1 + nu));
  *lambda = (E * nu) / ((1 + nu) * (1 - 2 * nu));
}

/* computeB3D.m:18-29 (Dm, DmInv), makeElements3D.m:11 (volume through
 * gptoolbox volume(), restated as det(Dm)/6, which equals gptoolbox's
 * -dot(a-d,cross(b-d,c-d))/6 for T=[a b c d]; gptoolbox is not vendored, so
 * that identity is parity unpinned against the library itself),
 * computeMass.m:50-65 (lumped mass elV*rho/4 per vertex DOF),
 * Mesh3D.m:445-460 (K, alpha = inv(K), beta = alpha1*K).
 * Pins: DmInv*Dm = I, sum of volumes = box volume, sum of mass = rho*V,
 * alpha*K = I (tests). */
int ora_precompute3d(int nV, int nT, const double *p, const int32_t *T,
                     const double *rho, const double *mu, const double *lam,
                     const double *alpha1, double *DmInv, double *vol,
                     double *mass, double *alpha, double *beta) {
  int bad = 0;
  memset(mass, 0, sizeof(double) * 3 * nV);
  for (int e = 0; e < nT; ++e) {
    const double *p0 = p + 3 * T[4 * e], *p1 = p + 3 * T[4 * e + 1];
    const double *p2 = p + 3 * T[4 * e + 2], *p3 = p + 3 * T[4 * e + 3];
    double Dm[9];
    for (int d = 0; d < 3; ++d) {
      Dm[d + 0] = p1[d] - p0[d];
      Dm[d + 3] = p2[d] - p0[d];
      Dm[d + 6] = p3[d] - p0[d];
    }
    bad |= inv_n(3, Dm, DmInv + 9 * e);
    double det = Dm[0] * (Dm[4] * Dm[8] - Dm[7] * Dm[5]) -
                 Dm[3] * (Dm[1] * Dm[8] - Dm[7] * Dm[2]) +
                 Dm[6] * (Dm[1] * Dm[5] - Dm[4] * Dm[2]);
    vol[e] = det / 6.0;
  }
  for (int e = 0; e < nT; ++e) {
    double mv = (vol[e] * rho[e]) / 4;
    for (int k = 0; k < 4; ++k) {
      int v = T[4 * e + k];
      for (int d = 0; d < 3; ++d) mass[3 * v + d] = mass[3 * v + d] + mv;
    }
  }
  for (int e = 0; e < nT; ++e) {
    double l = lam[e], diagMu = 2 * mu[e], diagVal = l + 2 * diagMu;
    double K[36];
    memset(K, 0, sizeof K);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) K[i + 6 * j] = (i == j) ? diagVal : l;
    for (int i = 3; i < 6; ++i) K[i + 6 * i] = diagMu;
    for (int k = 0; k < 36; ++k) beta[36 * e + k] = K[k] * alpha1[e];
    bad |= inv_n(6, K, alpha + 36 * e);
  }
  return bad;
}

/* Mesh.m:439-469 (computeB2D: Dm = [p0-p2, p1-p2], DmInv = inv(Dm)),
 * Mesh.m:481-492 (area = |det([pj-pi, pk-pi])|/2), @Mesh/computeMass.m:1-18
 * (mass (1/3)*area*rho per vertex DOF), Mesh.m:247-258 (K, alpha, beta).
 * Pins: DmInv*Dm = I, sum of areas = plate area, alpha*K = I (tests). */
int ora_precompute2d(int nV, int nT, const double *p, const int32_t *T,
                     const double *rho, const double *mu, const double *lam,
                     const double *alpha1, double *DmInv, double *area,
                     double *mass, double *alpha, double *beta) {
  int bad = 0;
  memset(mass, 0, sizeof(double) * 2 * nV);
  for (int e = 0; e < nT; ++e) {
    const double *pi = p + 2 * T[3 * e], *pj = p + 2 * T[3 * e + 1];
    const double *pk = p + 2 * T[3 * e + 2];
    double Dm[4] = {pi[0] - pk[0], pi[1] - pk[1], pj[0] - pk[0], pj[1] - pk[1]};
    bad |= inv_n(2, Dm, DmInv + 4 * e);
    double a0 = pj[0] - pi[0], a1 = pj[1] - pi[1];
    double b0 = pk[0] - pi[0], b1 = pk[1] - pi[1];
    area[e] = 1.0 / 2 * fabs(a0 * b1 - b0 * a1);
  }
  for (int e = 0; e < nT; ++e)
    for (int j = 0; j < 3; ++j) {
      int v = T[3 * e + j];
      mass[2 * v] = mass[2 * v] + (1.0 / 3) * area[e] * rho[e];
      mass[2 * v + 1] = mass[2 * v + 1] + (1.0 / 3) * area[e] * rho[e];
    }
  for (int e = 0; e < nT; ++e) {
    double l = lam[e], diagVal = l + 2 * mu[e];
    double K[9] = {diagVal, l, 0, l, diagVal, 0, 0, 0, 2 * mu[e]};
    for (int k = 0; k < 9; ++k) beta[9 * e + k] = K[k] * alpha1[e];
    bad |= inv_n(3, K, alpha + 9 * e);
  }
  return bad;
}

/* @Mesh3D/pin.m:6-10 and @Mesh/pin.m:7-10: pinnedTets = unique(
 * TetsPerParticle{pinnedInds}), i.e. every element touching a pinned
 * vertex (elementAdjacencyMatrix.m:26-35).  Returns the count. */
int ora_pinned_elements(const int32_t *T, int nT, int vpe, int nV,
                        const uint8_t *pinnedVert, uint8_t *forcedElastic) {
  (void)nV;
  int n = 0;
  for (int e = 0; e < nT; ++e) {
    int f = 0;
    for (int k = 0; k < vpe; ++k) f |= pinnedVert[T[e * vpe + k]] != 0;
    forcedElastic[e] = (uint8_t)f;
    n += f;
  }
  return n;
}

/* ------------------------------------------------------------------ */
/* layer building                                                       */
/* ------------------------------------------------------------------ */

/* Fa = mesh3D.B * p (xpbdLayer3D.m:606) with B from computeB3D.m:31-41:
 * B(9e+3d+j, 3*t_a+d) = D(a,j), D = [-sum(DmInv,1); DmInv], so
 * F[9e+3d+j] = sum_a D(a,j) p[3 t_a + d] -- the deformation gradient stored
 * row by row.  The sparse product accumulates in ascending column order,
 * i.e. in ascending vertex index; that order is kept here.
 * Pins: closed form F = Ds*DmInv (tests, numpy). */
void ora_b_times_p3d(const int32_t *T, int nT, const double *DmInv,
                     const double *p, double *F) {
  for (int e = 0; e < nT; ++e) {
    const double *Mi = DmInv + 9 * e;
    double D[4][3];
    for (int j = 0; j < 3; ++j) {
      D[0][j] = -((Mi[0 + 3 * j] + Mi[1 + 3 * j]) + Mi[2 + 3 * j]);
      for (int a = 1; a < 4; ++a) D[a][j] = Mi[(a - 1) + 3 * j];
    }
    int ord[4] = {0, 1, 2, 3};
    for (int i = 1; i < 4; ++i)
      for (int k = i; k > 0 && T[4 * e + ord[k - 1]] > T[4 * e + ord[k]]; --k) {
        int t = ord[k]; ord[k] = ord[k - 1]; ord[k - 1] = t;
      }
    for (int d = 0; d < 3; ++d)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int q = 0; q < 4; ++q) {
          int a = ord[q];
          s = s + D[a][j] * p[3 * T[4 * e + a] + d];
        }
        F[9 * e + 3 * d + j] = s;
      }
  }
}

/* Fa = mesh2d.B * p (xpbdLayer2D.m:770) with B from Mesh.m:458-467 and
 * mult2x2Stack.m:1-8: F = Ds*DmInv, Ds = [p0-p2, p1-p2], stored column-major
 * F[4e + i + 2j].  Entries of one row: p0 -> DmInv(0,j), p1 -> DmInv(1,j),
 * p2 -> -(DmInv(0,j)+DmInv(1,j)); accumulated in ascending column order. */
void ora_b_times_p2d(const int32_t *T, int nT, const double *DmInv,
                     const double *p, double *F) {
  for (int e = 0; e < nT; ++e) {
    const double *Mi = DmInv + 4 * e;
    double D[3][2];
    for (int j = 0; j < 2; ++j) {
      D[0][j] = Mi[0 + 2 * j];
      D[1][j] = Mi[1 + 2 * j];
      D[2][j] = (-1 * Mi[0 + 2 * j]) + (-1 * Mi[1 + 2 * j]);
    }
    int ord[3] = {0, 1, 2};
    for (int i = 1; i < 3; ++i)
      for (int k = i; k > 0 && T[3 * e + ord[k - 1]] > T[3 * e + ord[k]]; --k) {
        int t = ord[k]; ord[k] = ord[k - 1]; ord[k - 1] = t;
      }
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        double s = 0.0;
        for (int q = 0; q < 3; ++q) {
          int a = ord[q];
          s = s + D[a][j] * p[2 * T[3 * e + a] + i];
        }
        F[4 * e + i + 2 * j] = s;
      }
  }
}

static inline double sq(double x) { return x * x; }

/* mexEdiffNorm3D.cpp:35-61, same expression tree (pow(x,2.0) written x*x).
 * Pins: closed form ||(Ga - Gb)/2||_F^2 / h^2 with G = M M^T for the 3x3 M
 * read column-major from the 9 values (tests, numpy). */
void ora_ediff_norm3d(const double *Fa, const double *Fb, int nT, double h,
                      double *out) {
  for (int i = 0; i < nT; ++i) {
    const double *a = Fa + 9 * i, *b = Fb + 9 * i;
    double a11 = a[0], a21 = a[1], a31 = a[2], a12 = a[3], a22 = a[4], a32 = a[5],
           a13 = a[6], a23 = a[7], a33 = a[8];
    double b11 = b[0], b21 = b[1], b31 = b[2], b12 = b[3], b22 = b[4], b32 = b[5],
           b13 = b[6], b23 = b[7], b33 = b[8];
    double t2 = 1.0 / (h * h);
    double t0 =
        t2 * sq((a11 * a12) / 2.0 + (a21 * a22) / 2.0 + (a31 * a32) / 2.0 - (b11 * b12) / 2.0 - (b21 * b22) / 2.0 - (b31 * b32) / 2.0) * 2.0 +
        t2 * sq((a11 * a13) / 2.0 + (a21 * a23) / 2.0 + (a31 * a33) / 2.0 - (b11 * b13) / 2.0 - (b21 * b23) / 2.0 - (b31 * b33) / 2.0) * 2.0 +
        t2 * sq((a12 * a13) / 2.0 + (a22 * a23) / 2.0 + (a32 * a33) / 2.0 - (b12 * b13) / 2.0 - (b22 * b23) / 2.0 - (b32 * b33) / 2.0) * 2.0 +
        t2 * sq((a11 * a11) / 2.0 + (a21 * a21) / 2.0 + (a31 * a31) / 2.0 - (b11 * b11) / 2.0 - (b21 * b21) / 2.0 - (b31 * b31) / 2.0) +
        t2 * sq((a12 * a12) / 2.0 + (a22 * a22) / 2.0 + (a32 * a32) / 2.0 - (b12 * b12) / 2.0 - (b22 * b22) / 2.0 - (b32 * b32) / 2.0) +
        t2 * sq((a13 * a13) / 2.0 + (a23 * a23) / 2.0 + (a33 * a33) / 2.0 - (b13 * b13) / 2.0 - (b23 * b23) / 2.0 - (b33 * b33) / 2.0);
    out[i] = t0;
  }
}

/* mexEdiffNorm2D.cpp:35-52.  Pin: closed form (tests). */
void ora_ediff_norm2d(const double *Fa, const double *Fb, int nT, double h,
                      double *out) {
  for (int i = 0; i < nT; ++i) {
    const double *A = Fa + 4 * i, *B = Fb + 4 * i;
    double a = A[0] * A[0] + A[1] * A[1] - B[0] * B[0] - B[1] * B[1];
    double b = A[0] * A[2] + A[1] * A[3] - B[0] * B[2] - B[1] * B[3];
    double c = A[2] * A[2] + A[3] * A[3] - B[2] * B[2] - B[3] * B[3];
    out[i] = (0.25 * a * a + 0.5 * b * b + 0.25 * c * c) / h / h;
  }
}

/* Matlab sort(x,"ascend") index output (xpbdLayer3D.m:609): stable, NaN last.
 * Merge sort.  Pin: numpy argsort(kind="stable") (tests). */
static int key_less(double a, double b) {
  if (isnan(a)) return 0;
  if (isnan(b)) return 1;
  return a < b;
}
void ora_stable_sort_ascend(const double *keys, int n, int32_t *order) {
  int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) order[i] = i;
  for (int w = 1; w < n; w *= 2) {
    for (int lo = 0; lo < n; lo += 2 * w) {
      int mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
      int i = lo, j = mid, k = lo;
      while (i < mid && j < hi) {
        if (key_less(keys[order[j]], keys[order[i]])) tmp[k++] = order[j++];
        else tmp[k++] = order[i++];
      }
      while (i < mid) tmp[k++] = order[i++];
      while (j < hi) tmp[k++] = order[j++];
    }
    memcpy(order, tmp, sizeof(int32_t) * n);
  }
  free(tmp);
}

/* mexUnionFindByVertice.cpp:11-178, restated: union by size with path
 * compression; each inserted element merges with the last element that
 * claimed each of its vertices; per layer, body ids are handed out in
 * ascending element order (copyLayer :86-112).
 * order: element indices (0-based) in strain-rate order; sets: nT*L
 * column-major body ids (0 = elastic); numRigids: L.
 * Pins: bodies equal the vertex-sharing connected components of the inserted
 * elements, numbered by ascending minimum element (tests, BFS in Python). */
typedef struct { int *id, *sz, *claimed, count; } uf_t;
static int uf_find(uf_t *u, int p) {
  if (p == -1 || u->id[p] == -1) return -1;
  int root = p;
  while (root != u->id[root]) root = u->id[root];
  while (p != root) { int np = u->id[p]; u->id[p] = root; p = np; }
  return root;
}
static void uf_merge(uf_t *u, int x, int y) {
  int i = uf_find(u, x), j = uf_find(u, y);
  if (i == j || i == -1 || j == -1) return;
  if (u->sz[i] < u->sz[j]) { u->id[i] = j; u->sz[j] += u->sz[i]; }
  else { u->id[j] = i; u->sz[i] += u->sz[j]; }
  u->count--;
}
static void uf_copy_layer(uf_t *u, int nT, int layer, int32_t *sets,
                          int32_t *numRigids) {
  int *map = (int *)malloc(sizeof(int) * nT);
  for (int i = 0; i < nT; ++i) map[i] = 0;
  int counter = 1;
  numRigids[layer] = u->count;
  for (int i = 0; i < nT; ++i) {
    int32_t *dst = sets + (size_t)nT * layer + i;
    if (u->id[i] == -1) { *dst = 0; continue; }
    int root = uf_find(u, i);
    if (map[root] == 0) map[root] = counter++;
    *dst = map[root];
  }
  free(map);
}
int ora_union_find_by_vertice(const int32_t *order, int nT,
                              const double *sortedLayers, int L,
                              const uint8_t *forcedElastic, const int32_t *T,
                              int vpe, int nV, int32_t *sets,
                              int32_t *numRigids) {
  uf_t u;
  u.id = (int *)malloc(sizeof(int) * (nT > 0 ? nT : 1));
  u.sz = (int *)malloc(sizeof(int) * (nT > 0 ? nT : 1));
  u.claimed = (int *)malloc(sizeof(int) * (nV > 0 ? nV : 1));
  u.count = 0;
  for (int i = 0; i < nT; ++i) { u.id[i] = -1; u.sz[i] = 1; }
  for (int i = 0; i < nV; ++i) u.claimed[i] = -1;
  uf_copy_layer(&u, nT, 0, sets, numRigids);
  int layerStart = 0;
  for (int layer = 0; layer < L; ++layer) {
    int layerEnd = (int)floor((sortedLayers[layer] / 100.0) * nT);
    for (int e = layerStart; e < layerEnd; ++e) {
      int el = order[e];
      if (!forcedElastic[el]) {
        u.id[el] = el;
        u.count += 1;
        for (int k = 0; k < vpe; ++k) {
          int v = T[el * vpe + k];
          uf_merge(&u, el, u.claimed[v]);
          u.claimed[v] = el;
        }
      }
    }
    uf_copy_layer(&u, nT, layer, sets, numRigids);
    layerStart = layerEnd;
  }
  free(u.id); free(u.sz); free(u.claimed);
  return 0;
}

/* xpbdLayer2D.findBoundaryVertices (xpbdLayer2D.m:551-569), used by 3D too
 * (xpbdLayer3D.m:652).  Column l of each output belongs to sorted layer l.
 * Pins: definitions (tests). */
void ora_find_boundary_vertices(int nV, int nT, int vpe, const int32_t *T,
                                int L, const int32_t *sets,
                                uint8_t *isBoundary, uint8_t *isRigidVert,
                                int32_t *vertexRigidBody) {
  uint8_t *ofElastic = (uint8_t *)malloc(nV > 0 ? nV : 1);
  for (int l = 0; l < L; ++l) {
    uint8_t *rv = isRigidVert + (size_t)nV * l;
    int32_t *vb = vertexRigidBody + (size_t)nV * l;
    memset(rv, 0, nV);
    memset(ofElastic, 0, nV);
    for (int v = 0; v < nV; ++v) vb[v] = 0;
    for (int k = 0; k < vpe; ++k)
      for (int e = 0; e < nT; ++e) {
        int v = T[e * vpe + k];
        int32_t b = sets[(size_t)nT * l + e];
        if (b != 0) { rv[v] = 1; vb[v] = b; }
        else ofElastic[v] = 1;
      }
    for (int v = 0; v < nV; ++v) isBoundary[(size_t)nV * l + v] = rv[v] && ofElastic[v];
  }
  free(ofElastic);
}

/* setLayers (xpbdLayer3D.m:595-599): [sortedLayers,~,ic] = unique(layers).
 * map[i] = 0-based column of layers[i] in sorted.  Returns #sorted. */
int ora_sorted_layers(const double *layers, int n, double *sorted, int32_t *map) {
  int m = 0;
  for (int i = 0; i < n; ++i) {
    int dup = 0;
    for (int j = 0; j < m; ++j) if (sorted[j] == layers[i]) dup = 1;
    if (!dup) sorted[m++] = layers[i];
  }
  for (int i = 1; i < m; ++i)
    for (int k = i; k > 0 && sorted[k - 1] > sorted[k]; --k) {
      double t = sorted[k]; sorted[k] = sorted[k - 1]; sorted[k - 1] = t;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) if (sorted[j] == layers[i]) map[i] = j;
  return m;
}

/* ------------------------------------------------------------------ */
/* element constraints                                                  */
/* ------------------------------------------------------------------ */

/* StVKconstraintGrad3D.m:1-70 (the symbolic expression tree).  gradC is
 * 6x12 column-major (gradC[r + 6c]); C = [E11 E22 E33 E12 E13 E23] with
 * E = (F^T F - I)/2 and F built from (p0 - pk) differences.
 * Pins: central finite differences of C; C=0 at rest (tests). */
void ora_stvk_grad3d(const double *D, const double *x, double *gradC, double *C) {
  double D11 = D[0], D21 = D[1], D31 = D[2], D12 = D[3], D22 = D[4], D32 = D[5],
         D13 = D[6], D23 = D[7], D33 = D[8];
  double t2 = D11 + D21 + D31, t3 = D12 + D22 + D32, t4 = D13 + D23 + D33;
  double t5 = -x[3], t6 = -x[4], t7 = -x[5], t8 = -x[6], t9 = -x[7], t10 = -x[8],
         t11 = -x[9], t12 = -x[10], t13 = -x[11];
  double t14 = t5 + x[0], t15 = t6 + x[1], t16 = t8 + x[0], t17 = t7 + x[2],
         t18 = t9 + x[1], t19 = t11 + x[0], t20 = t10 + x[2], t21 = t12 + x[1],
         t22 = t13 + x[2];
  double t23 = D11 * t14, t24 = D12 * t14, t25 = D11 * t15, t26 = D13 * t14,
         t27 = D12 * t15, t28 = D11 * t17, t29 = D13 * t15, t30 = D12 * t17,
         t31 = D13 * t17, t32 = D21 * t16, t33 = D22 * t16, t34 = D21 * t18,
         t35 = D23 * t16, t36 = D22 * t18, t37 = D21 * t20, t38 = D23 * t18,
         t39 = D22 * t20, t40 = D23 * t20, t41 = D31 * t19, t42 = D32 * t19,
         t43 = D31 * t21, t44 = D33 * t19, t45 = D32 * t21, t46 = D31 * t22,
         t47 = D33 * t21, t48 = D32 * t22, t49 = D33 * t22;
  double t50 = t23 + t32 + t41, t51 = t24 + t33 + t42, t52 = t25 + t34 + t43,
         t53 = t26 + t35 + t44, t54 = t27 + t36 + t45, t55 = t28 + t37 + t46,
         t56 = t29 + t38 + t47, t57 = t30 + t39 + t48, t58 = t31 + t40 + t49;
  const double mh = -1.0 / 2.0;
  double out[78] = {
      t2 * t50, t2 * t52, t2 * t55, -D11 * t50, -D11 * t52, -D11 * t55,
      -D21 * t50, -D21 * t52, -D21 * t55, -D31 * t50, -D31 * t52, -D31 * t55,
      t3 * t51, t3 * t54, t3 * t57, -D12 * t51, -D12 * t54, -D12 * t57,
      -D22 * t51, -D22 * t54, -D22 * t57, -D32 * t51, -D32 * t54, -D32 * t57,
      t4 * t53, t4 * t56, t4 * t58, -D13 * t53, -D13 * t56, -D13 * t58,
      -D23 * t53, -D23 * t56, -D23 * t58, -D33 * t53, -D33 * t56, -D33 * t58,
      (t2 * t51) / 2.0 + (t3 * t50) / 2.0, (t3 * t52) / 2.0 + (t2 * t54) / 2.0,
      (t3 * t55) / 2.0 + (t2 * t57) / 2.0,
      D11 * t51 * mh - (D12 * t50) / 2.0, D12 * t52 * mh - (D11 * t54) / 2.0,
      D12 * t55 * mh - (D11 * t57) / 2.0, D21 * t51 * mh - (D22 * t50) / 2.0,
      D22 * t52 * mh - (D21 * t54) / 2.0, D22 * t55 * mh - (D21 * t57) / 2.0,
      D31 * t51 * mh - (D32 * t50) / 2.0, D32 * t52 * mh - (D31 * t54) / 2.0,
      D32 * t55 * mh - (D31 * t57) / 2.0,
      (t4 * t50) / 2.0 + (t2 * t53) / 2.0, (t4 * t52) / 2.0 + (t2 * t56) / 2.0,
      (t4 * t55) / 2.0 + (t2 * t58) / 2.0,
      D13 * t50 * mh - (D11 * t53) / 2.0, D13 * t52 * mh - (D11 * t56) / 2.0,
      D13 * t55 * mh - (D11 * t58) / 2.0, D23 * t50 * mh - (D21 * t53) / 2.0,
      D23 * t52 * mh - (D21 * t56) / 2.0, D23 * t55 * mh - (D21 * t58) / 2.0,
      D33 * t50 * mh - (D31 * t53) / 2.0, D33 * t52 * mh - (D31 * t56) / 2.0,
      D33 * t55 * mh - (D31 * t58) / 2.0,
      (t4 * t51) / 2.0 + (t3 * t53) / 2.0, (t4 * t54) / 2.0 + (t3 * t56) / 2.0,
      (t3 * t58) / 2.0 + (t4 * t57) / 2.0,
      D13 * t51 * mh - (D12 * t53) / 2.0, D13 * t54 * mh - (D12 * t56) / 2.0,
      D12 * t58 * mh - (D13 * t57) / 2.0, D23 * t51 * mh - (D22 * t53) / 2.0,
      D23 * t54 * mh - (D22 * t56) / 2.0, D22 * t58 * mh - (D23 * t57) / 2.0,
      D33 * t51 * mh - (D32 * t53) / 2.0, D33 * t54 * mh - (D32 * t56) / 2.0,
      D32 * t58 * mh - (D33 * t57) / 2.0,
      t50 * t50 / 2.0 + t52 * t52 / 2.0 + t55 * t55 / 2.0 - 1.0 / 2.0,
      t51 * t51 / 2.0 + t54 * t54 / 2.0 + t57 * t57 / 2.0 - 1.0 / 2.0,
      t53 * t53 / 2.0 + t56 * t56 / 2.0 + t58 * t58 / 2.0 - 1.0 / 2.0,
      (t50 * t51) / 2.0 + (t52 * t54) / 2.0 + (t55 * t57) / 2.0,
      (t50 * t53) / 2.0 + (t52 * t56) / 2.0 + (t55 * t58) / 2.0,
      (t51 * t53) / 2.0 + (t54 * t56) / 2.0 + (t57 * t58) / 2.0};
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 12; ++c) gradC[r + 6 * c] = out[12 * r + c];
  for (int r = 0; r < 6; ++r) C[r] = out[72 + r];
}

/* matStvkVoigtConstraint3D.m:1-22 with decomposeSolve :31-45.
 * alphaT = alpha/h^2 and betaT = beta*h^2 (6x6 column-major), as
 * xpbdLayer3D.m:318 passes them.  dx (12), dlambda (6), C (6, volume-scaled).
 * Pins: rest state gives C = 0 and dlambda = 0; the 3+3 solve against an
 * independent dense numpy solve of the same block system (tests). */
void ora_stvk_voigt_constraint3d(const double *x, const double *xprev,
                                 const double *DmInv, const double *aT,
                                 const double *lambda, const double *minv,
                                 double volume, const double *bT, double h,
                                 double *dx, double *dl, double *Cout) {
  double g[72], C[6];
  ora_stvk_grad3d(DmInv, x, g, C);
  for (int r = 0; r < 6; ++r) C[r] = C[r] * volume;
  for (int k = 0; k < 72; ++k) g[k] = g[k] * volume;
  double gM[72];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 12; ++c) gM[r + 6 * c] = g[r + 6 * c] * minv[c];
  double gam[36];
  matmul_n(6, aT, bT, gam);
  for (int k = 0; k < 36; ++k) gam[k] = gam[k] / h;
  double gdx[12];
  for (int c = 0; c < 12; ++c) gdx[c] = x[c] - xprev[c];
  double G2[72]; /* gamma * gradC (6x12) */
  for (int i = 0; i < 6; ++i)
    for (int c = 0; c < 12; ++c) {
      double s = gam[i] * g[6 * c];
      for (int k = 1; k < 6; ++k) s = s + gam[i + 6 * k] * g[k + 6 * c];
      G2[i + 6 * c] = s;
    }
  double numGamma[6];
  for (int i = 0; i < 6; ++i) {
    double s = G2[i] * gdx[0];
    for (int c = 1; c < 12; ++c) s = s + G2[i + 6 * c] * gdx[c];
    numGamma[i] = s;
  }
  double DG[36];
  for (int k = 0; k < 36; ++k) DG[k] = gam[k] + ((k % 7) == 0 ? 1.0 : 0.0);
  /* top 3x3 block */
  double topNum[3];
  for (int i = 0; i < 3; ++i) {
    double al = aT[i] * lambda[0];
    for (int k = 1; k < 3; ++k) al = al + aT[i + 6 * k] * lambda[k];
    topNum[i] = -C[i] - al - numGamma[i];
  }
  double A[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = gM[i] * g[j];
      for (int c = 1; c < 12; ++c) s = s + gM[i + 6 * c] * g[j + 6 * c];
      A[i + 3 * j] = s;
    }
  double topDen[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = DG[i] * A[3 * j];
      for (int k = 1; k < 3; ++k) s = s + DG[i + 6 * k] * A[k + 3 * j];
      topDen[i + 3 * j] = s + aT[i + 6 * j];
    }
  solve_n(3, topDen, topNum, dl);
  double a44 = aT[3 + 6 * 3], dg44 = DG[3 + 6 * 3];
  for (int i = 3; i < 6; ++i) {
    double num = -C[i] - a44 * lambda[i] - numGamma[i];
    double bottom = gM[i] * g[i];
    for (int c = 1; c < 12; ++c) bottom = bottom + gM[i + 6 * c] * g[i + 6 * c];
    double den = dg44 * bottom + a44;
    dl[i] = num / den;
  }
  for (int c = 0; c < 12; ++c) {
    double s = gM[6 * c] * dl[0];
    for (int i = 1; i < 6; ++i) s = s + gM[i + 6 * c] * dl[i];
    dx[c] = s;
  }
  if (Cout) for (int r = 0; r < 6; ++r) Cout[r] = C[r];
}

/* matStvkVoigtConstraint.m:53-78 (StVKAnalyticalConstraintGrad).  gradC is
 * 3x6 column-major; C = [E11 E22 E12] with F = [p0-p2, p1-p2]*DmInv.
 * Pins: finite differences; C=0 at rest (tests). */
void ora_stvk_grad2d(const double *D, const double *x, double *gradC, double *Cc) {
  double Ds[4] = {x[0] - x[4], x[1] - x[5], x[2] - x[4], x[3] - x[5]};
  double F[4];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) F[i + 2 * j] = Ds[i] * D[2 * j] + Ds[i + 2] * D[1 + 2 * j];
  double FtF[4];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) FtF[i + 2 * j] = F[2 * i] * F[2 * j] + F[1 + 2 * i] * F[1 + 2 * j];
  double E11 = 0.5 * (FtF[0] - 1.0), E22 = 0.5 * (FtF[3] - 1.0), E12 = 0.5 * (FtF[2] - 0.0);
  Cc[0] = E11; Cc[1] = E22; Cc[2] = E12;
  for (int i = 0; i < 2; ++i) {
    double b0 = D[0 + 2 * i] * F[0 + 2 * i], b1 = D[0 + 2 * i] * F[1 + 2 * i];
    double b2 = D[1 + 2 * i] * F[0 + 2 * i], b3 = D[1 + 2 * i] * F[1 + 2 * i];
    gradC[i + 0] = b0; gradC[i + 3] = b1; gradC[i + 6] = b2; gradC[i + 9] = b3;
    gradC[i + 12] = -(b0 + b2);
    gradC[i + 15] = -(b1 + b3);
  }
  double P11 = D[2], P21 = D[3], P12 = D[0], P22 = D[1];
  double s1 = P11 + P21, s2 = P12 + P22;
  double bp[6][2] = {{P11, P12}, {P11, P12}, {P21, P22}, {P21, P22}, {-s1, -s2}, {-s1, -s2}};
  double bf[6][2] = {{F[0], F[2]}, {F[1], F[3]}, {F[0], F[2]}, {F[1], F[3]}, {F[0], F[2]}, {F[1], F[3]}};
  for (int k = 0; k < 6; ++k) gradC[2 + 3 * k] = 0.5 * (bp[k][0] * bf[k][0] + bp[k][1] * bf[k][1]);
}

/* matStvkVoigtConstraint.m:1-22 with justSolve :24-34 (full 3x3 solve). */
void ora_stvk_voigt_constraint2d(const double *x, const double *xprev,
                                 const double *DmInv, const double *aT,
                                 const double *lambda, const double *minv,
                                 double volume, const double *bT, double h,
                                 double *dx, double *dl, double *Cout) {
  double g[18], C[3];
  ora_stvk_grad2d(DmInv, x, g, C);
  for (int k = 0; k < 18; ++k) g[k] = volume * g[k];
  for (int r = 0; r < 3; ++r) C[r] = volume * C[r];
  double gM[18];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 6; ++c) gM[r + 3 * c] = g[r + 3 * c] * minv[c];
  double gam[9];
  matmul_n(3, aT, bT, gam);
  for (int k = 0; k < 9; ++k) gam[k] = gam[k] / h;
  double gdx[6];
  for (int c = 0; c < 6; ++c) gdx[c] = x[c] - xprev[c];
  double G2[18];
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 6; ++c) {
      double s = gam[i] * g[3 * c];
      for (int k = 1; k < 3; ++k) s = s + gam[i + 3 * k] * g[k + 3 * c];
      G2[i + 3 * c] = s;
    }
  double num[3];
  for (int i = 0; i < 3; ++i) {
    double ng = G2[i] * gdx[0];
    for (int c = 1; c < 6; ++c) ng = ng + G2[i + 3 * c] * gdx[c];
    double al = aT[i] * lambda[0];
    for (int k = 1; k < 3; ++k) al = al + aT[i + 3 * k] * lambda[k];
    num[i] = -C[i] - al - ng;
  }
  double A[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = gM[i] * g[j];
      for (int c = 1; c < 6; ++c) s = s + gM[i + 3 * c] * g[j + 3 * c];
      A[i + 3 * j] = s;
    }
  double DG[9];
  for (int k = 0; k < 9; ++k) DG[k] = gam[k] + ((k % 4) == 0 ? 1.0 : 0.0);
  double den[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = DG[i] * A[3 * j];
      for (int k = 1; k < 3; ++k) s = s + DG[i + 3 * k] * A[k + 3 * j];
      den[i + 3 * j] = s + aT[i + 3 * j];
    }
  solve_n(3, den, num, dl);
  for (int c = 0; c < 6; ++c) {
    double s = gM[3 * c] * dl[0];
    for (int i = 1; i < 3; ++i) s = s + gM[i + 3 * c] * dl[i];
    dx[c] = s;
  }
  if (Cout) for (int r = 0; r < 3; ++r) Cout[r] = C[r];
}

/* expRodrigues.m:1-15 with crossProductMatrix.m:1-7.  R column-major.
 * Pins: scipy Rotation.from_rotvec; expRodrigues(0) = I (tests). */
void ora_exp_rodrigues(const double *omega, double h, double *R) {
  double mag = sqrt(omega[0] * omega[0] + omega[1] * omega[1] + omega[2] * omega[2]);
  double w[3];
  for (int i = 0; i < 3; ++i) { w[i] = omega[i] / mag; if (isnan(w[i])) w[i] = 0; }
  double theta = h * mag;
  double K[9] = {0, w[2], -w[1], -w[2], 0, w[0], w[1], -w[0], 0};
  double K2[9];
  matmul_n(3, K, K, K2);
  double s = sin(theta), c1 = 1 - cos(theta);
  for (int k = 0; k < 9; ++k) R[k] = ((k % 4) == 0 ? 1.0 : 0.0) + s * K[k] + c1 * K2[k];
}

/* ------------------------------------------------------------------ */
/* color sweeps                                                         */
/* ------------------------------------------------------------------ */

/* NeoHookean elements of a 3D mesh (defined with the 2D pair below). */
static void nh_project3d(const ora_mesh *m, int e, double h, double *lambda, double *xi);
static double nh_residual_sq3d(const ora_mesh *m, int e, double h, const double *lambda,
                               const double *xi);

/* xpbdLayer3D.stvkVoigtConstraints (xpbdLayer3D.m:347-379) with the
 * alpha/h^2 and beta*h^2 of the call site :318. */
void ora_stvk_voigt_constraints3d(const ora_mesh *m, const uint8_t *isElastic,
                                  double h, double *lambda, double *xi,
                                  const double *oldp) {
  for (int c = 1; c <= m->numColors; ++c) {
    /* the elements of one color share no vertex: with threads they run in
     * any order and still give the serial result bit for bit */
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
    for (int e = 0; e < m->nT; ++e) {
      if (m->color[e] != c || !isElastic[e]) continue;
      if (m->nhModel && m->nhModel[e]) {
        nh_project3d(m, e, h, lambda, xi);
        continue;
      }
      double aT[36], bT[36];
      double x[12], xo[12], mi[12], dx[12], dl[6];
      for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 3; ++d) {
          int dof = 3 * m->T[4 * e + a] + d;
          x[3 * a + d] = xi[dof];
          xo[3 * a + d] = oldp[dof];
          mi[3 * a + d] = 1.0 / m->mass[dof];
        }
      for (int k = 0; k < 36; ++k) {
        aT[k] = m->alpha[36 * e + k] / (h * h);
        bT[k] = m->beta[36 * e + k] * (h * h);
      }
      ora_stvk_voigt_constraint3d(x, xo, m->DmInv + 9 * e, aT, lambda + 6 * e, mi,
                                  m->vol[e], bT, h, dx, dl, NULL);
      for (int k = 0; k < 6; ++k) lambda[6 * e + k] = lambda[6 * e + k] + dl[k];
      for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 3; ++d) {
          int dof = 3 * m->T[4 * e + a] + d;
          xi[dof] = x[3 * a + d] + dx[3 * a + d];
        }
    }
  }
}

/* xpbdLayer2D.stvkVoigtConstraints (xpbdLayer2D.m:61-89), call site :178. */
void ora_stvk_voigt_constraints2d(const ora_mesh *m, const uint8_t *isElastic,
                                  double h, double *lambda, double *xi,
                                  const double *oldp) {
  double aT[9], bT[9];
  for (int c = 1; c <= m->numColors; ++c)
    for (int e = 0; e < m->nT; ++e) {
      if (m->color[e] != c || !isElastic[e]) continue;
      double x[6], xo[6], mi[6], dx[6], dl[3];
      for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 2; ++d) {
          int dof = 2 * m->T[3 * e + a] + d;
          x[2 * a + d] = xi[dof];
          xo[2 * a + d] = oldp[dof];
          mi[2 * a + d] = 1.0 / m->mass[dof];
        }
      for (int k = 0; k < 9; ++k) {
        aT[k] = m->alpha[9 * e + k] / (h * h);
        bT[k] = m->beta[9 * e + k] * (h * h);
      }
      ora_stvk_voigt_constraint2d(x, xo, m->DmInv + 4 * e, aT, lambda + 3 * e, mi,
                                  m->vol[e], bT, h, dx, dl, NULL);
      for (int k = 0; k < 3; ++k) lambda[3 * e + k] = lambda[3 * e + k] + dl[k];
      for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 2; ++d) {
          int dof = 2 * m->T[3 * e + a] + d;
          xi[dof] = x[2 * a + d] + dx[2 * a + d];
        }
    }
}

/* ------------------------------------------------------------------ */
/* residual                                                             */
/* ------------------------------------------------------------------ */

/* Contact residual of xpbdLayer3D.m:573-583.  contactCompliance arrives as a
 * 1x1xK page array and contactLambdas(:) as Kx1, so Matlab's implicit
 * expansion makes the correction K x 1 x K: term (i,j) uses compliance j and
 * lambda i.  Returns the sum of squares of all K*K terms. */
static double contact_residual_sq3d(const ora_contacts *c, const double *xi,
                                    const double *cl, double r, double h) {
  double acc = 0.0;
  double h2 = h * h;
  for (int i = 0; i < c->K; ++i) {
    int v = c->vertex[i];
    const double *n = c->normal + 3 * i, *pc = c->point + 3 * i;
    double phi = n[0] * (xi[3 * v] - pc[0]) + n[1] * (xi[3 * v + 1] - pc[1]) +
                 n[2] * (xi[3 * v + 2] - pc[2]);
    if (phi > 0) phi = 0;
    double aC = fabs(r * phi);
    for (int j = 0; j < c->K; ++j) {
      double corr = fabs((c->compliance[j] / h2) * cl[i]);
      double t = aC - (aC < corr ? aC : corr);
      acc = acc + t * t;
    }
  }
  return acc;
}

/* simpleEvaluateAllConstraintResiduals (xpbdLayer3D.m:546-586) over every
 * element, elastic or rigid.  The sum of squares is accumulated in element
 * order; the reference's norm() scales to avoid overflow, which changes
 * only the last bits. */
static double elem_residual_sq3d(const ora_mesh *m, int e, double h, const double *lambda,
                                 const double *xi) {
  if (m->nhModel && m->nhModel[e]) return nh_residual_sq3d(m, e, h, lambda, xi);
  double h2 = h * h, acc = 0.0;
  double x[12], g[72], C[6];
  for (int a = 0; a < 4; ++a)
    for (int d = 0; d < 3; ++d) x[3 * a + d] = xi[3 * m->T[4 * e + a] + d];
  ora_stvk_grad3d(m->DmInv + 9 * e, x, g, C);
  for (int i = 0; i < 6; ++i) {
    double Ci = C[i] * m->vol[e];
    double al = (m->alpha[36 * e + i] / h2) * lambda[6 * e];
    for (int k = 1; k < 6; ++k) al = al + (m->alpha[36 * e + i + 6 * k] / h2) * lambda[6 * e + k];
    double aC = fabs(Ci), corr = fabs(al);
    double t = aC - (aC < corr ? aC : corr);
    acc = acc + t * t;
  }
  return acc;
}

double ora_residual3d(const ora_mesh *m, double h, const double *lambda,
                      const double *xi, const double *oldp,
                      const ora_contacts *c, const double *cl,
                      double contactResistance) {
  (void)oldp;
  double acc = 0.0;
  if (g_threads > 1 && m->nT > 0) {
    /* element terms in parallel, summed in element order as below */
    double *terms = (double *)malloc(sizeof(double) * m->nT);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int e = 0; e < m->nT; ++e) terms[e] = elem_residual_sq3d(m, e, h, lambda, xi);
    for (int e = 0; e < m->nT; ++e) acc = acc + terms[e];
    free(terms);
  } else {
    for (int e = 0; e < m->nT; ++e) acc = acc + elem_residual_sq3d(m, e, h, lambda, xi);
  }
  if (c && c->K > 0) acc = acc + contact_residual_sq3d(c, xi, cl, contactResistance, h);
  return sqrt(acc);
}

/* xpbdLayer2D.simpleEvaluateAllConstraintResiduals (xpbdLayer2D.m:477-518).
 * Quirk kept: lambda pages are formed by reshape(lambdai,3,1,[]) of the
 * nT x 3 column-major array (:489), so page e component r is the linear
 * entry 3e+r of lambdai, not lambdai(e, r+1). */
double ora_residual2d(const ora_mesh *m, double h, const double *lambda,
                      const double *xi, const double *oldp,
                      const ora_contacts *c, const double *cl,
                      double contactResistance) {
  (void)oldp;
  double acc = 0.0, h2 = h * h;
  int nT = m->nT;
  for (int e = 0; e < nT; ++e) {
    double x[6], g[18], C[3], lp[3];
    for (int a = 0; a < 3; ++a)
      for (int d = 0; d < 2; ++d) x[2 * a + d] = xi[2 * m->T[3 * e + a] + d];
    ora_stvk_grad2d(m->DmInv + 4 * e, x, g, C);
    for (int r = 0; r < 3; ++r) {
      long L = 3L * e + r;
      lp[r] = lambda[(L % nT) * 3 + (L / nT)];
    }
    for (int i = 0; i < 3; ++i) {
      double Ci = m->vol[e] * C[i];
      double al = (m->alpha[9 * e + i] / h2) * lp[0];
      for (int k = 1; k < 3; ++k) al = al + (m->alpha[9 * e + i + 3 * k] / h2) * lp[k];
      double aC = fabs(Ci), corr = fabs(al);
      double t = aC - (aC < corr ? aC : corr);
      acc = acc + t * t;
    }
  }
  if (c && c->K > 0) {
    for (int i = 0; i < c->K; ++i) {
      int v = c->vertex[i];
      const double *n = c->normal + 2 * i, *pc = c->point + 2 * i;
      double phi = n[0] * (xi[2 * v] - pc[0]) + n[1] * (xi[2 * v + 1] - pc[1]);
      if (phi > 0) phi = 0;
      double aC = fabs(contactResistance * phi);
      double corr = fabs((c->compliance[i] / h2) * cl[i]);
      double t = aC - (aC < corr ? aC : corr);
      acc = acc + t * t;
    }
  }
  return sqrt(acc);
}

/* ------------------------------------------------------------------ */
/* contacts                                                             */
/* ------------------------------------------------------------------ */

/* PlaneContactFinder3D.findContacts (PlaneContactFinder3D.m:34-75) with
 * contactInfo3D (contactInfo3D.m:32-55): active where
 * distance = (x n1 + y n2 + z n3) - D <= 0 (strict < 0 in 2D,
 * PlaneContactFinder.m:27-30); pointCollider = p - n*distance.  normal must
 * already be unit length and D = dot(normal, position).  Returns K. */
int ora_plane_contacts(int dim, int nV, const double *p, const double *normal,
                       const double *position, double compliance, int strict,
                       int32_t *vertex, double *cnormal, double *cpoint,
                       double *ccompliance) {
  double D = 0;
  for (int d = 0; d < dim; ++d) D = D + normal[d] * position[d];
  int K = 0;
  for (int v = 0; v < nV; ++v) {
    double dist = p[dim * v] * normal[0] + p[dim * v + 1] * normal[1];
    if (dim == 3) dist = dist + p[dim * v + 2] * normal[2];
    dist = dist - D;
    int on = strict ? (dist < 0) : (dist <= 0);
    if (!on) continue;
    vertex[K] = v;
    for (int d = 0; d < dim; ++d) {
      cnormal[dim * K + d] = normal[d];
      cpoint[dim * K + d] = p[dim * v + d] - normal[d] * dist;
    }
    ccompliance[K] = compliance;
    ++K;
  }
  return K;
}

/* PlatformContactDetector.findContacts (3d/contact/PlatformContactDetector.m:65-120):
 * vertices inside the box, in ascending order; phi = -min(half extent -
 * |offset . axis|) over the axes (the first on ties, Matlab min), normal the
 * axis of that minimum signed by sign(offset . axis), collider point x -
 * normal phi (contactInfo3D.m:34-38).  axes: 3 unit rows.  Pins: closed
 * forms on axis-aligned and rotated boxes (tests). */
int ora_box_contacts(int nV, const double *p, const double *center, const double *axes,
                     const double *halfExtent, double compliance, int32_t *vertex,
                     double *cnormal, double *cpoint, double *ccompliance) {
  int K = 0;
  for (int v = 0; v < nV; ++v) {
    const double *x = p + 3 * v;
    double d[3] = {x[0] - center[0], x[1] - center[1], x[2] - center[2]};
    double dot[3];
    for (int a = 0; a < 3; ++a)
      dot[a] = d[0] * axes[3 * a] + d[1] * axes[3 * a + 1] + d[2] * axes[3 * a + 2];
    if (!(fabs(dot[0]) <= halfExtent[0] && fabs(dot[1]) <= halfExtent[1] &&
          fabs(dot[2]) <= halfExtent[2]))
      continue;
    double dist[3] = {halfExtent[0] - fabs(dot[0]), halfExtent[1] - fabs(dot[1]),
                      halfExtent[2] - fabs(dot[2])};
    int k = 0;
    if (dist[1] < dist[k]) k = 1;
    if (dist[2] < dist[k]) k = 2;
    double phi = -dist[k];
    double sg = dot[k] > 0 ? 1.0 : (dot[k] < 0 ? -1.0 : 0.0);
    vertex[K] = v;
    for (int a = 0; a < 3; ++a) {
      double n = sg * axes[3 * k + a];
      cnormal[3 * K + a] = n;
      cpoint[3 * K + a] = x[a] - n * phi;
    }
    ccompliance[K] = compliance;
    ++K;
  }
  return K;
}

/* Contacts against a signed distance sampled on a regular grid (values x
 * fastest, sample (i,j,k) at origin + spacing (i,j,k)), in place of
 * gptoolbox's signed_distance query of MeshSCD3D (3d/contact/MeshSCD3D.m:139,
 * contacts where S < 0 with a finite normal): trilinear interpolation in
 * the vertex's cell (none outside the grid), normal the unit gradient of the
 * interpolant.  Pins: a sampled plane (trilinear is exact there) against the
 * plane finder; a sampled sphere against its closed form (tests). */
int ora_sdf_contacts(int nV, const double *p, const double *grid, const int32_t *dims,
                     const double *origin, double spacing, double compliance, int32_t *vertex,
                     double *cnormal, double *cpoint, double *ccompliance) {
  int K = 0;
  const long sx = 1, sy = dims[0], sz = (long)dims[0] * dims[1];
  for (int v = 0; v < nV; ++v) {
    const double *x = p + 3 * v;
    int i0[3], out = 0;
    double t[3];
    for (int a = 0; a < 3; ++a) {
      double g = (x[a] - origin[a]) / spacing;
      if (!(g >= 0.0) || !(g <= (double)(dims[a] - 1))) { out = 1; break; }
      int i = (int)floor(g);
      if (i > dims[a] - 2) i = dims[a] - 2;
      i0[a] = i;
      t[a] = g - (double)i;
    }
    if (out) continue;
    const double *c = grid + i0[0] * sx + i0[1] * sy + i0[2] * sz;
    double c000 = c[0], c100 = c[sx], c010 = c[sy], c110 = c[sx + sy], c001 = c[sz],
           c101 = c[sx + sz], c011 = c[sy + sz], c111 = c[sx + sy + sz];
    double u = t[0], w = t[1], r = t[2];
    double c00 = c000 + u * (c100 - c000), c10 = c010 + u * (c110 - c010);
    double c01 = c001 + u * (c101 - c001), c11 = c011 + u * (c111 - c011);
    double c0 = c00 + w * (c10 - c00), c1 = c01 + w * (c11 - c01);
    double val = c0 + r * (c1 - c0);
    if (!(val < 0.0)) continue;
    double dx00 = c100 - c000, dx10 = c110 - c010, dx01 = c101 - c001, dx11 = c111 - c011;
    double dx0 = dx00 + w * (dx10 - dx00), dx1 = dx01 + w * (dx11 - dx01);
    double gx = (dx0 + r * (dx1 - dx0)) / spacing;
    double dy0 = c10 - c00, dy1 = c11 - c01;
    double gy = (dy0 + r * (dy1 - dy0)) / spacing;
    double gz = (c1 - c0) / spacing;
    double gn = sqrt(gx * gx + gy * gy + gz * gz);
    if (!(gn > 0.0)) continue;
    double n[3] = {gx / gn, gy / gn, gz / gn};
    vertex[K] = v;
    for (int a = 0; a < 3; ++a) {
      cnormal[3 * K + a] = n[a];
      cpoint[3 * K + a] = x[a] - n[a] * val;
    }
    ccompliance[K] = compliance;
    ++K;
  }
  return K;
}

/* xpbdLayer2D.contactVelocityUpdate (xpbdLayer2D.m:577-603), shared by 3D
 * (xpbdLayer3D.m:740).  The e = 0 branch tests ||v|| <= 2*g*h with g < 0 and
 * so never fires for a downward gravity; it is kept. */
void ora_contact_velocity_update(int dim, const ora_contacts *c, double *v,
                                 double h, double gravity, double e) {
  for (int i = 0; i < c->K; ++i) {
    int id = c->vertex[i];
    double vn[3], nrm = 0.0;
    for (int d = 0; d < dim; ++d) { vn[d] = v[dim * id + d]; nrm = nrm + vn[d] * vn[d]; }
    nrm = sqrt(nrm);
    double ei = e;
    if (nrm <= 2 * gravity * h) ei = 0;
    double vt = 0.0;
    for (int d = 0; d < dim; ++d) vt = vt + c->normal[dim * i + d] * vn[d];
    double mn = -ei * vt;
    if (mn > 0) mn = 0;
    for (int d = 0; d < dim; ++d) {
      double dv = c->normal[dim * i + d] * (-vn[d] + mn);
      v[dim * id + d] = vn[d] + dv;
    }
  }
}

/* ------------------------------------------------------------------ */
/* rigid layer (3D)                                                     */
/* ------------------------------------------------------------------ */

typedef struct {
  int nb;
  double *M, *COM, *COMDOT, *R, *I0, *I, *omega; /* per body */
} bodies_t;

static void bodies_alloc(bodies_t *b, int nb) {
  b->nb = nb;
  int n = nb > 0 ? nb : 1;
  b->M = (double *)calloc(n, sizeof(double));
  b->COM = (double *)calloc(3 * n, sizeof(double));
  b->COMDOT = (double *)calloc(3 * n, sizeof(double));
  b->R = (double *)calloc(9 * n, sizeof(double));
  b->I0 = (double *)calloc(9 * n, sizeof(double));
  b->I = (double *)calloc(9 * n, sizeof(double));
  b->omega = (double *)calloc(3 * n, sizeof(double));
}
static void bodies_free(bodies_t *b) {
  free(b->M); free(b->COM); free(b->COMDOT); free(b->R); free(b->I0);
  free(b->I); free(b->omega);
}

/* AdaptiveMesh3D.makeRigidsFromConnectivity (AdaptiveMesh3D.m:32-43) with
 * RigidBody3D.matComputeRigidBodyProperties3D (RigidBody3D.m:198-247) and
 * buildInertiaTensor (:249-265).  Sums run in ascending vertex order, as
 * accumarray does.  ri (3nV) is written for rigid vertices.
 * Pins: COM and inertia against numpy sums per body; L = I*omega (tests). */
static void rigid_properties3d(int nV, const int32_t *vb, const double *p,
                               const double *v, const double *mass,
                               bodies_t *b, double *ri) {
  int nb = b->nb;
  double *mx = (double *)calloc(3 * (nb > 0 ? nb : 1), sizeof(double));
  double *mv = (double *)calloc(3 * (nb > 0 ? nb : 1), sizeof(double));
  double *L = (double *)calloc(3 * (nb > 0 ? nb : 1), sizeof(double));
  for (int k = 0; k < nb; ++k) b->M[k] = 0;
  for (int i = 0; i < nV; ++i) {
    int k = vb[i] - 1;
    if (k < 0) continue;
    double m = mass[3 * i];
    b->M[k] = b->M[k] + m;
    for (int d = 0; d < 3; ++d) {
      mx[3 * k + d] = mx[3 * k + d] + p[3 * i + d] * m;
      mv[3 * k + d] = mv[3 * k + d] + v[3 * i + d] * m;
    }
  }
  for (int k = 0; k < nb; ++k)
    for (int d = 0; d < 3; ++d) {
      b->COM[3 * k + d] = mx[3 * k + d] / b->M[k];
      b->COMDOT[3 * k + d] = mv[3 * k + d] / b->M[k];
    }
  for (int k = 0; k < 9 * nb; ++k) b->I0[k] = 0;
  for (int i = 0; i < nV; ++i) {
    int k = vb[i] - 1;
    if (k < 0) continue;
    double m = mass[3 * i];
    double r[3], vd[3];
    for (int d = 0; d < 3; ++d) {
      r[d] = p[3 * i + d] - b->COM[3 * k + d];
      vd[d] = v[3 * i + d] - b->COMDOT[3 * k + d];
    }
    double rh[9] = {0, r[2], -r[1], -r[2], 0, r[0], r[1], -r[0], 0};
    double rr[9];
    matmul_n(3, rh, rh, rr);
    for (int q = 0; q < 9; ++q) b->I0[9 * k + q] = b->I0[9 * k + q] + (-m) * rr[q];
    double cr[3];
    cross3(r, vd, cr);
    for (int d = 0; d < 3; ++d) L[3 * k + d] = L[3 * k + d] + m * cr[d];
  }
  for (int k = 0; k < nb; ++k) {
    solve_n(3, b->I0 + 9 * k, L + 3 * k, b->omega + 3 * k);
    for (int q = 0; q < 9; ++q) b->R[9 * k + q] = (q % 4) == 0 ? 1.0 : 0.0;
    memcpy(b->I + 9 * k, b->I0 + 9 * k, sizeof(double) * 9);
  }
  for (int i = 0; i < nV; ++i) {
    int k = vb[i] - 1;
    if (k < 0) continue;
    for (int d = 0; d < 3; ++d) ri[3 * i + d] = p[3 * i + d] - b->COM[3 * k + d];
  }
  free(mx); free(mv); free(L);
}

/* RigidBody.pageRelativeVertexPosition (RigidBody3D.m:193-196):
 * pos = (R * ri^T)^T + COM. */
static void rigid_pos(const double *R, const double *ri, const double *COM, double *out) {
  for (int i = 0; i < 3; ++i) {
    double s = R[i] * ri[0];
    for (int k = 1; k < 3; ++k) s = s + R[i + 3 * k] * ri[k];
    out[i] = s + COM[i];
  }
}

/* inertia = R * (inertia0 * R^T) (xpbdLayer3D.m:414,479). */
static void inertia_update(const double *R, const double *I0, double *I) {
  double Rt[9], tmp[9];
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Rt[i + 3 * j] = R[j + 3 * i];
  matmul_n(3, I0, Rt, tmp);
  matmul_n(3, R, tmp, I);
}

/* rigidGeneralizedInverseMass.m:1-13: w = minv + (r x n) . (I \ (r x n)). */
static double rigid_gen_inv_mass(double minv, const double *r, const double *I,
                                 const double *n) {
  double ax[3], ic[3];
  cross3(r, n, ax);
  solve_n(3, I, ax, ic);
  return minv + (ax[0] * ic[0] + ax[1] * ic[1] + ax[2] * ic[2]);
}

/* Apply per-body accumulated dCOM and dOmega (xpbdLayer3D.m:400-414 and
 * :465-479): COM += sum dCOM; w = I \ sum dOmega; R = exp(h w) R;
 * I = R I0 R^T.  Bodies with nothing accumulated get w = 0 and keep R. */
static void bodies_apply(bodies_t *b, const double *dCOM, const double *dOm, double h) {
  for (int k = 0; k < b->nb; ++k) {
    for (int d = 0; d < 3; ++d) b->COM[3 * k + d] = b->COM[3 * k + d] + dCOM[3 * k + d];
    double w[3], dR[9], Rn[9];
    solve_n(3, b->I + 9 * k, dOm + 3 * k, w);
    ora_exp_rodrigues(w, h, dR);
    matmul_n(3, dR, b->R + 9 * k, Rn);
    memcpy(b->R + 9 * k, Rn, sizeof Rn);
    inertia_update(b->R + 9 * k, b->I0 + 9 * k, b->I + 9 * k);
  }
}

/* rigidElasticBoundaryAccumulationConstraints (xpbdLayer3D.m:381-420) with
 * rigidBoundaryConstraint.project (rigidBoundaryConstraint.m:4-35). */
static void boundary3d(bodies_t *b, int nbv, const int32_t *bverts, const int32_t *vb,
                       const double *ri, const double *mass, double compliance,
                       double *lamB, double *xi, double h) {
  if (nbv == 0) return;
  int nb = b->nb;
  double *dCOM = (double *)calloc(3 * (nb > 0 ? nb : 1), sizeof(double));
  double *dOm = (double *)calloc(3 * (nb > 0 ? nb : 1), sizeof(double));
  for (int q = 0; q < nbv; ++q) {
    int v = bverts[q], k = vb[v] - 1;
    double rx[3], r[3], dir[3], n[3];
    rigid_pos(b->R + 9 * k, ri + 3 * v, b->COM + 3 * k, rx);
    for (int d = 0; d < 3; ++d) r[d] = rx[d] - b->COM[3 * k + d];
    for (int d = 0; d < 3; ++d) dir[d] = rx[d] - xi[3 * v + d];
    double C = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    for (int d = 0; d < 3; ++d) { n[d] = dir[d] / C; if (isnan(n[d])) n[d] = 0; }
    double minvE = 1.0 / mass[3 * v + 2];
    double invM = 1.0 / b->M[k];
    double w = rigid_gen_inv_mass(invM, r, b->I + 9 * k, n);
    double num = -C - compliance * lamB[q];
    double den = (minvE + w) + compliance;
    double dl = num / den;
    double pp[3], ph[3], dom[3];
    for (int d = 0; d < 3; ++d) { pp[d] = dl * n[d]; ph[d] = pp[d] / h; }
    cross3(r, ph, dom);
    lamB[q] = lamB[q] + dl;
    for (int d = 0; d < 3; ++d) {
      dCOM[3 * k + d] = dCOM[3 * k + d] + pp[d] * invM;
      dOm[3 * k + d] = dOm[3 * k + d] + dom[d];
    }
  }
  bodies_apply(b, dCOM, dOm, h);
  for (int q = 0; q < nbv; ++q) {
    int v = bverts[q], k = vb[v] - 1;
    rigid_pos(b->R + 9 * k, ri + 3 * v, b->COM + 3 * k, xi + 3 * v);
  }
  free(dCOM); free(dOm);
}

/* contactConstraints (xpbdLayer3D.m:422-485) with penaltyConstraint
 * projectElastic (penaltyConstraint.m:4-28) and projectRigid (:30-63).
 * Elastic contacts are all computed from the incoming xi and written in
 * contact order (a repeated vertex keeps the last write, like the indexed
 * assignment :437). */
static void contacts3d(bodies_t *b, const ora_contacts *c, const int32_t *vb,
                       const double *ri, const double *mass, double *cl,
                       double *xi, double h, double resistance, int useRigid) {
  if (!c || c->K == 0) return;
  double h2 = h * h;
  double *newx = (double *)malloc(sizeof(double) * 3 * c->K);
  for (int i = 0; i < c->K; ++i) {
    int v = c->vertex[i];
    if (vb[v] != 0) continue;
    const double *n = c->normal + 3 * i, *pc = c->point + 3 * i;
    double x[3] = {xi[3 * v], xi[3 * v + 1], xi[3 * v + 2]};
    double phi = n[0] * (x[0] - pc[0]) + n[1] * (x[1] - pc[1]) + n[2] * (x[2] - pc[2]);
    if (phi > 0) phi = 0;
    double C = resistance * phi;
    double aT = c->compliance[i] / h2;
    double num = -C - aT * cl[i];
    double den = 2 * (1.0 / mass[3 * v]) + aT;
    double dl = num / den;
    for (int d = 0; d < 3; ++d) newx[3 * i + d] = x[d] + (1.0 / mass[3 * v + d]) * (dl * n[d]);
    cl[i] = cl[i] + dl;
  }
  for (int i = 0; i < c->K; ++i) {
    int v = c->vertex[i];
    if (vb[v] != 0) continue;
    for (int d = 0; d < 3; ++d) xi[3 * v + d] = newx[3 * i + d];
  }
  free(newx);
  int nr = 0;
  for (int i = 0; i < c->K; ++i) nr += vb[c->vertex[i]] != 0;
  if (nr < 1 || !useRigid) return;
  int nb = b->nb;
  double *dCOM = (double *)calloc(3 * (nb > 0 ? nb : 1), sizeof(double));
  double *dOm = (double *)calloc(3 * (nb > 0 ? nb : 1), sizeof(double));
  for (int i = 0; i < c->K; ++i) {
    int v = c->vertex[i];
    int k = vb[v] - 1;
    if (k < 0) continue;
    const double *n = c->normal + 3 * i, *pc = c->point + 3 * i;
    double rx[3], r[3];
    rigid_pos(b->R + 9 * k, ri + 3 * v, b->COM + 3 * k, rx);
    for (int d = 0; d < 3; ++d) r[d] = rx[d] - b->COM[3 * k + d];
    double phi = n[0] * (xi[3 * v] - pc[0]) + n[1] * (xi[3 * v + 1] - pc[1]) +
                 n[2] * (xi[3 * v + 2] - pc[2]);
    if (phi > 0) phi = 0;
    double C = resistance * phi;
    double invM = 1.0 / b->M[k];
    double w = rigid_gen_inv_mass(invM, r, b->I + 9 * k, n);
    double aT = (2 * c->compliance[i]) / h2;
    double num = -C - aT * cl[i];
    double den = 2 * w + aT;
    double dl = num / den;
    double pp[3], cr[3];
    for (int d = 0; d < 3; ++d) pp[d] = dl * n[d];
    cross3(r, pp, cr);
    cl[i] = cl[i] + dl;
    for (int d = 0; d < 3; ++d) {
      dCOM[3 * k + d] = dCOM[3 * k + d] + pp[d] * invM;
      dOm[3 * k + d] = dOm[3 * k + d] + (1 / h) * cr[d];
    }
  }
  bodies_apply(b, dCOM, dOm, h);
  for (int i = 0; i < c->K; ++i) {
    int v = c->vertex[i], k = vb[v] - 1;
    if (k < 0) continue;
    rigid_pos(b->R + 9 * k, ri + 3 * v, b->COM + 3 * k, xi + 3 * v);
  }
  free(dCOM); free(dOm);
}

/* rotateAllVerts (xpbdLayer3D.m:508-513): rigid vertices to R ri + COM. */
static void rotate_all_verts(int nV, const int32_t *vb, const double *ri,
                             const bodies_t *b, double *xi) {
  for (int i = 0; i < nV; ++i) {
    int k = vb[i] - 1;
    if (k < 0) continue;
    rigid_pos(b->R + 9 * k, ri + 3 * i, b->COM + 3 * k, xi + 3 * i);
  }
}

/* ------------------------------------------------------------------ */
/* monitor                                                              */
/* ------------------------------------------------------------------ */

typedef struct {
  double *arr; int cap, count;
  int errorChangeDecreased;
} monitor_t;

static void monitor_push(monitor_t *mo, double r) {
  if (mo->count < mo->cap) mo->arr[mo->count] = r;
  mo->count++;
}
static double monitor_at(const monitor_t *mo, int i) {
  return (i < mo->cap) ? mo->arr[i] : 0.0;
}

/* simpleMonitorError (xpbdLayer3D.m:515-544; 2D xpbdLayer2D.m:520-549).
 * is3d selects the 3D stop rule (1e-8 floor, normresv test) or the 2D one
 * (no floor, resvnorm < h^2 test). */
static void monitor_step(monitor_t *mo, int is3d, double newRes, ora_settings *s,
                         double *setPercentStop, int isLastLayer,
                         double normresv, double h, int *endLayer, int *stopSolve) {
  *endLayer = 0; *stopSolve = 0;
  if (!s->computeResiduals) return;
  monitor_push(mo, newRes);
  double last = fabs(monitor_at(mo, mo->count - 2));
  double nr = fabs(newRes);
  double change = nr - last;
  if (*setPercentStop > 0) {
    double ru = (1 - *setPercentStop / 100.0) * nr;
    if (is3d) ru = ru > 1e-8 ? ru : 1e-8;
    s->runUntil = ru;
    *setPercentStop = 0;
  }
  if (change < 0) mo->errorChangeDecreased = 1;
  if (s->giveUpEnabled && change > s->giveUpThreshold && mo->errorChangeDecreased &&
      (!isLastLayer || s->lastLayerGiveUpEnabled))
    *endLayer = 1;
  int stopHanging = mo->count > s->hangingStop && s->hangingStop >= 0;
  if (is3d) {
    if (stopHanging || (s->runUntil >= 0 && (s->runUntil >= nr && (isLastLayer || normresv < s->runUntil))))
      *stopSolve = 1;
  } else {
    if ((s->runUntil >= 0 && (s->runUntil >= nr) && (isLastLayer || normresv < h * h)) || stopHanging)
      *stopSolve = 1;
  }
}

/* ------------------------------------------------------------------ */
/* layer building (shared by the integrators)                           */
/* ------------------------------------------------------------------ */

typedef struct {
  int nSorted;
  double sorted[ORA_MAX_LAYERS];
  int32_t map[ORA_MAX_LAYERS];
  int32_t *sets, *numRigids, *vertexRigidBody;
  uint8_t *isBoundary, *isRigidVert;
} layers_t;

static void layers_free(layers_t *L) {
  free(L->sets); free(L->numRigids); free(L->vertexRigidBody);
  free(L->isBoundary); free(L->isRigidVert);
}

/* buildLayers (xpbdLayer3D.m:601-619; 2D xpbdLayer2D.m:740-780 ordering type
 * 0) followed by findBoundaryVertices. */
static void build_layers(const ora_mesh *m, const ora_settings *s, const double *pa,
                         const double *pb, double h, const double *keyScale,
                         const int32_t *givenSets, layers_t *L) {
  int nT = m->nT, nV = m->nV;
  L->nSorted = ora_sorted_layers(s->layers, s->nLayers, L->sorted, L->map);
  int ns = L->nSorted;
  if (givenSets) {
    L->sets = (int32_t *)malloc(sizeof(int32_t) * (size_t)nT * ns + 4);
    L->numRigids = (int32_t *)malloc(sizeof(int32_t) * (ns + 1));
    memcpy(L->sets, givenSets, sizeof(int32_t) * (size_t)nT * ns);
    for (int l = 0; l < ns; ++l) {
      int mx = 0;
      for (int e = 0; e < nT; ++e) mx = L->sets[(size_t)nT * l + e] > mx ? L->sets[(size_t)nT * l + e] : mx;
      L->numRigids[l] = mx;
    }
    L->vertexRigidBody = (int32_t *)malloc(sizeof(int32_t) * (size_t)nV * ns + 4);
    L->isBoundary = (uint8_t *)malloc((size_t)nV * ns + 4);
    L->isRigidVert = (uint8_t *)malloc((size_t)nV * ns + 4);
    ora_find_boundary_vertices(nV, nT, m->vpe, m->T, ns, L->sets, L->isBoundary,
                               L->isRigidVert, L->vertexRigidBody);
    return;
  }
  double *Fa = (double *)malloc(sizeof(double) * 9 * (nT > 0 ? nT : 1));
  double *Fb = (double *)malloc(sizeof(double) * 9 * (nT > 0 ? nT : 1));
  double *ed = (double *)malloc(sizeof(double) * (nT > 0 ? nT : 1));
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (nT > 0 ? nT : 1));
  if (m->dim == 3) {
    ora_b_times_p3d(m->T, nT, m->DmInv, pa, Fa);
    ora_b_times_p3d(m->T, nT, m->DmInv, pb, Fb);
    ora_ediff_norm3d(Fa, Fb, nT, h, ed);
    if (keyScale)
      for (int e = 0; e < nT; ++e) ed[e] = ed[e] * keyScale[e];
  } else {
    ora_b_times_p2d(m->T, nT, m->DmInv, pa, Fa);
    ora_b_times_p2d(m->T, nT, m->DmInv, pb, Fb);
    ora_ediff_norm2d(Fa, Fb, nT, h, ed);
  }
  ora_stable_sort_ascend(ed, nT, order);
  L->sets = (int32_t *)malloc(sizeof(int32_t) * (size_t)nT * ns + 4);
  L->numRigids = (int32_t *)malloc(sizeof(int32_t) * (ns + 1));
  ora_union_find_by_vertice(order, nT, L->sorted, ns, m->forcedElastic, m->T, m->vpe,
                            nV, L->sets, L->numRigids);
  L->vertexRigidBody = (int32_t *)malloc(sizeof(int32_t) * (size_t)nV * ns + 4);
  L->isBoundary = (uint8_t *)malloc((size_t)nV * ns + 4);
  L->isRigidVert = (uint8_t *)malloc((size_t)nV * ns + 4);
  ora_find_boundary_vertices(nV, nT, m->vpe, m->T, ns, L->sets, L->isBoundary,
                             L->isRigidVert, L->vertexRigidBody);
  free(Fa); free(Fb); free(ed); free(order);
}

/* ------------------------------------------------------------------ */
/* 3D substep                                                           */
/* ------------------------------------------------------------------ */

/* ------------------------------------------------------------------ */
/* learned warm start (no reference counterpart)                        */
/* ------------------------------------------------------------------ */

/* Residual of the equations the projection solves, over the elastic
 * elements (per constraint the numerator of its update: C + alpha~ lambda +
 * gamma gradC (x - x_old) for StVK-Voigt as matStvkVoigtConstraint3D.m:4-21
 * forms it, the normal rows with the 3x3 block of alpha~, the shear rows
 * with alpha~(4,4); C + alpha~ lambda for the NeoHookean pair) and the
 * contacts (r phi, their multipliers being zero when a layer starts).
 * Pins: zero at a projected fixed point and the closed forms of the tests. */
double ora_learn_eq_residual3d(const ora_mesh *m, double h, const uint8_t *isEl,
                               const double *lambda, const double *xi, const double *oldp,
                               const ora_contacts *c, double resistance) {
  double acc = 0;
  for (int e = 0; e < m->nT; ++e) {
    if (!isEl[e]) continue;
    double x[12], xo[12];
    for (int a = 0; a < 4; ++a)
      for (int d = 0; d < 3; ++d) {
        int dof = 3 * m->T[4 * e + a] + d;
        x[3 * a + d] = xi[dof];
        xo[3 * a + d] = oldp[dof];
      }
    const double *lam = lambda + 6 * e;
    if (m->nhModel && m->nhModel[e]) {
      double n, J;
      ora_neohookean_grad3d(m->DmInv + 9 * e, x, NULL, NULL, &n, &J);
      double iv = (1.0 / m->vol[e]) * (1.0 / (h * h));
      double rD = n + iv * (1.0 / m->nhMu[e]) * lam[0];
      double rH = (J - (1 + m->nhMu[e] / m->nhLam[e])) + iv * (1.0 / m->nhLam[e]) * lam[1];
      acc = acc + rD * rD + rH * rH;
      continue;
    }
    double g[72], C[6], aT[36], bT[36], gam[36];
    ora_stvk_grad3d(m->DmInv + 9 * e, x, g, C);
    for (int r = 0; r < 6; ++r) C[r] = C[r] * m->vol[e];
    for (int k = 0; k < 72; ++k) g[k] = g[k] * m->vol[e];
    for (int k = 0; k < 36; ++k) {
      aT[k] = m->alpha[36 * e + k] / (h * h);
      bT[k] = m->beta[36 * e + k] * (h * h);
    }
    matmul_n(6, aT, bT, gam);
    for (int k = 0; k < 36; ++k) gam[k] = gam[k] / h;
    double gdx[6];
    for (int i = 0; i < 6; ++i) {
      double sgd = 0;
      for (int q = 0; q < 12; ++q) sgd = sgd + g[i + 6 * q] * (x[q] - xo[q]);
      gdx[i] = sgd;
    }
    for (int i = 0; i < 6; ++i) {
      double ng = 0;
      for (int k = 0; k < 6; ++k) ng = ng + gam[i + 6 * k] * gdx[k];
      double al;
      if (i < 3) {
        al = aT[i] * lam[0];
        for (int k = 1; k < 3; ++k) al = al + aT[i + 6 * k] * lam[k];
      } else {
        al = aT[3 + 6 * 3] * lam[i];
      }
      double r = C[i] + al + ng;
      acc = acc + r * r;
    }
  }
  for (int i = 0; c && i < c->K; ++i) {
    const double *n = c->normal + 3 * i, *pc = c->point + 3 * i;
    const double *xv = xi + 3 * c->vertex[i];
    double phi = n[0] * (xv[0] - pc[0]) + n[1] * (xv[1] - pc[1]) + n[2] * (xv[2] - pc[2]);
    if (phi > 0) phi = 0;
    acc = acc + (resistance * phi) * (resistance * phi);
  }
  return sqrt(acc);
}

/* Warm start of the elastic elements whose multipliers are still zero:
 * lambda = lambda0 and x += M^-1 gradC(x)^T lambda0, gradients at the
 * positions before the move, pinned DOFs left in place.  Pins: the
 * displacement against central differences of lambda0 . C (tests). */
void ora_learn_displace3d(const ora_mesh *m, const uint8_t *isEl, const double *lam0,
                          double *lambda, double *xi, const uint8_t *pinnedDof) {
  int nV = m->nV, nT = m->nT;
  double *dx = (double *)calloc(3 * (nV > 0 ? nV : 1), sizeof(double));
  for (int e = 0; e < nT; ++e) {
    if (!isEl[e]) continue;
    int fresh = 1;
    for (int k = 0; k < 6; ++k) fresh = fresh && lambda[6 * e + k] == 0;
    if (!fresh) continue;
    double x[12];
    for (int a = 0; a < 4; ++a)
      for (int d = 0; d < 3; ++d) x[3 * a + d] = xi[3 * m->T[4 * e + a] + d];
    const double *l0 = lam0 + 6 * e;
    double step[12];
    if (m->nhModel && m->nhModel[e]) {
      double gD[12], gH[12];
      ora_neohookean_grad3d(m->DmInv + 9 * e, x, gD, gH, NULL, NULL);
      for (int q = 0; q < 12; ++q) step[q] = gD[q] * l0[0] + gH[q] * l0[1];
    } else {
      double g[72], C[6];
      ora_stvk_grad3d(m->DmInv + 9 * e, x, g, C);
      for (int q = 0; q < 12; ++q) {
        double sq_ = 0;
        for (int i = 0; i < 6; ++i) sq_ = sq_ + g[i + 6 * q] * m->vol[e] * l0[i];
        step[q] = sq_;
      }
    }
    for (int a = 0; a < 4; ++a)
      for (int d = 0; d < 3; ++d) {
        int dof = 3 * m->T[4 * e + a] + d;
        dx[dof] = dx[dof] + step[3 * a + d] / m->mass[dof];
      }
    for (int k = 0; k < 6; ++k) lambda[6 * e + k] = l0[k];
  }
  for (int q = 0; q < 3 * nV; ++q)
    if (!pinnedDof[q]) xi[q] = xi[q] + dx[q];
  free(dx);
}

/* The warm start with its judge: kept only if the equation residual does
 * not rise, else x and lambda are restored.  Returns 1 kept, 0 rejected;
 * res gets the residuals before and after. */
int ora_learn_warm_start3d(const ora_mesh *m, double h, const uint8_t *isEl,
                           const double *lam0, double *lambda, double *xi, const double *oldp,
                           const uint8_t *pinnedDof, const ora_contacts *c, double resistance,
                           double *res) {
  int nV = m->nV, nT = m->nT;
  double cold = ora_learn_eq_residual3d(m, h, isEl, lambda, xi, oldp, c, resistance);
  double *xs = (double *)malloc(sizeof(double) * 3 * (nV > 0 ? nV : 1));
  double *ls = (double *)malloc(sizeof(double) * 6 * (nT > 0 ? nT : 1));
  memcpy(xs, xi, sizeof(double) * 3 * nV);
  memcpy(ls, lambda, sizeof(double) * 6 * nT);
  ora_learn_displace3d(m, isEl, lam0, lambda, xi, pinnedDof);
  double warm = ora_learn_eq_residual3d(m, h, isEl, lambda, xi, oldp, c, resistance);
  res[0] = cold;
  res[1] = warm;
  int keep = warm <= cold;
  if (!keep) {
    memcpy(xi, xs, sizeof(double) * 3 * nV);
    memcpy(lambda, ls, sizeof(double) * 6 * nT);
  }
  free(xs); free(ls);
  return keep;
}

/* xpbdLayer3D.integrate (xpbdLayer3D.m:621-745) with layerSolve (:44-232)
 * and projectGS (:243-345).  Pins: the pieces above, plus an independent
 * numpy restatement on small scenes (tests). */
int ora_integrate3d(const ora_mesh *m, ora_settings *s, double h, int frame,
                    double *p, double *v, double *oldp, double *oldOldp,
                    double *pinnedx, const ora_contacts *c, ora_step_out *out) {
  int nV = m->nV, nT = m->nT, N3 = 3 * nV;
  /* the reference loops forever when the last layer runs until a residual
   * that is never computed (xpbdLayer3D.m:315 with computeResiduals false) */
  if (s->runUntil >= 0 && !s->computeResiduals) return -3;
  int anynan = 0;
  for (int i = 0; i < N3; ++i) anynan |= isnan(oldp[i]);
  if (anynan || frame == 1) {
    memcpy(oldp, p, sizeof(double) * N3);
    for (int q = 0; q < m->nPinnedDOFs; ++q) pinnedx[q] = p[m->pinnedDOFs[q]];
  }
  if (out) out->runUntilAtEntry = s->runUntil;
  /* forces: gravity on z (:638-643) */
  double *f = (double *)calloc(N3, sizeof(double));
  for (int i = 0; i < nV; ++i) f[3 * i + 2] = f[3 * i + 2] + m->mass[3 * i + 2] * s->gravity;
  layers_t L;
  build_layers(m, s, p, oldp, h, out ? out->learnKeyScale : NULL, out ? out->givenSets : NULL,
               &L);
  if (out) out->learnKept = -1;
  if (out && out->setsPerLayer)
    memcpy(out->setsPerLayer, L.sets, sizeof(int32_t) * (size_t)nT * L.nSorted);
  if (out && out->numRigids) memcpy(out->numRigids, L.numRigids, sizeof(int32_t) * L.nSorted);
  memcpy(oldOldp, oldp, sizeof(double) * N3);
  memcpy(oldp, p, sizeof(double) * N3);
  double *xi = (double *)malloc(sizeof(double) * N3);
  memcpy(xi, p, sizeof(double) * N3);
  uint8_t *pinnedDof = (uint8_t *)calloc(N3, 1);
  for (int q = 0; q < m->nPinnedDOFs; ++q) pinnedDof[m->pinnedDOFs[q]] = 1;
  double *rv = (double *)malloc(sizeof(double) * N3);
  for (int i = 0; i < N3; ++i) {
    rv[i] = v[i];
    if (!pinnedDof[i]) rv[i] = rv[i] + h * f[i] / m->mass[i];
    rv[i] = (1 - s->airDrag) * rv[i];
  }
  double *lambda = (double *)calloc(6 * (size_t)(nT > 0 ? nT : 1), sizeof(double));
  int K = c ? c->K : 0;
  double *cl = (double *)calloc(K > 0 ? K : 1, sizeof(double));
  double *clLayer = (double *)calloc(K > 0 ? K : 1, sizeof(double));
  monitor_t mo;
  double dummy[1];
  mo.arr = out && out->residualArray ? out->residualArray : dummy;
  mo.cap = out && out->residualArray ? out->residualCap : 0;
  mo.count = 0;
  mo.errorChangeDecreased = 0;
  monitor_push(&mo, 0.0);
  double *ri = (double *)calloc(N3, sizeof(double));
  double *xprev = (double *)malloc(sizeof(double) * N3);
  uint8_t *isEl = (uint8_t *)malloc(nT > 0 ? nT : 1);
  int32_t *bverts = (int32_t *)malloc(sizeof(int32_t) * (nV > 0 ? nV : 1));
  int64_t projections = 0;
  int layersRun = 0;
  for (int layer = 0; layer < s->nLayers; ++layer) {
    double setPercentStop = (layer == 0) ? s->stopAtImprovementPercent : 0;
    memcpy(xprev, xi, sizeof(double) * N3);
    int isLast = layer == s->nLayers - 1;
    int col = L.map[layer];
    const int32_t *sets = L.sets + (size_t)nT * col;
    const int32_t *vb = L.vertexRigidBody + (size_t)nV * col;
    const uint8_t *isB = L.isBoundary + (size_t)nV * col;
    const uint8_t *isRV = L.isRigidVert + (size_t)nV * col;
    int nRigidEl = 0, nElastic = 0;
    for (int e = 0; e < nT; ++e) { isEl[e] = sets[e] == 0; nRigidEl += !isEl[e]; nElastic += isEl[e]; }
    bodies_t b;
    bodies_alloc(&b, L.numRigids[col]);
    rigid_properties3d(nV, vb, xi, rv, m->mass, &b, ri);
    /* release elastic residual (:138-139) */
    for (int i = 0; i < nV; ++i) {
      if (isRV[i]) continue;
      for (int d = 0; d < 3; ++d) {
        xi[3 * i + d] = xi[3 * i + d] + h * rv[3 * i + d];
        rv[3 * i + d] = 0;
      }
    }
    double *Rstep = (double *)malloc(sizeof(double) * 9 * (b.nb > 0 ? b.nb : 1));
    for (int k = 0; k < 9 * b.nb; ++k) Rstep[k] = b.R[k];
    if (nRigidEl > 0) {
      /* rigidSymplecticImperativeStep (RigidBody3D.m:174-179) */
      for (int k = 0; k < b.nb; ++k) {
        for (int d = 0; d < 3; ++d) b.COM[3 * k + d] = b.COM[3 * k + d] + h * b.COMDOT[3 * k + d];
        double dR[9], Rn[9], tmp[9], Rt[9];
        ora_exp_rodrigues(b.omega + 3 * k, h, dR);
        matmul_n(3, dR, b.R + 9 * k, Rn);
        memcpy(b.R + 9 * k, Rn, sizeof Rn);
        matmul_n(3, b.R + 9 * k, b.I0 + 9 * k, tmp);
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Rt[i + 3 * j] = b.R[9 * k + j + 3 * i];
        matmul_n(3, tmp, Rt, b.I + 9 * k);
      }
      memcpy(Rstep, b.R, sizeof(double) * 9 * b.nb);
      rotate_all_verts(nV, vb, ri, &b, xi);
      for (int i = 0; i < nV; ++i) {
        int k = vb[i] - 1;
        if (k < 0) continue;
        double rr[3], z[3] = {0, 0, 0};
        rigid_pos(b.R + 9 * k, rv + 3 * i, z, rr);
        for (int d = 0; d < 3; ++d) {
          double rres = (xi[3 * i + d] - xprev[3 * i + d]) / h;
          rv[3 * i + d] = rr[d] - rres;
        }
      }
    }
    double normresv = 0;
    for (int i = 0; i < N3; ++i) normresv = normresv + rv[i] * rv[i];
    normresv = sqrt(normresv);
    if (out) out->normresv[layer] = normresv;
    if (isLast && out && out->learnLam0 && nRigidEl == 0 && b.nb == 0)
      out->learnKept = ora_learn_warm_start3d(m, h, isEl, out->learnLam0, lambda, xi, oldp,
                                              pinnedDof, c, s->contactResistance, out->learnRes);
    /* projectGS (:243-345) */
    int nbv = 0;
    for (int i = 0; i < nV; ++i) if (isB[i]) bverts[nbv++] = i;
    double *lamB = (double *)calloc(nbv > 0 ? nbv : 1, sizeof(double));
    memcpy(clLayer, cl, sizeof(double) * (K > 0 ? K : 1));
    int it = 1, stopSolve = 0, iters = 0;
    while (it <= s->iterations[layer] || (!stopSolve && s->runUntil >= 0 && isLast)) {
      it = it + 1;
      ora_stvk_voigt_constraints3d(m, isEl, h, lambda, xi, oldp);
      projections += nElastic;
      boundary3d(&b, nbv, bverts, vb, ri, m->mass, s->boundaryCompliance, lamB, xi, h);
      contacts3d(&b, c, vb, ri, m->mass, clLayer, xi, h, s->contactResistance, s->useRigidContacts);
      for (int q = 0; q < m->nPinnedDOFs; ++q) xi[m->pinnedDOFs[q]] = pinnedx[q];
      ++iters;
      int endLayer = 0;
      if (s->computeResiduals) {
        double r = ora_residual3d(m, h, lambda, xi, oldp, c, clLayer, s->contactResistance);
        monitor_step(&mo, 1, r, s, &setPercentStop, isLast, normresv, h, &endLayer, &stopSolve);
      }
      if (endLayer || stopSolve) break;
    }
    rotate_all_verts(nV, vb, ri, &b, xi);
    if (nRigidEl > 0) {
      /* deltaR = R_proj * R_step^T; rotateResidual (:217-221, :234-241) */
      for (int i = 0; i < nV; ++i) {
        int k = vb[i] - 1;
        if (k < 0) continue;
        double Rt[9], dR[9], z[3] = {0, 0, 0}, out3[3];
        for (int a = 0; a < 3; ++a) for (int bb = 0; bb < 3; ++bb) Rt[a + 3 * bb] = Rstep[9 * k + bb + 3 * a];
        matmul_n(3, b.R + 9 * k, Rt, dR);
        rigid_pos(dR, rv + 3 * i, z, out3);
        for (int d = 0; d < 3; ++d) rv[3 * i + d] = out3[d];
      }
    }
    if (out) out->itersPerLayer[layer] = iters;
    layersRun = layer + 1;
    free(lamB); free(Rstep);
    bodies_free(&b);
    if (stopSolve) break;
  }
  /* velocity update (:737-740) */
  for (int i = 0; i < N3; ++i) {
    v[i] = (xi[i] - oldp[i]) / h;
    p[i] = xi[i];
  }
  if (c && c->K > 0) ora_contact_velocity_update(3, c, v, h, s->gravity, s->restitution);
  if (out) {
    if (out->lambda) memcpy(out->lambda, lambda, sizeof(double) * 6 * nT);
    out->residualCount = mo.count;
    out->layersRun = layersRun;
    out->elasticProjections = projections;
  }
  free(f); free(xi); free(pinnedDof); free(rv); free(lambda); free(cl); free(clLayer);
  free(ri); free(xprev); free(isEl); free(bverts);
  layers_free(&L);
  return 0;
}

/* ------------------------------------------------------------------ */
/* 2D substep                                                           */
/* ------------------------------------------------------------------ */

/* ------------------------------------------------------------------ */
/* rigid layer (2D)                                                     */
/* ------------------------------------------------------------------ */

typedef struct {
  int nb;
  double *M, *COM, *angle, *I; /* per body: mass, COM (2), angle, scalar inertia */
} bodies2d_t;

static void bodies2d_alloc(bodies2d_t *b, int nb) {
  int n = nb > 0 ? nb : 1;
  b->nb = nb;
  b->M = (double *)calloc(n, sizeof(double));
  b->COM = (double *)calloc(2 * n, sizeof(double));
  b->angle = (double *)calloc(n, sizeof(double));
  b->I = (double *)calloc(n, sizeof(double));
}
static void bodies2d_free(bodies2d_t *b) { free(b->M); free(b->COM); free(b->angle); free(b->I); }

/* AdaptiveMesh.makeRigidsFromConnectivity (AdaptiveMesh.m:27-37) with
 * RigidBody.matComputeRigidBodyProperties2D (RigidBody.m:171-215): per body
 * mass (x-DOF masses), COM = sum(m p)/M, inertia sum(m |p - COM|^2),
 * angle = 0, R = I; ri = p - COM for rigid vertices.  The COM velocity and
 * angular velocity the routine also returns are unused by the 2D layerSolve
 * (xpbdLayer2D.m:681) and are not formed.  Sums in ascending vertex order. */
static void rigid_properties2d(int nV, const int32_t *vb, const double *p,
                               const double *mass, bodies2d_t *b, double *ri) {
  int nb = b->nb;
  double *mx = (double *)calloc(2 * (nb > 0 ? nb : 1), sizeof(double));
  for (int i = 0; i < nV; ++i) {
    int k = vb[i] - 1;
    if (k < 0) continue;
    double m = mass[2 * i];
    b->M[k] = b->M[k] + m;
    mx[2 * k] = mx[2 * k] + p[2 * i] * m;
    mx[2 * k + 1] = mx[2 * k + 1] + p[2 * i + 1] * m;
  }
  for (int k = 0; k < nb; ++k) {
    b->COM[2 * k] = mx[2 * k] / b->M[k];
    b->COM[2 * k + 1] = mx[2 * k + 1] / b->M[k];
    b->angle[k] = 0;
    b->I[k] = 0;
  }
  for (int i = 0; i < nV; ++i) {
    int k = vb[i] - 1;
    if (k < 0) continue;
    double rx = p[2 * i] - b->COM[2 * k], ry = p[2 * i + 1] - b->COM[2 * k + 1];
    ri[2 * i] = rx;
    ri[2 * i + 1] = ry;
    b->I[k] = b->I[k] + mass[2 * i] * (rx * rx + ry * ry);
  }
  free(mx);
}

/* RigidBody.pageRelativeVertexPosition with angle2rotm2D (RigidBody.m:166-169,
 * 2d/tools/math/angle2rotm2D.m:1-5): pos = [c -s; s c] ri + COM. */
static void rigid_pos2d(double angle, const double *ri, const double *COM, double *out) {
  double c = cos(angle), s = sin(angle);
  out[0] = (c * ri[0] + (-s) * ri[1]) + COM[0];
  out[1] = (s * ri[0] + c * ri[1]) + COM[1];
}

/* cross2D.m:1-7: v1x v2y - v1y v2x. */
static double cross2d(const double *a, const double *b) { return a[0] * b[1] - a[1] * b[0]; }

/* rigidElasticBoundaryAccumulationConstraints (xpbdLayer2D.m:296-330) with
 * rigidBoundaryConstraint2D.projectVectorized (rigidBoundaryConstraint2D.m:4-34):
 * every boundary vertex is projected against its body's pose at the start of
 * the call; the per-body sums of dCOM = p/M and dAngle = (Iinv/2) r x p are
 * applied; boundary vertices are then set to the rigid positions (the elastic
 * correction is discarded).  w splits the body mass over the body's
 * boundary-vertex count nb: w = (1/M)/nb + Iinv (r x n)^2. */
static void boundary2d(bodies2d_t *b, int nbv, const int32_t *bverts, const int32_t *vb,
                       const double *ri, const double *mass, double alpha,
                       double *lamB, double *xi, double h) {
  if (nbv == 0) return;
  int nb = b->nb;
  int n1 = nb > 0 ? nb : 1;
  int *cnt = (int *)calloc(n1, sizeof(int));
  double *dC = (double *)calloc(2 * n1, sizeof(double));
  double *dA = (double *)calloc(n1, sizeof(double));
  for (int q = 0; q < nbv; ++q) cnt[vb[bverts[q]] - 1]++;
  double h2 = h * h;
  double aT = alpha / h2;
  for (int q = 0; q < nbv; ++q) {
    int v = bverts[q], k = vb[v] - 1;
    double rx[2], r[2], dir[2], n[2];
    rigid_pos2d(b->angle[k], ri + 2 * v, b->COM + 2 * k, rx);
    r[0] = rx[0] - b->COM[2 * k];
    r[1] = rx[1] - b->COM[2 * k + 1];
    dir[0] = rx[0] - xi[2 * v];
    dir[1] = rx[1] - xi[2 * v + 1];
    double C = sqrt(dir[0] * dir[0] + dir[1] * dir[1]);
    n[0] = dir[0] / C;
    n[1] = dir[1] / C;
    if (C == 0) { n[0] = 0; n[1] = 0; }
    double Iinv = 1.0 / b->I[k];
    double invM = 1.0 / b->M[k];
    double invMs = invM / cnt[k];
    double crn = cross2d(r, n);
    double w = invMs + Iinv * (crn * crn);
    double minvSum = 1.0 / mass[2 * v + 1] + w;
    double dl = (-C - aT * lamB[q]) / (minvSum + aT);
    double pp[2] = {dl * n[0], dl * n[1]};
    lamB[q] = lamB[q] + dl;
    dC[2 * k] = dC[2 * k] + pp[0] * invM;
    dC[2 * k + 1] = dC[2 * k + 1] + pp[1] * invM;
    dA[k] = dA[k] + (0.5 * Iinv) * cross2d(r, pp);
  }
  for (int k = 0; k < nb; ++k) {
    b->COM[2 * k] = b->COM[2 * k] + dC[2 * k];
    b->COM[2 * k + 1] = b->COM[2 * k + 1] + dC[2 * k + 1];
    b->angle[k] = b->angle[k] + dA[k];
  }
  for (int q = 0; q < nbv; ++q) {
    int v = bverts[q], k = vb[v] - 1;
    rigid_pos2d(b->angle[k], ri + 2 * v, b->COM + 2 * k, xi + 2 * v);
  }
  free(cnt); free(dC); free(dA);
}

/* contactConstraints (xpbdLayer2D.m:388-444): elastic contacts by
 * penaltyConstraint2D.projectElastic (penaltyConstraint2D.m:4-26), computed
 * from the incoming xi and written in contact order; then, when rigid
 * contacts exist and useRigidConstacts is set, penaltyConstraint2D.projectRigid
 * (:28-58) for every rigid contact against its body's pose at that point,
 * the per-body sums applied (COM += sum p/M, angle += sum (Iinv/2) r x p),
 * and the rigid contact vertices set to their rigid positions.  nb is the
 * body's rigid-contact count; w = (1/M)/nb + Iinv (r x n)^2 and the
 * denominator is 2 w + alpha~. */
static void contacts2d(bodies2d_t *b, const ora_contacts *c, const int32_t *vb,
                       const double *ri, const double *mass, double *cl, double *xi,
                       double h, double resistance, int useRigid) {
  if (!c || c->K == 0) return;
  double h2 = h * h;
  double *newx = (double *)malloc(sizeof(double) * 2 * c->K);
  for (int i = 0; i < c->K; ++i) {
    int vv = c->vertex[i];
    if (vb[vv] != 0) continue;
    const double *n = c->normal + 2 * i, *pc = c->point + 2 * i;
    double x[2] = {xi[2 * vv], xi[2 * vv + 1]};
    double phi = n[0] * (x[0] - pc[0]) + n[1] * (x[1] - pc[1]);
    if (phi > 0) phi = 0;
    /* gradCminv = n*2*minv, den = gradCminv.n + alpha/h^2, dx = gradCminv * dlambda */
    double C = resistance * phi;
    double aT = c->compliance[i] / h2;
    double num = -(C + aT * cl[i]);
    double gm0 = n[0] * 2 * (1.0 / mass[2 * vv]);
    double gm1 = n[1] * 2 * (1.0 / mass[2 * vv + 1]);
    double den = (gm0 * n[0] + gm1 * n[1]) + aT;
    double dl = num / den;
    newx[2 * i] = x[0] + gm0 * dl;
    newx[2 * i + 1] = x[1] + gm1 * dl;
    cl[i] = cl[i] + dl;
  }
  for (int i = 0; i < c->K; ++i) {
    int vv = c->vertex[i];
    if (vb[vv] != 0) continue;
    xi[2 * vv] = newx[2 * i];
    xi[2 * vv + 1] = newx[2 * i + 1];
  }
  free(newx);
  int nr = 0;
  for (int i = 0; i < c->K; ++i) nr += vb[c->vertex[i]] != 0;
  if (nr < 1 || !useRigid) return;
  int nb = b->nb, n1 = nb > 0 ? nb : 1;
  int *cnt = (int *)calloc(n1, sizeof(int));
  double *dC = (double *)calloc(2 * n1, sizeof(double));
  double *dA = (double *)calloc(n1, sizeof(double));
  for (int i = 0; i < c->K; ++i) {
    int k = vb[c->vertex[i]] - 1;
    if (k >= 0) cnt[k]++;
  }
  for (int i = 0; i < c->K; ++i) {
    int v = c->vertex[i], k = vb[v] - 1;
    if (k < 0) continue;
    const double *n = c->normal + 2 * i, *pc = c->point + 2 * i;
    double rx[2], r[2];
    rigid_pos2d(b->angle[k], ri + 2 * v, b->COM + 2 * k, rx);
    r[0] = rx[0] - b->COM[2 * k];
    r[1] = rx[1] - b->COM[2 * k + 1];
    double phi = (xi[2 * v] - pc[0]) * n[0] + (xi[2 * v + 1] - pc[1]) * n[1];
    if (phi > 0) phi = 0;
    double C = resistance * phi;
    double Iinv = 1.0 / b->I[k];
    double invM = 1.0 / b->M[k];
    double invMs = invM / cnt[k];
    double crn = cross2d(r, n);
    double w = invMs + Iinv * (crn * crn);
    double aT = c->compliance[i] / h2;
    double dl = (-C - aT * cl[i]) / (2 * w + aT);
    double pp[2] = {dl * n[0], dl * n[1]};
    cl[i] = cl[i] + dl;
    dC[2 * k] = dC[2 * k] + pp[0] * invM;
    dC[2 * k + 1] = dC[2 * k + 1] + pp[1] * invM;
    dA[k] = dA[k] + (0.5 * Iinv) * cross2d(r, pp);
  }
  for (int k = 0; k < nb; ++k) {
    b->COM[2 * k] = b->COM[2 * k] + dC[2 * k];
    b->COM[2 * k + 1] = b->COM[2 * k + 1] + dC[2 * k + 1];
    b->angle[k] = b->angle[k] + dA[k];
  }
  for (int i = 0; i < c->K; ++i) {
    int v = c->vertex[i], k = vb[v] - 1;
    if (k < 0) continue;
    rigid_pos2d(b->angle[k], ri + 2 * v, b->COM + 2 * k, xi + 2 * v);
  }
  free(cnt); free(dC); free(dA);
}

/* rotateAllVerts2D.m:1-7: rigid vertices to R(angle) ri + COM. */
static void rotate_all_verts2d(int nV, const int32_t *vb, const double *ri,
                               const bodies2d_t *b, double *xi) {
  for (int i = 0; i < nV; ++i) {
    int k = vb[i] - 1;
    if (k < 0) continue;
    rigid_pos2d(b->angle[k], ri + 2 * i, b->COM + 2 * k, xi + 2 * i);
  }
}

/* ------------------------------------------------------------------ */
/* 2D substep                                                           */
/* ------------------------------------------------------------------ */

/* xpbdLayer2D.integrate (xpbdLayer2D.m:793-917) with layerSolve (:606-738)
 * and projectGS (:105-201).  Per layer the rigid bodies are rebuilt from the
 * current positions (:676-681); unlike 3D they are not stepped.  Contact
 * lambdas start at zero in every layer (projectGS does not return them).
 * Pins: the numpy restatement of the 2D layer (tests), plus the pieces'
 * closed forms. */
int ora_integrate2d(const ora_mesh *m, ora_settings *s, double h, double *p,
                    double *v, const double *prevp, const double *prevPrevp,
                    const ora_contacts *c, ora_step_out *out) {
  int nV = m->nV, nT = m->nT, N2 = 2 * nV;
  if (s->runUntil >= 0 && !s->computeResiduals) return -3;
  double *f = (double *)calloc(N2, sizeof(double));
  for (int i = 0; i < nV; ++i) f[2 * i + 1] = f[2 * i + 1] + m->mass[2 * i + 1] * s->gravity;
  layers_t L;
  build_layers(m, s, p, prevPrevp, h, NULL, NULL, &L);
  if (out && out->setsPerLayer)
    memcpy(out->setsPerLayer, L.sets, sizeof(int32_t) * (size_t)nT * L.nSorted);
  if (out && out->numRigids) memcpy(out->numRigids, L.numRigids, sizeof(int32_t) * L.nSorted);
  uint8_t *pinnedDof = (uint8_t *)calloc(N2, 1);
  for (int q = 0; q < m->nPinnedDOFs; ++q) pinnedDof[m->pinnedDOFs[q]] = 1;
  double *xi = (double *)malloc(sizeof(double) * N2);
  memcpy(xi, p, sizeof(double) * N2);
  for (int i = 0; i < N2; ++i) {
    if (pinnedDof[i]) continue;
    v[i] = v[i] + h * f[i] / m->mass[i];
    xi[i] = xi[i] + h * v[i];
  }
  double *lambda = (double *)calloc(3 * (size_t)(nT > 0 ? nT : 1), sizeof(double));
  int K = c ? c->K : 0;
  double *clL = (double *)calloc(K > 0 ? K : 1, sizeof(double));
  monitor_t mo;
  double dummy[1];
  mo.arr = out && out->residualArray ? out->residualArray : dummy;
  mo.cap = out && out->residualArray ? out->residualCap : 0;
  mo.count = 0;
  mo.errorChangeDecreased = 0;
  monitor_push(&mo, 0.0);
  uint8_t *isEl = (uint8_t *)malloc(nT > 0 ? nT : 1);
  double *ri = (double *)calloc(N2 > 0 ? N2 : 1, sizeof(double));
  int32_t *bverts = (int32_t *)malloc(sizeof(int32_t) * (nV > 0 ? nV : 1));
  int64_t projections = 0;
  int layersRun = 0;
  for (int layer = 0; layer < s->nLayers; ++layer) {
    double setPercentStop = (layer == 0) ? s->stopAtImprovementPercent : 0;
    int isLast = layer == s->nLayers - 1;
    int col = L.map[layer];
    const int32_t *sets = L.sets + (size_t)nT * col;
    const int32_t *vb = L.vertexRigidBody + (size_t)nV * col;
    const uint8_t *isB = L.isBoundary + (size_t)nV * col;
    int nElastic = 0;
    for (int e = 0; e < nT; ++e) { isEl[e] = sets[e] == 0; nElastic += isEl[e]; }
    bodies2d_t b;
    bodies2d_alloc(&b, L.numRigids[col]);
    rigid_properties2d(nV, vb, xi, m->mass, &b, ri);
    int nbv = 0;
    for (int i = 0; i < nV; ++i) if (isB[i]) bverts[nbv++] = i;
    double *lamB = (double *)calloc(nbv > 0 ? nbv : 1, sizeof(double));
    for (int i = 0; i < K; ++i) clL[i] = 0;
    int it = 1, stopSolve = 0, iters = 0;
    while (it <= s->iterations[layer] || (!stopSolve && s->runUntil >= 0 && isLast)) {
      it = it + 1;
      ora_stvk_voigt_constraints2d(m, isEl, h, lambda, xi, prevp);
      projections += nElastic;
      boundary2d(&b, nbv, bverts, vb, ri, m->mass, s->boundaryCompliance, lamB, xi, h);
      contacts2d(&b, c, vb, ri, m->mass, clL, xi, h, s->contactResistance, s->useRigidContacts);
      for (int q = 0; q < m->nPinnedDOFs; ++q) xi[m->pinnedDOFs[q]] = prevp[m->pinnedDOFs[q]];
      ++iters;
      int endLayer = 0;
      if (s->computeResiduals) {
        double r = ora_residual2d(m, h, lambda, xi, prevp, c, clL, s->contactResistance);
        monitor_step(&mo, 0, r, s, &setPercentStop, isLast, INFINITY, h, &endLayer, &stopSolve);
      }
      if (endLayer || stopSolve) break;
    }
    rotate_all_verts2d(nV, vb, ri, &b, xi);
    free(lamB);
    bodies2d_free(&b);
    if (out) out->itersPerLayer[layer] = iters;
    layersRun = layer + 1;
    if (stopSolve) break;
  }
  for (int i = 0; i < N2; ++i) {
    v[i] = (xi[i] - prevp[i]) / h;
    p[i] = xi[i];
  }
  if (c && c->K > 0) ora_contact_velocity_update(2, c, v, h, s->gravity, s->restitution);
  if (out) {
    if (out->lambda) memcpy(out->lambda, lambda, sizeof(double) * 3 * nT);
    out->residualCount = mo.count;
    out->layersRun = layersRun;
    out->elasticProjections = projections;
  }
  free(f); free(pinnedDof); free(xi); free(lambda); free(clL); free(isEl); free(ri); free(bverts);
  layers_free(&L);
  return 0;
}

/* ------------------------------------------------------------------ */
/* test hooks: the rigid-layer steps on explicit body state             */
/* ------------------------------------------------------------------ */

static void bodies_view(bodies_t *b, int nb, double *M, double *COM, double *COMDOT,
                        double *R, double *I0, double *I, double *omega) {
  b->nb = nb; b->M = M; b->COM = COM; b->COMDOT = COMDOT; b->R = R; b->I0 = I0;
  b->I = I; b->omega = omega;
}

/* makeRigidsFromConnectivity on explicit arrays (vb: 1-based body per
 * vertex, 0 = not rigid).  Outputs per body M, COM, COMDOT, I0 (= I),
 * omega, R (= eye) and ri (3nV). */
void ora_rigid_properties3d(int nV, int nb, const int32_t *vb, const double *p,
                            const double *v, const double *mass, double *M,
                            double *COM, double *COMDOT, double *R, double *I0,
                            double *I, double *omega, double *ri) {
  bodies_t b;
  bodies_view(&b, nb, M, COM, COMDOT, R, I0, I, omega);
  rigid_properties3d(nV, vb, p, v, mass, &b, ri);
}

/* rigidSymplecticImperativeStep on explicit arrays (one body). */
void ora_rigid_step3d(double h, double *COM, const double *COMDOT, double *R,
                      const double *omega, const double *I0, double *I) {
  for (int d = 0; d < 3; ++d) COM[d] = COM[d] + h * COMDOT[d];
  double dR[9], Rn[9], tmp[9], Rt[9];
  ora_exp_rodrigues(omega, h, dR);
  matmul_n(3, dR, R, Rn);
  memcpy(R, Rn, sizeof Rn);
  matmul_n(3, R, I0, tmp);
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Rt[i + 3 * j] = R[j + 3 * i];
  matmul_n(3, tmp, Rt, I);
}

void ora_boundary3d(int nb, double *M, double *COM, double *R, double *I0, double *I,
                    int nbv, const int32_t *bverts, const int32_t *vb, const double *ri,
                    const double *mass, double compliance, double *lamB, double *xi,
                    double h) {
  bodies_t b;
  bodies_view(&b, nb, M, COM, NULL, R, I0, I, NULL);
  boundary3d(&b, nbv, bverts, vb, ri, mass, compliance, lamB, xi, h);
}

void ora_contacts3d(int nb, double *M, double *COM, double *R, double *I0, double *I,
                    const ora_contacts *c, const int32_t *vb, const double *ri,
                    const double *mass, double *cl, double *xi, double h,
                    double resistance, int useRigid) {
  bodies_t b;
  bodies_view(&b, nb, M, COM, NULL, R, I0, I, NULL);
  contacts3d(&b, c, vb, ri, mass, cl, xi, h, resistance, useRigid);
}

/* ------------------------------------------------------------------ */
/* 2D NeoHookean constraint pair (row a7)                               */
/* ------------------------------------------------------------------ */

/* mexNeoHookeanDConstraint.cpp:46-138: C_D = ||F||_F with
 * F = [p0-p2, p1-p2] DmInv, alpha~ = 1/(area h^2 mu); one XPBD scalar step
 * dlambda = -(C + alpha~ lambda) / (sum_k m_k^-1 gradC_k^2 + alpha~),
 * dx_k = m_k^-1 gradC_k dlambda.  Expression tree of the MEX kept.
 * Pins: C against ||F||_F, gradient against finite differences, dlambda
 * against the scalar XPBD closed form (tests). */
void ora_neohookean_d2d(double h, const double *x, const double *D, double mu,
                        double area, double lambda, const double *mass,
                        double *dx, double *dlambda, double *C) {
  double t2 = D[0] + D[1], t3 = D[2] + D[3];
  double t4 = 1.0 / area, t5 = 1.0 / (h * h);
  double t6 = 1.0 / mass[0], t7 = 1.0 / mass[1], t8 = 1.0 / mass[2], t9 = 1.0 / mass[3],
         t10 = 1.0 / mass[4], t11 = 1.0 / mass[5];
  double t12 = 1.0 / mu;
  double t13 = -x[4], t14 = -x[5];
  double t15 = t13 + x[0], t16 = t14 + x[1], t17 = t13 + x[2], t18 = t14 + x[3];
  double t27 = t4 * t5 * t12;
  double t19 = D[0] * t15, t20 = D[2] * t15, t21 = D[0] * t16, t22 = D[2] * t16;
  double t23 = D[1] * t17, t24 = D[3] * t17, t25 = D[1] * t18, t26 = D[3] * t18;
  double t28 = lambda * t27;
  double t29 = t19 + t23, t30 = t20 + t24, t31 = t21 + t25, t32 = t22 + t26;
  double t33 = t29 * t29, t34 = t30 * t30, t35 = t31 * t31, t36 = t32 * t32;
  double t37 = D[0] * t29 * 2.0, t38 = D[2] * t30 * 2.0, t39 = D[0] * t31 * 2.0,
         t40 = D[2] * t32 * 2.0, t41 = D[1] * t29 * 2.0, t42 = D[3] * t30 * 2.0,
         t43 = D[1] * t31 * 2.0, t44 = D[3] * t32 * 2.0;
  double t45 = t2 * t29 * 2.0, t46 = t3 * t30 * 2.0, t47 = t2 * t31 * 2.0, t48 = t3 * t32 * 2.0;
  double t49 = t37 + t38, t50 = t39 + t40, t51 = t41 + t42, t52 = t43 + t44;
  double t57 = t45 + t46, t58 = t47 + t48;
  double t61 = t33 + t34 + t35 + t36;
  double t53 = t49 * t49, t54 = t50 * t50, t55 = t51 * t51, t56 = t52 * t52;
  double t59 = t57 * t57, t60 = t58 * t58;
  double t62 = 1.0 / t61, t63 = sqrt(t61), t64 = 1.0 / t63, t65 = t28 + t63;
  double t66 = (t6 * t53 * t62) / 4.0, t67 = (t7 * t54 * t62) / 4.0,
         t68 = (t8 * t55 * t62) / 4.0, t69 = (t9 * t56 * t62) / 4.0,
         t70 = (t10 * t59 * t62) / 4.0, t71 = (t11 * t60 * t62) / 4.0;
  double t72 = t27 + t66 + t67 + t68 + t69 + t70 + t71;
  double t73 = 1.0 / t72;
  dx[0] = t6 * t49 * t64 * t65 * t73 * (-1.0 / 2.0);
  dx[1] = t7 * t50 * t64 * t65 * t73 * (-1.0 / 2.0);
  dx[2] = t8 * t51 * t64 * t65 * t73 * (-1.0 / 2.0);
  dx[3] = t9 * t52 * t64 * t65 * t73 * (-1.0 / 2.0);
  dx[4] = (t10 * t57 * t64 * t65 * t73) / 2.0;
  dx[5] = (t11 * t58 * t64 * t65 * t73) / 2.0;
  *dlambda = -t65 * t73;
  if (C) *C = t63;
}

/* mexNeoHookeanHConstraint.cpp:47-138: C_H = det F - gamma,
 * gamma = 1 + mu/lambda (Mesh.m:244), alpha~ = 1/(area h^2 lambda). */
void ora_neohookean_h2d(double h, const double *x, const double *D, double lamL,
                        double area, double lambda, const double *mass, double gamma,
                        double *dx, double *dlambda, double *C) {
  double t2 = D[0] * D[3] * x[0], t3 = D[2] * D[1] * x[0], t4 = D[0] * D[3] * x[1],
         t5 = D[2] * D[1] * x[1], t6 = D[0] * D[3] * x[2], t7 = D[2] * D[1] * x[2],
         t8 = D[0] * D[3] * x[3], t9 = D[2] * D[1] * x[3], t10 = D[0] * D[3] * x[4],
         t11 = D[2] * D[1] * x[4], t12 = D[0] * D[3] * x[5], t13 = D[2] * D[1] * x[5];
  double t14 = 1.0 / area, t15 = 1.0 / (h * h);
  double t16 = 1.0 / mass[0], t17 = 1.0 / mass[1], t18 = 1.0 / mass[2], t19 = 1.0 / mass[3],
         t20 = 1.0 / mass[4], t21 = 1.0 / mass[5];
  double t22 = 1.0 / lamL;
  double t23 = t2 * x[3], t24 = t4 * x[2], t25 = t3 * x[3], t26 = t5 * x[2], t27 = t2 * x[5],
         t28 = t4 * x[4], t29 = t3 * x[5], t30 = t5 * x[4], t31 = t6 * x[5], t32 = t8 * x[4],
         t33 = t7 * x[5], t34 = t9 * x[4];
  double t35 = -t3, t36 = -t5, t37 = -t6, t38 = -t7, t39 = -t8, t40 = -t9, t41 = -t10, t42 = -t12;
  double t49 = t14 * t15 * t22;
  double t43 = -t23, t44 = -t26, t45 = -t28, t46 = -t29, t47 = -t31, t48 = -t34;
  double t50 = lambda * t49;
  double t52 = t2 + t7 + t35 + t37, t53 = t2 + t11 + t35 + t41, t54 = t4 + t9 + t36 + t39,
         t55 = t4 + t13 + t36 + t42, t56 = t6 + t11 + t38 + t41, t57 = t8 + t13 + t40 + t42;
  double t51 = -t50;
  double t58 = t52 * t52, t59 = t53 * t53, t60 = t54 * t54, t61 = t55 * t55, t62 = t56 * t56,
         t63 = t57 * t57;
  double t64 = t21 * t58, t65 = t19 * t59, t66 = t20 * t60, t67 = t17 * t62, t68 = t18 * t61,
         t69 = t16 * t63;
  double t70 = gamma + t24 + t25 + t27 + t30 + t32 + t33 + t43 + t44 + t45 + t46 + t47 + t48 + t51;
  double t71 = t49 + t64 + t65 + t66 + t67 + t68 + t69;
  double t72 = 1.0 / t71;
  dx[0] = t16 * t57 * t70 * t72;
  dx[1] = -t17 * t56 * t70 * t72;
  dx[2] = -t18 * t55 * t70 * t72;
  dx[3] = t19 * t53 * t70 * t72;
  dx[4] = t20 * t54 * t70 * t72;
  dx[5] = -t21 * t52 * t70 * t72;
  *dlambda = t70 * t72;
  if (C) *C = -gamma + t23 - t24 - t25 + t26 - t27 + t28 + t29 - t30 + t31 - t32 - t33 + t34;
}

/* xpbdLayer2D.neoHookeanConstraints (xpbdLayer2D.m:260-294): per color, per
 * element, the D and H steps are both evaluated at the positions the color
 * started from (xi is written back after the color), their dx summed.
 * lambda: nT x 2 row-major (lambda_D, lambda_H). */
void ora_neohookean_constraints2d(const ora_mesh *m, const uint8_t *isElastic, double h,
                                  const double *mu, const double *lamL, double *lambda,
                                  double *xi) {
  for (int c = 1; c <= m->numColors; ++c)
    for (int e = 0; e < m->nT; ++e) {
      if (m->color[e] != c || !isElastic[e]) continue;
      double x[6], ms[6], dxD[6], dxH[6], dlD, dlH;
      for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 2; ++d) {
          int dof = 2 * m->T[3 * e + a] + d;
          x[2 * a + d] = xi[dof];
          ms[2 * a + d] = m->mass[dof];
        }
      double gamma = 1 + mu[e] / lamL[e];
      ora_neohookean_d2d(h, x, m->DmInv + 4 * e, mu[e], m->vol[e], lambda[2 * e], ms, dxD, &dlD, NULL);
      ora_neohookean_h2d(h, x, m->DmInv + 4 * e, lamL[e], m->vol[e], lambda[2 * e + 1], ms, gamma,
                         dxH, &dlH, NULL);
      lambda[2 * e] = lambda[2 * e] + dlD;
      lambda[2 * e + 1] = lambda[2 * e + 1] + dlH;
      for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 2; ++d) {
          int dof = 2 * m->T[3 * e + a] + d;
          xi[dof] = (x[2 * a + d] + dxD[2 * a + d]) + dxH[2 * a + d];
        }
    }
}

/* NeoHookean pair on a tetrahedron (extension of the 2D pair above).
 * F[r][c] = sum_k (x_{k+1} - x_0)[r] DmInv(k, c), DmInv column-major.
 * d||F||/dF = F/||F||; d det F/dF = cof(F).  The gradient w.r.t. vertex
 * k+1 is (dC/dF) times row k of DmInv, and vertex 0 takes minus their sum.
 * Pins: central finite differences of ||F|| and det F; the rest state with
 * the equilibrium multipliers lambda_D = -sqrt(3) vol h^2 mu and
 * lambda_H = vol h^2 mu is a fixed point (tests). */
static void nh_F3(const double *D, const double *x, double F[3][3]) {
  double e[3][3];
  for (int k = 0; k < 3; ++k)
    for (int r = 0; r < 3; ++r) e[k][r] = x[3 * (k + 1) + r] - x[r];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      F[r][c] = e[0][r] * D[3 * c] + e[1][r] * D[1 + 3 * c] + e[2][r] * D[2 + 3 * c];
}

static void nh_vertex_grads(const double *D, double P[3][3], double *g) {
  for (int k = 0; k < 3; ++k)
    for (int r = 0; r < 3; ++r)
      g[3 * (k + 1) + r] = P[r][0] * D[k] + P[r][1] * D[k + 3] + P[r][2] * D[k + 6];
  for (int r = 0; r < 3; ++r) g[r] = -(g[3 + r] + g[6 + r] + g[9 + r]);
}

void ora_neohookean_grad3d(const double *D, const double *x, double *gD, double *gH,
                           double *CD, double *J) {
  double F[3][3], PD[3][3], PH[3][3];
  nh_F3(D, x, F);
  double n2 = 0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) n2 = n2 + F[r][c] * F[r][c];
  double n = sqrt(n2);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) PD[r][c] = F[r][c] / n;
  PH[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
  PH[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
  PH[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
  PH[1][0] = F[0][2] * F[2][1] - F[0][1] * F[2][2];
  PH[1][1] = F[0][0] * F[2][2] - F[0][2] * F[2][0];
  PH[1][2] = F[0][1] * F[2][0] - F[0][0] * F[2][1];
  PH[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
  PH[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
  PH[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
  if (gD) nh_vertex_grads(D, PD, gD);
  if (gH) nh_vertex_grads(D, PH, gH);
  if (CD) *CD = n;
  if (J) *J = F[0][0] * PH[0][0] + F[0][1] * PH[0][1] + F[0][2] * PH[0][2];
}

static void nh_step(double h, double C, double atil, double lambda, const double *g,
                    const double *mass, double *dx, double *dlambda) {
  double s = atil;
  for (int q = 0; q < 12; ++q) s = s + (1.0 / mass[q]) * g[q] * g[q];
  double dl = -(C + atil * lambda) / s;
  for (int q = 0; q < 12; ++q) dx[q] = (1.0 / mass[q]) * g[q] * dl;
  *dlambda = dl;
}

void ora_neohookean_d3d(double h, const double *x, const double *D, double mu, double vol,
                        double lambda, const double *mass, double *dx, double *dlambda,
                        double *C) {
  double g[12], n;
  ora_neohookean_grad3d(D, x, g, NULL, &n, NULL);
  double atil = (1.0 / vol) * (1.0 / (h * h)) * (1.0 / mu);
  nh_step(h, n, atil, lambda, g, mass, dx, dlambda);
  if (C) *C = n;
}

void ora_neohookean_h3d(double h, const double *x, const double *D, double lamL, double vol,
                        double lambda, const double *mass, double gamma, double *dx,
                        double *dlambda, double *C) {
  double g[12], J;
  ora_neohookean_grad3d(D, x, NULL, g, NULL, &J);
  double atil = (1.0 / vol) * (1.0 / (h * h)) * (1.0 / lamL);
  nh_step(h, J - gamma, atil, lambda, g, mass, dx, dlambda);
  if (C) *C = J - gamma;
}

static void nh_project3d(const ora_mesh *m, int e, double h, double *lambda, double *xi) {
  double x[12], ms[12], dxD[12], dxH[12], dlD, dlH;
  for (int a = 0; a < 4; ++a)
    for (int d = 0; d < 3; ++d) {
      int dof = 3 * m->T[4 * e + a] + d;
      x[3 * a + d] = xi[dof];
      ms[3 * a + d] = m->mass[dof];
    }
  double gamma = 1 + m->nhMu[e] / m->nhLam[e];
  ora_neohookean_d3d(h, x, m->DmInv + 9 * e, m->nhMu[e], m->vol[e], lambda[6 * e], ms, dxD,
                     &dlD, NULL);
  ora_neohookean_h3d(h, x, m->DmInv + 9 * e, m->nhLam[e], m->vol[e], lambda[6 * e + 1], ms,
                     gamma, dxH, &dlH, NULL);
  lambda[6 * e] = lambda[6 * e] + dlD;
  lambda[6 * e + 1] = lambda[6 * e + 1] + dlH;
  for (int a = 0; a < 4; ++a)
    for (int d = 0; d < 3; ++d) {
      int dof = 3 * m->T[4 * e + a] + d;
      xi[dof] = (x[3 * a + d] + dxD[3 * a + d]) + dxH[3 * a + d];
    }
}

void ora_neohookean_constraints3d(const ora_mesh *m, const uint8_t *isElastic, double h,
                                  double *lambda, double *xi) {
  for (int c = 1; c <= m->numColors; ++c)
    for (int e = 0; e < m->nT; ++e) {
      if (m->color[e] != c || !isElastic[e] || !m->nhModel || !m->nhModel[e]) continue;
      nh_project3d(m, e, h, lambda, xi);
    }
}

/* Residual terms of a NeoHookean element: max(|C| - |alpha~ lambda|, 0)^2
 * for C_D and C_H, as the StVK terms of ora_residual3d. */
static double nh_residual_sq3d(const ora_mesh *m, int e, double h, const double *lambda,
                               const double *xi) {
  double x[12], n, J;
  for (int a = 0; a < 4; ++a)
    for (int d = 0; d < 3; ++d) x[3 * a + d] = xi[3 * m->T[4 * e + a] + d];
  ora_neohookean_grad3d(m->DmInv + 9 * e, x, NULL, NULL, &n, &J);
  double iv = (1.0 / m->vol[e]) * (1.0 / (h * h));
  double C[2] = {n, J - (1 + m->nhMu[e] / m->nhLam[e])};
  double al[2] = {iv * (1.0 / m->nhMu[e]) * lambda[6 * e], iv * (1.0 / m->nhLam[e]) * lambda[6 * e + 1]};
  double acc = 0;
  for (int i = 0; i < 2; ++i) {
    double aC = fabs(C[i]), corr = fabs(al[i]);
    double t = aC - (aC < corr ? aC : corr);
    acc = acc + t * t;
  }
  return acc;
}

/* ------------------------------------------------------------------ */
/* preconditioned conjugate gradient (row a13)                          */
/* ------------------------------------------------------------------ */

static void csr_mv(int n, const int32_t *rp, const int32_t *ci, const double *v,
                   const double *x, double *y) {
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) s = s + v[k] * x[ci[k]];
    y[i] = s;
  }
}

static double dotn(int n, const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + a[i] * b[i];
  return s;
}

/* solvePCG.m:1-57 with a diagonal preconditioner z = pinv .* r: fixed
 * iteration count, alpha = 0 when d'Ad == 0, resvec(i) = ||A x - b||.
 * A is CSR (0-based).  x: in x0, out solution.  resvec may be NULL.
 * Pins: equals numpy.linalg.solve after n iterations on small SPD systems;
 * hand-computed first iteration; the d'Ad == 0 guard (tests). */
void ora_solve_pcg(int n, const int32_t *rp, const int32_t *ci, const double *v,
                   const double *b, const double *pinv, int iterations, double *x,
                   double *resvec) {
  double *r = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
  double *d = (double *)malloc(sizeof(double) * n), *Ap = (double *)malloc(sizeof(double) * n);
  csr_mv(n, rp, ci, v, x, Ap);
  for (int i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
  for (int i = 0; i < n; ++i) z[i] = pinv[i] * r[i];
  for (int i = 0; i < n; ++i) d[i] = z[i];
  double rTz = dotn(n, r, z);
  for (int it = 0; it < iterations; ++it) {
    csr_mv(n, rp, ci, v, d, Ap);
    double qAp = dotn(n, d, Ap);
    double alpha = (qAp == 0) ? 0.0 : rTz / qAp;
    for (int i = 0; i < n; ++i) x[i] = x[i] + alpha * d[i];
    for (int i = 0; i < n; ++i) r[i] = r[i] - alpha * Ap[i];
    for (int i = 0; i < n; ++i) z[i] = pinv[i] * r[i];
    double rTznew = dotn(n, r, z);
    double beta = rTznew / rTz;
    rTz = rTznew;
    for (int i = 0; i < n; ++i) d[i] = z[i] + beta * d[i];
    if (resvec) {
      csr_mv(n, rp, ci, v, x, Ap);
      double s = 0.0;
      for (int i = 0; i < n; ++i) s = s + (Ap[i] - b[i]) * (Ap[i] - b[i]);
      resvec[it] = sqrt(s);
    }
  }
  free(r); free(z); free(d); free(Ap);
}

/* ------------------------------------------------------------------ */
/* test hooks: the 2D rigid-layer steps on explicit body state          */
/* ------------------------------------------------------------------ */

void ora_rigid_properties2d(int nV, int nb, const int32_t *vb, const double *p,
                            const double *mass, double *M, double *COM, double *I,
                            double *ri) {
  bodies2d_t b;
  bodies2d_alloc(&b, nb);
  rigid_properties2d(nV, vb, p, mass, &b, ri);
  for (int k = 0; k < nb; ++k) {
    M[k] = b.M[k];
    COM[2 * k] = b.COM[2 * k];
    COM[2 * k + 1] = b.COM[2 * k + 1];
    I[k] = b.I[k];
  }
  bodies2d_free(&b);
}

static void bodies2d_wrap(bodies2d_t *b, int nb, double *M, double *COM, double *angle,
                          double *I) {
  b->nb = nb; b->M = M; b->COM = COM; b->angle = angle; b->I = I;
}

void ora_boundary2d(int nb, double *M, double *COM, double *angle, double *I, int nbv,
                    const int32_t *bverts, const int32_t *vb, const double *ri,
                    const double *mass, double alpha, double *lamB, double *xi, double h) {
  bodies2d_t b;
  bodies2d_wrap(&b, nb, M, COM, angle, I);
  boundary2d(&b, nbv, bverts, vb, ri, mass, alpha, lamB, xi, h);
}

void ora_contacts2d(int nb, double *M, double *COM, double *angle, double *I,
                    const ora_contacts *c, const int32_t *vb, const double *ri,
                    const double *mass, double *cl, double *xi, double h, double resistance,
                    int useRigid) {
  bodies2d_t b;
  bodies2d_wrap(&b, nb, M, COM, angle, I);
  contacts2d(&b, c, vb, ri, mass, cl, xi, h, resistance, useRigid);
}

