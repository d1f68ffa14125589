// Host-side parts of libmlxpbd: error reporting, scene precompute (graph
// coloring, rest-shape inverses, masses, compliance matrices) and the layer
// map construction (union-find over shared vertices, boundary vertices).
// These run once per scene or once per substep on small integer data; the
// per-iteration work is in mlx_solver.hip.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "mlx_host.h"
#include "mlxpbd.h"

namespace mlx {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

const char *last_error() { return g_err; }

// Greedy coloring in element order (naiveGraphColoring.m:1-19): each element
// takes the smallest color not yet used at any of its vertices.  Colors are
// tracked as a growable bitset per vertex.
int graph_coloring(const int32_t *T0, int64_t nT, int vpe, int64_t nV,
                   int32_t *color, int32_t *numColors) {
  int words = 1;
  std::vector<uint64_t> used((size_t)nV, 0);
  int maxc = 0;
  for (int64_t e = 0; e < nT; ++e) {
    const int32_t *v = T0 + e * vpe;
    int c = 0;
    for (;;) {
      int w = c / 64;
      if (w >= words) {
        int nw = words * 2;
        std::vector<uint64_t> grown((size_t)nV * nw, 0);
        for (int64_t i = 0; i < nV; ++i)
          std::memcpy(&grown[i * nw], &used[i * words], sizeof(uint64_t) * words);
        used.swap(grown);
        words = nw;
      }
      uint64_t mask = 0;
      for (int k = 0; k < vpe; ++k) mask |= used[(size_t)v[k] * words + w];
      uint64_t freebits = ~mask & (~0ull << (c % 64));
      if (freebits) {
        c = w * 64 + __builtin_ctzll(freebits);
        break;
      }
      c = (w + 1) * 64;
    }
    for (int k = 0; k < vpe; ++k) used[(size_t)v[k] * words + c / 64] |= 1ull << (c % 64);
    color[e] = c + 1;
    maxc = std::max(maxc, c + 1);
  }
  *numColors = maxc;
  return MLX_OK;
}

// 3x3 inverse by cofactors, column-major.
static void inv3(const double *A, double *R, double *detOut) {
  double a = A[0], b = A[3], c = A[6], d = A[1], e = A[4], f = A[7], g = A[2],
         h = A[5], i = A[8];
  double c00 = e * i - f * h, c01 = -(d * i - f * g), c02 = d * h - e * g;
  double det = a * c00 + b * c01 + c * c02;
  double id = 1.0 / det;
  R[0] = c00 * id;
  R[1] = c01 * id;
  R[2] = c02 * id;
  R[3] = -(b * i - c * h) * id;
  R[4] = (a * i - c * g) * id;
  R[5] = -(a * h - b * g) * id;
  R[6] = (b * f - c * e) * id;
  R[7] = -(a * f - c * d) * id;
  R[8] = (a * e - b * d) * id;
  if (detOut) *detOut = det;
}

// Inverse of the isotropic Voigt stiffness K = [[a b b 0 0 0]..; shear s]:
// the normal block [[a b b],[b a b],[b b a]] has inverse
// p I + q (11^T - I) with p = (a+b)/((a-b)(a+2b)), q = -b/((a-b)(a+2b)).
static void inv_voigt(int n, double a, double b, double s, double *alpha) {
  std::memset(alpha, 0, sizeof(double) * n * n);
  int nn = (n == 6) ? 3 : 2;
  double p, q;
  if (nn == 3) {
    double den = (a - b) * (a + 2 * b);
    p = (a + b) / den;
    q = -b / den;
  } else {
    double den = a * a - b * b;
    p = a / den;
    q = -b / den;
  }
  for (int i = 0; i < nn; ++i)
    for (int j = 0; j < nn; ++j) alpha[i + n * j] = (i == j) ? p : q;
  for (int i = nn; i < n; ++i) alpha[i + n * i] = 1.0 / s;
}

int precompute3d(const double *p, int64_t nV, const int32_t *T0, int64_t nT,
                 const double *rho, const double *mu, const double *lam,
                 const double *alpha1, double *DmInv, double *vol, double *mass,
                 double *alpha, double *beta) {
  std::fill(mass, mass + 3 * nV, 0.0);
  for (int64_t e = 0; e < nT; ++e) {
    const int32_t *t = T0 + 4 * e;
    double Dm[9];
    for (int a = 0; a < 3; ++a)
      for (int d = 0; d < 3; ++d) Dm[d + 3 * a] = p[3 * t[a + 1] + d] - p[3 * t[0] + d];
    double det;
    inv3(Dm, DmInv + 9 * e, &det);
    if (!(det != 0.0) || !std::isfinite(det)) {
      set_error("element %lld is degenerate (det(Dm) = %g)", (long long)e, det);
      return MLX_ERR_ARG;
    }
    vol[e] = det / 6.0;
  }
  for (int64_t e = 0; e < nT; ++e) {
    double mv = (vol[e] * rho[e]) / 4;
    for (int a = 0; a < 4; ++a)
      for (int d = 0; d < 3; ++d) mass[3 * T0[4 * e + a] + d] += mv;
  }
  for (int64_t e = 0; e < nT; ++e) {
    double l = lam[e], dm = 2 * mu[e], dv = l + 2 * dm;
    double *A = alpha + 36 * e, *B = beta + 36 * e;
    inv_voigt(6, dv, l, dm, A);
    std::memset(B, 0, sizeof(double) * 36);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) B[i + 6 * j] = ((i == j) ? dv : l) * alpha1[e];
    for (int i = 3; i < 6; ++i) B[i + 6 * i] = dm * alpha1[e];
  }
  return MLX_OK;
}

int precompute2d(const double *p, int64_t nV, const int32_t *T0, int64_t nT,
                 const double *rho, const double *mu, const double *lam,
                 const double *alpha1, double *DmInv, double *area, double *mass,
                 double *alpha, double *beta) {
  std::fill(mass, mass + 2 * nV, 0.0);
  for (int64_t e = 0; e < nT; ++e) {
    const int32_t *t = T0 + 3 * e;
    const double *pi = p + 2 * t[0], *pj = p + 2 * t[1], *pk = p + 2 * t[2];
    double a = pi[0] - pk[0], c = pi[1] - pk[1], b = pj[0] - pk[0], d = pj[1] - pk[1];
    double det = a * d - b * c;
    if (!(det != 0.0) || !std::isfinite(det)) {
      set_error("triangle %lld is degenerate", (long long)e);
      return MLX_ERR_ARG;
    }
    double *M = DmInv + 4 * e;
    M[0] = d / det;
    M[1] = -c / det;
    M[2] = -b / det;
    M[3] = a / det;
    double u0 = pj[0] - pi[0], u1 = pj[1] - pi[1], w0 = pk[0] - pi[0], w1 = pk[1] - pi[1];
    area[e] = 0.5 * std::fabs(u0 * w1 - w0 * u1);
  }
  for (int64_t e = 0; e < nT; ++e)
    for (int j = 0; j < 3; ++j) {
      int v = T0[3 * e + j];
      double mm = (1.0 / 3) * area[e] * rho[e];
      mass[2 * v] += mm;
      mass[2 * v + 1] += mm;
    }
  for (int64_t e = 0; e < nT; ++e) {
    double l = lam[e], dv = l + 2 * mu[e], s = 2 * mu[e];
    inv_voigt(3, dv, l, s, alpha + 9 * e);
    double *B = beta + 9 * e;
    double K[9] = {dv, l, 0, l, dv, 0, 0, 0, s};
    for (int k = 0; k < 9; ++k) B[k] = K[k] * alpha1[e];
  }
  return MLX_OK;
}

// Connected components of "shares a vertex" over the inserted elements,
// grown layer by layer (mexUnionFindByVertice.cpp:117-178).  The component
// representative is its smallest element, so numbering components by their
// representative reproduces the reference's ascending-element numbering.
int union_find_by_vertice(const int32_t *order0, int64_t nT,
                          const double *sortedLayers, int L,
                          const uint8_t *forcedElastic, const int32_t *T0,
                          int vpe, int64_t nV, int32_t *sets, int32_t *numRigids) {
  std::vector<int32_t> parent(nT, -1), lastAt(nV, -1), id(nT, 0);
  auto find = [&](int32_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  int64_t count = 0, start = 0;
  for (int l = 0; l < L; ++l) {
    int64_t end = (int64_t)std::floor((sortedLayers[l] / 100.0) * (double)nT);
    if (end > nT) end = nT;
    for (int64_t q = start; q < end; ++q) {
      int32_t e = order0[q];
      if (forcedElastic[e]) continue;
      parent[e] = e;
      ++count;
      for (int k = 0; k < vpe; ++k) {
        int32_t v = T0[(size_t)e * vpe + k];
        int32_t o = lastAt[v];
        lastAt[v] = e;
        if (o < 0) continue;
        int32_t a = find(e), b = find(o);
        if (a == b) continue;
        if (a < b) parent[b] = a; else parent[a] = b;
        --count;
      }
    }
    if (end > start) start = end;
    int32_t next = 1;
    int32_t *col = sets + (size_t)nT * l;
    for (int64_t e = 0; e < nT; ++e) {
      if (parent[e] < 0) { col[e] = 0; continue; }
      int32_t r = find((int32_t)e);
      if (r == e) id[e] = next++;
      col[e] = id[r];
    }
    numRigids[l] = (int32_t)count;
  }
  return MLX_OK;
}

int find_boundary_vertices(const int32_t *T0, int64_t nT, int vpe, int64_t nV,
                           int L, const int32_t *sets, uint8_t *isBoundary,
                           uint8_t *isRigidVert, int32_t *vertexBody) {
  std::vector<uint8_t> ofElastic(nV);
  for (int l = 0; l < L; ++l) {
    uint8_t *rv = isRigidVert + (size_t)nV * l;
    int32_t *vb = vertexBody + (size_t)nV * l;
    std::fill(rv, rv + nV, 0);
    std::fill(vb, vb + nV, 0);
    std::fill(ofElastic.begin(), ofElastic.end(), 0);
    const int32_t *col = sets + (size_t)nT * l;
    for (int64_t e = 0; e < nT; ++e) {
      int32_t b = col[e];
      for (int k = 0; k < vpe; ++k) {
        int32_t v = T0[(size_t)e * vpe + k];
        if (b) { rv[v] = 1; vb[v] = b; } else ofElastic[v] = 1;
      }
    }
    for (int64_t v = 0; v < nV; ++v) isBoundary[(size_t)nV * l + v] = rv[v] && ofElastic[v];
  }
  return MLX_OK;
}

// Stable ascending order of keys (Matlab sort(...,"ascend")), NaN last.
void stable_order(const double *keys, int64_t n, int32_t *order0) {
  std::iota(order0, order0 + n, 0);
  std::stable_sort(order0, order0 + n, [&](int32_t a, int32_t b) {
    double x = keys[a], y = keys[b];
    if (std::isnan(x)) return false;
    if (std::isnan(y)) return true;
    return x < y;
  });
}

}  // namespace mlx

// ------------------------------------------------------------------------
// C ABI (host-only entry points)
// ------------------------------------------------------------------------

static int to_zero_based(const int32_t *t, int64_t nT, int vpe, int64_t nV,
                         std::vector<int32_t> &T0) {
  T0.resize((size_t)nT * vpe);
  for (int64_t e = 0; e < nT; ++e)
    for (int k = 0; k < vpe; ++k) {
      int32_t v = t[e + nT * k] - 1;
      if (v < 0 || v >= nV) {
        mlx::set_error("t(%lld,%d) = %d is outside 1..%lld", (long long)e + 1, k + 1,
                       (int)t[e + nT * k], (long long)nV);
        return MLX_ERR_ARG;
      }
      T0[(size_t)e * vpe + k] = v;
    }
  return MLX_OK;
}

extern "C" {

const char *mlx_last_error(void) { return mlx::last_error(); }
int mlx_version(void) { return 100; }

int mlx_naive_graph_coloring(const int32_t *t, int64_t nT, int32_t vpe,
                             int64_t nV, int32_t *elementColor,
                             int32_t *numColors) {
  if (!t || !elementColor || !numColors || nT < 0 || vpe < 2 || vpe > 4) {
    mlx::set_error("mlx_naive_graph_coloring: bad arguments");
    return MLX_ERR_ARG;
  }
  std::vector<int32_t> T0;
  MLX_TRY(to_zero_based(t, nT, vpe, nV, T0));
  return mlx::graph_coloring(T0.data(), nT, vpe, nV, elementColor, numColors);
}

int mlx_mesh3d_precompute(const double *p, int64_t nV, const int32_t *t,
                          int64_t nT, const double *rho, const double *mu,
                          const double *lambda, const double *alpha1,
                          double *DmInv, double *elV, double *mass,
                          double *alpha, double *beta) {
  if (!p || !t || !rho || !mu || !lambda || !alpha1 || !DmInv || !elV || !mass ||
      !alpha || !beta || nV <= 0 || nT <= 0) {
    mlx::set_error("mlx_mesh3d_precompute: bad arguments");
    return MLX_ERR_ARG;
  }
  std::vector<int32_t> T0;
  MLX_TRY(to_zero_based(t, nT, 4, nV, T0));
  return mlx::precompute3d(p, nV, T0.data(), nT, rho, mu, lambda, alpha1, DmInv,
                           elV, mass, alpha, beta);
}

int mlx_mesh2d_precompute(const double *p, int64_t nV, const int32_t *t,
                          int64_t nT, const double *rho, const double *mu,
                          const double *lambda, const double *alpha1,
                          double *DmInv, double *elA, double *mass,
                          double *alpha, double *beta) {
  if (!p || !t || !rho || !mu || !lambda || !alpha1 || !DmInv || !elA || !mass ||
      !alpha || !beta || nV <= 0 || nT <= 0) {
    mlx::set_error("mlx_mesh2d_precompute: bad arguments");
    return MLX_ERR_ARG;
  }
  std::vector<int32_t> T0;
  MLX_TRY(to_zero_based(t, nT, 3, nV, T0));
  return mlx::precompute2d(p, nV, T0.data(), nT, rho, mu, lambda, alpha1, DmInv,
                           elA, mass, alpha, beta);
}

int mlx_union_find_by_vertice(const int32_t *strainInds, int64_t nT,
                              const double *sortedLayers, int32_t L,
                              const uint8_t *forcedElastic, const int32_t *t,
                              int32_t vpe, int64_t nV,
                              int32_t *rigidBodySetsPerLayer,
                              int32_t *numRigids) {
  if (!strainInds || !sortedLayers || !forcedElastic || !t || !rigidBodySetsPerLayer ||
      !numRigids || nT < 0 || L < 1 || vpe < 2 || vpe > 4) {
    mlx::set_error("mlx_union_find_by_vertice: bad arguments");
    return MLX_ERR_ARG;
  }
  for (int l = 1; l < L; ++l)
    if (!(sortedLayers[l] > sortedLayers[l - 1])) {
      mlx::set_error("sortedLayers must be strictly ascending");
      return MLX_ERR_ARG;
    }
  std::vector<int32_t> T0, order((size_t)nT);
  MLX_TRY(to_zero_based(t, nT, vpe, nV, T0));
  std::vector<uint8_t> seen((size_t)nT, 0);
  for (int64_t i = 0; i < nT; ++i) {
    int32_t e = strainInds[i] - 1;
    if (e < 0 || e >= nT || seen[e]) {
      mlx::set_error("strainInds is not a permutation of 1..nT");
      return MLX_ERR_ARG;
    }
    seen[e] = 1;
    order[i] = e;
  }
  return mlx::union_find_by_vertice(order.data(), nT, sortedLayers, L, forcedElastic,
                                    T0.data(), vpe, nV, rigidBodySetsPerLayer, numRigids);
}

int mlx_find_boundary_vertices(const int32_t *t, int64_t nT, int32_t vpe,
                               int64_t nV, int32_t L,
                               const int32_t *rigidBodySetsPerLayer,
                               uint8_t *isBoundaryVertex, uint8_t *isRigidVertex,
                               int32_t *vertexRigidBody) {
  if (!t || !rigidBodySetsPerLayer || !isBoundaryVertex || !isRigidVertex ||
      !vertexRigidBody || L < 1 || vpe < 2 || vpe > 4) {
    mlx::set_error("mlx_find_boundary_vertices: bad arguments");
    return MLX_ERR_ARG;
  }
  std::vector<int32_t> T0;
  MLX_TRY(to_zero_based(t, nT, vpe, nV, T0));
  return mlx::find_boundary_vertices(T0.data(), nT, vpe, nV, L, rigidBodySetsPerLayer,
                                     isBoundaryVertex, isRigidVertex, vertexRigidBody);
}

}  // extern "C"
