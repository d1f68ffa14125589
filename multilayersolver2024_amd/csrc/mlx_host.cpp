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

// Inverse of an n x n column-major matrix (n <= 6) as Matlab's inv /
// pageinv forms it: an LU factorisation with partial pivoting (LAPACK
// getrf), then one forward/back substitution per column of the identity.
// Returns false if a pivot is exactly zero.  The strain-rate keys of the
// layer maps (B*p at rest) are rounding noise, so the inverse's last bits
// decide the layer sets: every implementation that must reproduce them
// uses this elimination order.
static bool inv_lu(int n, const double *A, double *R) {
#pragma clang fp contract(off)
  double lu[36];
  int perm[6];
  bool ok = true;
  std::memcpy(lu, A, sizeof(double) * n * n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(lu[i + n * k]) > std::fabs(lu[piv + n * k])) piv = i;
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(lu[k + n * j], lu[piv + n * j]);
      std::swap(perm[k], perm[piv]);
    }
    const double d = lu[k + n * k];
    if (d == 0.0) ok = false;
    for (int i = k + 1; i < n; ++i) {
      const double l = lu[i + n * k] / d;
      lu[i + n * k] = l;
      for (int j = k + 1; j < n; ++j) lu[i + n * j] = lu[i + n * j] - l * lu[k + n * j];
    }
  }
  for (int c = 0; c < n; ++c) {
    // right-hand side: column c of the identity, rows permuted like the LU
    double y[6], x[6];
    for (int i = 0; i < n; ++i) y[i] = perm[i] == c ? 1.0 : 0.0;
    for (int i = 1; i < n; ++i)
      for (int j = 0; j < i; ++j) y[i] = y[i] - lu[i + n * j] * y[j];
    for (int i = n - 1; i >= 0; --i) {
      double t = y[i];
      for (int j = i + 1; j < n; ++j) t = t - lu[i + n * j] * x[j];
      x[i] = t / lu[i + n * i];
    }
    for (int i = 0; i < n; ++i) R[i + n * c] = x[i];
  }
  return ok;
}

// det of a 3x3 column-major matrix, cofactor expansion along the first row.
static double det3(const double *A) {
#pragma clang fp contract(off)
  return A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) +
         A[6] * (A[1] * A[5] - A[4] * A[2]);
}

// The isotropic Voigt stiffness K of updateMaterials (3D: normal block
// diag lambda + 2 (2 mu), off-diagonal lambda, shear 2 mu; 2D: diag
// lambda + 2 mu, off-diagonal lambda, shear 2 mu), column-major n x n.
static void voigt_k(int n, double diag, double off, double shear, double *K) {
  std::memset(K, 0, sizeof(double) * n * n);
  const int nn = (n == 6) ? 3 : 2;
  for (int i = 0; i < nn; ++i)
    for (int j = 0; j < nn; ++j) K[i + n * j] = (i == j) ? diag : off;
  for (int i = nn; i < n; ++i) K[i + n * i] = shear;
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
    const double det = det3(Dm);
    if (!(det != 0.0) || !std::isfinite(det) || !inv_lu(3, Dm, DmInv + 9 * e)) {
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
    double K[36];
    voigt_k(6, dv, l, dm, K);
    if (!inv_lu(6, K, alpha + 36 * e)) {
      set_error("element %lld: singular stiffness", (long long)e);
      return MLX_ERR_ARG;
    }
    for (int k = 0; k < 36; ++k) beta[36 * e + k] = K[k] * alpha1[e];
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
    const double Dm[4] = {a, c, b, d};
    inv_lu(2, Dm, DmInv + 4 * e);
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
    double K[9];
    voigt_k(3, dv, l, s, K);
    if (!inv_lu(3, K, alpha + 9 * e)) {
      set_error("triangle %lld: singular stiffness", (long long)e);
      return MLX_ERR_ARG;
    }
    for (int k = 0; k < 9; ++k) beta[9 * e + k] = K[k] * alpha1[e];
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
