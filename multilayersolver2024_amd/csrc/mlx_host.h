// Host-side helpers of libmlxpbd shared by the C ABI and the GPU
// orchestration.  All index arrays here are 0-based, elements row-major.
#pragma once
#include <cstdint>

#ifndef MLX_TRY
#define MLX_TRY(expr)            \
  do {                           \
    int r_ = (expr);             \
    if (r_ != 0) return r_;      \
  } while (0)
#endif

namespace mlx {
void set_error(const char *fmt, ...);
const char *last_error();
int graph_coloring(const int32_t *T0, int64_t nT, int vpe, int64_t nV,
                   int32_t *color, int32_t *numColors);
int precompute3d(const double *p, int64_t nV, const int32_t *T0, int64_t nT,
                 const double *rho, const double *mu, const double *lam,
                 const double *alpha1, double *DmInv, double *vol, double *mass,
                 double *alpha, double *beta);
int precompute2d(const double *p, int64_t nV, const int32_t *T0, int64_t nT,
                 const double *rho, const double *mu, const double *lam,
                 const double *alpha1, double *DmInv, double *area, double *mass,
                 double *alpha, double *beta);
int union_find_by_vertice(const int32_t *order0, int64_t nT,
                          const double *sortedLayers, int L,
                          const uint8_t *forcedElastic, const int32_t *T0,
                          int vpe, int64_t nV, int32_t *sets, int32_t *numRigids);
int find_boundary_vertices(const int32_t *T0, int64_t nT, int vpe, int64_t nV,
                           int L, const int32_t *sets, uint8_t *isBoundary,
                           uint8_t *isRigidVert, int32_t *vertexBody);
void stable_order(const double *keys, int64_t n, int32_t *order0);
}  // namespace mlx
