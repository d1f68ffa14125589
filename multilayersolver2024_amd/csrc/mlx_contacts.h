// Contact sample generation on the device beyond the ground plane (3D):
// the platform box of PlatformContactDetector.findContacts
// (3d/contact/PlatformContactDetector.m:65-120) and signed-distance samples
// of a collider given as values on a regular grid (the role of
// MeshSCD3D's signed_distance query, 3d/contact/MeshSCD3D.m:139), next to
// the plane of PlaneContactFinder3D.  Every finder flags the vertices it
// claims at the start-of-substep positions; the flags of all finders, one
// row of nV per finder, are compacted in one pass, so the contacts come out
// finder after finder and in ascending vertex order within a finder, the
// order in which the reference concatenates its finders' contact lists.
// The arithmetic is not contracted into fused multiply-adds, so active
// sets, normals and collider points are bit-identical to the same
// expressions evaluated on the host.  Included once, by mlx_solver.hip.
#pragma once

namespace mlx {

constexpr int kMaxFinders = 8;

// One finder's parameters on the device (MLX_CONTACT_* kinds).
struct FinderDev {
  int kind;
  double compliance;
  double n[3], D;                    // plane: unit normal, n . position
  double c[3], ax[9], he[3];         // box: center, unit axes (rows), half extents
  const double *sdf;                 // sdf: values, x fastest
  int dims[3];
  double origin[3], spacing;
};

// Platform box: the vertex is inside when its offset from the center
// projects within the half extents on every axis (:71-82); phi is minus the
// smallest distance to a face (:89-90, the first axis on ties) and the
// normal that axis, signed by the side of the center the vertex lies on.
__device__ __forceinline__ bool box_contact(const FinderDev &f, const double (&x)[3], double *phi,
                                            double (&nrm)[3]) {
#pragma clang fp contract(off)
  const double d[3] = {x[0] - f.c[0], x[1] - f.c[1], x[2] - f.c[2]};
  double dot[3];
  for (int a = 0; a < 3; ++a)
    dot[a] = d[0] * f.ax[3 * a] + d[1] * f.ax[3 * a + 1] + d[2] * f.ax[3 * a + 2];
  if (!(fabs(dot[0]) <= f.he[0] && fabs(dot[1]) <= f.he[1] && fabs(dot[2]) <= f.he[2])) return false;
  const double dist[3] = {f.he[0] - fabs(dot[0]), f.he[1] - fabs(dot[1]), f.he[2] - fabs(dot[2])};
  int k = 0;
  if (dist[1] < dist[k]) k = 1;
  if (dist[2] < dist[k]) k = 2;
  *phi = -dist[k];
  const double s = dot[k] > 0 ? 1.0 : dot[k] < 0 ? -1.0 : 0.0;
  for (int a = 0; a < 3; ++a) nrm[a] = s * f.ax[3 * k + a];
  return true;
}

// Signed distance from grid samples: trilinear interpolation in the cell
// that holds the vertex (none outside the grid) and the gradient of that
// interpolant; a contact where the value is negative and the gradient
// nonzero, its normal the unit gradient.
__device__ __forceinline__ bool sdf_contact(const FinderDev &f, const double (&x)[3], double *phi,
                                            double (&nrm)[3]) {
#pragma clang fp contract(off)
  int i0[3];
  double t[3];
  for (int a = 0; a < 3; ++a) {
    const double g = (x[a] - f.origin[a]) / f.spacing;
    if (!(g >= 0.0) || !(g <= (double)(f.dims[a] - 1))) return false;
    int i = (int)floor(g);
    if (i > f.dims[a] - 2) i = f.dims[a] - 2;
    i0[a] = i;
    t[a] = g - (double)i;
  }
  const int64_t sx = 1, sy = f.dims[0], sz = (int64_t)f.dims[0] * f.dims[1];
  const double *v = f.sdf + i0[0] * sx + i0[1] * sy + i0[2] * sz;
  const double c000 = v[0], c100 = v[sx], c010 = v[sy], c110 = v[sx + sy], c001 = v[sz],
               c101 = v[sx + sz], c011 = v[sy + sz], c111 = v[sx + sy + sz];
  const double u = t[0], w = t[1], r = t[2];
  // along x, then y, then z
  const double c00 = c000 + u * (c100 - c000), c10 = c010 + u * (c110 - c010);
  const double c01 = c001 + u * (c101 - c001), c11 = c011 + u * (c111 - c011);
  const double c0 = c00 + w * (c10 - c00), c1 = c01 + w * (c11 - c01);
  const double val = c0 + r * (c1 - c0);
  if (!(val < 0.0)) return false;
  // derivatives of the interpolant
  const double dx00 = c100 - c000, dx10 = c110 - c010, dx01 = c101 - c001, dx11 = c111 - c011;
  const double dx0 = dx00 + w * (dx10 - dx00), dx1 = dx01 + w * (dx11 - dx01);
  const double gx = (dx0 + r * (dx1 - dx0)) / f.spacing;
  const double dy0 = c10 - c00, dy1 = c11 - c01;
  const double gy = (dy0 + r * (dy1 - dy0)) / f.spacing;
  const double gz = (c1 - c0) / f.spacing;
  const double gn = sqrt(gx * gx + gy * gy + gz * gz);
  if (!(gn > 0.0)) return false;
  *phi = val;
  nrm[0] = gx / gn;
  nrm[1] = gy / gn;
  nrm[2] = gz / gn;
  return true;
}

__device__ __forceinline__ bool finder_contact(const FinderDev &f, const double (&x)[3],
                                               double *phi, double (&nrm)[3]) {
#pragma clang fp contract(off)
  if (f.kind == MLX_CONTACT_PLANE) {
    double dist = x[0] * f.n[0] + x[1] * f.n[1];
    dist = dist + x[2] * f.n[2];
    dist = dist - f.D;
    if (!(dist <= 0)) return false;  // PlaneContactFinder3D.m:40-43
    *phi = dist;
    for (int a = 0; a < 3; ++a) nrm[a] = f.n[a];
    return true;
  }
  if (f.kind == MLX_CONTACT_BOX) return box_contact(f, x, phi, nrm);
  return sdf_contact(f, x, phi, nrm);
}

// flag[f nV + v]: vertex v claimed by finder f.
template <typename T>
__global__ void k_finder_flags(int64_t nV, const vec4_t<T> *x, const FinderDev *finders, int nf,
                               int32_t *flag) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nV * nf) return;
  const int f = (int)(i / nV);
  const int64_t v = i - f * nV;
  const vec4_t<T> p = x[v];
  const double xs[3] = {(double)p.x, (double)p.y, (double)p.z};
  double phi, nrm[3];
  flag[i] = finder_contact(finders[f], xs, &phi, nrm) ? 1 : 0;
}

// Contact i (i below the count and the arrays' capacity) from its code
// f nV + v: vertex, normal, collider point x - n phi (contactInfo3D.m:34-38)
// and the finder's compliance.
template <typename T>
__global__ void k_finder_gather(const int32_t *count, int cap, int64_t nV, const int32_t *code,
                                const vec4_t<T> *x, const FinderDev *finders, int32_t *vert,
                                double *normal, double *point, double *compl_, double *lambda) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= count[0] || i >= cap) return;
  const int f = (int)(code[i] / nV);
  const int v = (int)(code[i] - f * nV);
  const vec4_t<T> p = x[v];
  const double xs[3] = {(double)p.x, (double)p.y, (double)p.z};
  double phi = 0, nrm[3] = {0, 0, 0};
  (void)finder_contact(finders[f], xs, &phi, nrm);
  vert[i] = v;
  {
#pragma clang fp contract(off)
    for (int d = 0; d < 3; ++d) {
      normal[3 * i + d] = nrm[d];
      point[3 * i + d] = xs[d] - nrm[d] * phi;
    }
  }
  compl_[i] = finders[f].compliance;
  lambda[i] = 0.0;
}

}  // namespace mlx
