// Layer maps on the device: buildLayers (xpbdLayer3D.m:601-619; 2D
// xpbdLayer2D.m:740-790, ordering type 0) and findBoundaryVertices
// (xpbdLayer2D.m:551-569).
//
// 1. Strain-rate keys (EDiff per element, already on the device) are sorted
//    ascending and stably, NaN last, like Matlab's sort(...,"ascend"): the
//    keys are mapped to order-preserving 64-bit integers and radix-sorted
//    with the element indices as values (LSD radix sort is stable).
// 2. For each sorted layer l the rigid candidates are the first
//    floor(layers(l)/100 * nT) elements of that order, minus the pinned
//    (forced elastic) ones.  mexUnionFindByVertice merges each inserted
//    element with the elements sharing its vertices and numbers the bodies
//    by ascending smallest element (copyLayer).  Here: every element is
//    hooked to the smallest inserted element of each of its vertices with a
//    lock-free union-find that always hooks the larger root under the
//    smaller one, so each tree's root is its component's smallest element;
//    an exclusive scan of the root flags numbers the bodies in that order.
// 3. Vertex maps: rigid vertices carry their body id, boundary vertices are
//    rigid vertices that also belong to an elastic element.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <climits>
#include <cstring>
#include <cmath>
#include <vector>

#include "mlx_internal.h"

namespace mlx {
namespace {

constexpr int kMapBlock = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kMapBlock - 1) / kMapBlock); }

// Order-preserving key: -0 and +0 compare equal in Matlab's sort and every
// NaN goes last in input order, so -0 becomes +0 and any NaN one quiet NaN
// before the usual sign-flip transform of IEEE doubles.
__global__ void k_sort_keys(int64_t n, const double *keys, uint64_t *out, int32_t *idx) {
  const int64_t i = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (i >= n) return;
  double k = keys[i];
  if (k != k) k = __longlong_as_double(0x7ff8000000000000LL);
  if (k == 0.0) k = 0.0;
  uint64_t b = (uint64_t)__double_as_longlong(k);
  b = (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
  out[i] = b;
  idx[i] = (int32_t)i;
}

__global__ void k_rank(int64_t n, const int32_t *order, int32_t *rank) {
  const int64_t i = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (i >= n) return;
  rank[order[i]] = (int32_t)i;
}

// parent[e] = e for inserted elements, -1 otherwise; vmin cleared.
__global__ void k_cc_init(int64_t nT, int64_t nV, const int32_t *rank, const uint8_t *forced,
                          int32_t nInsert, int32_t *parent, int32_t *vmin) {
  const int64_t i = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (i < nT) parent[i] = (rank[i] < nInsert && !forced[i]) ? (int32_t)i : -1;
  if (i < nV) vmin[i] = INT_MAX;
}

// vmin[v] = smallest inserted element containing v.
__global__ void k_cc_vmin(int64_t nT, int vpe, const int32_t *T, const int32_t *parent,
                          int32_t *vmin) {
  const int64_t e = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (e >= nT || parent[e] < 0) return;
  for (int k = 0; k < vpe; ++k) atomicMin(&vmin[T[e * vpe + k]], (int32_t)e);
}

__device__ __forceinline__ int32_t cc_load(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int32_t cc_find(const int32_t *parent, int32_t x) {
  int32_t p = cc_load(parent + x);
  while (p != x) {
    x = p;
    p = cc_load(parent + x);
  }
  return x;
}

// Root of x with path splitting: every non-root met on the way is pointed
// at its grandparent.  Only roots are ever hooked (k_cc_union's CAS), a
// non-root never becomes a root again, and a grandparent is an ancestor,
// so concurrent finds and unions keep a forest with the same components.
__device__ __forceinline__ int32_t cc_find_split(int32_t *parent, int32_t x) {
  int32_t cur = cc_load(parent + x);
  if (cur == x) return x;
  int32_t prev = x, next;
  while (cur != (next = cc_load(parent + cur))) {
    __hip_atomic_store(parent + prev, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    prev = cur;
    cur = next;
  }
  return cur;
}

// Join the trees of a and b.  Roots only ever get a smaller parent, so
// every retry either finds the two in one tree or hooks one root; the loop
// ends.
__device__ __forceinline__ void cc_union_pair(int32_t *parent, int32_t a, int32_t b) {
  while (true) {
    a = cc_find_split(parent, a);
    b = cc_find_split(parent, b);
    if (a == b) return;
    if (a < b) {
      const int32_t t = a;
      a = b;
      b = t;
    }
    // hook root a (larger) under b (smaller)
    int32_t expected = a;
    if (__hip_atomic_compare_exchange_strong(parent + a, &expected, b, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return;
    a = expected;
  }
}

// Union of e with vmin[v] for each of its vertices (every inserted element).
__global__ void k_cc_union(int64_t nT, int vpe, const int32_t *T, const int32_t *vmin,
                           int32_t *parent) {
  const int64_t e = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (e >= nT || cc_load(parent + e) < 0) return;
  for (int k = 0; k < vpe; ++k) cc_union_pair(parent, (int32_t)e, vmin[T[e * vpe + k]]);
}

// Incremental layers (sorted percentages ascending, so each layer's
// inserted set contains the previous one's): the elements with rank in
// [prev, n) (and not forced elastic) become singleton roots; the earlier
// ones keep their forest.
__global__ void k_cc_add(int64_t nT, const int32_t *rank, const uint8_t *forced, int32_t prev,
                         int32_t n, int32_t *parent) {
  const int64_t e = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (e >= nT) return;
  const int32_t r = rank[e];
  if (r >= prev && r < n && !forced[e]) parent[e] = (int32_t)e;
}

// vmin over the newly inserted elements (vmin already holds the minimum
// over the earlier ones).
__global__ void k_cc_vmin_new(int64_t nT, int vpe, const int32_t *T, const int32_t *rank,
                              const uint8_t *forced, int32_t prev, int32_t n, int32_t *vmin) {
  const int64_t e = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (e >= nT) return;
  const int32_t r = rank[e];
  if (r < prev || r >= n || forced[e]) return;
  for (int k = 0; k < vpe; ++k) atomicMin(&vmin[T[e * vpe + k]], (int32_t)e);
}

// Each newly inserted element starts under the smallest inserted element
// sharing one of its vertices (vmin, which already covers the new ones): a
// smaller id of the same component, so roots stay component minima.  The
// unions that follow then mostly find both sides in one tree instead of
// racing to hook the same few roots.
__global__ void k_cc_hook_new(int64_t nT, int vpe, const int32_t *T, const int32_t *rank,
                              const uint8_t *forced, int32_t prev, int32_t n, const int32_t *vmin,
                              int32_t *parent) {
  const int64_t e = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (e >= nT) return;
  const int32_t r = rank[e];
  if (r < prev || r >= n || forced[e]) return;
  int32_t lo = (int32_t)e;
  for (int k = 0; k < vpe; ++k) lo = min(lo, vmin[T[e * vpe + k]]);
  parent[e] = lo;
}

// One round of pointer jumping over the newly inserted elements: parent
// becomes the grandparent.  Concurrent rounds of other threads only move
// pointers to other ancestors, so every pointer stays in its tree.
__global__ void k_cc_jump_new(int64_t nT, const int32_t *rank, const uint8_t *forced, int32_t prev,
                              int32_t n, int32_t *parent) {
  const int64_t e = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (e >= nT) return;
  const int32_t r = rank[e];
  if (r < prev || r >= n || forced[e]) return;
  const int32_t p = cc_load(parent + e);
  const int32_t pp = cc_load(parent + p);
  if (pp != p) __hip_atomic_store(parent + e, pp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Union of a newly inserted element e with, for each of its vertices v,
// the smallest earlier element containing v (vminOld: every earlier element
// containing v is already in that one's tree) and the smallest element of
// the layer containing v (vmin: joins the new elements sharing v).  With
// the earlier layer's forest this gives the components of the whole
// inserted set: two earlier elements sharing a vertex were joined before.
// The roots of e and of all its partners are walked together, one level of
// every walk per round of loads, so a thread waits for the deepest walk
// instead of the sum of them; only partners found in another tree go
// through cc_union_pair.  A root found equal to e's is in e's component
// (trees only ever merge), and a root found different is an ancestor of
// its element, so cc_union_pair may start from it.
__global__ void k_cc_union_new(int64_t nT, int vpe, const int32_t *T, const int32_t *rank,
                               const uint8_t *forced, int32_t prev, int32_t n,
                               const int32_t *vminOld, const int32_t *vmin, int32_t *parent) {
  constexpr int kMaxWalk = 1 + 2 * 4;  // e, then (vminOld, vmin) per vertex
  const int64_t e = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (e >= nT) return;
  const int32_t r = rank[e];
  if (r < prev || r >= n || forced[e]) return;
  // fixed slots (the array stays in registers); unused slots walk e again
  int32_t cur[kMaxWalk];
#pragma unroll
  for (int j = 0; j < kMaxWalk; ++j) cur[j] = (int32_t)e;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < vpe) {
      const int32_t v = T[e * vpe + k];
      const int32_t o = vminOld[v];
      if (o != INT_MAX) cur[1 + 2 * k] = o;
      cur[2 + 2 * k] = vmin[v];
    }
  unsigned live = (1u << kMaxWalk) - 1;
  while (live) {
    int32_t nxt[kMaxWalk];
#pragma unroll
    for (int j = 0; j < kMaxWalk; ++j)
      if (live >> j & 1) nxt[j] = cc_load(parent + cur[j]);
#pragma unroll
    for (int j = 0; j < kMaxWalk; ++j)
      if (live >> j & 1) {
        if (nxt[j] == cur[j]) live &= ~(1u << j);
        else cur[j] = nxt[j];
      }
  }
#pragma unroll
  for (int j = 1; j < kMaxWalk; ++j)
    if (cur[j] != cur[0]) cc_union_pair(parent, cur[0], cur[j]);
}

// Elastic flags of every sorted layer in slot order (flags pre-zeroed, so
// padding slots stay 0) and elastic counts per (layer, color).
__global__ void k_slot_flags(int64_t nT, int ns, int64_t ld, const int32_t *sets,
                             const int32_t *slotOf, uint8_t *flags) {
  const int64_t e = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (e >= nT) return;
  const int64_t s = slotOf[e];
  for (int l = 0; l < ns; ++l) flags[l * ld + s] = sets[(int64_t)l * nT + e] == 0 ? 1 : 0;
}

// Elastic elements per (layer, color): slots are color-major, so one
// workgroup per (layer, color) sums the flags of that color's slot range.
__global__ void k_color_counts(int nc, int64_t ld, const int64_t *colorTab, const uint8_t *flags,
                               int32_t *counts) {
  __shared__ int32_t part[kMapBlock];
  const int l = blockIdx.x / nc, c = blockIdx.x % nc;
  const int64_t a = colorTab[c], e = a + colorTab[nc + c];
  int32_t acc = 0;
  for (int64_t s = a + threadIdx.x; s < e; s += kMapBlock) acc += flags[l * ld + s];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kMapBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[blockIdx.x] = part[0];
}

// Body count of a layer: last scan entry + last root flag.
__global__ void k_cc_count(int64_t nT, const int32_t *scan, const int32_t *isRoot,
                           int32_t *count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *count = scan[nT - 1] + isRoot[nT - 1];
}

// After all unions: each inserted element points at its root, and roots
// are flagged for the numbering scan.
__global__ void k_cc_flatten(int64_t nT, int32_t *parent, int32_t *isRoot) {
  const int64_t e = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (e >= nT) return;
  const int32_t p = parent[e];
  if (p < 0) {
    isRoot[e] = 0;
    return;
  }
  const int32_t r = cc_find_split(parent, (int32_t)e);
  if (r != e) __hip_atomic_store(parent + e, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  isRoot[e] = r == e ? 1 : 0;
}

__global__ void k_cc_label(int64_t nT, const int32_t *parent, const int32_t *scan,
                           int32_t *sets) {
  const int64_t e = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (e >= nT) return;
  const int32_t p = parent[e];
  if (p < 0) {
    sets[e] = 0;
    return;
  }
  sets[e] = scan[cc_find(parent, (int32_t)e)] + 1;
}

__global__ void k_vmaps_clear(int64_t nV, int32_t *vb, uint8_t *rigid, uint8_t *ofElastic) {
  const int64_t v = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (v >= nV) return;
  vb[v] = 0;
  rigid[v] = 0;
  ofElastic[v] = 0;
}

// Every rigid element writes its body id to its vertices (all rigid
// elements at a vertex share one body, so the writes agree).
__global__ void k_vmaps(int64_t nT, int vpe, const int32_t *T, const int32_t *sets, int32_t *vb,
                        uint8_t *rigid, uint8_t *ofElastic) {
  const int64_t e = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (e >= nT) return;
  const int32_t b = sets[e];
  for (int k = 0; k < vpe; ++k) {
    const int32_t v = T[e * vpe + k];
    if (b != 0) {
      rigid[v] = 1;
      vb[v] = b;
    } else {
      ofElastic[v] = 1;
    }
  }
}

__global__ void k_boundary_flags(int64_t nV, const uint8_t *rigid, const uint8_t *ofElastic,
                                 uint8_t *boundary) {
  const int64_t v = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (v >= nV) return;
  boundary[v] = rigid[v] && ofElastic[v];
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

int layer_maps_dev(mlx_mesh &m, const double *d_keys, const double *sorted, int ns,
                   HostMaps &maps, std::vector<int32_t> *setsOut, bool deviceFlags) {
  const int64_t nT = m.nT, nV = m.nV;
  const int vpe = m.vpe;
  hipStream_t st = m.stream;
  maps.nSorted = ns;
  maps.numRigids.assign(ns, 0);
  maps.deviceFlags = deviceFlags;
  if (!deviceFlags) maps.elastic.assign((size_t)ns * nT, 1);
  maps.vbody.assign((size_t)ns * nV, 0);
  maps.boundary.assign((size_t)ns * nV, 0);
  if (nT == 0) return MLX_OK;
  if (vpe > 4) {
    set_error("layer maps: %d vertices per element (tetrahedra and triangles only)", vpe);
    return MLX_ERR_ARG;
  }
  if (nT > INT_MAX - 1) {
    set_error("layer maps: more than 2^31 elements");
    return MLX_ERR_UNSUPPORTED;
  }
  if (m.d_forced.bytes < (size_t)nT) {
    MLX_TRY(m.d_forced.reserve((size_t)nT));
    MLX_CHECK_HIP(hipMemcpyAsync(m.d_forced.ptr, m.forcedElastic.data(), (size_t)nT,
                                 hipMemcpyHostToDevice, st));
  }
  // scratch layout
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  const size_t oKeyA = take(8 * nT), oKeyB = take(8 * nT), oIdxA = take(4 * nT),
               oOrder = take(4 * nT), oRank = take(4 * nT), oParent = take(4 * nT),
               oVmin = take(4 * nV), oRoot = take(4 * nT), oScan = take(4 * nT),
               oSets = take(4 * nT * (size_t)ns), oVb = take(4 * nV * (size_t)ns),
               oRig = take(nV), oOfEl = take(nV), oBnd = take(nV * (size_t)ns),
               oVminOld = take(4 * nV), oCount = take(4 * (size_t)ns),
               oColorCnt = take(4 * (size_t)ns * (size_t)std::max(m.numColors, 1));
  MLX_TRY(m.d_maps.reserve(off));
  char *base = m.d_maps.as<char>();
  auto at32 = [&](size_t o) { return reinterpret_cast<int32_t *>(base + o); };
  uint64_t *keyA = reinterpret_cast<uint64_t *>(base + oKeyA);
  uint64_t *keyB = reinterpret_cast<uint64_t *>(base + oKeyB);
  int32_t *idxA = at32(oIdxA), *order = at32(oOrder), *rank = at32(oRank),
          *parent = at32(oParent), *vmin = at32(oVmin), *isRoot = at32(oRoot),
          *scan = at32(oScan), *sets = at32(oSets), *vb = at32(oVb);
  uint8_t *rig = reinterpret_cast<uint8_t *>(base + oRig);
  uint8_t *ofEl = reinterpret_cast<uint8_t *>(base + oOfEl);
  uint8_t *bnd = reinterpret_cast<uint8_t *>(base + oBnd);
  int32_t *vminOld = at32(oVminOld), *count = at32(oCount);
  const int32_t *T = m.d_T.as<int32_t>();
  const uint8_t *forced = m.d_forced.as<uint8_t>();

  // 1. stable ascending sort of the keys
  hipLaunchKernelGGL(k_sort_keys, dim3(grid_for(nT)), dim3(kMapBlock), 0, st, nT, d_keys, keyA,
                     idxA);
  size_t tmpSort = 0, tmpScan = 0;
  MLX_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, tmpSort, keyA, keyB, idxA, order,
                                          (unsigned int)nT, 0, 64, st));
  MLX_CHECK_HIP(rocprim::exclusive_scan(nullptr, tmpScan, isRoot, scan, 0, (size_t)nT,
                                        rocprim::plus<int32_t>(), st));
  MLX_TRY(m.d_scan.reserve(std::max<size_t>(std::max(tmpSort, tmpScan), 256)));
  MLX_CHECK_HIP(rocprim::radix_sort_pairs(m.d_scan.ptr, tmpSort, keyA, keyB, idxA, order,
                                          (unsigned int)nT, 0, 64, st));
  hipLaunchKernelGGL(k_rank, dim3(grid_for(nT)), dim3(kMapBlock), 0, st, nT, order, rank);

  // 2. bodies per sorted layer.  The percentages are ascending, so each
  // layer's inserted set contains the previous one's: its forest is kept
  // and only the newly inserted elements are added and joined.
  const int64_t nMax = std::max(nT, nV);
  hipLaunchKernelGGL(k_cc_init, dim3(grid_for(nMax)), dim3(kMapBlock), 0, st, nT, nV, rank,
                     forced, 0, parent, vmin);
  int32_t prevInsert = 0;
  for (int l = 0; l < ns; ++l) {
    const double end = std::floor((sorted[l] / 100.0) * (double)nT);
    const int32_t nInsert = end <= 0 ? 0 : (end >= (double)nT ? (int32_t)nT : (int32_t)end);
    int32_t *setsL = sets + (size_t)nT * l;
    if (nInsert < prevInsert) {
      // percentages given out of order: start this layer from scratch
      hipLaunchKernelGGL(k_cc_init, dim3(grid_for(nMax)), dim3(kMapBlock), 0, st, nT, nV, rank,
                         forced, 0, parent, vmin);
      prevInsert = 0;
    }
    if (nInsert > prevInsert) {
      hipLaunchKernelGGL(k_cc_add, dim3(grid_for(nT)), dim3(kMapBlock), 0, st, nT, rank, forced,
                         prevInsert, nInsert, parent);
      MLX_CHECK_HIP(hipMemcpyAsync(vminOld, vmin, 4 * (size_t)nV, hipMemcpyDeviceToDevice, st));
      hipLaunchKernelGGL(k_cc_vmin_new, dim3(grid_for(nT)), dim3(kMapBlock), 0, st, nT, vpe, T,
                         rank, forced, prevInsert, nInsert, vmin);
      hipLaunchKernelGGL(k_cc_hook_new, dim3(grid_for(nT)), dim3(kMapBlock), 0, st, nT, vpe, T,
                         rank, forced, prevInsert, nInsert, vmin, parent);
      for (int round = 0; round < 6; ++round)
        hipLaunchKernelGGL(k_cc_jump_new, dim3(grid_for(nT)), dim3(kMapBlock), 0, st, nT, rank,
                           forced, prevInsert, nInsert, parent);
      hipLaunchKernelGGL(k_cc_union_new, dim3(grid_for(nT)), dim3(kMapBlock), 0, st, nT, vpe, T,
                         rank, forced, prevInsert, nInsert, vminOld, vmin, parent);
      prevInsert = nInsert;
    }
    hipLaunchKernelGGL(k_cc_flatten, dim3(grid_for(nT)), dim3(kMapBlock), 0, st, nT, parent,
                       isRoot);
    size_t tb = tmpScan;
    MLX_CHECK_HIP(rocprim::exclusive_scan(m.d_scan.ptr, tb, isRoot, scan, 0, (size_t)nT,
                                          rocprim::plus<int32_t>(), st));
    hipLaunchKernelGGL(k_cc_label, dim3(grid_for(nT)), dim3(kMapBlock), 0, st, nT, parent, scan,
                       setsL);
    // 3. vertex maps of this layer
    int32_t *vbL = vb + (size_t)nV * l;
    uint8_t *bndL = bnd + (size_t)nV * l;
    hipLaunchKernelGGL(k_vmaps_clear, dim3(grid_for(nV)), dim3(kMapBlock), 0, st, nV, vbL, rig,
                       ofEl);
    hipLaunchKernelGGL(k_vmaps, dim3(grid_for(nT)), dim3(kMapBlock), 0, st, nT, vpe, T, setsL,
                       vbL, rig, ofEl);
    hipLaunchKernelGGL(k_boundary_flags, dim3(grid_for(nV)), dim3(kMapBlock), 0, st, nV, rig, ofEl,
                       bndL);
    // body count = sum of root flags = last scan entry + last flag
    hipLaunchKernelGGL(k_cc_count, dim3(1), dim3(64), 0, st, nT, scan, isRoot, count + l);
  }
  MLX_CHECK_HIP(hipGetLastError());
  // everything read back goes through pinned staging, so the copies queue
  // behind the kernels and one synchronization ends them all
  auto take32 = [&](size_t n, int32_t **out) -> int {
    char *p = nullptr;
    MLX_TRY(stage_take(m, 4 * std::max<size_t>(n, 1), &p));
    *out = reinterpret_cast<int32_t *>(p);
    return MLX_OK;
  };
  int32_t *hNum = nullptr;
  MLX_TRY(take32((size_t)ns, &hNum));
  MLX_CHECK_HIP(hipMemcpyAsync(hNum, count, 4 * (size_t)ns, hipMemcpyDeviceToHost, st));
  std::vector<int32_t> colorCnt;
  int32_t *hCC = nullptr;
  if (deviceFlags) {
    // 4. elastic flags in slot order straight into m.d_flags
    const int nc = m.numColors;
    if (m.d_colortab.bytes < 16 * (size_t)nc) {
      // slot range of each color: colorOff | colorCnt (int64), uploaded once
      std::vector<int64_t> ct(2 * (size_t)nc);
      for (int c = 0; c < nc; ++c) {
        ct[c] = m.colorOff[c];
        ct[nc + c] = m.colorCnt[c];
      }
      MLX_TRY(m.d_colortab.reserve(16 * (size_t)nc));
      MLX_TRY(stage_upload(m, m.d_colortab.ptr, ct.data(), 16 * (size_t)nc));
    }
    MLX_TRY(m.d_flags.reserve((size_t)ns * m.ld));
    int32_t *cc = at32(oColorCnt);
    MLX_CHECK_HIP(hipMemsetAsync(m.d_flags.ptr, 0, (size_t)ns * m.ld, st));
    hipLaunchKernelGGL(k_slot_flags, dim3(grid_for(nT)), dim3(kMapBlock), 0, st, nT, ns, m.ld,
                       sets, m.d_slotOf.as<int32_t>(), m.d_flags.as<uint8_t>());
    if (ns * nc > 0)
      hipLaunchKernelGGL(k_color_counts, dim3((unsigned)(ns * nc)), dim3(kMapBlock), 0, st, nc,
                         m.ld, m.d_colortab.as<int64_t>(), m.d_flags.as<uint8_t>(), cc);
    colorCnt.resize((size_t)ns * nc);
    MLX_TRY(take32(colorCnt.size(), &hCC));
    MLX_CHECK_HIP(hipMemcpyAsync(hCC, cc, 4 * colorCnt.size(), hipMemcpyDeviceToHost, st));
  }
  std::vector<int32_t> hs(deviceFlags && !setsOut ? 0 : (size_t)nT * ns);
  int32_t *hSets = nullptr, *hVb = nullptr;
  if (!hs.empty()) {
    MLX_TRY(take32(hs.size(), &hSets));
    MLX_CHECK_HIP(hipMemcpyAsync(hSets, sets, 4 * hs.size(), hipMemcpyDeviceToHost, st));
  }
  maps.devVbody = vb;
  MLX_TRY(take32(maps.vbody.size(), &hVb));
  MLX_CHECK_HIP(hipMemcpyAsync(hVb, vb, 4 * maps.vbody.size(), hipMemcpyDeviceToHost, st));
  char *hBnd = nullptr;
  MLX_TRY(stage_take(m, std::max<size_t>(maps.boundary.size(), 1), &hBnd));
  MLX_CHECK_HIP(hipMemcpyAsync(hBnd, bnd, maps.boundary.size(), hipMemcpyDeviceToHost, st));
  MLX_CHECK_HIP(hipStreamSynchronize(st));
  std::memcpy(maps.numRigids.data(), hNum, 4 * (size_t)ns);
  if (hCC) std::memcpy(colorCnt.data(), hCC, 4 * colorCnt.size());
  if (hSets) std::memcpy(hs.data(), hSets, 4 * hs.size());
  std::memcpy(maps.vbody.data(), hVb, 4 * maps.vbody.size());
  std::memcpy(maps.boundary.data(), hBnd, maps.boundary.size());
  if (deviceFlags) {
    const int nc = m.numColors;
    maps.perColor.assign(colorCnt.begin(), colorCnt.end());
    maps.nElastic.assign(ns, 0);
    for (int l = 0; l < ns; ++l)
      for (int c = 0; c < nc; ++c) maps.nElastic[l] += colorCnt[(size_t)l * nc + c];
  } else {
    for (size_t i = 0; i < hs.size(); ++i) maps.elastic[i] = hs[i] == 0;
  }
  if (setsOut) setsOut->swap(hs);
  return MLX_OK;
}

namespace {

// Ordered compaction: flagged vertex v goes to slot pos[v] (pos = exclusive
// scan of the flags), so the list ascends in vertex id as a host loop over
// the flags would build it; its multiplier starts at 0.  The last vertex
// writes the count.
__global__ void k_compact_flags(int64_t n, const int32_t *flag, const int32_t *pos,
                                int32_t *list, double *lambda, int32_t *count) {
  const int64_t v = (int64_t)blockIdx.x * kMapBlock + threadIdx.x;
  if (v >= n) return;
  if (flag[v]) {
    list[pos[v]] = (int32_t)v;
    lambda[pos[v]] = 0.0;
  }
  if (v == n - 1) count[0] = pos[v] + flag[v];
}

}  // namespace

int compact_flags_dev(mlx_mesh &m, const int32_t *flag, int64_t n, int32_t *pos, int32_t *list,
                      double *lambda, int32_t *count, hipStream_t st) {
  if (n <= 0) {
    MLX_CHECK_HIP(hipMemsetAsync(count, 0, sizeof(int32_t), st));
    return MLX_OK;
  }
  size_t tmp = 0;
  MLX_CHECK_HIP(rocprim::exclusive_scan(nullptr, tmp, flag, pos, 0, (size_t)n,
                                        rocprim::plus<int32_t>(), st));
  MLX_TRY(m.d_scan.reserve(std::max<size_t>(tmp, 256)));
  MLX_CHECK_HIP(rocprim::exclusive_scan(m.d_scan.ptr, tmp, flag, pos, 0, (size_t)n,
                                        rocprim::plus<int32_t>(), st));
  hipLaunchKernelGGL(k_compact_flags, dim3(grid_for(n)), dim3(kMapBlock), 0, st, n, flag, pos,
                     list, lambda, count);
  return MLX_OK;
}

}  // namespace mlx
