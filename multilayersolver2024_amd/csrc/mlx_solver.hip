// GPU orchestration of the multi-layer XPBD substep (gfx950): scene upload
// into color-major slots, layerSolve (layers x iterations x colors), the
// integrate() prologue/epilogue and the C ABI entry points that touch the
// device.  Kernels are in mlx_kernels.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "mlx_host.h"
#include "mlx_internal.h"
#include "mlx_kernels.h"
#include "mlx_learn_kernels.h"
#include "mlx_batch_layered.h"
#include "mlx_contacts.h"

using namespace mlx;

namespace mlx {

// Work vectors (r, z, d, Ap) of mlx_pcg_solve_noalloc: one copy per device,
// held until the process ends (2 MiB).  Host code never passes this name to
// a kernel: it passes the device address resolved by pcg_scratch_resolve.
static constexpr int64_t kPcgScratchMaxN = 65536;
static __device__ __align__(256) double scratch[4 * kPcgScratchMaxN];

int DevBuf::reserve(size_t n) {
  if (n <= bytes && ptr) return MLX_OK;
  release();
  size_t want = n < 256 ? 256 : n;
  if (hipMalloc(&ptr, want) != hipSuccess) {
    ptr = nullptr;
    bytes = 0;
    set_error("hipMalloc(%zu) failed", want);
    return MLX_ERR_NOMEM;
  }
  bytes = want;
  return MLX_OK;
}

void DevBuf::release() {
  // a buffer may still be read by queued kernels: drain the device first
  if (ptr) {
    (void)hipDeviceSynchronize();
    (void)hipFree(ptr);
  }
  ptr = nullptr;
  bytes = 0;
}

static int g_device = 0;

static int check_device(int dev) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    set_error("no HIP device visible");
    return MLX_ERR_NODEVICE;
  }
  if (dev < 0 || dev >= n) {
    set_error("device %d out of range (0..%d)", dev, n - 1);
    return MLX_ERR_NODEVICE;
  }
  hipDeviceProp_t prop;
  MLX_CHECK_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error("device %d is %s; this library is built for gfx950", dev, prop.gcnArchName);
    return MLX_ERR_NODEVICE;
  }
  return MLX_OK;
}

// Device address of the no-allocation solve's scratch array, per device.
// Resolved once per device with that device current (mlx_init, or the
// entry point's first call on a device); host code uses only this address.
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices] = {};

static int pcg_scratch_resolve(int dev, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device %d out of range for the scratch table", dev);
    return MLX_ERR_NODEVICE;
  }
  if (!g_scratch_dev[dev]) {
    int cur = -1;
    MLX_CHECK_HIP(hipGetDevice(&cur));
    if (cur != dev) MLX_CHECK_HIP(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(mlx::scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      return MLX_ERR_HIP;
    }
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_dev[dev];
  return MLX_OK;
}

static inline unsigned blocks_for(int64_t n) {
  return (unsigned)((n + kBlock - 1) / kBlock);
}

// ---------------------------------------------------------------------------
// host-side layer data
// ---------------------------------------------------------------------------

struct HostContacts {
  int K = 0;
  std::vector<int32_t> vert;        // 0-based
  std::vector<double> compliance;
};

struct SolveParams {
  int dim = 3;
  int nLayers = 1;
  int32_t layersMap0[MLX_MAX_LAYERS];
  int32_t iterations[MLX_MAX_LAYERS];
  double h = 0.01;
  double boundaryCompliance = 1e-8;
  double contactResistance = 0.5;
  int computeResiduals = 0, hangingStop = 1000, giveUpEnabled = 0;
  int lastLayerGiveUpEnabled = 0, useRigidContacts = 1;
  double giveUpThreshold = 0, stopAtImprovementPercent = 0;
  double runUntil = -1, residualLast = 0;
};

struct SolveResult {
  double runUntil = -1;
  int residualCount = 1;
  std::vector<double> residuals;
  int itersPerLayer[MLX_MAX_LAYERS] = {0};
  int layersRun = 0;
  int64_t projections = 0;
};

static constexpr int kMaxPartials = 65536;
static constexpr int kXcds = 8;  // XCDs of an MI355X (workgroups are dealt round-robin over them)
static constexpr int kSweepWg = 1024;  // threads of the single-workgroup sweep (k_sweep_wg)

// The persistent-loop kernel for the mesh's dimension and barrier mode.
// The XCD-local form is 3D only.
static bool persist_local(const mlx_mesh &m) { return m.persistLocal && m.dim == 3; }

template <typename T>
static const void *persist_fn(const mlx_mesh &m) {
  if (persist_local(m)) return reinterpret_cast<const void *>(&k_persist<T, 3, true>);
  return m.dim == 3 ? reinterpret_cast<const void *>(&k_persist<T, 3, false>)
                    : reinterpret_cast<const void *>(&k_persist<T, 2, false>);
}

// Upload the 1-based column-major contact arrays of a layerSolve call.
// Contact arrays live at offsets fixed by the capacity (one contact per
// vertex at most), so captured graphs survive a change of the contact count.
static size_t contact_cap(const mlx_mesh &m) { return (size_t)std::max<int64_t>(m.nV, 1); }

static int reserve_contacts(mlx_mesh &m) {
  const size_t cap = contact_cap(m);
  return m.d_contacts.reserve(68 * cap + 64 + 256);
}

static ContactArrays contact_arrays(mlx_mesh &m, int K) {
  const size_t cap = contact_cap(m);
  char *base = m.d_contacts.as<char>();
  ContactArrays c;
  c.normal = reinterpret_cast<const double *>(base);
  c.point = reinterpret_cast<const double *>(base + 24 * cap);
  c.compliance = reinterpret_cast<const double *>(base + 48 * cap);
  c.lambda = reinterpret_cast<double *>(base + 56 * cap);
  c.vert = reinterpret_cast<const int32_t *>(base + 64 * cap);
  c.counts = reinterpret_cast<const int32_t *>(base + ((68 * cap + 15) & ~(size_t)15));
  c.K = K;
  return c;
}

static int upload_contacts(mlx_mesh &m, int dim, int K, const int32_t *vid1,
                           const double *normals, const double *points,
                           const double *compl_, HostContacts &hc) {
  hc.K = K;
  hc.vert.resize(K);
  hc.compliance.resize(K);
  std::vector<double> nrm(3 * (size_t)K, 0.0), pts(3 * (size_t)K, 0.0);
  std::vector<uint8_t> seen(m.nV, 0);
  for (int i = 0; i < K; ++i) {
    int v = vid1[i] - 1;
    if (v < 0 || v >= m.nV) {
      set_error("contactVertexID(%d) = %d out of range", i + 1, vid1[i]);
      return MLX_ERR_ARG;
    }
    if (seen[v]) {
      set_error("vertex %d has more than one contact; the GPU path requires one contact per vertex",
                v + 1);
      return MLX_ERR_UNSUPPORTED;
    }
    seen[v] = 1;
    hc.vert[i] = v;
    hc.compliance[i] = compl_[i];
    for (int d = 0; d < dim; ++d) {
      nrm[3 * i + d] = normals[i + (size_t)K * d];
      pts[3 * i + d] = points[i + (size_t)K * d];
    }
  }
  MLX_TRY(reserve_contacts(m));
  ContactArrays c = contact_arrays(m, K);
  if (K > 0) {
    MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
    MLX_CHECK_HIP(hipMemcpy(const_cast<double *>(c.normal), nrm.data(), 24 * (size_t)K,
                            hipMemcpyHostToDevice));
    MLX_CHECK_HIP(hipMemcpy(const_cast<double *>(c.point), pts.data(), 24 * (size_t)K,
                            hipMemcpyHostToDevice));
    MLX_CHECK_HIP(hipMemcpy(const_cast<double *>(c.compliance), hc.compliance.data(),
                            8 * (size_t)K, hipMemcpyHostToDevice));
    MLX_CHECK_HIP(hipMemcpy(const_cast<int32_t *>(c.vert), hc.vert.data(), 4 * (size_t)K,
                            hipMemcpyHostToDevice));
  }
  return MLX_OK;
}

// p[0..n) = v and dst = src, as kernels: a substep that is captured into a
// graph holds no memset or memcpy nodes.
template <typename V>
__global__ void k_fill(int64_t n, V *p, V v) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename V>
__global__ void k_copy(int64_t n, const V *src, V *dst) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Layout conversions of the layerSolve seam on the device.  The caller's
// arrays (double, Matlab layouts) travel through pinned staging as they
// are, and these kernels pack them: positions into vec4 with w = 1/m,
// lambda (nT x nc column-major, element order) into the slots, scalars to
// the working precision; the unpack kernels reverse it for the results.
template <typename T>
__global__ void k_pack_positions(int64_t nV, int dim, const double *src, const double *minv,
                                 vec4_t<T> *dst) {
  const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (v >= nV) return;
  vec4_t<T> q;
  q.x = (T)src[dim * v];
  q.y = (T)src[dim * v + 1];
  q.z = dim == 3 ? (T)src[dim * v + 2] : T(0);
  q.w = (T)minv[v];
  dst[v] = q;
}

template <typename T>
__global__ void k_unpack_positions(int64_t nV, int dim, const vec4_t<T> *src, double *dst) {
  const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (v >= nV) return;
  const vec4_t<T> q = src[v];
  dst[dim * v] = (double)q.x;
  dst[dim * v + 1] = (double)q.y;
  if (dim == 3) dst[dim * v + 2] = (double)q.z;
}

template <typename T>
__global__ void k_pack_lambda(int64_t nT, int nc, int64_t ld, const double *src,
                              const int32_t *slotOf, T *lam) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= nT) return;
  const int64_t s = slotOf[e];
  for (int i = 0; i < nc; ++i) lam[i * ld + s] = (T)src[e + nT * i];
}

template <typename T>
__global__ void k_unpack_lambda(int64_t nT, int nc, int64_t ld, const T *lam,
                                const int32_t *slotOf, double *dst) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= nT) return;
  const int64_t s = slotOf[e];
  for (int i = 0; i < nc; ++i) dst[e + nT * i] = (double)lam[i * ld + s];
}

template <typename S, typename D>
__global__ void k_cast(int64_t n, const S *src, D *dst) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = (D)src[i];
}

// Elastic flags of every layer from element order (0/1 per element and
// layer) into slot order; padding slots keep the 0 they were filled with.
__global__ void k_slot_flags_u8(int64_t nT, int ns, int64_t ld, const uint8_t *elastic,
                                const int32_t *slotOf, uint8_t *flags) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= nT) return;
  const int64_t s = slotOf[e];
  for (int l = 0; l < ns; ++l) flags[l * ld + s] = elastic[(int64_t)l * nT + e] ? 1 : 0;
}

template <typename T>
static SweepParams<T> sweep_params(mlx_mesh &m, double h, const Ctl *ctl) {
  SweepParams<T> P;
  P.vid = m.d_vid.as<int32_t>();
  P.dminv = m.d_dminv.as<T>();
  P.vol = m.d_vol.as<T>();
  P.mat = m.d_mat.as<T>();
  P.lam = m.d_lam.as<T>();
  P.flags = nullptr;
  P.x = m.d_x.as<vec4_t<T>>();
  P.xold = m.d_xold.as<vec4_t<T>>();
  P.ctl = ctl;
  P.ld = m.ld;
  P.begin = 0;
  P.end = 0;
  P.ih2 = (T)(1.0 / (h * h));
  P.ih = (T)(1.0 / h);
  P.nh = m.hasNH ? 1 : 0;
  return P;
}

// Profiling records one event pair around the color sweeps of an iteration
// (all launches back to back on the mesh's stream), so the measured time is
// kernel execution plus the dispatch gaps between consecutive launches, not
// the cost of an event pair per launch.  Events are created on demand
// without the system-scope fence and kept for the mesh's lifetime.
static hipEvent_t prof_event(mlx_mesh &m, size_t i) {
  while (m.prof.ev.size() <= i) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    m.prof.ev.push_back(e);
  }
  return m.prof.ev[i];
}

static constexpr size_t kProfPairs = 8192;

template <typename T>
static void enqueue_sweeps(mlx_mesh &m, SweepParams<T> P, const uint8_t *flags,
                           const int64_t *elasticPerColor = nullptr, int iter = 0) {
  P.flags = flags;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (m.prof.on && m.prof.used < kProfPairs) {
    e0 = prof_event(m, 2 * m.prof.used);
    e1 = prof_event(m, 2 * m.prof.used + 1);
    if (e0 && e1) {
      int64_t elems = 0, launches = 0;
      for (int c = 0; c < m.numColors; ++c) {
        if (m.colorCnt[c] == 0) continue;
        if (flags && elasticPerColor && elasticPerColor[c] == 0) continue;
        ++launches;
        elems += elasticPerColor ? elasticPerColor[c] : m.colorCnt[c];
      }
      m.prof.elems.resize(m.prof.used + 1);
      m.prof.nlaunch.resize(m.prof.used + 1);
      m.prof.iter.resize(m.prof.used + 1);
      m.prof.elems[m.prof.used] = elems;
      m.prof.nlaunch[m.prof.used] = launches;
      m.prof.iter[m.prof.used] = iter;
      m.prof.used++;
      (void)hipEventRecord(e0, m.stream);
    }
  }
  const int nb = m.sweepBlock;
  for (int c = 0; c < m.numColors; ++c) {
    P.begin = m.colorOff[c];
    P.end = m.colorOff[c] + m.colorCnt[c];
    if (P.end <= P.begin) continue;
    // a color with no elastic element in a layer with rigid ones projects
    // nothing: not launched
    if (flags && elasticPerColor && elasticPerColor[c] == 0) continue;
    const unsigned grid = (unsigned)((P.end - P.begin + nb - 1) / nb);
    if (m.dim == 3 && P.nh)
      hipLaunchKernelGGL((k_sweep3d<T, kMatMixed>), dim3(grid), dim3(nb), 0, m.stream, P);
    else if (m.dim == 3)
      hipLaunchKernelGGL((k_sweep3d<T, kMatStVK>), dim3(grid), dim3(nb), 0, m.stream, P);
    else
      hipLaunchKernelGGL(k_sweep2d<T>, dim3(grid), dim3(nb), 0, m.stream, P);
  }
  if (e0 && e1) (void)hipEventRecord(e1, m.stream);
}

// One iteration's sweep of a layer with few elastic elements per color:
// a single k_sweep_wg workgroup over the compacted color lists.  Not in the
// sweep timing of mlx_profile_read, which covers k_sweep3d/k_sweep2d.
template <typename T>
static void enqueue_sweep_wg(mlx_mesh &m, SweepParams<T> P) {
  const char *eb = m.d_elist.as<char>();
  const int nc = m.numColors;
  hipLaunchKernelGGL((k_sweep_wg<T, kSweepWg>), dim3(1), dim3(kSweepWg), 0, m.stream, P,
                     reinterpret_cast<const int64_t *>(eb),
                     reinterpret_cast<const int32_t *>(eb + 20 * (size_t)nc),
                     reinterpret_cast<const int32_t *>(eb + 16 * (size_t)nc), nc);
}

// After a stream sync: fold the recorded iterations whose index is below
// itersDone (the rest returned at once: the layer had stopped).
static void prof_harvest(mlx_mesh &m, int itersDone) {
  for (size_t i = 0; i < m.prof.used; ++i) {
    if (m.prof.iter[i] >= itersDone || m.prof.elems[i] == 0) continue;
    float ms = 0;
    if (hipEventElapsedTime(&ms, m.prof.ev[2 * i], m.prof.ev[2 * i + 1]) == hipSuccess) {
      m.prof.ms += ms;
      m.prof.launches += m.prof.nlaunch[i];
      m.prof.elements += m.prof.elems[i];
    }
  }
  m.prof.used = 0;
}

// Per-layer CSR lists, built on the host from the layer maps.
struct LayerCSR {
  int nb = 0, nbnd = 0, nrc = 0, nec = 0, nchunks = 0, nbchunks = 0;
  // rig_off | rig_vert | bnd_off | bnd_vert | rc_off | rc_idx | ec_idx | chk_off | bchk_off
  std::vector<int32_t> buf;
  size_t o_rig_off, o_rig_vert, o_bnd_off, o_bnd_vert, o_rc_off, o_rc_idx, o_ec, o_chk, o_bchk;
};

static void build_csr(const mlx_mesh &m, const HostMaps &maps, int col, const HostContacts &hc,
                      LayerCSR &L) {
  const int64_t nV = m.nV;
  const int32_t *vb = maps.vbody.data() + (size_t)nV * col;
  const uint8_t *bd = maps.boundary.data() + (size_t)nV * col;
  const int nb = maps.numRigids[col];
  L.nb = nb;
  std::vector<int32_t> cr(nb + 1, 0), cb(nb + 1, 0), cc(nb + 1, 0);
  int nrig = 0;
  for (int64_t v = 0; v < nV; ++v)
    if (vb[v] > 0) {
      cr[vb[v]]++;
      nrig++;
      if (bd[v]) cb[vb[v]]++;
    }
  int nec = 0, nrc = 0;
  for (int i = 0; i < hc.K; ++i) {
    int b = vb[hc.vert[i]];
    if (b > 0) {
      cc[b]++;
      nrc++;
    } else {
      nec++;
    }
  }
  int nbnd = 0;
  for (int k = 0; k < nb; ++k) {
    cr[k + 1] += cr[k];
    cb[k + 1] += cb[k];
    cc[k + 1] += cc[k];
  }
  nbnd = cb[nb];
  L.nbnd = nbnd;
  L.nrc = nrc;
  L.nec = nec;
  L.o_rig_off = 0;
  L.o_rig_vert = L.o_rig_off + nb + 1;
  L.o_bnd_off = L.o_rig_vert + nrig;
  L.o_bnd_vert = L.o_bnd_off + nb + 1;
  L.o_rc_off = L.o_bnd_vert + nbnd;
  L.o_rc_idx = L.o_rc_off + nb + 1;
  L.o_ec = L.o_rc_idx + nrc;
  L.o_chk = L.o_ec + nec + 1;
  L.o_bchk = L.o_chk + nb + 1;
  L.buf.assign(L.o_bchk + nb + 1, 0);
  int32_t *B = L.buf.data();
  for (int k = 0; k <= nb; ++k) {
    B[L.o_rig_off + k] = cr[k];
    B[L.o_bnd_off + k] = cb[k];
    B[L.o_rc_off + k] = cc[k];
  }
  // restriction chunks: kRestrictChunk rigid vertices each, at least one per body
  for (int k = 0; k < nb; ++k)
    B[L.o_chk + k + 1] =
        B[L.o_chk + k] + std::max(1, (cr[k + 1] - cr[k] + kRestrictChunk - 1) / kRestrictChunk);
  L.nchunks = B[L.o_chk + nb];
  // boundary chunks: kBoundaryChunk boundary vertices each, at least one per body
  for (int k = 0; k < nb; ++k)
    B[L.o_bchk + k + 1] =
        B[L.o_bchk + k] + std::max(1, (cb[k + 1] - cb[k] + kBoundaryChunk - 1) / kBoundaryChunk);
  L.nbchunks = B[L.o_bchk + nb];
  std::vector<int32_t> fr(cr.begin(), cr.end() - 1), fb(cb.begin(), cb.end() - 1),
      fc(cc.begin(), cc.end() - 1);
  for (int64_t v = 0; v < nV; ++v)
    if (vb[v] > 0) {
      const int k = vb[v] - 1;
      B[L.o_rig_vert + fr[k]++] = (int32_t)v;
      if (bd[v]) B[L.o_bnd_vert + fb[k]++] = (int32_t)v;
    }
  int e = 0;
  for (int i = 0; i < hc.K; ++i) {
    int b = vb[hc.vert[i]];
    if (b > 0)
      B[L.o_rc_idx + fc[b - 1]++] = i;
    else
      B[L.o_ec + e++] = i;
  }
}

// Pinned staging for the copies of one substep.  Copies from or into
// pinned memory are asynchronous, so enqueueing them never waits for the
// stream (a copy from pageable memory does).  Regions are handed out in
// order and stay valid until the next stage_begin, which drains the stream
// first if regions are in use; a buffer outgrown meanwhile is kept until
// then, since queued copies may still use it.
int stage_begin(mlx_mesh &m) {
  if (m.stageUsed || !m.stageRetired.empty()) MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  for (char *p : m.stageRetired) (void)hipHostFree(p);
  m.stageRetired.clear();
  m.stageUsed = 0;
  return MLX_OK;
}

int stage_take(mlx_mesh &m, size_t bytes, char **out) {
  const size_t off = (m.stageUsed + 255) & ~(size_t)255;
  if (off + bytes > m.stageCap) {
    const size_t cap = std::max<size_t>(2 * (off + bytes), (size_t)1 << 20);
    char *p = nullptr;
    MLX_CHECK_HIP(hipHostMalloc((void **)&p, cap, hipHostMallocDefault));
    if (m.h_stage) m.stageRetired.push_back(m.h_stage);
    m.h_stage = p;
    m.stageCap = cap;
    *out = p;
    m.stageUsed = bytes;
    return MLX_OK;
  }
  m.stageUsed = off + bytes;
  *out = m.h_stage + off;
  return MLX_OK;
}

int stage_upload(mlx_mesh &m, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return MLX_OK;
  char *p = nullptr;
  MLX_TRY(stage_take(m, bytes, &p));
  std::memcpy(p, src, bytes);
  MLX_CHECK_HIP(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, m.stream));
  return MLX_OK;
}

// Chunk dependences of the dataflow sweep (k_flow), built once per mesh.
// Chunk j of color c holds slots colorOff[c] + 64 j .. + 63 (one
// wavefront); its dependences are the chunks holding, for each vertex of
// its elements, the element of the largest color below the element's own
// among the vertex's elements (the vertex's previous updater in the color
// order), without repeats.
static int flow_prepare(mlx_mesh &m) {
  if (m.flowReady) return MLX_OK;
  const int64_t nV = m.nV, nT = m.nT;
  const int vpe = m.vpe, nc = m.numColors;
  std::vector<int32_t> chunkBase(nc + 1, 0);
  for (int c = 0; c < nc; ++c)
    chunkBase[c + 1] = chunkBase[c] + (int32_t)((m.colorCnt[c] + 63) / 64);
  const int nChunks = chunkBase[nc];
  auto chunk_of = [&](int32_t e) {
    const int c = m.color[e] - 1;
    const int64_t sl = m.slotOf[e];
    return chunkBase[c] + (int32_t)((sl - m.colorOff[c]) / 64);
  };
  // elements of each vertex
  std::vector<int32_t> off(nV + 1, 0);
  for (int64_t e = 0; e < nT; ++e)
    for (int a = 0; a < vpe; ++a) off[m.T[(size_t)e * vpe + a] + 1]++;
  for (int64_t v = 0; v < nV; ++v) off[v + 1] += off[v];
  std::vector<int32_t> el(off[nV]), fill(off.begin(), off.end() - 1);
  for (int64_t e = 0; e < nT; ++e)
    for (int a = 0; a < vpe; ++a) el[fill[m.T[(size_t)e * vpe + a]]++] = (int32_t)e;
  std::vector<std::vector<int32_t>> deps(nChunks);
  for (int64_t e = 0; e < nT; ++e) {
    const int32_t ce = m.color[e], g = chunk_of((int32_t)e);
    for (int a = 0; a < vpe; ++a) {
      const int32_t v = m.T[(size_t)e * vpe + a];
      int32_t best = -1, bestColor = 0;
      for (int32_t q = off[v]; q < off[v + 1]; ++q) {
        const int32_t f = el[q], cf = m.color[f];
        if (cf < ce && cf > bestColor) {
          bestColor = cf;
          best = f;
        }
      }
      if (best >= 0) deps[g].push_back(chunk_of(best));
    }
  }
  std::vector<int32_t> buf;
  buf.reserve((size_t)nc + 2 + nChunks + 8 * (size_t)nChunks);
  buf.insert(buf.end(), chunkBase.begin(), chunkBase.end());
  const size_t o_depOff = buf.size();
  buf.resize(o_depOff + nChunks + 1, 0);
  for (int g = 0; g < nChunks; ++g) {
    std::sort(deps[g].begin(), deps[g].end());
    deps[g].erase(std::unique(deps[g].begin(), deps[g].end()), deps[g].end());
    buf[o_depOff + g + 1] = buf[o_depOff + g] + (int32_t)deps[g].size();
  }
  for (int g = 0; g < nChunks; ++g) buf.insert(buf.end(), deps[g].begin(), deps[g].end());
  MLX_TRY(m.d_flowdeps.reserve(sizeof(int32_t) * buf.size()));
  MLX_TRY(m.d_flowflags.reserve(sizeof(uint32_t) * (size_t)std::max(nChunks, 1)));
  MLX_TRY(m.d_flowsync.reserve(sizeof(uint32_t) * 64));
  if (!m.h_flowerr) MLX_CHECK_HIP(hipHostMalloc((void **)&m.h_flowerr, sizeof(int32_t)));
  MLX_CHECK_HIP(hipMemcpy(m.d_flowdeps.ptr, buf.data(), sizeof(int32_t) * buf.size(),
                          hipMemcpyHostToDevice));
  m.flowChunks = nChunks;
  m.flowDeps = (int)(buf.size() - o_depOff - nChunks - 1);
  m.flowReady = true;
  return MLX_OK;
}

static int read_ctl(mlx_mesh &m, Ctl &out) {
  MLX_CHECK_HIP(hipMemcpyAsync(m.h_ctl, m.d_ctl.ptr, sizeof(Ctl), hipMemcpyDeviceToHost, m.stream));
  MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  out = *m.h_ctl;
  return MLX_OK;
}

// layerSolve on the device (xpbdLayer3D.m:44-232, xpbdLayer2D.m:606-738).
// Expects d_x (xi), d_xold (oldp), d_lam, d_rv (3D), d_pinnedx and the
// contact arrays to be on the device already.
// ---------------------------------------------------------------------------
// learned warm start and key correction (3D)
// ---------------------------------------------------------------------------

// The network on every slot from the state at the start of a substep
// (positions, the previous substep's positions and multipliers): lambda_0
// and the key factors, and the raw features when feat is given.
template <typename T>
static int learn_infer(mlx_mesh &m, double h, float *feat) {
  mlx_mesh::Learn &L = m.learn;
  SweepParams<T> P = sweep_params<T>(m, h, nullptr);
  const unsigned nb = (unsigned)((m.ld + kLearnBlock - 1) / kLearnBlock);
  hipLaunchKernelGGL(k_learn_infer<T>, dim3(nb), dim3(kLearnBlock), 0, m.stream, P,
                     mlp_device_view(L.model), (const vec4_t<T> *)m.d_xold.as<vec4_t<T>>(),
                     (float)(1.0 / h), (float)(1.0 / (h * h)), L.d_lam0.as<T>(),
                     L.d_keyScale.as<double>(), L.kappaMax, feat);
  MLX_CHECK_HIP(hipGetLastError());
  return MLX_OK;
}

// Warm start of the last layer with its fallback: residual of the cold
// start, copies of positions and multipliers, the displacement of every
// fresh elastic slot gathered per vertex, residual of the warm start, and
// the choice (restoring the copies when the warm start is worse).
template <typename T>
static int learn_warm_start(mlx_mesh &m, SweepParams<T> P, const uint8_t *fl,
                            const ContactArrays &ca, int nE, int nC, const SolveParams &sp) {
  mlx_mesh::Learn &L = m.learn;
  hipStream_t st = m.stream;
  const int64_t ld = m.ld, nV = m.nV, nLam = (int64_t)m.nc * ld;
  double *res = L.d_res.as<double>(), *part = res + 2;
  P.flags = fl;
  auto residual = [&](double *out) {
    hipLaunchKernelGGL(k_learn_res_part<T>, dim3(nE + nC), dim3(kBlock), 0, st, P, nE, ca,
                       sp.contactResistance, part);
    hipLaunchKernelGGL(k_learn_res_final, dim3(1), dim3(kBlock), 0, st, (const double *)part,
                       nE + nC, out);
  };
  residual(res);
  vec4_t<T> *xs = L.d_save.as<vec4_t<T>>();
  T *lams = reinterpret_cast<T *>(xs + nV);
  MLX_CHECK_HIP(hipMemcpyAsync(xs, P.x, sizeof(vec4_t<T>) * nV, hipMemcpyDeviceToDevice, st));
  MLX_CHECK_HIP(hipMemcpyAsync(lams, P.lam, sizeof(T) * nLam, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_learn_disp<T>, dim3(blocks_for(ld)), dim3(kBlock), 0, st, P,
                     (const T *)L.d_lam0.as<T>(), L.d_dx.as<T>());
  const int32_t *inc = L.d_inc.as<int32_t>();
  hipLaunchKernelGGL(k_learn_gather<T>, dim3(blocks_for(nV)), dim3(kBlock), 0, st, nV, ld, inc,
                     inc + nV + 1, (const T *)L.d_dx.as<T>(),
                     (const uint8_t *)m.d_pinMask.as<uint8_t>(), P.x);
  residual(res + 1);
  hipLaunchKernelGGL(k_learn_select<T>, dim3(blocks_for(std::max(nV, nLam))), dim3(kBlock), 0, st,
                     nV, nLam, (const double *)res, P.x, (const vec4_t<T> *)xs, P.lam,
                     (const T *)lams, L.d_counts.as<unsigned long long>());
  MLX_CHECK_HIP(hipGetLastError());
  return MLX_OK;
}

template <typename T>
static int layer_solve_dev(mlx_mesh &m, const HostMaps &maps, const SolveParams &sp,
                           const HostContacts &hc, SolveResult &res) {
  const int64_t nV = m.nV, nT = m.nT, ld = m.ld;
  const int ns = maps.nSorted;
  hipStream_t st = m.stream;
  // elastic flags in slot order, one column per sorted layer (already on
  // the device when the maps were built there)
  std::vector<int64_t> nElastic(ns, 0), perColor((size_t)ns * m.numColors, 0);
  if (maps.deviceFlags) {
    nElastic = maps.nElastic;
    perColor = maps.perColor;
  } else {
    // counts by a sequential pass over the elements; the flags are put in
    // slot order on the device (k_slot_flags_u8 below)
    for (int l = 0; l < ns; ++l) {
      const uint8_t *el = maps.elastic.data() + (size_t)nT * l;
      int64_t *pc = perColor.data() + (size_t)m.numColors * l;
      for (int64_t e = 0; e < nT; ++e)
        if (el[e]) {
          ++nElastic[l];
          ++pc[m.color[e] - 1];
        }
    }
  }
  // layers with rigid elements and at most kSweepWg elastic elements in
  // every color sweep all colors in one workgroup (k_sweep_wg)
  std::vector<uint8_t> wgLayer(ns, 0);
  bool anyWg = false;
  if (m.useSweepWg && m.dim == 3) {
    for (int l = 0; l < ns; ++l) {
      int64_t mx = 0;
      for (int c = 0; c < m.numColors; ++c) mx = std::max(mx, perColor[(size_t)m.numColors * l + c]);
      wgLayer[l] = nElastic[l] < nT && mx <= kSweepWg;
      anyWg = anyWg || wgLayer[l];
    }
  }
  if (anyWg) {
    const size_t nc = (size_t)m.numColors;
    MLX_TRY(m.d_elist.reserve(20 * nc + 4 * (size_t)ld));
    std::vector<int64_t> oc(2 * nc);
    for (size_t c = 0; c < nc; ++c) {
      oc[c] = m.colorOff[c];
      oc[nc + c] = m.colorCnt[c];
    }
    MLX_TRY(stage_upload(m, m.d_elist.ptr, oc.data(), 16 * nc));
  }
  if (!maps.deviceFlags) {
    MLX_TRY(m.d_flags.reserve((size_t)ns * ld));
    MLX_TRY(m.d_elastic.reserve(std::max<size_t>(maps.elastic.size(), 1)));
    MLX_TRY(stage_upload(m, m.d_elastic.ptr, maps.elastic.data(), maps.elastic.size()));
    MLX_CHECK_HIP(hipMemsetAsync(m.d_flags.ptr, 0, (size_t)ns * ld, st));
    hipLaunchKernelGGL(k_slot_flags_u8, dim3(blocks_for(nT)), dim3(kBlock), 0, st, nT, ns, ld,
                       (const uint8_t *)m.d_elastic.as<uint8_t>(), m.d_slotOf.as<int32_t>(),
                       m.d_flags.as<uint8_t>());
  }
  MLX_TRY(m.d_vbody.reserve(sizeof(int32_t) * maps.vbody.size()));
  if (maps.devVbody)
    MLX_CHECK_HIP(hipMemcpyAsync(m.d_vbody.ptr, maps.devVbody, sizeof(int32_t) * maps.vbody.size(),
                                 hipMemcpyDeviceToDevice, st));
  else
    MLX_TRY(stage_upload(m, m.d_vbody.ptr, maps.vbody.data(), sizeof(int32_t) * maps.vbody.size()));
  int maxNb = 1;
  for (int l = 0; l < ns; ++l) maxNb = std::max(maxNb, maps.numRigids[l]);
  MLX_TRY(m.d_bodies.reserve(sizeof(Body) * maxNb));
  MLX_TRY(m.d_ri.reserve(sizeof(double) * 3 * nV));
  MLX_TRY(m.d_partials.reserve(sizeof(double) * kMaxPartials));
  const int residCap = 1 << 16;
  MLX_TRY(m.d_resid.reserve(sizeof(double) * residCap));
  MLX_TRY(m.d_ctl.reserve(sizeof(Ctl)));
  if (!m.d_count.ptr) {
    // arrival counter of k_residual_tail: zero here, re-armed by the kernel
    MLX_TRY(m.d_count.reserve(256));
    MLX_CHECK_HIP(hipMemsetAsync(m.d_count.ptr, 0, 256, st));
  }
  const int K = hc.K;
  ContactArrays ca = contact_arrays(m, K);

  // distinct contact compliances and their multiplicities (3D residual)
  std::vector<double> uniq, cnt;
  for (int i = 0; i < K; ++i) {
    auto it = std::find(uniq.begin(), uniq.end(), hc.compliance[i]);
    if (it == uniq.end()) {
      uniq.push_back(hc.compliance[i]);
      cnt.push_back(1.0);
    } else {
      cnt[it - uniq.begin()] += 1.0;
    }
  }
  const int nUniq = (int)uniq.size();
  // fixed layout (capacity, not count) so captured graphs stay valid
  const int uniqCap = std::max(64, nUniq);
  MLX_TRY(m.d_tmp.reserve(sizeof(double) * 2 * uniqCap + sizeof(int32_t) * 3 * (m.dim == 2 ? ld : 1)));
  double *d_uniq = m.d_tmp.as<double>();
  double *d_cnt = d_uniq + uniqCap;
  int32_t *d_lamSlot2d = reinterpret_cast<int32_t *>(d_cnt + uniqCap);
  if (nUniq > 0) {
    MLX_TRY(stage_upload(m, d_uniq, uniq.data(), 8 * (size_t)nUniq));
    MLX_TRY(stage_upload(m, d_cnt, cnt.data(), 8 * (size_t)nUniq));
  }
  std::vector<int32_t> lamSlot;
  if (m.dim == 2 && sp.computeResiduals) {
    // page e, entry r of reshape(lambdai,3,1,[]) is lambdai(L mod nT, L div nT)
    lamSlot.assign(3 * (size_t)ld, 0);
    for (int64_t s = 0; s < ld; ++s) {
      const int32_t e = m.perm[s];
      if (e < 0) continue;
      for (int r = 0; r < 3; ++r) {
        const int64_t Lin = 3 * (int64_t)e + r;
        const int64_t el = Lin % nT, comp = Lin / nT;
        lamSlot[3 * s + r] = (int32_t)(m.slotOf[el] * 3 + comp);
      }
    }
    MLX_TRY(stage_upload(m, d_lamSlot2d, lamSlot.data(), sizeof(int32_t) * lamSlot.size()));
  }

  Ctl c0;
  std::memset(&c0, 0, sizeof c0);
  c0.residualCount = 1;
  c0.stopLayer = -1;
  c0.runUntil = sp.runUntil;
  c0.lastResidual = sp.residualLast;
  MLX_TRY(stage_upload(m, m.d_ctl.ptr, &c0, sizeof c0));
  MLX_TRY(stage_upload(m, m.d_resid.ptr, &sp.residualLast, 8));
  Ctl *ctl = m.d_ctl.as<Ctl>();
  double *partials = m.d_partials.as<double>();
  Body *bodies = m.d_bodies.as<Body>();
  double *ri = m.d_ri.as<double>();
  vec4_t<T> *x = m.d_x.as<vec4_t<T>>();
  T *rv = m.d_rv.as<T>();
  const double *minv = m.d_minv64.as<double>();
  const double *mass = m.d_mass64.as<double>();
  SweepParams<T> P = sweep_params<T>(m, sp.h, ctl);
  MonitorParams mp;
  mp.is3d = m.dim == 3;
  mp.computeResiduals = sp.computeResiduals;
  mp.hangingStop = sp.hangingStop;
  mp.giveUpEnabled = sp.giveUpEnabled;
  mp.lastLayerGiveUpEnabled = sp.lastLayerGiveUpEnabled;
  mp.giveUpThreshold = sp.giveUpThreshold;
  mp.h = sp.h;
  const int nResElemBlocks = (int)blocks_for(ld);
  const int nResContactBlocks = K > 0 ? (int)blocks_for((int64_t)contact_cap(m)) : 0;
  if (nResElemBlocks + nResContactBlocks > kMaxPartials) {
    set_error("scene too large for the residual partial buffer");
    return MLX_ERR_UNSUPPORTED;
  }
  // k_residual_tail: kResPerThread slots per thread
  const int64_t resSpan = (int64_t)kBlock * kResPerThread;
  const int nTailElemBlocks = (int)((ld + resSpan - 1) / resSpan);

  const int64_t nPin = (int64_t)m.pinnedDOF.size();
  // persistent loop: color table on the device, barrier counter, grid size
  int persistG = 0, persistMaxCnt = 0;
  if (m.usePersistent) {
    std::vector<int32_t> ct(2 * (size_t)m.numColors);
    int64_t maxCnt = 1;
    for (int c = 0; c < m.numColors; ++c) {
      ct[c] = (int32_t)m.colorOff[c];
      ct[m.numColors + c] = (int32_t)m.colorCnt[c];
      maxCnt = std::max<int64_t>(maxCnt, m.colorCnt[c]);
    }
    persistMaxCnt = (int)maxCnt;
    MLX_TRY(m.d_colors.reserve(sizeof(int32_t) * ct.size() + 4));
    MLX_TRY(stage_upload(m, m.d_colors.ptr, ct.data(), sizeof(int32_t) * ct.size()));
    MLX_TRY(m.d_sync.reserve(256));
    // local: persistG workgroups per XCD are launched on each of the 8 XCDs
    // and one XCD's group runs the loop; otherwise persistG in all
    persistG = m.persistWG > 0 ? m.persistWG
               : persist_local(m)
                   ? (int)std::min<int64_t>(32, std::max<int64_t>(4, (ld + 8 * kBlock - 1) / (8 * kBlock)))
                   : (int)std::min<int64_t>(32, std::max<int64_t>(4, (maxCnt + kBlock - 1) / kBlock));
    int perCU = 0, dev = 0;
    MLX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, persist_fn<T>(m), kBlock, 0));
    MLX_CHECK_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MLX_CHECK_HIP(hipGetDeviceProperties(&prop, dev));
    persistG = std::min(persistG * (persist_local(m) ? kXcds : 1), perCU * prop.multiProcessorCount);
    if (persistG > kMaxPartials || persistG < 1) {
      set_error("persistent loop: no valid grid size");
      return MLX_ERR_STATE;
    }
    if (m.prof.on)
      for (auto &e : m.prof.pev)
        if (!e) MLX_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  }
  // dataflow sweep: per-slot order (once per mesh), the color table, the
  // group size (one slot per thread per color) and the grid (that many
  // workgroups on each of the 8 XCDs; one XCD's group runs)
  int flowG = 0;
  if (m.useFlow && m.dim == 3) {
    MLX_TRY(flow_prepare(m));
    std::vector<int32_t> ct(2 * (size_t)m.numColors);
    int64_t maxCnt = 1;
    for (int c = 0; c < m.numColors; ++c) {
      ct[c] = (int32_t)m.colorOff[c];
      ct[m.numColors + c] = (int32_t)m.colorCnt[c];
      maxCnt = std::max<int64_t>(maxCnt, m.colorCnt[c]);
    }
    MLX_TRY(m.d_colors.reserve(sizeof(int32_t) * ct.size() + 4));
    MLX_TRY(stage_upload(m, m.d_colors.ptr, ct.data(), sizeof(int32_t) * ct.size()));
    flowG = (int)std::min<int64_t>(32, std::max<int64_t>(4, (maxCnt + kBlock - 1) / kBlock));
    int perCU = 0, dev = 0;
    const void *fn = m.hasNH ? reinterpret_cast<const void *>(&k_flow<T, kMatMixed>)
                             : reinterpret_cast<const void *>(&k_flow<T, kMatStVK>);
    MLX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, kBlock, 0));
    MLX_CHECK_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MLX_CHECK_HIP(hipGetDeviceProperties(&prop, dev));
    if ((int64_t)flowG * kXcds > (int64_t)perCU * prop.multiProcessorCount) {
      set_error("dataflow sweep: %d workgroups do not fit the device", flowG * kXcds);
      return MLX_ERR_STATE;
    }
    if (m.prof.on)
      for (auto &e : m.prof.pev)
        if (!e) MLX_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  }
  res.projections = 0;
  res.layersRun = 0;
  Ctl hc_ctl;
  std::memset(&hc_ctl, 0, sizeof hc_ctl);
  // every layer's CSR lists, uploaded before the first layer runs, each in
  // a region of its own at a fixed offset (captured graphs keep their
  // pointers).  Bodies are vertex-disjoint with at least one vertex each
  // and a vertex has at most one contact, so a layer needs at most
  // 4 (nV + 1) + 3 nV + 1 entries whatever the substep's contacts.
  const size_t csrCap = 7 * (size_t)nV + 16;
  std::vector<LayerCSR> Ls(sp.nLayers);
  MLX_TRY(m.d_csr.reserve(sizeof(int32_t) * csrCap * (size_t)sp.nLayers));
  for (int layer = 0; layer < sp.nLayers; ++layer) {
    build_csr(m, maps, sp.layersMap0[layer], hc, Ls[layer]);
    if (Ls[layer].buf.size() > csrCap) {
      set_error("layer lists exceed their region (%zu > %zu entries)", Ls[layer].buf.size(), csrCap);
      return MLX_ERR_STATE;
    }
    MLX_TRY(stage_upload(m, m.d_csr.as<int32_t>() + csrCap * (size_t)layer, Ls[layer].buf.data(),
                         sizeof(int32_t) * Ls[layer].buf.size()));
  }
  for (int layer = 0; layer < sp.nLayers; ++layer) {
    const int col = sp.layersMap0[layer];
    const bool isLast = layer == sp.nLayers - 1;
    const uint8_t *fl = nElastic[col] == nT ? nullptr : m.d_flags.as<uint8_t>() + (size_t)ld * col;
    const int32_t *vb = m.d_vbody.as<int32_t>() + (size_t)nV * col;
    const LayerCSR &L = Ls[layer];
    const bool hasRigidEl = nElastic[col] < nT;
    const int32_t *cb = m.d_csr.as<int32_t>() + csrCap * (size_t)layer;
    LayerLists LL;
    LL.rig_off = cb + L.o_rig_off;
    LL.rig_vert = cb + L.o_rig_vert;
    LL.bnd_off = cb + L.o_bnd_off;
    LL.bnd_vert = cb + L.o_bnd_vert;
    LL.rc_off = cb + L.o_rc_off;
    LL.rc_idx = cb + L.o_rc_idx;
    LL.ec_idx = cb + L.o_ec;
    LL.chk_off = L.nb > 0 ? cb + L.o_chk : nullptr;
    LL.bchk_off = L.nb > 0 ? cb + L.o_bchk : nullptr;
    LL.nchunks = L.nchunks;
    LL.nbchunks = L.nbchunks;
    LL.nb = L.nb;
    LL.nbnd = L.nbnd;
    LL.nrc = L.nrc;
    LL.nec = L.nec;
    MLX_TRY(m.d_lamB.reserve(sizeof(double) * (L.nbnd + 1)));
    if (L.nbnd > 0)
      MLX_CHECK_HIP(hipMemsetAsync(m.d_lamB.ptr, 0, sizeof(double) * L.nbnd, st));
    if (K > 0) MLX_CHECK_HIP(hipMemsetAsync(ca.lambda, 0, sizeof(double) * K, st));
    // contact counts of this layer, read on the device by the iteration kernels
    m.h_counts[0] = K;
    m.h_counts[1] = L.nec;
    m.h_counts[2] = L.nrc;
    m.h_counts[3] = nUniq;
    MLX_CHECK_HIP(hipMemcpyAsync(const_cast<int32_t *>(ca.counts), m.h_counts, 4 * sizeof(int32_t),
                                 hipMemcpyHostToDevice, st));

    if (m.dim == 3) {
      if (L.nb > 0) {
        // per-body arrival counters of the chunked restriction and boundary
        // kernels: they start at zero and each body's last chunk re-arms its own
        const size_t need = sizeof(unsigned) * ((size_t)L.nb + 1);
        if (m.d_bodycnt.bytes < need) {
          MLX_TRY(m.d_bodycnt.reserve(need));
          MLX_CHECK_HIP(hipMemsetAsync(m.d_bodycnt.ptr, 0, m.d_bodycnt.bytes, st));
        }
      }
      if (L.nb > 0 && 12 * (int64_t)L.nchunks <= kMaxPartials) {
        // bodies split in chunks over many workgroups (the coarse layers'
        // bodies span most of the mesh)
        hipLaunchKernelGGL(k_restrict_mass<T>, dim3(L.nchunks), dim3(kBlock), 0, st, LL, x, rv,
                           mass, bodies, partials, m.d_bodycnt.as<unsigned>());
        hipLaunchKernelGGL(k_restrict_inertia<T>, dim3(L.nchunks), dim3(kBlock), 0, st, LL, x, rv,
                           mass, bodies, ri, partials, m.d_bodycnt.as<unsigned>());
      } else if (L.nb > 0) {
        hipLaunchKernelGGL(k_restrict<T>, dim3(L.nb), dim3(kBlock), 0, st, LL, x, rv, mass,
                           bodies, ri);
      }
      const int step = hasRigidEl && L.nb > 0;
      if (step)
        hipLaunchKernelGGL(k_body_step, dim3(blocks_for(L.nb)), dim3(kBlock), 0, st, bodies,
                           L.nb, sp.h);
      hipLaunchKernelGGL(k_layer_vertices<T>, dim3(blocks_for(nV)), dim3(kBlock), 0, st, nV, vb,
                         x, rv, bodies, ri, sp.h, step);
      hipLaunchKernelGGL(k_sumsq<T>, dim3(256), dim3(kBlock), 0, st, rv, 3 * nV, partials);
      hipLaunchKernelGGL(k_layer_begin, dim3(1), dim3(kBlock), 0, st, ctl, partials, 256, layer,
                         (int)isLast, layer == 0 ? sp.stopAtImprovementPercent : 0.0);
      if (isLast && L.nb == 0 && !hasRigidEl && m.learn.model && m.learn.ready &&
          (m.learn.flags & MLX_LEARN_WARMSTART))
        MLX_TRY(learn_warm_start<T>(m, P, fl, ca, nResElemBlocks, nResContactBlocks, sp));
    } else {
      // bodies from the current positions, not stepped (xpbdLayer2D.m:676-681)
      if (L.nb > 0)
        hipLaunchKernelGGL(k_restrict2d<T>, dim3(L.nb), dim3(kBlock), 0, st, LL, x, mass, bodies,
                           ri);
      hipLaunchKernelGGL(k_layer_begin, dim3(1), dim3(kBlock), 0, st, ctl, partials, 0, layer,
                         (int)isLast, layer == 0 ? sp.stopAtImprovementPercent : 0.0);
      // 2D layerSolve passes resvnorm = inf to projectGS (xpbdLayer2D.m:735)
      static const double inf = INFINITY;
      MLX_TRY(stage_upload(m, reinterpret_cast<char *>(m.d_ctl.ptr) + offsetof(Ctl, normresv), &inf,
                           sizeof(double)));
    }

    hc_ctl.iters = 0;  // k_layer_begin cleared the device count
    const bool sweepWg = wgLayer[col] != 0;
    // dataflow sweep (k_flow) for a layer without rigid elements: the
    // chunk flags and the error word start at zero
    const bool flow = flowG > 0 && !hasRigidEl && L.nb == 0 && !sweepWg;
    unsigned flowBase = 0;
    unsigned *flowSync = m.d_flowsync.as<unsigned>();
    if (flow) {
      hipLaunchKernelGGL(k_fill<unsigned>, dim3(blocks_for(m.flowChunks)), dim3(kBlock), 0, st,
                         (int64_t)m.flowChunks, m.d_flowflags.as<unsigned>(), 0u);
      hipLaunchKernelGGL(k_fill<unsigned>, dim3(1), dim3(kBlock), 0, st, (int64_t)64, flowSync, 0u);
    }
    if (sweepWg) {
      char *eb = m.d_elist.as<char>();
      const size_t nc = (size_t)m.numColors;
      hipLaunchKernelGGL(k_color_lists, dim3(m.numColors), dim3(kBlock), 0, st,
                         reinterpret_cast<const int64_t *>(eb),
                         reinterpret_cast<const int64_t *>(eb + 8 * nc), fl,
                         reinterpret_cast<int32_t *>(eb + 20 * nc),
                         reinterpret_cast<int32_t *>(eb + 16 * nc));
    }
    int iterEnqueued = 0;
    auto enqueue_flow = [&]() -> int {
      FlowParams<T> Q;
      Q.P = P;
      Q.P.flags = nullptr;
      Q.colorOff = m.d_colors.as<int32_t>();
      Q.colorCnt = m.d_colors.as<int32_t>() + m.numColors;
      Q.numColors = m.numColors;
      Q.chunkBase = m.d_flowdeps.as<int32_t>();
      Q.depOff = Q.chunkBase + m.numColors + 1;
      Q.dep = Q.depOff + m.flowChunks + 1;
      Q.flags = m.d_flowflags.as<unsigned>();
      Q.base = flowBase++;
      Q.ca = ca;
      Q.ecIdx = LL.ec_idx;
      Q.nec = L.nec;
      Q.minv = minv;
      Q.resistance = sp.contactResistance;
      Q.h = sp.h;
      Q.pinDof = m.d_pinned.as<int32_t>();
      Q.pinVal = m.d_pinnedx.as<T>();
      Q.nPin = nPin;
      Q.err = reinterpret_cast<int *>(flowSync);
      Q.bar = flowSync + 15;
      Q.reg = flowSync + 16;
      Q.spinLimit = 1L << 22;
      // the barrier and XCD-group counters start at zero in every launch
      hipLaunchKernelGGL(k_fill<unsigned>, dim3(1), dim3(kBlock), 0, st, (int64_t)18, flowSync + 15,
                         0u);
      void *args[] = {&Q};
      const void *fn = m.hasNH ? reinterpret_cast<const void *>(&k_flow<T, kMatMixed>)
                               : reinterpret_cast<const void *>(&k_flow<T, kMatStVK>);
      const bool timed = m.prof.on && m.prof.pev[0] && m.prof.pev[1];
      if (timed) MLX_CHECK_HIP(hipEventRecord(m.prof.pev[0], st));
      MLX_CHECK_HIP(hipLaunchCooperativeKernel(fn, dim3(flowG * kXcds), dim3(kBlock), args, 0, st));
      if (timed) {
        MLX_CHECK_HIP(hipEventRecord(m.prof.pev[1], st));
        // one launch is one iteration's sweep when the layer has not stopped
        float ms = 0;
        MLX_CHECK_HIP(hipEventSynchronize(m.prof.pev[1]));
        MLX_CHECK_HIP(hipEventElapsedTime(&ms, m.prof.pev[0], m.prof.pev[1]));
        Ctl c;
        MLX_TRY(read_ctl(m, c));
        if (!c.stopSolve && !c.endLayer) {
          m.prof.pms += ms;
          m.prof.plaunches += 1;
          m.prof.piters += 1;
          m.prof.pelements += nElastic[col];
        }
      }
      return MLX_OK;
    };
    int flowErr = MLX_OK;  // first failure of a flow launch in this layer
    auto enqueue_iteration = [&]() {
      if (flow) {
        const int r = enqueue_flow();
        if (r != MLX_OK && flowErr == MLX_OK) flowErr = r;
      } else if (sweepWg)
        enqueue_sweep_wg<T>(m, P);
      else
        enqueue_sweeps<T>(m, P, fl, perColor.data() + (size_t)m.numColors * col, iterEnqueued++);
      if (m.dim == 3 && L.nbnd > 0 && 12 * (int64_t)L.nbchunks <= kMaxPartials) {
        // the partials are free here: the restriction and the layer's norm
        // consumed theirs before the iterations, the residual tail writes its
        // own after this
        hipLaunchKernelGGL(k_boundary_chunk<T>, dim3(L.nbchunks), dim3(kBlock), 0, st, LL, x,
                           minv, bodies, ri, m.d_lamB.as<double>(), sp.boundaryCompliance, sp.h,
                           ctl, partials, m.d_bodycnt.as<unsigned>());
        hipLaunchKernelGGL(k_boundary_snap<T>, dim3(blocks_for(L.nbnd)), dim3(kBlock), 0, st, LL,
                           x, bodies, ri, ctl);
      } else if (m.dim == 3 && L.nbnd > 0) {
        hipLaunchKernelGGL(k_boundary<T>, dim3(L.nb), dim3(kBlock), 0, st, LL, x, minv, bodies,
                           ri, m.d_lamB.as<double>(), sp.boundaryCompliance, sp.h, ctl);
      }
      if (m.dim == 2 && L.nbnd > 0)
        hipLaunchKernelGGL(k_boundary2d<T>, dim3(L.nb), dim3(kBlock), 0, st, LL, x, minv, bodies,
                           ri, m.d_lamB.as<double>(), sp.boundaryCompliance / (sp.h * sp.h),
                           ctl);
      if (L.nec > 0 && !flow) {
        if (m.dim == 3)
          hipLaunchKernelGGL((k_contacts_elastic<T, 3>), dim3(blocks_for((int64_t)contact_cap(m))), dim3(kBlock), 0,
                             st, LL, ca, x, minv, sp.contactResistance, sp.h, ctl);
        else
          hipLaunchKernelGGL((k_contacts_elastic<T, 2>), dim3(blocks_for((int64_t)contact_cap(m))), dim3(kBlock), 0,
                             st, LL, ca, x, minv, sp.contactResistance, sp.h, ctl);
      }
      if (m.dim == 3 && L.nrc > 0 && sp.useRigidContacts)
        hipLaunchKernelGGL(k_contacts_rigid<T>, dim3(L.nb), dim3(kBlock), 0, st, LL, ca, x, bodies,
                           ri, sp.contactResistance, sp.h, ctl);
      if (m.dim == 2 && L.nrc > 0 && sp.useRigidContacts)
        hipLaunchKernelGGL(k_contacts_rigid2d<T>, dim3(L.nb), dim3(kBlock), 0, st, LL, ca, x,
                           bodies, ri, sp.contactResistance, sp.h, ctl);
      if (nPin > 0 && !flow) {
        if (m.dim == 3)
          hipLaunchKernelGGL((k_pins<T, 3>), dim3(blocks_for(nPin)), dim3(kBlock), 0, st, nPin,
                             m.d_pinned.as<int32_t>(), m.d_pinnedx.as<T>(), x, ctl);
        else
          hipLaunchKernelGGL((k_pins<T, 2>), dim3(blocks_for(nPin)), dim3(kBlock), 0, st, nPin,
                             m.d_pinned.as<int32_t>(), m.d_pinnedx.as<T>(), x, ctl);
      }
      if (sp.computeResiduals) {
        const int nBlk = nTailElemBlocks + (K > 0 ? nResContactBlocks : 0);
        if (m.dim == 3)
          hipLaunchKernelGGL((k_residual_tail<T, 3>), dim3(nBlk), dim3(kBlock), 0, st, P, ld,
                             (const int32_t *)nullptr, nTailElemBlocks, ca, sp.contactResistance,
                             sp.h, d_uniq, d_cnt, partials, m.d_count.as<unsigned>(), ctl, mp,
                             m.d_resid.as<double>(), residCap);
        else
          hipLaunchKernelGGL((k_residual_tail<T, 2>), dim3(nBlk), dim3(kBlock), 0, st, P, ld,
                             (const int32_t *)d_lamSlot2d, nTailElemBlocks, ca,
                             sp.contactResistance, sp.h, d_uniq, d_cnt, partials,
                             m.d_count.as<unsigned>(), ctl, mp, m.d_resid.as<double>(),
                             residCap);
      } else {
        hipLaunchKernelGGL(k_iter_end, dim3(1), dim3(kBlock), 0, st, ctl, mp, partials, 0,
                           m.d_resid.as<double>(), residCap);
      }
    };

    // Iterations run as replays of a captured chunk of G iterations when
    // graphs are on (not while profiling: the timing events are per launch).
    auto graph_for = [&](int G, bool capture) -> hipGraphExec_t {
      auto bits = [](double d) {
        uint64_t u;
        std::memcpy(&u, &d, 8);
        return u;
      };
      std::vector<uint64_t> sig = {
          (uint64_t)col, (uint64_t)layer, (uint64_t)csrCap, (uint64_t)G, (uint64_t)m.sweepBlock,
          (uint64_t)L.nb, (uint64_t)L.nbnd,
          (uint64_t)(L.nec > 0),
          (uint64_t)L.nrc, (uint64_t)(K > 0), (uint64_t)uniqCap, (uint64_t)L.o_rig_vert,
          (uint64_t)L.o_bnd_off, (uint64_t)L.o_bnd_vert, (uint64_t)L.o_rc_off,
          (uint64_t)L.o_rc_idx, (uint64_t)L.o_ec, (uint64_t)(L.nb > 0 ? L.o_chk : 0), (uint64_t)L.nbchunks,
          (uint64_t)sp.computeResiduals,
          (uint64_t)sp.useRigidContacts, (uint64_t)(uintptr_t)fl, (uint64_t)(uintptr_t)m.d_csr.ptr,
          (uint64_t)(uintptr_t)m.d_contacts.ptr, (uint64_t)(uintptr_t)m.d_tmp.ptr,
          (uint64_t)(uintptr_t)m.d_lamB.ptr, (uint64_t)(uintptr_t)m.d_bodies.ptr,
          (uint64_t)(uintptr_t)m.d_ri.ptr, (uint64_t)(uintptr_t)m.d_x.ptr,
          (uint64_t)(uintptr_t)m.d_xold.ptr, (uint64_t)(uintptr_t)m.d_lam.ptr,
          (uint64_t)(uintptr_t)m.d_partials.ptr, (uint64_t)(uintptr_t)m.d_resid.ptr,
          (uint64_t)(uintptr_t)m.d_pinnedx.ptr, (uint64_t)(uintptr_t)m.d_ctl.ptr,
          (uint64_t)(uintptr_t)m.d_count.ptr, (uint64_t)sweepWg, (uint64_t)(uintptr_t)m.d_elist.ptr,
          (uint64_t)(uintptr_t)m.d_vbody.ptr, (uint64_t)(uintptr_t)m.d_bodycnt.ptr, bits(sp.h), bits(sp.boundaryCompliance),
          bits(sp.contactResistance), bits(sp.giveUpThreshold), (uint64_t)sp.hangingStop,
          (uint64_t)sp.giveUpEnabled, (uint64_t)sp.lastLayerGiveUpEnabled};
      if (fl) {
        // colors skipped for having no elastic element (bit per color)
        const int64_t *pc = perColor.data() + (size_t)m.numColors * col;
        for (int c0 = 0; c0 < m.numColors; c0 += 64) {
          uint64_t word = 0;
          for (int c = c0; c < std::min(m.numColors, c0 + 64); ++c)
            if (pc[c] == 0) word |= (uint64_t)1 << (c - c0);
          sig.push_back(word);
        }
      }
      for (auto &g : m.graphs)
        if (g.exec && g.sig == sig) return g.exec;
      if (!capture) return nullptr;
      hipGraph_t graph = nullptr;
      if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
      for (int i = 0; i < G; ++i) enqueue_iteration();
      if (hipStreamEndCapture(st, &graph) != hipSuccess || !graph) return nullptr;
      hipGraphExec_t ex = nullptr;
      const hipError_t ie = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) return nullptr;
      auto &slot = m.graphs[m.graphNext++ % 8];
      if (slot.exec) (void)hipGraphExecDestroy(slot.exec);
      slot.exec = ex;
      slot.sig = sig;
      return ex;
    };
    // Layers without rigid elements: the iterations run inside one
    // cooperative launch (k_persist) per call.
    // (StVK-Voigt meshes only: the persistent loop has no NeoHookean body)
    const bool persist = m.usePersistent && !hasRigidEl && L.nb == 0 && !m.hasNH && !flow;
    auto run_persistent = [&](int n) -> int {
      if (n <= 0) return MLX_OK;
      MLX_CHECK_HIP(hipMemsetAsync(m.d_sync.ptr, 0, 256, st));
      PersistParams<T> Q;
      Q.P = P;
      Q.P.flags = nullptr;
      Q.colorOff = m.d_colors.as<int32_t>();
      Q.colorCnt = m.d_colors.as<int32_t>() + m.numColors;
      Q.numColors = m.numColors;
      Q.lamSlot2d = d_lamSlot2d;
      Q.ca = ca;
      Q.ecIdx = LL.ec_idx;
      Q.nec = L.nec;
      Q.minv = minv;
      Q.resistance = sp.contactResistance;
      Q.h = sp.h;
      Q.uniq = d_uniq;
      Q.ucnt = d_cnt;
      Q.nUniq = nUniq;
      Q.pinDof = m.d_pinned.as<int32_t>();
      Q.pinVal = m.d_pinnedx.as<T>();
      Q.nPin = nPin;
      Q.mp = mp;
      Q.partials = partials;
      Q.resid = m.d_resid.as<double>();
      Q.residCap = residCap;
      Q.bar = m.d_sync.as<unsigned>();
      Q.err = reinterpret_cast<int *>(m.d_sync.as<unsigned>() + 1);
      Q.reg = m.d_sync.as<unsigned>() + 16;
      Q.maxIters = n;
      Q.maxColorCnt = persistMaxCnt;
      void *args[] = {&Q};
      const void *fn = persist_fn<T>(m);
      const bool timed = m.prof.on && m.prof.pev[0] && m.prof.pev[1];
      if (timed) MLX_CHECK_HIP(hipEventRecord(m.prof.pev[0], st));
      MLX_CHECK_HIP(hipLaunchCooperativeKernel(fn, dim3(persistG), dim3(kBlock), args, 0, st));
      if (timed) MLX_CHECK_HIP(hipEventRecord(m.prof.pev[1], st));
      const int before = hc_ctl.iters;
      int32_t sync[2];
      MLX_CHECK_HIP(hipMemcpyAsync(sync, m.d_sync.ptr, 8, hipMemcpyDeviceToHost, st));
      MLX_TRY(read_ctl(m, hc_ctl));
      if (sync[1] != 0) {
        set_error("persistent iteration loop: a workgroup did not reach the grid barrier");
        return MLX_ERR_STATE;
      }
      if (timed) {
        float ms = 0;
        MLX_CHECK_HIP(hipEventElapsedTime(&ms, m.prof.pev[0], m.prof.pev[1]));
        m.prof.pms += ms;
        m.prof.plaunches += 1;
        m.prof.piters += hc_ctl.iters - before;
        m.prof.pelements += (int64_t)(hc_ctl.iters - before) * nElastic[col];
      }
      return MLX_OK;
    };
    // n iterations as replays of captured chunks of 16, 8, 4 and 2
    // iterations (at most MLX_OPT_GRAPH_ITERS), the rest launched directly.
    // Graphs are captured and replayed only for layers without bodies: their
    // launch parameters do not change from substep to substep (contact counts
    // are read on the device), so one capture serves the whole run.  Layers
    // with bodies always launch directly (their lists, body counters and
    // chunk tables change with every substep's layer maps).
    auto run_iterations = [&](int n) -> int {
      if (persist) return run_persistent(n);
      int left = n;
      if (m.useGraphs && !m.prof.on && L.nb == 0 && !flow) {
        for (int G = 16; G >= 2; G /= 2) {
          if (G > m.graphIters) continue;
          while (left >= G) {
            hipGraphExec_t ex = graph_for(G, true);
            if (!ex) break;
            MLX_CHECK_HIP(hipGraphLaunch(ex, st));
            left -= G;
          }
        }
      }
      for (int it = 0; it < left; ++it) enqueue_iteration();
      return MLX_OK;
    };
    // The last layer keeps iterating while the relative stop is not met.
    // When runUntil is already known (set in an earlier layer), the
    // iterations the previous substep needed past the fixed count are
    // enqueued with the fixed ones: kernels return at once after the
    // monitor stops, so a guess that is too long costs only empty launches.
    // In contact the count repeats from substep to substep, so the guess is
    // the previous count itself (one read of the control block when it
    // holds).
    int first = sp.iterations[layer];
    const bool predict = isLast && layer > 0 && hc_ctl.runUntil >= 0 && !persist;
    if (predict) first += m.prevExtraIters;
    MLX_TRY(run_iterations(first));
    MLX_TRY(read_ctl(m, hc_ctl));
    if (m.prof.on) prof_harvest(m, hc_ctl.iters);
    if (isLast) {
      // while it <= iterations || (~stopSolve && runUntil >= 0 && isLastLayer)
      int64_t guard = 0;
      // after a prediction the remaining need is usually one or two
      // iterations: start small (an iteration enqueued past the stop still
      // costs its ~45 early-exit launches, more than a read of the control
      // block) and double
      int chunk = predict ? 1 : 2;
      while (!hc_ctl.stopSolve && !hc_ctl.endLayer && hc_ctl.runUntil >= 0) {
        MLX_TRY(run_iterations(chunk));
        MLX_TRY(read_ctl(m, hc_ctl));
        if (m.prof.on) prof_harvest(m, hc_ctl.iters);
        guard += chunk;
        chunk = persist ? 512 : std::min(chunk * 2, 32);
        if (guard > 10000000) {
          set_error("runUntil loop exceeded 1e7 iterations without converging (hangingStop < 0?)");
          return MLX_ERR_STATE;
        }
      }
    }
    if (flow) {
      if (flowErr != MLX_OK) return flowErr;
      MLX_CHECK_HIP(hipMemcpyAsync(m.h_flowerr, flowSync, sizeof(int32_t), hipMemcpyDeviceToHost, st));
      MLX_CHECK_HIP(hipStreamSynchronize(st));
      if (*m.h_flowerr != 0) {
        set_error("dataflow sweep: a wait or barrier of k_flow did not complete");
        return MLX_ERR_STATE;
      }
    }
    if (m.dim == 2 && L.nb > 0)
      hipLaunchKernelGGL(k_rotate_all2d<T>, dim3(blocks_for(nV)), dim3(kBlock), 0, st, nV, vb, x,
                         bodies, ri);
    if (m.dim == 3 && L.nb > 0) {
      hipLaunchKernelGGL(k_rotate_all<T>, dim3(blocks_for(nV)), dim3(kBlock), 0, st, nV, vb, x,
                         bodies, ri);
      if (hasRigidEl)
        hipLaunchKernelGGL(k_rotate_residual<T>, dim3(blocks_for(nV)), dim3(kBlock), 0, st, nV, vb,
                           rv, bodies);
    }
    MLX_CHECK_HIP(hipGetLastError());
    res.itersPerLayer[layer] = hc_ctl.iters;
    if (isLast) m.prevExtraIters = std::max(0, hc_ctl.iters - sp.iterations[layer]);
    res.projections += (int64_t)hc_ctl.iters * nElastic[col];
    res.layersRun = layer + 1;
    if (hc_ctl.stopSolve) break;
  }
  MLX_TRY(read_ctl(m, hc_ctl));
  res.runUntil = hc_ctl.runUntil;
  res.residualCount = hc_ctl.residualCount;
  const int n = std::min(hc_ctl.residualCount, residCap);
  res.residuals.resize(n);
  MLX_CHECK_HIP(hipMemcpy(res.residuals.data(), m.d_resid.ptr, sizeof(double) * n,
                          hipMemcpyDeviceToHost));
  return MLX_OK;
}

// ---------------------------------------------------------------------------
// uploads of per-call arrays
// ---------------------------------------------------------------------------

template <typename T>
static int upload_positions(mlx_mesh &m, DevBuf &dst, const double *p) {
  std::vector<vec4_t<T>> h(m.nV);
  for (int64_t v = 0; v < m.nV; ++v) {
    vec4_t<T> q;
    q.x = (T)p[m.dim * v];
    q.y = (T)p[m.dim * v + 1];
    q.z = m.dim == 3 ? (T)p[m.dim * v + 2] : T(0);
    q.w = (T)(1.0 / m.massV[v]);
    h[v] = q;
  }
  MLX_TRY(dst.reserve(sizeof(vec4_t<T>) * m.nV));
  MLX_CHECK_HIP(hipMemcpy(dst.ptr, h.data(), sizeof(vec4_t<T>) * m.nV, hipMemcpyHostToDevice));
  return MLX_OK;
}

template <typename T>
static int download_positions(mlx_mesh &m, const DevBuf &src, double *p) {
  std::vector<vec4_t<T>> h(m.nV);
  MLX_CHECK_HIP(hipMemcpy(h.data(), src.ptr, sizeof(vec4_t<T>) * m.nV, hipMemcpyDeviceToHost));
  for (int64_t v = 0; v < m.nV; ++v) {
    p[m.dim * v] = (double)h[v].x;
    p[m.dim * v + 1] = (double)h[v].y;
    if (m.dim == 3) p[m.dim * v + 2] = (double)h[v].z;
  }
  return MLX_OK;
}

// lambdai nT x nc column-major (element order) <-> slots
template <typename T>
static int upload_lambda(mlx_mesh &m, const double *lam) {
  std::vector<T> h((size_t)m.nc * m.ld, T(0));
  if (lam)
    for (int64_t e = 0; e < m.nT; ++e)
      for (int i = 0; i < m.nc; ++i) h[(size_t)i * m.ld + m.slotOf[e]] = (T)lam[e + m.nT * i];
  MLX_CHECK_HIP(hipMemcpy(m.d_lam.ptr, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return MLX_OK;
}

template <typename T>
static int download_lambda(mlx_mesh &m, double *lam) {
  std::vector<T> h((size_t)m.nc * m.ld);
  MLX_CHECK_HIP(hipMemcpy(h.data(), m.d_lam.ptr, sizeof(T) * h.size(), hipMemcpyDeviceToHost));
  for (int64_t e = 0; e < m.nT; ++e)
    for (int i = 0; i < m.nc; ++i) lam[e + m.nT * i] = (double)h[(size_t)i * m.ld + m.slotOf[e]];
  return MLX_OK;
}

template <typename T>
static int upload_scalars(DevBuf &dst, const double *src, size_t n) {
  std::vector<T> h(n);
  for (size_t i = 0; i < n; ++i) h[i] = (T)src[i];
  MLX_TRY(dst.reserve(sizeof(T) * (n > 0 ? n : 1)));
  if (n) MLX_CHECK_HIP(hipMemcpy(dst.ptr, h.data(), sizeof(T) * n, hipMemcpyHostToDevice));
  return MLX_OK;
}

template <typename T>
static int download_scalars(const DevBuf &src, double *dst, size_t n) {
  std::vector<T> h(n);
  if (n) MLX_CHECK_HIP(hipMemcpy(h.data(), src.ptr, sizeof(T) * n, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) dst[i] = (double)h[i];
  return MLX_OK;
}

// ---------------------------------------------------------------------------
// mesh creation
// ---------------------------------------------------------------------------

template <typename T>
static int upload_mesh_arrays(mlx_mesh &m, const mlx_mesh_desc &d) {
  const int64_t ld = m.ld, nT = m.nT;
  const int dd = m.dim * m.dim;
  std::vector<int32_t> vid((size_t)m.vpe * ld, -1);
  std::vector<T> dm((size_t)dd * ld, T(0)), vol(ld, T(0)), mat(4 * (size_t)ld, T(0));
  for (int64_t s = 0; s < ld; ++s) {
    const int32_t e = m.perm[s];
    if (e < 0) continue;
    for (int k = 0; k < m.vpe; ++k) vid[(size_t)k * ld + s] = m.T[(size_t)e * m.vpe + k];
    for (int k = 0; k < dd; ++k) dm[(size_t)k * ld + s] = (T)d.DmInv[(size_t)dd * e + k];
    vol[s] = (T)d.volume[e];
    for (int k = 0; k < 4; ++k) mat[(size_t)k * ld + s] = (T)m.mat[4 * (size_t)e + k];
  }
  MLX_TRY(m.d_vid.reserve(sizeof(int32_t) * vid.size()));
  MLX_TRY(m.d_dminv.reserve(sizeof(T) * dm.size()));
  MLX_TRY(m.d_vol.reserve(sizeof(T) * vol.size()));
  MLX_TRY(m.d_mat.reserve(sizeof(T) * mat.size()));
  MLX_TRY(m.d_lam.reserve(sizeof(T) * m.nc * ld));
  MLX_CHECK_HIP(hipMemcpy(m.d_vid.ptr, vid.data(), sizeof(int32_t) * vid.size(), hipMemcpyHostToDevice));
  MLX_CHECK_HIP(hipMemcpy(m.d_dminv.ptr, dm.data(), sizeof(T) * dm.size(), hipMemcpyHostToDevice));
  MLX_CHECK_HIP(hipMemcpy(m.d_vol.ptr, vol.data(), sizeof(T) * vol.size(), hipMemcpyHostToDevice));
  MLX_CHECK_HIP(hipMemcpy(m.d_mat.ptr, mat.data(), sizeof(T) * mat.size(), hipMemcpyHostToDevice));
  MLX_CHECK_HIP(hipMemset(m.d_lam.ptr, 0, sizeof(T) * m.nc * ld));
  (void)nT;
  return MLX_OK;
}

}  // namespace mlx

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int mlx_init(int32_t device) {
  MLX_TRY(check_device(device));
  g_device = device;
  MLX_CHECK_HIP(hipSetDevice(device));
  double *scr = nullptr;
  MLX_TRY(pcg_scratch_resolve(device, &scr));
  return MLX_OK;
}

}  // extern "C"

// Scene upload.  nScenes > 1 lays slots out scene by scene (then color by
// color, buckets padded to multiples of `pad`), for the batched kernel.
static int mesh_create_impl(const mlx_mesh_desc *d, int32_t precision, int nScenes,
                            const int64_t *sceneElemOff, int pad, mlx_mesh **out) {
  if (!d || !out || (d->dim != 2 && d->dim != 3) || d->nV <= 0 || d->nT <= 0 || !d->t ||
      !d->DmInv || !d->alpha || !d->beta || !d->volume || !d->mass || !d->elementColor ||
      !d->forcedElastic || (d->nPinnedDOFs > 0 && !d->pinnedDOFs) ||
      (precision != MLX_FP64 && precision != MLX_FP32)) {
    set_error("mlx_mesh_create: bad arguments");
    return MLX_ERR_ARG;
  }
  MLX_TRY(check_device(g_device));
  MLX_CHECK_HIP(hipSetDevice(g_device));
  mlx_mesh *m = new mlx_mesh();
  auto fail = [&](int code) {
    mlx_mesh_destroy(m);
    return code;
  };
  m->device = g_device;
  m->dim = d->dim;
  m->vpe = d->dim + 1;
  m->nc = d->dim == 3 ? 6 : 3;
  m->precision = precision;
  m->nV = d->nV;
  m->nT = d->nT;
  const int64_t nV = d->nV, nT = d->nT;
  const int vpe = m->vpe, dim = m->dim, nc = m->nc, dd = dim * dim;
  m->T.resize((size_t)nT * vpe);
  for (int64_t e = 0; e < nT; ++e)
    for (int k = 0; k < vpe; ++k) {
      const int32_t v = d->t[e + nT * k] - 1;
      if (v < 0 || v >= nV) {
        set_error("t(%lld,%d) out of range", (long long)e + 1, k + 1);
        return fail(MLX_ERR_ARG);
      }
      m->T[(size_t)e * vpe + k] = v;
    }
  m->massV.resize(nV);
  for (int64_t v = 0; v < nV; ++v) {
    const double mv = d->mass[dim * v];
    for (int c = 1; c < dim; ++c)
      if (d->mass[dim * v + c] != mv) {
        set_error("vertex %lld has unequal masses on its DOFs", (long long)v + 1);
        return fail(MLX_ERR_UNSUPPORTED);
      }
    if (!(mv > 0)) {
      set_error("vertex %lld has non-positive mass", (long long)v + 1);
      return fail(MLX_ERR_ARG);
    }
    m->massV[v] = mv;
  }
  // material: alpha = inv(K) of the isotropic Voigt form, beta = alpha1 K
  m->mat.resize(4 * (size_t)nT);
  const int nn = dim == 3 ? 3 : 2;
  for (int64_t e = 0; e < nT; ++e) {
    if (d->neohookean && d->neohookean[e]) {
      // NeoHookean tetrahedron: -1/mu, 1/lambda, gamma = 1 + mu/lambda, 0
      // (the negative first scalar tags the slot for the kernels)
      if (dim != 3 || !d->mu || !d->lambda) {
        set_error("NeoHookean elements need a 3D mesh and the mu/lambda arrays");
        return fail(MLX_ERR_UNSUPPORTED);
      }
      const double mu = d->mu[e], la = d->lambda[e];
      if (!(mu > 0) || !(la > 0) || !std::isfinite(mu) || !std::isfinite(la)) {
        set_error("element %lld: NeoHookean mu and lambda must be positive", (long long)e + 1);
        return fail(MLX_ERR_ARG);
      }
      m->mat[4 * e] = -1.0 / mu;
      m->mat[4 * e + 1] = 1.0 / la;
      m->mat[4 * e + 2] = 1.0 + mu / la;
      m->mat[4 * e + 3] = 0.0;
      m->hasNH = true;
      continue;
    }
    const double *A = d->alpha + (size_t)nc * nc * e, *B = d->beta + (size_t)nc * nc * e;
    const double ad = A[0], ao = A[nc], as = A[nn + nc * nn];
    double a1 = 0;
    for (int k = 0; k < nc; ++k) a1 += A[nc * k] * B[k];
    double scale = std::max(std::fabs(ad), std::fabs(as));
    bool ok = std::isfinite(ad) && std::isfinite(as) && scale > 0;
    for (int i = 0; i < nc && ok; ++i)
      for (int j = 0; j < nc && ok; ++j) {
        double expect = 0;
        if (i < nn && j < nn) expect = (i == j) ? ad : ao;
        else if (i == j) expect = as;
        if (std::fabs(A[i + nc * j] - expect) > 1e-9 * scale) ok = false;
        double ab = 0;
        for (int k = 0; k < nc; ++k) ab += A[i + nc * k] * B[k + nc * j];
        if (std::fabs(ab - (i == j ? a1 : 0.0)) > 1e-9 * std::max(std::fabs(a1), 1e-300)) ok = false;
      }
    if (!ok) {
      set_error("element %lld: alpha/beta are not the isotropic StVK-Voigt pair "
                "(alpha = inv(K), beta = alpha1*K) that updateMaterials builds",
                (long long)e + 1);
      return fail(MLX_ERR_UNSUPPORTED);
    }
    m->mat[4 * e] = ad;
    m->mat[4 * e + 1] = ao;
    m->mat[4 * e + 2] = as;
    m->mat[4 * e + 3] = a1;
  }
  m->DmInv.assign(d->DmInv, d->DmInv + (size_t)dd * nT);
  // colors -> slots, each color padded to a multiple of 64 slots
  int32_t nColors = 0;
  m->color.resize(nT);
  for (int64_t e = 0; e < nT; ++e) {
    const int32_t c = d->elementColor[e];
    if (c < 1) {
      set_error("elementColor(%lld) = %d < 1", (long long)e + 1, c);
      return fail(MLX_ERR_ARG);
    }
    m->color[e] = c;
    nColors = std::max(nColors, c);
  }
  // check the coloring is valid: no vertex twice in one color
  {
    std::vector<int64_t> byColor(nColors + 1, 0);
    for (int64_t e = 0; e < nT; ++e) byColor[m->color[e]]++;
    std::vector<int64_t> start(nColors + 2, 0);
    for (int c = 1; c <= nColors; ++c) start[c + 1] = start[c] + byColor[c];
    std::vector<int32_t> order(nT);
    std::vector<int64_t> fill(start);
    for (int64_t e = 0; e < nT; ++e) order[fill[m->color[e]]++] = (int32_t)e;
    std::vector<int32_t> mark((size_t)nV, 0);
    for (int c = 1; c <= nColors; ++c)
      for (int64_t q = start[c]; q < start[c + 1]; ++q) {
        const int32_t e = order[q];
        for (int k = 0; k < vpe; ++k) {
          const int32_t v = m->T[(size_t)e * vpe + k];
          if (mark[v] == c) {
            set_error("elementColor is not a valid coloring: vertex %d appears twice in color %d",
                      v + 1, c);
            return fail(MLX_ERR_ARG);
          }
          mark[v] = c;
        }
      }
    if (nScenes <= 1) {
      m->numColors = nColors;
      m->colorOff.assign(nColors + 1, 0);
      m->colorCnt.assign(nColors, 0);
      int64_t off = 0;
      for (int c = 1; c <= nColors; ++c) {
        m->colorOff[c - 1] = off;
        m->colorCnt[c - 1] = byColor[c];
        off += (byColor[c] + pad - 1) / pad * pad;
      }
      m->colorOff[nColors] = off;
      // colorOff marks padded starts; the sweep range ends at start + count
      m->ld = std::max<int64_t>(off, 64);
      m->perm.assign(m->ld, -1);
      m->slotOf.assign(nT, -1);
      std::vector<int64_t> cur(m->colorOff.begin(), m->colorOff.end());
      for (int c = 1; c <= nColors; ++c)
        for (int64_t q = start[c]; q < start[c + 1]; ++q) {
          const int32_t e = order[q];
          const int64_t s = cur[c - 1]++;
          m->perm[s] = e;
          m->slotOf[e] = (int32_t)s;
        }
      m->nScenes = 1;
      m->sceneColor0.assign(1, 0);
      m->sceneNColors.assign(1, nColors);
      m->sceneSlot0 = {0, m->ld};
    } else {
      // scene-major: per scene, its colors in ascending order
      m->nScenes = nScenes;
      m->sceneColor0.assign(nScenes, 0);
      m->sceneNColors.assign(nScenes, 0);
      m->sceneSlot0.assign(nScenes + 1, 0);
      m->colorOff.clear();
      m->colorCnt.clear();
      int64_t off = 0;
      std::vector<std::pair<int64_t, int64_t>> placed;  // (slot start, element)
      m->slotOf.assign(nT, -1);
      std::vector<int32_t> permv;
      for (int sc = 0; sc < nScenes; ++sc) {
        const int64_t e0 = sceneElemOff[sc], e1 = sceneElemOff[sc + 1];
        int mc = 0;
        for (int64_t e = e0; e < e1; ++e) mc = std::max(mc, m->color[e]);
        std::vector<std::vector<int32_t>> buck(mc + 1);
        for (int64_t e = e0; e < e1; ++e) buck[m->color[e]].push_back((int32_t)e);
        m->sceneColor0[sc] = (int32_t)m->colorCnt.size();
        m->sceneSlot0[sc] = off;
        for (int c = 1; c <= mc; ++c) {
          m->colorOff.push_back(off);
          m->colorCnt.push_back((int64_t)buck[c].size());
          for (size_t q = 0; q < buck[c].size(); ++q) {
            permv.resize(off + q + 1, -1);
            permv[off + q] = buck[c][q];
            m->slotOf[buck[c][q]] = (int32_t)(off + q);
          }
          off += ((int64_t)buck[c].size() + pad - 1) / pad * pad;
        }
        m->sceneNColors[sc] = mc;
      }
      m->sceneSlot0[nScenes] = off;
      m->numColors = (int)m->colorCnt.size();
      m->colorOff.push_back(off);
      m->ld = std::max<int64_t>(off, 64);
      permv.resize(m->ld, -1);
      m->perm = permv;
    }
  }
  m->pinnedDOF.resize(d->nPinnedDOFs);
  for (int64_t q = 0; q < d->nPinnedDOFs; ++q) {
    const int32_t dof = d->pinnedDOFs[q] - 1;
    if (dof < 0 || dof >= dim * nV) {
      set_error("pinnedDOFs(%lld) out of range", (long long)q + 1);
      return fail(MLX_ERR_ARG);
    }
    m->pinnedDOF[q] = dof;
  }
  m->forcedElastic.assign(d->forcedElastic, d->forcedElastic + nT);

  if (hipStreamCreate(&m->stream) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(MLX_ERR_HIP);
  }
  if (hipHostMalloc((void **)&m->h_ctl, sizeof(Ctl), hipHostMallocDefault) != hipSuccess) {
    m->h_ctl = nullptr;
    set_error("hipHostMalloc failed");
    return fail(MLX_ERR_HIP);
  }
  if (hipHostMalloc((void **)&m->h_counts, 4 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
    m->h_counts = nullptr;
    set_error("hipHostMalloc failed");
    return fail(MLX_ERR_HIP);
  }
  int r = precision == MLX_FP64 ? upload_mesh_arrays<double>(*m, *d) : upload_mesh_arrays<float>(*m, *d);
  if (r != MLX_OK) return fail(r);
  // element-order arrays for layer building, per-vertex masses, pins
  if ((r = m->d_T.reserve(sizeof(int32_t) * m->T.size())) != MLX_OK) return fail(r);
  if ((r = m->d_dminv64.reserve(sizeof(double) * m->DmInv.size())) != MLX_OK) return fail(r);
  if ((r = m->d_mass64.reserve(sizeof(double) * nV)) != MLX_OK) return fail(r);
  if ((r = m->d_minv64.reserve(sizeof(double) * nV)) != MLX_OK) return fail(r);
  if ((r = m->d_pinned.reserve(sizeof(int32_t) * (m->pinnedDOF.size() + 1))) != MLX_OK) return fail(r);
  if ((r = m->d_slotOf.reserve(sizeof(int32_t) * nT)) != MLX_OK) return fail(r);
  std::vector<double> minv(nV);
  for (int64_t v = 0; v < nV; ++v) minv[v] = 1.0 / m->massV[v];
  if (hipMemcpy(m->d_T.ptr, m->T.data(), sizeof(int32_t) * m->T.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_dminv64.ptr, m->DmInv.data(), sizeof(double) * m->DmInv.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_mass64.ptr, m->massV.data(), sizeof(double) * nV, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_minv64.ptr, minv.data(), sizeof(double) * nV, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_slotOf.ptr, m->slotOf.data(), sizeof(int32_t) * nT, hipMemcpyHostToDevice) != hipSuccess ||
      (!m->pinnedDOF.empty() &&
       hipMemcpy(m->d_pinned.ptr, m->pinnedDOF.data(), sizeof(int32_t) * m->pinnedDOF.size(),
                 hipMemcpyHostToDevice) != hipSuccess)) {
    set_error("upload of mesh arrays failed");
    return fail(MLX_ERR_HIP);
  }
  const size_t tsz = precision == MLX_FP64 ? 8 : 4;
  if ((r = m->d_x.reserve(4 * tsz * nV)) != MLX_OK) return fail(r);
  if ((r = m->d_xold.reserve(4 * tsz * nV)) != MLX_OK) return fail(r);
  if ((r = m->d_rv.reserve(tsz * 3 * nV)) != MLX_OK) return fail(r);
  if ((r = m->d_pinnedx.reserve(tsz * (m->pinnedDOF.size() + 1))) != MLX_OK) return fail(r);
  *out = m;
  return MLX_OK;
}

namespace mlx {

// The seam's inputs: the caller's arrays go up through pinned staging
// unchanged and are packed on the device (xi, oldp, lambdai, the residual
// velocity in 3D, the pin targets).
template <typename T>
static int seam_upload(mlx_mesh &m, const mlx_layer_solve_args *a, const double *pinv,
                       size_t nPin) {
  const int64_t nV = m.nV, nT = m.nT, ld = m.ld;
  const int dim = m.dim, nc = m.nc;
  hipStream_t st = m.stream;
  const size_t nPos = (size_t)dim * nV, nLam = (size_t)nc * nT, nRv = dim == 3 ? 3 * (size_t)nV : 0;
  MLX_TRY(m.d_seam.reserve(sizeof(double) * (2 * nPos + nLam + nRv + nPin + 8)));
  double *rx = m.d_seam.as<double>(), *ro = rx + nPos, *rl = ro + nPos, *rr = rl + nLam,
         *rp = rr + nRv;
  MLX_TRY(stage_upload(m, rx, a->xi, sizeof(double) * nPos));
  MLX_TRY(stage_upload(m, ro, a->oldp, sizeof(double) * nPos));
  MLX_TRY(stage_upload(m, rl, a->lambdai, sizeof(double) * nLam));
  if (nRv) MLX_TRY(stage_upload(m, rr, a->residualv, sizeof(double) * nRv));
  if (nPin) MLX_TRY(stage_upload(m, rp, pinv, sizeof(double) * nPin));
  MLX_TRY(m.d_x.reserve(sizeof(vec4_t<T>) * nV));
  MLX_TRY(m.d_xold.reserve(sizeof(vec4_t<T>) * nV));
  MLX_TRY(m.d_rv.reserve(sizeof(T) * (nRv > 0 ? nRv : 1)));
  MLX_TRY(m.d_pinnedx.reserve(sizeof(T) * (nPin > 0 ? nPin : 1)));
  const double *minv = m.d_minv64.as<double>();
  hipLaunchKernelGGL(k_pack_positions<T>, dim3(blocks_for(nV)), dim3(kBlock), 0, st, nV, dim,
                     (const double *)rx, minv, m.d_x.as<vec4_t<T>>());
  hipLaunchKernelGGL(k_pack_positions<T>, dim3(blocks_for(nV)), dim3(kBlock), 0, st, nV, dim,
                     (const double *)ro, minv, m.d_xold.as<vec4_t<T>>());
  hipLaunchKernelGGL(k_fill<T>, dim3(blocks_for((int64_t)nc * ld)), dim3(kBlock), 0, st,
                     (int64_t)nc * ld, m.d_lam.as<T>(), T(0));
  hipLaunchKernelGGL(k_pack_lambda<T>, dim3(blocks_for(nT)), dim3(kBlock), 0, st, nT, nc, ld,
                     (const double *)rl, m.d_slotOf.as<int32_t>(), m.d_lam.as<T>());
  if (nRv)
    hipLaunchKernelGGL((k_cast<double, T>), dim3(blocks_for((int64_t)nRv)), dim3(kBlock), 0, st,
                       (int64_t)nRv, (const double *)rr, m.d_rv.as<T>());
  if (nPin)
    hipLaunchKernelGGL((k_cast<double, T>), dim3(blocks_for((int64_t)nPin)), dim3(kBlock), 0, st,
                       (int64_t)nPin, (const double *)rp, m.d_pinnedx.as<T>());
  return MLX_OK;
}

// The seam's outputs: unpacked on the device, read back through pinned
// staging with one synchronization, copied into the caller's arrays.
template <typename T>
static int seam_download(mlx_mesh &m, const mlx_layer_solve_args *a) {
  const int64_t nV = m.nV, nT = m.nT, ld = m.ld;
  const int dim = m.dim, nc = m.nc;
  hipStream_t st = m.stream;
  const size_t nPos = (size_t)dim * nV, nLam = (size_t)nc * nT, nRv = dim == 3 ? 3 * (size_t)nV : 0;
  MLX_TRY(m.d_seam.reserve(sizeof(double) * (nPos + nLam + nRv + 8)));
  double *dx = m.d_seam.as<double>(), *dl = dx + nPos, *dr = dl + nLam;
  hipLaunchKernelGGL(k_unpack_positions<T>, dim3(blocks_for(nV)), dim3(kBlock), 0, st, nV, dim,
                     (const vec4_t<T> *)m.d_x.as<vec4_t<T>>(), dx);
  hipLaunchKernelGGL(k_unpack_lambda<T>, dim3(blocks_for(nT)), dim3(kBlock), 0, st, nT, nc, ld,
                     (const T *)m.d_lam.as<T>(), m.d_slotOf.as<int32_t>(), dl);
  if (nRv)
    hipLaunchKernelGGL((k_cast<T, double>), dim3(blocks_for((int64_t)nRv)), dim3(kBlock), 0, st,
                       (int64_t)nRv, (const T *)m.d_rv.as<T>(), dr);
  char *hp = nullptr;
  MLX_TRY(stage_take(m, sizeof(double) * (nPos + nLam + nRv), &hp));
  MLX_CHECK_HIP(hipMemcpyAsync(hp, dx, sizeof(double) * (nPos + nLam + nRv),
                               hipMemcpyDeviceToHost, st));
  MLX_CHECK_HIP(hipStreamSynchronize(st));
  const double *h = reinterpret_cast<const double *>(hp);
  std::memcpy(a->xi, h, sizeof(double) * nPos);
  std::memcpy(a->lambdai, h + nPos, sizeof(double) * nLam);
  if (nRv) std::memcpy(a->residualv, h + nPos + nLam, sizeof(double) * nRv);
  return MLX_OK;
}

}  // namespace mlx

extern "C" {

int mlx_mesh_create(const mlx_mesh_desc *d, int32_t precision, mlx_mesh **out) {
  return mesh_create_impl(d, precision, 1, nullptr, 64, out);
}

int mlx_mesh_destroy(mlx_mesh *m) {
  if (!m) return MLX_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  mlx::DevBuf *bufs[] = {&m->d_vid, &m->d_dminv, &m->d_vol, &m->d_mat, &m->d_lam,
                         &m->d_dminv64, &m->d_T, &m->d_minv64, &m->d_mass64, &m->d_pinned,
                         &m->d_slotOf, &m->d_x, &m->d_xold, &m->d_rv, &m->d_pinnedx,
                         &m->d_flags, &m->d_vbody, &m->d_ri, &m->d_bodies, &m->d_csr,
                         &m->d_lamB, &m->d_contacts, &m->d_partials, &m->d_ctl, &m->d_resid,
                         &m->d_tmp, &m->d_nh, &m->d_colors, &m->d_sync, &m->d_forced, &m->d_colortab,
                         &m->d_maps, &m->d_scan, &m->d_count, &m->d_elist, &m->d_pinMask,
                         &m->d_bodycnt, &m->d_seam, &m->d_elastic, &m->learn.d_lam0,
                         &m->learn.d_keyScale, &m->learn.d_save, &m->learn.d_dx,
                         &m->learn.d_inc, &m->learn.d_res, &m->learn.d_counts,
                         &m->learn.d_feat, &m->d_finder, &m->d_flowdeps, &m->d_flowflags,
                         &m->d_flowsync};
  for (auto *b : bufs) b->release();
  for (auto e : m->prof.ev) (void)hipEventDestroy(e);
  for (auto e : m->prof.pev)
    if (e) (void)hipEventDestroy(e);
  for (auto &g : m->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (m->h_ctl) (void)hipHostFree(m->h_ctl);
  if (m->h_counts) (void)hipHostFree(m->h_counts);
  if (m->h_flowerr) (void)hipHostFree(m->h_flowerr);
  if (m->h_stage) (void)hipHostFree(m->h_stage);
  for (char *p : m->stageRetired) (void)hipHostFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return MLX_OK;
}

int mlx_mesh_info(const mlx_mesh *m, int32_t *numColors, int64_t *bucketSizes) {
  if (!m || !numColors) {
    set_error("mlx_mesh_info: bad arguments");
    return MLX_ERR_ARG;
  }
  *numColors = m->numColors;
  if (bucketSizes) {
    std::vector<int64_t> cnt(m->numColors, 0);
    for (int64_t e = 0; e < m->nT; ++e) cnt[m->color[e] - 1]++;
    for (int c = 0; c < m->numColors; ++c) bucketSizes[c] = cnt[c];
  }
  return MLX_OK;
}

int mlx_stvk_voigt_constraints(mlx_mesh *m, const uint8_t *isElastic, double h, double *lambdai,
                               double *xi, const double *oldp) {
  if (!m || !isElastic || !lambdai || !xi || !oldp || !(h > 0)) {
    set_error("mlx_stvk_voigt_constraints: bad arguments");
    return MLX_ERR_ARG;
  }
  MLX_CHECK_HIP(hipSetDevice(m->device));
  const bool f64 = m->precision == MLX_FP64;
  std::vector<uint8_t> fl(m->ld, 0);
  for (int64_t s = 0; s < m->ld; ++s)
    if (m->perm[s] >= 0) fl[s] = isElastic[m->perm[s]] ? 1 : 0;
  MLX_TRY(m->d_flags.reserve(m->ld));
  MLX_CHECK_HIP(hipMemcpy(m->d_flags.ptr, fl.data(), m->ld, hipMemcpyHostToDevice));
  if (f64) {
    MLX_TRY(upload_positions<double>(*m, m->d_x, xi));
    MLX_TRY(upload_positions<double>(*m, m->d_xold, oldp));
    MLX_TRY(upload_lambda<double>(*m, lambdai));
    enqueue_sweeps<double>(*m, sweep_params<double>(*m, h, nullptr), m->d_flags.as<uint8_t>());
  } else {
    MLX_TRY(upload_positions<float>(*m, m->d_x, xi));
    MLX_TRY(upload_positions<float>(*m, m->d_xold, oldp));
    MLX_TRY(upload_lambda<float>(*m, lambdai));
    enqueue_sweeps<float>(*m, sweep_params<float>(*m, h, nullptr), m->d_flags.as<uint8_t>());
  }
  MLX_CHECK_HIP(hipGetLastError());
  MLX_CHECK_HIP(hipStreamSynchronize(m->stream));
  if (f64) {
    MLX_TRY(download_positions<double>(*m, m->d_x, xi));
    MLX_TRY(download_lambda<double>(*m, lambdai));
  } else {
    MLX_TRY(download_positions<float>(*m, m->d_x, xi));
    MLX_TRY(download_lambda<float>(*m, lambdai));
  }
  return MLX_OK;
}

int mlx_layer_solve(mlx_mesh *m, const mlx_layer_solve_args *a, mlx_layer_solve_out *out) {
  if (!m || !a || !out || a->nLayers < 1 || a->nLayers > MLX_MAX_LAYERS || a->nSorted < 1 ||
      !a->layersMap || !a->iterations || !a->isElasticElement || !a->vertexRigidBody ||
      !a->isBoundaryVertex || !a->numRigids || !a->xi || !a->lambdai || !a->oldp ||
      !(a->h > 0) || (a->nContacts > 0 && (!a->contactVertexID || !a->contactNormals ||
                                           !a->pointColliders || !a->contactCompliance)) ||
      (m->dim == 3 && (!a->residualv || (!m->pinnedDOF.empty() && !a->pinnedx)))) {
    set_error("mlx_layer_solve: bad arguments");
    return MLX_ERR_ARG;
  }
  MLX_CHECK_HIP(hipSetDevice(m->device));
  const int64_t nV = m->nV, nT = m->nT;
  HostMaps maps;
  maps.nSorted = a->nSorted;
  maps.elastic.assign(a->isElasticElement, a->isElasticElement + (size_t)nT * a->nSorted);
  maps.vbody.assign(a->vertexRigidBody, a->vertexRigidBody + (size_t)nV * a->nSorted);
  maps.boundary.assign(a->isBoundaryVertex, a->isBoundaryVertex + (size_t)nV * a->nSorted);
  maps.numRigids.assign(a->numRigids, a->numRigids + a->nSorted);
  for (size_t i = 0; i < maps.vbody.size(); ++i) {
    const int col = (int)(i / nV);
    if (maps.vbody[i] < 0 || maps.vbody[i] > maps.numRigids[col]) {
      set_error("vertexRigidBody out of 0..numRigids");
      return MLX_ERR_ARG;
    }
  }
  SolveParams sp;
  sp.dim = m->dim;
  sp.nLayers = a->nLayers;
  for (int l = 0; l < a->nLayers; ++l) {
    sp.layersMap0[l] = a->layersMap[l] - 1;
    sp.iterations[l] = a->iterations[l];
    if (sp.layersMap0[l] < 0 || sp.layersMap0[l] >= a->nSorted) {
      set_error("layersMap(%d) out of range", l + 1);
      return MLX_ERR_ARG;
    }
  }
  sp.h = a->h;
  sp.boundaryCompliance = a->boundaryCompliance;
  sp.contactResistance = a->contactResistance;
  sp.computeResiduals = a->computeResiduals;
  sp.hangingStop = a->hangingStop;
  sp.giveUpEnabled = a->giveUpEnabled;
  sp.giveUpThreshold = a->giveUpThreshold;
  sp.lastLayerGiveUpEnabled = a->lastLayerGiveUpEnabled;
  sp.useRigidContacts = a->useRigidContacts;
  sp.stopAtImprovementPercent = a->stopAtImprovementPercent;
  sp.runUntil = a->runUntilResSmallerThan;
  sp.residualLast = a->residualArrayLast;
  if (sp.runUntil >= 0 && !sp.computeResiduals) {
    set_error("runUntilResSmallerThan >= 0 needs computeResiduals (the reference would not terminate)");
    return MLX_ERR_ARG;
  }
  HostContacts hc;
  MLX_TRY(stage_begin(*m));
  MLX_TRY(upload_contacts(*m, m->dim, a->nContacts, a->contactVertexID, a->contactNormals,
                          a->pointColliders, a->contactCompliance, hc));
  const bool f64 = m->precision == MLX_FP64;
  const size_t nPin = m->pinnedDOF.size();
  std::vector<double> pinv(nPin);
  for (size_t q = 0; q < nPin; ++q)
    pinv[q] = m->dim == 3 ? a->pinnedx[q] : a->oldp[m->pinnedDOF[q]];
  SolveResult res;
  int r;
  if (f64) {
    MLX_TRY(seam_upload<double>(*m, a, pinv.data(), nPin));
    r = layer_solve_dev<double>(*m, maps, sp, hc, res);
  } else {
    MLX_TRY(seam_upload<float>(*m, a, pinv.data(), nPin));
    r = layer_solve_dev<float>(*m, maps, sp, hc, res);
  }
  if (r != MLX_OK) return r;
  if (f64) MLX_TRY(seam_download<double>(*m, a));
  else MLX_TRY(seam_download<float>(*m, a));
  out->runUntilResSmallerThan = res.runUntil;
  out->residualCount = res.residualCount;
  if (out->residualArray)
    for (int i = 0; i < std::min<int>(out->residualCap, (int)res.residuals.size()); ++i)
      out->residualArray[i] = res.residuals[i];
  for (int l = 0; l < MLX_MAX_LAYERS; ++l) out->itersPerLayer[l] = res.itersPerLayer[l];
  out->layersRun = res.layersRun;
  out->elasticProjections = res.projections;
  if (out->contactLambdas && a->nContacts > 0) {
    ContactArrays ca = contact_arrays(*m, a->nContacts);
    MLX_CHECK_HIP(hipMemcpy(out->contactLambdas, ca.lambda, sizeof(double) * a->nContacts,
                            hipMemcpyDeviceToHost));
  }
  return MLX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// EDiff natives and the device-resident integrate()
// ---------------------------------------------------------------------------

namespace mlx {

// mexEdiffNorm3D / mexEdiffNorm2D on given F arrays (same expression trees as
// the fused kernels, no FMA contraction).
__global__ void k_ediff_from_f3d(int64_t nT, const double *Fa, const double *Fb, double h,
                                 double *out) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nT) return;
  const double *a = Fa + 9 * i, *b = Fb + 9 * i;
  const double a11 = a[0], a21 = a[1], a31 = a[2], a12 = a[3], a22 = a[4], a32 = a[5],
               a13 = a[6], a23 = a[7], a33 = a[8];
  const double b11 = b[0], b21 = b[1], b31 = b[2], b12 = b[3], b22 = b[4], b32 = b[5],
               b13 = b[6], b23 = b[7], b33 = b[8];
  const double t2 = 1.0 / (h * h);
  double q1 = (a11 * a12) / 2.0 + (a21 * a22) / 2.0 + (a31 * a32) / 2.0 - (b11 * b12) / 2.0 - (b21 * b22) / 2.0 - (b31 * b32) / 2.0;
  double q2 = (a11 * a13) / 2.0 + (a21 * a23) / 2.0 + (a31 * a33) / 2.0 - (b11 * b13) / 2.0 - (b21 * b23) / 2.0 - (b31 * b33) / 2.0;
  double q3 = (a12 * a13) / 2.0 + (a22 * a23) / 2.0 + (a32 * a33) / 2.0 - (b12 * b13) / 2.0 - (b22 * b23) / 2.0 - (b32 * b33) / 2.0;
  double q4 = (a11 * a11) / 2.0 + (a21 * a21) / 2.0 + (a31 * a31) / 2.0 - (b11 * b11) / 2.0 - (b21 * b21) / 2.0 - (b31 * b31) / 2.0;
  double q5 = (a12 * a12) / 2.0 + (a22 * a22) / 2.0 + (a32 * a32) / 2.0 - (b12 * b12) / 2.0 - (b22 * b22) / 2.0 - (b32 * b32) / 2.0;
  double q6 = (a13 * a13) / 2.0 + (a23 * a23) / 2.0 + (a33 * a33) / 2.0 - (b13 * b13) / 2.0 - (b23 * b23) / 2.0 - (b33 * b33) / 2.0;
  out[i] = t2 * (q1 * q1) * 2.0 + t2 * (q2 * q2) * 2.0 + t2 * (q3 * q3) * 2.0 + t2 * (q4 * q4) +
           t2 * (q5 * q5) + t2 * (q6 * q6);
}

__global__ void k_ediff_from_f2d(int64_t nT, const double *Fa, const double *Fb, double h,
                                 double *out) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nT) return;
  const double *A = Fa + 4 * i, *B = Fb + 4 * i;
  const double a = A[0] * A[0] + A[1] * A[1] - B[0] * B[0] - B[1] * B[1];
  const double b = A[0] * A[2] + A[1] * A[3] - B[0] * B[2] - B[1] * B[3];
  const double c = A[2] * A[2] + A[3] * A[3] - B[2] * B[2] - B[3] * B[3];
  out[i] = (0.25 * a * a + 0.5 * b * b + 0.25 * c * c) / h / h;
}

static int ediff_from_f(int dim, const double *Fa, const double *Fb, int64_t nT, double h,
                        double *out) {
  MLX_TRY(check_device(g_device));
  MLX_CHECK_HIP(hipSetDevice(g_device));
  const size_t nf = (size_t)(dim == 3 ? 9 : 4) * nT;
  DevBuf a, b, o;
  int r = a.reserve(8 * nf);
  if (r == MLX_OK) r = b.reserve(8 * nf);
  if (r == MLX_OK) r = o.reserve(8 * (size_t)nT);
  if (r != MLX_OK) {
    a.release(); b.release(); o.release();
    return r;
  }
  hipError_t e = hipMemcpy(a.ptr, Fa, 8 * nf, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b.ptr, Fb, 8 * nf, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    if (dim == 3)
      hipLaunchKernelGGL(k_ediff_from_f3d, dim3(blocks_for(nT)), dim3(kBlock), 0, 0, nT,
                         a.as<double>(), b.as<double>(), h, o.as<double>());
    else
      hipLaunchKernelGGL(k_ediff_from_f2d, dim3(blocks_for(nT)), dim3(kBlock), 0, 0, nT,
                         a.as<double>(), b.as<double>(), h, o.as<double>());
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, o.ptr, 8 * (size_t)nT, hipMemcpyDeviceToHost);
  a.release(); b.release(); o.release();
  if (e != hipSuccess) {
    set_error("mlx_ediff_norm: %s", hipGetErrorString(e));
    return MLX_ERR_HIP;
  }
  return MLX_OK;
}

template <typename T, int DIM>
__global__ void k_gather_pins(int64_t n, const int32_t *dof, const vec4_t<T> *x, T *out) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= n) return;
  const int d = dof[q];
  const T *px = reinterpret_cast<const T *>(&x[d / DIM]);
  out[q] = px[d % DIM];
}

// cache.oldp = p and pinnedx = p(pinned) when frame == 1 or oldp has a NaN
// (xpbdLayer3D.m:629-632).
template <typename T, int DIM>
__global__ void k_reset_oldp(int64_t nV, const int *flag, int force, const vec4_t<T> *x,
                             vec4_t<T> *xold, int64_t nPin, const int32_t *dof, T *pinnedx) {
  if (!force && !*flag) return;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < nV) xold[i] = x[i];
  if (i < nPin) {
    const int d = dof[i];
    const T *px = reinterpret_cast<const T *>(&x[d / DIM]);
    pinnedx[i] = px[d % DIM];
  }
}

// Contact arrays of the plane finder: pointCollider = p - n * distance.
template <typename T, int DIM>
__device__ __forceinline__ void plane_gather_one(int i, const int32_t *vert, const vec4_t<T> *x,
                                                 double n0, double n1, double n2, double D,
                                                 double compliance, double *normal, double *point,
                                                 double *compl_) {
  const vec4_t<T> p = x[vert[i]];
  const double xs[3] = {(double)p.x, (double)p.y, DIM == 3 ? (double)p.z : 0.0};
  const double nn[3] = {n0, n1, DIM == 3 ? n2 : 0.0};
  double dist = xs[0] * n0 + xs[1] * n1;
  if (DIM == 3) dist = dist + xs[2] * n2;
  dist = dist - D;
  for (int d = 0; d < 3; ++d) {
    normal[3 * i + d] = nn[d];
    point[3 * i + d] = d < DIM ? xs[d] - nn[d] * dist : 0.0;
  }
  compl_[i] = compliance;
}

template <typename T, int DIM>
__global__ void k_plane_gather(int K, const int32_t *vert, const vec4_t<T> *x, double n0,
                               double n1, double n2, double D, double compliance,
                               double *normal, double *point, double *compl_) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= K) return;
  plane_gather_one<T, DIM>(i, vert, x, n0, n1, n2, D, compliance, normal, point, compl_);
}

// The same with the contact count read from the device.
template <typename T, int DIM>
__global__ void k_plane_gather_dev(const int32_t *count, const int32_t *vert, const vec4_t<T> *x,
                                   double n0, double n1, double n2, double D, double compliance,
                                   double *normal, double *point, double *compl_) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= count[0]) return;
  plane_gather_one<T, DIM>(i, vert, x, n0, n1, n2, D, compliance, normal, point, compl_);
}

// Contact finders on the device (mlx_contacts.h): parameters and grid
// samples staged to the device, flags of every finder over every vertex,
// one compaction (finder-major codes f nV + v), the contact arrays from the
// codes; the count and the codes are copied to pinned host memory (hList:
// count, then up to nV codes) without a wait of their own.
template <typename T>
static int enqueue_finders(mlx_mesh &m, const mlx_contact_finder *fi, int nf, int32_t **hListOut) {
  const int64_t nV = m.nV;
  hipStream_t s = m.stream;
  std::vector<FinderDev> fd(nf);
  std::vector<size_t> sdfOff(nf, 0);
  size_t sdfTotal = 0;
  for (int f = 0; f < nf; ++f) {
    const mlx_contact_finder &c = fi[f];
    FinderDev &d = fd[f];
    std::memset(&d, 0, sizeof d);
    d.kind = c.kind;
    d.compliance = c.compliance;
    if (c.kind == MLX_CONTACT_PLANE) {
      double nn = 0;
      for (int a = 0; a < 3; ++a) nn += c.normal[a] * c.normal[a];
      nn = std::sqrt(nn);
      if (!(nn > 0)) {
        set_error("contact finder %d: zero plane normal", f);
        return MLX_ERR_ARG;
      }
      double D = 0;
      for (int a = 0; a < 3; ++a) {
        d.n[a] = c.normal[a] / nn;
        D += c.normal[a] * c.position[a];
      }
      d.D = D;
    } else if (c.kind == MLX_CONTACT_BOX) {
      for (int a = 0; a < 3; ++a) {
        d.c[a] = c.center[a];
        d.he[a] = c.halfExtent[a];
      }
      for (int k = 0; k < 9; ++k) d.ax[k] = c.axes[k];
    } else if (c.kind == MLX_CONTACT_SDF) {
      if (!c.sdf || c.dims[0] < 2 || c.dims[1] < 2 || c.dims[2] < 2 || !(c.spacing > 0)) {
        set_error("contact finder %d: an SDF grid needs samples, dims >= 2 and spacing > 0", f);
        return MLX_ERR_ARG;
      }
      for (int a = 0; a < 3; ++a) {
        d.dims[a] = c.dims[a];
        d.origin[a] = c.origin[a];
      }
      d.spacing = c.spacing;
      sdfOff[f] = sdfTotal;
      sdfTotal += (size_t)c.dims[0] * c.dims[1] * c.dims[2];
    } else {
      set_error("contact finder %d: unknown kind %d", f, c.kind);
      return MLX_ERR_ARG;
    }
  }
  const size_t n = (size_t)nf * nV;
  auto al = [](size_t o) { return (o + 255) / 256 * 256; };
  const size_t oPar = 0, oSdf = al(sizeof(FinderDev) * nf), oFlag = al(oSdf + 8 * sdfTotal),
               oPos = al(oFlag + 4 * n), oCode = al(oPos + 4 * n), oLam = al(oCode + 4 * n),
               oEnd = al(oLam + 8 * n);
  MLX_TRY(m.d_finder.reserve(oEnd));
  char *base = m.d_finder.as<char>();
  const double *sdfDev = reinterpret_cast<const double *>(base + oSdf);
  for (int f = 0; f < nf; ++f)
    if (fd[f].kind == MLX_CONTACT_SDF) {
      fd[f].sdf = sdfDev + sdfOff[f];
      MLX_TRY(stage_upload(m, base + oSdf + 8 * sdfOff[f], fi[f].sdf,
                           8 * (size_t)fi[f].dims[0] * fi[f].dims[1] * fi[f].dims[2]));
    }
  MLX_TRY(stage_upload(m, base + oPar, fd.data(), sizeof(FinderDev) * nf));
  int32_t *flag = reinterpret_cast<int32_t *>(base + oFlag);
  int32_t *pos = reinterpret_cast<int32_t *>(base + oPos);
  int32_t *code = reinterpret_cast<int32_t *>(base + oCode);
  double *lamTmp = reinterpret_cast<double *>(base + oLam);
  const FinderDev *fdev = reinterpret_cast<const FinderDev *>(base + oPar);
  vec4_t<T> *x = m.d_x.as<vec4_t<T>>();
  ContactArrays ca = contact_arrays(m, 0);
  int32_t *cnt = const_cast<int32_t *>(ca.counts);
  hipLaunchKernelGGL(k_finder_flags<T>, dim3(blocks_for((int64_t)n)), dim3(kBlock), 0, s, nV,
                     (const vec4_t<T> *)x, fdev, nf, flag);
  MLX_TRY(compact_flags_dev(m, flag, (int64_t)n, pos, code, lamTmp, cnt, s));
  const int cap = (int)contact_cap(m);
  hipLaunchKernelGGL(k_finder_gather<T>, dim3(blocks_for(cap)), dim3(kBlock), 0, s,
                     (const int32_t *)cnt, cap, nV, (const int32_t *)code, (const vec4_t<T> *)x,
                     fdev, const_cast<int32_t *>(ca.vert), const_cast<double *>(ca.normal),
                     const_cast<double *>(ca.point), const_cast<double *>(ca.compliance),
                     ca.lambda);
  char *hp = nullptr;
  MLX_TRY(stage_take(m, 4 * ((size_t)nV + 1), &hp));
  int32_t *hList = reinterpret_cast<int32_t *>(hp);
  MLX_CHECK_HIP(hipMemcpyAsync(hList, cnt, 4, hipMemcpyDeviceToHost, s));
  MLX_CHECK_HIP(hipMemcpyAsync(hList + 1, code, 4 * (size_t)nV, hipMemcpyDeviceToHost, s));
  *hListOut = hList;
  return MLX_OK;
}

// Host view of the finders' contacts after the stream has drained: vertex
// list and per-contact compliance from the codes; a vertex claimed twice is
// refused.
static int finder_contacts_host(const mlx_mesh &m, const mlx_contact_finder *fi, int nf,
                                const int32_t *hList, HostContacts &hc) {
  const int64_t nV = m.nV;
  const int K = hList[0];
  if (K > (int)contact_cap(m)) {
    set_error("contact finders: %d contacts for %lld vertices (a vertex claimed twice)", K,
              (long long)nV);
    return MLX_ERR_UNSUPPORTED;
  }
  hc.K = K;
  hc.vert.resize(K);
  hc.compliance.resize(K);
  std::vector<uint8_t> seen(nV, 0);
  for (int i = 0; i < K; ++i) {
    const int32_t f = (int32_t)(hList[1 + i] / nV), v = (int32_t)(hList[1 + i] % nV);
    if (seen[v]) {
      set_error("contact finders: vertex %d is claimed by two finders", v + 1);
      return MLX_ERR_UNSUPPORTED;
    }
    seen[v] = 1;
    hc.vert[i] = v;
    hc.compliance[i] = fi[f].compliance;
  }
  return MLX_OK;
}

// The layer maps of integrate(): strain-rate order (device EDiff, host
// stable sort), union-find over shared vertices, boundary vertices.
template <typename T>
static int build_layer_maps(mlx_state &S, const double *sorted, int ns, double h, HostMaps &maps) {
  mlx_mesh &m = *S.mesh;
  const int64_t nT = m.nT, nV = m.nV;
  maps.nSorted = ns;
  maps.numRigids.assign(ns, 0);
  maps.elastic.assign((size_t)ns * nT, 1);
  maps.vbody.assign((size_t)ns * nV, 0);
  maps.boundary.assign((size_t)ns * nV, 0);
  bool any = false;
  for (int l = 0; l < ns; ++l) any |= std::floor((sorted[l] / 100.0) * (double)nT) > 0;
  if (!any) return MLX_OK;
  MLX_TRY(S.d_ediff.reserve(sizeof(double) * nT));
  if (m.dim == 3) {
    hipLaunchKernelGGL(k_ediff3d<T>, dim3(blocks_for(nT)), dim3(kBlock), 0, m.stream, nT,
                       m.d_T.as<int32_t>(), m.d_dminv64.as<double>(), m.d_x.as<vec4_t<T>>(),
                       m.d_xold.as<vec4_t<T>>(), h, S.d_ediff.as<double>());
    if (m.learn.model && m.learn.ready && (m.learn.flags & MLX_LEARN_KEYS))
      hipLaunchKernelGGL(k_learn_keys, dim3(blocks_for(nT)), dim3(kBlock), 0, m.stream, nT,
                         (const int32_t *)m.d_slotOf.as<int32_t>(),
                         (const double *)m.learn.d_keyScale.as<double>(), S.d_ediff.as<double>());
  } else
    hipLaunchKernelGGL(k_ediff2d<T>, dim3(blocks_for(nT)), dim3(kBlock), 0, m.stream, nT,
                       m.d_T.as<int32_t>(), m.d_dminv64.as<double>(), m.d_x.as<vec4_t<T>>(),
                       m.d_xold.as<vec4_t<T>>(), h, S.d_ediff.as<double>());
  if (m.deviceMaps)
    return layer_maps_dev(m, S.d_ediff.as<double>(), sorted, ns, maps, nullptr, true);
  S.h_ediff.resize(nT);
  MLX_CHECK_HIP(hipMemcpyAsync(S.h_ediff.data(), S.d_ediff.ptr, sizeof(double) * nT,
                               hipMemcpyDeviceToHost, m.stream));
  MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  std::vector<int32_t> order(nT), sets((size_t)nT * ns);
  stable_order(S.h_ediff.data(), nT, order.data());
  MLX_TRY(union_find_by_vertice(order.data(), nT, sorted, ns, m.forcedElastic.data(), m.T.data(),
                                m.vpe, nV, sets.data(), maps.numRigids.data()));
  std::vector<uint8_t> isRigidVert((size_t)ns * nV);
  MLX_TRY(find_boundary_vertices(m.T.data(), nT, m.vpe, nV, ns, sets.data(), maps.boundary.data(),
                                 isRigidVert.data(), maps.vbody.data()));
  for (size_t i = 0; i < sets.size(); ++i) maps.elastic[i] = sets[i] == 0;
  return MLX_OK;
}

template <typename T>
static int integrate_impl(mlx_state &S, mlx_settings *st, double h, int frame,
                          const mlx_plane *plane, mlx_layer_solve_out *out,
                          const mlx_contact_finder *finders = nullptr, int nFinders = 0) {
  mlx_mesh &m = *S.mesh;
  const int64_t nV = m.nV, nT = m.nT;
  const int dim = m.dim;
  hipStream_t s = m.stream;
  vec4_t<T> *x = m.d_x.as<vec4_t<T>>();
  vec4_t<T> *xo = m.d_xold.as<vec4_t<T>>();
  const int64_t nPin = (int64_t)m.pinnedDOF.size();
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  // staging regions of the previous call are reclaimed while the stream is
  // still idle (every call ends with the stream drained)
  MLX_TRY(stage_begin(m));
  MLX_TRY(m.d_tmp.reserve(4096));
  // 1. frame-1 / NaN reset of cache.oldp and pinnedx (3D)
  if (dim == 3) {
    int *flag = reinterpret_cast<int *>(m.d_tmp.as<char>() + 4000);
    MLX_CHECK_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
    if (frame != 1)
      hipLaunchKernelGGL(k_any_nan<T>, dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, xo, flag);
    hipLaunchKernelGGL((k_reset_oldp<T, 3>), dim3(blocks_for(std::max(nV, nPin))), dim3(kBlock), 0,
                       s, nV, flag, frame == 1, x, xo, nPin, m.d_pinned.as<int32_t>(),
                       m.d_pinnedx.as<T>());
    // the network reads p, cache.oldp and the previous substep's multipliers
    m.learn.ready = false;
    if (m.learn.model) {
      MLX_TRY(learn_infer<T>(m, h, nullptr));
      m.learn.ready = true;
    }
  }
  // 2. contacts at the start-of-step positions: the plane test and the list
  // in ascending vertex order are built on the device; the list is read
  // back together with the layer maps, without a wait of its own
  HostContacts hc;
  double nrm[3] = {0, 0, 0}, Dp = 0;
  const bool planeOn = plane && plane->enabled;
  MLX_TRY(reserve_contacts(m));
  ContactArrays ca = contact_arrays(m, 0);
  int32_t *hList = nullptr;  // pinned: the count, then the list
  if (planeOn) {
    double nn = 0;
    for (int d = 0; d < dim; ++d) nn += plane->normal[d] * plane->normal[d];
    nn = std::sqrt(nn);
    for (int d = 0; d < dim; ++d) {
      nrm[d] = plane->normal[d] / nn;
      Dp += plane->normal[d] * plane->position[d];
    }
    MLX_TRY(S.d_ediff.reserve(sizeof(double) * std::max<int64_t>(nT, nV)));
    int32_t *fl = S.d_ediff.as<int32_t>(), *pos = fl + nV;
    if (dim == 3)
      hipLaunchKernelGGL((k_plane_flags<T, 3>), dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, x,
                         nrm[0], nrm[1], nrm[2], Dp, 0, fl);
    else
      hipLaunchKernelGGL((k_plane_flags<T, 2>), dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, x,
                         nrm[0], nrm[1], nrm[2], Dp, 1, fl);
    int32_t *cnt = const_cast<int32_t *>(ca.counts);
    MLX_TRY(compact_flags_dev(m, fl, nV, pos, const_cast<int32_t *>(ca.vert), ca.lambda, cnt, s));
    if (dim == 3)
      hipLaunchKernelGGL((k_plane_gather_dev<T, 3>), dim3(blocks_for(nV)), dim3(kBlock), 0, s, cnt,
                         ca.vert, x, nrm[0], nrm[1], nrm[2], Dp, plane->compliance,
                         const_cast<double *>(ca.normal), const_cast<double *>(ca.point),
                         const_cast<double *>(ca.compliance));
    else
      hipLaunchKernelGGL((k_plane_gather_dev<T, 2>), dim3(blocks_for(nV)), dim3(kBlock), 0, s, cnt,
                         ca.vert, x, nrm[0], nrm[1], nrm[2], Dp, plane->compliance,
                         const_cast<double *>(ca.normal), const_cast<double *>(ca.point),
                         const_cast<double *>(ca.compliance));
    char *hp = nullptr;
    MLX_TRY(stage_take(m, 4 * ((size_t)nV + 1), &hp));
    hList = reinterpret_cast<int32_t *>(hp);
    MLX_CHECK_HIP(hipMemcpyAsync(hList, cnt, 4, hipMemcpyDeviceToHost, s));
    MLX_CHECK_HIP(hipMemcpyAsync(hList + 1, ca.vert, 4 * (size_t)nV, hipMemcpyDeviceToHost, s));
  } else if (nFinders > 0) {
    MLX_TRY(reserve_contacts(m));
    MLX_TRY(enqueue_finders<T>(m, finders, nFinders, &hList));
  }
  // 3. layers (buildLayers + findBoundaryVertices) from p and cache.oldp
  double sorted[MLX_MAX_LAYERS];
  int32_t map0[MLX_MAX_LAYERS];
  int ns = 0;
  for (int i = 0; i < st->nLayers; ++i) {
    bool dup = false;
    for (int j = 0; j < ns; ++j) dup |= sorted[j] == st->layers[i];
    if (!dup) sorted[ns++] = st->layers[i];
  }
  std::sort(sorted, sorted + ns);
  for (int i = 0; i < st->nLayers; ++i)
    for (int j = 0; j < ns; ++j)
      if (sorted[j] == st->layers[i]) map0[i] = j;
  HostMaps maps;
  MLX_TRY(build_layer_maps<T>(S, sorted, ns, h, maps));
  if (planeOn) {
    // the maps have usually drained the stream already
    MLX_CHECK_HIP(hipStreamSynchronize(s));
    hc.K = hList[0];
    hc.vert.assign(hList + 1, hList + 1 + hc.K);
    hc.compliance.assign(hc.K, plane->compliance);
  } else if (nFinders > 0) {
    MLX_CHECK_HIP(hipStreamSynchronize(s));
    MLX_TRY(finder_contacts_host(m, finders, nFinders, hList, hc));
  }
  const int K = hc.K;
  ca = contact_arrays(m, K);
  S.lastContacts = K;
  const auto t1 = clk::now();
  // 4. shift: cache.oldp = p (3D), prevp = p (2D)
  MLX_CHECK_HIP(hipMemcpyAsync(xo, x, sizeof(vec4_t<T>) * nV, hipMemcpyDeviceToDevice, s));
  // 5. residual velocity (3D) or explicit predictor (2D)
  MLX_TRY(m.d_flags.reserve(std::max<size_t>(m.d_flags.bytes, 1)));
  if (!m.d_pinMask.ptr) {
    std::vector<uint8_t> pinMask((size_t)dim * nV, 0);
    for (auto d : m.pinnedDOF) pinMask[d] = 1;
    MLX_TRY(m.d_pinMask.reserve(std::max<size_t>(pinMask.size(), 1)));
    MLX_TRY(stage_upload(m, m.d_pinMask.ptr, pinMask.data(), pinMask.size()));
  }
  DevBuf &pm = m.d_pinMask;
  if (dim == 3) {
    hipLaunchKernelGGL(k_residual_velocity3d<T>, dim3(blocks_for(3 * nV)), dim3(kBlock), 0, s, nV,
                       S.d_v.as<T>(), m.d_rv.as<T>(), m.d_mass64.as<double>(), pm.as<uint8_t>(),
                       h, st->gravity, st->airDrag);
  } else {
    hipLaunchKernelGGL(k_predict2d<T>, dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, S.d_v.as<T>(),
                       x, m.d_mass64.as<double>(), pm.as<uint8_t>(), h, st->gravity);
    // 2D pins snap to prevp
    if (nPin > 0)
      hipLaunchKernelGGL((k_gather_pins<T, 2>), dim3(blocks_for(nPin)), dim3(kBlock), 0, s, nPin,
                         m.d_pinned.as<int32_t>(), xo, m.d_pinnedx.as<T>());
  }
  MLX_CHECK_HIP(hipMemsetAsync(m.d_lam.ptr, 0, sizeof(T) * m.nc * m.ld, s));
  // 6. layerSolve
  SolveParams sp;
  sp.dim = dim;
  sp.nLayers = st->nLayers;
  for (int l = 0; l < st->nLayers; ++l) {
    sp.layersMap0[l] = map0[l];
    sp.iterations[l] = st->iterations[l];
  }
  sp.h = h;
  sp.boundaryCompliance = st->boundaryCompliance;
  sp.contactResistance = st->contactResistance;
  sp.computeResiduals = st->computeResiduals;
  sp.hangingStop = st->hangingStop;
  sp.giveUpEnabled = st->giveUpEnabled;
  sp.giveUpThreshold = st->giveUpThreshold;
  sp.lastLayerGiveUpEnabled = st->lastLayerGiveUpEnabled;
  sp.useRigidContacts = st->useRigidContacts;
  sp.stopAtImprovementPercent = st->stopAtImprovementPercent;
  sp.runUntil = st->runUntilResSmallerThan;
  sp.residualLast = 0;
  if (sp.runUntil >= 0 && !sp.computeResiduals) {
    set_error("runUntilResSmallerThan >= 0 needs computeResiduals");
    return MLX_ERR_ARG;
  }
  SolveResult res;
  const auto t2 = clk::now();
  MLX_TRY(layer_solve_dev<T>(m, maps, sp, hc, res));
  const auto t3 = clk::now();
  st->runUntilResSmallerThan = res.runUntil;
  // 7. velocity update and contact velocity update
  if (dim == 3)
    hipLaunchKernelGGL((k_velocity<T, 3>), dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, x, xo,
                       S.d_v.as<T>(), h);
  else
    hipLaunchKernelGGL((k_velocity<T, 2>), dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, x, xo,
                       S.d_v.as<T>(), h);
  if (K > 0) {
    ca = contact_arrays(m, K);
    if (dim == 3)
      hipLaunchKernelGGL((k_contact_velocity<T, 3>), dim3(blocks_for(K)), dim3(kBlock), 0, s, ca,
                         S.d_v.as<T>(), h, st->gravity, st->restitution);
    else
      hipLaunchKernelGGL((k_contact_velocity<T, 2>), dim3(blocks_for(K)), dim3(kBlock), 0, s, ca,
                         S.d_v.as<T>(), h, st->gravity, st->restitution);
  }
  MLX_CHECK_HIP(hipGetLastError());
  MLX_CHECK_HIP(hipStreamSynchronize(s));
  m.learn.ready = false;
  {
    auto ms = [](clk::time_point a, clk::time_point b) {
      return std::chrono::duration<double, std::milli>(b - a).count();
    };
    m.phase.layersMs += ms(t0, t1);
    m.phase.solveMs += ms(t2, t3);
    m.phase.totalMs += ms(t0, clk::now());
    m.phase.substeps += 1;
  }
  if (out) {
    out->runUntilResSmallerThan = res.runUntil;
    out->residualCount = res.residualCount;
    if (out->residualArray)
      for (int i = 0; i < std::min<int>(out->residualCap, (int)res.residuals.size()); ++i)
        out->residualArray[i] = res.residuals[i];
    for (int l = 0; l < MLX_MAX_LAYERS; ++l) out->itersPerLayer[l] = res.itersPerLayer[l];
    out->layersRun = res.layersRun;
    out->elasticProjections = res.projections;
    if (out->contactLambdas && K > 0)
      MLX_CHECK_HIP(hipMemcpy(out->contactLambdas, ca.lambda, sizeof(double) * K,
                              hipMemcpyDeviceToHost));
  }
  return MLX_OK;
}

}  // namespace mlx

extern "C" {

int mlx_ediff_norm3d(const double *Fa, const double *Fb, int64_t nT, double h, double *out) {
  if (!Fa || !Fb || !out || nT < 0 || !(h > 0)) {
    set_error("mlx_ediff_norm3d: bad arguments");
    return MLX_ERR_ARG;
  }
  if (nT == 0) return MLX_OK;
  return ediff_from_f(3, Fa, Fb, nT, h, out);
}

int mlx_ediff_norm2d(const double *Fa, const double *Fb, int64_t nT, double h, double *out) {
  if (!Fa || !Fb || !out || nT < 0 || !(h > 0)) {
    set_error("mlx_ediff_norm2d: bad arguments");
    return MLX_ERR_ARG;
  }
  if (nT == 0) return MLX_OK;
  return ediff_from_f(2, Fa, Fb, nT, h, out);
}

int mlx_state_create(mlx_mesh *m, const double *p, const double *v, mlx_state **out) {
  if (!m || !p || !v || !out) {
    set_error("mlx_state_create: bad arguments");
    return MLX_ERR_ARG;
  }
  MLX_CHECK_HIP(hipSetDevice(m->device));
  mlx_state *S = new mlx_state();
  S->mesh = m;
  int r;
  const size_t n = (size_t)m->dim * m->nV;
  if (m->precision == MLX_FP64) {
    r = upload_positions<double>(*m, m->d_x, p);
    if (r == MLX_OK) r = upload_positions<double>(*m, m->d_xold, p);
    if (r == MLX_OK) r = upload_scalars<double>(S->d_v, v, n);
  } else {
    r = upload_positions<float>(*m, m->d_x, p);
    if (r == MLX_OK) r = upload_positions<float>(*m, m->d_xold, p);
    if (r == MLX_OK) r = upload_scalars<float>(S->d_v, v, n);
  }
  if (r == MLX_OK) {
    const size_t tsz = m->precision == MLX_FP64 ? 8 : 4;
    r = m->d_rv.reserve(tsz * 3 * m->nV);
    if (r == MLX_OK && hipMemset(m->d_rv.ptr, 0, tsz * 3 * m->nV) != hipSuccess) r = MLX_ERR_HIP;
  }
  if (r != MLX_OK) {
    mlx_state_destroy(S);
    return r;
  }
  *out = S;
  return MLX_OK;
}

int mlx_state_destroy(mlx_state *S) {
  if (!S) return MLX_OK;
  S->d_v.release();
  S->d_ediff.release();
  delete S;
  return MLX_OK;
}

int mlx_state_download(mlx_state *S, double *p, double *v, double *lambdai, int32_t *nContacts) {
  if (!S) {
    set_error("mlx_state_download: null state");
    return MLX_ERR_ARG;
  }
  mlx_mesh &m = *S->mesh;
  MLX_CHECK_HIP(hipSetDevice(m.device));
  MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  const size_t n = (size_t)m.dim * m.nV;
  if (m.precision == MLX_FP64) {
    if (p) MLX_TRY(download_positions<double>(m, m.d_x, p));
    if (v) MLX_TRY(download_scalars<double>(S->d_v, v, n));
    if (lambdai) MLX_TRY(download_lambda<double>(m, lambdai));
  } else {
    if (p) MLX_TRY(download_positions<float>(m, m.d_x, p));
    if (v) MLX_TRY(download_scalars<float>(S->d_v, v, n));
    if (lambdai) MLX_TRY(download_lambda<float>(m, lambdai));
  }
  if (nContacts) *nContacts = S->lastContacts;
  return MLX_OK;
}

int mlx_state_set(mlx_state *S, const double *p, const double *v, const double *oldp,
                  const double *lambdai) {
  if (!S) {
    set_error("mlx_state_set: null state");
    return MLX_ERR_ARG;
  }
  mlx_mesh &m = *S->mesh;
  MLX_CHECK_HIP(hipSetDevice(m.device));
  MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  const size_t n = (size_t)m.dim * m.nV;
  if (m.precision == MLX_FP64) {
    if (p) MLX_TRY(upload_positions<double>(m, m.d_x, p));
    if (oldp) MLX_TRY(upload_positions<double>(m, m.d_xold, oldp));
    if (v) MLX_TRY(upload_scalars<double>(S->d_v, v, n));
    if (lambdai) MLX_TRY(upload_lambda<double>(m, lambdai));
  } else {
    if (p) MLX_TRY(upload_positions<float>(m, m.d_x, p));
    if (oldp) MLX_TRY(upload_positions<float>(m, m.d_xold, oldp));
    if (v) MLX_TRY(upload_scalars<float>(S->d_v, v, n));
    if (lambdai) MLX_TRY(upload_lambda<float>(m, lambdai));
  }
  MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  return MLX_OK;
}

int mlx_state_get_oldp(mlx_state *S, double *oldp) {
  if (!S || !oldp) {
    set_error("mlx_state_get_oldp: bad arguments");
    return MLX_ERR_ARG;
  }
  mlx_mesh &m = *S->mesh;
  MLX_CHECK_HIP(hipSetDevice(m.device));
  MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  if (m.precision == MLX_FP64) MLX_TRY(download_positions<double>(m, m.d_xold, oldp));
  else MLX_TRY(download_positions<float>(m, m.d_xold, oldp));
  return MLX_OK;
}

int mlx_state_set_pinned_targets(mlx_state *S, const double *pinnedx) {
  if (!S || !pinnedx) {
    set_error("mlx_state_set_pinned_targets: bad arguments");
    return MLX_ERR_ARG;
  }
  mlx_mesh &m = *S->mesh;
  if (m.dim != 3) {
    set_error("mlx_state_set_pinned_targets: 3D meshes only (2D pins follow oldp)");
    return MLX_ERR_UNSUPPORTED;
  }
  const int64_t nPin = (int64_t)m.pinnedDOF.size();
  if (nPin == 0) return MLX_OK;
  MLX_CHECK_HIP(hipSetDevice(m.device));
  MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  if (m.precision == MLX_FP64) MLX_TRY(upload_scalars<double>(m.d_pinnedx, pinnedx, nPin));
  else MLX_TRY(upload_scalars<float>(m.d_pinnedx, pinnedx, nPin));
  return MLX_OK;
}

int mlx_integrate(mlx_state *S, mlx_settings *st, double h, int32_t frame, const mlx_plane *plane,
                  mlx_layer_solve_out *out) {
  if (!S || !st || !(h > 0) || frame < 1 || st->nLayers < 1 || st->nLayers > MLX_MAX_LAYERS) {
    set_error("mlx_integrate: bad arguments");
    return MLX_ERR_ARG;
  }
  MLX_CHECK_HIP(hipSetDevice(S->mesh->device));
  if (S->mesh->precision == MLX_FP64) return integrate_impl<double>(*S, st, h, frame, plane, out);
  return integrate_impl<float>(*S, st, h, frame, plane, out);
}

}  // extern "C"

extern "C" {

int mlx_integrate_finders(mlx_state *S, mlx_settings *st, double h, int32_t frame,
                          const mlx_contact_finder *finders, int32_t nFinders,
                          mlx_layer_solve_out *out) {
  if (!S || !st || !(h > 0) || frame < 1 || st->nLayers < 1 || st->nLayers > MLX_MAX_LAYERS ||
      nFinders < 0 || nFinders > MLX_MAX_FINDERS || (nFinders > 0 && !finders)) {
    set_error("mlx_integrate_finders: bad arguments");
    return MLX_ERR_ARG;
  }
  if (S->mesh->dim != 3 || S->mesh->nScenes != 1) {
    set_error("mlx_integrate_finders: 3D single-scene meshes");
    return MLX_ERR_UNSUPPORTED;
  }
  MLX_CHECK_HIP(hipSetDevice(S->mesh->device));
  if (S->mesh->precision == MLX_FP64)
    return integrate_impl<double>(*S, st, h, frame, nullptr, out, finders, nFinders);
  return integrate_impl<float>(*S, st, h, frame, nullptr, out, finders, nFinders);
}

int mlx_find_contacts(mlx_state *S, const mlx_contact_finder *finders, int32_t nFinders,
                      int32_t *K, int32_t cap, int32_t *vertex, double *normals, double *points,
                      double *compliance) {
  if (!S || !K || nFinders < 1 || nFinders > MLX_MAX_FINDERS || !finders || cap < 0) {
    set_error("mlx_find_contacts: bad arguments");
    return MLX_ERR_ARG;
  }
  mlx_mesh &m = *S->mesh;
  if (m.dim != 3) {
    set_error("mlx_find_contacts: 3D meshes");
    return MLX_ERR_UNSUPPORTED;
  }
  MLX_CHECK_HIP(hipSetDevice(m.device));
  MLX_TRY(stage_begin(m));
  MLX_TRY(reserve_contacts(m));
  int32_t *hList = nullptr;
  const int r = m.precision == MLX_FP64 ? enqueue_finders<double>(m, finders, nFinders, &hList)
                                        : enqueue_finders<float>(m, finders, nFinders, &hList);
  if (r != MLX_OK) return r;
  MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  HostContacts hc;
  MLX_TRY(finder_contacts_host(m, finders, nFinders, hList, hc));
  *K = hc.K;
  const int n = std::min(hc.K, cap);
  ContactArrays ca = contact_arrays(m, hc.K);
  if (vertex)
    for (int i = 0; i < n; ++i) vertex[i] = hc.vert[i];
  if (normals && n > 0)
    MLX_CHECK_HIP(hipMemcpy(normals, ca.normal, 24 * (size_t)n, hipMemcpyDeviceToHost));
  if (points && n > 0)
    MLX_CHECK_HIP(hipMemcpy(points, ca.point, 24 * (size_t)n, hipMemcpyDeviceToHost));
  if (compliance)
    for (int i = 0; i < n; ++i) compliance[i] = hc.compliance[i];
  return MLX_OK;
}

int mlx_profile_enable(mlx_mesh *m, int32_t on) {
  if (!m) {
    set_error("mlx_profile_enable: null mesh");
    return MLX_ERR_ARG;
  }
  m->prof.on = on != 0;
  m->prof.used = 0;
  return MLX_OK;
}

int mlx_profile_phases(mlx_mesh *m, double *layersMs, double *solveMs, double *totalMs,
                       int64_t *substeps) {
  if (!m) {
    set_error("mlx_profile_phases: null mesh");
    return MLX_ERR_ARG;
  }
  if (layersMs) *layersMs = m->phase.layersMs;
  if (solveMs) *solveMs = m->phase.solveMs;
  if (totalMs) *totalMs = m->phase.totalMs;
  if (substeps) *substeps = m->phase.substeps;
  m->phase = mlx_mesh::Phase();
  return MLX_OK;
}

int mlx_layer_maps(mlx_mesh *m, const double *keys, const double *sortedLayers, int32_t nSorted,
                   int32_t *sets, int32_t *numRigids, uint8_t *isBoundary,
                   int32_t *vertexRigidBody) {
  if (!m || (!keys && m->nT > 0) || !sortedLayers || nSorted < 1 || nSorted > MLX_MAX_LAYERS ||
      !sets || !numRigids || !isBoundary || !vertexRigidBody) {
    set_error("mlx_layer_maps: bad arguments");
    return MLX_ERR_ARG;
  }
  for (int l = 0; l < nSorted; ++l)
    if (!(sortedLayers[l] >= 0 && sortedLayers[l] <= 100) || (l > 0 && !(sortedLayers[l] > sortedLayers[l - 1]))) {
      set_error("mlx_layer_maps: sortedLayers must be ascending percentages");
      return MLX_ERR_ARG;
    }
  MLX_CHECK_HIP(hipSetDevice(m->device));
  DevBuf dk;
  MLX_TRY(dk.reserve(sizeof(double) * std::max<int64_t>(m->nT, 1)));
  MLX_CHECK_HIP(hipMemcpy(dk.ptr, keys, sizeof(double) * m->nT, hipMemcpyHostToDevice));
  HostMaps maps;
  std::vector<int32_t> hs;
  MLX_TRY(stage_begin(*m));
  const int r = layer_maps_dev(*m, dk.as<double>(), sortedLayers, nSorted, maps, &hs);
  dk.release();
  if (r != MLX_OK) return r;
  std::copy(hs.begin(), hs.end(), sets);
  std::copy(maps.numRigids.begin(), maps.numRigids.end(), numRigids);
  std::copy(maps.boundary.begin(), maps.boundary.end(), isBoundary);
  std::copy(maps.vbody.begin(), maps.vbody.end(), vertexRigidBody);
  return MLX_OK;
}

int mlx_profile_read_persistent(mlx_mesh *m, double *ms, int64_t *launches, int64_t *elements,
                                int64_t *iterations) {
  if (!m) {
    set_error("mlx_profile_read_persistent: null mesh");
    return MLX_ERR_ARG;
  }
  if (ms) *ms = m->prof.pms;
  if (launches) *launches = m->prof.plaunches;
  if (elements) *elements = m->prof.pelements;
  if (iterations) *iterations = m->prof.piters;
  m->prof.pms = 0;
  m->prof.plaunches = m->prof.pelements = m->prof.piters = 0;
  return MLX_OK;
}

int mlx_profile_read(mlx_mesh *m, double *ms, int64_t *launches, int64_t *elements) {
  if (!m) {
    set_error("mlx_profile_read: null mesh");
    return MLX_ERR_ARG;
  }
  if (ms) *ms = m->prof.ms;
  if (launches) *launches = m->prof.launches;
  if (elements) *elements = m->prof.elements;
  m->prof.ms = 0;
  m->prof.launches = 0;
  m->prof.elements = 0;
  return MLX_OK;
}

int mlx_mesh_set_option(mlx_mesh *m, int32_t option, int64_t value) {
  if (!m) {
    set_error("mlx_mesh_set_option: null mesh");
    return MLX_ERR_ARG;
  }
  switch (option) {
    case MLX_OPT_GRAPHS:
      m->useGraphs = value != 0;
      return MLX_OK;
    case MLX_OPT_GRAPH_ITERS:
      if (value < 1 || value > 256) break;
      m->graphIters = (int)value;
      return MLX_OK;
    case MLX_OPT_PERSISTENT:
      m->usePersistent = value != 0;
      return MLX_OK;
    case MLX_OPT_PERSIST_WG:
      if (value < 0 || value > 256) break;
      m->persistWG = (int)value;
      return MLX_OK;
    case MLX_OPT_DEVICE_MAPS:
      m->deviceMaps = value != 0;
      return MLX_OK;
    case MLX_OPT_PERSIST_LOCAL:
      m->persistLocal = value != 0;
      return MLX_OK;
    case MLX_OPT_FLOW:
      m->useFlow = value != 0;
      return MLX_OK;
    case MLX_OPT_SWEEP_WG:
      m->useSweepWg = value != 0;
      return MLX_OK;
    case MLX_OPT_SWEEP_BLOCK:
      if (value != 64 && value != 128 && value != 256) break;
      m->sweepBlock = (int)value;
      return MLX_OK;
    default:
      break;
  }
  set_error("mlx_mesh_set_option: bad option %d or value %lld", option, (long long)value);
  return MLX_ERR_ARG;
}

// Attaches a network to a mesh (single scene or a batch's merged mesh):
// the warm start's buffers and the vertex incidences.
static int learn_attach(mlx_mesh *m, mlx_mlp *model, int32_t flags, float kappaMax) {
  mlx_mesh::Learn &L = m->learn;
  if (!model) {
    L.model = nullptr;
    L.flags = 0;
    L.ready = false;
    return MLX_OK;
  }
  if (m->dim != 3) {
    set_error("mlx_mesh_set_learned: needs a 3D mesh");
    return MLX_ERR_UNSUPPORTED;
  }
  if ((flags & ~(MLX_LEARN_WARMSTART | MLX_LEARN_KEYS)) != 0 || !(kappaMax > 0) ||
      !std::isfinite(kappaMax)) {
    set_error("mlx_mesh_set_learned: bad flags %d or kappaMax %g", flags, (double)kappaMax);
    return MLX_ERR_ARG;
  }
  if (mlp_device_of(model) != m->device) {
    set_error("mlx_mesh_set_learned: model on device %d, mesh on device %d",
              mlp_device_of(model), m->device);
    return MLX_ERR_ARG;
  }
  MLX_CHECK_HIP(hipSetDevice(m->device));
  const int64_t ld = m->ld, nV = m->nV, nT = m->nT;
  const size_t tb = m->precision == MLX_FP64 ? sizeof(double) : sizeof(float);
  const size_t nLam = (size_t)m->nc * ld;
  MLX_TRY(L.d_lam0.reserve(tb * nLam));
  MLX_TRY(L.d_keyScale.reserve(sizeof(double) * ld));
  MLX_TRY(L.d_save.reserve(4 * tb * nV + tb * nLam));
  MLX_TRY(L.d_dx.reserve(tb * 12 * ld));
  MLX_TRY(L.d_res.reserve(sizeof(double) * (2 + kMaxPartials)));
  if (!L.d_counts.ptr) {
    MLX_TRY(L.d_counts.reserve(16));
    MLX_CHECK_HIP(hipMemsetAsync(L.d_counts.ptr, 0, 16, m->stream));
  }
  // incidences (vertex a of slot s, as a ld + s) grouped by vertex, each
  // vertex's in ascending slot order
  std::vector<int32_t> off(nV + 1, 0), inc(4 * (size_t)nT);
  for (int64_t s = 0; s < ld; ++s) {
    const int32_t e = m->perm[s];
    if (e < 0) continue;
    for (int a = 0; a < 4; ++a) off[m->T[4 * (size_t)e + a] + 1] += 1;
  }
  for (int64_t v = 0; v < nV; ++v) off[v + 1] += off[v];
  std::vector<int32_t> fill(off.begin(), off.end() - 1);
  for (int64_t s = 0; s < ld; ++s) {
    const int32_t e = m->perm[s];
    if (e < 0) continue;
    for (int a = 0; a < 4; ++a) inc[fill[m->T[4 * (size_t)e + a]]++] = (int32_t)(a * ld + s);
  }
  if (4 * ld > INT32_MAX) {
    set_error("mlx_mesh_set_learned: slot count too large");
    return MLX_ERR_UNSUPPORTED;
  }
  MLX_TRY(L.d_inc.reserve(sizeof(int32_t) * (off.size() + inc.size())));
  MLX_CHECK_HIP(hipMemcpyAsync(L.d_inc.ptr, off.data(), sizeof(int32_t) * off.size(),
                               hipMemcpyHostToDevice, m->stream));
  MLX_CHECK_HIP(hipMemcpyAsync(L.d_inc.as<int32_t>() + off.size(), inc.data(),
                               sizeof(int32_t) * inc.size(), hipMemcpyHostToDevice, m->stream));
  MLX_CHECK_HIP(hipStreamSynchronize(m->stream));
  L.model = model;
  L.flags = flags;
  L.kappaMax = kappaMax;
  L.ready = false;
  return MLX_OK;
}

int mlx_mesh_set_learned(mlx_mesh *m, mlx_mlp *model, int32_t flags, float kappaMax) {
  if (!m) {
    set_error("mlx_mesh_set_learned: null mesh");
    return MLX_ERR_ARG;
  }
  if (model && (m->dim != 3 || m->nScenes != 1)) {
    set_error("mlx_mesh_set_learned: needs a 3D single-scene mesh (batches: "
              "mlx_batch_set_learned)");
    return MLX_ERR_UNSUPPORTED;
  }
  return learn_attach(m, model, flags, kappaMax);
}

int mlx_learn_stats(mlx_mesh *m, int64_t *kept, int64_t *rejected) {
  if (!m || !m->learn.d_counts.ptr) {
    set_error("mlx_learn_stats: no network attached");
    return MLX_ERR_STATE;
  }
  MLX_CHECK_HIP(hipSetDevice(m->device));
  unsigned long long c[2];
  MLX_CHECK_HIP(hipMemcpyAsync(c, m->learn.d_counts.ptr, 16, hipMemcpyDeviceToHost, m->stream));
  MLX_CHECK_HIP(hipMemsetAsync(m->learn.d_counts.ptr, 0, 16, m->stream));
  MLX_CHECK_HIP(hipStreamSynchronize(m->stream));
  if (kept) *kept = (int64_t)c[0];
  if (rejected) *rejected = (int64_t)c[1];
  return MLX_OK;
}

int mlx_learn_outputs(mlx_mesh *m, double *lambda0, double *keyScale) {
  if (!m || !m->learn.d_lam0.ptr) {
    set_error("mlx_learn_outputs: no network attached");
    return MLX_ERR_STATE;
  }
  MLX_CHECK_HIP(hipSetDevice(m->device));
  const int64_t ld = m->ld, nT = m->nT;
  if (lambda0) {
    std::vector<double> lam(6 * (size_t)ld);
    if (m->precision == MLX_FP64) {
      MLX_CHECK_HIP(hipMemcpy(lam.data(), m->learn.d_lam0.ptr, sizeof(double) * 6 * ld,
                              hipMemcpyDeviceToHost));
    } else {
      std::vector<float> lf(6 * (size_t)ld);
      MLX_CHECK_HIP(hipMemcpy(lf.data(), m->learn.d_lam0.ptr, sizeof(float) * 6 * ld,
                              hipMemcpyDeviceToHost));
      for (size_t i = 0; i < lf.size(); ++i) lam[i] = lf[i];
    }
    for (int64_t e = 0; e < nT; ++e)
      for (int i = 0; i < 6; ++i) lambda0[6 * e + i] = lam[(size_t)i * ld + m->slotOf[e]];
  }
  if (keyScale) {
    std::vector<double> ks(ld);
    MLX_CHECK_HIP(hipMemcpy(ks.data(), m->learn.d_keyScale.ptr, sizeof(double) * ld,
                            hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < nT; ++e) keyScale[e] = ks[m->slotOf[e]];
  }
  return MLX_OK;
}

int mlx_learn_features(mlx_state *S, double h, float *features) {
  if (!S || !S->mesh || !features || !(h > 0)) {
    set_error("mlx_learn_features: bad arguments");
    return MLX_ERR_ARG;
  }
  mlx_mesh &m = *S->mesh;
  if (!m.learn.model) {
    set_error("mlx_learn_features: no network attached");
    return MLX_ERR_STATE;
  }
  MLX_CHECK_HIP(hipSetDevice(m.device));
  const int64_t ld = m.ld, nT = m.nT;
  MLX_TRY(m.learn.d_feat.reserve(sizeof(float) * kMlpIn * ld));
  if (m.precision == MLX_FP64)
    MLX_TRY(learn_infer<double>(m, h, m.learn.d_feat.as<float>()));
  else
    MLX_TRY(learn_infer<float>(m, h, m.learn.d_feat.as<float>()));
  std::vector<float> f((size_t)kMlpIn * ld);
  MLX_CHECK_HIP(hipMemcpyAsync(f.data(), m.learn.d_feat.ptr, sizeof(float) * f.size(),
                               hipMemcpyDeviceToHost, m.stream));
  MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  for (int64_t e = 0; e < nT; ++e)
    std::memcpy(features + (size_t)kMlpIn * e, f.data() + (size_t)kMlpIn * m.slotOf[e],
                sizeof(float) * kMlpIn);
  return MLX_OK;
}

int mlx_synchronize(void) {
  MLX_CHECK_HIP(hipSetDevice(g_device));
  MLX_CHECK_HIP(hipDeviceSynchronize());
  return MLX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batched independent scenes (one persistent launch per substep)
// ---------------------------------------------------------------------------

struct mlx_batch {
  mlx_mesh *m = nullptr;
  int nScenes = 0;
  int block = 512;
  std::vector<int64_t> vertOff, elemOff, cont0, pin0;
  DevBuf d_tables;   // color0 | ncolor | colorOff | colorCnt | slot0 | cont0 | pin0
  DevBuf d_v;        // velocity T 3nV
  DevBuf d_out;      // per-scene outputs
  DevBuf d_pinMask;  // uint8 3nV
  DevBuf d_sceneNH;  // uint8 per scene: holds NeoHookean tetrahedra (allocated only if any)
  DevBuf d_flags;    // int32 nV (plane test)
  DevBuf d_split;    // split schedule, per-scene barriers and flags, partials
  int split = 0;     // target tets per workgroup (0: one workgroup per scene)
  int splitWGs = 0;  // workgroups of the last split launch
  int64_t splitFallbacks = 0;  // split launches redone unsplit (XCD placement differed)
  int useLds = 1;              // scenes that fit keep their positions in LDS (k_batch_plain, k_batch_layered)
  int ldsVerts = 0;            // vertex limit of the LDS path in the last launch
  bool tablesReady = false;    // static part of d_tables uploaded
  int useGraph = 1;            // substeps after frame 1 replay a captured graph
  hipGraphExec_t gexec = nullptr;
  std::vector<uint64_t> gsig;  // what the captured kernels bake in
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // coarse layer with rigid clusters (mlx_batch_set_rigid)
  bool layered = false;
  int nBodies = 0;
  DevBuf d_layer;    // body0 | rigOff | rigVert | bndOff | bndVert | vbody | flags0
  size_t o_body0 = 0, o_rigOff = 0, o_rigVert = 0, o_bndOff = 0, o_bndVert = 0, o_vbody = 0,
         o_flags0 = 0;
  DevBuf d_bodies, d_ri, d_rv, d_lamB;
  DevBuf d_citers;                   // int32 per scene: iterations of the coarse layer
  std::vector<int64_t> elastic0;     // per scene: elastic elements of the coarse layer
  // per-bucket launches (mlx_batch_autotune): scenes grouped by their mean
  // elements per color, each group launched with its own block size on its
  // own stream, the streams forked from and joined to the batch's stream
  static constexpr int kBuckets = 3;
  bool useBuckets = false;
  int bucketBlock[kBuckets] = {512, 512, 512};
  int bucketOff[kBuckets + 1] = {0, 0, 0, 0};
  int64_t bucketMaxVerts[kBuckets] = {0, 0, 0};
  DevBuf d_bucket;                   // int32 scene lists, bucket after bucket
  hipStream_t bstream[kBuckets] = {nullptr, nullptr, nullptr};
  hipEvent_t bfork = nullptr, bjoin[kBuckets] = {nullptr, nullptr, nullptr};
  int tuneOnly = -1;                 // while tuning: launch this bucket only
};

namespace mlx {

// residual velocity and release of every (elastic) vertex for the plain
// method (xpbdLayer3D.m:662-664 then :138-139).
template <typename T>
__global__ void k_batch_prologue(int64_t nV, const T *v, vec4_t<T> *x, const double *mass,
                                 const uint8_t *pinnedDof, double h, double g, double airDrag) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nV) return;
  vec4_t<T> p = x[i];
  T *px = reinterpret_cast<T *>(&p);
  const double m = mass[i];
  for (int d = 0; d < 3; ++d) {
    double r = (double)v[3 * i + d];
    if (!pinnedDof[3 * i + d]) {
      const double f = d == 2 ? 0.0 + m * g : 0.0;
      r = r + h * f / m;
    }
    r = (1 - airDrag) * r;
    px[d] = (T)((double)px[d] + h * (double)(T)r);
  }
  x[i] = p;
}

template <typename T, int NT>
static void launch_batch(const BatchParams<T> &B, int n, size_t ldsBytes, hipStream_t s) {
  hipLaunchKernelGGL((k_batch_plain<T, NT>), dim3(n), dim3(NT), ldsBytes, s, B);
}

// Scenes with at most this many vertices keep their positions in LDS
// (3 values of T per vertex), within the device's per-workgroup LDS less
// the largest static LDS of the kernels fns; 0 disables the LDS path.
template <int NF>
static int lds_verts_for(const void *const (&fns)[NF], size_t elem, int64_t maxSceneVerts,
                         int enabled, size_t *bytes) {
  *bytes = 0;
  if (!enabled) return 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
  size_t stat = 0;
  for (const void *fn : fns) {
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, fn) != hipSuccess) return 0;
    stat = std::max(stat, fa.sharedSizeBytes);
  }
  if (prop.sharedMemPerBlock <= stat + 1024) return 0;
  const size_t avail = prop.sharedMemPerBlock - stat - 256;
  const int64_t fit = (int64_t)(avail / (3 * elem));
  const int64_t nv = std::min<int64_t>(fit, maxSceneVerts);
  if (nv <= 0) return 0;
  *bytes = (size_t)nv * 3 * elem;
  if (*bytes > 65536)
    for (const void *fn : fns)
      if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)*bytes) !=
          hipSuccess) {
        *bytes = 0;
        return 0;
      }
  return (int)nv;
}

template <typename T, int NT>
static int batch_lds_verts(int64_t maxSceneVerts, int enabled, size_t *bytes) {
  const void *const fns[2] = {reinterpret_cast<const void *>(&k_batch_plain<T, NT>),
                              reinterpret_cast<const void *>(&k_batch_split<T, NT>)};
  return lds_verts_for(fns, sizeof(T), maxSceneVerts, enabled, bytes);
}

template <typename T, int NT, bool LEARN = false>
static int layered_lds_verts(int64_t maxSceneVerts, int enabled, size_t *bytes) {
  const void *const fns[1] = {reinterpret_cast<const void *>(&k_batch_layered<T, NT, LEARN>)};
  return lds_verts_for(fns, sizeof(T), maxSceneVerts, enabled, bytes);
}

// k_batch_layered over n scenes; scenes of at most maxSceneVerts vertices
// that fit the LDS keep their positions there.
template <typename T, int NT, bool LEARN = false>
static void launch_layered(BatchParams<T> B, const LayeredParams<T> &Lp, int n,
                           int64_t maxSceneVerts, int useLds, hipStream_t s) {
  size_t lb = 0;
  B.ldsVerts = layered_lds_verts<T, NT, LEARN>(maxSceneVerts, useLds, &lb);
  hipLaunchKernelGGL((k_batch_layered<T, NT, LEARN>), dim3(n), dim3(NT), lb, s, B, Lp);
}

// First contact of each scene: the number of flagged vertices before the
// scene's first vertex (pos: exclusive scan of the plane flags); entry ns
// is the total.
__global__ void k_scene_contacts(int ns, const int64_t *vertOff, int64_t nV, const int32_t *pos,
                                 const int32_t *count, int64_t *cont0) {
  const int sc = blockIdx.x * kBlock + threadIdx.x;
  if (sc > ns) return;
  cont0[sc] = (sc == ns || vertOff[sc] >= nV) ? (int64_t)count[0] : (int64_t)pos[vertOff[sc]];
}

template <typename T>
static int batch_integrate_impl(mlx_batch &b, mlx_settings *st, double h, int frame,
                                const mlx_plane *plane, mlx_batch_out *out) {
  mlx_mesh &m = *b.m;
  const int64_t nV = m.nV;
  hipStream_t s = m.stream;
  vec4_t<T> *x = m.d_x.as<vec4_t<T>>();
  vec4_t<T> *xo = m.d_xold.as<vec4_t<T>>();
  const int64_t nPin = (int64_t)m.pinnedDOF.size();
  if (b.layered) {
    // the coarse layer's rigid elements are the clusters of
    // mlx_batch_set_rigid, then the all-elastic layer
    if (st->nLayers != 2 || st->layers[1] != 0 || st->layers[0] == 0) {
      set_error("a batch with rigid clusters runs two layers: layers = {coarse, 0}");
      return MLX_ERR_UNSUPPORTED;
    }
    if (st->useRigidContacts) {
      set_error("the batched path projects elastic contacts only (useRigidContacts = false)");
      return MLX_ERR_UNSUPPORTED;
    }
  } else {
    for (int i = 0; i < st->nLayers; ++i)
      if (st->layers[i] != 0) {
        set_error("the batched path runs the plain method: every layer must be 0 %% rigid "
                  "(rigid clusters: mlx_batch_set_rigid)");
        return MLX_ERR_UNSUPPORTED;
      }
  }
  if (st->runUntilResSmallerThan >= 0 || st->stopAtImprovementPercent > 0) {
    if (!st->computeResiduals || st->hangingStop < 0) {
      set_error("a residual stop in the batched path needs computeResiduals and hangingStop >= 0");
      return MLX_ERR_UNSUPPORTED;
    }
  }
  // learned warm start of the all-elastic layer (mlx_batch_set_learned)
  const bool learned = m.learn.model && (m.learn.flags & MLX_LEARN_WARMSTART);
  if (learned) {
    if (!b.layered) {
      set_error("the batched learned warm start runs with rigid clusters (mlx_batch_set_rigid)");
      return MLX_ERR_UNSUPPORTED;
    }
    if (!b.useBuckets && b.block != 256 && b.block != 512) {
      set_error("the batched learned warm start runs at 256 or 512 threads per scene");
      return MLX_ERR_UNSUPPORTED;
    }
    for (int i = 0; b.useBuckets && i < mlx_batch::kBuckets; ++i)
      if (b.bucketBlock[i] != 256 && b.bucketBlock[i] != 512) {
        set_error("the batched learned warm start runs at 256 or 512 threads per scene");
        return MLX_ERR_UNSUPPORTED;
      }
  }
  MLX_TRY(m.d_tmp.reserve(4096));
  int *flag = reinterpret_cast<int *>(m.d_tmp.as<char>() + 4000);
  // per-scene tables: color0 | ncolor | colorOff | colorCnt | slot0 | cont0 | pin0 | vert0.
  // All but cont0 are fixed for the batch and uploaded once; cont0 is
  // written on the device every substep.
  const int ns = b.nScenes, ncT = (int)m.colorCnt.size();
  const size_t o_color0 = 0, o_ncolor = o_color0 + 4 * (size_t)ns,
               o_coff = (o_ncolor + 4 * (size_t)ns + 7) / 8 * 8, o_ccnt = o_coff + 8 * (size_t)ncT,
               o_slot = o_ccnt + 8 * (size_t)ncT, o_cont = o_slot + 8 * (size_t)(ns + 1),
               o_pin = o_cont + 8 * (size_t)(ns + 1), o_vert = o_pin + 8 * (size_t)(ns + 1),
               o_end = o_vert + 8 * (size_t)(ns + 1);
  if (!b.tablesReady) {
    std::vector<char> tab(o_end, 0);
    std::memcpy(tab.data() + o_color0, m.sceneColor0.data(), 4 * (size_t)ns);
    std::memcpy(tab.data() + o_ncolor, m.sceneNColors.data(), 4 * (size_t)ns);
    std::memcpy(tab.data() + o_coff, m.colorOff.data(), 8 * (size_t)ncT);
    std::memcpy(tab.data() + o_ccnt, m.colorCnt.data(), 8 * (size_t)ncT);
    std::memcpy(tab.data() + o_slot, m.sceneSlot0.data(), 8 * (size_t)(ns + 1));
    std::memcpy(tab.data() + o_pin, b.pin0.data(), 8 * (size_t)(ns + 1));
    std::memcpy(tab.data() + o_vert, b.vertOff.data(), 8 * (size_t)(ns + 1));
    MLX_TRY(b.d_tables.reserve(o_end));
    MLX_CHECK_HIP(hipMemcpy(b.d_tables.ptr, tab.data(), o_end, hipMemcpyHostToDevice));
    b.tablesReady = true;
  }
  int64_t *cont0Dev = reinterpret_cast<int64_t *>(b.d_tables.as<char>() + o_cont);
  const int64_t *vertOffDev = reinterpret_cast<const int64_t *>(b.d_tables.as<char>() + o_vert);
  MLX_TRY(reserve_contacts(m));
  ContactArrays ca = contact_arrays(m, 0);
  int32_t *countDev = const_cast<int32_t *>(ca.counts);
  const bool planeOn = plane && plane->enabled;
  double nrm[3] = {0, 0, 0}, Dp = 0;
  if (planeOn) {
    double nn = 0;
    for (int d = 0; d < 3; ++d) nn += plane->normal[d] * plane->normal[d];
    nn = std::sqrt(nn);
    for (int d = 0; d < 3; ++d) {
      nrm[d] = plane->normal[d] / nn;
      Dp += plane->normal[d] * plane->position[d];
    }
    MLX_TRY(b.d_flags.reserve(2 * sizeof(int32_t) * std::max<int64_t>(nV, 1)));
  }
  int32_t *fl = planeOn ? b.d_flags.as<int32_t>() : nullptr;
  int32_t *pos = planeOn ? fl + nV : nullptr;
  T *lam = m.d_lam.as<T>();
  const int64_t nLam = (int64_t)m.nc * m.ld;
  // Device work before the solve: NaN / frame-1 reset of cache.oldp,
  // contacts at the start-of-step positions (the plane test, the list in
  // ascending vertex order and each scene's range; the count stays in
  // ca.counts[0] for the kernels that need it), the position shift, the
  // prologue and lambda = 0.  Fill and copy kernels instead of memset and
  // memcpy nodes, so the sequence can be captured.
  auto enqueue_pre = [&]() -> int {
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(kBlock), 0, s, (int64_t)1, flag, 0);
    if (frame != 1)
      hipLaunchKernelGGL(k_any_nan<T>, dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, xo, flag);
    hipLaunchKernelGGL((k_reset_oldp<T, 3>), dim3(blocks_for(std::max(nV, nPin))), dim3(kBlock), 0,
                       s, nV, flag, frame == 1, x, xo, nPin, m.d_pinned.as<int32_t>(),
                       m.d_pinnedx.as<T>());
    // the network reads p, cache.oldp and the previous substep's multipliers
    if (learned) MLX_TRY(learn_infer<T>(m, h, nullptr));
    if (planeOn) {
      hipLaunchKernelGGL((k_plane_flags<T, 3>), dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, x,
                         nrm[0], nrm[1], nrm[2], Dp, 0, fl);
      MLX_TRY(compact_flags_dev(m, fl, nV, pos, const_cast<int32_t *>(ca.vert), ca.lambda,
                                countDev, s));
      hipLaunchKernelGGL((k_plane_gather_dev<T, 3>), dim3(blocks_for(nV)), dim3(kBlock), 0, s,
                         countDev, ca.vert, x, nrm[0], nrm[1], nrm[2], Dp, plane->compliance,
                         const_cast<double *>(ca.normal), const_cast<double *>(ca.point),
                         const_cast<double *>(ca.compliance));
      hipLaunchKernelGGL(k_scene_contacts, dim3(blocks_for(ns + 1)), dim3(kBlock), 0, s, ns,
                         vertOffDev, nV, pos, countDev, cont0Dev);
    } else {
      hipLaunchKernelGGL(k_fill<int32_t>, dim3(1), dim3(kBlock), 0, s, (int64_t)1, countDev, 0);
      hipLaunchKernelGGL(k_fill<int64_t>, dim3(blocks_for(ns + 1)), dim3(kBlock), 0, s,
                         (int64_t)(ns + 1), cont0Dev, (int64_t)0);
    }
    hipLaunchKernelGGL(k_copy<vec4_t<T>>, dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV,
                       (const vec4_t<T> *)x, xo);
    if (b.layered)
      hipLaunchKernelGGL(k_residual_velocity3d<T>, dim3(blocks_for(3 * nV)), dim3(kBlock), 0, s,
                         nV, (const T *)b.d_v.as<T>(), b.d_rv.as<T>(), m.d_mass64.as<double>(),
                         (const uint8_t *)b.d_pinMask.as<uint8_t>(), h, st->gravity,
                         st->airDrag);
    else
      hipLaunchKernelGGL(k_batch_prologue<T>, dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV,
                         b.d_v.as<T>(), x, m.d_mass64.as<double>(), b.d_pinMask.as<uint8_t>(), h,
                         st->gravity, st->airDrag);
    hipLaunchKernelGGL(k_fill<T>, dim3(blocks_for(nLam)), dim3(kBlock), 0, s, nLam, lam, T(0));
    return MLX_OK;
  };
  // after the solve: velocity update and contactVelocityUpdate
  auto enqueue_post = [&]() -> int {
    hipLaunchKernelGGL((k_velocity<T, 3>), dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, x, xo,
                       b.d_v.as<T>(), h);
    if (planeOn)
      hipLaunchKernelGGL((k_contact_velocity_dev<T, 3>), dim3(blocks_for(nV)), dim3(kBlock), 0, s,
                         ca, b.d_v.as<T>(), h, st->gravity, st->restitution);
    return MLX_OK;
  };
  // the persistent solve
  MLX_TRY(b.d_out.reserve((4 + 8 + 8 + 4) * (size_t)ns + 64));
  char *ob = b.d_out.as<char>();
  BatchParams<T> B;
  B.P = sweep_params<T>(m, h, nullptr);
  const char *tb = b.d_tables.as<char>();
  B.color0 = reinterpret_cast<const int32_t *>(tb + o_color0);
  B.ncolor = reinterpret_cast<const int32_t *>(tb + o_ncolor);
  B.colorOff = reinterpret_cast<const int64_t *>(tb + o_coff);
  B.colorCnt = reinterpret_cast<const int64_t *>(tb + o_ccnt);
  B.slot0 = reinterpret_cast<const int64_t *>(tb + o_slot);
  B.cont0 = reinterpret_cast<const int64_t *>(tb + o_cont);
  B.pin0 = reinterpret_cast<const int64_t *>(tb + o_pin);
  B.vert0 = reinterpret_cast<const int64_t *>(tb + o_vert);
  int64_t maxSceneVerts = 0;
  for (int sc = 0; sc < ns; ++sc)
    maxSceneVerts = std::max<int64_t>(maxSceneVerts, b.vertOff[sc + 1] - b.vertOff[sc]);
  size_t ldsBytes = 0;
  B.ldsVerts = b.block == 256    ? batch_lds_verts<T, 256>(maxSceneVerts, b.useLds, &ldsBytes)
               : b.block == 768  ? batch_lds_verts<T, 768>(maxSceneVerts, b.useLds, &ldsBytes)
               : b.block == 1024 ? batch_lds_verts<T, 1024>(maxSceneVerts, b.useLds, &ldsBytes)
                                 : batch_lds_verts<T, 512>(maxSceneVerts, b.useLds, &ldsBytes);
  b.ldsVerts = B.ldsVerts;
  B.sceneNH = m.hasNH ? b.d_sceneNH.as<uint8_t>() : nullptr;
  B.pinDof = m.d_pinned.as<int32_t>();
  B.pinVal = m.d_pinnedx.as<T>();
  B.c = ca;
  B.minv = m.d_minv64.as<double>();
  B.mp.is3d = 1;
  B.mp.computeResiduals = st->computeResiduals;
  B.mp.hangingStop = st->hangingStop;
  B.mp.giveUpEnabled = st->giveUpEnabled;
  B.mp.lastLayerGiveUpEnabled = st->lastLayerGiveUpEnabled;
  B.mp.giveUpThreshold = st->giveUpThreshold;
  B.mp.h = h;
  int iters = 0;
  for (int i = 0; i < st->nLayers; ++i) iters += st->iterations[i];
  B.iterations = iters;
  B.runUntil0 = st->runUntilResSmallerThan;
  B.stopPct = st->stopAtImprovementPercent;
  B.contactResistance = st->contactResistance;
  B.contactCompliance = plane && plane->enabled ? plane->compliance : 0.0;
  B.h = h;
  B.hardCap = std::max(iters, st->hangingStop >= 0 ? st->hangingStop + 1 : iters);
  B.iters = reinterpret_cast<int32_t *>(ob);
  B.lastResidual = reinterpret_cast<double *>(ob + (4 * (size_t)ns + 7) / 8 * 8);
  B.runUntilOut = B.lastResidual + ns;
  B.residualCount = reinterpret_cast<int32_t *>(B.runUntilOut + ns);
  B.sceneList = nullptr;
  if (!b.ev0) {
    MLX_CHECK_HIP(hipEventCreate(&b.ev0));
    MLX_CHECK_HIP(hipEventCreate(&b.ev1));
  }
  LayeredParams<T> Lp;
  if (b.layered) {
    const char *lb = b.d_layer.as<char>();
    Lp.body0 = reinterpret_cast<const int32_t *>(lb + b.o_body0);
    Lp.rigOff = reinterpret_cast<const int32_t *>(lb + b.o_rigOff);
    Lp.rigVert = reinterpret_cast<const int32_t *>(lb + b.o_rigVert);
    Lp.bndOff = reinterpret_cast<const int32_t *>(lb + b.o_bndOff);
    Lp.bndVert = reinterpret_cast<const int32_t *>(lb + b.o_bndVert);
    Lp.vbody = reinterpret_cast<const int32_t *>(lb + b.o_vbody);
    Lp.flags0 = reinterpret_cast<const uint8_t *>(lb + b.o_flags0);
    Lp.bodies = b.d_bodies.as<Body>();
    Lp.ri = b.d_ri.as<double>();
    Lp.rv = b.d_rv.as<T>();
    Lp.lamB = b.d_lamB.as<double>();
    Lp.mass = m.d_mass64.as<double>();
    Lp.coarseIters = b.d_citers.as<int32_t>();
    Lp.it0 = st->iterations[0];
    Lp.it1 = st->iterations[1];
    Lp.boundaryCompliance = st->boundaryCompliance;
    if (learned) {
      mlx_mesh::Learn &Ln = m.learn;
      Lp.learn.lam0 = Ln.d_lam0.as<T>();
      Lp.learn.dx = Ln.d_dx.as<T>();
      Lp.learn.xsave = Ln.d_save.as<vec4_t<T>>();
      Lp.learn.lamsave = reinterpret_cast<T *>(Lp.learn.xsave + nV);
      Lp.learn.incOff = Ln.d_inc.as<int32_t>();
      Lp.learn.inc = Lp.learn.incOff + nV + 1;
      Lp.learn.pinMask = b.d_pinMask.as<uint8_t>();
      Lp.learn.counts = Ln.d_counts.as<unsigned long long>();
    }
  }
  // one bucket's launch: its scene list, block size and LDS share
  auto launch_bucket = [&](int i, hipStream_t bs) {
    BatchParams<T> Bi = B;
    Bi.sceneList = b.d_bucket.as<int32_t>() + b.bucketOff[i];
    const int n = b.bucketOff[i + 1] - b.bucketOff[i];
    const int blk = b.bucketBlock[i];
    if (b.layered) {
      if (learned) {
        if (blk == 256) launch_layered<T, 256, true>(Bi, Lp, n, b.bucketMaxVerts[i], b.useLds, bs);
        else launch_layered<T, 512, true>(Bi, Lp, n, b.bucketMaxVerts[i], b.useLds, bs);
      } else {
        if (blk == 256) launch_layered<T, 256>(Bi, Lp, n, b.bucketMaxVerts[i], b.useLds, bs);
        else launch_layered<T, 512>(Bi, Lp, n, b.bucketMaxVerts[i], b.useLds, bs);
      }
      return;
    }
    size_t lb = 0;
    Bi.ldsVerts = blk == 256    ? batch_lds_verts<T, 256>(b.bucketMaxVerts[i], b.useLds, &lb)
                  : blk == 768  ? batch_lds_verts<T, 768>(b.bucketMaxVerts[i], b.useLds, &lb)
                  : blk == 1024 ? batch_lds_verts<T, 1024>(b.bucketMaxVerts[i], b.useLds, &lb)
                                : batch_lds_verts<T, 512>(b.bucketMaxVerts[i], b.useLds, &lb);
    if (blk == 256) launch_batch<T, 256>(Bi, n, lb, bs);
    else if (blk == 768) launch_batch<T, 768>(Bi, n, lb, bs);
    else if (blk == 1024) launch_batch<T, 1024>(Bi, n, lb, bs);
    else launch_batch<T, 512>(Bi, n, lb, bs);
  };
  auto launch_unsplit = [&]() {
    if (b.useBuckets) {
      (void)hipEventRecord(b.bfork, s);
      for (int i = 0; i < mlx_batch::kBuckets; ++i) {
        if (b.bucketOff[i + 1] == b.bucketOff[i] || (b.tuneOnly >= 0 && b.tuneOnly != i)) continue;
        (void)hipStreamWaitEvent(b.bstream[i], b.bfork, 0);
        launch_bucket(i, b.bstream[i]);
        (void)hipEventRecord(b.bjoin[i], b.bstream[i]);
        (void)hipStreamWaitEvent(s, b.bjoin[i], 0);
      }
      return;
    }
    if (b.layered) {
      if (learned) {  // 256 or 512 threads (checked above)
        if (b.block == 256) launch_layered<T, 256, true>(B, Lp, ns, maxSceneVerts, b.useLds, s);
        else launch_layered<T, 512, true>(B, Lp, ns, maxSceneVerts, b.useLds, s);
        return;
      }
      if (b.block == 256) launch_layered<T, 256>(B, Lp, ns, maxSceneVerts, b.useLds, s);
      else if (b.block == 768) launch_layered<T, 768>(B, Lp, ns, maxSceneVerts, b.useLds, s);
      else if (b.block == 1024) launch_layered<T, 1024>(B, Lp, ns, maxSceneVerts, b.useLds, s);
      else launch_layered<T, 512>(B, Lp, ns, maxSceneVerts, b.useLds, s);
      return;
    }
    if (b.block == 256) launch_batch<T, 256>(B, ns, ldsBytes, s);
    else if (b.block == 768) launch_batch<T, 768>(B, ns, ldsBytes, s);
    else if (b.block == 1024) launch_batch<T, 1024>(B, ns, ldsBytes, s);
    else launch_batch<T, 512>(B, ns, ldsBytes, s);
  };
  int *splitErr = nullptr, *splitMismatch = nullptr;
  bool replayed = false;
  // the split schedule has a StVK-Voigt plain body only
  if (b.split > 0 && !m.hasNH && !b.layered && !b.useBuckets) {
    MLX_TRY(enqueue_pre());
    // split schedule: scene sc gets G = round(nT_sc / split) workgroups (1..64)
    // on one XCD; scenes go to XCDs largest first, each to the XCD with the
    // fewest tets so far; P = the largest XCD's workgroup count, and the grid
    // of 8 P workgroups must be resident (else the split is coarsened)
    const void *fn = b.block == 256    ? reinterpret_cast<const void *>(&k_batch_split<T, 256>)
                     : b.block == 768  ? reinterpret_cast<const void *>(&k_batch_split<T, 768>)
                     : b.block == 1024 ? reinterpret_cast<const void *>(&k_batch_split<T, 1024>)
                                       : reinterpret_cast<const void *>(&k_batch_split<T, 512>);
    int dev = 0;
    MLX_CHECK_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MLX_CHECK_HIP(hipGetDeviceProperties(&prop, dev));
    std::vector<int> order(ns);
    for (int sc = 0; sc < ns; ++sc) order[sc] = sc;
    std::stable_sort(order.begin(), order.end(), [&](int p_, int q_) {
      return b.elemOff[p_ + 1] - b.elemOff[p_] > b.elemOff[q_ + 1] - b.elemOff[q_];
    });
    std::vector<int32_t> G(ns), xcdOf(ns);
    int64_t P = 0;
    bool fits = false;
    // the LDS of single-workgroup scenes limits residency: halve it (fewer
    // scenes keep positions in LDS) until a schedule fits
    for (;;) {
      int perCU = 0;
      MLX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, b.block, ldsBytes));
      const int64_t room = (int64_t)perCU * prop.multiProcessorCount / kXcds;
      double per = (double)b.split;
      for (int guard = 0; guard < 64 && !fits; ++guard, per *= 1.25) {
        int64_t tets[kXcds] = {0}, wgs[kXcds] = {0};
        bool ok = true;
        for (int sc : order) {
          const double nt = (double)(b.elemOff[sc + 1] - b.elemOff[sc]);
          G[sc] = (int32_t)std::max(1.0, std::min(64.0, std::floor(nt / per + 0.5)));
          int best = -1;
          for (int x = 0; x < kXcds; ++x)
            if (wgs[x] + G[sc] <= room && (best < 0 || tets[x] < tets[best])) best = x;
          if (best < 0) {
            ok = false;
            break;
          }
          xcdOf[sc] = best;
          tets[best] += (int64_t)nt;
          wgs[best] += G[sc];
        }
        if (!ok) continue;
        P = *std::max_element(wgs, wgs + kXcds);
        fits = P >= 1;
      }
      if (fits || ldsBytes == 0) break;
      ldsBytes = ldsBytes / 2 < 4096 ? 0 : ldsBytes / 2;
      B.ldsVerts = (int)(ldsBytes / (3 * sizeof(T)));
      b.ldsVerts = B.ldsVerts;
    }
    if (!fits) {
      set_error("batched split: no schedule fits the resident workgroups");
      return MLX_ERR_UNSUPPORTED;
    }
    std::vector<int32_t> entry((size_t)kXcds * P, -1), pbase(ns);
    std::vector<int64_t> fill(kXcds, 0);
    int64_t np = 0;
    for (int sc = 0; sc < ns; ++sc) {
      pbase[sc] = (int32_t)np;
      np += G[sc];
      const int x = xcdOf[sc];
      for (int q = 0; q < G[sc]; ++q) entry[(size_t)x * P + fill[x]++] = sc * 64 + q;
    }
    const size_t nE = entry.size();
    const size_t o_entry = 0, o_G = o_entry + 4 * nE, o_pb = o_G + 4 * (size_t)ns,
                 o_zero = (o_pb + 4 * (size_t)ns + 255) / 256 * 256, o_reg = o_zero,
                 o_bar = o_reg + 4 * 32, o_stop = o_bar + 4 * (size_t)ns,
                 o_err = o_stop + 4 * (size_t)ns, o_mis = o_err + 4,
                 o_zend = (o_mis + 4 + 7) / 8 * 8, o_ru = o_zend, o_part = o_ru + 8 * (size_t)ns,
                 o_end = o_part + 8 * (size_t)np;
    MLX_TRY(b.d_split.reserve(o_end));
    char *sb = b.d_split.as<char>();
    MLX_CHECK_HIP(hipMemcpy(sb + o_entry, entry.data(), 4 * nE, hipMemcpyHostToDevice));
    MLX_CHECK_HIP(hipMemcpy(sb + o_G, G.data(), 4 * (size_t)ns, hipMemcpyHostToDevice));
    MLX_CHECK_HIP(hipMemcpy(sb + o_pb, pbase.data(), 4 * (size_t)ns, hipMemcpyHostToDevice));
    MLX_CHECK_HIP(hipMemsetAsync(sb + o_zero, 0, o_zend - o_zero, s));
    SplitParams SP;
    SP.entry = reinterpret_cast<const int32_t *>(sb + o_entry);
    SP.sceneG = reinterpret_cast<const int32_t *>(sb + o_G);
    SP.pbase = reinterpret_cast<const int32_t *>(sb + o_pb);
    SP.P = (int)P;
    SP.reg = reinterpret_cast<unsigned *>(sb + o_reg);
    SP.bar = reinterpret_cast<unsigned *>(sb + o_bar);
    SP.sceneStop = reinterpret_cast<int32_t *>(sb + o_stop);
    SP.err = reinterpret_cast<int *>(sb + o_err);
    SP.mismatch = reinterpret_cast<int *>(sb + o_mis);
    SP.sceneRunUntil = reinterpret_cast<double *>(sb + o_ru);
    SP.partials = reinterpret_cast<double *>(sb + o_part);
    splitErr = SP.err;
    splitMismatch = SP.mismatch;
    b.splitWGs = (int)(kXcds * P);
    void *args[] = {&B, &SP};
    MLX_CHECK_HIP(hipEventRecord(b.ev0, s));
    MLX_CHECK_HIP(hipLaunchCooperativeKernel(fn, dim3((unsigned)(kXcds * P)), dim3(b.block), args,
                                             ldsBytes, s));
    MLX_CHECK_HIP(hipEventRecord(b.ev1, s));
    // workgroups dealt other than round-robin over the XCDs: nothing was
    // written; run the substep with one workgroup per scene instead
    int mis = 0;
    MLX_CHECK_HIP(hipMemcpyAsync(&mis, splitMismatch, 4, hipMemcpyDeviceToHost, s));
    MLX_CHECK_HIP(hipStreamSynchronize(s));
    if (mis) {
      ++b.splitFallbacks;
      MLX_CHECK_HIP(hipEventRecord(b.ev0, s));
      if (b.block == 256) launch_batch<T, 256>(B, ns, ldsBytes, s);
      else if (b.block == 768) launch_batch<T, 768>(B, ns, ldsBytes, s);
      else if (b.block == 1024) launch_batch<T, 1024>(B, ns, ldsBytes, s);
      else launch_batch<T, 512>(B, ns, ldsBytes, s);
      MLX_CHECK_HIP(hipEventRecord(b.ev1, s));
    }
    MLX_TRY(enqueue_post());
  } else if (b.useGraph && frame != 1) {
    // the whole substep as one graph, captured again when anything baked
    // into its kernels changes
    auto bits = [](double d) {
      uint64_t u;
      std::memcpy(&u, &d, 8);
      return u;
    };
    const std::vector<uint64_t> sig = {
        (uint64_t)ns, (uint64_t)nV, (uint64_t)b.block, (uint64_t)ldsBytes, (uint64_t)B.ldsVerts,
        (uint64_t)planeOn, bits(nrm[0]), bits(nrm[1]), bits(nrm[2]), bits(Dp),
        bits(planeOn ? plane->compliance : 0.0), bits(h), bits(st->gravity), bits(st->airDrag),
        bits(st->restitution), (uint64_t)B.iterations, bits(B.runUntil0), bits(B.stopPct),
        bits(B.contactResistance), (uint64_t)B.hardCap, (uint64_t)B.mp.computeResiduals,
        (uint64_t)B.mp.hangingStop, (uint64_t)B.mp.giveUpEnabled,
        (uint64_t)B.mp.lastLayerGiveUpEnabled, bits(B.mp.giveUpThreshold),
        (uint64_t)(uintptr_t)m.d_x.ptr, (uint64_t)(uintptr_t)m.d_xold.ptr,
        (uint64_t)(uintptr_t)m.d_lam.ptr, (uint64_t)(uintptr_t)b.d_v.ptr,
        (uint64_t)(uintptr_t)b.d_flags.ptr, (uint64_t)(uintptr_t)m.d_scan.ptr,
        (uint64_t)(uintptr_t)m.d_contacts.ptr, (uint64_t)(uintptr_t)b.d_tables.ptr,
        (uint64_t)(uintptr_t)b.d_out.ptr, (uint64_t)(uintptr_t)m.d_tmp.ptr,
        (uint64_t)(uintptr_t)b.d_pinMask.ptr, (uint64_t)(uintptr_t)m.d_pinnedx.ptr,
        (uint64_t)b.layered, (uint64_t)(uintptr_t)b.d_layer.ptr,
        (uint64_t)(uintptr_t)b.d_bodies.ptr, (uint64_t)(uintptr_t)b.d_ri.ptr,
        (uint64_t)(uintptr_t)b.d_rv.ptr, (uint64_t)(uintptr_t)b.d_lamB.ptr,
        (uint64_t)(b.layered ? Lp.it0 : 0), (uint64_t)(b.layered ? Lp.it1 : 0),
        bits(b.layered ? Lp.boundaryCompliance : 0.0), (uint64_t)b.useBuckets,
        (uint64_t)b.bucketBlock[0], (uint64_t)b.bucketBlock[1], (uint64_t)b.bucketBlock[2],
        (uint64_t)(uintptr_t)b.d_bucket.ptr, (uint64_t)b.useLds};
    if (!b.gexec || b.gsig != sig) {
      if (b.gexec) {
        (void)hipGraphExecDestroy(b.gexec);
        b.gexec = nullptr;
      }
      hipGraph_t graph = nullptr;
      MLX_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int r = enqueue_pre();
      if (r == MLX_OK) launch_unsplit();
      if (r == MLX_OK) r = enqueue_post();
      const hipError_t ce = hipStreamEndCapture(s, &graph);
      if (r != MLX_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return r;
      }
      MLX_CHECK_HIP(ce);
      const hipError_t ie = hipGraphInstantiate(&b.gexec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) b.gexec = nullptr;
      MLX_CHECK_HIP(ie);
      b.gsig = sig;
    }
    MLX_CHECK_HIP(hipEventRecord(b.ev0, s));
    MLX_CHECK_HIP(hipGraphLaunch(b.gexec, s));
    MLX_CHECK_HIP(hipEventRecord(b.ev1, s));
    replayed = true;
  } else {
    MLX_TRY(enqueue_pre());
    MLX_CHECK_HIP(hipEventRecord(b.ev0, s));
    launch_unsplit();
    MLX_CHECK_HIP(hipEventRecord(b.ev1, s));
    MLX_TRY(enqueue_post());
  }
  MLX_CHECK_HIP(hipGetLastError());
  MLX_CHECK_HIP(hipStreamSynchronize(s));
  b.cont0.assign(ns + 1, 0);
  MLX_CHECK_HIP(hipMemcpy(b.cont0.data(), cont0Dev, 8 * (size_t)(ns + 1), hipMemcpyDeviceToHost));
  if (splitErr) {
    int e = 0;
    MLX_CHECK_HIP(hipMemcpy(&e, splitErr, 4, hipMemcpyDeviceToHost));
    if (e) {
      set_error("batched split: a workgroup did not reach its scene barrier");
      return MLX_ERR_STATE;
    }
  }
  std::vector<int32_t> it(ns), rc(ns);
  std::vector<double> lr(ns);
  MLX_CHECK_HIP(hipMemcpy(it.data(), B.iters, 4 * (size_t)ns, hipMemcpyDeviceToHost));
  MLX_CHECK_HIP(hipMemcpy(lr.data(), B.lastResidual, 8 * (size_t)ns, hipMemcpyDeviceToHost));
  float kms = 0;
  // direct launches: the solve kernel(s); a replay: the whole graph (its
  // small pre- and post-solve kernels included)
  MLX_CHECK_HIP(hipEventElapsedTime(&kms, b.ev0, b.ev1));
  (void)replayed;
  std::vector<int32_t> cit(ns, 0);
  if (b.layered)
    MLX_CHECK_HIP(hipMemcpy(cit.data(), b.d_citers.ptr, 4 * (size_t)ns, hipMemcpyDeviceToHost));
  if (out) {
    int64_t proj = 0;
    for (int sc = 0; sc < ns; ++sc) {
      // the coarse layer projects its elastic elements only
      proj += (int64_t)(it[sc] - cit[sc]) * (b.elemOff[sc + 1] - b.elemOff[sc]) +
              (b.layered ? (int64_t)cit[sc] * b.elastic0[sc] : 0);
      if (out->iterations) out->iterations[sc] = it[sc];
      if (out->lastResidual) out->lastResidual[sc] = lr[sc];
      if (out->contacts) out->contacts[sc] = (int32_t)(b.cont0[sc + 1] - b.cont0[sc]);
    }
    out->elasticProjections = proj;
    out->kernelMs = kms;
  }
  return MLX_OK;
}

}  // namespace mlx

extern "C" {

int mlx_batch_create(const mlx_mesh_desc *d, int32_t nScenes, const int64_t *vertOff,
                     const int64_t *elemOff, int32_t precision, mlx_batch **out) {
  if (!d || !out || nScenes < 1 || !vertOff || !elemOff || d->dim != 3 ||
      vertOff[0] != 0 || elemOff[0] != 0 || vertOff[nScenes] != d->nV || elemOff[nScenes] != d->nT) {
    set_error("mlx_batch_create: bad arguments (3D scenes, offsets must span the merged mesh)");
    return MLX_ERR_ARG;
  }
  for (int sc = 0; sc < nScenes; ++sc)
    if (vertOff[sc + 1] < vertOff[sc] || elemOff[sc + 1] <= elemOff[sc]) {
      set_error("mlx_batch_create: scene %d has a bad range", sc);
      return MLX_ERR_ARG;
    }
  // every element's vertices must lie in its scene
  for (int sc = 0; sc < nScenes; ++sc)
    for (int64_t e = elemOff[sc]; e < elemOff[sc + 1]; ++e)
      for (int k = 0; k < 4; ++k) {
        const int64_t v = (int64_t)d->t[e + d->nT * k] - 1;
        if (v < vertOff[sc] || v >= vertOff[sc + 1]) {
          set_error("element %lld uses a vertex outside scene %d", (long long)e + 1, sc);
          return MLX_ERR_ARG;
        }
      }
  mlx_batch *b = new mlx_batch();
  int r = mesh_create_impl(d, precision, nScenes, elemOff, 16, &b->m);
  if (r != MLX_OK) {
    delete b;
    return r;
  }
  b->nScenes = nScenes;
  b->vertOff.assign(vertOff, vertOff + nScenes + 1);
  b->elemOff.assign(elemOff, elemOff + nScenes + 1);
  // pinned DOFs are ascending, so each scene's are contiguous
  b->pin0.assign(nScenes + 1, 0);
  {
    const auto &pd = b->m->pinnedDOF;
    for (size_t q = 1; q < pd.size(); ++q)
      if (pd[q] <= pd[q - 1]) {
        mlx_batch_destroy(b);
        set_error("pinnedDOFs must be strictly ascending for a batch");
        return MLX_ERR_ARG;
      }
    size_t q = 0;
    for (int sc = 0; sc < nScenes; ++sc) {
      b->pin0[sc] = (int64_t)q;
      while (q < pd.size() && pd[q] < 3 * vertOff[sc + 1]) ++q;
    }
    b->pin0[nScenes] = (int64_t)pd.size();
  }
  if (b->m->hasNH) {
    // a scene is all StVK-Voigt or all NeoHookean: the batched kernel runs
    // each scene with the body specialised for its one material
    std::vector<uint8_t> snh(nScenes, 0);
    for (int sc = 0; sc < nScenes; ++sc) {
      int64_t n = 0;
      for (int64_t e = elemOff[sc]; e < elemOff[sc + 1]; ++e) n += d->neohookean[e] != 0;
      if (n > 0 && n < elemOff[sc + 1] - elemOff[sc]) {
        mlx_batch_destroy(b);
        set_error("mlx_batch_create: scene %d mixes StVK-Voigt and NeoHookean elements", sc);
        return MLX_ERR_UNSUPPORTED;
      }
      snh[sc] = n > 0;
    }
    if ((r = b->d_sceneNH.reserve(snh.size())) != MLX_OK ||
        hipMemcpy(b->d_sceneNH.ptr, snh.data(), snh.size(), hipMemcpyHostToDevice) != hipSuccess) {
      mlx_batch_destroy(b);
      return r != MLX_OK ? r : MLX_ERR_HIP;
    }
  }
  std::vector<uint8_t> mask(3 * (size_t)d->nV, 0);
  for (auto q : b->m->pinnedDOF) mask[q] = 1;
  if ((r = b->d_pinMask.reserve(mask.size())) != MLX_OK ||
      hipMemcpy(b->d_pinMask.ptr, mask.data(), mask.size(), hipMemcpyHostToDevice) != hipSuccess) {
    mlx_batch_destroy(b);
    return r != MLX_OK ? r : MLX_ERR_HIP;
  }
  *out = b;
  return MLX_OK;
}

int mlx_batch_destroy(mlx_batch *b) {
  if (!b) return MLX_OK;
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->d_tables.release();
  b->d_v.release();
  b->d_out.release();
  b->d_pinMask.release();
  b->d_sceneNH.release();
  b->d_flags.release();
  b->d_split.release();
  b->d_layer.release();
  b->d_bodies.release();
  b->d_ri.release();
  b->d_rv.release();
  b->d_lamB.release();
  b->d_citers.release();
  b->d_bucket.release();
  for (int i = 0; i < mlx_batch::kBuckets; ++i) {
    if (b->bjoin[i]) (void)hipEventDestroy(b->bjoin[i]);
    if (b->bstream[i]) (void)hipStreamDestroy(b->bstream[i]);
  }
  if (b->bfork) (void)hipEventDestroy(b->bfork);
  mlx_mesh_destroy(b->m);
  delete b;
  return MLX_OK;
}

int mlx_batch_set_lds(mlx_batch *b, int32_t enable) {
  if (!b) {
    set_error("mlx_batch_set_lds: null batch");
    return MLX_ERR_ARG;
  }
  b->useLds = enable != 0;
  return MLX_OK;
}

int mlx_batch_set_graph(mlx_batch *b, int32_t enable) {
  if (!b) {
    set_error("mlx_batch_set_graph: null batch");
    return MLX_ERR_ARG;
  }
  b->useGraph = enable != 0;
  return MLX_OK;
}

int mlx_batch_set_split(mlx_batch *b, int32_t tetsPerWorkgroup) {
  if (!b || tetsPerWorkgroup < 0) {
    set_error("mlx_batch_set_split: bad arguments");
    return MLX_ERR_ARG;
  }
  b->split = tetsPerWorkgroup;
  return MLX_OK;
}

int mlx_batch_set_learned(mlx_batch *b, mlx_mlp *model, float kappaMax) {
  if (!b) {
    set_error("mlx_batch_set_learned: null batch");
    return MLX_ERR_ARG;
  }
  MLX_CHECK_HIP(hipSetDevice(b->m->device));
  // the captured substep changes
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  b->gsig.clear();
  return learn_attach(b->m, model, MLX_LEARN_WARMSTART, kappaMax);
}

int mlx_batch_learn_stats(mlx_batch *b, int64_t *kept, int64_t *rejected) {
  if (!b) {
    set_error("mlx_batch_learn_stats: null batch");
    return MLX_ERR_ARG;
  }
  return mlx_learn_stats(b->m, kept, rejected);
}

int mlx_batch_set_rigid(mlx_batch *b, const uint8_t *isRigidElement) {
  if (!b) {
    set_error("mlx_batch_set_rigid: null batch");
    return MLX_ERR_ARG;
  }
  mlx_mesh &m = *b->m;
  MLX_CHECK_HIP(hipSetDevice(m.device));
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  b->gsig.clear();
  if (!isRigidElement) {
    b->layered = false;
    return MLX_OK;
  }
  const int64_t nT = m.nT, nV = m.nV, ld = m.ld;
  const int ns = b->nScenes;
  // rigid elements: the mask less the forced-elastic (pinned) ones
  std::vector<uint8_t> rigid(nT);
  for (int64_t e = 0; e < nT; ++e) rigid[e] = isRigidElement[e] && !m.forcedElastic[e];
  // bodies: rigid elements sharing a vertex, numbered by ascending
  // smallest element (the numbering of mexUnionFindByVertice)
  std::vector<int32_t> parent(nT), owner(nV, -1);
  for (int64_t e = 0; e < nT; ++e) parent[e] = (int32_t)e;
  auto find = [&](int32_t a) {
    while (parent[a] != a) {
      parent[a] = parent[parent[a]];
      a = parent[a];
    }
    return a;
  };
  for (int64_t e = 0; e < nT; ++e) {
    if (!rigid[e]) continue;
    for (int a = 0; a < 4; ++a) {
      const int32_t v = m.T[4 * (size_t)e + a];
      if (owner[v] < 0) {
        owner[v] = (int32_t)e;
        continue;
      }
      int32_t r1 = find((int32_t)e), r2 = find(owner[v]);
      if (r1 != r2) parent[std::max(r1, r2)] = std::min(r1, r2);
    }
  }
  std::vector<int32_t> sets(nT, 0), bodyOfRoot(nT, 0);
  int nb = 0;
  for (int64_t e = 0; e < nT; ++e) {
    if (!rigid[e]) continue;
    const int32_t r = find((int32_t)e);
    if (!bodyOfRoot[r]) bodyOfRoot[r] = ++nb;
    sets[e] = bodyOfRoot[r];
  }
  std::vector<uint8_t> isB(nV), isRV(nV);
  std::vector<int32_t> vbody(nV);
  MLX_TRY(find_boundary_vertices(m.T.data(), nT, 4, nV, 1, sets.data(), isB.data(), isRV.data(),
                                 vbody.data()));
  // bodies per scene: a body lies in one scene (scenes share no vertex) and
  // bodies are numbered in ascending smallest element, so each scene's
  // bodies are consecutive
  std::vector<int32_t> body0(ns + 1, 0);
  b->elastic0.assign(ns, 0);
  for (int sc = 0; sc < ns; ++sc) {
    int32_t lo = INT32_MAX, hi = 0;
    for (int64_t e = b->elemOff[sc]; e < b->elemOff[sc + 1]; ++e) {
      if (sets[e]) {
        lo = std::min(lo, sets[e]);
        hi = std::max(hi, sets[e]);
      } else {
        b->elastic0[sc] += 1;
      }
    }
    const int32_t cnt = hi ? hi - lo + 1 : 0;
    if (cnt > kMaxSceneBodies) {
      set_error("mlx_batch_set_rigid: scene %d has %d rigid bodies (at most %d)", sc, cnt,
                kMaxSceneBodies);
      return MLX_ERR_UNSUPPORTED;
    }
    if (cnt && lo - 1 != body0[sc]) {
      set_error("mlx_batch_set_rigid: bodies of scene %d are not consecutive", sc);
      return MLX_ERR_STATE;
    }
    body0[sc + 1] = body0[sc] + cnt;
  }
  // rigid and boundary vertices of each body, ascending
  std::vector<int32_t> rigOff(nb + 1, 0), bndOff(nb + 1, 0);
  for (int64_t v = 0; v < nV; ++v)
    if (vbody[v]) {
      rigOff[vbody[v]] += 1;
      if (isB[v]) bndOff[vbody[v]] += 1;
    }
  for (int k = 0; k < nb; ++k) {
    rigOff[k + 1] += rigOff[k];
    bndOff[k + 1] += bndOff[k];
  }
  std::vector<int32_t> rigVert(std::max(rigOff[nb], 1)), bndVert(std::max(bndOff[nb], 1));
  {
    std::vector<int32_t> fr(rigOff.begin(), rigOff.end() - 1), fb(bndOff.begin(), bndOff.end() - 1);
    for (int64_t v = 0; v < nV; ++v)
      if (vbody[v]) {
        rigVert[fr[vbody[v] - 1]++] = (int32_t)v;
        if (isB[v]) bndVert[fb[vbody[v] - 1]++] = (int32_t)v;
      }
  }
  std::vector<uint8_t> flags0(ld, 1);
  for (int64_t e = 0; e < nT; ++e) flags0[m.slotOf[e]] = sets[e] == 0;
  auto al = [](size_t o) { return (o + 255) / 256 * 256; };
  b->o_body0 = 0;
  b->o_rigOff = al(b->o_body0 + 4 * body0.size());
  b->o_rigVert = al(b->o_rigOff + 4 * rigOff.size());
  b->o_bndOff = al(b->o_rigVert + 4 * rigVert.size());
  b->o_bndVert = al(b->o_bndOff + 4 * bndOff.size());
  b->o_vbody = al(b->o_bndVert + 4 * bndVert.size());
  b->o_flags0 = al(b->o_vbody + 4 * (size_t)nV);
  const size_t total = al(b->o_flags0 + (size_t)ld);
  std::vector<char> buf(total, 0);
  std::memcpy(buf.data() + b->o_body0, body0.data(), 4 * body0.size());
  std::memcpy(buf.data() + b->o_rigOff, rigOff.data(), 4 * rigOff.size());
  std::memcpy(buf.data() + b->o_rigVert, rigVert.data(), 4 * rigVert.size());
  std::memcpy(buf.data() + b->o_bndOff, bndOff.data(), 4 * bndOff.size());
  std::memcpy(buf.data() + b->o_bndVert, bndVert.data(), 4 * bndVert.size());
  std::memcpy(buf.data() + b->o_vbody, vbody.data(), 4 * (size_t)nV);
  std::memcpy(buf.data() + b->o_flags0, flags0.data(), (size_t)ld);
  MLX_TRY(b->d_layer.reserve(total));
  MLX_CHECK_HIP(hipMemcpy(b->d_layer.ptr, buf.data(), total, hipMemcpyHostToDevice));
  const size_t tb = m.precision == MLX_FP64 ? 8 : 4;
  MLX_TRY(b->d_bodies.reserve(sizeof(Body) * std::max(nb, 1)));
  MLX_TRY(b->d_ri.reserve(sizeof(double) * 3 * (size_t)nV));
  MLX_TRY(b->d_rv.reserve(tb * 3 * (size_t)nV));
  MLX_TRY(b->d_lamB.reserve(sizeof(double) * std::max<size_t>(bndOff[nb], 1)));
  MLX_TRY(b->d_citers.reserve(4 * (size_t)ns));
  MLX_CHECK_HIP(hipMemset(b->d_citers.ptr, 0, 4 * (size_t)ns));
  b->nBodies = nb;
  b->layered = true;
  return MLX_OK;
}

int mlx_batch_set_block(mlx_batch *b, int32_t threads) {
  if (!b || (threads != 256 && threads != 512 && threads != 768 && threads != 1024)) {
    set_error("mlx_batch_set_block: threads must be 256, 512, 768 or 1024");
    return MLX_ERR_ARG;
  }
  b->block = threads;
  b->useBuckets = false;
  return MLX_OK;
}

int mlx_batch_set_state(mlx_batch *b, const double *p, const double *v) {
  if (!b || !p || !v) {
    set_error("mlx_batch_set_state: bad arguments");
    return MLX_ERR_ARG;
  }
  mlx_mesh &m = *b->m;
  MLX_CHECK_HIP(hipSetDevice(m.device));
  MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  const size_t n = 3 * (size_t)m.nV;
  if (m.precision == MLX_FP64) {
    MLX_TRY(upload_positions<double>(m, m.d_x, p));
    MLX_TRY(upload_positions<double>(m, m.d_xold, p));
    MLX_TRY(upload_scalars<double>(b->d_v, v, n));
  } else {
    MLX_TRY(upload_positions<float>(m, m.d_x, p));
    MLX_TRY(upload_positions<float>(m, m.d_xold, p));
    MLX_TRY(upload_scalars<float>(b->d_v, v, n));
  }
  return MLX_OK;
}

int mlx_batch_download(mlx_batch *b, double *p, double *v, double *lambdai) {
  if (!b) {
    set_error("mlx_batch_download: null batch");
    return MLX_ERR_ARG;
  }
  mlx_mesh &m = *b->m;
  MLX_CHECK_HIP(hipSetDevice(m.device));
  MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  const size_t n = 3 * (size_t)m.nV;
  if (m.precision == MLX_FP64) {
    if (p) MLX_TRY(download_positions<double>(m, m.d_x, p));
    if (v) MLX_TRY(download_scalars<double>(b->d_v, v, n));
    if (lambdai) MLX_TRY(download_lambda<double>(m, lambdai));
  } else {
    if (p) MLX_TRY(download_positions<float>(m, m.d_x, p));
    if (v) MLX_TRY(download_scalars<float>(b->d_v, v, n));
    if (lambdai) MLX_TRY(download_lambda<float>(m, lambdai));
  }
  return MLX_OK;
}

int mlx_batch_autotune(mlx_batch *b, mlx_settings *st, double h, int32_t frame,
                       const mlx_plane *plane, int32_t *bucketBlocks, float *bucketMs,
                       int32_t *bucketScenes) {
  if (!b || !st || !(h > 0) || frame < 1) {
    set_error("mlx_batch_autotune: bad arguments");
    return MLX_ERR_ARG;
  }
  mlx_mesh &m = *b->m;
  MLX_CHECK_HIP(hipSetDevice(m.device));
  if (!b->d_v.ptr) {
    set_error("mlx_batch_autotune: call mlx_batch_set_state first");
    return MLX_ERR_STATE;
  }
  const int ns = b->nScenes, K = mlx_batch::kBuckets;
  // buckets by mean elements per color: < 256, < 1024, the rest
  std::vector<int32_t> lists[K];
  for (int i = 0; i < K; ++i) b->bucketMaxVerts[i] = 0;
  for (int sc = 0; sc < ns; ++sc) {
    const double per = (double)(b->elemOff[sc + 1] - b->elemOff[sc]) /
                       std::max(1, m.sceneNColors[sc]);
    const int i = per < 256 ? 0 : per < 1024 ? 1 : 2;
    lists[i].push_back(sc);
    b->bucketMaxVerts[i] = std::max<int64_t>(b->bucketMaxVerts[i],
                                             b->vertOff[sc + 1] - b->vertOff[sc]);
  }
  std::vector<int32_t> all;
  for (int i = 0; i < K; ++i) {
    b->bucketOff[i] = (int)all.size();
    all.insert(all.end(), lists[i].begin(), lists[i].end());
  }
  b->bucketOff[K] = (int)all.size();
  MLX_TRY(b->d_bucket.reserve(sizeof(int32_t) * all.size()));
  MLX_CHECK_HIP(hipMemcpy(b->d_bucket.ptr, all.data(), sizeof(int32_t) * all.size(),
                          hipMemcpyHostToDevice));
  for (int i = 0; i < K; ++i) {
    if (!b->bstream[i]) MLX_CHECK_HIP(hipStreamCreateWithFlags(&b->bstream[i], hipStreamNonBlocking));
    if (!b->bjoin[i]) MLX_CHECK_HIP(hipEventCreateWithFlags(&b->bjoin[i], hipEventDisableTiming));
  }
  if (!b->bfork) MLX_CHECK_HIP(hipEventCreateWithFlags(&b->bfork, hipEventDisableTiming));
  // the state every candidate starts from
  const size_t tb = m.precision == MLX_FP64 ? 8 : 4;
  DevBuf sx, sv;
  MLX_TRY(sx.reserve(4 * tb * (size_t)m.nV));
  MLX_TRY(sv.reserve(3 * tb * (size_t)m.nV));
  MLX_CHECK_HIP(hipMemcpy(sx.ptr, m.d_x.ptr, 4 * tb * (size_t)m.nV, hipMemcpyDeviceToDevice));
  MLX_CHECK_HIP(hipMemcpy(sv.ptr, b->d_v.ptr, 3 * tb * (size_t)m.nV, hipMemcpyDeviceToDevice));
  auto restore = [&]() -> int {
    MLX_CHECK_HIP(hipMemcpy(m.d_x.ptr, sx.ptr, 4 * tb * (size_t)m.nV, hipMemcpyDeviceToDevice));
    MLX_CHECK_HIP(hipMemcpy(b->d_v.ptr, sv.ptr, 3 * tb * (size_t)m.nV, hipMemcpyDeviceToDevice));
    return MLX_OK;
  };
  const int useGraph = b->useGraph;
  b->useGraph = 0;  // candidates launch directly: their kernel time is measured
  b->useBuckets = true;
  const std::vector<int> cands = b->layered ? std::vector<int>{256, 512}
                                            : std::vector<int>{256, 512, 768, 1024};
  mlx_settings st0 = *st;
  int r = MLX_OK;
  for (int i = 0; i < K && r == MLX_OK; ++i) {
    if (bucketScenes) bucketScenes[i] = b->bucketOff[i + 1] - b->bucketOff[i];
    if (bucketBlocks) bucketBlocks[i] = 0;
    if (bucketMs) bucketMs[i] = 0;
    if (b->bucketOff[i + 1] == b->bucketOff[i]) continue;
    int best = cands[0];
    float bestMs = 0;
    for (int c : cands) {
      if ((r = restore()) != MLX_OK) break;
      b->bucketBlock[i] = c;
      b->tuneOnly = i;
      mlx_settings sti = st0;
      mlx_batch_out o;
      std::memset(&o, 0, sizeof o);
      r = m.precision == MLX_FP64 ? batch_integrate_impl<double>(*b, &sti, h, frame, plane, &o)
                                  : batch_integrate_impl<float>(*b, &sti, h, frame, plane, &o);
      if (r != MLX_OK) break;
      if (bestMs == 0 || o.kernelMs < bestMs) {
        bestMs = o.kernelMs;
        best = c;
      }
    }
    b->bucketBlock[i] = best;
    if (bucketBlocks) bucketBlocks[i] = best;
    if (bucketMs) bucketMs[i] = bestMs;
  }
  b->tuneOnly = -1;
  b->useGraph = useGraph;
  const int r2 = restore();
  sx.release();
  sv.release();
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  b->gsig.clear();
  return r != MLX_OK ? r : r2;
}

int mlx_batch_integrate(mlx_batch *b, mlx_settings *st, double h, int32_t frame,
                        const mlx_plane *plane, mlx_batch_out *out) {
  if (!b || !st || !(h > 0) || frame < 1 || st->nLayers < 1 || st->nLayers > MLX_MAX_LAYERS) {
    set_error("mlx_batch_integrate: bad arguments");
    return MLX_ERR_ARG;
  }
  MLX_CHECK_HIP(hipSetDevice(b->m->device));
  if (!b->d_v.ptr) {
    set_error("mlx_batch_integrate: call mlx_batch_set_state first");
    return MLX_ERR_STATE;
  }
  if (b->m->precision == MLX_FP64) return batch_integrate_impl<double>(*b, st, h, frame, plane, out);
  return batch_integrate_impl<float>(*b, st, h, frame, plane, out);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// 2D NeoHookean constraint sweep (row a7)
// ---------------------------------------------------------------------------

namespace mlx {

template <typename T>
static int neohookean2d_impl(mlx_mesh &m, const uint8_t *isElastic, double h, const double *mu,
                             const double *lamL, double *lambdai, double *xi) {
  const int64_t ld = m.ld, nT = m.nT;
  std::vector<T> nh(4 * (size_t)ld, T(0));
  std::vector<uint8_t> fl(ld, 0);
  for (int64_t s = 0; s < ld; ++s) {
    const int32_t e = m.perm[s];
    if (e < 0) continue;
    nh[s] = (T)mu[e];
    nh[ld + s] = (T)lamL[e];
    nh[2 * ld + s] = (T)lambdai[e];
    nh[3 * ld + s] = (T)lambdai[nT + e];
    fl[s] = isElastic[e] ? 1 : 0;
  }
  MLX_TRY(m.d_nh.reserve(sizeof(T) * nh.size()));
  MLX_TRY(m.d_flags.reserve(ld));
  MLX_CHECK_HIP(hipMemcpy(m.d_nh.ptr, nh.data(), sizeof(T) * nh.size(), hipMemcpyHostToDevice));
  MLX_CHECK_HIP(hipMemcpy(m.d_flags.ptr, fl.data(), ld, hipMemcpyHostToDevice));
  MLX_TRY(upload_positions<T>(m, m.d_x, xi));
  SweepParams<T> P = sweep_params<T>(m, h, nullptr);
  P.flags = m.d_flags.as<uint8_t>();
  for (int c = 0; c < m.numColors; ++c) {
    P.begin = m.colorOff[c];
    P.end = m.colorOff[c] + m.colorCnt[c];
    if (P.end <= P.begin) continue;
    hipLaunchKernelGGL(k_neohookean2d<T>, dim3(blocks_for(P.end - P.begin)), dim3(kBlock), 0,
                       m.stream, P, m.d_nh.as<T>());
  }
  MLX_CHECK_HIP(hipGetLastError());
  MLX_CHECK_HIP(hipStreamSynchronize(m.stream));
  MLX_CHECK_HIP(hipMemcpy(nh.data(), m.d_nh.ptr, sizeof(T) * nh.size(), hipMemcpyDeviceToHost));
  for (int64_t s = 0; s < ld; ++s) {
    const int32_t e = m.perm[s];
    if (e < 0) continue;
    lambdai[e] = (double)nh[2 * ld + s];
    lambdai[nT + e] = (double)nh[3 * ld + s];
  }
  return download_positions<T>(m, m.d_x, xi);
}

}  // namespace mlx

extern "C" int mlx_neohookean_constraints2d(mlx_mesh *m, const uint8_t *isElastic, double h,
                                            const double *elMu, const double *elLambda,
                                            double *lambdai, double *xi) {
  if (!m || m->dim != 2 || !isElastic || !elMu || !elLambda || !lambdai || !xi || !(h > 0)) {
    set_error("mlx_neohookean_constraints2d: bad arguments (2D mesh required)");
    return MLX_ERR_ARG;
  }
  MLX_CHECK_HIP(hipSetDevice(m->device));
  if (m->precision == MLX_FP64)
    return neohookean2d_impl<double>(*m, isElastic, h, elMu, elLambda, lambdai, xi);
  return neohookean2d_impl<float>(*m, isElastic, h, elMu, elLambda, lambdai, xi);
}

// ---------------------------------------------------------------------------
// preconditioned conjugate gradient for the coarse layer (row a13)
// ---------------------------------------------------------------------------

namespace mlx {

__device__ __forceinline__ double wg_sum1024(double v, double *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// solvePCG.m:1-57 in one workgroup (coarse systems are small): SpMV,
// FP64 dot products with a fixed reduction tree, diagonal preconditioner.
// r, z, d, Ap: n-vectors of scratch, every entry written before it is read.
__device__ __forceinline__ void pcg_run(int n, const int32_t *rp, const int32_t *ci,
                                        const double *v, const double *b, const double *pinv,
                                        int iters, double *x, double *r, double *z, double *d,
                                        double *Ap, double *resvec, double *sh) {
  const int t = threadIdx.x;
  double acc = 0;
  for (int i = t; i < n; i += 1024) {
    double s = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) s = s + v[k] * x[ci[k]];
    r[i] = b[i] - s;
    z[i] = pinv[i] * r[i];
    d[i] = z[i];
    acc += r[i] * z[i];
  }
  double rTz = wg_sum1024(acc, sh);
  for (int it = 0; it < iters; ++it) {
    acc = 0;
    for (int i = t; i < n; i += 1024) {
      double s = 0.0;
      for (int k = rp[i]; k < rp[i + 1]; ++k) s = s + v[k] * d[ci[k]];
      Ap[i] = s;
      acc += d[i] * s;
    }
    const double qAp = wg_sum1024(acc, sh);
    const double alpha = (qAp == 0) ? 0.0 : rTz / qAp;
    acc = 0;
    for (int i = t; i < n; i += 1024) {
      x[i] = x[i] + alpha * d[i];
      r[i] = r[i] - alpha * Ap[i];
      z[i] = pinv[i] * r[i];
      acc += r[i] * z[i];
    }
    const double rTznew = wg_sum1024(acc, sh);
    const double beta = rTznew / rTz;
    rTz = rTznew;
    for (int i = t; i < n; i += 1024) d[i] = z[i] + beta * d[i];
    __syncthreads();
    if (resvec) {
      acc = 0;
      for (int i = t; i < n; i += 1024) {
        double s = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) s = s + v[k] * x[ci[k]];
        acc += (s - b[i]) * (s - b[i]);
      }
      const double ss = wg_sum1024(acc, sh);
      if (t == 0) resvec[it] = sqrt(ss);
    }
  }
}

__global__ __launch_bounds__(1024) void k_pcg(int n, const int32_t *rp, const int32_t *ci,
                                              const double *v, const double *b,
                                              const double *pinv, int iters, double *x,
                                              double *r, double *z, double *d, double *Ap,
                                              double *resvec) {
  __shared__ double sh[1024];
  pcg_run(n, rp, ci, v, b, pinv, iters, x, r, z, d, Ap, resvec, sh);
}

// The solve of k_pcg on caller-owned device arrays, with the work vectors
// in `work` (4n doubles).  The caller's compressed structure is device
// memory the host cannot inspect, so it is checked here first: colptr starts
// at 0, never decreases and ends at nnz, row indices lie in [0, n).  An
// invalid structure leaves x untouched and sets status[0] = 1 (0 after a
// solve); status may be null.
__global__ __launch_bounds__(1024) void k_pcg_dev(int n, int nnz, const int32_t *rp,
                                                  const int32_t *ci, const double *v,
                                                  const double *b, const double *pinv, int iters,
                                                  double *x, double *work, double *resvec,
                                                  int32_t *status) {
  __shared__ double sh[1024];
  __shared__ int bad;
  if (threadIdx.x == 0) bad = (rp[0] != 0 || rp[n] != nnz);
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 1024)
    if (rp[i + 1] < rp[i]) bad = 1;
  __syncthreads();
  if (!bad)
    for (int k = threadIdx.x; k < nnz; k += 1024)
      if (ci[k] < 0 || ci[k] >= n) bad = 1;
  __syncthreads();
  const int invalid = bad;
  if (threadIdx.x == 0 && status) status[0] = invalid;
  if (invalid) return;
  pcg_run(n, rp, ci, v, b, pinv, iters, x, work, work + n, work + 2 * n, work + 3 * n, resvec,
          sh);
}

}  // namespace mlx

extern "C" int mlx_pcg_solve(int64_t n, const int32_t *colptr, const int32_t *rowind,
                             const double *val, const double *b, const double *pinv,
                             int32_t iterations, double *x, double *resvec) {
  if (n <= 0 || n > (1 << 26) || !colptr || !rowind || !val || !b || !pinv || !x || iterations < 0) {
    set_error("mlx_pcg_solve: bad arguments");
    return MLX_ERR_ARG;
  }
  const int64_t nnz = colptr[n];
  for (int64_t i = 0; i < n; ++i)
    if (colptr[i + 1] < colptr[i]) {
      set_error("mlx_pcg_solve: column pointers must be non-decreasing");
      return MLX_ERR_ARG;
    }
  for (int64_t k = 0; k < nnz; ++k)
    if (rowind[k] < 0 || rowind[k] >= n) {
      set_error("mlx_pcg_solve: row index out of range");
      return MLX_ERR_ARG;
    }
  MLX_TRY(check_device(g_device));
  MLX_CHECK_HIP(hipSetDevice(g_device));
  DevBuf buf;
  const size_t bytes = 4 * (n + 1) + 4 * nnz + 8 * nnz + 8 * (7 * n + iterations + 1) + 64;
  MLX_TRY(buf.reserve(bytes));
  char *p = buf.as<char>();
  int32_t *d_cp = reinterpret_cast<int32_t *>(p);
  int32_t *d_ri = d_cp + (n + 1);
  double *d_v = reinterpret_cast<double *>(p + ((4 * (n + 1) + 4 * nnz + 7) / 8) * 8);
  double *d_b = d_v + nnz, *d_pinv = d_b + n, *d_x = d_pinv + n, *d_r = d_x + n, *d_z = d_r + n,
         *d_d = d_z + n, *d_Ap = d_d + n, *d_res = d_Ap + n;
  hipError_t e = hipMemcpy(d_cp, colptr, 4 * (n + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_ri, rowind, 4 * nnz, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_v, val, 8 * nnz, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_b, b, 8 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_pinv, pinv, 8 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_x, x, 8 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    // A is symmetric: the column-compressed arrays of a Matlab sparse are its
    // row-compressed arrays too.
    hipLaunchKernelGGL(k_pcg, dim3(1), dim3(1024), 0, 0, (int)n, d_cp, d_ri, d_v, d_b, d_pinv,
                       iterations, d_x, d_r, d_z, d_d, d_Ap, resvec ? d_res : nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(x, d_x, 8 * n, hipMemcpyDeviceToHost);
  if (e == hipSuccess && resvec && iterations > 0)
    e = hipMemcpy(resvec, d_res, 8 * (size_t)iterations, hipMemcpyDeviceToHost);
  buf.release();
  if (e != hipSuccess) {
    set_error("mlx_pcg_solve: %s", hipGetErrorString(e));
    return MLX_ERR_HIP;
  }
  return MLX_OK;
}

// Allocation-free variant of mlx_pcg_solve on caller-owned device arrays.
extern "C" int mlx_pcg_solve_noalloc(int64_t n, int64_t nnz, const int32_t *colptrDev,
                                     const int32_t *rowindDev, const double *valDev,
                                     const double *bDev, const double *pinvDev,
                                     int32_t iterations, double *xDev, double *resvecDev,
                                     int32_t *statusDev, void *stream) {
  if (n <= 0 || n > kPcgScratchMaxN) {
    set_error("mlx_pcg_solve_noalloc: n = %lld outside 1..%lld (the size of its static "
              "scratch array)", (long long)n, (long long)kPcgScratchMaxN);
    return MLX_ERR_ARG;
  }
  if (nnz < 0 || nnz > INT32_MAX || !colptrDev || (nnz > 0 && (!rowindDev || !valDev)) ||
      !bDev || !pinvDev || !xDev || iterations < 0) {
    set_error("mlx_pcg_solve_noalloc: bad arguments");
    return MLX_ERR_ARG;
  }
  int dev = -1;
  MLX_CHECK_HIP(hipGetDevice(&dev));
  double *work = nullptr;
  MLX_TRY(pcg_scratch_resolve(dev, &work));
  hipLaunchKernelGGL(k_pcg_dev, dim3(1), dim3(1024), 0, static_cast<hipStream_t>(stream),
                     (int)n, (int)nnz, colptrDev, rowindDev, valDev, bDev, pinvDev, iterations,
                     xDev, work, resvecDev, statusDev);
  MLX_CHECK_HIP(hipGetLastError());
  return MLX_OK;
}
