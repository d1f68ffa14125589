// Internal declarations of libmlxpbd: device data layout, kernel launchers
// and the host-side scene/state objects behind the C ABI in mlxpbd.h.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "mlxpbd.h"

struct mlx_mlp;

namespace mlx {

void set_error(const char *fmt, ...);

// Device view of a network (mlx_learn.hip) and the device it lives on.
struct MlpDev;
MlpDev mlp_device_view(const mlx_mlp *M);
int mlp_device_of(const mlx_mlp *M);

#define MLX_CHECK_HIP(call)                                                   \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::mlx::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                   \
      return MLX_ERR_HIP;                                                     \
    }                                                                         \
  } while (0)

#ifndef MLX_TRY
#define MLX_TRY(expr)            \
  do {                           \
    int r_ = (expr);             \
    if (r_ != 0) return r_;      \
  } while (0)
#endif

template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };
template <typename T> using vec4_t = typename Vec4<T>::type;

// Control block shared by the kernels of one layerSolve.  Iteration kernels
// return immediately once endLayer or stopSolve is set, so the host can
// enqueue iterations ahead of the monitor's decision.
struct Ctl {
  int stopSolve;
  int endLayer;
  int iters;          // iterations run in the current layer
  int errorChangeDecreased;
  int residualCount;  // entries in the residual array (entry 0 included)
  int stopLayer;      // layer index where stopSolve was raised, else -1
  int isLastLayer;
  int layer;
  double runUntil;
  double setPercentStop;
  double normresv;
  double lastResidual;
};

// Iteration-invariant parameters of the monitor (simpleMonitorError).
struct MonitorParams {
  int is3d;
  int computeResiduals;
  int hangingStop;
  int giveUpEnabled;
  int lastLayerGiveUpEnabled;
  double giveUpThreshold;
  double h;
};

// Device buffer with grow-only capacity.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  int reserve(size_t n);
  void release();
  template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

}  // namespace mlx

// Scene: static data, laid out color-major ("slots") so that each color's
// elements are contiguous and every per-element field is read coalesced.
struct mlx_mesh {
  int device = 0;
  int dim = 3, vpe = 4, nc = 6, precision = MLX_FP64;
  int64_t nV = 0, nT = 0, ld = 0;  // ld: padded slot count (multiple of 64)
  int numColors = 0;
  hipStream_t stream = nullptr;

  // host copies (element order, 0-based)
  std::vector<int32_t> T;           // nT*vpe
  std::vector<double> DmInv;        // nT*dim*dim pages
  std::vector<double> massV;        // per vertex
  std::vector<int32_t> pinnedDOF;   // 0-based
  std::vector<uint8_t> forcedElastic;
  bool hasNH = false;               // some element is a NeoHookean tetrahedron
  std::vector<int32_t> color;       // 1-based
  std::vector<int64_t> colorOff;    // numColors+1 padded slot offsets
  std::vector<int64_t> colorCnt;    // elements per color
  std::vector<int32_t> perm;        // slot -> element (-1 for padding)
  int nScenes = 1;                  // > 1: scene-major slot layout (batch)
  std::vector<int32_t> sceneColor0, sceneNColors;
  std::vector<int64_t> sceneSlot0;  // nScenes+1
  std::vector<int32_t> slotOf;      // element -> slot
  // per-element material in element order: alpha_d, alpha_o, alpha_s, alpha1
  std::vector<double> mat;

  // device: slot-major SoA arrays (precision T)
  mlx::DevBuf d_vid;      // int32  vpe*ld
  mlx::DevBuf d_dminv;    // T      dim*dim*ld
  mlx::DevBuf d_vol;      // T      ld
  mlx::DevBuf d_mat;      // T      4*ld
  mlx::DevBuf d_lam;      // T      nc*ld
  mlx::DevBuf d_dminv64;  // double nT*dim*dim, element order (layer building)
  mlx::DevBuf d_T;        // int32  nT*vpe element order
  mlx::DevBuf d_minv64;   // double nV (1/mass per vertex)
  mlx::DevBuf d_mass64;   // double nV
  mlx::DevBuf d_pinned;   // int32  nPinnedDOF
  mlx::DevBuf d_slotOf;   // int32  nT

  // work buffers (sized on demand)
  mlx::DevBuf d_x, d_xold;        // vec4<T> nV: position (w = 1/mass), oldp
  mlx::DevBuf d_rv;               // T 3nV residual velocity
  mlx::DevBuf d_pinnedx;          // T nPinnedDOF
  mlx::DevBuf d_flags;            // uint8 nSorted*ld (slot order)
  mlx::DevBuf d_vbody;            // int32 nSorted*nV
  mlx::DevBuf d_ri;               // double 3nV
  mlx::DevBuf d_bodies;           // double per body state
  mlx::DevBuf d_csr;              // int32 per-layer CSR lists
  mlx::DevBuf d_lamB;             // double boundary lambdas
  mlx::DevBuf d_contacts;         // contact arrays
  mlx::DevBuf d_finder;           // contact finders: parameters | sdf samples | flags | scan | codes
  mlx::DevBuf d_partials;         // double reduction partials
  mlx::DevBuf d_ctl;              // Ctl
  mlx::DevBuf d_resid;            // double residual array
  mlx::DevBuf d_tmp;              // misc
  mlx::DevBuf d_count;            // uint32 arrival counter of k_residual_tail (kept at 0)
  mlx::DevBuf d_bodycnt;          // uint32 per-body arrival counters of the chunked restriction (kept at 0)
  mlx::DevBuf d_nh;               // T 4*ld: 2D NeoHookean mu | lambda | lambda_D | lambda_H
  mlx::Ctl *h_ctl = nullptr;      // pinned host mirror
  int32_t *h_counts = nullptr;    // pinned {K, nec, nrc, nUniq} staging for the device counts
  // pinned staging of one substep's host-to-device copies (and of the
  // contact list read back), handed out in order, reclaimed once the
  // stream has drained
  char *h_stage = nullptr;
  size_t stageCap = 0, stageUsed = 0;
  std::vector<char *> stageRetired;  // outgrown staging buffers, freed at stage_begin
  mlx::DevBuf d_pinMask;           // uint8 dim*nV: pinned DOFs (built once)
  mlx::DevBuf d_seam;              // double: the layerSolve seam's arrays in caller layout
  mlx::DevBuf d_elastic;           // uint8 nSorted*nT: caller's elastic flags (element order)

  // replayable graphs of iteration chunks (hipGraph), keyed by everything
  // the captured kernels bake into their parameters
  struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> sig;
  };
  GraphSlot graphs[8];
  int graphNext = 0;
  int useGraphs = 1;
  int graphIters = 16;             // largest captured chunk (MLX_OPT_GRAPH_ITERS)
  int prevExtraIters = 0;          // last layer: iterations past the fixed count, previous call
  // persistent iteration loop (layers without rigid elements)
  int usePersistent = 0;
  int persistWG = 0;               // 0 = from the largest color (local: from the slot count)
  int persistLocal = 1;            // MLX_OPT_PERSIST_LOCAL: one XCD's workgroups, xcd_barrier
  int sweepBlock = 64;             // MLX_OPT_SWEEP_BLOCK: threads per workgroup of a color sweep
  int useSweepWg = 0;              // MLX_OPT_SWEEP_WG: one-workgroup sweeps for sparse layers
  // dataflow sweep of layers without rigid elements (MLX_OPT_FLOW, k_flow)
  int useFlow = 0;
  bool flowReady = false;          // d_need / d_deg built for the mesh
  mlx::DevBuf d_flowdeps;          // int32 chunkBase[numColors+1] | depOff[nChunks+1] | dep
  mlx::DevBuf d_flowflags;         // uint32 nChunks: iterations each chunk has finished
  int flowChunks = 0, flowDeps = 0;
  mlx::DevBuf d_flowsync;          // uint32 64: err | ... | barrier (15) | XCD group counters (16..32)
  int32_t *h_flowerr = nullptr;    // pinned copy of the error word
  mlx::DevBuf d_elist;             // colorOff i64 | colorCnt i64 | count i32 | elastic slot lists i32
  mlx::DevBuf d_colors;            // int32 colorOff[numColors] | colorCnt[numColors]
  // layer maps on the device (mlx_maps.hip)
  int deviceMaps = 1;              // MLX_OPT_DEVICE_MAPS
  mlx::DevBuf d_forced;            // uint8 nT forcedElastic
  mlx::DevBuf d_maps;              // sort keys/values, ranks, parents, labels, outputs
  mlx::DevBuf d_scan;              // rocPRIM temporary storage
  mlx::DevBuf d_sync;              // uint32 barrier counter | int32 error flag
  mlx::DevBuf d_colortab;          // int64 colorOff[numColors] | colorCnt[numColors]

  // sweep timing (mlx_profile_enable)
  struct Prof {
    bool on = false;
    std::vector<hipEvent_t> ev;          // 2 per recorded iteration of sweeps
    std::vector<int64_t> elems;          // elastic elements projected in it
    std::vector<int64_t> nlaunch;        // sweep launches in it
    std::vector<int32_t> iter;           // its iteration index
    size_t used = 0;
    double ms = 0;
    int64_t launches = 0, elements = 0;
    hipEvent_t pev[2] = {nullptr, nullptr};  // persistent launch pair
    double pms = 0;
    int64_t plaunches = 0, pelements = 0, piters = 0;
  } prof;

  // host wall time of mlx_integrate phases (mlx_profile_phases)
  struct Phase {
    double layersMs = 0, solveMs = 0, totalMs = 0;
    int64_t substeps = 0;
  } phase;

  // learned warm start / key correction (mlx_mesh_set_learned, 3D)
  struct Learn {
    mlx_mlp *model = nullptr;
    int flags = 0;
    float kappaMax = 2.0f;
    bool ready = false;     // inference ran for the current substep
    mlx::DevBuf d_lam0;     // T nc*ld: warm-start multipliers (slot-major like d_lam)
    mlx::DevBuf d_keyScale; // double ld: strain-rate key factors 10^c
    mlx::DevBuf d_save;     // vec4<T> nV positions | T nc*ld multipliers (fallback copy)
    mlx::DevBuf d_dx;       // T 12*ld per (vertex, slot) displacements
    mlx::DevBuf d_inc;      // int32 nV+1 offsets | 4*nT (a*ld + slot) incidences
    mlx::DevBuf d_res;      // double 2 (cold, warm) | partials
    mlx::DevBuf d_counts;   // uint64 2: warm starts kept, rejected
    mlx::DevBuf d_feat;     // float 21*ld features (mlx_learn_features)
  } learn;
};

struct mlx_state {
  mlx_mesh *mesh = nullptr;
  int frame_seen = 0;
  mlx::DevBuf d_v;        // T dim*nV velocity
  mlx::DevBuf d_ediff;    // double nT
  std::vector<double> h_ediff;
  int32_t lastContacts = 0;
};

namespace mlx {

// Layer maps of one substep in the reference's layouts, column l for sorted
// layer l: elastic flags (element order), vertex body ids, boundary flags,
// body counts.
struct HostMaps {
  int nSorted = 0;
  std::vector<uint8_t> elastic;   // nSorted x nT (element order)
  std::vector<int32_t> vbody;     // nSorted x nV
  std::vector<uint8_t> boundary;  // nSorted x nV
  std::vector<int32_t> numRigids; // nSorted
  // set by layer_maps_dev(deviceFlags): elastic flags already in m.d_flags
  // (slot order, one column of ld per sorted layer); elastic element counts
  // per (layer, color) and per layer; `elastic` is then left empty
  bool deviceFlags = false;
  std::vector<int64_t> perColor;  // nSorted x numColors
  std::vector<int64_t> nElastic;  // nSorted
  // set by layer_maps_dev: the vertex body ids (nSorted x nV) on the device
  const int32_t *devVbody = nullptr;
};

// buildLayers + findBoundaryVertices on the device from per-element
// strain-rate keys already in device memory (mlx_maps.hip): stable radix
// sort, connected components per sorted layer with bodies numbered by
// ascending minimum element, vertex maps.  Results are copied to `maps`
// (and the per-layer body ids to *setsOut, nT x ns, when given).
int layer_maps_dev(mlx_mesh &m, const double *d_keys, const double *sorted, int ns,
                   HostMaps &maps, std::vector<int32_t> *setsOut = nullptr,
                   bool deviceFlags = false);

// Pinned staging of one substep's copies (mlx_solver.hip): stage_begin
// reclaims every region (draining the stream if regions are in use),
// stage_take hands out a region valid until the next stage_begin,
// stage_upload copies host data through a region to the device on the
// mesh's stream.
int stage_begin(mlx_mesh &m);
int stage_take(mlx_mesh &m, size_t bytes, char **out);
int stage_upload(mlx_mesh &m, void *dst, const void *src, size_t bytes);

// Ordered list of the flagged entries of flag[0..n) (int32 0/1) on stream
// st: list[k] = the k-th flagged index in ascending order, lambda[k] = 0,
// count[0] = number flagged.  pos: n int32 of scratch (the exclusive scan).
int compact_flags_dev(mlx_mesh &m, const int32_t *flag, int64_t n, int32_t *pos, int32_t *list,
                      double *lambda, int32_t *count, hipStream_t st);

}  // namespace mlx
