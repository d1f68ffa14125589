/*
 * mlxpbd.h -- C ABI of the MI355X-native multi-layer XPBD substep.
 *
 * Drop-in boundary for the hot path of AlexandreMercierAubin/MultiLayerSolver2024:
 * the entry points below replace, one for one, the Matlab functions and MEX
 * files cited next to them, so that a MEX gateway (or ctypes) can route the
 * reference's simulate/simulate3D scenes through the GPU without touching the
 * scene scripts, SimulationSettings or the layer logic around them.
 *
 * Conventions (those of the Matlab call sites):
 *   - arrays are column-major; per-element matrices are pages (M(i,j,e) at
 *     i + n*j + n*n*e);
 *   - index arrays are int32 and 1-based (vertex, DOF, color, body ids);
 *   - DOFs are xyz (xy in 2D) interleaved: vertex v owns DOFs dim*v-dim+1..dim*v;
 *   - logical arrays are uint8 (0/1);
 *   - all pointers are host pointers unless the name says "dev".
 * Every function returns MLX_OK (0) or a negative status; mlx_last_error()
 * returns a message for the last failure on the calling thread.  No function
 * calls back into Matlab.  Calls on one mlx_mesh are not reentrant; each
 * mesh owns one HIP stream and every call returns after that stream is idle.
 */
#ifndef MLXPBD_H
#define MLXPBD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MLX_OK 0
#define MLX_ERR_ARG -1          /* bad argument (null, size, range) */
#define MLX_ERR_HIP -2          /* HIP runtime failure */
#define MLX_ERR_NOMEM -3        /* allocation failure */
#define MLX_ERR_UNSUPPORTED -4  /* input outside what the GPU path implements */
#define MLX_ERR_NODEVICE -5     /* no gfx950 device visible */
#define MLX_ERR_STATE -6        /* call sequence error */

#define MLX_FP64 0              /* all arithmetic in double (parity build) */
#define MLX_FP32 1              /* bulk projection in float, rigid layer and
                                   residual reductions in double */

#define MLX_MAX_LAYERS 16

const char *mlx_last_error(void);
int mlx_version(void);

/* Selects the HIP device used by meshes created afterwards (default 0) and
 * checks that it is a gfx950 part. */
int mlx_init(int32_t device);

/* ------------------------------------------------------------------ */
/* Scene precompute (host)                                              */
/* ------------------------------------------------------------------ */

/* Replaces naiveGraphColoring(T) (2d/tools/functionalities/naiveGraphColoring.m:1-19).
 * t: nT x vpe int32 1-based, column-major (mesh.t).  elementColor: nT, 1-based.
 * numColors: max color.  Identical output to the reference on identical input. */
int mlx_naive_graph_coloring(const int32_t *t, int64_t nT, int32_t vpe,
                             int64_t nV, int32_t *elementColor,
                             int32_t *numColors);

/* Replaces the precompute of the Mesh3D constructor on the XPBD path:
 * computeB3D (3d/tools/functionalities/computeB3D.m:18-29) -> DmInv,
 * makeElements3D (3d/@Mesh3D/makeElements3D.m:11) -> elV (signed volume
 * det(Dm)/6), computeMass (3d/@Mesh3D/computeMass.m:50-65) -> mass,
 * updateMaterials (3d/@Mesh3D/Mesh3D.m:445-460) -> alpha = inv(K),
 * beta = alpha1*K.  p: 3nV positions; per-element rho, mu, lambda, alpha1.
 * Outputs: DmInv 3x3xnT, elV nT, mass 3nV, alpha/beta 6x6xnT. */
int mlx_mesh3d_precompute(const double *p, int64_t nV, const int32_t *t,
                          int64_t nT, const double *rho, const double *mu,
                          const double *lambda, const double *alpha1,
                          double *DmInv, double *elV, double *mass,
                          double *alpha, double *beta);

/* Replaces the 2D Mesh constructor precompute: computeB2D (2d/@Mesh/Mesh.m:439-469),
 * makeElements area (Mesh.m:481-492), computeMass (2d/@Mesh/computeMass.m:1-18),
 * updateMaterials (Mesh.m:247-258).  Outputs DmInv 2x2xnT, elA nT, mass 2nV,
 * alpha/beta 3x3xnT. */
int mlx_mesh2d_precompute(const double *p, int64_t nV, const int32_t *t,
                          int64_t nT, const double *rho, const double *mu,
                          const double *lambda, const double *alpha1,
                          double *DmInv, double *elA, double *mass,
                          double *alpha, double *beta);

/* ------------------------------------------------------------------ */
/* Layer building natives                                               */
/* ------------------------------------------------------------------ */

/* Replaces mexEdiffNorm3D(Fa, Fb, h) (3d/mex/mexEdiffNorm3D.cpp:25-61).
 * Fa, Fb: 9nT (mesh3D.B*p layout).  out: nT.  Runs on the GPU; bit-identical
 * to the same expression evaluated without fused multiply-adds. */
int mlx_ediff_norm3d(const double *Fa, const double *Fb, int64_t nT, double h,
                     double *out);
/* Replaces mexEdiffNorm2D(Fa, Fb, h) (2d/mex/mexEdiffNorm2D.cpp:25-52). */
int mlx_ediff_norm2d(const double *Fa, const double *Fb, int64_t nT, double h,
                     double *out);

/* Replaces mexUnionFindByVertice(strainInds, sortedLayers, forcedElastic, t, N)
 * (2d/mex/mexUnionFindByVertice.cpp:117-178).  strainInds: nT 1-based element
 * ids in strain-rate order; sortedLayers: L ascending percentages;
 * forcedElastic: nT logical; t: nT x vpe 1-based column-major.  Outputs:
 * rigidBodySetsPerLayer nT x L column-major (0 = elastic, else 1-based body id
 * in ascending element order), numRigids L.  Bit-identical maps. */
int mlx_union_find_by_vertice(const int32_t *strainInds, int64_t nT,
                              const double *sortedLayers, int32_t L,
                              const uint8_t *forcedElastic, const int32_t *t,
                              int32_t vpe, int64_t nV,
                              int32_t *rigidBodySetsPerLayer,
                              int32_t *numRigids);

/* Replaces xpbdLayer2D.findBoundaryVertices (2d/integrator/xpbdLayer2D.m:551-569),
 * also used by 3D (3d/integrator/xpbdLayer3D.m:652).  Outputs nV x L column-major. */
int mlx_find_boundary_vertices(const int32_t *t, int64_t nT, int32_t vpe,
                               int64_t nV, int32_t L,
                               const int32_t *rigidBodySetsPerLayer,
                               uint8_t *isBoundaryVertex,
                               uint8_t *isRigidVertex,
                               int32_t *vertexRigidBody);

/* ------------------------------------------------------------------ */
/* Device-resident scene                                                */
/* ------------------------------------------------------------------ */

typedef struct mlx_mesh mlx_mesh;

/* The static arrays layerSolve receives (3d/integrator/xpbdLayer3D.m:44-93,
 * 2d/integrator/xpbdLayer2D.m:606-641).  alpha and beta must have the
 * isotropic StVK-Voigt form that Mesh3D.updateMaterials / Mesh.updateMaterials
 * build (alpha = inv(K), beta = alpha1*K); anything else is rejected with
 * MLX_ERR_UNSUPPORTED.  Masses must be equal on the dim DOFs of a vertex. */
typedef struct {
  int32_t dim;                 /* 2 or 3 */
  int64_t nV, nT;
  const int32_t *t;            /* nT x (dim+1), 1-based */
  const double *DmInv;         /* dim x dim x nT */
  const double *alpha;         /* nc x nc x nT (nc = 3 in 2D, 6 in 3D) */
  const double *beta;          /* nc x nc x nT */
  const double *volume;        /* nT: mesh3D.elV / mesh2d.elA */
  const double *mass;          /* dim*nV */
  const int32_t *elementColor; /* nT, 1-based */
  const int32_t *pinnedDOFs;   /* 1-based, ascending */
  int64_t nPinnedDOFs;
  const uint8_t *forcedElastic;/* nT: pinned elements (pinnedTets) */
  /* Optional (3D): elements projected with the NeoHookean pair instead of
   * StVK-Voigt -- the deviatoric C_D = ||F||_F (compliance 1/(vol h^2 mu))
   * and volumetric C_H = det F - (1 + mu/lambda) (compliance
   * 1/(vol h^2 lambda)) constraints of mexNeoHookeanDConstraint.cpp:46-138 /
   * mexNeoHookeanHConstraint.cpp:47-138 (xpbdLayer2D.neoHookeanConstraints,
   * 2d/integrator/xpbdLayer2D.m:260-294) lifted to tetrahedra, both
   * evaluated at the positions the projection starts from, displacements
   * summed.  No counterpart in the reference's 3D integrator, which projects
   * StVK-Voigt only (3d/integrator/xpbdLayer3D.m:318,347-379).  Their
   * multipliers are lambdai(:,1:2); alpha/beta of such elements are not
   * read.  neohookean: nT logical or NULL (all StVK-Voigt); mu, lambda: nT
   * Lame parameters (mesh.elMu / elLambda), read for the flagged elements. */
  const uint8_t *neohookean;
  const double *mu;
  const double *lambda;
} mlx_mesh_desc;

int mlx_mesh_create(const mlx_mesh_desc *desc, int32_t precision,
                    mlx_mesh **out);
int mlx_mesh_destroy(mlx_mesh *mesh);
/* Number of graph colors and color bucket sizes (for tuning and tests). */
int mlx_mesh_info(const mlx_mesh *mesh, int32_t *numColors,
                  int64_t *bucketSizes /* numColors, may be NULL */);

/* Replaces xpbdLayer3D.stvkVoigtConstraints(numColors, elementDOFs, elementColor,
 * DmInv, alpha/h^2, mass, volume, lambdai, xi, elasticElements, beta*h^2, oldp, h)
 * (3d/integrator/xpbdLayer3D.m:347-379) and its 2D twin
 * (2d/integrator/xpbdLayer2D.m:61-89): one Gauss-Seidel sweep over all colors
 * (elements flagged neohookean in the mesh take the NeoHookean pair).
 * isElasticElement: nT logical.  lambdai: nT x nc column-major, in/out.
 * xi: dim*nV in/out.  oldp: dim*nV. */
int mlx_stvk_voigt_constraints(mlx_mesh *mesh, const uint8_t *isElasticElement,
                               double h, double *lambdai, double *xi,
                               const double *oldp);

/* Replaces xpbdLayer2D.neoHookeanConstraints(mesh2d, lambdai, xi,
 * elasticElements, h) (2d/integrator/xpbdLayer2D.m:260-294) with its per-element
 * MEX pair mexNeoHookeanDConstraint / mexNeoHookeanHConstraint
 * (2d/mex/mexNeoHookeanDConstraint.cpp:46-138, mexNeoHookeanHConstraint.cpp:47-138):
 * one sweep over all colors of the deviatoric (||F||, compliance 1/(A mu h^2))
 * and volumetric (det F - (1 + mu/lambda), compliance 1/(A lambda h^2))
 * constraints, both evaluated at the positions each color starts from.
 * 2D mesh only.  elMu, elLambda: nT Lame parameters (mesh2d.elMu, elLambda).
 * lambdai: nT x 2 column-major (lambda_D, lambda_H), in/out.  xi: 2nV in/out. */
int mlx_neohookean_constraints2d(mlx_mesh *mesh, const uint8_t *isElasticElement,
                                 double h, const double *elMu,
                                 const double *elLambda, double *lambdai,
                                 double *xi);

/* Arguments of layerSolve (3d/integrator/xpbdLayer3D.m:44-93; 2D
 * 2d/integrator/xpbdLayer2D.m:606-641), minus those already in the mesh. */
typedef struct {
  int32_t nLayers;                /* numel(layers) */
  const int32_t *layersMap;       /* 1-based map column per layer (setLayers) */
  const int32_t *iterations;      /* per layer */
  int32_t nSorted;                /* columns of the map arrays below */
  const uint8_t *isElasticElement;/* nT x nSorted */
  const int32_t *vertexRigidBody; /* nV x nSorted (0 = not rigid) */
  const uint8_t *isBoundaryVertex;/* nV x nSorted */
  const uint8_t *isRigidVertex;   /* nV x nSorted */
  const int32_t *numRigids;       /* nSorted */
  double h;
  double boundaryCompliance;
  double contactResistance;
  int32_t computeResiduals;
  int32_t hangingStop;
  int32_t giveUpEnabled;
  double giveUpThreshold;
  int32_t lastLayerGiveUpEnabled;
  int32_t useRigidContacts;
  double stopAtImprovementPercent;
  double runUntilResSmallerThan;  /* value on entry */
  double residualArrayLast;       /* residualArray(end) on entry (0) */
  double *xi;                     /* dim*nV in/out */
  double *lambdai;                /* nT x nc in/out */
  const double *oldp;             /* dim*nV (cache.oldp / mesh2d.prevp) */
  const double *pinnedx;          /* nPinnedDOFs (3D; 2D pins to oldp) */
  double *residualv;              /* 3nV in/out (3D only) */
  int32_t nContacts;
  const int32_t *contactVertexID; /* K, 1-based */
  const double *contactNormals;   /* K x dim */
  const double *pointColliders;   /* K x dim */
  const double *contactCompliance;/* K */
} mlx_layer_solve_args;

typedef struct {
  double *residualArray;          /* optional, capacity residualCap: appended
                                     residuals (entry 0 = residualArrayLast) */
  int32_t residualCap;
  int32_t residualCount;          /* entries produced, incl. entry 0 */
  double runUntilResSmallerThan;  /* value on exit */
  int32_t itersPerLayer[MLX_MAX_LAYERS];
  int32_t layersRun;
  int64_t elasticProjections;     /* sum over iterations of elastic elements */
  double *contactLambdas;         /* optional K: contact lambdas of the last
                                     layer run (the reference discards them) */
} mlx_layer_solve_out;

/* Replaces xpbdLayer3D.layerSolve / xpbdLayer2D.layerSolve.  Returns final xi
 * and lambdai (the reference discards lambdai; it is returned here). */
int mlx_layer_solve(mlx_mesh *mesh, const mlx_layer_solve_args *args,
                    mlx_layer_solve_out *out);

/* ------------------------------------------------------------------ */
/* Whole substep with device-resident state                             */
/* ------------------------------------------------------------------ */

typedef struct mlx_state mlx_state;

/* Integrator properties (xpbdLayer3D.m:5-40, xpbdLayer2D.m:7-44) used by
 * integrate(). */
typedef struct {
  int32_t nLayers;
  double layers[MLX_MAX_LAYERS];  /* as passed to setLayers */
  int32_t iterations[MLX_MAX_LAYERS];
  double gravity;                 /* Integrator.Gravity */
  double airDrag;                 /* 3D only */
  double boundaryCompliance;
  double contactResistance;
  double restitution;             /* contactVelocityUpdate: 1e-2 (3D), 1e-5 (2D) */
  int32_t computeResiduals;
  int32_t hangingStop;
  int32_t giveUpEnabled;
  double giveUpThreshold;
  int32_t lastLayerGiveUpEnabled;
  int32_t useRigidContacts;
  double stopAtImprovementPercent;
  double runUntilResSmallerThan;  /* in/out, persists like the property */
} mlx_settings;

/* Plane contact finder (3d/contact/PlaneContactFinder3D.m:34-75,
 * 2d/contact/PlaneContactFinder.m:22-59), evaluated on the device each substep
 * at the start-of-step positions (3d/simulate3D.m:163). normal need not be
 * unit length; D = dot(normal, position) as in the reference constructor. */
typedef struct {
  int32_t enabled;
  double normal[3];
  double position[3];
  double compliance;              /* ContactFinder.xpbdContactCompliance (1e-4) */
} mlx_plane;

/* Contact finders of a 3D substep (mlx_integrate_finders), evaluated on the
 * device at the start-of-step positions; their contacts are concatenated
 * in finder order, each finder's in ascending vertex order.  A vertex
 * claimed twice (by two finders) is refused with MLX_ERR_UNSUPPORTED.
 *   MLX_CONTACT_PLANE: PlaneContactFinder3D (3d/contact/PlaneContactFinder3D.m:34-75),
 *     normal and position as in mlx_plane.
 *   MLX_CONTACT_BOX: PlatformContactDetector (3d/contact/PlatformContactDetector.m:65-120):
 *     center, axes (3 unit rows: the rows of the box's rotation) and
 *     halfExtent (width, depth, height halves); a vertex inside the box
 *     gets phi = -(distance to the nearest face) and that face's outward
 *     normal (the first axis on ties).
 *   MLX_CONTACT_SDF: a collider's signed distance sampled on a regular grid
 *     (values x fastest, dims[0] x dims[1] x dims[2], sample (i,j,k) at
 *     origin + spacing (i,j,k)), the samples in place of gptoolbox's
 *     signed_distance query (3d/contact/MeshSCD3D.m:139): trilinear value
 *     and gradient in the vertex's cell, a contact where the value is
 *     negative, normal the unit gradient, none outside the grid.
 * The collider point is x - normal phi (contactInfo3D.m:34-38). */
#define MLX_CONTACT_PLANE 0
#define MLX_CONTACT_BOX 1
#define MLX_CONTACT_SDF 2
#define MLX_MAX_FINDERS 8
typedef struct {
  int32_t kind;
  double compliance;              /* xpbdContactCompliance */
  double normal[3], position[3];  /* plane */
  double center[3], axes[9], halfExtent[3];  /* box */
  const double *sdf;              /* sdf: host array of dims[0]*dims[1]*dims[2] */
  int32_t dims[3];
  double origin[3], spacing;
} mlx_contact_finder;

/* p, v: dim*nV initial state (mesh.p, mesh.v). */
int mlx_state_create(mlx_mesh *mesh, const double *p, const double *v,
                     mlx_state **out);
int mlx_state_destroy(mlx_state *state);
/* Copies state to the host; any pointer may be NULL.  lambdai: nT x nc of the
 * last substep; contacts: number of active contacts of the last substep. */
int mlx_state_download(mlx_state *state, double *p, double *v,
                       double *lambdai, int32_t *nContacts);
/* Overwrites the state between substeps; any pointer may be NULL (kept).
 * p, v, oldp: dim*nV; oldp is the integrator's position history
 * (xpbdLayer3D's cache.oldp, the start-of-substep positions of the last
 * substep, xpbdLayer3D.m:629-635 / :660); lambdai: nT x nc multipliers of
 * the last substep.  With mlx_state_get_oldp it restarts a run mid-way on
 * another mesh object (another precision, a learned mode) from the same
 * state; continuing with frame >= 2 keeps oldp and pinnedx (set the
 * latter with mlx_state_set_pinned_targets). */
int mlx_state_set(mlx_state *state, const double *p, const double *v, const double *oldp,
                  const double *lambdai);
int mlx_state_get_oldp(mlx_state *state, double *oldp);

/* One xpbdLayer3D.integrate (3d/integrator/xpbdLayer3D.m:621-745) or
 * xpbdLayer2D.integrate (2d/integrator/xpbdLayer2D.m:793-917) substep, with
 * the simulate3D/simulate bookkeeping around it (contacts at start of step,
 * position history shift).  frame is 1-based; frame 1 freezes pinnedx and
 * cache.oldp like the reference. */
int mlx_integrate(mlx_state *state, mlx_settings *settings, double h,
                  int32_t frame, const mlx_plane *plane,
                  mlx_layer_solve_out *out);

/* mlx_integrate with the contact finders above (3D; nFinders 0..8) in
 * place of the single plane. */
int mlx_integrate_finders(mlx_state *state, mlx_settings *settings, double h, int32_t frame,
                          const mlx_contact_finder *finders, int32_t nFinders,
                          mlx_layer_solve_out *out);
/* The finders' contacts at the state's current positions, as the next
 * substep would find them: K, then per contact (cap at most) the 0-based
 * vertex, normal (3), collider point (3) and compliance.  Any output but K
 * may be NULL. */
int mlx_find_contacts(mlx_state *state, const mlx_contact_finder *finders, int32_t nFinders,
                      int32_t *K, int32_t cap, int32_t *vertex, double *normals, double *points,
                      double *compliance);

/* Prescribed values of the pinned DOFs (nPinnedDOFs, in pinnedDOFs order)
 * for the following substeps of frame >= 2 (3D).  No counterpart in the
 * reference, which freezes pinnedx at frame 1 (xpbdLayer3D.m:629-632); it
 * drives scripted pins such as a twisted end cap.  Frame 1 still freezes
 * pinnedx from the current positions. */
int mlx_state_set_pinned_targets(mlx_state *state, const double *pinnedx);

/* ------------------------------------------------------------------ */
/* Coarse-layer linear solve                                            */
/* ------------------------------------------------------------------ */

/* Preconditioned conjugate gradient of solvePCG(iterations, A, b, x0,
 * preconditioner) (2d/tools/solver/solvePCG.m:1-57) with the diagonal
 * preconditioner z = pinv .* r, on the GPU in FP64 (one workgroup: sized for
 * coarse-layer systems).  A: n x n symmetric, given by the compressed arrays
 * of a Matlab sparse (mxGetJc colptr n+1, mxGetIr rowind, 0-based as Matlab
 * stores them, values).  x: x0 in, solution out.  resvec (optional,
 * iterations): ||A x - b|| after each iteration.  Not on the reference's
 * XPBD path, which solves its coarse layer by per-body accumulation. */
int mlx_pcg_solve(int64_t n, const int32_t *colptr, const int32_t *rowind,
                  const double *val, const double *b, const double *pinv,
                  int32_t iterations, double *x, double *resvec);

/* Allocation-free variant of mlx_pcg_solve.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Every array is caller-owned DEVICE memory on the current HIP device, with
 * the meaning it has in mlx_pcg_solve (colptr n+1 and rowind nnz 0-based,
 * val nnz, b, pinv, x n; resvecDev iterations or NULL).  The work vectors
 * live in a static device array of the library (4 x 65,536 doubles, 2 MiB
 * per device, held until the process ends): n > 65,536 returns MLX_ERR_ARG.
 * The call only enqueues one kernel on `stream` (a hipStream_t; NULL = the
 * null stream) and returns; results are ready when the stream reaches it.
 * It calls no hipMalloc/hipFree, no memset and no synchronization.  The
 * compressed structure is checked on the device before the solve: an
 * invalid one leaves x untouched and sets *statusDev = 1 (0 after a solve;
 * statusDev may be NULL).  All calls on one device share its scratch array:
 * calls that overlap in time on one device (other streams, other threads)
 * race and must be serialized by the caller. */
int mlx_pcg_solve_noalloc(int64_t n, int64_t nnz, const int32_t *colptrDev,
                          const int32_t *rowindDev, const double *valDev,
                          const double *bDev, const double *pinvDev,
                          int32_t iterations, double *xDev, double *resvecDev,
                          int32_t *statusDev, void *stream);

/* ------------------------------------------------------------------ */
/* Many independent scenes in one launch                                */
/* ------------------------------------------------------------------ */

typedef struct mlx_batch mlx_batch;

/* merged: the concatenation of nScenes 3D scenes in the layout of
 * mlx_mesh_desc, as Mesh3D.mergeMesh builds it (3d/@Mesh3D/Mesh3D.m:300-402:
 * vertex ids offset per scene, each scene keeps its own coloring, pinned DOFs
 * ascending).  sceneVertOff / sceneElemOff: nScenes+1 offsets (0-based).
 * Each scene is integrated independently, with its own residual and stop
 * rule; no extension of the reference (whose merged mesh shares one
 * residual) beyond that.  Plain method only (every layer 0 %% rigid).  A
 * scene is all StVK-Voigt or all NeoHookean (mlx_mesh_desc.neohookean);
 * a scene that mixes them is rejected with MLX_ERR_UNSUPPORTED. */
int mlx_batch_create(const mlx_mesh_desc *merged, int32_t nScenes,
                     const int64_t *sceneVertOff, const int64_t *sceneElemOff,
                     int32_t precision, mlx_batch **out);
int mlx_batch_destroy(mlx_batch *batch);
/* Threads per scene workgroup: 256, 512 (default), 768 or 1024. */
int mlx_batch_set_block(mlx_batch *batch, int32_t threads);
/* Rigid clusters: isRigidElement (nT, merged element order) marks the
 * elements that are rigid in the coarse layer (pinned elements stay
 * elastic); elements sharing a vertex form one body, numbered by ascending
 * smallest element as mexUnionFindByVertice numbers them
 * (2d/mex/mexUnionFindByVertice.cpp:117-178), at most 8 bodies per scene.
 * Every substep then runs, per scene, the two-layer layerSolve
 * (xpbdLayer3D.m:109-226): the coarse layer with these bodies (restriction,
 * rigid step, boundary accumulation, prolongation) for iterations[0], then
 * the all-elastic layer for iterations[1] and the stop rule; settings must
 * have layers = {x, 0} with x != 0 (the coarse layer's rigid set comes from
 * here, not from the strain-rate percentage) and useRigidContacts off.
 * NULL returns to the plain method.  The layered launch keeps positions in
 * HBM and runs one workgroup per scene. */
int mlx_batch_set_rigid(mlx_batch *batch, const uint8_t *isRigidElement);
/* Block-size autotuner per color bucket: scenes are grouped by their mean
 * elements per color (< 256, < 1024, more), and each group is launched with
 * its own block size on its own stream (forked from and joined to the
 * batch's stream, captured with the rest of the substep).  Every candidate
 * (256, 512, 768, 1024 threads; 256 and 512 with rigid clusters) runs one
 * substep of its group alone from the current state, timed with HIP events;
 * the fastest is kept per group and the state restored.  Outputs (3 each,
 * optional): chosen threads (0: empty group), its kernel ms, scenes per
 * group.  mlx_batch_set_block returns to one launch at one block size. */
int mlx_batch_autotune(mlx_batch *batch, mlx_settings *settings, double h, int32_t frame,
                       const mlx_plane *plane, int32_t *bucketBlocks, float *bucketMs,
                       int32_t *bucketScenes);
/* (Re)starts all scenes from p, v (3nV of the merged mesh). */
int mlx_batch_set_state(mlx_batch *batch, const double *p, const double *v);
int mlx_batch_download(mlx_batch *batch, double *p, double *v, double *lambdai);

typedef struct {
  int32_t *iterations;        /* optional nScenes: iterations run */
  double *lastResidual;       /* optional nScenes */
  int32_t *contacts;          /* optional nScenes */
  int64_t elasticProjections; /* all scenes */
  double kernelMs;            /* duration of the batched solve (a replay: of the whole graph) */
} mlx_batch_out;

/* One integrate() substep of every scene (xpbdLayer3D.m:621-745 per scene),
 * the whole projectGS loop of all scenes in one kernel launch. */
/* Splits large scenes over several workgroups (tetsPerWorkgroup > 0): a
 * scene of n tets gets round(n / tetsPerWorkgroup) workgroups (1..64), all
 * on one XCD (placed by the XCC_ID register), which meet at a per-scene
 * barrier after every color; the whole grid is one cooperative launch (the
 * target grows until every workgroup is resident).  If the workgroups were
 * not dealt over the XCDs as planned, the launch does nothing and the
 * substep runs with one workgroup per scene.  0 (default): one workgroup per
 * scene.  Results follow the same per-element updates either way; the
 * residual is summed in another order (last-bit differences). */
int mlx_batch_set_split(mlx_batch *batch, int32_t tetsPerWorkgroup);

/* With one workgroup per scene, scenes whose positions fit in the
 * workgroup's LDS (3 values per vertex) keep them there from the first
 * color to the last residual (enable = 1, default); 0 keeps every scene's
 * positions in HBM.  Same updates in the same order either way. */
int mlx_batch_set_lds(mlx_batch *batch, int32_t enable);
/* Substeps after frame 1 replay a captured hipGraph of the whole substep
 * (contacts, prologue, the solve, velocity update) when the batch is not
 * split; it is captured again whenever a setting baked into its kernels
 * changes.  enable = 0 launches every kernel directly (the default is 1).
 * mlx_batch_out.kernelMs is the solve kernel's time for directly launched
 * substeps and the whole graph's for a replay.  No counterpart in the
 * reference. */
int mlx_batch_set_graph(mlx_batch *batch, int32_t enable);

int mlx_batch_integrate(mlx_batch *batch, mlx_settings *settings, double h,
                        int32_t frame, const mlx_plane *plane,
                        mlx_batch_out *out);

/* buildLayers + findBoundaryVertices on the device (xpbdLayer3D.m:601-619,
 * xpbdLayer2D.m:551-569), from the strain-rate keys that mexEdiffNorm3D/2D
 * return (nT doubles, element order): stable ascending sort with NaN last,
 * the union-find bodies of mexUnionFindByVertice for each sorted layer
 * (pinned elements of the mesh stay elastic), and the vertex maps.  Outputs
 * have the layouts of mlx_union_find_by_vertice (sets nT x nSorted,
 * numRigids) and mlx_find_boundary_vertices (isBoundary, vertexRigidBody
 * nV x nSorted), column-major; the values are identical to theirs. */
int mlx_layer_maps(mlx_mesh *mesh, const double *keys, const double *sortedLayers,
                   int32_t nSorted, int32_t *sets, int32_t *numRigids, uint8_t *isBoundary,
                   int32_t *vertexRigidBody);

/* ------------------------------------------------------------------ */
/* Measurement                                                          */
/* ------------------------------------------------------------------ */

/* Times the element-sweep launches of the mesh (on = 1) with one pair of
 * HIP events (no system-scope fence) around the back-to-back color launches
 * of each iteration on the mesh's stream, or stops (on = 0).  Iterations
 * run as direct launches while it is on.  mlx_profile_read returns the
 * accumulated time (ms), the number of sweep launches covered and the
 * elastic elements they projected, then clears the totals. */
int mlx_profile_enable(mlx_mesh *mesh, int32_t on);
int mlx_profile_read(mlx_mesh *mesh, double *sweepMs, int64_t *sweepLaunches,
                     int64_t *sweepElements);

/* Host wall time of the mlx_integrate calls on states of this mesh since the
 * last read: contact detection plus layer maps (buildLayers and
 * findBoundaryVertices), the layerSolve proper, and the whole call
 * (including the velocity updates and the final stream synchronization);
 * substeps counts the calls.  Clears the totals. */
int mlx_profile_phases(mlx_mesh *mesh, double *layersMs, double *solveMs, double *totalMs,
                       int64_t *substeps);

/* Time of the persistent-loop launches (MLX_OPT_PERSISTENT) while profiling
 * is on: one HIP event pair around each launch on the mesh's stream.
 * Returns the accumulated ms, launches, elastic elements projected and
 * iterations run, then clears the totals. */
int mlx_profile_read_persistent(mlx_mesh *mesh, double *ms, int64_t *launches,
                                int64_t *elements, int64_t *iterations);

/* Execution options of a mesh.  MLX_OPT_GRAPHS (default 1): iterations of
 * layers without rigid bodies run as replays of captured hipGraphs of 16,
 * 8, 4 and 2 iterations, the largest at most MLX_OPT_GRAPH_ITERS (default
 * 16, 1..256); 0 launches every kernel directly.  Results are the
 * same either way.  MLX_OPT_PERSISTENT (default 0): layers without rigid
 * elements run their iterations inside one cooperative launch per chunk of
 * iterations (colors separated by grid barriers); positions and lambdas
 * follow the same updates as with per-color launches, the residual is
 * summed in another order (last-bit differences).  On the 81,920-tet beam
 * it is slower than graph replays (every barrier's acquire drops the L2),
 * hence off by default.  MLX_OPT_PERSIST_WG: workgroups of that launch
 * (0 = from the largest color, 1..256). */
#define MLX_OPT_GRAPHS 1
#define MLX_OPT_GRAPH_ITERS 2
#define MLX_OPT_PERSISTENT 3
#define MLX_OPT_PERSIST_WG 4
/* MLX_OPT_DEVICE_MAPS (default 1): mlx_integrate builds the layer maps on
 * the device (mlx_layer_maps); 0 downloads the strain-rate keys and builds
 * them on the host (mlx_union_find_by_vertice, mlx_find_boundary_vertices).
 * The maps are identical either way. */
#define MLX_OPT_DEVICE_MAPS 5
/* MLX_OPT_PERSIST_LOCAL (default 1): the persistent loop runs on the
 * workgroups of one XCD (MLX_OPT_PERSIST_WG is then the size of that group;
 * 8 times as many are launched, one group per XCD, and all but one group
 * return at once), which meet at a barrier without the L2 write-back of
 * an agent-scope release (stores acknowledged, one cache invalidate per
 * workgroup).  0: all launched workgroups take part, with agent-scope
 * fences. */
#define MLX_OPT_PERSIST_LOCAL 6
/* MLX_OPT_SWEEP_BLOCK (64, 128 or 256; default 64): threads per workgroup
 * of the per-color sweep launches (default 64: a color of ~2,000 elements
 * spreads over ~32 CUs).  Same results for every value. */
#define MLX_OPT_SWEEP_BLOCK 7
/* MLX_OPT_SWEEP_WG (default 0): a 3D layer with rigid elements whose every
 * color has at most 1,024 elastic elements runs each iteration's sweep of
 * all colors in one workgroup (compacted per-color lists of elastic slots,
 * colors separated by workgroup barriers) instead of one launch per color.
 * Same projections in the same order.  Off by default: on the 81,920-tet
 * beam it is slower (8.62 against 8.31 ms per substep), a workgroup's
 * dependent gather-project-store chain per color costing as much as a
 * pipelined launch. */
#define MLX_OPT_SWEEP_WG 8
/* MLX_OPT_FLOW (default 0; 3D): in a layer without rigid elements each
 * iteration's color sweep, elastic contacts and pins run in one
 * cooperative launch on the workgroups of one XCD, where every chunk of 64
 * slots of a color waits only for the chunks that last updated its
 * vertices instead of for the whole previous color.  Same updates in the
 * same order per vertex as the per-color launches: bit-identical results.
 * Off by default: on the 81,920-tet beam a launch takes 160-185 us for the
 * 40 colors (about 4 us per color against 3.5 us per color launch). */
#define MLX_OPT_FLOW 9
int mlx_mesh_set_option(mlx_mesh *mesh, int32_t option, int64_t value);

/* ---- learned warm start (3D) ----
 * No counterpart in the reference: its substep restarts every multiplier at
 * zero (xpbdLayer3D.m:666) and orders the layers by the strain rate of the
 * last substep (xpbdLayer3D.m:601-619).  A per-element network, 21 inputs
 * -> 64 -> 64 -> 2 outputs with SiLU activations (BF16 weights, FP32
 * accumulation on the matrix cores), reads at the start of each substep
 *   F (9, row-major) of the current positions, (F - F_prev)/h (9) with
 *   F_prev from the previous substep's, log10(|lambda_prev| / (vol h^2 E)
 *   + 1e-12), log10 E and nu (E, nu from the element's material)
 * and gives kappa and c.  MLX_LEARN_WARMSTART: at the start of the last
 * layer (when it has no rigid body), every elastic element whose
 * multipliers are still zero starts from lambda_0 = clamp(kappa, 0,
 * kappaMax) lambda_prev, its vertices displaced by (1/m_a) grad_a C^T
 * lambda_0 (pinned vertices stay); the warm start is kept only if the
 * residual of the equations the projection solves (per constraint of the
 * layer's elastic elements C + alpha~ lambda + gamma grad C (x - x_old),
 * per contact r phi) does not rise over the cold start's, else positions
 * and multipliers are restored exactly.  (The monitor's measure,
 * xpbdLayer3D.m:546-586, cannot judge it: it vanishes wherever
 * |alpha~ lambda| >= |C|, however far lambda overshoots.)  MLX_LEARN_KEYS: each strain-rate key is multiplied by
 * 10^clamp(c, -3, 3) before the layers are built (the rate predicted for
 * the coming substep).  Neither changes the stop rule. */
typedef struct mlx_mlp mlx_mlp;
#define MLX_LEARN_WARMSTART 1
#define MLX_LEARN_KEYS 2

/* A network on the current device: zero weights and biases, identity
 * normalisation and output map. */
int mlx_mlp_create(mlx_mlp **out);
int mlx_mlp_destroy(mlx_mlp *model);
/* Parameters: W1 64x21, b1 64, W2 64x64, b2 64, W3 2x64, b3 2 (row-major,
 * out x in); features are normalised as (x - featMean) * featInvStd (21
 * each), outputs mapped as y = raw * outScale + outShift (2 each).  Resets
 * the optimiser state.  Non-finite values are rejected. */
int mlx_mlp_set(mlx_mlp *model, const float *W1, const float *b1, const float *W2,
                const float *b2, const float *W3, const float *b3, const float *featMean,
                const float *featInvStd, const float *outScale, const float *outShift);
int mlx_mlp_get(mlx_mlp *model, float *W1, float *b1, float *W2, float *b2, float *W3,
                float *b3);
/* Outputs Y (n x 2) of raw features X (n x 21), host arrays. */
int mlx_mlp_infer(mlx_mlp *model, const float *X, int64_t n, float *Y);
/* Gradient of the weighted mean squared error sum_s w_s |y_s - t_s|^2 /
 * sum_s w_s over X (n x 21), targets T (n x 2), weights W (n, or null for
 * ones), in the layout dW1 64x21 | db1 | dW2 | db2 | dW3 2x64 | db3 (7,426
 * floats); *loss the error (optional). */
int mlx_mlp_grad(mlx_mlp *model, const float *X, const float *T, const float *W, int64_t n,
                 float *grad, double *loss);
/* Adam (0.9, 0.999, 1e-8) over shuffled minibatches, the shuffle from
 * seed; epochLoss (optional, epochs) the mean error of each epoch. */
int mlx_mlp_train(mlx_mlp *model, const float *X, const float *T, const float *W, int64_t n,
                  int32_t epochs, int32_t batch, float lr, uint64_t seed, float *epochLoss);

/* Attach a network to a 3D single-scene mesh (null detaches); flags
 * MLX_LEARN_*; kappaMax bounds kappa (> 0).  The model must live on the
 * mesh's device and outlive the attachment. */
int mlx_mesh_set_learned(mlx_mesh *mesh, mlx_mlp *model, int32_t flags, float kappaMax);
/* Warm starts kept and rejected since the last call (then cleared). */
int mlx_learn_stats(mlx_mesh *mesh, int64_t *kept, int64_t *rejected);
/* The last substep's network outputs in element order: lambda_0 (nT x 6)
 * and the key factors 10^c (nT); either may be null. */
int mlx_learn_outputs(mlx_mesh *mesh, double *lambda0, double *keyScale);
/* The network's 21 raw features of every element (nT x 21, element order)
 * from the state as the next mlx_integrate call will see it (positions,
 * the previous substep's positions and multipliers), step h.  Needs a
 * mesh with a network attached. */
int mlx_learn_features(mlx_state *state, double h, float *features);

/* The learned warm start on a batch with rigid clusters (the batched
 * workload; no counterpart in the reference): the network runs on every
 * slot at the start of each substep, and each scene's all-elastic layer
 * starts from its lambda_0 and displacement as mlx_mesh_set_learned's
 * MLX_LEARN_WARMSTART does for a single scene, judged per scene by the
 * same equation residual (kept, or positions and multipliers of that scene
 * restored exactly).  256 or 512 threads per scene.  Null detaches. */
int mlx_batch_set_learned(mlx_batch *batch, mlx_mlp *model, float kappaMax);
/* Warm starts kept and rejected over all scenes since the last call
 * (then cleared). */
int mlx_batch_learn_stats(mlx_batch *batch, int64_t *kept, int64_t *rejected);

/* hipDeviceSynchronize on the library's device. */
int mlx_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif
