/*
 * mlx_oracle.h -- CPU FP64 restatement of the multi-layer XPBD substep of
 * AlexandreMercierAubin/MultiLayerSolver2024.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only to check or to time the CPU
 * baseline.  The product path never links, loads or calls it.
 *
 * Conventions: C arrays, 0-based indices, element arrays row-major
 * (T[e*vpe + k]); small per-element matrices stored like a Matlab page,
 * i.e. column-major (M[i + n*j]).  All arithmetic is double; the file is
 * compiled with -ffp-contract=off so that no FMA changes the expression
 * trees that the reference evaluates.
 */
#ifndef MLX_ORACLE_H
#define MLX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_MAX_LAYERS 16

/* Static scene data (the fields xpbdLayer{2,3}D.layerSolve receives). */
typedef struct {
  int dim;            /* 2 or 3 */
  int nV, nT, vpe;    /* vertices, elements, vertices per element (dim+1) */
  const int32_t *T;   /* nT*vpe, 0-based */
  const double *DmInv;/* nT*dim*dim column-major pages */
  const double *alpha;/* nT*nc*nc (nc = 3 in 2D, 6 in 3D) */
  const double *beta; /* nT*nc*nc */
  const double *vol;  /* nT (area in 2D) */
  const double *mass; /* dim*nV, per DOF */
  const int32_t *color; /* nT, 1-based graph colors */
  int numColors;
  const int32_t *pinnedDOFs; int nPinnedDOFs; /* 0-based, ascending */
  const uint8_t *forcedElastic;                /* nT (pinned elements) */
  /* 3D only, may be NULL: elements projected with the NeoHookean pair
   * instead of StVK-Voigt, and their Lame parameters (nT each) */
  const uint8_t *nhModel;
  const double *nhMu, *nhLam;
} ora_mesh;

typedef struct {
  int nLayers;
  double layers[ORA_MAX_LAYERS];   /* rigid percentage per layer, as given */
  int iterations[ORA_MAX_LAYERS];
  double boundaryCompliance;
  double contactResistance;
  double gravity;
  double airDrag;
  int computeResiduals;
  double runUntil;                 /* runUntilResSmallerThan (in/out) */
  int hangingStop;
  int giveUpEnabled;
  double giveUpThreshold;
  int lastLayerGiveUpEnabled;
  double stopAtImprovementPercent;
  int useRigidContacts;
  double restitution;              /* contactVelocityUpdate e */
} ora_settings;

typedef struct {
  int K;                    /* number of contacts */
  const int32_t *vertex;    /* K, 0-based vertex ids */
  const double *normal;     /* K*dim row-major */
  const double *point;      /* K*dim pointCollider row-major */
  const double *compliance; /* K */
} ora_contacts;

/* Per-substep diagnostics. */
typedef struct {
  double *lambda;          /* out: nT*nc row-major (nc per element), optional */
  double *residualArray;   /* out: capacity cap, entry 0 is the initial 0 */
  int residualCap;
  int residualCount;       /* out */
  int itersPerLayer[ORA_MAX_LAYERS]; /* out */
  int layersRun;           /* out */
  int64_t elasticProjections; /* out: elastic element updates summed */
  int32_t *setsPerLayer;   /* out optional: nT*nSorted column-major */
  int32_t *numRigids;      /* out optional: nSorted */
  double normresv[ORA_MAX_LAYERS]; /* out: norm of the residual velocity each
                                      layer's monitor compares (3D) */
  double runUntilAtEntry;  /* out: runUntilResSmallerThan as layerSolve received it */
  /* in, optional (3D): the learned warm start of the last layer and the
   * strain-rate key factors, applied as the product's learned mode applies
   * them (no reference counterpart); learnKept out: 1 kept, 0 rejected,
   * -1 not applied */
  const double *learnLam0;     /* nT*6 row-major */
  const double *learnKeyScale; /* nT */
  int learnKept;
  double learnRes[2];          /* out: equation residual before / after */
  /* in, optional (3D): the layers' rigid body ids (nT*nSorted, column l for
   * sorted layer l, 0 = elastic, bodies 1-based) used instead of the
   * strain-rate layers: layerSolve with maps handed in, as its signature
   * allows (xpbdLayer3D.m:44-93) */
  const int32_t *givenSets;
} ora_step_out;

/* Threads for the per-color element loop of the 3D sweep and the 3D
 * residual's element terms (OpenMP; default 1).  Results are the serial
 * ones bit for bit: elements of one color share no vertex, and the residual
 * terms are summed in element order. */
void ora_set_threads(int n);

/* ---- precompute (Mesh3D / Mesh constructors) ---- */
int ora_naive_graph_coloring(const int32_t *T, int nT, int vpe, int nV,
                             int32_t *color);
void ora_to_lame(double nu, double E, double *mu, double *lambda);
int ora_precompute3d(int nV, int nT, const double *p, const int32_t *T,
                     const double *rho, const double *mu, const double *lam,
                     const double *alpha1, double *DmInv, double *vol,
                     double *mass, double *alpha, double *beta);
int ora_precompute2d(int nV, int nT, const double *p, const int32_t *T,
                     const double *rho, const double *mu, const double *lam,
                     const double *alpha1, double *DmInv, double *area,
                     double *mass, double *alpha, double *beta);
int ora_pinned_elements(const int32_t *T, int nT, int vpe, int nV,
                        const uint8_t *pinnedVert, uint8_t *forcedElastic);

/* ---- layer building ---- */
void ora_b_times_p3d(const int32_t *T, int nT, const double *DmInv,
                     const double *p, double *F);
void ora_b_times_p2d(const int32_t *T, int nT, const double *DmInv,
                     const double *p, double *F);
void ora_ediff_norm3d(const double *Fa, const double *Fb, int nT, double h,
                      double *out);
void ora_ediff_norm2d(const double *Fa, const double *Fb, int nT, double h,
                      double *out);
void ora_stable_sort_ascend(const double *keys, int n, int32_t *order);
int ora_union_find_by_vertice(const int32_t *order, int nT,
                              const double *sortedLayers, int L,
                              const uint8_t *forcedElastic, const int32_t *T,
                              int vpe, int nV, int32_t *sets,
                              int32_t *numRigids);
void ora_find_boundary_vertices(int nV, int nT, int vpe, const int32_t *T,
                                int L, const int32_t *sets,
                                uint8_t *isBoundary, uint8_t *isRigidVert,
                                int32_t *vertexRigidBody);
int ora_sorted_layers(const double *layers, int n, double *sorted,
                      int32_t *map);

/* ---- single-element kernels ---- */
void ora_stvk_grad3d(const double *DmInv, const double *x, double *gradC,
                     double *C);
void ora_stvk_voigt_constraint3d(const double *x, const double *xprev,
                                 const double *DmInv, const double *alphaT,
                                 const double *lambda, const double *minv,
                                 double volume, const double *betaT, double h,
                                 double *dx, double *dlambda, double *C);
void ora_stvk_grad2d(const double *DmInv, const double *x, double *gradC,
                     double *C);
void ora_stvk_voigt_constraint2d(const double *x, const double *xprev,
                                 const double *DmInv, const double *alphaT,
                                 const double *lambda, const double *minv,
                                 double volume, const double *betaT, double h,
                                 double *dx, double *dlambda, double *C);
void ora_exp_rodrigues(const double *omega, double h, double *R);

/* ---- sweeps ---- */
void ora_stvk_voigt_constraints3d(const ora_mesh *m, const uint8_t *isElastic,
                                  double h, double *lambda, double *xi,
                                  const double *oldp);
void ora_stvk_voigt_constraints2d(const ora_mesh *m, const uint8_t *isElastic,
                                  double h, double *lambda, double *xi,
                                  const double *oldp);
double ora_residual3d(const ora_mesh *m, double h, const double *lambda,
                      const double *xi, const double *oldp,
                      const ora_contacts *c, const double *contactLambda,
                      double contactResistance);
double ora_residual2d(const ora_mesh *m, double h, const double *lambda,
                      const double *xi, const double *oldp,
                      const ora_contacts *c, const double *contactLambda,
                      double contactResistance);

/* ---- contacts ---- */
int ora_plane_contacts(int dim, int nV, const double *p, const double *normal,
                       const double *position, double compliance,
                       int strict, int32_t *vertex, double *cnormal,
                       double *cpoint, double *ccompliance);
void ora_contact_velocity_update(int dim, const ora_contacts *c, double *v,
                                 double h, double gravity, double e);

/* ---- rigid layer on explicit state (test hooks) ---- */
void ora_rigid_properties3d(int nV, int nb, const int32_t *vb, const double *p,
                            const double *v, const double *mass, double *M,
                            double *COM, double *COMDOT, double *R, double *I0,
                            double *I, double *omega, double *ri);
void ora_rigid_step3d(double h, double *COM, const double *COMDOT, double *R,
                      const double *omega, const double *I0, double *I);
void ora_boundary3d(int nb, double *M, double *COM, double *R, double *I0, double *I,
                    int nbv, const int32_t *bverts, const int32_t *vb, const double *ri,
                    const double *mass, double compliance, double *lamB, double *xi,
                    double h);
void ora_contacts3d(int nb, double *M, double *COM, double *R, double *I0, double *I,
                    const ora_contacts *c, const int32_t *vb, const double *ri,
                    const double *mass, double *cl, double *xi, double h,
                    double resistance, int useRigid);

/* ---- 2D NeoHookean D/H pair ---- */
void ora_neohookean_d2d(double h, const double *x, const double *DmInv, double mu,
                        double area, double lambda, const double *mass,
                        double *dx, double *dlambda, double *C);
void ora_neohookean_h2d(double h, const double *x, const double *DmInv, double lamL,
                        double area, double lambda, const double *mass, double gamma,
                        double *dx, double *dlambda, double *C);
void ora_neohookean_constraints2d(const ora_mesh *m, const uint8_t *isElastic, double h,
                                  const double *mu, const double *lamL, double *lambda,
                                  double *xi);

/* NeoHookean pair on a tetrahedron: the 2D reference's deviatoric and
 * volumetric constraints (mexNeoHookeanDConstraint.cpp:46-138,
 * mexNeoHookeanHConstraint.cpp:47-138) with F = Ds DmInv of a tetrahedron,
 * Ds = [x1-x0, x2-x0, x3-x0].  No counterpart in the reference, whose 3D
 * integrator projects StVK-Voigt only.
 * ora_neohookean_grad3d: gD, gH (12 each, d/dx of ||F||_F and det F),
 * CD = ||F||_F, J = det F.
 * ora_neohookean_d3d: C = ||F||_F, compliance 1/(vol h^2 mu);
 * ora_neohookean_h3d: C = det F - gamma, compliance 1/(vol h^2 lambda);
 * both dlambda = -(C + alpha~ lambda) / (sum_a w_a |grad_a C|^2 + alpha~),
 * dx_a = w_a grad_a C dlambda. */
void ora_neohookean_grad3d(const double *DmInv, const double *x, double *gD, double *gH,
                           double *CD, double *J);
void ora_neohookean_d3d(double h, const double *x, const double *DmInv, double mu,
                        double vol, double lambda, const double *mass, double *dx,
                        double *dlambda, double *C);
void ora_neohookean_h3d(double h, const double *x, const double *DmInv, double lamL,
                        double vol, double lambda, const double *mass, double gamma,
                        double *dx, double *dlambda, double *C);
/* One sweep over all colors of the NeoHookean elements (m->nhModel), both
 * constraints of an element evaluated at the positions its color started
 * from and their dx summed, as the 2D pair (ora_neohookean_constraints2d).
 * gamma = 1 + mu/lambda.  lambda: nT x 6 row-major, (lambda_D, lambda_H) in
 * the first two entries. */
void ora_neohookean_constraints3d(const ora_mesh *m, const uint8_t *isElastic, double h,
                                  double *lambda, double *xi);

/* ---- PCG (solvePCG.m) ---- */
void ora_solve_pcg(int n, const int32_t *rowptr, const int32_t *col, const double *val,
                   const double *b, const double *pinv, int iterations, double *x,
                   double *resvec);

/* ---- full substep ---- */
/* 3D: p,v (3nV) in/out; oldp = cache.oldp in/out; oldOldp out;
 * pinnedx (nPinnedDOFs) in/out; frame is 1-based. */
int ora_integrate3d(const ora_mesh *m, ora_settings *s, double h, int frame,
                    double *p, double *v, double *oldp, double *oldOldp,
                    double *pinnedx, const ora_contacts *c,
                    ora_step_out *out);
/* Learned warm start of the last layer (no reference counterpart; the
 * product's learned mode): the residual of the projection's equations over
 * the elastic elements and contacts, the displacement of the fresh elastic
 * elements for multipliers lam0 (nT*6), and both with the keep / restore
 * judge (res: residual before, after). */
double ora_learn_eq_residual3d(const ora_mesh *m, double h, const uint8_t *isEl,
                               const double *lambda, const double *xi, const double *oldp,
                               const ora_contacts *c, double resistance);
void ora_learn_displace3d(const ora_mesh *m, const uint8_t *isEl, const double *lam0,
                          double *lambda, double *xi, const uint8_t *pinnedDof);
int ora_learn_warm_start3d(const ora_mesh *m, double h, const uint8_t *isEl,
                           const double *lam0, double *lambda, double *xi, const double *oldp,
                           const uint8_t *pinnedDof, const ora_contacts *c, double resistance,
                           double *res);
/* Platform box and grid-sampled signed-distance contact finders (3D);
 * return K and fill the contact arrays (capacity nV). */
int ora_box_contacts(int nV, const double *p, const double *center, const double *axes,
                     const double *halfExtent, double compliance, int32_t *vertex,
                     double *cnormal, double *cpoint, double *ccompliance);
int ora_sdf_contacts(int nV, const double *p, const double *grid, const int32_t *dims,
                     const double *origin, double spacing, double compliance, int32_t *vertex,
                     double *cnormal, double *cpoint, double *ccompliance);
/* 2D rigid-layer steps on explicit body state (test hooks): bodies from
 * connectivity, boundary accumulation and contacts (xpbdLayer2D.m:296-444). */
void ora_rigid_properties2d(int nV, int nb, const int32_t *vb, const double *p,
                            const double *mass, double *M, double *COM, double *I,
                            double *ri);
void ora_boundary2d(int nb, double *M, double *COM, double *angle, double *I, int nbv,
                    const int32_t *bverts, const int32_t *vb, const double *ri,
                    const double *mass, double alpha, double *lamB, double *xi, double h);
void ora_contacts2d(int nb, double *M, double *COM, double *angle, double *I,
                    const ora_contacts *c, const int32_t *vb, const double *ri,
                    const double *mass, double *cl, double *xi, double h, double resistance,
                    int useRigid);
/* 2D: prevp/prevPrevp are shifted by the caller (simulate.m:332-334). */
int ora_integrate2d(const ora_mesh *m, ora_settings *s, double h,
                    double *p, double *v, const double *prevp,
                    const double *prevPrevp, const ora_contacts *c,
                    ora_step_out *out);

#ifdef __cplusplus
}
#endif
#endif
