/*
 * C ABI of the MI355X geometric-multigrid PCG solver for static linear
 * elasticity on voxel hex8 meshes.
 *
 * It replaces the reference's MATLAB solve path (paths relative to the
 * reference's src/):
 *   gmgs_create    common/DiscretizeDesignDomain.m:29-120 + kernel/BuildingMeshHierarchy.m
 *                  + kernel/SetupSolver.m:8-244 (mesh, hierarchy, coarse operators, smoothers)
 *   gmgs_set_bc    common/ApplyBoundaryCondition.m:7-23 + kernel/SetupSolver.m:246-259
 *   gmgs_solve     kernel/GetDeformation.m:18-21 -> kernel/CG_solver.m:1-56 (U_, iterationHist_)
 *   gmgs_stress    kernel/GetCartesianStress.m:1-42 (cartesianStressField_)
 *   gmgs_atx       kernel/AtX.m:1-58          (the @AtX operator handle)
 *   gmgs_vcycle    kernel/Vcycle.m:1-36       (the @Vcycle preconditioner handle)
 *   gmgs_smooth    kernel/JacobiSmoother.m:1-10
 *   gmgs_restrict  kernel/RestrictResidual.m:1-18
 *   gmgs_interpolate kernel/InterpolationDeviation.m:1-18
 *   gmgs_query     read-back of meshHierarchy_ fields (kernel/MeshStruct.m:2-33)
 * gmgs_set_modulus and gmgs_adjoint have no reference counterpart (per-element
 * stiffness scaling and the adjoint gradient).
 *
 * Conventions: host buffers are caller-owned and copied; node ids are the
 * reference's compact 1-based ids; vectors are node-interleaved
 * [ux1 uy1 uz1 ux2 ...] in compact node order; 2-D arrays are column-major
 * (MATLAB layout) unless stated otherwise.  Every call returns a gmgs_status;
 * gmgs_last_error() gives the message.  A context owns one HIP stream on one
 * device and is not re-entrant.
 */
#ifndef GMGS_H
#define GMGS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gmgs_ctx gmgs_ctx;

typedef enum gmgs_status {
  GMGS_OK = 0,
  GMGS_ERR_ARGUMENT = 1,        /* bad pointer, size or parameter */
  GMGS_ERR_RESOLUTION = 2,      /* too few voxels for the hierarchy (DiscretizeDesignDomain.m:30-32) */
  GMGS_ERR_LEVELS = 3,          /* fewer than 2 levels (SetupSolver.m:19-21) */
  GMGS_ERR_NO_LOADS = 4,        /* ApplyBoundaryCondition.m:11-12 */
  GMGS_ERR_NO_CONSTRAINTS = 5,  /* ApplyBoundaryCondition.m:22 */
  GMGS_ERR_NODE_ID = 6,         /* node id out of range or not an integer */
  GMGS_ERR_NONFINITE = 7,       /* NaN/Inf in an input */
  GMGS_ERR_NOT_SPD = 8,         /* coarsest Cholesky failed (ignored by SetupSolver.m:259) */
  GMGS_ERR_NOT_CONVERGED = 9,   /* maxIT_ exceeded (CG_solver.m:52-55); U is still written */
  GMGS_ERR_STATE = 10,          /* call order (no BCs, no solution yet, ...) */
  GMGS_ERR_DEVICE = 11,         /* HIP runtime error / out of device memory */
  GMGS_ERR_BREAKDOWN = 12       /* NaN in the PCG recursion */
} gmgs_status;

enum { GMGS_PREC_FP64 = 0, GMGS_PREC_MIXED = 1 };

enum {
  GMGS_FLAG_EXACT_GALERKIN = 1,   /* true Galerkin coarse operators (no B3/B12 quirks) */
  GMGS_FLAG_NO_INDEX_TABLES = 2,  /* skip eNodMat/transferMat/... read-back tables */
  GMGS_FLAG_FR_BETA = 4,          /* Fletcher-Reeves beta also in mixed precision */
  GMGS_FLAG_PR_BETA = 8,          /* Polak-Ribiere (flexible) beta also in FP64 */
  GMGS_FLAG_SIMPLE_KERNELS = 16,  /* one-wave-per-tile element-sum operator everywhere
                                     (no z-marching stencil kernel); for cross-checks */
  GMGS_FLAG_AUTOTUNE = 32,        /* time column-tile shapes per level at set-up */
  GMGS_FLAG_NO_GRAPHS = 64,       /* launch PCG iterations one kernel at a time */
  GMGS_FLAG_SYM_VCYCLE = 128,     /* symmetric V(nu, nu) cycle: residual recomputed after the
                                     coarse correction, nu = vcycle_sweeps (no reference
                                     counterpart; kernel/Vcycle.m:27-33 reuses the stale one) */
  GMGS_FLAG_NO_JACOBI_STEP = 256, /* omit the Jacobi step opening every PCG iteration
                                     (CG_solver.m:18-19) */
  GMGS_FLAG_CHEBYSHEV = 512,      /* with GMGS_FLAG_SYM_VCYCLE: the nu sweeps of a level use
                                     the reciprocal roots of the degree-nu Chebyshev polynomial
                                     on [0.055, 1.1] x lambda_max(D^-1 A), estimated per level by
                                     20 Jacobi-CG (Lanczos) steps when the BCs are set (no
                                     reference counterpart; JacobiSmoother.m uses one weight) */
  GMGS_FLAG_ASSEMBLED_COARSE = 1024, /* coarse smoothed levels apply an assembled 27-point
                                     stencil (243 coefficients per node, set-up time) wherever it
                                     fits, instead of the z-marching or element-sum kernels */
  GMGS_FLAG_SPLIT_COARSE = 2048,  /* coarse smoothed levels with a universal element: the
                                     z-marching kernel on the nodes whose eight elements are all
                                     present and universal, the assembled rows on the others (one
                                     lane per row); elsewhere as GMGS_FLAG_ASSEMBLED_COARSE.
                                     GMGS_FLAG_AUTOTUNE also times this operator per level */
  GMGS_FLAG_UNFUSED_PROLONG = 4096, /* symmetric V(1, 1) cycle on a z-marching level 1: store
                                     e = P x_c with a prolongation kernel of its own before the
                                     residual kernel reads it, instead of forming each value of e
                                     while the residual kernel stages it (same results; for
                                     cross-checks and timing) */
  GMGS_FLAG_UNFUSED_RESIDUAL = 8192 /* mixed solver without the Jacobi step, symmetric cycle on
                                     a z-marching level 1: update the PCG residual
                                     r = r - alpha q (and its norm) in a pointwise kernel of its
                                     own after the matvec, instead of in the next iteration's
                                     first level-1 sweep, which stages r - alpha q (same
                                     iterates; the norm's partial sums are grouped
                                     differently) */
};

typedef struct gmgs_params {
  double youngs_modulus;               /* moduls_            GlobalVariables.m:8  */
  double poisson_ratio;                /* poissonRatio_      GlobalVariables.m:9  */
  double tol;                          /* tol_               GlobalVariables.m:3  */
  double jacobi_weight;                /* weightFactorJacobi_ GlobalVariables.m:5 */
  double coarsest_resolution_control;  /* coarsestResolutionControl_ GlobalVariables.m:23 */
  int32_t max_it;                      /* maxIT_             GlobalVariables.m:4  */
  int32_t num_levels;                  /* 0 = level rule of DiscretizeDesignDomain.m:33-41 */
  int32_t precision;                   /* GMGS_PREC_FP64 or GMGS_PREC_MIXED */
  int32_t flags;                       /* GMGS_FLAG_* */
  int32_t device;                      /* HIP device ordinal */
  int32_t vcycle_sweeps;               /* nu of GMGS_FLAG_SYM_VCYCLE on level 1: 1..4 (0 = 2) */
  int32_t vcycle_sweeps_coarse;        /* nu on levels 2..numLevels-1: 1..4 (0 = vcycle_sweeps) */
  int32_t vcycle_sweeps_level2;        /* nu on level 2 alone: 1..4 (0 = vcycle_sweeps_coarse) */
  int32_t reserved[4];
} gmgs_params;

typedef struct gmgs_solve_info {
  int32_t its;             /* CG iterations (CG_solver.m:22) */
  int32_t converged;       /* relres < tol reached */
  double relres;           /* recursive ||r||/||b|| at exit (CG_solver.m:35) */
  double true_relres;      /* ||b - A U||/||b|| recomputed in FP64 at exit */
  double solve_seconds;    /* device wall time of the PCG loop */
  int64_t hist_len;        /* rows of iterationHist_ (non-final iterations) */
} gmgs_solve_info;

/* query keys (level = 1..numLevels, 1 = finest) */
enum {
  GMGS_Q_NUM_LEVELS = 0,        /* int64 */
  GMGS_Q_RES = 1,               /* int64[3] resX resY resZ */
  GMGS_Q_NUM_ELEMENTS = 2,      /* int64 */
  GMGS_Q_NUM_NODES = 3,         /* int64 */
  GMGS_Q_ELE_MAP_BACK = 4,      /* int32[numElements] */
  GMGS_Q_ELE_MAP_FORWARD = 5,   /* int32[resX*resY*resZ] */
  GMGS_Q_NOD_MAP_BACK = 6,      /* int32[numNodes] (finest-grid ids on levels >= 2) */
  GMGS_Q_NOD_MAP_FORWARD = 7,   /* int32[(resX+1)(resY+1)(resZ+1)] */
  GMGS_Q_ENODMAT = 8,           /* int32[numElements x 8] */
  GMGS_Q_NUM_NOD2ELES = 9,      /* int32[numNodes] */
  GMGS_Q_NODES_ON_BOUNDARY = 10,    /* int32[n] */
  GMGS_Q_ELEMENTS_ON_BOUNDARY = 11, /* int32[n] */
  GMGS_Q_ELEMENT_UPWARD_MAP = 12,   /* int32[numElements x 8] (level >= 2) */
  GMGS_Q_TRANSFER_MAT = 13,         /* int32[27 x numElements] */
  GMGS_Q_TRANSFER_MAT_COEFFI = 14,  /* double[intermediateNumNodes] */
  GMGS_Q_SOLID_NODE_MAP_C2F = 15,   /* int32[numNodes of level-1] */
  GMGS_Q_INTERMEDIATE_NUM_NODES = 16, /* int64 */
  GMGS_Q_STORING_STATE = 17,        /* int64 */
  GMGS_Q_UNIQUE_ELEMENTS = 18,      /* int32[n] (storing state 2) */
  GMGS_Q_KS = 19,                   /* double[numElements x 576]: each element's 24x24 matrix */
  GMGS_Q_DIAGK = 20,                /* double[numDOFs] (levels < numLevels) */
  GMGS_Q_FIXED_DOFS = 21,           /* int32[n] 1-based, after gmgs_set_bc.  Level 1: the
                                       reference's fixedDOFs, [3f-2 3f-1 3f] per fixingCond_
                                       entry f in input order, duplicates kept
                                       (ApplyBoundaryCondition.m:15-17); levels >= 2: sorted
                                       (SetupSolver.m:250-253) */
  GMGS_Q_FREE_DOFS = 22,            /* int32[n] sorted 1-based */
  GMGS_Q_KE0 = 23,                  /* double[576] */
  GMGS_Q_BBOX = 24,                 /* double[6] adjusted bounding box [xmin ymin zmin xmax ymax zmax] */
  GMGS_Q_F = 25,                    /* double[numDOFs] load vector F_ */
  GMGS_Q_NUM_TILES = 26,            /* int64 active 16x4x1 tiles of the level */
  GMGS_Q_NUM_FREE_COARSEST = 27,    /* int64 */
  GMGS_Q_TILE_CONFIG = 28,          /* int64[5]: level kernel (0 tile kernel with element sums,
                                       1 z-marching, 2 tile kernel with the assembled stencil,
                                       3 split: z-marching on F_FAST nodes + assembled rows,
                                       4 level 1 z-marching whose symmetric V(1, 1) residual
                                       kernel forms e = P x_c while staging it),
                                       z-marching tile TX, z-chunk, then TX and z-chunk of the
                                       FP64 matvec of the PCG (level 1 of a mixed context tunes
                                       its own; the level's otherwise) */
  GMGS_Q_SMOOTHER = 29,             /* double[5] (levels < numLevels, after gmgs_set_bc):
                                       lambda_max(D^-1 A) estimate (0 = none), the weights of the
                                       level's nu sweeps, 0 beyond nu */
  GMGS_Q_MAX_IT = 30                /* int64 maxIT_ of the context (sizes iterationHist_:
                                       at most maxIT_ + 1 rows) */
};

void gmgs_default_params(gmgs_params* p);
const char* gmgs_status_string(int status);

int gmgs_create(const gmgs_params* p, const int32_t* vox, int64_t ny, int64_t nx, int64_t nz,
                const double* bbox, gmgs_ctx** out);
int gmgs_set_bc(gmgs_ctx* c, const int32_t* fixed_nodes, int64_t k, const double* loads,
                int64_t m);
int gmgs_set_modulus(gmgs_ctx* c, const double* modulus, int64_t n);
int gmgs_solve(gmgs_ctx* c, double* U, gmgs_solve_info* info, double* hist, int64_t hist_cap);
int gmgs_stress(gmgs_ctx* c, const double* U, double* S);
int gmgs_adjoint(gmgs_ctx* c, const double* g, const double* U, double* lambda, double* dJdE,
                 gmgs_solve_info* info);

/* Batched solve (no reference counterpart): solves the problems of n distinct
 * contexts (same device, BCs set) together.  One captured graph holds a PCG
 * iteration pair of every problem as parallel branches and is replayed until
 * every problem's device-side convergence flag is set; once half of a graph's
 * problems have converged the replays continue on a graph of the others.  The
 * graphs are kept for the same list of contexts until one of them changes
 * (set_bc, set_modulus, destroy).  U[i] (may be NULL) and infos[i] receive
 * each problem's result. */
int gmgs_batch_solve(gmgs_ctx** ctxs, int n, double** U, gmgs_solve_info* infos);

int gmgs_atx(gmgs_ctx* c, int level, const double* x, double* y);

/* gmgs_atx on device-resident vectors, allocating nothing (no counterpart in
 * the reference; for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * x_dev and y_dev are caller-owned device buffers of 3 * numNodes(level)
 * doubles, node-interleaved in compact node order, on the context's device.
 * The two padded level vectors the operator works on (48 bytes per padded
 * node of the level) come from a file-scope device array of
 * GMGS_NOALLOC_SCRATCH_MIB MiB that each GPU holds until the process ends;
 * a larger level returns GMGS_ERR_ARGUMENT.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (two contexts' streams, two host
 * threads) race.  Runs on the context's stream and returns when y_dev is
 * written.  Needs gmgs_set_bc (the fixed DOFs). */
#define GMGS_NOALLOC_SCRATCH_MIB 64
int gmgs_atx_device_noalloc(gmgs_ctx* c, int level, const double* x_dev, double* y_dev);
int gmgs_vcycle(gmgs_ctx* c, const double* r, double* z);
int gmgs_smooth(gmgs_ctx* c, int level, const double* r, double* dx, double* r_out);
int gmgs_restrict(gmgs_ctx* c, int level, const double* r_fine, double* r_coarse);
int gmgs_interpolate(gmgs_ctx* c, int level, const double* x_coarse, double* x_fine);

/* Kernel timing for benchmarks (no reference counterpart): when enabled,
 * every level-1 launch of the PCG iteration is bracketed by HIP events on the
 * context's stream; reads give the mean duration over iterations that ran. */
enum {
  GMGS_K_MATVEC64 = 0,    /* q = A p, FP64 (CG_solver.m:29) */
  GMGS_K_JACOBI64 = 1,    /* Jacobi step r - A(omega r/d), FP64 (CG_solver.m:18-19) */
  GMGS_K_VSMOOTH1 = 2,    /* V-cycle level-1 pre-smoothing (Vcycle.m:14) */
  GMGS_K_RESTRICT1 = 3,   /* level 1 -> 2 restriction (Vcycle.m:17) */
  GMGS_K_PROLONG1 = 4,    /* level 2 -> 1 prolongation + post-smoothing (Vcycle.m:29-32) */
  GMGS_K_ITERATION = 5,   /* one whole PCG iteration */
  GMGS_K_MATVEC64_P = 6   /* q = A p with the direction update fused into it: p = z + beta p_old,
                             and in every other iteration the two deferred x += alpha p updates
                             (CG_solver.m:23-31; mixed precision without the Jacobi step,
                             z-marching level 1) */
};
int gmgs_kernel_timing(gmgs_ctx* c, int enable);
/* Plain device copy of nbytes (read + write), reps times; mean duration.  Used
 * to calibrate HBM traffic counters against a kernel of known size. */
int gmgs_copy_probe(int device, int64_t nbytes, int reps, double* avg_ms);
int gmgs_kernel_timing_read(gmgs_ctx* c, int kernel, double* avg_ms, int64_t* count);
/* Benchmark hook (no reference counterpart): launches one level-1 FP64
 * operator of the PCG iteration reps times back to back on the context's
 * stream, between one pair of HIP events, and returns the mean duration.
 * kernel = GMGS_K_MATVEC64 (q = A p with its p'q partials),
 * GMGS_K_JACOBI64 (the Jacobi step into scratch vectors) or
 * GMGS_K_MATVEC64_P (the fused direction update and matvec, on scratch
 * vectors, alternating as in the iteration between the launch that applies
 * the deferred x updates and the one that does not; GMGS_ERR_STATE where the
 * solver does not use it).  Needs
 * gmgs_set_bc; leaves the solution untouched. */
int gmgs_time_operator(gmgs_ctx* c, int kernel, int reps, double* avg_ms);

int gmgs_query_size(gmgs_ctx* c, int key, int level, int64_t* count);
int gmgs_query(gmgs_ctx* c, int key, int level, void* out, int64_t count);

const char* gmgs_last_error(const gmgs_ctx* c);
void gmgs_destroy(gmgs_ctx* c);

#ifdef __cplusplus
}
#endif

#endif /* GMGS_H */
