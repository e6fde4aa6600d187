/* ela.h -- C ABI of the MI355X (gfx950) nodal elasticity solve.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (0 = ELA_OK).  On error, ela_last_error() returns a thread-local message.
 *
 * Each entry point replaces one piece of the reference's
 * Operator::Elastic<T> / Operator<Grid::Node> / Solver::Nonlocal interface
 * (paths relative to the reference checkout); the citation is on each
 * declaration.  The reference signals errors with Util::Abort (SIGABRT,
 * src/Util/Util.H:72-90); this ABI returns a status instead and never aborts.
 *
 * Field layout on the device.  Every multigrid level of a hierarchy is one
 * box of (nx, ny, nz) nodes (nz = 1 in 2D).  A field with c components is
 * stored component-major: element (i, j, k, comp) of a field at base pointer
 * p lives at p[comp*cstride + k*plane + j*pitch + i], with pitch, plane and
 * cstride from ela_level_info().  Host arrays handed to the upload/download
 * helpers and to the model setters are dense (pitch = nx, plane = nx*ny,
 * cstride = nx*ny*nz).
 *
 * Stiffness storage (one record per node, stored as ncomp planes):
 *   ELA_MODEL_ISO   2 values (lambda, mu)            -- Set/Matrix4_Isotropic.H:15
 *   ELA_MODEL_MM    21 (3D) / 6 (2D) values of C_ijkl in the upper-triangular
 *                   Voigt order (00,01,02,11,12,22)  -- Set/Matrix4_MajorMinor.H:19-31,148-189
 *   ELA_MODEL_DIAG  dim*dim values A(i,j), row-major -- Set/Matrix4_Diagonal.H:12-21
 *
 * Boundary conditions: type[face][comp] and value[face][comp] with faces
 * numbered as BC::Operator::Elastic::Face (src/BC/Operator/Elastic.H:24-42):
 * 26 classes in 3D, 8 in 2D; comp stride 3.
 */
#ifndef ELA_H
#define ELA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ELA_OK = 0,
    ELA_ERR_BADARG = 1,
    ELA_ERR_OOM = 2,
    ELA_ERR_HIP = 3,
    ELA_ERR_NONCONVERGED = 4,
    ELA_ERR_NAN = 5,
    ELA_ERR_STATE = 6
};

enum { ELA_MODEL_ISO = 0, ELA_MODEL_MM = 1, ELA_MODEL_DIAG = 2 };
enum { ELA_BC_DISP = 0, ELA_BC_TRAC = 1, ELA_BC_PERIODIC = 2, ELA_BC_NEUMANN = 3 };
enum { ELA_F64 = 0, ELA_F32 = 1 };
enum { ELA_SOLVE_FP64 = 0, ELA_SOLVE_MIXED = 1, ELA_ALLOC_BOTH = 2 };
enum { ELA_CYCLE_F = 0, ELA_CYCLE_V = 1 };

typedef struct ela_ctx ela_ctx;
typedef struct ela_hier ela_hier;

typedef struct {
    int dim;      /* 2 or 3 */
    int n[3];     /* cells per direction */
    double dx[3]; /* cell size per direction */
} ela_domain;

typedef struct {
    int dim;
    int n[3];        /* cells */
    int nodes[3];    /* n + 1 (1 in the unused third direction in 2D) */
    int64_t pitch;   /* elements between j rows */
    int64_t plane;   /* elements between k planes */
    int64_t cstride; /* elements between components */
    double dx[3];
} ela_level_layout;

typedef struct {
    int iters;          /* outer iterations (MG cycles) performed */
    int converged;      /* 1 if |r|inf <= max(tol_abs, tol_rel |r0|inf) */
    int mg_levels;
    int mode;           /* ELA_SOLVE_FP64 or ELA_SOLVE_MIXED */
    double res0;        /* |b - A u0|inf */
    double res;         /* final |b - A u|inf (FP64, on the device) */
    double seconds;     /* wall time of the iteration loop */
    double setup_seconds;
} ela_stats;

const char* ela_last_error(void);

/* Context: one device, with an HBM budget that hierarchies are admitted against. */
int ela_ctx_create(int device, size_t hbm_budget_bytes, ela_ctx** out);
int ela_ctx_destroy(ela_ctx* ctx);
size_t ela_ctx_bytes_in_use(const ela_ctx* ctx);

/* Multigrid hierarchy on one box.  Replaces Operator::Elastic<T>::define
 * (src/Operator/Elastic.cpp:30-55) and Operator<Node>::define
 * (src/Operator/Operator.cpp:215-258): the MG levels are built here.  The
 * domain is coarsened by 2 while every n is even and >= 8 (max_mg_levels
 * <= 0: no further limit).  mode selects which MG precisions are allocated:
 * ELA_SOLVE_FP64 (FP64 levels), ELA_SOLVE_MIXED (FP32 levels plus the FP64
 * level-0 residual) or ELA_ALLOC_BOTH. */
int ela_hier_create(ela_ctx* ctx, const ela_domain* dom, int model_kind, int max_mg_levels, int mode,
                    ela_hier** out);
int ela_hier_destroy(ela_hier* h);
int ela_hier_levels(const ela_hier* h);
int ela_level_info(const ela_hier* h, int mglev, ela_level_layout* out);
size_t ela_hier_bytes(const ela_hier* h);

/* SetModel(const FabArray<BaseFab<T>>&) (src/Operator/Elastic.cpp:84-118):
 * a per-node stiffness field, dense host SoA of ncomp planes. */
int ela_set_model_field(ela_hier* h, const double* C_host);

/* Per-node stiffness evaluation of a cubic polycrystal
 * (Model::Solid::Linear::Cubic::Define, src/Model/Solid/Linear/Cubic.H:30-59,
 * evaluated per node as the reference's callers fill model fields,
 * e.g. src/Integrator/PhaseFieldMicrostructure.cpp:529-543): each node takes
 * the grain of its nearest Voronoi seed.  Seeds are int64 triples in units of
 * (cell size / q); R holds one row-major 3x3 rotation per grain. */
int ela_set_model_voronoi_cubic(ela_hier* h, int nseeds, const int64_t* seeds, int q, const double* R,
                                double C11, double C12, double C44);

/* Stiffness representation of level 0 for every kernel that follows:
 * 0 = per-node planes (the reference's per-node model field, default);
 * 1 = grain table: a per-node 16-bit grain id and one record per grain
 *     (available after ela_set_model_voronoi_cubic).  The same values, so
 *     the same operator bit for bit; coarse MG levels always use planes
 *     (their stiffness is the full-weighting average). */
int ela_set_stiffness_mode(ela_hier* h, int mode);

/* Phase-field-degraded isotropic stiffness (Degradable::Isotropic::DegradeModulus,
 * src/Model/Solid/LinearElastic/Degradable/Isotropic.H:74-104, with a general
 * degradation g): lambda = g(eta) lambda0, mu = g(eta) mu0 with
 * g(eta) = a*eta^2 + b*eta + c, eta a dense host array of node values. */
int ela_set_model_degraded_iso(ela_hier* h, double lambda0, double mu0, const double* eta, double a, double b,
                               double c);

/* SetBC (src/Operator/Elastic.H:58-66); values are used by ela_bc_init. */
int ela_set_bc(ela_hier* h, const int* types, const double* values);

/* SetUniform (src/Operator/Elastic.H:111): uniform != 0 drops the grad(C) term. */
int ela_set_uniform(ela_hier* h, int uniform);

/* prepareForSolve (src/Operator/Elastic.H:131-136, src/Operator/Operator.cpp:445-452):
 * coarsen the stiffness to every MG level (averageDownCoeffsSameAmrLevel,
 * Elastic.cpp:722-809), compute the diagonals (Diagonal, Elastic.cpp:264-352)
 * and factor the coarsest level. */
int ela_prepare(ela_hier* h);

/* Device field helpers on a level's layout. */
int ela_field_alloc(ela_hier* h, int mglev, int prec, int ncomp, void** dev);
int ela_field_free(ela_hier* h, void* dev);
int ela_upload(ela_hier* h, int mglev, int prec, int ncomp, const double* host_dense, void* dev);
int ela_download(ela_hier* h, int mglev, int prec, int ncomp, const void* dev, double* host_dense);
/* Set every element of a device field (layout incl. padding) to zero, on the library's stream. */
int ela_field_zero(ela_hier* h, int mglev, int prec, int ncomp, void* dev);

/* Fapply (src/Operator/Elastic.cpp:151-260): f = A u on MG level mglev. */
int ela_fapply(ela_hier* h, int mglev, int prec, const void* u, void* f, void* stream);

/* Diagonal (src/Operator/Elastic.cpp:264-352) into a caller buffer. */
int ela_diagonal(ela_hier* h, int mglev, int prec, void* diag, void* stream);

/* Fsmooth (src/Operator/Operator.cpp:90-162): nsweeps weighted-Jacobi sweeps
 * x <- x + omega (b - A x) / diag (the reference uses omega = 2/3, 2 sweeps). */
int ela_fsmooth(ela_hier* h, int mglev, int prec, void* x, const void* b, int nsweeps, double omega, void* stream);

/* restriction (src/Operator/Operator.cpp:454-554): nodal full weighting from
 * MG level cmglev-1 onto cmglev. */
int ela_restriction(ela_hier* h, int cmglev, int prec, const void* fine, void* crse, void* stream);

/* interpolation (src/Operator/Operator.cpp:556-624): fine += P crse, from
 * MG level fmglev+1 onto fmglev. */
int ela_interpolation(ela_hier* h, int fmglev, int prec, void* fine, const void* crse, void* stream);

/* solutionResidual / correctionResidual (src/Operator/Operator.cpp:841-863):
 * r = b - A x and |r|inf (synchronizes the stream to return the norm). */
int ela_residual(ela_hier* h, int mglev, int prec, const void* x, const void* b, void* r, double* norm_inf,
                 void* stream);

/* BC::Operator::Elastic::Init (src/BC/Operator/Elastic.H:106-181): write the
 * prescribed boundary values into a dense host rhs of level 0. */
int ela_bc_init(ela_hier* h, double* rhs_host_dense, int homogeneous);

/* Solver::Nonlocal::Linear::solve (src/Solver/Nonlocal/Linear.H:102-112),
 * which for linear models from u0 = 0 is also Newton<T>::solve
 * (src/Solver/Nonlocal/Newton.H:142-209): multigrid on A u = b to
 * |r|inf <= max(tol_abs, tol_rel |r0|inf).  b and u are dense host arrays of
 * level 0 (u holds the initial guess on entry).  mode: ELA_SOLVE_FP64 (every
 * kernel FP64) or ELA_SOLVE_MIXED (FP32 smoothing/transfer inside FP64
 * iterative refinement).  Returns ELA_ERR_NONCONVERGED if max_iter is reached
 * and ELA_ERR_NAN on a non-finite residual. */
int ela_solve(ela_hier* h, const double* b_host, double* u_host, double tol_rel, double tol_abs, int mode,
              int max_iter, ela_stats* stats);

/* The same solve on device-resident level-0 fields (FP64, level layout). */
int ela_solve_device(ela_hier* h, const void* b_dev, void* u_dev, double tol_rel, double tol_abs, int mode,
                     int max_iter, ela_stats* stats);

/* Options of the cycle (defaults mirror Linear.H:23-29 and AMReX MLMG:
 * F-cycles throughout, nu1 = nu2 = 2 smoother calls of 2 sweeps each). */
int ela_set_cycle(ela_hier* h, int cycle, int nu1, int nu2);

/* Timing of the most recent solve's kernels, measured with HIP events on the
 * solve's stream: average duration (ms) of the level-0 FP64 residual kernel
 * and of the level-0 smoother sweep, and their launch counts. */
int ela_kernel_times(const ela_hier* h, double* ms_resid64, int* n_resid64, double* ms_smooth0, int* n_smooth0);

/* Accelerate the multigrid iteration with flexible GMRES(restart), one MG
 * cycle per preconditioner application (0, the default: the plain MG
 * iteration of the reference).  Same discrete solution; max_iter then counts
 * MG cycles. */
int ela_set_krylov(ela_hier* h, int restart);

/* Record HIP events around the level-0 kernels of subsequent solves (on != 0). */
int ela_set_timing(ela_hier* h, int on);

/* Time nrep back-to-back launches of Fapply on level 0 with HIP events on
 * the library's stream; returns the average ms per launch. */
int ela_time_fapply(ela_hier* h, int prec, int nrep, double* avg_ms);

/* Measurement helper with no counterpart in the reference: nrep launches of a
 * plain streaming copy of `bytes` on the device (reads and writes `bytes`
 * each), average ms per launch.  It calibrates the HBM byte counters. */
int ela_probe_copy(ela_ctx* ctx, size_t bytes, int nrep, double* avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* ELA_H */
