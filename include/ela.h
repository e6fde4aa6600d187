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
 * <= 0: no further limit).  The coarsest level is solved exactly by a dense
 * inverse when it has at most 3 * 9^3 unknowns, otherwise by BiCGStab to a
 * relative 1e-4 (AMReX's bottom solver and MLMG's bottom tolerance; at most
 * 200 steps), so any n is accepted.  mode selects which MG precisions are allocated:
 * ELA_SOLVE_FP64 (FP64 levels), ELA_SOLVE_MIXED (FP32 levels plus the FP64
 * level-0 residual) or ELA_ALLOC_BOTH. */
int ela_hier_create(ela_ctx* ctx, const ela_domain* dom, int model_kind, int max_mg_levels, int mode,
                    ela_hier** out);
/* The same with the coarsening floor as a parameter: a level is coarsened
 * while every n is even and n / 2 >= coarse_min (ela_hier_create uses 4).
 * Problems with Displacement rows on every face may go down to 2 cells. */
int ela_hier_create_ex(ela_ctx* ctx, const ela_domain* dom, int model_kind, int max_mg_levels, int coarse_min,
                       int mode, ela_hier** out);
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

/* Stiffness representation for every kernel that follows:
 * 0 = per-node planes (the reference's per-node model field, default);
 * 1 = grain table: a per-node 16-bit grain id and one record per grain
 *     (available after ela_set_model_voronoi_cubic).
 * Level 0 holds the same values in both, so its operator is the same bit for
 * bit.  The coarse MG levels differ: plane form coarsens the stiffness by
 * full weighting (averageDownCoeffsSameAmrLevel, Elastic.cpp:722-809), grain
 * form by injection (a coarse node keeps the grain id and record of the fine
 * node it coincides with).  The preconditioner, and so the iteration count,
 * can therefore differ between the two; the converged solution is that of
 * the same level-0 operator. */
int ela_set_stiffness_mode(ela_hier* h, int mode);

/* Degradable::Isotropic::DegradeModulus (src/Model/Solid/LinearElastic/
 * Degradable/Isotropic.H:74-104) at every level-0 node of an isotropic
 * hierarchy with undegraded moduli (lambda0, mu0): neta = 1 gives
 * lambda = lambda0 (1 - eta), mu = mu0 (1 - eta); neta = 2 gives
 * lambda = lambda0 (1 - eta0), mu = mu0 (1 - eta1).  eta: dense host node
 * values, neta components.  Where the reference aborts, this returns, with
 * the model left as it was: ELA_ERR_NAN for a NaN eta or an infinite eta0
 * ("Nans in eta" / "Infs in eta"), ELA_ERR_BADARG for a value outside [0, 1]
 * (an infinite eta1 included: the reference's Inf test reads eta0 only). */
int ela_set_model_degraded_iso(ela_hier* h, double lambda0, double mu0, int neta, const double* eta);

/* PolymerDegradation::DegradeMaterial (src/Integrator/PolymerDegradation.cpp:
 * 793-833) on the device.  eta_cells: device array of cell-centred damage,
 * neta components of (n0+2)(n1+2)[(n2+2)] cells each (cells -1..n per
 * direction, one ghost layer, x fastest).  Each node takes the mean of the
 * 2^dim cells around it; the reference calls the one-eta DegradeModulus once
 * per component, each from (lambda0, mu0), so the last component sets the
 * moduli, and every component's node mean must lie in [0, 1]
 * (ELA_ERR_BADARG; a NaN gives ELA_ERR_NAN, which the reference would meet as
 * its NaN-diagonal abort, Elastic.cpp:347).  stream NULL: the library's. */
int ela_degrade_material(ela_hier* h, double lambda0, double mu0, int neta, const double* eta_cells, void* stream);

/* The model fill of PhaseFieldMicrostructure (src/Integrator/
 * PhaseFieldMicrostructure.cpp:529-543) with Model::Solid::Affine::Cubic::
 * Combine (src/Model/Solid/Affine/Cubic.H:70-94), on the device, into a
 * major/minor hierarchy's level 0.  nph models: ddw (host, nph records of 21
 * values in 3D / 6 in 2D, the ELA_MODEL_MM order), F0 (host, nph row-major
 * 3x3, or NULL for zero).  eta_cells: device cell array of nph order
 * parameters laid out as for ela_degrade_material; a node's eta_n is the mean
 * of the 4 cells (i, j), (i, j-1), (i-1, j), (i-1, j-1) of its x-y plane at
 * cell layer k (the reference's 2D average, also in 3D).  Each node's record
 * and F0 are sum_n ddw_n eta_n / sum(eta), accumulated in model order.
 * F0_out: device, dim*dim FP64 planes (row-major) in the level-0 layout, the
 * F0 input of ela_newton_rhs / ela_newton_w_dw, or NULL.  A node whose
 * weights sum to zero makes a NaN record: ELA_ERR_NAN (the model is written). */
int ela_combine_cubic(ela_hier* h, int nph, const double* ddw, const double* F0, const double* eta_cells,
                      void* F0_out, void* stream);

/* SetBC (src/Operator/Elastic.H:58-66); values are used by ela_bc_init. */
int ela_set_bc(ela_hier* h, const int* types, const double* values);

/* SetUniform (src/Operator/Elastic.H:111): uniform != 0 drops the grad(C) term. */
int ela_set_uniform(ela_hier* h, int uniform);

/* prepareForSolve (src/Operator/Elastic.H:131-136, src/Operator/Operator.cpp:445-452):
 * coarsen the stiffness to every MG level (averageDownCoeffsSameAmrLevel,
 * Elastic.cpp:722-809), compute the diagonals (Diagonal, Elastic.cpp:264-352)
 * and factor the coarsest level. */
int ela_prepare(ela_hier* h);

/* Raw device buffers, counted against the context's HBM budget (for inputs
 * that are not level fields, e.g. the cell arrays of ela_degrade_material),
 * and a copy between any host or device pointers (hipMemcpyDefault). */
int ela_dev_alloc(ela_ctx* ctx, size_t bytes, void** out);
int ela_dev_free(ela_ctx* ctx, void* p);
int ela_memcpy(void* dst, const void* src, size_t bytes);

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

/* Post-processing of a level-0 displacement field u (FP64, level layout):
 * what = 0 Strain (src/Operator/Elastic.cpp:399-447): sym(grad u);
 * what = 1 Stress (Elastic.cpp:451-499): C : grad u;
 * what = 2 Energy (Elastic.cpp:502-549): W(grad u) + sigma : eps / 2 with
 *          W = tr(grad u^T sigma) / 2, the sum the reference forms (for the
 *          symmetric sigma of these models it equals sigma : eps).
 * grad u uses central differences inside and one-sided ones on domain walls.
 * Output (FP64, level layout): voigt != 0 gives 6 components (00,11,22,12,20,01)
 * in 3D, 3 (00,11,01) in 2D; otherwise dim*dim row-major; energy has 1. */
int ela_postprocess(ela_hier* h, int what, int voigt, const void* u, void* out, void* stream);

/* Newton<T>::prepareForSolve (src/Solver/Nonlocal/Newton.H:30-140) for the
 * models whose kinematic variable is grad u (Linear and Affine Isotropic /
 * Cubic, Laplacian): with dw = DDW : (grad u - F0) per node (one-sided
 * gradient on the walls), rhs = b - BC row(u, grad u, dw) on the boundary and
 * rhs = b - div dw (central differences of dw) inside.  u, b, rhs: FP64
 * level-0 fields (dim components); F0: dim*dim FP64 planes (row-major) in the
 * level-0 layout, or NULL for F0 = 0.  Newton's step then solves A du = rhs. */
int ela_newton_rhs(ela_hier* h, const void* u, const void* b, const void* F0, void* rhs, void* stream);

/* Newton<T>::W and Newton<T>::DW (src/Solver/Nonlocal/Newton.H:233-333) for
 * the models whose kinematic variable is grad u, at every level-0 node with
 * the domain stencil (one-sided on the walls): dw = DDW : (grad u - F0)
 * (dim*dim FP64 planes, row-major) and
 * W = trace((wfac (grad u - F0))^T dw), where wfac = 0.5 for the Linear
 * models and Affine::Cubic and wfac = 1 for Affine::Isotropic (its W has no
 * 1/2, src/Model/Solid/Affine/Isotropic.H:31-34).  u: FP64 level-0 field; F0:
 * dim*dim planes or NULL; w (1 plane) or dw may be NULL.  stream NULL: the
 * call completes before returning. */
int ela_newton_w_dw(ela_hier* h, const void* u, const void* F0, double wfac, void* w, void* dw, void* stream);

/* normalize (src/Operator/Operator.cpp:164-195), as the reference computes it:
 * its per-component lambda indexes neither x nor diag by component, so
 * component 0 of x is divided by component 0 of the diagonal once per
 * operator component (dim times) and the other components are unchanged.
 * Requires ela_prepare (the reference aborts unless Diagonal was computed). */
int ela_normalize(ela_hier* h, int mglev, int prec, void* x, void* stream);

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

/* ela_residual without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels, arguments and results as ela_residual, bit for bit; the per-block
 * maxima of the norm and the norm word live in a file-scope device array
 * (1 MiB + 256 B per GPU, set up by ela_ctx_create and held until the process
 * ends) instead of the hierarchy's buffers.  A level needing more than 131072
 * block maxima (more nodes than a 257^3 level) returns ELA_ERR_BADARG.  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (other
 * host threads, other streams) race, and must be serialized by the caller.
 * Nothing in the array needs to start at zero: every call writes each block
 * maximum it reads. */
int ela_residual_noalloc(ela_hier* h, int mglev, int prec, const void* x, const void* b, void* r, double* norm_inf,
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

/* Kernel options (no counterpart in the reference):
 * ELA_OPT_MARCH  1 (default): grain-form stiffness (a finest level in grain
 *                form, up to 256 grains) runs the LDS-tiled z-marching stencil
 *                kernels; 0: the one-thread-per-node kernels.  The arithmetic
 *                is the same, so are the results, bit for bit.
 * ELA_OPT_KZ     planes marched per tile (default 16).
 * ELA_OPT_XCD    1: consecutive tiles of a marching launch run on one XCD
 *                (shared halo lines stay in its L2); 0: hardware order.
 *                Default 1 for a hierarchy, 0 for a batch.
 * ELA_OPT_MASK   per-node-plane levels: skip the derivative of C where the
 *                records at a node's two neighbours are bitwise equal (its
 *                terms are exactly zero, so results keep their bits up to
 *                the sign of a zero), and launch the nodes with constant C
 *                around them first (a level-wide node list, or regrouped
 *                inside each block), so that whole wavefronts skip.  0 never,
 *                1 on the levels where it pays (default), 2 every plane level
 *                with node lists, 3 every plane level with in-block
 *                regrouping.  Hierarchies only.
 * ELA_OPT_GRAPH  1: the cycles' launches on levels >= 1 (about 255 of an
 *                F-cycle's 270) are captured into HIP graphs, one per launch
 *                configuration and buffer state, and replayed; level 0's
 *                launches stay on the stream.  Same kernels in the same
 *                order, so the same results.  0 (default): stream launches
 *                (on MI355X the replay measured no faster: the small levels'
 *                cost is each kernel's own latency).  Hierarchies only.
 *                At most 8 graphs are kept per hierarchy; an evicted one is
 *                destroyed only after the stream has drained.
 * ELA_OPT_FOLD   1 (default): in a Krylov step on a grain-form level 0 the
 *                F-cycle's last smoothing sweep writes the FP64 vector
 *                (its value + the held correction, rounded as the separate
 *                addition rounds) instead of storing its own value; 0: the
 *                sweep stores its value and a separate kernel adds.  Same
 *                bits either way.  Hierarchies only.
 * ELA_OPT_ROWS   rows (j) per thread of level 0's FP32 marching tiles: 1
 *                (default; 32 x 8 nodes per tile) or 2 (32 x 16: the halo
 *                is a smaller share of the tile, and a thread's two stencils
 *                share LDS reads).  Same results, bit for bit.  Hierarchies
 *                only.
 * ELA_OPT_PREFETCH planes of u and grain ids a marching tile loads ahead of
 *                the plane it computes: 1 (default) or 2 (the loads of plane
 *                k+3 and the rhs/diagonal of plane k+1 are issued while plane
 *                k is computed, so more bytes are in flight per block).  Same
 *                results, bit for bit.  Hierarchies only.
 * ELA_OPT_GJ     1 (default): ela_prepare inverts the dense coarsest operator
 *                by Gauss-Jordan in one launch a column (every block finds
 *                the pivot itself; the row swaps are kept as a permutation);
 *                0: a pivot launch and an elimination launch a column.  Same
 *                pivots and bits either way.  Hierarchies only.
 * ELA_OPT_GRAIN_DIAG 1 (default): the marching tiles of a grain-form level,
 *                the grain-table node kernels of its coarse levels and the
 *                first Jacobi sweep from zero on every level take
 *                an interior node's Jacobi diagonal from its grain (3 values a
 *                grain and level; the tiles stage them in LDS with the table:
 *                sum_b C_{p b p b} (-2 / h_b^2) depends on the node's own
 *                record only) instead of reading the stored diagonal field;
 *                0: they read the field.  Same bits.  Hierarchies only. */
enum {
    ELA_OPT_MARCH = 0,
    ELA_OPT_KZ = 1,
    ELA_OPT_XCD = 2,
    ELA_OPT_MASK = 3,
    ELA_OPT_GRAPH = 4,
    ELA_OPT_FOLD = 5,
    ELA_OPT_ROWS = 6,
    ELA_OPT_PREFETCH = 7,
    ELA_OPT_GJ = 8,
    ELA_OPT_GRAIN_DIAG = 9,
    ELA_OPT_PROLONG = 10
};
int ela_set_option(ela_hier* h, int option, int value);

/* Autotuner of the marching kernels (no reference counterpart): times nrep
 * level-0 Jacobi sweeps on this hierarchy's data for every tile height in
 * {8, 16, 32}, both tile orders, (FP32 sweeps) 1 or 2 rows per thread and
 * 1 or 2 planes loaded ahead, and keeps the fastest (ELA_OPT_KZ, ELA_OPT_XCD,
 * ELA_OPT_ROWS, ELA_OPT_PREFETCH); then gives every coarse marching level the
 * fastest of the heights {4, 8, 16, 32} for its own sweeps (ELA_OPT_KZ resets
 * those to the hierarchy's).  ms (24 doubles, may be NULL) receives level 0's
 * average sweep times at index 12 (prefetch - 1) + 6 (rows - 1) + 2 (kz
 * 8/16/32 -> 0/1/2) + order; 0 for rows 2 of an FP64 hierarchy.  Requires
 * ela_prepare and grain form. */
int ela_autotune(ela_hier* h, int nrep, double* ms);

/* Record HIP events around the level-0 kernels of subsequent solves (on != 0). */
int ela_set_timing(ela_hier* h, int on);

/* Time nrep back-to-back launches of Fapply on level 0 with HIP events on
 * the library's stream; returns the average ms per launch. */
int ela_time_fapply(ela_hier* h, int prec, int nrep, double* avg_ms);

/* Measurement helper with no counterpart in the reference: nrep launches of a
 * plain streaming copy of `bytes` on the device (reads and writes `bytes`
 * each), average ms per launch.  It calibrates the HBM byte counters. */
int ela_probe_copy(ela_ctx* ctx, size_t bytes, int nrep, double* avg_ms);
/* The same with a copy variant: 0 the kernel above (8192 blocks, one 16-B
 * access per lane and step); 1-3 four 16-B loads in flight per lane and
 * non-temporal stores on 2048 / 8192 / 32768 blocks; 4-6 eight loads in
 * flight, same grids.  The bandwidth probe reports the fastest. */
int ela_probe_copy_ex(ela_ctx* ctx, size_t bytes, int nrep, int variant, double* avg_ms);
/* Measurement helper with no counterpart in the reference: the cost of a
 * device-wide barrier among `blocks` (1..1024) co-resident blocks of 256
 * threads (one device-scope atomic counter, bounded spin; `rounds` barriers in
 * one launch, microseconds each), and of a back-to-back empty launch of the
 * same grid (microseconds each, over 1000 launches).  ELA_ERR_STATE if the
 * barrier timed out. */
int ela_probe_barrier(ela_ctx* ctx, int blocks, int rounds, double* us_per_barrier, double* us_per_launch);

/* ------------------------------------------------------------------ AMR level transfers
 *
 * An AMR level's refined region is one box of nodes of the finer level, with
 * its own layout (`fine`, e.g. the level 0 of a hierarchy built on the box);
 * fine_origin is the fine-level node index of its first node (even).  The
 * coarse level is described by `crse` (its node counts are the coarse
 * domain's).  Fields on the device in those layouts, ncomp components.
 * stream NULL: the call completes before returning. */

/* reflux (src/Operator/Operator.cpp:728-839): every coarse node of the
 * coarsened fine box takes the full-weighting restriction of the fine
 * residual (the corner / edge / face / interior branches of the coarse
 * domain); fine values outside the box are ghost nodes, zero. */
int ela_reflux(const ela_level_layout* crse, void* crse_res, const ela_level_layout* fine, const void* fine_res,
               const int fine_origin[3], int ncomp, int prec, void* stream);

/* The same two transfers into a window of the coarse level (e.g. an AMR
 * level's own arrays, which cover only its grown box): crse describes the
 * coarse domain (node counts; they decide the domain-boundary branches),
 * crse_win the window's array (node counts and strides) and crse_origin the
 * coarse node index of its first node; the window must hold every coarse
 * node under the fine box. */
int ela_reflux_window(const ela_level_layout* crse, const ela_level_layout* crse_win, const int crse_origin[3],
                      void* crse_res, const ela_level_layout* fine, const void* fine_res, const int fine_origin[3],
                      int ncomp, int prec, void* stream);
int ela_average_down_window(const ela_level_layout* crse, const ela_level_layout* crse_win, const int crse_origin[3],
                            void* crse_sol, const ela_level_layout* fine, const void* fine_sol,
                            const int fine_origin[3], int ncomp, int prec, void* stream);

/* averageDownSolutionRHS (src/Operator/Operator.cpp:626-641): coarse nodes of
 * the coarsened fine box take the coincident fine solution value (nodal
 * amrex::average_down). */
int ela_average_down(const ela_level_layout* crse, void* crse_sol, const ela_level_layout* fine,
                     const void* fine_sol, const int fine_origin[3], int ncomp, int prec, void* stream);

/* ------------------------------------------------------------------ composite (AMR) solve
 *
 * Operator::Elastic<T>::define on several AMR levels (src/Operator/Elastic.cpp:30-55,
 * src/Operator/Operator.cpp:215-258) and the multi-level
 * Solver::Nonlocal::Linear::solve (src/Solver/Nonlocal/Linear.H:102-112).
 * Level 0 is the domain; level l >= 1 (ratio 2) is the union of a list of boxes
 * of level-l nodes [lo, hi] (even corners; boxes may share faces, and the union
 * need not be a box -- the reference's BoxArray of the level, e.g. a region
 * chopped into boxes of at most max_grid_size cells).  Every box, coarsened and
 * grown by 2 coarse cells (clipped to the domain), must lie in the coarser
 * level's union.  Fields of level l live on the bounding box of its union grown
 * by 2 nodes and clipped to the domain (ela_amr_level_info: layout and the
 * global index of its first node); the boxes share that array, so their shared
 * nodes are stored once (AMReX's halo fills and nodalSync are identities).
 * Nodes outside the union and on its coarse/fine boundary hold the
 * interpolated coarser solution, as the reference's fine ghost nodes do
 * (Operator.cpp:556-624, Linear.H:23-29).  Coarse nodes under a finer union
 * take its values (averageDownSolutionRHS, Operator.cpp:626-641); at coarse
 * nodes on the coarse/fine interface the equation is the reflux of the fine
 * residual (Operator.cpp:728-839).  The solve is flexible GMRES(restart) in
 * FP64 preconditioned by one FAC cycle: level 0, and a refined level whose
 * union is one box (e.g. a box chopped into max_grid_size boxes) that
 * coarsens by 2 to a dense-solvable level, correct with a cycle of a
 * multigrid hierarchy on that box (FP32 or FP64 by mode); any other level
 * with weighted-Jacobi smoothing (FP64), as AMReX's MLMG smooths fine AMR
 * levels. */
typedef struct ela_amr ela_amr;
enum { ELA_AMR_U = 0, ELA_AMR_B = 1, ELA_AMR_R = 2 };

/* nboxes: (nlev - 1) box counts; boxes: the boxes of level 1, then of level 2,
 * ..., each lo[3], hi[3] (level-l node indices).  Level 0's multigrid hierarchy
 * coarsens down to coarse_min cells.  mode: ELA_SOLVE_FP64 or ELA_SOLVE_MIXED. */
int ela_amr_create_boxes(ela_ctx* ctx, const ela_domain* dom, int nlev, const int* nboxes, const int* boxes,
                         int model_kind, int coarse_min, int mode, ela_amr** out);
/* The same with one box per refined level: boxes holds (nlev - 1) records. */
int ela_amr_create(ela_ctx* ctx, const ela_domain* dom, int nlev, const int* boxes, int model_kind, int coarse_min,
                   int mode, ela_amr** out);
int ela_amr_destroy(ela_amr* a);
size_t ela_amr_bytes(const ela_amr* a);
int ela_amr_level_info(const ela_amr* a, int lev, ela_level_layout* out, int origin[3]);
/* SetModel(amrlev, FabArray) (src/Operator/Elastic.H:49-54): dense host SoA
 * over the level's grown box. */
int ela_amr_set_model_field(ela_amr* a, int lev, const double* C_host);
/* SetBC / SetUniform / cycle / Krylov options as for a single box. */
int ela_amr_set_bc(ela_amr* a, const int* types, const double* values);
/* BC::Operator::Elastic::Init (src/BC/Operator/Elastic.H:106-181) on level lev:
 * nodes of the grown box on the domain boundary take their class's value. */
int ela_amr_bc_init(ela_amr* a, int lev, double* rhs_host_dense, int homogeneous);
int ela_amr_set_uniform(ela_amr* a, int uniform);
int ela_amr_set_cycle(ela_amr* a, int cycle, int nu1, int nu2);
/* The cycle of one level only (ela_amr_set_cycle sets every level).  A level
 * with a hierarchy (its union is one box): that hierarchy's cycle (defaults:
 * F-cycle with nu1 = nu2 = 2 on level 0, V-cycle with nu1 = nu2 = 1 on refined
 * levels).  Any other refined level: nu1 / nu2 Fsmooth calls (2 Jacobi sweeps
 * each) before / after the coarse correction (default 2 / 2; cycle unused). */
int ela_amr_set_level_cycle(ela_amr* a, int lev, int cycle, int nu1, int nu2);
/* restart 0: the plain FAC iteration u += FAC(r). */
int ela_amr_set_krylov(ela_amr* a, int restart);
/* prepareForSolve on every level (Operator.cpp:445-452). */
int ela_amr_prepare(ela_amr* a);
/* Resident fields (ELA_AMR_U solution, ELA_AMR_B rhs, ELA_AMR_R composite
 * residual), dense host arrays over a level's grown box. */
int ela_amr_upload(ela_amr* a, int lev, int field, const double* host);
int ela_amr_download(ela_amr* a, int lev, int field, double* host);
/* Composite residual of the resident u and b into ELA_AMR_R (zero off the
 * unknowns), with its max norm (solutionResidual + reflux, Operator.cpp:728-863). */
int ela_amr_residual(ela_amr* a, double* norm_inf);
/* Solve from the resident u to |r|inf <= max(tol_abs, tol_rel |r0|inf);
 * max_iter counts FAC cycles. */
int ela_amr_solve(ela_amr* a, double tol_rel, double tol_abs, int max_iter, ela_stats* stats);

/* ------------------------------------------------------------------ batched independent solves
 *
 * The reference solves one problem per Solver::Nonlocal::Linear::solve call
 * (src/Solver/Nonlocal/Linear.H:102-112).  A batch holds many independent
 * members ("RVEs": cubes of n cells per side on [0,1]^3, cubic Voronoi
 * polycrystals) with their multigrid hierarchies resident on the device, and
 * runs each kernel of a cycle as one launch over all members.  Per member it
 * is the same discrete problem and the same F-cycle iteration as ela_solve
 * with coarse_min = 2; a member's iterates do not depend on the other
 * members, so its solution is bitwise the same alone and in any mix.
 *
 * Members: any n in 4..1024.  A member's levels halve n while it is even and
 * n / 2 >= 2; its coarsest level of c cells is solved by a dense inverse when
 * 3 (c+1)^3 <= 1536 (c <= 7), else by BiCGStab to 1e-4 (|r|inf relative, at
 * most 200 iterations; the reference's bottom solver, Linear.H:23-29), with
 * each member's scalars from its own partial sums.  Fields on the host are
 * dense (x fastest), 3 components. */
typedef struct ela_batch ela_batch;
typedef struct ela_rve ela_rve;

/* mode: ELA_SOLVE_MIXED (FP32 cycles inside FP64 refinement) or ELA_SOLVE_FP64. */
int ela_batch_create(ela_ctx* ctx, int mode, ela_batch** out);
int ela_batch_destroy(ela_batch* b);
size_t ela_batch_bytes_in_use(const ela_batch* b);

/* Device bytes a member needs (admission control against the context budget). */
size_t ela_rve_bytes(int n, int mode, int max_grains, int with_factor);

/* A member of n cells per side, its hierarchy allocated against the context's
 * HBM budget (ELA_ERR_OOM when it does not fit).  Replaces, per member,
 * Operator::Elastic<T>::define (src/Operator/Elastic.cpp:30-55).  with_factor
 * reserves a per-node degradation factor (ela_rve_set_crack).  Every BC type
 * starts as Displacement, the reference's default (BC/Operator/Elastic.H:46-55). */
int ela_rve_create(ela_batch* b, int n, int max_grains, int with_factor, ela_rve** out);
int ela_rve_destroy(ela_rve* r);
size_t ela_rve_bytes_in_use(const ela_rve* r);
int ela_rve_levels(const ela_rve* r);

/* SetBC types (src/Operator/Elastic.H:58-66): types[26][3], faces as ela_set_bc. */
int ela_rve_set_bc_types(ela_rve* r, const int* types);

/* SetUniform (src/Operator/Elastic.H:111): uniform != 0 drops the grad(C) term (default 0). */
int ela_rve_set_uniform(ela_rve* r, int uniform);

/* Per-node cubic stiffness of a Voronoi polycrystal (Cubic::Define,
 * src/Model/Solid/Linear/Cubic.H:30-59; the per-node fill of
 * src/Integrator/PhaseFieldMicrostructure.cpp:529-543), stored as a 16-bit
 * grain id per node and one record per grain.  Seeds as in
 * ela_set_model_voronoi_cubic. */
int ela_rve_set_voronoi_cubic(ela_rve* r, int nseeds, const int64_t* seeds, int q, const double* R, double C11,
                              double C12, double C44);

/* Phase-field degradation of the stiffness (the role of
 * LinearElastic::Degradable::Isotropic::DegradeModulus,
 * src/Model/Solid/LinearElastic/Degradable/Isotropic.H:74-104, applied to the
 * cubic polycrystal): C(x) = g(eta(x)) C_grain(x), g(eta) = a eta^2 + b eta + c,
 * eta = (1 - tanh(d / width)) / 2 with d the distance to the crack
 * {x : nrm.(x - p) = 0, tv.(x - p) <= 0} (a half-plane with front through p). */
int ela_rve_set_crack(ela_rve* r, const double p[3], const double nrm[3], const double tv[3], double width, double a,
                      double b, double c);

/* Right-hand side for a prescribed affine displacement u = G x (G row-major
 * 3x3) on the Displacement rows of boundary nodes and 0 elsewhere
 * (BC::Operator::Elastic::Init, src/BC/Operator/Elastic.H:106-181, with
 * node-dependent values).  Uses the member's current BC types. */
int ela_rve_set_rhs_affine(ela_rve* r, const double G[9]);

/* Host <-> member fields of the finest level: 0 = u, 1 = b, 2 = r (output of
 * ela_batch_fapply), 3 = degradation factor (1 component, download only). */
int ela_rve_upload(ela_rve* r, int field, const double* host);
int ela_rve_download(const ela_rve* r, int field, double* host);
int ela_rve_grain_ids(const ela_rve* r, uint16_t* host);

/* prepareForSolve for a set of members (src/Operator/Operator.cpp:445-452):
 * coarse stiffness (grain ids and factors injected level by level, as a
 * grain-form hierarchy coarsens), diagonals and the dense coarsest inverses,
 * batched. */
int ela_batch_prepare(ela_batch* b, int count, ela_rve* const* rves);

/* Fapply (src/Operator/Elastic.cpp:151-260) on the finest level of each
 * member, one launch: r = A u (prec ELA_F64), or u rounded to FP32 through the
 * FP32 operator (ELA_F32, mixed batches). */
int ela_batch_fapply(ela_batch* b, int count, ela_rve* const* rves, int prec);

/* Linear::solve per member (src/Solver/Nonlocal/Linear.H:102-112): submit
 * queues a member (u on the device is the initial guess, zero after
 * create); each step admits the queued members and runs one F-cycle on every
 * active member; members leave the set when |r|inf <= max(tol_abs,
 * tol_rel |r0|inf), on max_cycles or on a non-finite residual.  n_active
 * counts members still queued or active.  ela_batch_solve submits count
 * members and steps until all are done. */
int ela_batch_submit(ela_batch* b, ela_rve* r, double tol_rel, double tol_abs, int max_cycles);
int ela_batch_step(ela_batch* b, int* n_active);
int ela_batch_solve(ela_batch* b, int count, ela_rve* const* rves, double tol_rel, double tol_abs, int max_cycles);

/* state: 0 idle, 1 queued, 2 active, 3 done; status of a finished solve
 * (ELA_OK, ELA_ERR_NONCONVERGED, ELA_ERR_NAN); stats as for ela_solve
 * (seconds = submit to done). */
int ela_rve_status(const ela_rve* r, int* state, int* status, ela_stats* stats);

/* One request of the serving loop: a Voronoi cubic polycrystal member of n
 * cells per side (any n in 4..1024), optionally cracked (ela_rve_set_crack
 * with g(eta) = crack_g[0] eta^2 + crack_g[1] eta + crack_g[2]), loaded by the
 * affine displacement u = G x on every face.  u_out (3 (n+1)^3 doubles) or
 * NULL receives the solution. */
typedef struct {
    double arrival; /* seconds from the start of ela_serve */
    int n;
    int nseeds;
    const int64_t* seeds;
    int q;
    const double* R;
    double C11, C12, C44;
    int crack;
    double crack_p[3], crack_nrm[3], crack_tv[3];
    double crack_width;
    double crack_g[3];
    double G[9];
    double* u_out;
    int max_cycles; /* > 0: this request's cycle limit; 0: the loop's */
} ela_request;

typedef struct {
    double admitted; /* seconds from the start of ela_serve; -1 if never admitted */
    double finished; /* -1 if not finished within max_seconds */
    int cycles;
    int status;      /* ELA_OK, ELA_ERR_NONCONVERGED, ELA_ERR_NAN, an admission error, or -1 (not served) */
    double res0, res;
} ela_result;

/* Where a serving run spent its time (seconds of host wall time). */
typedef struct {
    int steps;         /* cycles run over the members in flight */
    int groups;        /* admission groups (one set-up each) */
    int admitted;
    int peak_active;   /* most members in flight at once */
    size_t peak_bytes; /* most member bytes resident at once */
    double t_admit;    /* member creation, model and right-hand side evaluation */
    double t_prepare;  /* batched set-up of admitted groups */
    double t_step;     /* cycles (including the FP64 residual and its norm read-back) */
    double t_retire;   /* copy-out and freeing of finished members */
    double t_idle;     /* waiting for arrivals with nothing in flight */
} ela_serve_summary;

/* Serving loop with continuous batching (no counterpart in the reference):
 * requests sorted by arrival are admitted first-come first-served while the
 * context's HBM budget and max_active allow; each step is one F-cycle over
 * every member in flight; members leave as they converge and free their
 * memory.  A request that fails (see ela_server_start) gets its own status
 * while the others continue.  Stops when all requests are done or after
 * max_seconds (> 0).  summary (may be NULL) receives the time breakdown. */
int ela_serve(ela_batch* b, int nreq, const ela_request* req, double tol_rel, double tol_abs, int max_cycles,
              int max_active, double max_seconds, ela_result* out, ela_serve_summary* summary);

/* Online serving (no counterpart in the reference): a host thread runs the
 * continuous-batching loop of ela_serve on batch b.  Requests submitted at any
 * time are admitted first come, first served as the context's HBM budget and
 * max_active allow, join the members in flight at the next cycle and leave as
 * they finish.  A request whose member cannot be made (bad size, a model or
 * set-up that is not finite, larger than the whole budget) or that fails in
 * flight (ELA_ERR_NAN, ELA_ERR_NONCONVERGED on its cycle limit) gets its own
 * status; the others are unaffected and keep the bits they have alone.
 * submit copies the request and its seed and rotation arrays (u_out, if set,
 * must stay valid until the request is done) and returns a ticket (> 0);
 * poll sets done = 1 and the result once the ticket is finished (result times
 * are seconds from ela_server_start; a result is reported once).  stop with
 * drain != 0 finishes the requests submitted so far, otherwise the waiting and
 * in-flight ones end unserved (status -1); it joins the thread and frees the
 * server, returning the loop's fatal error if one stopped it.  One server per
 * batch; the batch takes no other calls while it runs. */
typedef struct ela_server ela_server;
int ela_server_start(ela_batch* b, double tol_rel, double tol_abs, int max_cycles, int max_active, ela_server** out);
int ela_server_submit(ela_server* s, const ela_request* req, int64_t* ticket);
int ela_server_poll(ela_server* s, int64_t ticket, int* done, ela_result* out);
int ela_server_stop(ela_server* s, int drain);

/* Kernel options of a batch (set while no member is queued or active); see
 * ela_set_option.  ELA_OPT_MARCH, ELA_OPT_KZ, ELA_OPT_XCD and ELA_OPT_GRAPH:
 * 1 captures a step's F-cycle (every launch on every level) into a HIP graph
 * and replays it while the active set is unchanged; a change of the set (a
 * member admitted or retired) re-captures.  Cycles with a BiCGStab coarsest
 * level (their host reads the members' states) and timed cycles run on the
 * stream.  Same kernels in the same order: the same bits.  Default 0.
 * ELA_OPT_GJ: the members' coarsest Gauss-Jordan in one launch a column (1,
 * default) or in pivot/eliminate launch pairs (0); the same bits.
 * ELA_OPT_GRAIN_DIAG: the marching tiles of members without factors take an
 * interior node's Jacobi diagonal from its grain (1, default) or read the
 * stored diagonal field (0); the same bits.
 * ELA_OPT_PROLONG: prolongation with one thread per coarse node, writing its
 * up to 8 fine nodes (1, default), or one thread per fine node (0); the same
 * bits.  Batches only. */
int ela_batch_set_option(ela_batch* b, int option, int value);
/* Captures and replays of the batch's F-cycle graph since its creation (no
 * reference counterpart). */
int ela_batch_graph_stats(const ela_batch* b, int* captures, int* replays);

/* The autotuner of ela_autotune over the finest levels of the given prepared
 * members (a representative set of the batch's sizes); keeps the fastest
 * tile height and order for the batch. */
int ela_batch_autotune(ela_batch* b, int count, ela_rve* const* rves, int nrep, double* ms);

/* HIP-event timing of every smoother launch of later steps (on != 0), and
 * the totals since: milliseconds, nodes swept, launches. */
int ela_batch_set_timing(ela_batch* b, int on);
int ela_batch_smoother_times(ela_batch* b, double* ms, long long* nodes, int* launches);

/* ------------------------------------------------------------------ plot files
 *
 * Nodal plot file in AMReX's plotfile layout (HyperCLaw-V1.1 Header,
 * Level_<l>/Cell_H, Level_<l>/Cell_D_00000), the output of the reference's
 * Integrator::WritePlotFile (src/Integrator/Integrator.cpp:582-687) for its
 * node fields, so results can be compared with a reference run (AMReX's
 * fcompare).  Host arrays; refinement ratio 2 between levels. */
typedef struct {
    int domain[3];      /* cells of the level's domain per direction */
    int origin[3];      /* node index of the first node of the data array */
    int nodes[3];       /* node counts of the data array (1 in the unused direction in 2D) */
    const double* data; /* ncomp components of nodes[0]*nodes[1]*nodes[2] values, x fastest */
    int nboxes;
    const int* boxes;   /* nboxes records of node lo[3], hi[3], inside the data array */
} ela_plot_level;
int ela_write_plotfile(const char* dir, int dim, int nlev, const ela_plot_level* levels, int ncomp,
                       const char* const* names, double time, int step, const double* prob_lo,
                       const double* prob_hi);

#ifdef __cplusplus
}
#endif

#endif /* ELA_H */
