// ela_composite.hip -- composite (AMR) solve of the nodal elasticity problem
// over a nest of boxes: Operator::Elastic<T> defined on several AMR levels
// (src/Operator/Elastic.cpp:30-55) and solved by Solver::Nonlocal::Linear
// (src/Solver/Nonlocal/Linear.H:102-112) with the reference's AMR pieces:
// reflux (src/Operator/Operator.cpp:728-839), averageDownSolutionRHS
// (:626-641), trilinear interpolation (:556-624), and fine ghost nodes that
// hold the interpolated coarse solution (the ghost-node coarse/fine strategy
// selected at Linear.H:23-29).
//
// Levels.  Level 0 is the domain.  Level l >= 1 is the union of a list of
// boxes of level-l nodes [lo, hi] (even corners; the boxes may share faces and
// their union need not be a box, e.g. an L shape, or one region chopped into
// boxes of at most max_grid_size cells).  On one GPU the boxes of a level share
// one array over G_l, the bounding box of the union grown by 2 nodes and clipped
// to the domain: a node shared by two boxes is stored once, so AMReX's
// box-to-box halo fills and nodalSync are identities, and every kernel of a
// level is one launch over G_l.  Node classes of level l, built on the host
// from the boxes and kept as one byte per node of G_l:
//   fill      not in the union, or on the union's boundary away from the domain
//             boundary (a slave: some cell touching it is in the domain but not
//             in the union);
//   covered   every in-domain cell touching it lies under level l+1;
//   interface some but not all of those cells lie under level l+1;
//   unknown   every other node of the union (interface nodes included).
// Fill nodes take the trilinear interpolation of level l-1; covered nodes the
// coincident level-(l+1) value.  The equations are b - A_l u = 0 at unknown
// nodes (A_l the operator on G_l with the domain's boundary rows on domain
// faces) and, at interface nodes, the reflux of level l+1's residual = 0 --
// the problem restated by the oracle (oracle/amr.py).
//
// Solver.  Flexible GMRES(m) in FP64 on the composite unknowns, right-
// preconditioned by one multiplicative FAC cycle over the AMR levels: from the
// finest level down, a correction of the level's residual equation from zero,
// whose residual is restricted into the coarser level under the union; the
// level-0 multigrid cycle; then upwards, the coarse correction interpolated
// onto the finer level (every node) and one more correction.  A level whose
// union is one box (level 0, or a region chopped into max_grid_size boxes)
// corrects with a V(1,1) cycle of a multigrid hierarchy on that box
// (homogeneous Dirichlet rows on its coarse/fine faces) when the box coarsens
// to a dense-solvable level; any other level
// smooths, as AMReX's MLMG does on fine AMR levels: nu1 / nu2 Fsmooth calls (2
// weighted-Jacobi sweeps each, omega = 2/3, on the level's non-fill nodes, fill
// nodes held).  In mixed precision the hierarchy cycles and the smoothing run
// in FP32 (FP32 stiffness and inverse diagonal of the level, the right-hand
// side and the iterate rounded in, the result widened out); everything else
// in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ela.h"
#include "ela_hooks.hpp"
#include "ela_internal.hpp"
#include "ela_kernels.hpp"

namespace ela {
namespace {

constexpr int CX = 64, CY = 4;
enum { NODE_UNKNOWN = 0, NODE_FILL = 1, NODE_COVERED = 2 };

// Node-class bits of the per-level mask (one byte per node of G).
enum { CLS_FILL = 1, CLS_COVERED = 2, CLS_IFACE = 4, CLS_UNDER = 8 };

// Geometry of one level (global node indices of the level).
struct Geo {
    int dim;
    int glo[3];                 // first node of G
    int gn[3];                  // nodes of G per direction
    int N[3];                   // domain cells of the level
    int has_fine;
    int ulo[3], uhi[3];         // bounding box of the nodes under level l+1 (has_fine)
    const unsigned char* cls;   // class bits per node of G (dense, x fastest)
};

__device__ __forceinline__ unsigned char cls_at(const Geo& g, const int p[3]) {
    const int i = p[0] - g.glo[0], j = p[1] - g.glo[1], k = g.dim == 3 ? p[2] - g.glo[2] : 0;
    return g.cls[i + (long long)g.gn[0] * (j + (long long)g.gn[1] * k)];
}

__device__ __forceinline__ int node_class(const Geo& g, const int p[3], bool* iface) {
    const unsigned char m = cls_at(g, p);
    *iface = (m & CLS_IFACE) != 0;
    if (m & CLS_FILL) return NODE_FILL;
    if (m & CLS_COVERED) return NODE_COVERED;
    return NODE_UNKNOWN;
}

__device__ __forceinline__ long long at(const Lay& L, int i, int j, int k) { return i + L.px * j + L.pl * k; }

// Fine level (layout Lf, geometry gf) from the coarser level (Lc, first node gclo):
// mode 0 assigns the interpolation at slave and ghost nodes, mode 1 adds it at every node.
template <typename T>
__global__ __launch_bounds__(CX* CY) void k_fill(Lay Lf, Geo gf, Lay Lc, int gclo0, int gclo1, int gclo2,
                                                 const T* __restrict__ crse, T* __restrict__ fine, int nc, int mode) {
    const int i = blockIdx.x * CX + threadIdx.x, j = blockIdx.y * CY + threadIdx.y, k = blockIdx.z;
    if (i >= Lf.nx || j >= Lf.ny) return;
    const int p[3] = {gf.glo[0] + i, gf.glo[1] + j, gf.glo[2] + k};
    if (mode == 0) {
        bool iface;
        if (node_class(gf, p, &iface) != NODE_FILL) return;
    }
    const long long oc = at(Lc, (p[0] >> 1) - gclo0, (p[1] >> 1) - gclo1, (p[2] >> 1) - gclo2);
    const long long of = at(Lf, i, j, k);
    for (int n = 0; n < nc; ++n) {
        const T v = prolong_node<T>(Lc, p[0], p[1], p[2], crse + n * Lc.cs + oc);
        if (mode == 0) fine[n * Lf.cs + of] = v;
        else fine[n * Lf.cs + of] += v;
    }
}

// Coarse nodes under level l+1 (coarse geometry gc) take the coincident fine value.
template <typename T>
__global__ __launch_bounds__(CX* CY) void k_inject(Lay Lc, Geo gc, Lay Lf, int gflo0, int gflo1, int gflo2,
                                                   const T* __restrict__ fine, T* __restrict__ crse, int nc) {
    const int a = blockIdx.x * CX + threadIdx.x, b = blockIdx.y * CY + threadIdx.y, c = blockIdx.z;
    const int I = gc.ulo[0] + a, J = gc.ulo[1] + b, K = gc.ulo[2] + c;
    if (I > gc.uhi[0] || J > gc.uhi[1]) return;
    const int p[3] = {I, J, K};
    if (!(cls_at(gc, p) & CLS_UNDER)) return;
    const long long oc = at(Lc, I - gc.glo[0], J - gc.glo[1], K - gc.glo[2]);
    const long long of = at(Lf, 2 * I - gflo0, 2 * J - gflo1, 2 * K - gflo2);
    for (int n = 0; n < nc; ++n) crse[n * Lc.cs + oc] = fine[n * Lf.cs + of];
}

// Reflux into the coarse nodes under level l+1: the full weighting of the fine
// field v (read on G_{l+1}), branches by the coarse domain (Operator.cpp:779-822).
// mode 0: crse = R v.  mode 1 (a FAC down-sweep): crse = (interface ? crse : R v) - R w.
template <typename T, int DIM>
__global__ __launch_bounds__(CX* CY) void k_reflux_c(Lay Lc, Geo gc, Lay Lf, int gflo0, int gflo1, int gflo2,
                                                     const T* __restrict__ v, const T* __restrict__ w,
                                                     T* __restrict__ crse, int nc, int mode) {
    const int a = blockIdx.x * CX + threadIdx.x, b = blockIdx.y * CY + threadIdx.y, c = blockIdx.z;
    const int I = gc.ulo[0] + a, J = gc.ulo[1] + b, K = gc.ulo[2] + c;
    if (I > gc.uhi[0] || J > gc.uhi[1]) return;
    {
        const int q[3] = {I, J, K};
        if (!(cls_at(gc, q) & CLS_UNDER)) return;
    }
    Lay Ld = Lc;  // the branches test the coarse domain boundary
    Ld.nx = gc.N[0] + 1;
    Ld.ny = gc.N[1] + 1;
    Ld.nz = DIM == 3 ? gc.N[2] + 1 : 1;
    const int p[3] = {I, J, K};
    bool iface = false;
    (void)node_class(gc, p, &iface);
    const long long oc = at(Lc, I - gc.glo[0], J - gc.glo[1], K - gc.glo[2]);
    const int fx = 2 * I - gflo0, fy = 2 * J - gflo1, fz = 2 * K - gflo2;
    for (int n = 0; n < nc; ++n) {
        auto getw = [&](int di, int dj, int dk) -> T { return w[n * Lf.cs + at(Lf, fx + di, fy + dj, fz + dk)]; };
        auto getv = [&](int di, int dj, int dk) -> T { return v[n * Lf.cs + at(Lf, fx + di, fy + dj, fz + dk)]; };
        T out;
        if (mode == 0) out = restrict_node<T, DIM>(Ld, I, J, K, getv);
        else {
            const T base = iface ? crse[n * Lc.cs + oc] : restrict_node<T, DIM>(Ld, I, J, K, getv);
            out = base - restrict_node<T, DIM>(Ld, I, J, K, getw);
        }
        crse[n * Lc.cs + oc] = out;
    }
}

// Zero every component at the nodes that are not composite unknowns.
template <typename T>
__global__ __launch_bounds__(CX* CY) void k_mask(Lay L, Geo g, T* __restrict__ x, int nc) {
    const int i = blockIdx.x * CX + threadIdx.x, j = blockIdx.y * CY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    const int p[3] = {g.glo[0] + i, g.glo[1] + j, g.glo[2] + k};
    bool iface;
    if (node_class(g, p, &iface) == NODE_UNKNOWN) return;
    for (int n = 0; n < nc; ++n) x[n * L.cs + at(L, i, j, k)] = T(0);
}

// G_l field -> right-hand side of the level's multigrid hierarchy on its union
// box (first node `lo`, global), converted to TO; zero at fill nodes (slaves:
// their rows are homogeneous Dirichlet rows of the box problem).
template <typename TO>
__global__ __launch_bounds__(CX* CY) void k_to_box(Lay Lg, Geo g, int lo0, int lo1, int lo2,
                                                   const double* __restrict__ src, Lay Lb, TO* __restrict__ dst,
                                                   int nc) {
    const int i = blockIdx.x * CX + threadIdx.x, j = blockIdx.y * CY + threadIdx.y, k = blockIdx.z;
    if (i >= Lb.nx || j >= Lb.ny) return;
    const int p[3] = {lo0 + i, lo1 + j, lo2 + k};
    const bool fill = (cls_at(g, p) & CLS_FILL) != 0;
    const long long os = at(Lg, p[0] - g.glo[0], p[1] - g.glo[1], p[2] - g.glo[2]);
    for (int n = 0; n < nc; ++n) dst[n * Lb.cs + at(Lb, i, j, k)] = fill ? TO(0) : (TO)src[n * Lg.cs + os];
}

// The hierarchy's correction -> G_l on the union box (assign or add).
template <typename TI>
__global__ __launch_bounds__(CX* CY) void k_from_box(Lay Lb, const TI* __restrict__ src, Lay Lg, int o0, int o1,
                                                     int o2, double* __restrict__ dst, int nc, int add) {
    const int i = blockIdx.x * CX + threadIdx.x, j = blockIdx.y * CY + threadIdx.y, k = blockIdx.z;
    if (i >= Lb.nx || j >= Lb.ny) return;
    const long long od = at(Lg, i + o0, j + o1, k + o2);
    for (int n = 0; n < nc; ++n) {
        const double v = (double)src[n * Lb.cs + at(Lb, i, j, k)];
        if (add) dst[n * Lg.cs + od] += v;
        else dst[n * Lg.cs + od] = v;
    }
}

// The union box of G's stiffness planes (offset o) into the hierarchy's level 0.
__global__ __launch_bounds__(CX* CY) void k_subbox(Lay Ls, int o0, int o1, int o2, const double* __restrict__ src,
                                                   Lay Ld, double* __restrict__ dst, int nc) {
    const int i = blockIdx.x * CX + threadIdx.x, j = blockIdx.y * CY + threadIdx.y, k = blockIdx.z;
    if (i >= Ld.nx || j >= Ld.ny) return;
    for (int n = 0; n < nc; ++n) dst[n * Ld.cs + at(Ld, i, j, k)] = src[n * Ls.cs + at(Ls, i + o0, j + o1, k + o2)];
}

// A Jacobi sweep's output keeps the input at fill nodes (they are the sweep's
// Dirichlet data: zero in a down-sweep, the interpolated coarse correction in an up-sweep).
template <typename T>
__global__ __launch_bounds__(CX* CY) void k_hold_fill(Lay L, Geo g, const T* __restrict__ xin, T* __restrict__ xout,
                                                      int nc) {
    const int i = blockIdx.x * CX + threadIdx.x, j = blockIdx.y * CY + threadIdx.y, k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    const int p[3] = {g.glo[0] + i, g.glo[1] + j, g.glo[2] + k};
    if (!(cls_at(g, p) & CLS_FILL)) return;
    const long long o = at(L, i, j, k);
    for (int n = 0; n < nc; ++n) xout[n * L.cs + o] = xin[n * L.cs + o];
}

// max |x| over a flat array into *out (bit pattern of a non-negative double).
__global__ __launch_bounds__(256) void k_maxabs(const double* __restrict__ x, long long n,
                                                 unsigned long long* __restrict__ out) {
    __shared__ double s[256];
    double m = 0;
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
        const double a = fabs(x[q]);
        m = (a > m || a != a) ? a : m;
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            const double o = s[threadIdx.x + st];
            if (o > s[threadIdx.x] || o != o) s[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double r = s[0];
        if (r != r) r = __builtin_huge_val();
        atomicMax(out, (unsigned long long)__double_as_longlong(r));
    }
}

// y += x over a flat array
__global__ __launch_bounds__(256) void k_add_flat(double* __restrict__ y, const double* __restrict__ x, long long n) {
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) y[q] += x[q];
}

dim3 grid_of(const Lay& L) { return dim3((L.nx + CX - 1) / CX, (L.ny + CY - 1) / CY, L.nz); }

// Faces of BC class c (numbering of BC::Operator::Elastic::Face): side[d] = 0 (low),
// 1 (high) or -1 (not on a face in direction d).
void class_sides(int dim, int c, int side[3]) {
    side[0] = side[1] = side[2] = -1;
    if (dim == 2) {
        static const int tab[8][2] = {{0, -1}, {-1, 0}, {1, -1}, {-1, 1}, {0, 0}, {0, 1}, {1, 0}, {1, 1}};
        side[0] = tab[c][0];
        side[1] = tab[c][1];
        return;
    }
    if (c < 6) {
        side[c % 3] = c < 3 ? 0 : 1;
    } else if (c < 10) {
        side[1] = ((c - 6) >> 1) & 1;
        side[2] = (c - 6) & 1;
    } else if (c < 14) {
        side[2] = ((c - 10) >> 1) & 1;
        side[0] = (c - 10) & 1;
    } else if (c < 18) {
        side[0] = ((c - 14) >> 1) & 1;
        side[1] = (c - 14) & 1;
    } else {
        side[0] = ((c - 18) >> 2) & 1;
        side[1] = ((c - 18) >> 1) & 1;
        side[2] = (c - 18) & 1;
    }
}

struct CLev {
    Geo geo{};
    Lay L{};
    double h[3] = {1, 1, 1};
    size_t off = 0;     // element offset of the level's segment in composite vectors
    size_t elems = 0;   // dim * L.cs
    BCTab bcG{};        // boundary rows of G (domain classes; Displacement on artificial faces)
    BCTab* dbcG = nullptr;  // its device copy (the z-marching kernels read the table from memory)
    double* C = nullptr;
    bool model_set = false;
    std::vector<int> boxes;          // lo[3], hi[3] per box (level l >= 1)
    unsigned char* cls = nullptr;    // node classes of G on the device
    double* dg = nullptr;            // diagonal of A_l on G (levels >= 1, FP64)
    // mixed precision, a level that smooths: its stiffness and inverse diagonal in FP32, and FP32 work
    // arrays for the iterate, the sweep output and the right-hand side
    float *C32 = nullptr, *dg32 = nullptr, *x32 = nullptr, *t32 = nullptr, *r32 = nullptr;
    int nu1 = 2, nu2 = 2;            // Fsmooth calls of the FAC cycle (smoothed levels)
    // a level whose union is one box (level 0: the domain) carries a multigrid
    // hierarchy on that box; its FAC step is a cycle of it instead of smoothing
    ela_hier* H = nullptr;
    Lay Lb{};                        // the hierarchy's level 0
    int ulo[3] = {0, 0, 0};          // the union's bounding box (global nodes)
    int uhi[3] = {0, 0, 0};
};

}  // namespace
}  // namespace ela

using namespace ela;

struct ela_amr {
    ela_ctx* ctx = nullptr;
    int dim = 3, kind = 0, nc = 2, mode = 0, nlev = 1, cprec = ELA_F64;
    int uniform = 0;
    std::vector<CLev> lev;
    BCTab bc{};
    int bctypes[26][3] = {};
    double bcval[26][3] = {};
    hipStream_t stream = nullptr;
    size_t ntot = 0;
    double *u = nullptr, *b = nullptr, *r = nullptr, *z = nullptr, *w = nullptr, *t = nullptr, *q = nullptr;
    int krylov = 8;
    std::vector<double*> V, Z;
    std::vector<float*> V32;  // mixed precision: the Krylov basis in FP32
    double* kw = nullptr;     // and w = A z in FP64
    double *kdev = nullptr, *kpart = nullptr;
    unsigned long long* dnorm = nullptr;
    unsigned long long* hnorm = nullptr;
    std::vector<std::pair<void*, size_t>> allocs;
    size_t bytes = 0;
    bool prepared = false;
};

namespace {

int afail(int code, const std::string& m) { return set_error(code, m); }

int aalloc(ela_amr* a, size_t bytes, void** out) {
    *out = nullptr;
    if (int rc = budget_reserve(a->ctx, bytes)) return rc;
    if (hipMalloc(out, bytes) != hipSuccess) {
        budget_release(a->ctx, bytes);
        *out = nullptr;
        return afail(ELA_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    if (hipMemsetAsync(*out, 0, bytes, a->stream) != hipSuccess) return afail(ELA_ERR_HIP, "hipMemsetAsync failed");
    a->allocs.push_back({*out, bytes});
    a->bytes += bytes;
    return ELA_OK;
}

template <typename T>
int aalloc_t(ela_amr* a, size_t count, T** out) {
    void* p = nullptr;
    int rc = aalloc(a, count * sizeof(T), &p);
    *out = (T*)p;
    return rc;
}

void afree_all(ela_amr* a) {
    for (auto& pr : a->allocs) {
        (void)hipFree(pr.first);
        budget_release(a->ctx, pr.second);
    }
    a->allocs.clear();
    a->bytes = 0;
}

double* seg(double* base, const CLev& v) { return base + v.off; }

// A level's operator on G: 3D isotropic planes run the z-marching tiles
// (k_plane_sweep: the node kernels' expressions, so the same bits), anything
// else the node kernels.
bool level_march(const ela_amr* a, const CLev& v) {
    return a->dim == 3 && a->kind == K_ISO && v.dbcG && grain_offsets_fit(v.L) && plane_offsets_fit(v.L, 2, 8);
}
constexpr int kLevelKz = 16;  // planes per marching tile on the levels' arrays
template <typename T = double>
GrainBox<T> level_box(const ela_amr* a, const CLev& v) {
    GrainBox<T> g{};
    g.L = v.L;
    g.iv = make_inv<T>(v.h);
    g.bc = v.dbcG;
    g.uniform = a->uniform;
    if constexpr (sizeof(T) == 8) g.C = v.C;
    else g.C = v.C32;
    g.sm = nullptr;
    g.xcd = 1;
    g.ry = 1;
    g.pf = 1;
    g.nt = false;
    g.xface = true;
    return g;
}
void lev_apply(const ela_amr* a, const CLev& v, const double* x, double* y) {
    if (level_march(a, v)) {
        GrainBox<double> g = level_box(a, v);
        g.in = x;
        g.out = y;
        launch_plane<double, 3, double>(g, kLevelKz, nullptr, nullptr, a->stream);
    } else {
        launch_fapply<double>(v.L, a->kind, v.C, x, y, v.bcG, v.h, a->uniform, a->stream);
    }
}
void lev_residual(const ela_amr* a, const CLev& v, const double* x, const double* b, double* r) {
    if (level_march(a, v)) {
        GrainBox<double> g = level_box(a, v);
        g.in = x;
        g.rhs = b;
        g.out = r;
        launch_plane<double, 1, double>(g, kLevelKz, nullptr, nullptr, a->stream);
    } else {
        launch_residual<double, double>(v.L, a->kind, v.C, x, b, r, (double*)nullptr, nullptr, nullptr, nullptr,
                                        v.bcG, v.h, a->uniform, a->stream);
    }
}
void lev_jacobi(const ela_amr* a, const CLev& v, const double* x, const double* b, double* out) {
    if (level_march(a, v)) {
        GrainBox<double> g = level_box(a, v);
        g.in = x;
        g.rhs = b;
        g.out = out;
        g.dg = v.dg;
        g.omega = 2.0 / 3.0;
        launch_plane<double, 0, double>(g, kLevelKz, nullptr, nullptr, a->stream);
    } else {
        launch_jacobi<double>(v.L, a->kind, v.C, v.dg, x, b, out, v.bcG, v.h, a->uniform, 2.0 / 3.0, a->stream);
    }
}
void lev_jacobi32(const ela_amr* a, const CLev& v, const float* x, const float* b, float* out) {
    if (level_march(a, v)) {
        GrainBox<float> g = level_box<float>(a, v);
        g.in = x;
        g.rhs = b;
        g.out = out;
        g.dg = v.dg32;
        g.omega = 2.0f / 3.0f;
        launch_plane<float, 0, float>(g, kLevelKz, nullptr, nullptr, a->stream);
    } else {
        launch_jacobi<float>(v.L, a->kind, v.C32, v.dg32, x, b, out, v.bcG, v.h, a->uniform, 2.0 / 3.0, a->stream);
    }
}

// Launch grid over the bounding box of a level's nodes under the next finer level.
dim3 under_grid(const ela_amr* a, const CLev& c) {
    return dim3((c.geo.uhi[0] - c.geo.ulo[0] + CX) / CX, (c.geo.uhi[1] - c.geo.ulo[1] + CY) / CY,
                a->dim == 3 ? c.geo.uhi[2] - c.geo.ulo[2] + 1 : 1);
}

// slaves and ghosts of every finer level from the coarser one, then covered nodes
// from the finer one
void fill(ela_amr* a, double* x) {
    for (int l = 1; l < a->nlev; ++l) {
        CLev &f = a->lev[l], &c = a->lev[l - 1];
        hipLaunchKernelGGL(k_fill<double>, grid_of(f.L), dim3(CX, CY), 0, a->stream, f.L, f.geo, c.L, c.geo.glo[0],
                           c.geo.glo[1], c.geo.glo[2], (const double*)seg(x, c), seg(x, f), a->dim, 0);
    }
    for (int l = a->nlev - 1; l >= 1; --l) {
        CLev &f = a->lev[l], &c = a->lev[l - 1];
        hipLaunchKernelGGL(k_inject<double>, under_grid(a, c), dim3(CX, CY), 0, a->stream, c.L, c.geo, f.L,
                           f.geo.glo[0], f.geo.glo[1], f.geo.glo[2], (const double*)seg(x, f), seg(x, c), a->dim);
    }
}

void reflux(ela_amr* a, int l, const double* v, const double* w, double* crse, int mode) {
    CLev &f = a->lev[l], &c = a->lev[l - 1];
    const dim3 g = under_grid(a, c);
    if (a->dim == 3)
        hipLaunchKernelGGL((k_reflux_c<double, 3>), g, dim3(CX, CY), 0, a->stream, c.L, c.geo, f.L, f.geo.glo[0],
                           f.geo.glo[1], f.geo.glo[2], v, w, crse, a->dim, mode);
    else
        hipLaunchKernelGGL((k_reflux_c<double, 2>), g, dim3(CX, CY), 0, a->stream, c.L, c.geo, f.L, f.geo.glo[0],
                           f.geo.glo[1], f.geo.glo[2], v, w, crse, a->dim, mode);
}

void mask(ela_amr* a, double* x) {
    for (auto& v : a->lev)
        hipLaunchKernelGGL(k_mask<double>, grid_of(v.L), dim3(CX, CY), 0, a->stream, v.L, v.geo, seg(x, v), a->dim);
}

void maxabs(ela_amr* a, const double* x) {
    launch_zero_u64(a->dnorm, 1, a->stream);
    const long long n = (long long)a->ntot;
    const int blocks = (int)std::min<long long>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_maxabs, dim3(blocks), dim3(256), 0, a->stream, x, n, a->dnorm);
}

double read_norm(ela_amr* a) {
    (void)hipMemcpyAsync(a->hnorm, a->dnorm, sizeof(unsigned long long), hipMemcpyDeviceToHost, a->stream);
    (void)hipStreamSynchronize(a->stream);
    double v;
    std::memcpy(&v, a->hnorm, sizeof v);
    return v;
}

// r = composite residual of x (masked), |r|inf into dnorm
void residual(ela_amr* a, double* x, const double* bb, double* r) {
    fill(a, x);
    for (auto& v : a->lev) lev_residual(a, v, seg(x, v), bb + v.off, seg(r, v));
    for (int l = a->nlev - 1; l >= 1; --l) reflux(a, l, seg(r, a->lev[l]), nullptr, seg(r, a->lev[l - 1]), 0);
    mask(a, r);
    maxabs(a, r);
}

// y = A_c x (masked)
void apply(ela_amr* a, double* x, double* y) {
    fill(a, x);
    for (auto& v : a->lev) lev_apply(a, v, seg(x, v), seg(y, v));
    for (int l = a->nlev - 1; l >= 1; --l) reflux(a, l, seg(y, a->lev[l]), nullptr, seg(y, a->lev[l - 1]), 0);
    mask(a, y);
}

// dst (+)= one multigrid cycle of the level's box hierarchy on the level's residual src
int box_correct(ela_amr* a, int l, const double* src, double* dst, int add) {
    CLev& v = a->lev[l];
    const dim3 gb = grid_of(v.Lb);
    void* res = hier_res0(v.H, a->cprec);
    if (a->cprec == ELA_F32)
        hipLaunchKernelGGL(k_to_box<float>, gb, dim3(CX, CY), 0, a->stream, v.L, v.geo, v.ulo[0], v.ulo[1], v.ulo[2],
                           src, v.Lb, (float*)res, a->dim);
    else
        hipLaunchKernelGGL(k_to_box<double>, gb, dim3(CX, CY), 0, a->stream, v.L, v.geo, v.ulo[0], v.ulo[1],
                           v.ulo[2], src, v.Lb, (double*)res, a->dim);
    if (int rc = hier_cycle(v.H, a->cprec)) return rc;
    const void* cor = hier_cor0(v.H, a->cprec);
    const int o0 = v.ulo[0] - v.geo.glo[0], o1 = v.ulo[1] - v.geo.glo[1], o2 = v.ulo[2] - v.geo.glo[2];
    if (a->cprec == ELA_F32)
        hipLaunchKernelGGL(k_from_box<float>, gb, dim3(CX, CY), 0, a->stream, v.Lb, (const float*)cor, v.L, o0, o1,
                           o2, dst, a->dim, add);
    else
        hipLaunchKernelGGL(k_from_box<double>, gb, dim3(CX, CY), 0, a->stream, v.Lb, (const double*)cor, v.L, o0, o1,
                           o2, dst, a->dim, add);
    return ELA_OK;
}

// ncalls Fsmooth calls (2 weighted-Jacobi sweeps each, Operator.cpp:90-162) of
// A_l x = rhs on level l >= 1: the non-fill nodes are updated, the fill nodes
// keep their values.  tmp: work segment of the level.
void level_smooth(ela_amr* a, int l, double* x, const double* rhs, double* tmp, int ncalls) {
    CLev& f = a->lev[l];
    const dim3 g = grid_of(f.L);
    if (f.x32) {  // mixed: the sweeps in FP32 (fill nodes held at their rounded values, widened back out)
        const long long n = (long long)f.elems;
        launch_convert<double, float>(n, rhs, f.r32, a->stream);
        launch_convert<double, float>(n, x, f.x32, a->stream);
        for (int c = 0; c < ncalls; ++c) {
            lev_jacobi32(a, f, f.x32, f.r32, f.t32);
            hipLaunchKernelGGL(k_hold_fill<float>, g, dim3(CX, CY), 0, a->stream, f.L, f.geo, (const float*)f.x32,
                               f.t32, a->dim);
            lev_jacobi32(a, f, f.t32, f.r32, f.x32);
            hipLaunchKernelGGL(k_hold_fill<float>, g, dim3(CX, CY), 0, a->stream, f.L, f.geo, (const float*)f.t32,
                               f.x32, a->dim);
        }
        launch_convert<float, double>(n, f.x32, x, a->stream);
        return;
    }
    for (int c = 0; c < ncalls; ++c) {
        lev_jacobi(a, f, x, rhs, tmp);
        hipLaunchKernelGGL(k_hold_fill<double>, g, dim3(CX, CY), 0, a->stream, f.L, f.geo, (const double*)x, tmp,
                           a->dim);
        lev_jacobi(a, f, tmp, rhs, x);
        hipLaunchKernelGGL(k_hold_fill<double>, g, dim3(CX, CY), 0, a->stream, f.L, f.geo, (const double*)tmp, x,
                           a->dim);
    }
}

// z = FAC(v): one multiplicative FAC cycle over the AMR levels (t, w, q: work vectors)
int fac(ela_amr* a, const double* v, double* z, const float* v32 = nullptr) {
    hipStream_t s = a->stream;
    const int L = a->nlev - 1;
    if (hipMemsetAsync(z, 0, a->ntot * sizeof(double), s) != hipSuccess) return afail(ELA_ERR_HIP, "memset");
    if (v32) launch_convert<float, double>((long long)a->ntot, v32, a->t, s);
    else if (hipMemcpyAsync(a->t, v, a->ntot * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess)
        return afail(ELA_ERR_HIP, "copy");
    for (int l = L; l >= 1; --l) {
        CLev& f = a->lev[l];
        if (f.H) {
            if (int rc = box_correct(a, l, seg(a->t, f), seg(z, f), 0)) return rc;
        } else {
            level_smooth(a, l, seg(z, f), seg(a->t, f), seg(a->q, f), f.nu1);
        }
        lev_apply(a, f, seg(z, f), seg(a->w, f));
        reflux(a, l, seg(a->t, f), seg(a->w, f), seg(a->t, a->lev[l - 1]), 1);
    }
    if (int rc = box_correct(a, 0, seg(a->t, a->lev[0]), seg(z, a->lev[0]), 0)) return rc;
    for (int l = 1; l <= L; ++l) {
        CLev &f = a->lev[l], &c = a->lev[l - 1];
        hipLaunchKernelGGL(k_fill<double>, grid_of(f.L), dim3(CX, CY), 0, s, f.L, f.geo, c.L, c.geo.glo[0],
                           c.geo.glo[1], c.geo.glo[2], (const double*)seg(z, c), seg(z, f), a->dim, 1);
        if (f.H) {
            lev_residual(a, f, seg(z, f), seg(a->t, f), seg(a->w, f));
            if (int rc = box_correct(a, l, seg(a->w, f), seg(z, f), 1)) return rc;
        } else {
            level_smooth(a, l, seg(z, f), seg(a->t, f), seg(a->q, f), f.nu2);
        }
    }
    return ELA_OK;
}

int krylov_alloc(ela_amr* a) {
    const int m = a->krylov;
    const bool b32 = a->cprec == ELA_F32;
    if ((int)a->Z.size() == m && (b32 ? (int)a->V32.size() : (int)a->V.size()) == m + 1) return ELA_OK;
    if (m > kMaxKrylov - 1) return afail(ELA_ERR_BADARG, "restart too long");
    a->V.clear();
    a->V32.clear();
    a->Z.clear();
    int rc = ELA_OK;
    for (int q = 0; q <= m && !rc; ++q) {
        if (b32) {
            float* p = nullptr;
            rc = aalloc_t(a, a->ntot, &p);
            a->V32.push_back(p);
        } else {
            double* p = nullptr;
            rc = aalloc_t(a, a->ntot, &p);
            a->V.push_back(p);
        }
    }
    if (!rc && b32 && !a->kw) rc = aalloc_t(a, a->ntot, &a->kw);
    for (int q = 0; q < m && !rc; ++q) {
        double* p = nullptr;
        rc = aalloc_t(a, a->ntot, &p);
        a->Z.push_back(p);
    }
    if (!rc && !a->kdev) rc = aalloc_t(a, 4 * (size_t)(kMaxKrylov + 2), &a->kdev);
    if (!rc && !a->kpart) rc = aalloc_t(a, (size_t)mdot_partials(kMaxKrylov), &a->kpart);
    return rc;
}

std::vector<double> host_vec(ela_amr* a, int count) {
    std::vector<double> out(count);
    (void)hipMemcpyAsync(out.data(), a->kdev, sizeof(double) * count, hipMemcpyDeviceToHost, a->stream);
    (void)hipStreamSynchronize(a->stream);
    return out;
}

// Right-preconditioned flexible GMRES(m) on the composite unknowns (classical
// Gram-Schmidt with a second pass when the new vector lost most of its norm);
// m = 0 runs the plain FAC iteration u += FAC(r).  Mixed precision keeps the
// basis vectors in FP32.
int solve(ela_amr* a, double tol_rel, double tol_abs, int max_iter, ela_stats* st) {
    hipStream_t s = a->stream;
    const long long n = (long long)a->ntot;
    residual(a, a->u, a->b, a->r);
    double rn = read_norm(a);
    const double r0 = rn;
    if (!(r0 == r0)) return afail(ELA_ERR_NAN, "initial residual is not finite");
    const double target = std::max(tol_abs, std::max(tol_rel, 1e-16) * r0);
    int total = 0, status = ELA_OK, restarts = 0;
    double best = r0;
    const int m = a->krylov;
    if (m > 0)
        if (int rc = krylov_alloc(a)) return rc;
    std::vector<double> H((size_t)(m + 1) * std::max(m, 1)), cs(m), sn(m), g(m + 1);
    std::vector<double> sv(m + 1, 1.0);  // FP32 basis: stored vector q = sv[q] x the unit vector
    while (rn > target) {
        if (total >= max_iter) {
            status = afail(ELA_ERR_NONCONVERGED, "max_iter reached at " + std::to_string(rn / r0) + " relative");
            break;
        }
        if (m == 0) {
            if (int rc = fac(a, a->r, a->z)) return rc;
            hipLaunchKernelGGL(k_add_flat, dim3((int)std::min<long long>(4096, (n + 255) / 256)), dim3(256), 0, s,
                               a->u, (const double*)a->z, n);
            ++total;
        } else {
            // mixed precision keeps the basis in FP32 (w = A z, the dots, the updates and the solution in FP64)
            const bool b32 = a->cprec == ELA_F32;
            VecList one{};
            one.v[0] = a->r;
            launch_mdot(a->r, one, 1, n, a->kpart, a->kdev, s);
            const double beta = std::sqrt(host_vec(a, 1)[0]);
            if (!(beta > 0)) break;
            if (b32) launch_scale_to2f(a->r, 1.0 / beta, a->V32[0], nullptr, n, s);
            else launch_scale_flat(a->V[0], a->r, 1.0 / beta, n, s);
            sv[0] = 1.0;
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            int jend = 0;
            for (int j = 0; j < m && total < max_iter; ++j) {
                if (int rc = b32 ? fac(a, nullptr, a->Z[j], a->V32[j]) : fac(a, a->V[j], a->Z[j])) return rc;
                double* const w = b32 ? a->kw : a->V[j + 1];
                apply(a, a->Z[j], w);
                VecList Vl{};
                VecList32 Vl32{};
                for (int q = 0; q <= j; ++q) {
                    if (b32) Vl32.v[q] = a->V32[q];
                    else Vl.v[q] = a->V[q];
                }
                // h1 and |w|^2 in one pass (w last in the list); |w'|^2 with the update
                if (b32) {
                    launch_mdot_self(w, Vl32, j + 1, n, a->kpart, a->kdev, s);
                } else {
                    VecList Vw = Vl;
                    Vw.v[j + 1] = w;
                    launch_mdot(w, Vw, j + 2, n, a->kpart, a->kdev, s);
                }
                std::vector<double> h1 = host_vec(a, j + 2), h2(j + 1, 0.0);
                const double wn0 = std::sqrt(h1[j + 1]);
                h1.resize(j + 1);
                // the FP32 basis holds s v (the single box's solve_fgmres): unit-vector dots, coefficients / s^2
                CoefMul mul = coef_ones();
                if (b32)
                    for (int q = 0; q <= j; ++q) {
                        h1[q] /= sv[q];
                        mul.f[q] = 1.0 / (sv[q] * sv[q]);
                    }
                double hn2 = wn0 * wn0;
                for (int q = 0; q <= j; ++q) hn2 -= h1[q] * h1[q];
                double* const norm2 = a->kdev + 2 * (kMaxKrylov + 2);  // beyond the coefficients
                auto pow2_inv = [](double e2, double fallback2) {
                    const double e = std::sqrt(e2 > 0 ? e2 : fallback2);
                    return e > 0 && std::isfinite(e) ? std::exp2(-std::floor(std::log2(e))) : 1.0;
                };
                double osc = 1.0;
                if (!(hn2 > kReorth2 * wn0 * wn0)) {
                    if (b32) {
                        launch_maxpy32(w, Vl32, j + 1, a->kdev, mul, -1.0, n, s);
                        launch_mdot_self(w, Vl32, j + 1, n, a->kpart, a->kdev, s);
                        std::vector<double> r2 = host_vec(a, j + 2);
                        double e2 = r2[j + 1];
                        for (int q = 0; q <= j; ++q) e2 -= (r2[q] / sv[q]) * (r2[q] / sv[q]);
                        osc = pow2_inv(e2, r2[j + 1]);
                        launch_maxpy_norm32_to2f(w, Vl32, j + 1, a->kdev, mul, -1.0, n, osc, a->V32[j + 1], nullptr,
                                                 a->kpart, norm2, s);
                        h2 = host_vec(a, j + 1);
                        for (int q = 0; q <= j; ++q) h2[q] /= sv[q];
                    } else {
                        launch_maxpy(w, Vl, j + 1, a->kdev, -1.0, n, s);
                        launch_mdot(w, Vl, j + 1, n, a->kpart, a->kdev, s);
                        launch_maxpy_norm(w, Vl, j + 1, a->kdev, -1.0, n, a->kpart, norm2, s);
                        h2 = host_vec(a, j + 1);
                    }
                } else {
                    if (b32) {
                        osc = pow2_inv(hn2, wn0 * wn0);
                        launch_maxpy_norm32_to2f(w, Vl32, j + 1, a->kdev, mul, -1.0, n, osc, a->V32[j + 1], nullptr,
                                                 a->kpart, norm2, s);
                    } else {
                        launch_maxpy_norm(w, Vl, j + 1, a->kdev, -1.0, n, a->kpart, norm2, s);
                    }
                }
                double hn2_dev = 0.0;
                ELA_HIPCHK(hipMemcpyAsync(&hn2_dev, norm2, sizeof(double), hipMemcpyDeviceToHost, s));
                ELA_HIPCHK(hipStreamSynchronize(s));
                const double hn = std::sqrt(hn2_dev);
                ++total;
                for (int q = 0; q <= j; ++q) H[(size_t)q * m + j] = h1[q] + h2[q];
                H[(size_t)(j + 1) * m + j] = hn;
                if (b32) sv[j + 1] = osc * hn;  // the update wrote osc hn v_{j+1}
                else if (hn > 0) launch_scale_flat(a->V[j + 1], a->V[j + 1], 1.0 / hn, n, s);
                for (int q = 0; q < j; ++q) {
                    const double x = H[(size_t)q * m + j], y = H[(size_t)(q + 1) * m + j];
                    H[(size_t)q * m + j] = cs[q] * x + sn[q] * y;
                    H[(size_t)(q + 1) * m + j] = -sn[q] * x + cs[q] * y;
                }
                const double x = H[(size_t)j * m + j], y = H[(size_t)(j + 1) * m + j];
                const double rr = std::hypot(x, y);
                cs[j] = rr > 0 ? x / rr : 1.0;
                sn[j] = rr > 0 ? y / rr : 0.0;
                H[(size_t)j * m + j] = rr;
                H[(size_t)(j + 1) * m + j] = 0.0;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                jend = j + 1;
                const double est = std::fabs(g[j + 1]);
                if (!(est == est)) break;
                if (est <= 0.5 * target || hn == 0) break;
            }
            std::vector<double> y(jend);
            for (int q = jend - 1; q >= 0; --q) {
                double acc = g[q];
                for (int k = q + 1; k < jend; ++k) acc -= H[(size_t)q * m + k] * y[k];
                y[q] = acc / H[(size_t)q * m + q];
            }
            if (jend > 0) {
                ELA_HIPCHK(hipMemcpyAsync(a->kdev, y.data(), sizeof(double) * jend, hipMemcpyHostToDevice, s));
                VecList Zl{};
                for (int q = 0; q < jend; ++q) Zl.v[q] = a->Z[q];
                launch_maxpy(a->u, Zl, jend, a->kdev, 1.0, n, s);
            }
        }
        residual(a, a->u, a->b, a->r);
        rn = read_norm(a);
        ++restarts;
        if (!(rn == rn)) { status = afail(ELA_ERR_NAN, "residual became non-finite"); break; }
        if (rn <= target) break;
        if (rn < 0.5 * best) best = rn;
        else if (restarts >= (m == 0 ? 12 : 4) && rn > 0.9 * best) {
            status = afail(ELA_ERR_NONCONVERGED, "residual stagnated at " + std::to_string(rn / r0) + " relative");
            break;
        }
    }
    fill(a, a->u);
    if (hipStreamSynchronize(s) != hipSuccess) return afail(ELA_ERR_HIP, "composite solve failed");
    if (st) {
        st->iters = total;
        st->converged = rn <= target;
        st->res0 = r0;
        st->res = rn;
        st->mg_levels = a->nlev;
        st->mode = a->mode;
    }
    return status;
}

int check_lev(const ela_amr* a, int l) {
    if (!a) return afail(ELA_ERR_BADARG, "null composite");
    if (l < 0 || l >= a->nlev) return afail(ELA_ERR_BADARG, "level out of range");
    return ELA_OK;
}

// dense host array (nc planes over G_l) <-> the level's layout
int xfer(ela_amr* a, const CLev& v, int nc, const double* host, double* dev, double* host_out, const double* dev_in) {
    const size_t nn = (size_t)v.L.nx * v.L.ny * v.L.nz;
    void* stage = nullptr;
    if (int rc = aalloc(a, nn * nc * sizeof(double), &stage)) return rc;
    hipError_t e = hipSuccess;
    if (host) {
        e = hipMemcpyAsync(stage, host, nn * nc * sizeof(double), hipMemcpyHostToDevice, a->stream);
        if (e == hipSuccess) launch_pack<double>(v.L, nc, (const double*)stage, dev, 1, a->stream);
    } else {
        launch_pack<double>(v.L, nc, (const double*)stage, (double*)dev_in, 0, a->stream);
        e = hipMemcpyAsync(host_out, stage, nn * nc * sizeof(double), hipMemcpyDeviceToHost, a->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(a->stream);
    (void)hipFree(stage);
    budget_release(a->ctx, nn * nc * sizeof(double));
    a->allocs.pop_back();
    a->bytes -= nn * nc * sizeof(double);
    if (e != hipSuccess) return afail(ELA_ERR_HIP, std::string("transfer: ") + hipGetErrorString(e));
    return ELA_OK;
}

double* field_of(ela_amr* a, int which) {
    switch (which) {
        case ELA_AMR_U: return a->u;
        case ELA_AMR_B: return a->b;
        case ELA_AMR_R: return a->r;
        default: return nullptr;
    }
}

// ---------------------------------------------------------------- level geometry (host)
// The boxes of one level and its array box G (the union's bounding box grown
// by 2 nodes, clipped to the domain).
struct LevelBoxes {
    int dim = 3;
    int N[3] = {0, 0, 0};      // domain cells of the level
    std::vector<int> boxes;    // lo[3], hi[3] per box
    int glo[3] = {0, 0, 0};    // first node of G
    int gn[3] = {1, 1, 1};     // nodes of G
    int gc(int d) const { return d < dim ? gn[d] - 1 : 1; }  // cells of G
};

int level_extent(LevelBoxes& B, int l) {
    for (int d = 0; d < 3; ++d) {
        if (d >= B.dim) {
            B.glo[d] = 0;
            B.gn[d] = 1;
            continue;
        }
        int lo = B.N[d], hi = 0;
        for (size_t q = 0; q < B.boxes.size(); q += 6) {
            lo = std::min(lo, B.boxes[q + d]);
            hi = std::max(hi, B.boxes[q + 3 + d]);
        }
        const int glo = l == 0 ? 0 : std::max(lo - 2, 0), ghi = l == 0 ? B.N[d] : std::min(hi + 2, B.N[d]);
        B.glo[d] = glo;
        B.gn[d] = ghi - glo + 1;
    }
    return ELA_OK;
}

// Whether a refined level's union box [lo, hi] coarsens (by 2 while every side
// is even and its half >= 2 cells) to a level whose dense inverse the
// hierarchy builds (at most 3 * 9^3 unknowns).  Otherwise the level smooths.
bool dense_coarsest(const int lo[3], const int hi[3], int dim) {
    int n[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) n[d] = hi[d] - lo[d];
    for (;;) {
        bool can = true;
        for (int d = 0; d < dim; ++d) can = can && n[d] % 2 == 0 && n[d] / 2 >= 2;
        if (!can) break;
        for (int d = 0; d < dim; ++d) n[d] /= 2;
    }
    long long unk = dim;
    for (int d = 0; d < dim; ++d) unk *= n[d] + 1;
    return unk <= 3 * 9 * 9 * 9;
}

// Cells of G covered by the level's boxes (cell c spans nodes c..c+1), dense, x fastest.
void union_cells(const LevelBoxes& B, std::vector<unsigned char>& cells) {
    const long long cx = B.gc(0), cy = B.gc(1), cz = B.gc(2);
    cells.assign((size_t)(cx * cy * cz), 0);
    for (size_t q = 0; q < B.boxes.size(); q += 6) {
        int lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = d < B.dim ? B.boxes[q + d] - B.glo[d] : 0;
            hi[d] = d < B.dim ? B.boxes[q + 3 + d] - B.glo[d] : 1;
        }
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) cells[(size_t)(i + cx * (j + cy * k))] = 1;
    }
}

// Global cell (x, y, z) of the level in `cells` (false outside G).
bool cell_in(const LevelBoxes& B, const std::vector<unsigned char>& cells, const int c[3]) {
    long long idx = 0, stride = 1;
    for (int d = 0; d < 3; ++d) {
        const int loc = d < B.dim ? c[d] - B.glo[d] : 0;
        if (loc < 0 || loc >= B.gc(d)) return false;
        idx += loc * stride;
        stride *= B.gc(d);
    }
    return cells[(size_t)idx] != 0;
}

// Proper nesting: every fine box, coarsened and grown by 2 coarse cells (clipped
// to the domain), lies in the coarse level's union.
bool nested(const LevelBoxes& C, const std::vector<unsigned char>& ccells, const LevelBoxes& F) {
    for (size_t q = 0; q < F.boxes.size(); q += 6) {
        int lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
        for (int d = 0; d < F.dim; ++d) {
            lo[d] = std::max(F.boxes[q + d] / 2 - 2, 0);
            hi[d] = std::min(F.boxes[q + 3 + d] / 2 + 2, C.N[d]);
        }
        for (int k = lo[2]; k < hi[2]; ++k)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int i = lo[0]; i < hi[0]; ++i) {
                    const int c[3] = {i, j, F.dim == 3 ? k : 0};
                    if (!cell_in(C, ccells, c)) return false;
                }
    }
    return true;
}

// Class bits of every node of G (see the top of the file).  A node's touching
// cells are the 2^dim cells it is a corner of; cells outside the domain do not
// count.  F / fcells: the next finer level (null on the finest).  ulo / uhi
// receive the bounding box of the nodes under F (global node indices).
void node_classes(const LevelBoxes& B, const std::vector<unsigned char>& cells, const LevelBoxes* F,
                  const std::vector<unsigned char>* fcells, bool refined, std::vector<unsigned char>& mask,
                  int ulo[3], int uhi[3]) {
    const int dim = B.dim;
    const long long nx = B.gn[0], ny = B.gn[1], nz = B.gn[2];
    mask.assign((size_t)(nx * ny * nz), 0);
    for (int d = 0; d < 3; ++d) {
        ulo[d] = d < dim ? B.N[d] : 0;
        uhi[d] = d < dim ? -1 : 0;
    }
    const int nk = dim == 3 ? 2 : 1;
    for (long long k = 0; k < nz; ++k)
        for (long long j = 0; j < ny; ++j)
            for (long long i = 0; i < nx; ++i) {
                const int p[3] = {(int)(B.glo[0] + i), (int)(B.glo[1] + j), (int)(B.glo[2] + k)};
                bool any_in = false, all_in = true, any_u = false, all_u = true;
                for (int dk = 0; dk < nk; ++dk)
                    for (int dj = 0; dj < 2; ++dj)
                        for (int di = 0; di < 2; ++di) {
                            const int c[3] = {p[0] - 1 + di, p[1] - 1 + dj, dim == 3 ? p[2] - 1 + dk : 0};
                            bool indom = true;
                            for (int d = 0; d < dim; ++d) indom = indom && c[d] >= 0 && c[d] < B.N[d];
                            if (!indom) continue;
                            const bool in = cell_in(B, cells, c);
                            any_in = any_in || in;
                            all_in = all_in && in;
                            if (F) {
                                const int fc[3] = {2 * c[0], 2 * c[1], dim == 3 ? 2 * c[2] : 0};
                                const bool u = cell_in(*F, *fcells, fc);
                                any_u = any_u || u;
                                all_u = all_u && u;
                            }
                        }
                unsigned char m = 0;
                if (!any_in || (refined && !all_in)) m |= CLS_FILL;
                if (F && any_u) {
                    m |= CLS_UNDER;
                    if (all_u && any_in) m |= CLS_COVERED;
                    else m |= CLS_IFACE;
                    for (int d = 0; d < dim; ++d) {
                        ulo[d] = std::min(ulo[d], p[d]);
                        uhi[d] = std::max(uhi[d], p[d]);
                    }
                }
                mask[(size_t)(i + nx * (j + ny * k))] = m;
            }
    if (!F)
        for (int d = 0; d < 3; ++d) ulo[d] = uhi[d] = 0;
}

}  // namespace

extern "C" {

int ela_amr_create(ela_ctx* ctx, const ela_domain* dom, int nlev, const int* boxes, int model_kind, int coarse_min,
                   int mode, ela_amr** out) {
    std::vector<int> counts(nlev > 1 ? nlev - 1 : 0, 1);
    return ela_amr_create_boxes(ctx, dom, nlev, counts.data(), boxes, model_kind, coarse_min, mode, out);
}

int ela_amr_create_boxes(ela_ctx* ctx, const ela_domain* dom, int nlev, const int* nboxes, const int* boxes,
                         int model_kind, int coarse_min, int mode, ela_amr** out) {
    if (!ctx || !dom || !out || (nlev > 1 && (!boxes || !nboxes))) return afail(ELA_ERR_BADARG, "null argument");
    *out = nullptr;
    if (dom->dim != 2 && dom->dim != 3) return afail(ELA_ERR_BADARG, "dim must be 2 or 3");
    if (nlev < 1 || nlev > 8) return afail(ELA_ERR_BADARG, "1 to 8 levels");
    if (mode != ELA_SOLVE_FP64 && mode != ELA_SOLVE_MIXED) return afail(ELA_ERR_BADARG, "mode must be FP64 or MIXED");
    if (model_kind < 0 || model_kind > 2) return afail(ELA_ERR_BADARG, "unknown model kind");
    const int dim = dom->dim;
    for (int d = 0; d < dim; ++d)
        if (dom->n[d] < 2 || !(dom->dx[d] > 0)) return afail(ELA_ERR_BADARG, "bad level-0 domain");
    // boxes per level, their bounding boxes and the arrays G_l
    std::vector<std::vector<int>> bx(nlev);
    std::vector<LevelBoxes> lb(nlev);
    size_t pos = 0;
    for (int l = 0; l < nlev; ++l) {
        LevelBoxes& B = lb[l];
        B.dim = dim;
        for (int d = 0; d < 3; ++d) B.N[d] = d < dim ? dom->n[d] << l : 0;
        if (l == 0) {
            bx[0] = {0, 0, 0, B.N[0], B.N[1], B.N[2]};
        } else {
            const int nb = nboxes[l - 1];
            if (nb < 1 || nb > 1 << 20) return afail(ELA_ERR_BADARG, "level " + std::to_string(l) + ": no boxes");
            bx[l].assign(boxes + 6 * pos, boxes + 6 * (pos + nb));
            pos += nb;
            for (int q = 0; q < nb; ++q)
                for (int d = 0; d < 3; ++d) {
                    int& lo = bx[l][6 * q + d];
                    int& hi = bx[l][6 * q + 3 + d];
                    if (d >= dim) {
                        lo = hi = 0;
                        continue;
                    }
                    if (lo % 2 || hi % 2 || lo < 0 || hi <= lo || hi > B.N[d])
                        return afail(ELA_ERR_BADARG, "level " + std::to_string(l) +
                                                         ": box corners must be even and inside the domain");
                }
        }
        B.boxes = bx[l];
        if (int rc = level_extent(B, l)) return rc;
    }
    std::vector<std::vector<unsigned char>> cells(nlev), masks(nlev);
    for (int l = 0; l < nlev; ++l) union_cells(lb[l], cells[l]);
    for (int l = 1; l < nlev; ++l)
        if (!nested(lb[l - 1], cells[l - 1], lb[l]))
            return afail(ELA_ERR_BADARG, "level " + std::to_string(l) + ": every box must sit 2 coarse cells inside "
                                         "its parent's union (or reach the domain boundary with it)");
    std::vector<Geo> geo(nlev);
    for (int l = 0; l < nlev; ++l) {
        Geo& g = geo[l];
        std::memset(&g, 0, sizeof g);
        g.dim = dim;
        for (int d = 0; d < 3; ++d) {
            g.glo[d] = lb[l].glo[d];
            g.gn[d] = lb[l].gn[d];
            g.N[d] = lb[l].N[d];
        }
        node_classes(lb[l], cells[l], l + 1 < nlev ? &lb[l + 1] : nullptr, l + 1 < nlev ? &cells[l + 1] : nullptr,
                     l > 0, masks[l], g.ulo, g.uhi);
        g.has_fine = l + 1 < nlev;
    }
    ela_amr* a = new ela_amr();
    a->ctx = ctx;
    a->dim = dim;
    a->kind = model_kind;
    a->nc = model_kind == ELA_MODEL_ISO ? 2 : (model_kind == ELA_MODEL_MM ? (dim == 3 ? 21 : 6) : dim * dim);
    a->mode = mode;
    a->cprec = mode == ELA_SOLVE_MIXED ? ELA_F32 : ELA_F64;
    a->nlev = nlev;
    a->lev.resize(nlev);
    int rc = ELA_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) rc = afail(ELA_ERR_HIP, "hipSetDevice failed");
    size_t off = 0;
    for (int l = 0; l < nlev && !rc; ++l) {
        CLev& v = a->lev[l];
        v.geo = geo[l];
        v.boxes = bx[l];
        for (int d = 0; d < 3; ++d) v.h[d] = d < dim ? dom->dx[d] / (double)(1 << l) : 1.0;
        v.L.dim = dim;
        v.L.nx = v.geo.gn[0];
        v.L.ny = v.geo.gn[1];
        v.L.nz = dim == 3 ? v.geo.gn[2] : 1;
        v.L.px = (v.L.nx + 31) / 32 * 32;
        v.L.pl = v.L.px * v.L.ny;
        v.L.cs = (v.L.pl * v.L.nz + 255) / 256 * 256;
        v.elems = (size_t)dim * v.L.cs;
        v.off = off;
        off += v.elems;
        // the union's bounding box; a union that fills it is one box and gets a multigrid hierarchy
        size_t bbox_cells = 1, ucells = 0;
        for (int d = 0; d < 3; ++d) {
            v.ulo[d] = d < dim ? v.geo.N[d] : 0;
            v.uhi[d] = 0;
            for (size_t q = 0; q < bx[l].size(); q += 6) {
                v.ulo[d] = std::min(v.ulo[d], bx[l][q + d]);
                v.uhi[d] = std::max(v.uhi[d], bx[l][q + 3 + d]);
            }
            if (d < dim) bbox_cells *= (size_t)(v.uhi[d] - v.ulo[d]);
        }
        for (unsigned char c : cells[l]) ucells += c;
        if (l == 0 || (ucells == bbox_cells && dense_coarsest(v.ulo, v.uhi, dim))) {
            // its own stream until the composite's is shared
            ela_domain bd{};
            bd.dim = dim;
            for (int d = 0; d < 3; ++d) {
                bd.n[d] = d < dim ? v.uhi[d] - v.ulo[d] : 0;
                bd.dx[d] = v.h[d];
            }
            rc = ela_hier_create_ex(ctx, &bd, model_kind, 0, l == 0 ? std::max(coarse_min, 2) : 2, mode, &v.H);
            if (!rc) rc = hier_level0(v.H, &v.Lb);
            if (!rc && l == 0) a->stream = hier_stream(v.H);
            if (!rc && l > 0) rc = hier_share_stream(v.H, a->stream);
            // a refined box (Displacement rows on its coarse/fine faces, a well-conditioned local problem)
            // takes a V(1,1) cycle: the 3-level crack nest on a 128^3 base converges in 31 instead of 30
            // iterations at 0.65x the time per solve
            if (!rc && l > 0) rc = ela_set_cycle(v.H, ELA_CYCLE_V, 1, 1);
        }
    }
    a->ntot = off;
    for (int l = 0; l < nlev && !rc; ++l) {
        CLev& v = a->lev[l];
        rc = aalloc_t(a, (size_t)a->nc * v.L.cs, &v.C);
        if (!rc) rc = aalloc_t(a, 1, &v.dbcG);
        if (!rc) rc = aalloc_t(a, masks[l].size(), &v.cls);
        if (!rc && hipMemcpyAsync(v.cls, masks[l].data(), masks[l].size(), hipMemcpyHostToDevice, a->stream) != hipSuccess)
            rc = afail(ELA_ERR_HIP, "mask upload failed");
        v.geo.cls = v.cls;
        if (!rc && !v.H) rc = aalloc_t(a, (size_t)dim * v.L.cs, &v.dg);
        if (!rc && !v.H && a->cprec == ELA_F32) {
            rc = aalloc_t(a, (size_t)a->nc * v.L.cs, &v.C32);
            for (float** p : {&v.dg32, &v.x32, &v.t32, &v.r32})
                if (!rc) rc = aalloc_t(a, (size_t)dim * v.L.cs, p);
        }
    }
    for (double** p : {&a->u, &a->b, &a->r, &a->z, &a->w, &a->t, &a->q})
        if (!rc) rc = aalloc_t(a, a->ntot, p);
    if (!rc) rc = aalloc_t(a, 4, &a->dnorm);
    if (!rc && hipHostMalloc((void**)&a->hnorm, 64) != hipSuccess) rc = afail(ELA_ERR_OOM, "hipHostMalloc failed");
    if (!rc && a->stream && hipStreamSynchronize(a->stream) != hipSuccess) rc = afail(ELA_ERR_HIP, "zeroing failed");
    if (rc) {
        std::string keep = ela_last_error();
        ela_amr_destroy(a);
        set_error(rc, keep);
        return rc;
    }
    // default BCs: Displacement everywhere (BC/Operator/Elastic.H:46-55)
    int types[26 * 3] = {};
    ela_amr_set_bc(a, types, nullptr);
    *out = a;
    return ELA_OK;
}

int ela_amr_destroy(ela_amr* a) {
    if (!a) return ELA_OK;
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    afree_all(a);
    if (a->hnorm) (void)hipHostFree(a->hnorm);
    for (int l = (int)a->lev.size() - 1; l >= 0; --l)  // level 0 owns the shared stream
        if (a->lev[l].H) ela_hier_destroy(a->lev[l].H);
    delete a;
    return ELA_OK;
}

size_t ela_amr_bytes(const ela_amr* a) {
    if (!a) return 0;
    size_t s = a->bytes;
    for (auto& v : a->lev) s += v.H ? ela_hier_bytes(v.H) : 0;
    return s;
}

int ela_amr_level_info(const ela_amr* a, int lev, ela_level_layout* out, int origin[3]) {
    if (int rc = check_lev(a, lev)) return rc;
    if (!out) return afail(ELA_ERR_BADARG, "null output");
    const CLev& v = a->lev[lev];
    out->dim = a->dim;
    out->nodes[0] = v.L.nx;
    out->nodes[1] = v.L.ny;
    out->nodes[2] = v.L.nz;
    for (int d = 0; d < 3; ++d) {
        out->n[d] = d < a->dim ? out->nodes[d] - 1 : 0;
        out->dx[d] = v.h[d];
        if (origin) origin[d] = v.geo.glo[d];
    }
    out->pitch = v.L.px;
    out->plane = v.L.pl;
    out->cstride = v.L.cs;
    return ELA_OK;
}

int ela_amr_set_model_field(ela_amr* a, int lev, const double* C_host) {
    if (int rc = check_lev(a, lev)) return rc;
    if (!C_host) return afail(ELA_ERR_BADARG, "null model");
    CLev& v = a->lev[lev];
    if (int rc = xfer(a, v, a->nc, C_host, v.C, nullptr, nullptr)) return rc;
    v.model_set = true;
    a->prepared = false;
    return ELA_OK;
}

int ela_amr_set_bc(ela_amr* a, const int* types, const double* values) {
    if (!a || !types) return afail(ELA_ERR_BADARG, "null argument");
    const int nf = a->dim == 3 ? 26 : 8;
    for (int f = 0; f < nf; ++f)
        for (int c = 0; c < a->dim; ++c)
            if (types[3 * f + c] < 0 || types[3 * f + c] > 3) return afail(ELA_ERR_BADARG, "unknown boundary type");
    std::memset(&a->bc, 0, sizeof a->bc);
    for (int f = 0; f < nf; ++f)
        for (int c = 0; c < 3; ++c) {
            a->bctypes[f][c] = c < a->dim ? types[3 * f + c] : 0;
            a->bcval[f][c] = (values && c < a->dim) ? values[3 * f + c] : 0.0;
            a->bc.t[f][c] = (unsigned char)a->bctypes[f][c];
        }
    // per level: the table of G (artificial faces -> Displacement) and of the box
    // hierarchy, if any (coarse/fine faces of the union box -> homogeneous Displacement)
    for (int l = 0; l < a->nlev; ++l) {
        CLev& v = a->lev[l];
        std::vector<int> bt(26 * 3, 0);
        v.bcG = a->bc;
        for (int f = 0; f < nf; ++f) {
            int side[3];
            class_sides(a->dim, f, side);
            bool art = false, cf = false;
            for (int d = 0; d < a->dim; ++d) {
                if (side[d] < 0) continue;
                const int gend = side[d] ? v.geo.glo[d] + v.geo.gn[d] - 1 : v.geo.glo[d];
                art = art || (side[d] ? gend < v.geo.N[d] : gend > 0);
                cf = cf || (side[d] ? v.uhi[d] < v.geo.N[d] : v.ulo[d] > 0);
            }
            for (int c = 0; c < 3; ++c) {
                if (art) v.bcG.t[f][c] = ELA_BC_DISP;
                bt[3 * f + c] = cf ? ELA_BC_DISP : a->bctypes[f][c];
            }
        }
        if (v.dbcG && hipMemcpyAsync(v.dbcG, &v.bcG, sizeof v.bcG, hipMemcpyHostToDevice, a->stream) != hipSuccess)
            return afail(ELA_ERR_HIP, "hipMemcpyAsync failed");
        if (v.H)
            if (int rc = ela_set_bc(v.H, bt.data(), nullptr)) return rc;
    }
    a->prepared = false;
    return ELA_OK;
}

int ela_amr_bc_init(ela_amr* a, int lev, double* rhs, int homogeneous) {
    if (int rc = check_lev(a, lev)) return rc;
    if (!rhs) return afail(ELA_ERR_BADARG, "null rhs");
    const CLev& v = a->lev[lev];
    const Geo& g = v.geo;
    const long long nn = (long long)v.L.nx * v.L.ny * v.L.nz;
    for (int k = 0; k < v.L.nz; ++k)
        for (int j = 0; j < v.L.ny; ++j)
            for (int i = 0; i < v.L.nx; ++i) {
                const int px = g.glo[0] + i, py = g.glo[1] + j, pz = g.glo[2] + k;
                const int face = a->dim == 3 ? face_of<3>(px, py, pz, g.N[0], g.N[1], g.N[2])
                                             : face_of<2>(px, py, 0, g.N[0], g.N[1], 0);
                if (face < 0) continue;
                const long long id = i + (long long)v.L.nx * (j + (long long)v.L.ny * k);
                for (int d = 0; d < a->dim; ++d)
                    rhs[d * nn + id] = (homogeneous && a->bctypes[face][d] == ELA_BC_DISP) ? 0.0 : a->bcval[face][d];
            }
    return ELA_OK;
}

int ela_amr_set_uniform(ela_amr* a, int uniform) {
    if (!a) return afail(ELA_ERR_BADARG, "null composite");
    a->uniform = uniform ? 1 : 0;
    for (auto& v : a->lev)
        if (v.H) ela_set_uniform(v.H, a->uniform);
    a->prepared = false;
    return ELA_OK;
}

int ela_amr_set_krylov(ela_amr* a, int restart) {
    if (!a) return afail(ELA_ERR_BADARG, "null composite");
    if (restart < 0 || restart > kMaxKrylov - 1) return afail(ELA_ERR_BADARG, "restart out of range");
    a->krylov = restart;
    return ELA_OK;
}

int ela_amr_set_cycle(ela_amr* a, int cycle, int nu1, int nu2) {
    if (!a) return afail(ELA_ERR_BADARG, "null composite");
    for (int l = 0; l < a->nlev; ++l)
        if (int rc = ela_amr_set_level_cycle(a, l, cycle, nu1, nu2)) return rc;
    return ELA_OK;
}

int ela_amr_set_level_cycle(ela_amr* a, int lev, int cycle, int nu1, int nu2) {
    if (int rc = check_lev(a, lev)) return rc;
    if (a->lev[lev].H) return ela_set_cycle(a->lev[lev].H, cycle, nu1, nu2);
    if (nu1 < 0 || nu2 < 0 || nu1 + nu2 < 1 || nu1 > 64 || nu2 > 64)
        return afail(ELA_ERR_BADARG, "smoothing calls must be in [0, 64], not both zero");
    a->lev[lev].nu1 = nu1;
    a->lev[lev].nu2 = nu2;
    return ELA_OK;
}

int ela_amr_prepare(ela_amr* a) {
    if (!a) return afail(ELA_ERR_BADARG, "null composite");
    for (int l = 0; l < a->nlev; ++l)
        if (!a->lev[l].model_set)
            return afail(ELA_ERR_STATE, "Attempting to use operator before calling SetModel (level " + std::to_string(l) + ")");
    for (auto& v : a->lev) {
        if (v.H) {  // the union box's stiffness into its hierarchy
            hipLaunchKernelGGL(k_subbox, grid_of(v.Lb), dim3(CX, CY), 0, a->stream, v.L, v.ulo[0] - v.geo.glo[0],
                               v.ulo[1] - v.geo.glo[1], v.ulo[2] - v.geo.glo[2], (const double*)v.C, v.Lb,
                               hier_C0(v.H), a->nc);
            hier_model_set(v.H);
            if (int rc = ela_prepare(v.H)) return rc;
        } else {  // the diagonal of A_l on G_l for the Jacobi sweeps (its values at fill nodes are never used)
            launch_diag(v.L, a->kind, v.C, v.dg, v.dg32, v.bcG, v.h, a->dnorm, a->stream);
            if (v.C32) launch_convert<double, float>((long long)a->nc * v.L.cs, v.C, v.C32, a->stream);
        }
    }
    ELA_HIPCHK(hipStreamSynchronize(a->stream));
    a->prepared = true;
    return ELA_OK;
}

int ela_amr_upload(ela_amr* a, int lev, int field, const double* host) {
    if (int rc = check_lev(a, lev)) return rc;
    double* f = field_of(a, field);
    if (!f || !host) return afail(ELA_ERR_BADARG, "bad field or null host array");
    return xfer(a, a->lev[lev], a->dim, host, seg(f, a->lev[lev]), nullptr, nullptr);
}

int ela_amr_download(ela_amr* a, int lev, int field, double* host) {
    if (int rc = check_lev(a, lev)) return rc;
    double* f = field_of(a, field);
    if (!f || !host) return afail(ELA_ERR_BADARG, "bad field or null host array");
    return xfer(a, a->lev[lev], a->dim, nullptr, nullptr, host, seg(f, a->lev[lev]));
}

int ela_amr_residual(ela_amr* a, double* norm_inf) {
    if (!a) return afail(ELA_ERR_BADARG, "null composite");
    if (!a->prepared) return afail(ELA_ERR_STATE, "call ela_amr_prepare first");
    residual(a, a->u, a->b, a->r);
    const double v = read_norm(a);
    if (norm_inf) *norm_inf = v;
    return ELA_OK;
}

int ela_amr_solve(ela_amr* a, double tol_rel, double tol_abs, int max_iter, ela_stats* st) {
    if (!a) return afail(ELA_ERR_BADARG, "null composite");
    if (!a->prepared) return afail(ELA_ERR_STATE, "call ela_amr_prepare first");
    if (max_iter < 1) return afail(ELA_ERR_BADARG, "max_iter must be >= 1");
    if (hipSetDevice(a->ctx->device) != hipSuccess) return afail(ELA_ERR_HIP, "hipSetDevice failed");
    return solve(a, tol_rel, tol_abs, max_iter, st);
}

}  // extern "C"
