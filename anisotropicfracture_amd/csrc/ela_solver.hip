// ela_solver.hip -- host side of the gfx950 nodal elasticity solve: device
// memory of the multigrid hierarchy, set-up (stiffness coarsening, diagonals,
// dense coarsest operator), the F/V-cycle driver with FP64 or FP32 inner
// cycles inside an FP64 outer iteration, and the C ABI of include/ela.h.
//
// Cycle structure (restating AMReX MLMG as the reference drives it from
// Solver::Nonlocal::Linear, src/Solver/Nonlocal/Linear.H:23-29,102-112):
//   outer: r = b - A u (FP64); cor = cycle(r); u += cor; until
//          |r|inf <= max(tol_abs, tol_rel |r0|inf).
//   F-cycle: restrict r to every level, solve the coarsest, then per level
//          upward: cor = P cor_c, r = r - A cor, V-cycle, cor += saved.
//   V-cycle: per level nu1 calls of Fsmooth (2 Jacobi sweeps each,
//          omega = 2/3, Operator.cpp:90-162), rescor = r - A cor, restrict;
//          the coarsest level exactly; then prolong-add and nu2 Fsmooth calls.
// The coarsest level is solved with a dense inverse formed on the device
// (exact, where the reference's BiCGStab stops at a relative 1e-4); the
// converged solution of A u = b is the same.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ela.h"
#include "ela_internal.hpp"
#include "ela_kernels.hpp"
#include "ela_hooks.hpp"

using namespace ela;

// ---------------------------------------------------------------- allocation-free residual
// Scratch of ela_residual_noalloc: kScratchPartials per-block maxima followed
// by the norm word, one copy per GPU, held until the process ends.  Host code
// never passes the array's name to a kernel or a copy: it is a host-side
// stand-in there.  g_scratch[device] is the device address from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr long long kScratchPartials = 1LL << 17;  // 1 MiB of doubles
static __device__ __align__(256) double scratch[kScratchPartials + 32];
namespace {
constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return ela::set_error(ELA_ERR_BADARG, "device index too large");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch[device]) return ELA_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return ela::set_error(ELA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = (double*)p;
    return ELA_OK;
}
}  // namespace

namespace {
thread_local std::string g_err;
}  // namespace

namespace ela {

int set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int budget_reserve(ela_ctx* ctx, size_t bytes) {
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->budget && ctx->in_use + bytes > ctx->budget)
        return set_error(ELA_ERR_OOM, "HBM budget exceeded: need " + std::to_string(bytes) + " more bytes, " +
                                          std::to_string(ctx->budget - ctx->in_use) + " left");
    ctx->in_use += bytes;
    return ELA_OK;
}

void budget_release(ela_ctx* ctx, size_t bytes) {
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->in_use -= bytes;
}

}  // namespace ela

namespace {

int fail(int code, const std::string& msg) { return set_error(code, msg); }

#define HIPCHK(expr) ELA_HIPCHK(expr)

constexpr int kMaxDenseUnknowns = 3 * 9 * 9 * 9;
constexpr int kMaxEvents = 8192;

struct Level {
    Lay L{};
    int n[3] = {0, 0, 0};
    double h[3] = {1, 1, 1};
    int nodes = 0;
    double* C64 = nullptr;
    float* C32 = nullptr;
    double* d64 = nullptr;
    float* d32 = nullptr;  // inverse diagonal of the FP32 smoothers
    unsigned char* sm = nullptr;  // same-direction mask of the stiffness planes (launch_same_mask)
    unsigned short* gid = nullptr;  // coarse levels of a grain-form hierarchy: grain ids by injection
    int kz = 0;                     // coarse marching level: tile height picked by ela_autotune (0: the hierarchy's)
    bool sm_ok = false;           // sm was built by the last ela_prepare
    bool use_sm = false;          // enough wavefronts skip a derivative of C for the mask to pay
    unsigned long long* nl = nullptr;  // node list (launch_node_list): constant-C interior nodes first
    int* nlcnt = nullptr;              // its scratch
    long long* nltot = nullptr;        // its part sizes (device)
    bool use_list = false;             // enough constant-C nodes for the list order to pay
    bool use_group = false;            // enough for regrouping them inside the blocks to pay
    double *cor64 = nullptr, *tmp64 = nullptr, *res64 = nullptr, *rc64 = nullptr, *hold64 = nullptr;
    float *cor32 = nullptr, *tmp32 = nullptr, *res32 = nullptr, *rc32 = nullptr, *hold32 = nullptr;
};

template <typename T>
struct LB {
    T *C, *dg, **cor, **tmp, **res, **rc, **hold;
};

LB<double> lb64(Level& v) { return {v.C64, v.d64, &v.cor64, &v.tmp64, &v.res64, &v.rc64, &v.hold64}; }
LB<float> lb32(Level& v) { return {v.C32, v.d32, &v.cor32, &v.tmp32, &v.res32, &v.rc32, &v.hold32}; }
template <typename T>
LB<T> lb(Level& v);
template <>
LB<double> lb<double>(Level& v) { return lb64(v); }
template <>
LB<float> lb<float>(Level& v) { return lb32(v); }

}  // namespace

struct ela_hier {
    ela_ctx* ctx = nullptr;
    int dim = 3, kind = 0, nc = 2, mode = 0;
    bool has64 = false, has32 = false;
    std::vector<Level> lev;
    BCTab bc{};
    double bcval[26][3] = {};
    int uniform = 0;
    bool model_set = false, prepared = false;
    hipStream_t stream = nullptr;
    bool own_stream = true;  // false once a composite solve shares its stream with the hierarchy
    std::vector<std::pair<void*, size_t>> allocs;
    size_t bytes = 0;
    // level-0 solve fields (FP64)
    double *u64 = nullptr, *b64 = nullptr;
    // coarsest level dense inverse [A | A^-1], n x 2n
    double* M = nullptr;
    int nunk = 0;
    // device scalars
    unsigned long long* dnorm = nullptr;
    double* partials = nullptr;  // per-block maxima of the residual kernels
    double* dscale = nullptr;
    int* dflag = nullptr;
    unsigned long long* mstats = nullptr;  // per level: wavefronts, skipping (wavefront, direction) pairs
    int mask_mode = 1;                     // ELA_OPT_MASK: 0 never, 1 where it pays, 2 / 3 every level
    // ELA_OPT_GRAPH: the launches of levels >= 1 replayed from captured HIP
    // graphs, one per launch configuration (the key holds everything those
    // launches read on the host)
    bool graph_on = false;
    bool fold = true;  // ELA_OPT_FOLD: the F-cycle's last level-0 sweep writes the Krylov vector itself
    long long epoch = 0;  // bumped by every ela_prepare
    struct CycleGraph {
        std::vector<unsigned long long> key;
        hipGraphExec_t exec = nullptr;
        long long used = 0;
    };
    std::vector<CycleGraph> graphs;
    long long graph_clock = 0;
    unsigned long long* hnorm = nullptr;  // pinned
    // BiCGStab bottom solve of a coarsest level too large for the dense inverse:
    // r, r-hat, p, v, s, t on the coarsest layout per precision, reduction scratch
    double* bv64[8] = {};  // coarsest BiCGStab: r, rh, p, v, s, t, D^-1 p, D^-1 s
    float* bv32[8] = {};
    double *bpart = nullptr, *bout = nullptr, *bhost = nullptr;  // bhost: pinned
    int bottom_iters = 0;                                           // BiCGStab steps of the last solve
    // grain-table stiffness of level 0 (Voronoi models): per-node grain id, per-grain records
    unsigned short* gid = nullptr;
    double *tab64 = nullptr, *dgt64 = nullptr;
    float *tab32 = nullptr, *dgt32 = nullptr;  // dgt32: inverse of the per-grain interior diagonal
    // (dgt64, dgt32: one table of 3 values a grain for each MG level, at 3 ngrains l for level l)
    int ngrains = 0;
    bool has_table = false, use_table = false;
    // z-marching kernels for the grain-form level 0 (planes per tile), and the
    // BC table on the device for them
    bool grain_march = true;
    int kz = 16;
    bool xcd = true;  // XCD-contiguous tile order of the marching kernels
    int ry = 1;       // rows per thread of level 0's marching tiles (FP32 sweeps)
    int pf = 1;       // planes loaded ahead by the marching tiles (ELA_OPT_PREFETCH)
    bool gj_one = true;  // Gauss-Jordan of the dense coarsest operator in one launch a column (ELA_OPT_GJ)
    bool tile_dgt = true;  // marching tiles take the interior diagonal per grain (ELA_OPT_GRAIN_DIAG)
    BCTab* dbc = nullptr;
    // dw = DDW : (grad u - F0) of Newton's right-hand side (dim*dim FP64 planes, allocated on first use)
    double* newton_dw = nullptr;
    // flexible GMRES around the MG cycle (restart length; 0 = plain MG iteration)
    int krylov = 0;
    std::vector<double*> V, Z;
    std::vector<float*> V32;  // FP32 Krylov basis of the mixed solve
    double* kw = nullptr;     // its FP64 work vector w = A z
    double *kr = nullptr, *kdev = nullptr, *kpart = nullptr;
    // cycle options
    int cycle = ELA_CYCLE_F, nu1 = 2, nu2 = 2;
    int pitch_align = 32;  // elements; rows start on 128 B (FP32) / 256 B (FP64) boundaries
    double omega = 2.0 / 3.0;
    // kernel timing of the last solve
    bool timing = false;
    std::vector<hipEvent_t> ev;
    int ev_used = 0;
    std::vector<std::pair<int, int>> ev_resid, ev_smooth;  // event index pairs
    double ms_resid = 0, ms_smooth = 0;
    int n_resid = 0, n_smooth = 0;
};

namespace {

int halloc(ela_hier* h, size_t bytes, void** out) {
    *out = nullptr;
    if (bytes == 0) return ELA_OK;
    if (int rc = budget_reserve(h->ctx, bytes)) return rc;
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        budget_release(h->ctx, bytes);
        *out = nullptr;
        return fail(ELA_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    e = hipMemsetAsync(*out, 0, bytes, h->stream);
    if (e != hipSuccess) return fail(ELA_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    h->allocs.push_back({*out, bytes});
    h->bytes += bytes;
    return ELA_OK;
}

template <typename T>
int halloc_t(ela_hier* h, size_t count, T** out) {
    void* p = nullptr;
    int rc = halloc(h, count * sizeof(T), &p);
    *out = (T*)p;
    return rc;
}

void hfree(ela_hier* h, void* p) {
    for (size_t q = 0; q < h->allocs.size(); ++q)
        if (h->allocs[q].first == p) {
            (void)hipFree(p);
            budget_release(h->ctx, h->allocs[q].second);
            h->bytes -= h->allocs[q].second;
            h->allocs.erase(h->allocs.begin() + q);
            return;
        }
}

int ncomp_of(int dim, int kind) { return kind == ELA_MODEL_ISO ? 2 : (kind == ELA_MODEL_MM ? (dim == 3 ? 21 : 6) : dim * dim); }

size_t field_elems(const Level& v, int nc) { return (size_t)v.L.cs * nc; }

// Operator application on a level: the LDS-staged marching kernels where the
// level is large enough (3D), the one-thread-per-node kernels otherwise.
template <typename T>
void A_apply(ela_hier* h, Level& v, const T* C, const T* u, T* f, hipStream_t s);
template <typename T>
void A_jacobi(ela_hier* h, Level& v, const T* C, const T* dg, const T* xin, const T* b, T* xout, hipStream_t s);
template <typename T>
void grain_jacobi_to64(ela_hier* h, Level& v, const T* dg, const T* xin, const T* b, const T* hold, double* out64,
                       hipStream_t s);
bool grain_on(ela_hier* h, const Level& v);
bool grain_levels(const ela_hier* h);
const unsigned short* level_gid(const ela_hier* h, const Level& v);
template <typename T>
const T* level_dgt(const ela_hier* h, const Level& v);
// part: scratch for the per-block maxima of the norm (null: the hierarchy's own).
template <typename T, typename TO>
void A_residual(ela_hier* h, Level& v, const T* C, const T* x, const T* b, T* r, TO* rs, const double* scale,
                unsigned long long* norm, hipStream_t s, double* part = nullptr);

// Kernel timing around level-0 launches.
int ev_begin(ela_hier* h) {
    if (!h->timing || h->ev_used + 2 > (int)h->ev.size()) return -1;
    const int a = h->ev_used;
    h->ev_used += 2;
    (void)hipEventRecord(h->ev[a], h->stream);
    return a;
}
void ev_end(ela_hier* h, int a, std::vector<std::pair<int, int>>& into) {
    if (a < 0) return;
    (void)hipEventRecord(h->ev[a + 1], h->stream);
    into.push_back({a, a + 1});
}

// ------------------------------------------------------------------ cycle pieces
// One Fsmooth call (2 sweeps cor -> tmp -> cor).  With out64 (a grain-form
// level 0 only), the second sweep writes (T) (its value + hold) into out64 in
// FP64 instead of into cor (MODE 4: k_add_to64 folded into the sweep); that
// sweep is left out of the level-0 sweep timing, whose bytes it does not match.
template <typename T>
void smooth_call(ela_hier* h, int l, bool zero_start, double* out64 = nullptr) {
    Level& v = h->lev[l];
    LB<T> b = lb<T>(v);
    const bool t0 = (l == 0);
    if (zero_start) {
        // a grain-form level's interior diagonal per grain (the field's values there)
        const bool pg = grain_levels(h) && h->tile_dgt && h->dim == 3;
        launch_jacobi0<T>(v.L, b.dg, *b.res, *b.tmp, h->omega, h->stream, pg ? level_gid(h, v) : nullptr,
                          pg ? level_dgt<T>(h, v) : nullptr);
    }
    else {
        const int e = t0 ? ev_begin(h) : -1;
        A_jacobi<T>(h, v, b.C, b.dg, *b.cor, *b.res, *b.tmp, h->stream);
        if (t0) ev_end(h, e, h->ev_smooth);
    }
    if (out64) {
        grain_jacobi_to64<T>(h, v, b.dg, *b.tmp, *b.res, *b.hold, out64, h->stream);
        return;
    }
    const int e = t0 ? ev_begin(h) : -1;
    A_jacobi<T>(h, v, b.C, b.dg, *b.tmp, *b.res, *b.cor, h->stream);
    if (t0) ev_end(h, e, h->ev_smooth);
}

template <typename T>
void res_of_cor(ela_hier* h, int l) {
    Level& v = h->lev[l];
    LB<T> b = lb<T>(v);
    A_residual<T, T>(h, v, b.C, *b.cor, *b.res, *b.rc, (T*)nullptr, nullptr, nullptr, h->stream);
}

// BiCGStab of a coarsest level without a dense inverse, from a zero correction
// to |r|inf <= kBottomTol |res|inf or kBottomMaxIter steps: AMReX's bottom
// solver (MLCGSolver's BiCGStab, the bottom Linear selects, Linear.H:23-29,
// with MLMG's bottom tolerance 1e-4), right-preconditioned by the level's
// Jacobi diagonal (A applied to D^-1 p and D^-1 s, as the reference's
// operators scale them through normalize, Operator.cpp:164-195).  Vectors in
// the cycle precision, dot products and norms in FP64; the coefficients go
// through the host.
constexpr double kBottomTol = 1e-4;
constexpr int kBottomMaxIter = 200;
template <typename T>
T* const* bicg_vecs(ela_hier* h);
template <>
double* const* bicg_vecs<double>(ela_hier* h) { return h->bv64; }
template <>
float* const* bicg_vecs<float>(ela_hier* h) { return h->bv32; }

template <typename T>
void bicg_bottom(ela_hier* h) {
    Level& v = h->lev.back();
    LB<T> lbv = lb<T>(v);
    T* x = *lbv.cor;
    const T* b = *lbv.res;
    T* const* W = bicg_vecs<T>(h);
    T *r = W[0], *rh = W[1], *p = W[2], *vv = W[3], *sv = W[4], *t = W[5], *ph = W[6], *sh = W[7];
    const T* dg = lbv.dg;  // right preconditioner: the level's Jacobi diagonal (FP32: its inverse)
    const long long n = (long long)field_elems(v, h->dim);
    hipStream_t st = h->stream;
    auto reduce = [&](const T* a, const T* bb, const T* c, const T* d, const T* e, double q[3]) {
        launch_bicg_reduce<T>(n, a, bb, c, d, e, h->bpart, h->bout, st);
        (void)hipMemcpyAsync(h->bhost, h->bout, 3 * sizeof(double), hipMemcpyDeviceToHost, st);
        (void)hipStreamSynchronize(st);
        for (int k = 0; k < 3; ++k) q[k] = h->bhost[k];
    };
    (void)hipMemsetAsync(x, 0, n * sizeof(T), st);
    (void)hipMemsetAsync(p, 0, n * sizeof(T), st);
    (void)hipMemsetAsync(vv, 0, n * sizeof(T), st);
    (void)hipMemcpyAsync(r, b, n * sizeof(T), hipMemcpyDeviceToDevice, st);
    (void)hipMemcpyAsync(rh, b, n * sizeof(T), hipMemcpyDeviceToDevice, st);
    double q[3];
    reduce(rh, r, nullptr, nullptr, r, q);
    double rho = q[0];
    const double rnorm0 = q[2];
    if (!(rnorm0 > 0)) return;  // zero right-hand side: zero correction
    const double tol = kBottomTol * rnorm0;
    double rho1 = 1.0, alpha = 1.0, omega = 1.0;
    int it = 0;
    for (; it < kBottomMaxIter; ++it) {
        if (!(rho != 0.0) || !std::isfinite(rho)) break;
        const double beta = it == 0 ? 0.0 : (rho / rho1) * (alpha / omega);
        launch_bicg_update<T>(n, 0, 0.0, beta, omega, x, r, p, vv, sv, t, dg, ph, sh, v.L, st);
        A_apply<T>(h, v, lbv.C, ph, vv, st);
        reduce(rh, vv, nullptr, nullptr, nullptr, q);
        if (!(q[0] != 0.0) || !std::isfinite(q[0])) break;
        alpha = rho / q[0];
        launch_bicg_update<T>(n, 1, alpha, 0.0, 0.0, x, r, p, vv, sv, t, dg, ph, sh, v.L, st);
        reduce(nullptr, nullptr, nullptr, nullptr, sv, q);
        if (q[2] <= tol) {
            launch_bicg_update<T>(n, 3, alpha, 0.0, 0.0, x, r, p, vv, sv, t, dg, ph, sh, v.L, st);
            ++it;
            break;
        }
        A_apply<T>(h, v, lbv.C, sh, t, st);
        reduce(t, t, t, sv, nullptr, q);
        omega = q[0] > 0.0 ? q[1] / q[0] : 0.0;
        if (!(omega != 0.0) || !std::isfinite(omega)) {
            launch_bicg_update<T>(n, 3, alpha, 0.0, 0.0, x, r, p, vv, sv, t, dg, ph, sh, v.L, st);
            ++it;
            break;
        }
        launch_bicg_update<T>(n, 2, alpha, 0.0, omega, x, r, p, vv, sv, t, dg, ph, sh, v.L, st);
        reduce(rh, r, nullptr, nullptr, r, q);
        rho1 = rho;
        rho = q[0];
        if (q[2] <= tol || !std::isfinite(q[2])) {
            ++it;
            break;
        }
    }
    h->bottom_iters += it;
}

template <typename T>
void bottom(ela_hier* h) {
    Level& v = h->lev.back();
    LB<T> b = lb<T>(v);
    if (h->M) launch_bottom<T>(v.L, v.nodes, h->dim, h->M, *b.res, *b.cor, h->stream);
    else bicg_bottom<T>(h);
}

// The V-cycle's way down on level l (pre-smoothing, correction residual,
// restriction) and its way up (prolongation, post-smoothing).
template <typename T>
void v_down(ela_hier* h, int l) {
    Level& v = h->lev[l];
    LB<T> b = lb<T>(v);
    if (h->nu1 == 0) (void)hipMemsetAsync(*b.cor, 0, field_elems(v, h->dim) * sizeof(T), h->stream);
    for (int s = 0; s < h->nu1; ++s) smooth_call<T>(h, l, s == 0);
    res_of_cor<T>(h, l);
    Level& c = h->lev[l + 1];
    launch_restrict<T>(c.L, v.L, h->dim, *b.rc, *lb<T>(c).res, h->stream);
}
// out64: the last sweep writes cor + hold into out64 instead (smooth_call).
template <typename T>
void v_up(ela_hier* h, int l, double* out64 = nullptr) {
    Level& v = h->lev[l];
    Level& c = h->lev[l + 1];
    launch_prolong<T>(v.L, c.L, h->dim, *lb<T>(c).cor, *lb<T>(v).cor, 1, h->stream);
    for (int s = 0; s < h->nu2; ++s) smooth_call<T>(h, l, false, s == h->nu2 - 1 ? out64 : nullptr);
}

template <typename T>
void vcycle(ela_hier* h, int top) {
    const int bot = (int)h->lev.size() - 1;
    for (int l = top; l < bot; ++l) v_down<T>(h, l);
    bottom<T>(h);
    for (int l = bot - 1; l >= top; --l) v_up<T>(h, l);
}

// An F-cycle's part on levels 1..bot: the restrictions of level 0's rhs down
// to the coarsest, the coarsest solve, and the V-cycles from levels bot-1 ..
// 1 with their initial guesses and corrections.  Each level l swaps res with
// rc and cor with hold on the host (fcycle_swaps repeats that for a replay).
template <typename T>
void fcycle_upper(ela_hier* h) {
    const int bot = (int)h->lev.size() - 1;
    for (int l = 1; l <= bot; ++l)
        launch_restrict<T>(h->lev[l].L, h->lev[l - 1].L, h->dim, *lb<T>(h->lev[l - 1]).res, *lb<T>(h->lev[l]).res,
                           h->stream);
    bottom<T>(h);
    for (int l = bot - 1; l >= 1; --l) {
        Level& v = h->lev[l];
        Level& c = h->lev[l + 1];
        LB<T> b = lb<T>(v);
        launch_prolong<T>(v.L, c.L, h->dim, *lb<T>(c).cor, *b.cor, 0, h->stream);
        res_of_cor<T>(h, l);
        std::swap(*b.res, *b.rc);
        std::swap(*b.cor, *b.hold);
        vcycle<T>(h, l);
        launch_axpy<T, T>(v.L, h->dim, *b.hold, *b.cor, nullptr, 1.0, 0, h->stream);
    }
}
template <typename T>
void fcycle_swaps(ela_hier* h) {
    for (int l = (int)h->lev.size() - 2; l >= 1; --l) {
        LB<T> b = lb<T>(h->lev[l]);
        std::swap(*b.res, *b.rc);
        std::swap(*b.cor, *b.hold);
    }
}

// Everything the launches of levels >= 1 read on the host: buffers,
// stiffness, per-level options, the set-up epoch (masks, node lists, kernel
// choices), the boundary table.
template <typename T>
std::vector<unsigned long long> cycle_key(ela_hier* h, int kind) {
    std::vector<unsigned long long> k;
    auto put = [&](const void* p) { k.push_back((unsigned long long)(uintptr_t)p); };
    k.push_back((unsigned long long)kind);
    k.push_back(sizeof(T));
    k.push_back((unsigned long long)h->epoch);
    k.push_back(((unsigned long long)h->pf << 48) | ((unsigned long long)h->ry << 40) | ((unsigned long long)h->kz << 32) |
                ((unsigned long long)h->xcd << 16) | ((unsigned long long)h->mask_mode << 8) |
                (unsigned long long)h->grain_march);
    k.push_back(((unsigned long long)h->nu1 << 32) | (unsigned long long)h->nu2);
    unsigned long long om;
    std::memcpy(&om, &h->omega, sizeof om);
    k.push_back(om);
    k.push_back(((unsigned long long)h->tile_dgt << 16) | ((unsigned long long)h->uniform << 8) |
                (unsigned long long)h->use_table);
    put(h->stream);
    put(h->M);
    put(h->partials);
    put(h->dscale);
    put(h->dnorm);
    put(h->dbc);
    put(h->gid);
    put(h->tab64);
    put(h->tab32);
    put(h->dgt64);
    put(h->dgt32);
    for (int f = 0; f < 26; ++f)
        k.push_back((unsigned long long)h->bc.t[f][0] | ((unsigned long long)h->bc.t[f][1] << 8) |
                    ((unsigned long long)h->bc.t[f][2] << 16));
    for (Level& v : h->lev) {
        LB<T> b = lb<T>(v);
        put(b.C);
        put(b.dg);
        put(*b.res);
        put(*b.rc);
        put(*b.cor);
        put(*b.hold);
        put(*b.tmp);
        put(v.gid);
        k.push_back((unsigned long long)v.kz);
    }
    return k;
}

// The launches of levels >= 1 (kind 0: fcycle_upper, kind 1: the V-cycle
// from level 1) replayed from a HIP graph captured on first use of a key;
// false when graphs are off or the capture failed (the caller then launches
// them).  Level 0's launches, the long ones, stay outside with their timing
// events.  Graphs hold kernel nodes only: a cycle without pre-smoothing
// clears cor with hipMemsetAsync and is never captured.
constexpr int kMaxCycleGraphs = 8;
template <typename T>
bool graph_run(ela_hier* h, int kind) {
    // no capture around a BiCGStab bottom: its coefficients go through the host
    if (!h->graph_on || h->nu1 == 0 || h->lev.size() < 2 || !h->M) return false;
    auto body = [&] {
        if (kind == 0) fcycle_upper<T>(h);
        else vcycle<T>(h, 1);
    };
    auto swaps = [&] {
        if (kind == 0) fcycle_swaps<T>(h);
    };
    std::vector<unsigned long long> key = cycle_key<T>(h, kind);
    for (auto& g : h->graphs)
        if (g.key == key) {
            if (hipGraphLaunch(g.exec, h->stream) != hipSuccess) return false;
            g.used = ++h->graph_clock;
            swaps();
            return true;
        }
    if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        h->graph_on = false;
        return false;
    }
    body();  // records the launches and makes the host swaps
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const bool ok = hipStreamEndCapture(h->stream, &graph) == hipSuccess && graph &&
                    hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) {
        (void)hipGetLastError();
        swaps();  // nothing ran: back to the state before the capture
        h->graph_on = false;
        return false;
    }
    if ((int)h->graphs.size() >= kMaxCycleGraphs) {
        auto lru = std::min_element(h->graphs.begin(), h->graphs.end(),
                                    [](const auto& x, const auto& y) { return x.used < y.used; });
        // A launch of the evicted graph may still be queued on the stream:
        // its executable (kernel nodes and their argument buffers) must
        // outlive it, so drain the stream before destroying it.
        (void)hipStreamSynchronize(h->stream);
        (void)hipGraphExecDestroy(lru->exec);
        h->graphs.erase(lru);
    }
    h->graphs.push_back({key, exec, ++h->graph_clock});
    // the recorded launches use the state before the swaps
    if (hipGraphLaunch(exec, h->stream) != hipSuccess) {
        (void)hipGetLastError();
        swaps();
        h->graph_on = false;
        return false;
    }
    return true;
}

void drop_graphs(ela_hier* h) {
    if (!h->graphs.empty() && h->stream) (void)hipStreamSynchronize(h->stream);  // no launch of them in flight
    for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
    h->graphs.clear();
}

// The V-cycle from level 0, its part on levels >= 1 from a graph when on.
template <typename T>
void vcycle0(ela_hier* h, double* out64 = nullptr) {
    if (h->lev.size() < 2) {
        bottom<T>(h);
        return;
    }
    v_down<T>(h, 0);
    if (!graph_run<T>(h, 1)) vcycle<T>(h, 1);
    v_up<T>(h, 0, out64);
}

// One MG cycle on level 0's res into cor; with out64, an F-cycle leaves its
// correction cor + hold in out64 (FP64, the preconditioned vector of a
// Krylov step) instead of in cor, and one_cycle returns true.  When the
// last sweep is folded (grain-form level 0, ELA_OPT_FOLD on) that sweep
// writes out64 directly and never stores its own value: level 0's cor is
// then undefined after the call, and only out64 holds the correction.
template <typename T>
bool one_cycle(ela_hier* h, double* out64 = nullptr) {
    if (h->cycle != ELA_CYCLE_F || h->lev.size() < 2) {
        vcycle0<T>(h);
        return false;
    }
    if (!graph_run<T>(h, 0)) fcycle_upper<T>(h);
    Level& v = h->lev[0];
    Level& c = h->lev[1];
    LB<T> b = lb<T>(v);
    launch_prolong<T>(v.L, c.L, h->dim, *lb<T>(c).cor, *b.cor, 0, h->stream);
    res_of_cor<T>(h, 0);
    std::swap(*b.res, *b.rc);
    std::swap(*b.cor, *b.hold);
    // a grain-form level 0 folds the final cor + hold into its last sweep
    const bool fold = out64 && h->fold && h->nu2 > 0 && grain_on(h, v);
    vcycle0<T>(h, fold ? out64 : nullptr);
    if (fold) return true;
    if (out64) launch_add_to64<T>(v.L, h->dim, *b.hold, *b.cor, out64, h->stream);
    else launch_axpy<T, T>(v.L, h->dim, *b.hold, *b.cor, nullptr, 1.0, 0, h->stream);
    return out64 != nullptr;
}

// FP64 solution residual of level 0 into the T right-hand side of the next
// cycle (scaled by *dscale), with |r|inf into dnorm.
template <typename T>
void solution_residual(ela_hier* h, const double* u, const double* b) {
    Level& v = h->lev[0];
    const int e = ev_begin(h);
    A_residual<double, T>(h, v, v.C64, u, b, (double*)nullptr, *lb<T>(v).res, h->dscale, h->dnorm, h->stream);
    ev_end(h, e, h->ev_resid);
}

double read_norm(ela_hier* h) {
    (void)hipMemcpyAsync(h->hnorm, h->dnorm, sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream);
    (void)hipStreamSynchronize(h->stream);
    double r;
    std::memcpy(&r, h->hnorm, sizeof r);
    return r;
}

template <typename T>
int solve_loop(ela_hier* h, const double* b, double* u, double tol_rel, double tol_abs, int max_iter, ela_stats* st) {
    const double one = 1.0;
    HIPCHK(hipMemcpyAsync(h->dscale, &one, sizeof one, hipMemcpyHostToDevice, h->stream));
    solution_residual<T>(h, u, b);
    const double r0 = read_norm(h);
    if (!(r0 == r0)) return fail(ELA_ERR_NAN, "initial residual is not finite");
    const double target = std::max(tol_abs, std::max(tol_rel, 1e-16) * r0);
    double rn = r0, best = r0;
    int it = 0, stall = 0;
    int status = ELA_OK;
    if (r0 > target) {
        for (it = 0; it < max_iter;) {
            one_cycle<T>(h);
            launch_axpy<T, double>(h->lev[0].L, h->dim, *lb<T>(h->lev[0]).cor, u, h->dscale, 1.0, 0, h->stream);
            launch_scale_from_norm(h->dnorm, h->dscale, h->stream);
            solution_residual<T>(h, u, b);
            rn = read_norm(h);
            ++it;
            if (!(rn == rn)) { status = fail(ELA_ERR_NAN, "residual became non-finite"); break; }
            if (rn <= target) break;
            if (rn < 0.5 * best) { best = rn; stall = 0; }
            else if (++stall >= 12) {
                status = fail(ELA_ERR_NONCONVERGED, "residual stagnated at " + std::to_string(rn / r0) + " relative");
                break;
            }
            if (rn > 1e8 * r0) { status = fail(ELA_ERR_NONCONVERGED, "multigrid diverged"); break; }
        }
        if (status == ELA_OK && rn > target)
            status = fail(ELA_ERR_NONCONVERGED, "max_iter reached at " + std::to_string(rn / r0) + " relative");
    }
    if (st) {
        st->iters = it;
        st->converged = rn <= target;
        st->res0 = r0;
        st->res = rn;
        st->mg_levels = (int)h->lev.size();
    }
    return status;
}

// Krylov storage: m FP64 preconditioned vectors z, the residual, the device
// scalars, and the basis: m + 1 FP64 vectors, or (b32) m + 1 FP32 vectors and
// one FP64 work vector.
int krylov_alloc(ela_hier* h, bool b32) {
    const int m = h->krylov;
    const bool basis_ok = b32 ? ((int)h->V32.size() == m + 1 && h->kw && h->V.empty())
                              : ((int)h->V.size() == m + 1 && h->V32.empty());
    if ((int)h->Z.size() == m && basis_ok) return ELA_OK;
    for (double* p : h->V) hfree(h, p);
    for (float* p : h->V32) hfree(h, p);
    for (double* p : h->Z) hfree(h, p);
    h->V.clear();
    h->V32.clear();
    h->Z.clear();
    for (double** p : {&h->kr, &h->kdev, &h->kpart, &h->kw}) {
        if (*p) hfree(h, *p);
        *p = nullptr;
    }
    const size_t fe = field_elems(h->lev[0], h->dim);
    int rc = ELA_OK;
    for (int q = 0; q <= m && !rc; ++q) {
        if (b32) {
            float* p = nullptr;
            rc = halloc_t(h, fe, &p);
            h->V32.push_back(p);
        } else {
            double* p = nullptr;
            rc = halloc_t(h, fe, &p);
            h->V.push_back(p);
        }
    }
    for (int q = 0; q < m && !rc; ++q) {
        double* p = nullptr;
        rc = halloc_t(h, fe, &p);
        h->Z.push_back(p);
    }
    if (!rc && b32) rc = halloc_t(h, fe, &h->kw);
    if (!rc) rc = halloc_t(h, fe, &h->kr);
    if (!rc) rc = halloc_t(h, 4 * (size_t)(m + 2), &h->kdev);
    if (!rc) rc = halloc_t(h, (size_t)mdot_partials(m + 1), &h->kpart);
    return rc;
}

// Right-preconditioned flexible GMRES(m) with classical Gram-Schmidt and a
// selective re-orthogonalization; the preconditioner is one MG cycle in
// precision T (FP32 cycles make it vary slightly, which the flexible variant
// allows).  The inner iteration stops when the Arnoldi residual estimate
// |r|_2 falls below the inf-norm target (so |r|_inf does too); convergence is
// then confirmed on the true FP64 residual, from which every restart begins.
// FP64 solves keep the basis in FP64.  Mixed solves keep it in FP32 (w = A z,
// the dot products, the updates and the solution stay FP64): a basis vector
// rounded to FP32 perturbs the Arnoldi relation by ~1e-7 of the restart's
// residual, well below the reduction one restart asks for, and the vectors
// are the MG cycle's FP32 right-hand sides as they stand.  The Gram-Schmidt
// update writes the new FP32 vector (and the next cycle's right-hand side)
// itself, before its norm is known: scaled by a power of two from the norm's
// estimate |w|^2 - |h|^2, so it is stored as s v with v of unit norm and s
// (near 1) kept on the host.  Its dot products are then divided by s and its
// update coefficients by s^2; the cycle's z = M(s v) is a preconditioned
// direction like any other (the flexible variant takes any), and A z enters
// the Hessenberg column as computed.
template <typename T>
int solve_fgmres(ela_hier* h, const double* b, double* u, double tol_rel, double tol_abs, int max_iter,
                 ela_stats* st) {
    constexpr bool B32 = sizeof(T) == 4;
    if (int rc = krylov_alloc(h, B32)) return rc;
    const int m = h->krylov;
    Level& v0 = h->lev[0];
    const long long n = (long long)field_elems(v0, h->dim);
    hipStream_t s = h->stream;
    auto true_residual = [&]() {
        const int e = ev_begin(h);
        A_residual<double, double>(h, v0, v0.C64, u, b, h->kr, (double*)nullptr, nullptr, h->dnorm, s);
        ev_end(h, e, h->ev_resid);
        return read_norm(h);
    };
    auto host_vec = [&](int count) {
        std::vector<double> out(count);
        (void)hipMemcpyAsync(out.data(), h->kdev, sizeof(double) * count, hipMemcpyDeviceToHost, s);
        (void)hipStreamSynchronize(s);
        return out;
    };
    double rn = true_residual();
    const double r0 = rn;
    if (!(r0 == r0)) return fail(ELA_ERR_NAN, "initial residual is not finite");
    const double target = std::max(tol_abs, std::max(tol_rel, 1e-16) * r0);
    int total = 0, status = ELA_OK, restarts = 0;
    double best = r0;
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1);
    std::vector<double> sv(m + 1, 1.0);  // FP32 basis: stored vector q = sv[q] x the unit vector
    while (rn > target) {
        if (total >= max_iter) {
            status = fail(ELA_ERR_NONCONVERGED, "max_iter reached at " + std::to_string(rn / r0) + " relative");
            break;
        }
        // beta = |r|_2, V0 = r / beta
        VecList one{};
        one.v[0] = h->kr;
        launch_mdot(h->kr, one, 1, n, h->kpart, h->kdev, s);
        const double beta = std::sqrt(host_vec(1)[0]);
        if (!(beta > 0)) break;
        // V0 and, in the same pass, the first cycle's right-hand side (T) V0
        if constexpr (B32) launch_scale_to2f(h->kr, 1.0 / beta, h->V32[0], *lb<T>(v0).res, n, s);
        else launch_scale_flat_to<T>(h->V[0], h->kr, 1.0 / beta, *lb<T>(v0).res, n, s);
        sv[0] = 1.0;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int jend = 0;
        for (int j = 0; j < m && total < max_iter; ++j) {
            LB<T> lbv = lb<T>(v0);
            // z_j = M^-1 v_j: one MG cycle on level 0 from zero; res already holds
            // (T) v_j, written with v_j; an F-cycle writes z_j itself
            if (!one_cycle<T>(h, h->Z[j])) launch_axpy<T, double>(v0.L, h->dim, *lbv.cor, h->Z[j], nullptr, 1.0, 1, s);
            // w = A z_j (into V[j+1] with an FP64 basis)
            double* const w = B32 ? h->kw : h->V[j + 1];
            {
                const int e = ev_begin(h);
                A_apply<double>(h, v0, v0.C64, h->Z[j], w, s);
                ev_end(h, e, h->ev_resid);
            }
            VecList Vl{};
            VecList32 Vl32{};
            for (int q = 0; q <= j; ++q) {
                if constexpr (B32) Vl32.v[q] = h->V32[q];
                else Vl.v[q] = h->V[q];
            }
            // classical Gram-Schmidt, h1 = V^T w, w -= V h1, with a second pass only
            // when w lost most of its norm (|w'| < 0.1 |w|, kReorth2: the test of
            // Daniel et al. with a smaller bound).
            // One pass over w gives h1 and |w|^2 (w is the last vector of the
            // list); the update w -= V h1 also sums |w'|^2.
            if constexpr (B32) {
                launch_mdot_self(w, Vl32, j + 1, n, h->kpart, h->kdev, s);
            } else {
                VecList Vw = Vl;
                Vw.v[j + 1] = w;
                launch_mdot(w, Vw, j + 2, n, h->kpart, h->kdev, s);
            }
            std::vector<double> h1 = host_vec(j + 2);
            const double wn0 = std::sqrt(h1[j + 1]);
            h1.resize(j + 1);
            // the FP32 basis' scaled vectors: unit-vector dot products, and the
            // update's coefficient multipliers 1 / s^2 (1 for a unit vector)
            CoefMul mul = coef_ones();
            if constexpr (B32)
                for (int q = 0; q <= j; ++q) {
                    h1[q] /= sv[q];
                    mul.f[q] = 1.0 / (sv[q] * sv[q]);
                }
            std::vector<double> h2(j + 1, 0.0);
            double hn2 = wn0 * wn0;
            for (int q = 0; q <= j; ++q) hn2 -= h1[q] * h1[q];
            double* const norm2 = h->kdev + 2 * (m + 2);  // beyond the coefficients the update reads
            // the FP32 vector's power-of-two scale, from the estimate of its norm
            // (the squared norm of w after the first pass with a second pass)
            auto pow2_inv = [](double e2, double fallback2) {
                const double e = std::sqrt(e2 > 0 ? e2 : fallback2);
                return e > 0 && std::isfinite(e) ? std::exp2(-std::floor(std::log2(e))) : 1.0;
            };
            double osc = 1.0;
            float* const vnext = B32 ? h->V32[j + 1] : nullptr;
            if (!(hn2 > kReorth2 * wn0 * wn0)) {
                if constexpr (B32) {
                    launch_maxpy32(w, Vl32, j + 1, h->kdev, mul, -1.0, n, s);
                    launch_mdot_self(w, Vl32, j + 1, n, h->kpart, h->kdev, s);
                    std::vector<double> r2 = host_vec(j + 2);
                    double e2 = r2[j + 1];
                    for (int q = 0; q <= j; ++q) e2 -= (r2[q] / sv[q]) * (r2[q] / sv[q]);
                    osc = pow2_inv(e2, r2[j + 1]);
                    launch_maxpy_norm32_to2f(w, Vl32, j + 1, h->kdev, mul, -1.0, n, osc, vnext, *lb<T>(v0).res,
                                             h->kpart, norm2, s);
                    h2 = host_vec(j + 1);
                    for (int q = 0; q <= j; ++q) h2[q] /= sv[q];
                } else {
                    launch_maxpy(w, Vl, j + 1, h->kdev, -1.0, n, s);
                    launch_mdot(w, Vl, j + 1, n, h->kpart, h->kdev, s);
                    launch_maxpy_norm(w, Vl, j + 1, h->kdev, -1.0, n, h->kpart, norm2, s);
                    h2 = host_vec(j + 1);
                }
            } else {
                if constexpr (B32) {
                    osc = pow2_inv(hn2, wn0 * wn0);
                    launch_maxpy_norm32_to2f(w, Vl32, j + 1, h->kdev, mul, -1.0, n, osc, vnext, *lb<T>(v0).res,
                                             h->kpart, norm2, s);
                } else {
                    launch_maxpy_norm(w, Vl, j + 1, h->kdev, -1.0, n, h->kpart, norm2, s);
                }
            }
            double hn2_dev = 0.0;
            HIPCHK(hipMemcpyAsync(&hn2_dev, norm2, sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            const double hn = std::sqrt(hn2_dev);
            ++total;
            for (int q = 0; q <= j; ++q) H[(size_t)q * m + j] = h1[q] + h2[q];
            H[(size_t)(j + 1) * m + j] = hn;
            // v_{j+1}: the FP32 basis holds it already as osc hn v_{j+1} (and the
            // next cycle's right-hand side with it); FP64 normalizes it here
            if constexpr (B32) sv[j + 1] = osc * hn;
            else if (hn > 0) launch_scale_flat_to<T>(h->V[j + 1], w, 1.0 / hn, *lb<T>(v0).res, n, s);
            // Givens rotations
            for (int q = 0; q < j; ++q) {
                const double a = H[(size_t)q * m + j], c = H[(size_t)(q + 1) * m + j];
                H[(size_t)q * m + j] = cs[q] * a + sn[q] * c;
                H[(size_t)(q + 1) * m + j] = -sn[q] * a + cs[q] * c;
            }
            const double a = H[(size_t)j * m + j], c = H[(size_t)(j + 1) * m + j];
            const double rr = std::hypot(a, c);
            cs[j] = rr > 0 ? a / rr : 1.0;
            sn[j] = rr > 0 ? c / rr : 0.0;
            H[(size_t)j * m + j] = rr;
            H[(size_t)(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            jend = j + 1;
            const double est = std::fabs(g[j + 1]);
            if (!(est == est)) break;
            if (est <= 0.5 * target || hn == 0) break;
        }
        // y = R^-1 g, u += Z y
        std::vector<double> y(jend);
        for (int q = jend - 1; q >= 0; --q) {
            double acc = g[q];
            for (int k = q + 1; k < jend; ++k) acc -= H[(size_t)q * m + k] * y[k];
            y[q] = acc / H[(size_t)q * m + q];
        }
        HIPCHK(hipMemcpyAsync(h->kdev, y.data(), sizeof(double) * jend, hipMemcpyHostToDevice, s));
        VecList Zl{};
        for (int q = 0; q < jend; ++q) Zl.v[q] = h->Z[q];
        launch_maxpy(u, Zl, jend, h->kdev, 1.0, n, s);
        rn = true_residual();
        ++restarts;
        if (!(rn == rn)) { status = fail(ELA_ERR_NAN, "residual became non-finite"); break; }
        if (rn <= target) break;
        if (rn < 0.5 * best) best = rn;
        else if (restarts >= 4 && rn > 0.9 * best) {
            status = fail(ELA_ERR_NONCONVERGED, "residual stagnated at " + std::to_string(rn / r0) + " relative");
            break;
        }
    }
    if (st) {
        st->iters = total;
        st->converged = rn <= target;
        st->res0 = r0;
        st->res = rn;
        st->mg_levels = (int)h->lev.size();
    }
    return status;
}

void collect_times(ela_hier* h) {
    h->ms_resid = h->ms_smooth = 0;
    h->n_resid = h->n_smooth = 0;
    if (!h->timing) return;
    (void)hipStreamSynchronize(h->stream);
    for (auto& p : h->ev_resid) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, h->ev[p.first], h->ev[p.second]) == hipSuccess) { h->ms_resid += ms; ++h->n_resid; }
    }
    for (auto& p : h->ev_smooth) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, h->ev[p.first], h->ev[p.second]) == hipSuccess) { h->ms_smooth += ms; ++h->n_smooth; }
    }
    if (h->n_resid) h->ms_resid /= h->n_resid;
    if (h->n_smooth) h->ms_smooth /= h->n_smooth;
}

int check_level(const ela_hier* h, int mglev) {
    if (!h) return fail(ELA_ERR_BADARG, "null hierarchy");
    if (mglev < 0 || mglev >= (int)h->lev.size()) return fail(ELA_ERR_BADARG, "MG level out of range");
    return ELA_OK;
}

int check_prec(const ela_hier* h, int prec) {
    if (prec == ELA_F64 && !h->has64) return fail(ELA_ERR_STATE, "hierarchy has no FP64 levels");
    if (prec == ELA_F32 && !h->has32) return fail(ELA_ERR_STATE, "hierarchy has no FP32 levels");
    if (prec != ELA_F64 && prec != ELA_F32) return fail(ELA_ERR_BADARG, "prec must be ELA_F64 or ELA_F32");
    return ELA_OK;
}

hipStream_t strm(ela_hier* h, void* s) { return s ? (hipStream_t)s : h->stream; }

template <typename T>
TabDesc<T> tab_of(ela_hier* h);
// Per-grain interior Jacobi diagonal of level v (the smoothers' form).
template <typename T>
const T* level_dgt(const ela_hier* h, const Level& v) {
    const T* d;
    if constexpr (sizeof(T) == 8) d = h->dgt64;
    else d = h->dgt32;
    return d ? d + (size_t)3 * h->ngrains * (size_t)(&v - &h->lev[0]) : nullptr;
}
template <>
TabDesc<double> tab_of<double>(ela_hier* h) { return {h->gid, h->tab64, h->dgt64}; }
template <>
TabDesc<float> tab_of<float>(ela_hier* h) { return {h->gid, h->tab32, h->dgt32}; }
bool table_on(ela_hier* h, const Level& v) { return h->use_table && h->has_table && &v == &h->lev[0]; }
// Coarse levels of a grain-form hierarchy below the marching size, from
// kTabMinNodes nodes a side: the node kernels read the injected grain ids and
// the table (50 B a node in FP32) instead of the 21 injected stiffness planes
// (132 B); the records are the planes' values, so the bits are the plane
// kernels' (up to the sign of a zero).  Measured on MI355X (256^3
// polycrystal, FP32): the 65^3 sweep 24.1 -> 18.7 us; at 33^3 and below the
// grain-id -> record gather lengthens the latency-bound sweeps (7.7 -> 8.8 us).
constexpr int kTabMinNodes = 65;
bool coarse_tab_on(ela_hier* h, const Level& v) {
    return h->use_table && h->has_table && &v != &h->lev[0] && v.gid && h->dim == 3 && h->kind == K_MM &&
           v.L.nx >= kTabMinNodes && v.L.ny >= kTabMinNodes && v.L.nz >= kTabMinNodes;
}
template <typename T>
TabDesc<T> coarse_tab(ela_hier* h, const Level& v) {
    TabDesc<T> t = tab_of<T>(h);
    t.gid = v.gid;
    t.dgt = h->tile_dgt ? level_dgt<T>(h, v) : nullptr;
    return t;
}
// The grain form holds on every level of a grain-form hierarchy: the coarse
// levels take grain ids and stiffness by injection (so a coarse node's record
// is its grain's), not the full weighting of the per-node planes.
bool grain_levels(const ela_hier* h) { return h->use_table && h->has_table; }
const unsigned short* level_gid(const ela_hier* h, const Level& v) { return &v == &h->lev[0] ? h->gid : v.gid; }

// MODE 2 (FP64 solution residual) exists for FP64 operands only.
template <typename T, typename TO>
void grain_resid64(const GrainBox<T>&, int, double*, unsigned long long*, hipStream_t) {}
template <>
void grain_resid64<double, double>(const GrainBox<double>& g, int kz, double* p, unsigned long long* n, hipStream_t s) {
    launch_grain<double, 2, double>(g, kz, p, n, s);
}
template <>
void grain_resid64<double, float>(const GrainBox<double>& g, int kz, double* p, unsigned long long* n, hipStream_t s) {
    launch_grain<double, 2, float>(g, kz, p, n, s);
}

template <typename T, typename TO>
void plane_resid64(const GrainBox<T>&, int, double*, unsigned long long*, hipStream_t) {}
template <>
void plane_resid64<double, double>(const GrainBox<double>& g, int kz, double* p, unsigned long long* n, hipStream_t s) {
    launch_plane<double, 2, double>(g, kz, p, n, s);
}
template <>
void plane_resid64<double, float>(const GrainBox<double>& g, int kz, double* p, unsigned long long* n, hipStream_t s) {
    launch_plane<double, 2, float>(g, kz, p, n, s);
}

// Share of (wavefront, direction) pairs that must skip the derivative of C
// for a level to read the mask: measured on MI355X, the mask speeds up a
// 257^3 Voronoi level (28 % of the pairs skip) and slows down its coarser
// levels (under 1 %), where the branch and the mask load buy nothing.
constexpr double kMaskMinSkip = 0.2;
// Share of a level's nodes with constant C around them for the level to run
// its FP64 plane stencils in node-list order (those nodes first, so whole
// wavefronts skip the derivative of C; the list order gives up the XCD-slab
// order's L2 reuse).  Measured on MI355X with a 64-grain Voronoi field: its
// 257^3 level (96 %) gains, its 129^3 level (79 %) loses.
constexpr double kListMinPure = 0.9;
// Share for regrouping them inside each block instead (the launch order and
// its L2 reuse stay): a block of 256 nodes then fills whole wavefronts with them.
constexpr double kGroupMinPure = 0.4;

// Mask and node list of a level's planes: valid from ela_prepare until the
// model changes (every model setter clears `prepared`).  ELA_OPT_MASK 2 uses
// the mask and node lists on every plane level, 3 the mask and in-block
// regrouping, 1 each where it pays, 0 neither.
template <typename T>
PlaneOpt plane_opt(const ela_hier* h, const Level& v) {
    PlaneOpt po;
    if (!h->prepared || !v.sm_ok || h->mask_mode == 0) return po;
    // Both orders pay in FP64 only.  Measured on MI355X (256^3 polycrystal):
    // FP64 solves gain 1.07 -> 0.98 s (planes) and 0.62 -> 0.56 s (grain form,
    // FP64 coarse levels); the FP32 stencils do not gain from skipping the
    // derivative of C (their 129^3 sweep: 137 us plain, 154 us regrouped,
    // 221 us in list order).
    const bool fp64 = sizeof(T) == 8;
    const bool list = h->mask_mode == 2 || (h->mask_mode == 1 && v.use_list && fp64);
    const bool group = !list && (h->mask_mode == 3 || (h->mask_mode == 1 && v.use_group && fp64));
    if (list || group || v.use_sm) po.sm = v.sm;
    if (list) {
        po.list = v.nl;
        po.nlist = v.nodes;
    }
    po.group = group;
    return po;
}

// The grain-form levels through the z-marching kernels: level 0, and coarse
// levels of at least kMarchMinNodes nodes a side.  On smaller ones the shell
// and too few tiles dominate and the plane kernels run (measured on MI355X:
// a 129^3 level's FP32 sweep 137 -> 59 us marching, a 65^3 one 28 -> 43 us).
constexpr int kMarchMinNodes = 129;
bool grain_on(ela_hier* h, const Level& v) {
    if (!grain_levels(h) || !h->grain_march || h->ngrains > kMarchMaxGrains || !h->dbc || !grain_offsets_fit(v.L))
        return false;
    if (&v == &h->lev[0]) return true;
    return v.gid && h->dim == 3 && v.L.nx >= kMarchMinNodes && v.L.ny >= kMarchMinNodes && v.L.nz >= kMarchMinNodes;
}

// Marching tile height of a level: level 0 and untuned levels use the
// hierarchy's (ELA_OPT_KZ or the autotuner's level-0 choice).
int level_kz(const ela_hier* h, const Level& v) { return &v != &h->lev[0] && v.kz > 0 ? v.kz : h->kz; }

// Levels from this many nodes (a 257^3 level, not a 129^3 one) take the
// non-temporal epilogue accesses of the marching kernels (GrainBox::nt).
constexpr long long kNtMinNodes = 8LL << 20;
template <typename T>
GrainBox<T> grain_box(ela_hier* h, const Level& v) {
    GrainBox<T> g{};
    g.L = v.L;
    g.iv = make_inv<T>(v.h);
    g.bc = h->dbc;
    g.uniform = h->uniform;
    g.gid = level_gid(h, v);
    g.tab = tab_of<T>(h).tab;
    g.sc = nullptr;
    g.dgt = h->tile_dgt ? level_dgt<T>(h, v) : nullptr;
    g.ngr = h->ngrains;
    g.xcd = h->xcd;
    g.ry = &v == &h->lev[0] ? h->ry : 1;
    g.pf = h->pf;
    g.nt = (long long)v.nodes >= kNtMinNodes;
    g.xface = true;
    return g;
}

// Isotropic per-node-plane levels through the z-marching kernels (3D): level
// 0, and levels of at least kMarchMinNodes nodes a side, as in grain form.
// The stencil is the node kernels' expression, so the bits are theirs.
// Major/minor planes keep the node kernels (launch_plane).
bool plane_march_on(ela_hier* h, const Level& v) {
    if (h->dim != 3 || !h->grain_march || !h->dbc || h->kind != K_ISO) return false;
    if (grain_on(h, v) || table_on(h, v) || coarse_tab_on(h, v)) return false;
    if (!grain_offsets_fit(v.L) || !plane_offsets_fit(v.L, 2, 8)) return false;
    if (&v == &h->lev[0]) return true;
    return v.L.nx >= kMarchMinNodes && v.L.ny >= kMarchMinNodes && v.L.nz >= kMarchMinNodes;
}

// The marching operands of a plane level: its stiffness planes C and, where
// the last ela_prepare built one, its same-direction mask.
template <typename T>
GrainBox<T> plane_box(ela_hier* h, const Level& v, const T* C) {
    GrainBox<T> g{};
    g.L = v.L;
    g.iv = make_inv<T>(v.h);
    g.bc = h->dbc;
    g.uniform = h->uniform;
    g.C = C;
    g.sm = h->prepared && v.sm_ok && h->mask_mode != 0 ? v.sm : nullptr;
    g.xcd = h->xcd;
    g.ry = 1;
    g.pf = h->pf;
    g.nt = (long long)v.nodes >= kNtMinNodes;
    g.xface = true;
    return g;
}

template <typename T>
void A_apply(ela_hier* h, Level& v, const T* C, const T* u, T* f, hipStream_t s) {
    if (plane_march_on(h, v)) {
        GrainBox<T> g = plane_box<T>(h, v, C);
        g.in = u;
        g.out = f;
        launch_plane<T, 3, T>(g, level_kz(h, v), nullptr, nullptr, s);
    } else if (grain_on(h, v)) {
        GrainBox<T> g = grain_box<T>(h, v);
        g.in = u;
        g.out = f;
        launch_grain<T, 3, T>(g, level_kz(h, v), nullptr, nullptr, s);
    } else if (table_on(h, v)) launch_fapply_tab<T>(v.L, tab_of<T>(h), u, f, h->bc, v.h, h->uniform, s);
    else if (coarse_tab_on(h, v)) launch_fapply_tab<T>(v.L, coarse_tab<T>(h, v), u, f, h->bc, v.h, h->uniform, s);
    else launch_fapply<T>(v.L, h->kind, C, u, f, h->bc, v.h, h->uniform, s, plane_opt<T>(h, v));
}

template <typename T>
void A_jacobi(ela_hier* h, Level& v, const T* C, const T* dg, const T* xin, const T* b, T* xout, hipStream_t s) {
    if (plane_march_on(h, v)) {
        GrainBox<T> g = plane_box<T>(h, v, C);
        g.in = xin;
        g.rhs = b;
        g.out = xout;
        g.dg = dg;
        g.omega = (T)h->omega;
        launch_plane<T, 0, T>(g, level_kz(h, v), nullptr, nullptr, s);
    } else if (grain_on(h, v)) {
        GrainBox<T> g = grain_box<T>(h, v);
        g.in = xin;
        g.rhs = b;
        g.out = xout;
        g.dg = dg;
        g.omega = (T)h->omega;
        launch_grain<T, 0, T>(g, level_kz(h, v), nullptr, nullptr, s);
    } else if (table_on(h, v)) launch_jacobi_tab<T>(v.L, tab_of<T>(h), dg, xin, b, xout, h->bc, v.h, h->uniform, h->omega, s);
    else if (coarse_tab_on(h, v)) launch_jacobi_tab<T>(v.L, coarse_tab<T>(h, v), dg, xin, b, xout, h->bc, v.h, h->uniform, h->omega, s);
    else launch_jacobi<T>(v.L, h->kind, C, dg, xin, b, xout, h->bc, v.h, h->uniform, h->omega, s, plane_opt<T>(h, v));
}

template <typename T>
void grain_jacobi_to64(ela_hier* h, Level& v, const T* dg, const T* xin, const T* b, const T* hold, double* out64,
                       hipStream_t s) {
    GrainBox<T> g = grain_box<T>(h, v);
    g.in = xin;
    g.rhs = b;
    g.dg = dg;
    g.hold = hold;
    g.out2 = out64;
    g.omega = (T)h->omega;
    launch_grain<T, 4, T>(g, level_kz(h, v), nullptr, nullptr, s);
}

template <typename T, typename TO>
void A_residual(ela_hier* h, Level& v, const T* C, const T* x, const T* b, T* r, TO* rs, const double* scale,
                unsigned long long* norm, hipStream_t s, double* part) {
    double* const P = part ? part : h->partials;
    const bool plain = !rs && !norm;
    if (plane_march_on(h, v) && (plain || sizeof(T) == 8)) {
        GrainBox<T> g = plane_box<T>(h, v, C);
        g.in = x;
        g.rhs = b;
        if (plain) {
            g.out = r;
            launch_plane<T, 1, T>(g, level_kz(h, v), nullptr, nullptr, s);
        } else {
            g.out = rs;
            g.out2 = (double*)r;
            g.scale = scale;
            plane_resid64<T, TO>(g, level_kz(h, v), P, norm, s);
        }
    } else if (grain_on(h, v) && (plain || sizeof(T) == 8)) {
        GrainBox<T> g = grain_box<T>(h, v);
        g.in = x;
        g.rhs = b;
        if (plain) {
            g.out = r;
            launch_grain<T, 1, T>(g, level_kz(h, v), nullptr, nullptr, s);
        } else {
            g.out = rs;
            g.out2 = (double*)r;
            g.scale = scale;
            grain_resid64<T, TO>(g, level_kz(h, v), P, norm, s);
        }
    } else if (table_on(h, v)) launch_residual_tab<T, TO>(v.L, tab_of<T>(h), x, b, r, rs, scale, P, norm, h->bc, v.h, h->uniform, s);
    else if (coarse_tab_on(h, v)) launch_residual_tab<T, TO>(v.L, coarse_tab<T>(h, v), x, b, r, rs, scale, P, norm, h->bc, v.h, h->uniform, s);
    else launch_residual<T, TO>(v.L, h->kind, C, x, b, r, rs, scale, P, norm, h->bc, v.h, h->uniform, s,
                                plane_opt<T>(h, v));
}

}  // namespace

// ====================================================================== composite-solve hooks
namespace ela {

int hier_level0(ela_hier* h, Lay* L) {
    if (!h || !L) return fail(ELA_ERR_BADARG, "null argument");
    *L = h->lev[0].L;
    return ELA_OK;
}

void* hier_res0(ela_hier* h, int prec) {
    Level& v = h->lev[0];
    return prec == ELA_F32 ? (void*)v.res32 : (void*)v.res64;
}

void* hier_cor0(ela_hier* h, int prec) {
    Level& v = h->lev[0];
    return prec == ELA_F32 ? (void*)v.cor32 : (void*)v.cor64;
}

int hier_cycle(ela_hier* h, int prec) {
    if (!h) return fail(ELA_ERR_BADARG, "null hierarchy");
    if (!h->prepared) return fail(ELA_ERR_STATE, "call ela_prepare first");
    if (prec == ELA_F32) {
        if (!h->has32) return fail(ELA_ERR_STATE, "no FP32 levels");
        one_cycle<float>(h);
    } else {
        if (!h->has64) return fail(ELA_ERR_STATE, "no FP64 levels");
        one_cycle<double>(h);
    }
    return ELA_OK;
}

int hier_share_stream(ela_hier* h, hipStream_t s) {
    if (!h || !s) return fail(ELA_ERR_BADARG, "null argument");
    if (h->stream == s) return ELA_OK;
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->own_stream) HIPCHK(hipStreamDestroy(h->stream));
    h->stream = s;
    h->own_stream = false;
    return ELA_OK;
}

hipStream_t hier_stream(const ela_hier* h) { return h->stream; }

double* hier_C0(ela_hier* h) { return h->lev[0].C64; }

void hier_model_set(ela_hier* h) {
    h->has_table = false;
    h->model_set = true;
    h->prepared = false;
}

}  // namespace ela

// ====================================================================== C ABI
extern "C" {

const char* ela_last_error(void) { return g_err.c_str(); }

int ela_ctx_create(int device, size_t hbm_budget_bytes, ela_ctx** out) {
    if (!out) return fail(ELA_ERR_BADARG, "null out");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(ELA_ERR_BADARG, "no such device");
    HIPCHK(hipSetDevice(device));
    if (int rc = scratch_init(device)) return rc;
    ela_ctx* c = new ela_ctx();
    c->device = device;
    c->budget = hbm_budget_bytes;
    *out = c;
    return ELA_OK;
}

int ela_ctx_destroy(ela_ctx* ctx) {
    delete ctx;
    return ELA_OK;
}

size_t ela_ctx_bytes_in_use(const ela_ctx* ctx) { return ctx ? ctx->in_use : 0; }

int ela_dev_alloc(ela_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(ELA_ERR_BADARG, "null argument");
    *out = nullptr;
    if (bytes == 0) return ELA_OK;
    HIPCHK(hipSetDevice(ctx->device));
    if (int rc = budget_reserve(ctx, bytes)) return rc;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        budget_release(ctx, bytes);
        return fail(ELA_ERR_OOM, "hipMalloc failed");
    }
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->raw[p] = bytes;
    }
    *out = p;
    return ELA_OK;
}

int ela_dev_free(ela_ctx* ctx, void* p) {
    if (!ctx) return fail(ELA_ERR_BADARG, "null context");
    if (!p) return ELA_OK;
    size_t bytes = 0;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        auto it = ctx->raw.find(p);
        if (it == ctx->raw.end()) return fail(ELA_ERR_BADARG, "not a buffer of ela_dev_alloc on this context");
        bytes = it->second;
        ctx->raw.erase(it);
    }
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipFree(p));
    budget_release(ctx, bytes);
    return ELA_OK;
}

int ela_memcpy(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return ELA_OK;
    if (!dst || !src) return fail(ELA_ERR_BADARG, "null pointer");
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    return ELA_OK;
}

int ela_hier_create(ela_ctx* ctx, const ela_domain* dom, int model_kind, int max_mg_levels, int mode,
                    ela_hier** out) {
    return ela_hier_create_ex(ctx, dom, model_kind, max_mg_levels, 4, mode, out);
}

int ela_hier_create_ex(ela_ctx* ctx, const ela_domain* dom, int model_kind, int max_mg_levels, int coarse_min,
                       int mode, ela_hier** out) {
    if (!ctx || !dom || !out) return fail(ELA_ERR_BADARG, "null argument");
    if (coarse_min < 2) return fail(ELA_ERR_BADARG, "coarse_min must be >= 2");
    *out = nullptr;
    if (dom->dim != 2 && dom->dim != 3) return fail(ELA_ERR_BADARG, "dim must be 2 or 3");
    for (int d = 0; d < dom->dim; ++d) {
        if (dom->n[d] < 2) return fail(ELA_ERR_BADARG, "need at least 2 cells per direction");
        if (!(dom->dx[d] > 0)) return fail(ELA_ERR_BADARG, "cell size must be positive");
        if (dom->n[d] > (1 << 14)) return fail(ELA_ERR_BADARG, "too many cells per direction");
    }
    if (model_kind < 0 || model_kind > 2) return fail(ELA_ERR_BADARG, "unknown model kind");
    if (mode < 0 || mode > 2) return fail(ELA_ERR_BADARG, "unknown mode");
    HIPCHK(hipSetDevice(ctx->device));
    ela_hier* h = new ela_hier();
    h->ctx = ctx;
    h->dim = dom->dim;
    h->kind = model_kind;
    h->nc = ncomp_of(dom->dim, model_kind);
    h->mode = mode;
    h->has64 = (mode == ELA_SOLVE_FP64 || mode == ELA_ALLOC_BOTH);
    h->has32 = (mode == ELA_SOLVE_MIXED || mode == ELA_ALLOC_BOTH);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(ELA_ERR_HIP, "hipStreamCreate failed");
    }
    int n[3] = {dom->n[0], dom->n[1], dom->dim == 3 ? dom->n[2] : 0};
    double dx[3] = {dom->dx[0], dom->dx[1], dom->dim == 3 ? dom->dx[2] : 1.0};
    for (;;) {
        Level v;
        for (int d = 0; d < 3; ++d) { v.n[d] = n[d]; v.h[d] = dx[d]; }
        v.L.dim = dom->dim;
        v.L.nx = n[0] + 1;
        v.L.ny = n[1] + 1;
        v.L.nz = dom->dim == 3 ? n[2] + 1 : 1;
        v.L.px = (v.L.nx + h->pitch_align - 1) / h->pitch_align * h->pitch_align;
        v.L.pl = v.L.px * v.L.ny;
        v.L.cs = (v.L.pl * v.L.nz + 255) / 256 * 256;
        v.nodes = v.L.nx * v.L.ny * v.L.nz;
        h->lev.push_back(v);
        bool can = true;
        for (int d = 0; d < dom->dim; ++d) can = can && (n[d] % 2 == 0) && (n[d] / 2 >= coarse_min);
        if (max_mg_levels > 0 && (int)h->lev.size() >= max_mg_levels) can = false;
        if (!can) break;
        for (int d = 0; d < dom->dim; ++d) { n[d] /= 2; dx[d] *= 2; }
    }
    h->nunk = h->dim * h->lev.back().nodes;
    int rc = ELA_OK;
    for (size_t l = 0; l < h->lev.size() && rc == ELA_OK; ++l) {
        Level& v = h->lev[l];
        const size_t fe = field_elems(v, h->dim);
        rc = halloc_t(h, field_elems(v, h->nc), &v.C64);
        if (rc == ELA_OK) rc = halloc_t(h, field_elems(v, 1), &v.sm);
        if (rc == ELA_OK) rc = halloc_t(h, (size_t)v.nodes, &v.nl);
        if (rc == ELA_OK) rc = halloc_t(h, 4 * (size_t)node_list_blocks(v.L), &v.nlcnt);
        if (rc == ELA_OK) rc = halloc_t(h, 2, &v.nltot);
        if (rc == ELA_OK && h->has64) {
            rc = halloc_t(h, fe, &v.d64);
            if (!rc) rc = halloc_t(h, fe, &v.cor64);
            if (!rc) rc = halloc_t(h, fe, &v.tmp64);
            if (!rc) rc = halloc_t(h, fe, &v.res64);
            if (!rc) rc = halloc_t(h, fe, &v.rc64);
            if (!rc) rc = halloc_t(h, fe, &v.hold64);
        }
        if (rc == ELA_OK && h->has32) {
            rc = halloc_t(h, field_elems(v, h->nc), &v.C32);
            if (!rc) rc = halloc_t(h, fe, &v.d32);
            if (!rc) rc = halloc_t(h, fe, &v.cor32);
            if (!rc) rc = halloc_t(h, fe, &v.tmp32);
            if (!rc) rc = halloc_t(h, fe, &v.res32);
            if (!rc) rc = halloc_t(h, fe, &v.rc32);
            if (!rc) rc = halloc_t(h, fe, &v.hold32);
        }
    }
    if (!rc) rc = halloc_t(h, field_elems(h->lev[0], h->dim), &h->u64);
    if (!rc) rc = halloc_t(h, field_elems(h->lev[0], h->dim), &h->b64);
    if (!rc && h->nunk <= kMaxDenseUnknowns) rc = halloc_t(h, (size_t)h->nunk * 2 * h->nunk, &h->M);
    if (!rc && !h->M) {  // the coarsest level is solved by BiCGStab
        const size_t fe = field_elems(h->lev.back(), h->dim);
        for (int k = 0; k < 8 && !rc; ++k) {
            if (h->has64) rc = halloc_t(h, fe, &h->bv64[k]);
            if (!rc && h->has32) rc = halloc_t(h, fe, &h->bv32[k]);
        }
        if (!rc) rc = halloc_t(h, (size_t)bicg_partials(), &h->bpart);
        if (!rc) rc = halloc_t(h, 4, &h->bout);
        if (!rc && hipHostMalloc((void**)&h->bhost, 64) != hipSuccess) rc = fail(ELA_ERR_OOM, "hipHostMalloc failed");
    }
    if (!rc) rc = halloc_t(h, 4, &h->dnorm);
    long long np = 0;
    for (auto& v : h->lev) np = std::max(np, residual_partials(v.L));
    if (h->dim == 3) np = std::max(np, grain_partials(h->lev[0].L, 1));  // every z-marching tile height
    if (!rc) rc = halloc_t(h, 1, &h->dbc);
    if (!rc) rc = halloc_t(h, (size_t)np, &h->partials);
    if (!rc) rc = halloc_t(h, 4, &h->dscale);
    if (!rc) rc = halloc_t(h, 4, &h->dflag);
    if (!rc) rc = halloc_t(h, 2 * h->lev.size(), &h->mstats);
    if (!rc && hipHostMalloc((void**)&h->hnorm, 64) != hipSuccess) rc = fail(ELA_ERR_OOM, "hipHostMalloc failed");
    if (rc) {
        std::string keep = g_err;
        ela_hier_destroy(h);
        g_err = keep;
        return rc;
    }
    // default BCs: Displacement everywhere (BC/Operator/Elastic.H:46-55)
    std::memset(&h->bc, 0, sizeof h->bc);
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
        ela_hier_destroy(h);
        return fail(ELA_ERR_HIP, "zeroing the hierarchy failed");
    }
    *out = h;
    return ELA_OK;
}

int ela_hier_destroy(ela_hier* h) {
    if (!h) return ELA_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    for (auto& e : h->ev) (void)hipEventDestroy(e);
    while (!h->allocs.empty()) hfree(h, h->allocs.back().first);
    if (h->hnorm) (void)hipHostFree(h->hnorm);
    if (h->bhost) (void)hipHostFree(h->bhost);
    if (h->stream && h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return ELA_OK;
}

int ela_hier_levels(const ela_hier* h) { return h ? (int)h->lev.size() : 0; }
size_t ela_hier_bytes(const ela_hier* h) { return h ? h->bytes : 0; }

int ela_level_info(const ela_hier* h, int mglev, ela_level_layout* out) {
    if (int rc = check_level(h, mglev)) return rc;
    if (!out) return fail(ELA_ERR_BADARG, "null out");
    const Level& v = h->lev[mglev];
    out->dim = h->dim;
    for (int d = 0; d < 3; ++d) { out->n[d] = v.n[d]; out->dx[d] = v.h[d]; }
    out->nodes[0] = v.L.nx;
    out->nodes[1] = v.L.ny;
    out->nodes[2] = v.L.nz;
    out->pitch = v.L.px;
    out->plane = v.L.pl;
    out->cstride = v.L.cs;
    return ELA_OK;
}

// Copy a dense host array (nc planes) into level mglev's layout on the device.
static int upload_dense(ela_hier* h, const Level& v, int nc, const double* host, void* dev, int prec) {
    const size_t n = (size_t)v.nodes * nc;
    void* stage = nullptr;
    int rc = halloc(h, n * sizeof(double), &stage);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(stage, host, n * sizeof(double), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        if (prec == ELA_F64) launch_pack<double>(v.L, nc, (const double*)stage, (double*)dev, 1, h->stream);
        else launch_pack<float>(v.L, nc, (const double*)stage, (float*)dev, 1, h->stream);
        e = hipStreamSynchronize(h->stream);
    }
    hfree(h, stage);
    if (e != hipSuccess) return fail(ELA_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    return ELA_OK;
}

static int download_dense(ela_hier* h, const Level& v, int nc, const void* dev, double* host, int prec) {
    const size_t n = (size_t)v.nodes * nc;
    void* stage = nullptr;
    int rc = halloc(h, n * sizeof(double), &stage);
    if (rc) return rc;
    if (prec == ELA_F64) launch_pack<double>(v.L, nc, (const double*)stage, (double*)dev, 0, h->stream);
    else launch_pack<float>(v.L, nc, (const double*)stage, (float*)dev, 0, h->stream);
    hipError_t e = hipMemcpyAsync(host, stage, n * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hfree(h, stage);
    if (e != hipSuccess) return fail(ELA_ERR_HIP, std::string("download: ") + hipGetErrorString(e));
    return ELA_OK;
}

int ela_set_model_field(ela_hier* h, const double* C_host) {
    if (!h || !C_host) return fail(ELA_ERR_BADARG, "null argument");
    h->has_table = false;
    if (int rc = upload_dense(h, h->lev[0], h->nc, C_host, h->lev[0].C64, ELA_F64)) return rc;
    h->model_set = true;
    h->prepared = false;
    return ELA_OK;
}

int ela_set_model_voronoi_cubic(ela_hier* h, int nseeds, const int64_t* seeds, int q, const double* R, double C11,
                                double C12, double C44) {
    if (!h || !seeds || !R) return fail(ELA_ERR_BADARG, "null argument");
    if (h->kind != ELA_MODEL_MM || h->dim != 3) return fail(ELA_ERR_STATE, "Voronoi cubic needs a 3D major/minor hierarchy");
    if (nseeds < 1 || q < 1) return fail(ELA_ERR_BADARG, "need at least one seed and q >= 1");
    const Level& v = h->lev[0];
    for (int s = 0; s < nseeds; ++s)
        for (int d = 0; d < 3; ++d) {
            const long long lim = (long long)v.n[d] * q;
            if (seeds[3 * s + d] < -lim || seeds[3 * s + d] > 2 * lim) return fail(ELA_ERR_BADARG, "seed outside the domain");
        }
    if (nseeds > 65535) return fail(ELA_ERR_BADARG, "at most 65535 grains");
    // per-grain records and the per-node grain id stay resident (grain-table mode)
    for (void* p : {(void*)h->gid, (void*)h->tab64, (void*)h->dgt64, (void*)h->tab32, (void*)h->dgt32})
        if (p) hfree(h, p);
    h->gid = nullptr;
    for (size_t l = 1; l < h->lev.size(); ++l) {
        if (h->lev[l].gid) hfree(h, h->lev[l].gid);
        h->lev[l].gid = nullptr;
    }
    h->tab64 = h->dgt64 = nullptr;
    h->tab32 = h->dgt32 = nullptr;
    h->has_table = false;
    void *dseeds = nullptr, *dR = nullptr;
    int rc = halloc(h, sizeof(long long) * 3 * nseeds, &dseeds);
    if (!rc) rc = halloc(h, sizeof(double) * 9 * nseeds, &dR);
    if (!rc) rc = halloc_t(h, (size_t)v.L.cs, &h->gid);
    for (size_t l = 1; l < h->lev.size() && !rc; ++l) rc = halloc_t(h, (size_t)h->lev[l].L.cs, &h->lev[l].gid);
    if (!rc) rc = halloc_t(h, (size_t)21 * nseeds, &h->tab64);
    if (!rc) rc = halloc_t(h, (size_t)21 * nseeds, &h->tab32);
    const size_t ndg = (size_t)3 * nseeds * h->lev.size();
    if (!rc) rc = halloc_t(h, ndg, &h->dgt64);
    if (!rc) rc = halloc_t(h, ndg, &h->dgt32);
    hipError_t e = hipSuccess;
    if (!rc) {
        e = hipMemcpyAsync(dseeds, seeds, sizeof(long long) * 3 * nseeds, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dR, R, sizeof(double) * 9 * nseeds, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) {
            launch_cubic_table(nseeds, (const double*)dR, C11, C12, C44, h->tab64, h->stream);
            launch_convert<double, float>(21LL * nseeds, h->tab64, h->tab32, h->stream);
            for (size_t l = 0; l < h->lev.size(); ++l)  // the coarse levels' records are the grains' (injection)
                launch_grain_diag(nseeds, h->tab64, h->lev[l].h, h->dgt64 + 3 * nseeds * l,
                                  h->dgt32 + 3 * nseeds * l, h->stream);
            launch_voronoi_cubic(v.L, nseeds, (const long long*)dseeds, q, h->tab64, v.C64, h->gid, h->stream);
            for (size_t l = 1; l < h->lev.size(); ++l)
                launch_inject_level<unsigned short>(h->lev[l].L, h->lev[l - 1].L, 1, level_gid(h, h->lev[l - 1]),
                                                    h->lev[l].gid, h->stream);
            e = hipStreamSynchronize(h->stream);
        }
    }
    if (dseeds) hfree(h, dseeds);
    if (dR) hfree(h, dR);
    if (rc) return rc;
    if (e != hipSuccess) return fail(ELA_ERR_HIP, std::string("voronoi stiffness: ") + hipGetErrorString(e));
    h->ngrains = nseeds;
    h->has_table = true;
    h->model_set = true;
    h->prepared = false;
    return ELA_OK;
}

// Cell-valued order parameter with one ghost layer on level 0.
static EtaSrc cell_src(const Level& v, int dim, const double* p, int avg) {
    EtaSrc e{p, avg, v.n[0] + 2, v.n[1] + 2, 0};
    e.cs = (long long)e.gx * e.gy * (dim == 3 ? v.n[2] + 2 : 1);
    return e;
}

// Run the eta checks of DegradeModulus on the device and turn them into a status.
static int eta_status(ela_hier* h, const Level& v, const EtaSrc& e, int c0, int c1, hipStream_t s) {
    HIPCHK(hipMemsetAsync(h->dflag, 0, sizeof(int), s));
    launch_eta_check(v.L, e, c0, c1, h->dflag, s);
    int f = 0;
    HIPCHK(hipMemcpyAsync(&f, h->dflag, sizeof f, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (f & 1) return fail(ELA_ERR_NAN, "Nans in eta");
    if (f & 2) return fail(ELA_ERR_NAN, "Infs in eta");
    if (f & 4) return fail(ELA_ERR_BADARG, "Invalid value of eta (outside [0, 1])");
    return ELA_OK;
}

int ela_set_model_degraded_iso(ela_hier* h, double lambda0, double mu0, int neta, const double* eta) {
    if (!h || !eta) return fail(ELA_ERR_BADARG, "null argument");
    if (h->kind != ELA_MODEL_ISO) return fail(ELA_ERR_STATE, "degraded isotropic needs an isotropic hierarchy");
    if (neta != 1 && neta != 2) return fail(ELA_ERR_BADARG, "one or two eta fields");
    const Level& v = h->lev[0];
    double* deta = nullptr;
    int rc = halloc_t(h, (size_t)v.nodes * neta, &deta);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(deta, eta, sizeof(double) * v.nodes * neta, hipMemcpyHostToDevice, h->stream));
    const EtaSrc e{deta, ETA_NODES, 0, 0, (long long)v.nodes};
    rc = eta_status(h, v, e, 0, neta, h->stream);
    if (!rc) {
        launch_degrade_iso(v.L, e, 0, neta - 1, lambda0, mu0, v.C64, h->stream);
        if (hipStreamSynchronize(h->stream) != hipSuccess) rc = fail(ELA_ERR_HIP, "degraded stiffness kernel failed");
    }
    hfree(h, deta);
    if (rc) return rc;
    h->has_table = false;
    h->model_set = true;
    h->prepared = false;
    return ELA_OK;
}

int ela_degrade_material(ela_hier* h, double lambda0, double mu0, int neta, const double* eta_cells, void* stream) {
    if (!h || !eta_cells) return fail(ELA_ERR_BADARG, "null argument");
    if (h->kind != ELA_MODEL_ISO) return fail(ELA_ERR_STATE, "DegradeMaterial needs an isotropic hierarchy");
    if (neta < 1) return fail(ELA_ERR_BADARG, "number_of_eta must be positive");
    const Level& v = h->lev[0];
    hipStream_t s = strm(h, stream);
    const EtaSrc e = cell_src(v, h->dim, eta_cells, ETA_CELLS);
    if (int rc = eta_status(h, v, e, 0, neta, s)) return rc;
    // each component calls DegradeModulus(temp) from (lambda0, mu0): the last one stays
    launch_degrade_iso(v.L, e, neta - 1, neta - 1, lambda0, mu0, v.C64, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    h->has_table = false;
    h->model_set = true;
    h->prepared = false;
    return ELA_OK;
}

int ela_combine_cubic(ela_hier* h, int nph, const double* ddw, const double* F0, const double* eta_cells,
                      void* F0_out, void* stream) {
    if (!h || !ddw || !eta_cells) return fail(ELA_ERR_BADARG, "null argument");
    if (h->kind != ELA_MODEL_MM) return fail(ELA_ERR_STATE, "Combine needs a major/minor hierarchy");
    if (nph < 1 || nph > 4096) return fail(ELA_ERR_BADARG, "1 to 4096 models");
    const Level& v = h->lev[0];
    hipStream_t s = strm(h, stream);
    const int nc = h->nc;
    double* dtab = nullptr;
    int rc = halloc_t(h, (size_t)nph * (nc + 9), &dtab);
    if (rc) return rc;
    std::vector<double> tab((size_t)nph * (nc + 9), 0.0);
    std::copy(ddw, ddw + (size_t)nph * nc, tab.begin());
    if (F0) std::copy(F0, F0 + (size_t)nph * 9, tab.begin() + (size_t)nph * nc);
    hipError_t e = hipMemcpyAsync(dtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(h->dflag, 0, sizeof(int), s);
    if (e == hipSuccess) {
        launch_combine(v.L, cell_src(v, h->dim, eta_cells, ETA_XY4), nph, dtab, F0 ? dtab + (size_t)nph * nc : nullptr,
                       v.C64, (double*)F0_out, h->dflag, s);
        e = hipGetLastError();
    }
    int f = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&f, h->dflag, sizeof f, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hfree(h, dtab);
    if (e != hipSuccess) return fail(ELA_ERR_HIP, std::string("Combine: ") + hipGetErrorString(e));
    h->has_table = false;
    h->model_set = true;
    h->prepared = false;
    if (f) return fail(ELA_ERR_NAN, "Combine: non-finite stiffness (the eta weights sum to zero at a node)");
    return ELA_OK;
}

int ela_set_bc(ela_hier* h, const int* types, const double* values) {
    if (!h || !types) return fail(ELA_ERR_BADARG, "null argument");
    const int nf = h->dim == 3 ? 26 : 8;
    for (int f = 0; f < nf; ++f)
        for (int c = 0; c < h->dim; ++c) {
            const int t = types[3 * f + c];
            if (t < 0 || t > 3) return fail(ELA_ERR_BADARG, "unknown boundary type");
        }
    std::memset(&h->bc, 0, sizeof h->bc);
    for (int f = 0; f < nf; ++f)
        for (int c = 0; c < 3; ++c) {
            h->bc.t[f][c] = (unsigned char)(c < h->dim ? types[3 * f + c] : 0);
            h->bcval[f][c] = values ? values[3 * f + c] : 0.0;
        }
    if (h->dbc) HIPCHK(hipMemcpyAsync(h->dbc, &h->bc, sizeof h->bc, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->prepared = false;
    return ELA_OK;
}

int ela_set_uniform(ela_hier* h, int uniform) {
    if (!h) return fail(ELA_ERR_BADARG, "null hierarchy");
    h->uniform = uniform ? 1 : 0;
    h->prepared = false;
    return ELA_OK;
}

int ela_set_cycle(ela_hier* h, int cycle, int nu1, int nu2) {
    if (!h) return fail(ELA_ERR_BADARG, "null hierarchy");
    if ((cycle != ELA_CYCLE_F && cycle != ELA_CYCLE_V) || nu1 < 0 || nu2 < 0 || nu1 > 16 || nu2 > 16)
        return fail(ELA_ERR_BADARG, "bad cycle options");
    h->cycle = cycle;
    h->nu1 = nu1;
    h->nu2 = nu2;
    return ELA_OK;
}

int ela_prepare(ela_hier* h) {
    if (!h) return fail(ELA_ERR_BADARG, "null hierarchy");
    if (!h->model_set) return fail(ELA_ERR_STATE, "Attempting to use operator before calling SetModel");
    ++h->epoch;  // captured cycles of the previous set-up no longer apply
    // stiffness on every MG level (FP64), then FP32 copies
    for (size_t l = 1; l < h->lev.size(); ++l) {
        if (grain_levels(h))  // the records of the coarse grain ids (Elastic.cpp:759-803 averages instead)
            launch_inject_level<double>(h->lev[l].L, h->lev[l - 1].L, h->nc, h->lev[l - 1].C64, h->lev[l].C64,
                                        h->stream);
        else
            launch_restrict<double>(h->lev[l].L, h->lev[l - 1].L, h->nc, h->lev[l - 1].C64, h->lev[l].C64,
                                    h->stream);
    }
    if (h->has32)
        for (auto& v : h->lev) launch_convert<double, float>((long long)field_elems(v, h->nc), v.C64, v.C32, h->stream);
    // where C is constant across a node, the plane-form stencils may skip its
    // derivative: worth it on a level where whole wavefronts skip often
    // (not on a level that runs in grain form, where the grain ids say it)
    const size_t nl = h->lev.size();
    std::vector<char> masked(nl, 0);
    HIPCHK(hipMemsetAsync(h->mstats, 0, 2 * nl * sizeof(unsigned long long), h->stream));
    for (size_t l = 0; l < nl; ++l) {
        Level& v = h->lev[l];
        if (table_on(h, v) || grain_on(h, v)) continue;
        launch_same_mask(v.L, h->nc, v.C64, v.sm, h->mstats + 2 * l, h->stream);
        launch_node_list(v.L, v.sm, v.nlcnt, v.nltot, v.nl, h->stream);
        masked[l] = 1;
    }
    // diagonals
    HIPCHK(hipMemsetAsync(h->dflag, 0, sizeof(unsigned long long), h->stream));
    for (auto& v : h->lev)
        launch_diag(v.L, h->kind, v.C64, v.d64, v.d32, h->bc, v.h, (unsigned long long*)h->dflag, h->stream);
    unsigned long long flag = 0;
    std::vector<unsigned long long> ms(2 * nl);
    HIPCHK(hipMemcpyAsync(ms.data(), h->mstats, ms.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          h->stream));
    std::vector<long long> npure(nl, 0);
    for (size_t l = 0; l < nl; ++l)
        if (masked[l])
            HIPCHK(hipMemcpyAsync(&npure[l], h->lev[l].nltot, sizeof(long long), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(&flag, h->dflag, sizeof flag, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t l = 0; l < nl; ++l) {
        Level& v = h->lev[l];
        v.sm_ok = masked[l] != 0;
        v.use_sm = masked[l] && ms[2 * l] > 0 &&
                   (double)ms[2 * l + 1] >= kMaskMinSkip * (double)h->dim * (double)ms[2 * l];
        v.use_list = masked[l] && (double)npure[l] >= kListMinPure * (double)v.nodes;
        v.use_group = masked[l] && h->dim == 3 && (double)npure[l] >= kGroupMinPure * (double)v.nodes;
    }
    if (flag) return fail(ELA_ERR_NAN, "diagonal is nan or zero (Elastic.cpp:347 aborts here)");
    // dense coarsest operator: columns A e_q, then Gauss-Jordan [A | I] -> A^-1 (left half)
    if (!h->M) {  // the coarsest level is solved by BiCGStab: nothing to factor
        h->prepared = true;
        return ELA_OK;
    }
    Level& v = h->lev.back();
    const int n = h->nunk;
    const size_t fe = field_elems(v, h->dim);
    // all columns in one launch when n unit fields fit in 256 MB of scratch
    const bool all = (size_t)n * fe <= ((size_t)1 << 25);
    double *E = nullptr, *F = nullptr;
    int rc = halloc_t(h, all ? (size_t)n * fe : fe, &E);
    if (!rc && !all) rc = halloc_t(h, fe, &F);
    if (rc) {
        if (E) hfree(h, E);
        return rc;
    }
    HIPCHK(hipMemsetAsync(h->M, 0, sizeof(double) * (size_t)n * 2 * n, h->stream));
    if (all) {  // E comes zeroed from halloc
        launch_probe_all(v.L, h->kind, v.nodes, v.C64, E, (long long)fe, h->bc, v.h, h->uniform, h->M, 2LL * n,
                         h->stream);
    } else {
        for (int q = 0; q < n; ++q) {
            launch_unit(v.L, v.nodes, E, q - 1, q, h->stream);
            launch_fapply<double>(v.L, h->kind, v.C64, E, F, h->bc, v.h, h->uniform, h->stream);
            launch_assemble_col(v.L, v.nodes, F, q, 2LL * n, h->M, h->dim, h->stream);
        }
    }
    launch_identity_right(h->M, n, h->stream);
    int gj = 0;
    HIPCHK(hipMemsetAsync(h->dflag, 0, sizeof(int), h->stream));
    // one-launch-a-column scratch: pivots (n doubles), two permutations (2n ints), two candidate columns (2n doubles)
    double* G = nullptr;
    if (h->gj_one && halloc_t(h, 4 * (size_t)n, &G) == ELA_OK)
        launch_gauss_jordan_col(h->M, n, h->dflag, (int*)(G + n), G, G + 2 * (size_t)n, h->stream);
    else
        launch_gauss_jordan(h->M, n, h->dflag, h->stream);
    HIPCHK(hipMemcpyAsync(&gj, h->dflag, sizeof gj, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    hfree(h, E);
    if (F) hfree(h, F);
    if (G) hfree(h, G);
    if (gj) return fail(ELA_ERR_NAN, "coarsest-level operator is singular");
    h->prepared = true;
    return ELA_OK;
}

int ela_field_alloc(ela_hier* h, int mglev, int prec, int ncomp, void** dev) {
    if (int rc = check_level(h, mglev)) return rc;
    if (!dev || ncomp < 1 || ncomp > 64) return fail(ELA_ERR_BADARG, "bad field request");
    const size_t es = prec == ELA_F32 ? sizeof(float) : sizeof(double);
    return halloc(h, field_elems(h->lev[mglev], ncomp) * es, dev);
}

int ela_field_free(ela_hier* h, void* dev) {
    if (!h) return fail(ELA_ERR_BADARG, "null hierarchy");
    (void)hipStreamSynchronize(h->stream);
    hfree(h, dev);
    return ELA_OK;
}

int ela_upload(ela_hier* h, int mglev, int prec, int ncomp, const double* host_dense, void* dev) {
    if (int rc = check_level(h, mglev)) return rc;
    if (!host_dense || !dev) return fail(ELA_ERR_BADARG, "null argument");
    return upload_dense(h, h->lev[mglev], ncomp, host_dense, dev, prec);
}

int ela_download(ela_hier* h, int mglev, int prec, int ncomp, const void* dev, double* host_dense) {
    if (int rc = check_level(h, mglev)) return rc;
    if (!host_dense || !dev) return fail(ELA_ERR_BADARG, "null argument");
    return download_dense(h, h->lev[mglev], ncomp, dev, host_dense, prec);
}

int ela_field_zero(ela_hier* h, int mglev, int prec, int ncomp, void* dev) {
    if (int rc = check_level(h, mglev)) return rc;
    if (!dev || ncomp < 1) return fail(ELA_ERR_BADARG, "bad field");
    const size_t es = prec == ELA_F32 ? sizeof(float) : sizeof(double);
    HIPCHK(hipMemsetAsync(dev, 0, field_elems(h->lev[mglev], ncomp) * es, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return ELA_OK;
}

int ela_fapply(ela_hier* h, int mglev, int prec, const void* u, void* f, void* stream) {
    if (int rc = check_level(h, mglev)) return rc;
    if (int rc = check_prec(h, prec)) return rc;
    if (!h->prepared && mglev > 0) return fail(ELA_ERR_STATE, "call ela_prepare before using coarse levels");
    if (!h->model_set) return fail(ELA_ERR_STATE, "Attempting to use operator before calling SetModel");
    if (!u || !f) return fail(ELA_ERR_BADARG, "null field");
    Level& v = h->lev[mglev];
    if (prec == ELA_F32 && !h->prepared) return fail(ELA_ERR_STATE, "call ela_prepare before FP32 use");
    if (prec == ELA_F64) A_apply<double>(h, v, v.C64, (const double*)u, (double*)f, strm(h, stream));
    else A_apply<float>(h, v, v.C32, (const float*)u, (float*)f, strm(h, stream));
    HIPCHK(hipGetLastError());
    return ELA_OK;
}

int ela_newton_rhs(ela_hier* h, const void* u, const void* b, const void* F0, void* rhs, void* stream) {
    if (!h || !u || !b || !rhs) return fail(ELA_ERR_BADARG, "null argument");
    if (!h->model_set) return fail(ELA_ERR_STATE, "Attempting to use operator before calling SetModel");
    Level& v = h->lev[0];
    if (!h->newton_dw)
        if (int rc = halloc_t(h, field_elems(v, h->dim * h->dim), &h->newton_dw)) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    launch_newton_rhs(v.L, h->kind, v.C64, (const double*)F0, (const double*)u, (const double*)b, h->newton_dw,
                      (double*)rhs, h->bc, v.h, s);
    HIPCHK(hipGetLastError());
    if (!stream) HIPCHK(hipStreamSynchronize(s));
    return ELA_OK;
}

int ela_newton_w_dw(ela_hier* h, const void* u, const void* F0, double wfac, void* w, void* dw, void* stream) {
    if (!h || !u) return fail(ELA_ERR_BADARG, "null argument");
    if (!h->model_set) return fail(ELA_ERR_STATE, "Attempting to use operator before calling SetModel");
    Level& v = h->lev[0];
    if (!v.C64) return fail(ELA_ERR_STATE, "no FP64 stiffness on level 0");
    if (!w && !dw) return ELA_OK;
    launch_newton_w_dw(v.L, h->kind, v.C64, (const double*)F0, (const double*)u, wfac, (double*)w, (double*)dw, v.h,
                       strm(h, stream));
    HIPCHK(hipGetLastError());
    if (!stream) HIPCHK(hipStreamSynchronize(h->stream));
    return ELA_OK;
}

int ela_postprocess(ela_hier* h, int what, int voigt, const void* u, void* out, void* stream) {
    if (!h) return fail(ELA_ERR_BADARG, "null hierarchy");
    if (!h->model_set) return fail(ELA_ERR_STATE, "Attempting to use operator before calling SetModel");
    if (what < 0 || what > 2) return fail(ELA_ERR_BADARG, "what: 0 strain, 1 stress, 2 energy");
    if (!u || !out) return fail(ELA_ERR_BADARG, "null field");
    Level& v = h->lev[0];
    if (!v.C64) return fail(ELA_ERR_STATE, "no FP64 stiffness on level 0");
    launch_post(v.L, h->kind, v.C64, (const double*)u, (double*)out, v.h, what, voigt != 0, strm(h, stream));
    HIPCHK(hipGetLastError());
    return ELA_OK;
}

int ela_normalize(ela_hier* h, int mglev, int prec, void* x, void* stream) {
    if (int rc = check_level(h, mglev)) return rc;
    if (int rc = check_prec(h, prec)) return rc;
    if (!h->prepared) return fail(ELA_ERR_STATE, "Operator::Diagonal() must be called before using normalize");
    if (!x) return fail(ELA_ERR_BADARG, "null field");
    Level& v = h->lev[mglev];
    // the FP64 diagonal (the FP32 smoothers keep only its inverse): the
    // level's, or evaluated into a temporary
    const double* dg = v.d64;
    void* tmp = nullptr;
    hipStream_t s = strm(h, stream);
    if (!dg) {
        if (int rc = halloc(h, field_elems(v, h->dim) * 8, &tmp)) return rc;
        launch_diag(v.L, h->kind, v.C64, (double*)tmp, nullptr, h->bc, v.h, (unsigned long long*)h->dflag, s);
        dg = (const double*)tmp;
    }
    if (prec == ELA_F64) launch_normalize<double>(v.L, h->dim, dg, (double*)x, s);
    else launch_normalize<float>(v.L, h->dim, dg, (float*)x, s);
    HIPCHK(hipGetLastError());
    if (tmp) {
        HIPCHK(hipStreamSynchronize(s));
        hfree(h, tmp);
    }
    return ELA_OK;
}

int ela_diagonal(ela_hier* h, int mglev, int prec, void* diag, void* stream) {
    if (int rc = check_level(h, mglev)) return rc;
    if (!h->prepared) return fail(ELA_ERR_STATE, "call ela_prepare first");
    if (!diag) return fail(ELA_ERR_BADARG, "null field");
    Level& v = h->lev[mglev];
    const size_t fe = field_elems(v, h->dim);
    if (prec == ELA_F64) {
        if (v.d64) HIPCHK(hipMemcpyAsync(diag, v.d64, fe * 8, hipMemcpyDeviceToDevice, strm(h, stream)));
        else launch_diag(v.L, h->kind, v.C64, (double*)diag, nullptr, h->bc, v.h, (unsigned long long*)h->dflag, strm(h, stream));
    } else {
        // the FP64 diagonal rounded to FP32 (the level's FP32 array holds the
        // smoothers' inverse)
        hipStream_t s = strm(h, stream);
        void* tmp = nullptr;
        const double* dg = v.d64;
        if (!dg) {
            if (int rc = halloc(h, fe * 8, &tmp)) return rc;
            launch_diag(v.L, h->kind, v.C64, (double*)tmp, nullptr, h->bc, v.h, (unsigned long long*)h->dflag, s);
            dg = (const double*)tmp;
        }
        launch_convert<double, float>((long long)fe, dg, (float*)diag, s);
        HIPCHK(hipGetLastError());
        if (tmp) {
            HIPCHK(hipStreamSynchronize(s));
            hfree(h, tmp);
        }
    }
    return ELA_OK;
}

int ela_fsmooth(ela_hier* h, int mglev, int prec, void* x, const void* b, int nsweeps, double omega, void* stream) {
    if (int rc = check_level(h, mglev)) return rc;
    if (int rc = check_prec(h, prec)) return rc;
    if (!h->prepared) return fail(ELA_ERR_STATE, "Operator::Diagonal() must be called before using Fsmooth");
    if (!x || !b || nsweeps < 0) return fail(ELA_ERR_BADARG, "bad argument");
    Level& v = h->lev[mglev];
    hipStream_t s = strm(h, stream);
    void* tmp = nullptr;
    const size_t es = prec == ELA_F32 ? 4 : 8;
    if (int rc = halloc(h, field_elems(v, h->dim) * es, &tmp)) return rc;
    for (int w = 0; w < nsweeps; ++w) {
        if (prec == ELA_F64) {
            launch_jacobi<double>(v.L, h->kind, v.C64, v.d64, (const double*)x, (const double*)b, (double*)tmp, h->bc, v.h, h->uniform, omega, s, plane_opt<double>(h, v));  // ela_fsmooth keeps its own omega
            HIPCHK(hipMemcpyAsync(x, tmp, field_elems(v, h->dim) * es, hipMemcpyDeviceToDevice, s));
        } else {
            launch_jacobi<float>(v.L, h->kind, v.C32, v.d32, (const float*)x, (const float*)b, (float*)tmp, h->bc, v.h, h->uniform, omega, s, plane_opt<float>(h, v));
            HIPCHK(hipMemcpyAsync(x, tmp, field_elems(v, h->dim) * es, hipMemcpyDeviceToDevice, s));
        }
    }
    HIPCHK(hipStreamSynchronize(s));
    hfree(h, tmp);
    return ELA_OK;
}

int ela_restriction(ela_hier* h, int cmglev, int prec, const void* fine, void* crse, void* stream) {
    if (int rc = check_level(h, cmglev)) return rc;
    if (cmglev < 1) return fail(ELA_ERR_BADARG, "cmglev must be >= 1");
    if (!fine || !crse) return fail(ELA_ERR_BADARG, "null field");
    Level& c = h->lev[cmglev];
    Level& f = h->lev[cmglev - 1];
    if (prec == ELA_F64) launch_restrict<double>(c.L, f.L, h->dim, (const double*)fine, (double*)crse, strm(h, stream));
    else launch_restrict<float>(c.L, f.L, h->dim, (const float*)fine, (float*)crse, strm(h, stream));
    HIPCHK(hipGetLastError());
    return ELA_OK;
}

int ela_interpolation(ela_hier* h, int fmglev, int prec, void* fine, const void* crse, void* stream) {
    if (int rc = check_level(h, fmglev)) return rc;
    if (fmglev + 1 >= (int)h->lev.size()) return fail(ELA_ERR_BADARG, "no coarser level");
    if (!fine || !crse) return fail(ELA_ERR_BADARG, "null field");
    Level& f = h->lev[fmglev];
    Level& c = h->lev[fmglev + 1];
    if (prec == ELA_F64) launch_prolong<double>(f.L, c.L, h->dim, (const double*)crse, (double*)fine, 1, strm(h, stream));
    else launch_prolong<float>(f.L, c.L, h->dim, (const float*)crse, (float*)fine, 1, strm(h, stream));
    HIPCHK(hipGetLastError());
    return ELA_OK;
}

int ela_residual(ela_hier* h, int mglev, int prec, const void* x, const void* b, void* r, double* norm_inf,
                 void* stream) {
    if (int rc = check_level(h, mglev)) return rc;
    if (int rc = check_prec(h, prec)) return rc;
    if (!x || !b) return fail(ELA_ERR_BADARG, "null field");
    if (!h->prepared && (mglev > 0 || prec == ELA_F32)) return fail(ELA_ERR_STATE, "call ela_prepare first");
    Level& v = h->lev[mglev];
    hipStream_t s = strm(h, stream);
    if (prec == ELA_F64)
        A_residual<double, double>(h, v, v.C64, (const double*)x, (const double*)b, (double*)r, (double*)nullptr, nullptr, h->dnorm, s);
    else
        A_residual<float, float>(h, v, v.C32, (const float*)x, (const float*)b, (float*)r, (float*)nullptr, nullptr, h->dnorm, s);
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, h->dnorm, sizeof bits, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (norm_inf) std::memcpy(norm_inf, &bits, sizeof(double));
    return ELA_OK;
}

int ela_residual_noalloc(ela_hier* h, int mglev, int prec, const void* x, const void* b, void* r, double* norm_inf,
                         void* stream) {
    if (int rc = check_level(h, mglev)) return rc;
    if (int rc = check_prec(h, prec)) return rc;
    if (!x || !b) return fail(ELA_ERR_BADARG, "null field");
    if (!h->prepared && (mglev > 0 || prec == ELA_F32)) return fail(ELA_ERR_STATE, "call ela_prepare first");
    Level& v = h->lev[mglev];
    long long need = residual_partials(v.L);
    if (v.L.dim == 3) need = std::max(need, grain_partials(v.L, level_kz(h, v)));
    if (need > kScratchPartials)
        return fail(ELA_ERR_BADARG, "ela_residual_noalloc: the level needs " + std::to_string(need) +
                                        " block maxima, more than the " + std::to_string(kScratchPartials) +
                                        " of the static scratch (use ela_residual)");
    const int dev = h->ctx->device;
    double* part = (dev >= 0 && dev < kMaxDevices) ? g_scratch[dev] : nullptr;
    if (!part) return fail(ELA_ERR_STATE, "scratch not set up for this device (ela_ctx_create does it)");
    unsigned long long* nw = (unsigned long long*)(part + kScratchPartials);
    hipStream_t s = strm(h, stream);
    if (prec == ELA_F64)
        A_residual<double, double>(h, v, v.C64, (const double*)x, (const double*)b, (double*)r, (double*)nullptr,
                                   nullptr, nw, s, part);
    else
        A_residual<float, float>(h, v, v.C32, (const float*)x, (const float*)b, (float*)r, (float*)nullptr, nullptr,
                                 nw, s, part);
    HIPCHK(hipGetLastError());
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, nw, sizeof bits, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (norm_inf) std::memcpy(norm_inf, &bits, sizeof(double));
    return ELA_OK;
}

int ela_bc_init(ela_hier* h, double* rhs, int homogeneous) {
    if (!h || !rhs) return fail(ELA_ERR_BADARG, "null argument");
    const Level& v = h->lev[0];
    const int hx = v.L.nx - 1, hy = v.L.ny - 1, hz = v.L.nz - 1;
    const long long nn = (long long)v.nodes;
    for (int k = 0; k < v.L.nz; ++k)
        for (int j = 0; j < v.L.ny; ++j)
            for (int i = 0; i < v.L.nx; ++i) {
                int face = -1;
                const bool xl = i == 0, xh = i == hx, yl = j == 0, yh = j == hy;
                if (h->dim == 2) {
                    if (xl && yl) face = 4; else if (xl && yh) face = 5; else if (xh && yl) face = 6;
                    else if (xh && yh) face = 7; else if (xl) face = 0; else if (xh) face = 2;
                    else if (yl) face = 1; else if (yh) face = 3;
                } else {
                    const bool zl = k == 0, zh = k == hz;
                    if ((xl || xh) && (yl || yh) && (zl || zh)) face = 18 + (xh ? 4 : 0) + (yh ? 2 : 0) + (zh ? 1 : 0);
                    else if ((yl || yh) && (zl || zh)) face = 6 + (yh ? 2 : 0) + (zh ? 1 : 0);
                    else if ((zl || zh) && (xl || xh)) face = 10 + (zh ? 2 : 0) + (xh ? 1 : 0);
                    else if ((xl || xh) && (yl || yh)) face = 14 + (xh ? 2 : 0) + (yh ? 1 : 0);
                    else if (xl) face = 0; else if (xh) face = 3; else if (yl) face = 1;
                    else if (yh) face = 4; else if (zl) face = 2; else if (zh) face = 5;
                }
                if (face < 0) continue;
                const long long id = i + (long long)v.L.nx * (j + (long long)v.L.ny * k);
                for (int d = 0; d < h->dim; ++d)
                    rhs[d * nn + id] = (homogeneous && h->bc.t[face][d] == ELA_BC_DISP) ? 0.0 : h->bcval[face][d];
            }
    return ELA_OK;
}

static int solve_common(ela_hier* h, double tol_rel, double tol_abs, int mode, int max_iter, ela_stats* st) {
    if (mode == ELA_SOLVE_FP64 && !h->has64) return fail(ELA_ERR_STATE, "hierarchy has no FP64 levels");
    if (mode == ELA_SOLVE_MIXED && !h->has32) return fail(ELA_ERR_STATE, "hierarchy has no FP32 levels");
    if (mode != ELA_SOLVE_FP64 && mode != ELA_SOLVE_MIXED) return fail(ELA_ERR_BADARG, "unknown solve mode");
    if (max_iter < 0) return fail(ELA_ERR_BADARG, "max_iter < 0");
    if (h->timing && h->ev.empty()) {
        h->ev.resize(kMaxEvents);
        for (auto& e : h->ev) HIPCHK(hipEventCreate(&e));
    }
    h->ev_used = 0;
    h->ev_resid.clear();
    h->ev_smooth.clear();
    auto t0 = std::chrono::steady_clock::now();
    int rc;
    if (h->krylov > 0)
        rc = mode == ELA_SOLVE_FP64 ? solve_fgmres<double>(h, h->b64, h->u64, tol_rel, tol_abs, max_iter, st)
                                    : solve_fgmres<float>(h, h->b64, h->u64, tol_rel, tol_abs, max_iter, st);
    else
        rc = mode == ELA_SOLVE_FP64 ? solve_loop<double>(h, h->b64, h->u64, tol_rel, tol_abs, max_iter, st)
                                    : solve_loop<float>(h, h->b64, h->u64, tol_rel, tol_abs, max_iter, st);
    (void)hipStreamSynchronize(h->stream);
    auto t1 = std::chrono::steady_clock::now();
    collect_times(h);
    if (st) {
        st->mode = mode;
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ELA_ERR_HIP, std::string("solve: ") + hipGetErrorString(e));
    return rc;
}

int ela_solve(ela_hier* h, const double* b_host, double* u_host, double tol_rel, double tol_abs, int mode,
              int max_iter, ela_stats* st) {
    if (!h || !b_host || !u_host) return fail(ELA_ERR_BADARG, "null argument");
    if (st) std::memset(st, 0, sizeof *st);
    auto t0 = std::chrono::steady_clock::now();
    if (!h->prepared)
        if (int rc = ela_prepare(h)) return rc;
    const Level& v = h->lev[0];
    if (int rc = upload_dense(h, v, h->dim, b_host, h->b64, ELA_F64)) return rc;
    if (int rc = upload_dense(h, v, h->dim, u_host, h->u64, ELA_F64)) return rc;
    auto t1 = std::chrono::steady_clock::now();
    int rc = solve_common(h, tol_rel, tol_abs, mode, max_iter, st);
    if (st) st->setup_seconds = std::chrono::duration<double>(t1 - t0).count();
    if (rc == ELA_OK || rc == ELA_ERR_NONCONVERGED) {
        std::string keep = g_err;
        if (int rc2 = download_dense(h, v, h->dim, h->u64, u_host, ELA_F64)) return rc2;
        g_err = keep;
    }
    return rc;
}

int ela_solve_device(ela_hier* h, const void* b_dev, void* u_dev, double tol_rel, double tol_abs, int mode,
                     int max_iter, ela_stats* st) {
    if (!h || !b_dev || !u_dev) return fail(ELA_ERR_BADARG, "null argument");
    if (st) std::memset(st, 0, sizeof *st);
    if (!h->prepared)
        if (int rc = ela_prepare(h)) return rc;
    const size_t bytes = field_elems(h->lev[0], h->dim) * sizeof(double);
    HIPCHK(hipMemcpyAsync(h->b64, b_dev, bytes, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->u64, u_dev, bytes, hipMemcpyDeviceToDevice, h->stream));
    int rc = solve_common(h, tol_rel, tol_abs, mode, max_iter, st);
    HIPCHK(hipMemcpyAsync(u_dev, h->u64, bytes, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return rc;
}

int ela_set_stiffness_mode(ela_hier* h, int mode) {
    if (!h) return fail(ELA_ERR_BADARG, "null hierarchy");
    if (mode != 0 && mode != 1) return fail(ELA_ERR_BADARG, "mode must be 0 (planes) or 1 (grain table)");
    if (mode == 1 && !h->has_table) return fail(ELA_ERR_STATE, "no grain table: set a Voronoi model first");
    h->use_table = mode == 1;
    return ELA_OK;
}

int ela_set_option(ela_hier* h, int option, int value) {
    if (!h) return fail(ELA_ERR_BADARG, "null hierarchy");
    if (option == ELA_OPT_MARCH) {
        h->grain_march = value != 0;
    } else if (option == ELA_OPT_KZ) {
        if (value < 1 || value > 1024) return fail(ELA_ERR_BADARG, "planes per marching tile must be 1..1024");
        h->kz = value;
        for (Level& v : h->lev) v.kz = 0;
    } else if (option == ELA_OPT_XCD) {
        h->xcd = value != 0;
    } else if (option == ELA_OPT_MASK) {
        if (value < 0 || value > 3) return fail(ELA_ERR_BADARG, "mask mode must be 0, 1, 2 or 3");
        h->mask_mode = value;
    } else if (option == ELA_OPT_GRAPH) {
        h->graph_on = value != 0;
    } else if (option == ELA_OPT_FOLD) {
        h->fold = value != 0;
    } else if (option == ELA_OPT_ROWS) {
        if (value != 1 && value != 2) return fail(ELA_ERR_BADARG, "rows per thread must be 1 or 2");
        h->ry = value;
    } else if (option == ELA_OPT_PREFETCH) {
        if (value != 1 && value != 2) return fail(ELA_ERR_BADARG, "planes loaded ahead must be 1 or 2");
        h->pf = value;
    } else if (option == ELA_OPT_GJ) {
        h->gj_one = value != 0;
        h->prepared = false;
    } else if (option == ELA_OPT_GRAIN_DIAG) {
        h->tile_dgt = value != 0;
    } else {
        return fail(ELA_ERR_BADARG, "unknown option");
    }
    return ELA_OK;
}

int ela_set_krylov(ela_hier* h, int restart) {
    if (!h) return fail(ELA_ERR_BADARG, "null hierarchy");
    if (restart < 0 || restart > kMaxKrylov - 1) return fail(ELA_ERR_BADARG, "restart must be in [0, 32]");
    h->krylov = restart;
    return ELA_OK;
}

int ela_set_timing(ela_hier* h, int on) {
    if (!h) return fail(ELA_ERR_BADARG, "null hierarchy");
    h->timing = on != 0;
    return ELA_OK;
}

int ela_kernel_times(const ela_hier* h, double* ms_resid64, int* n_resid64, double* ms_smooth0, int* n_smooth0) {
    if (!h) return fail(ELA_ERR_BADARG, "null hierarchy");
    if (ms_resid64) *ms_resid64 = h->ms_resid;
    if (n_resid64) *n_resid64 = h->n_resid;
    if (ms_smooth0) *ms_smooth0 = h->ms_smooth;
    if (n_smooth0) *n_smooth0 = h->n_smooth;
    return ELA_OK;
}

}  // extern "C"

namespace {
template <typename T>
int autotune_t(ela_hier* h, int nrep, double* ms) {
    Level& v = h->lev[0];
    LB<T> b = lb<T>(v);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const int kzs[3] = {8, 16, 32};
    double best = 1e300;
    int bkz = h->kz, bx = h->xcd ? 1 : 0, bry = h->ry, bpf = h->pf;
    const int nry = sizeof(T) == 4 ? 2 : 1;  // two rows per thread exist for FP32 sweeps
    if (ms)
        for (int q = 0; q < 24; ++q) ms[q] = 0.0;
    const int npf = sizeof(T) == 4 ? 2 : 1;  // two planes ahead: FP32 tiles of one row per thread
    for (int pf = 1; pf <= npf; ++pf)
    for (int ry = 1; ry <= (pf == 2 ? 1 : nry); ++ry)
    for (int a = 0; a < 3; ++a)
        for (int x = 0; x < 2; ++x) {
            h->kz = kzs[a];
            h->xcd = x != 0;
            h->ry = ry;
            h->pf = pf;
            A_jacobi<T>(h, v, b.C, b.dg, *b.cor, *b.res, *b.tmp, h->stream);  // first launch of the shape
            HIPCHK(hipEventRecord(e0, h->stream));
            for (int r = 0; r < nrep; ++r) A_jacobi<T>(h, v, b.C, b.dg, *b.cor, *b.res, *b.tmp, h->stream);
            HIPCHK(hipEventRecord(e1, h->stream));
            HIPCHK(hipEventSynchronize(e1));
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            const double avg = (double)t / nrep;
            if (ms) ms[12 * (pf - 1) + 6 * (ry - 1) + 2 * a + x] = avg;
            if (avg < best) {
                best = avg;
                bkz = kzs[a];
                bx = x;
                bry = ry;
                bpf = pf;
            }
        }
    h->kz = bkz;
    h->xcd = bx != 0;
    h->ry = bry;
    h->pf = bpf;
    // the coarse marching levels: their own tile height (fewer tiles per
    // level, so a lower height can fill the chip where level 0's would not)
    const int ckz[4] = {4, 8, 16, 32};
    for (size_t l = 1; l < h->lev.size(); ++l) {
        Level& c = h->lev[l];
        if (!grain_on(h, c)) continue;
        LB<T> cb = lb<T>(c);
        double cbest = 1e300;
        int ck = 0;
        for (int a = 0; a < 4; ++a) {
            c.kz = ckz[a];
            A_jacobi<T>(h, c, cb.C, cb.dg, *cb.cor, *cb.res, *cb.tmp, h->stream);
            HIPCHK(hipEventRecord(e0, h->stream));
            for (int r = 0; r < nrep; ++r) A_jacobi<T>(h, c, cb.C, cb.dg, *cb.cor, *cb.res, *cb.tmp, h->stream);
            HIPCHK(hipEventRecord(e1, h->stream));
            HIPCHK(hipEventSynchronize(e1));
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            if (t < cbest) {
                cbest = t;
                ck = ckz[a];
            }
        }
        c.kz = ck;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ELA_OK;
}
}  // namespace

extern "C" {

int ela_autotune(ela_hier* h, int nrep, double* ms) {
    if (!h || nrep < 1) return fail(ELA_ERR_BADARG, "bad argument");
    if (!h->prepared) return fail(ELA_ERR_STATE, "call ela_prepare first");
    if (!grain_on(h, h->lev[0])) return fail(ELA_ERR_STATE, "level 0 does not run the marching kernels");
    return h->has32 ? autotune_t<float>(h, nrep, ms) : autotune_t<double>(h, nrep, ms);
}

int ela_time_fapply(ela_hier* h, int prec, int nrep, double* avg_ms) {
    if (!h || !avg_ms || nrep < 1) return fail(ELA_ERR_BADARG, "bad argument");
    if (int rc = check_prec(h, prec)) return rc;
    if (prec == ELA_F32 && !h->prepared) return fail(ELA_ERR_STATE, "call ela_prepare first");
    Level& v = h->lev[0];
    void *u = nullptr, *f = nullptr;
    const size_t es = prec == ELA_F32 ? 4 : 8;
    int rc = halloc(h, field_elems(v, h->dim) * es, &u);
    if (!rc) rc = halloc(h, field_elems(v, h->dim) * es, &f);
    if (rc) return rc;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    auto launch = [&]() {
        if (prec == ELA_F64) A_apply<double>(h, v, v.C64, (const double*)u, (double*)f, h->stream);
        else A_apply<float>(h, v, v.C32, (const float*)u, (float*)f, h->stream);
    };
    launch();
    HIPCHK(hipEventRecord(a, h->stream));
    for (int r = 0; r < nrep; ++r) launch();
    HIPCHK(hipEventRecord(b, h->stream));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    *avg_ms = ms / nrep;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    hfree(h, u);
    hfree(h, f);
    return ELA_OK;
}

int ela_probe_copy(ela_ctx* ctx, size_t bytes, int nrep, double* avg_ms) {
    return ela_probe_copy_ex(ctx, bytes, nrep, 0, avg_ms);
}

int ela_probe_copy_ex(ela_ctx* ctx, size_t bytes, int nrep, int variant, double* avg_ms) {
    if (!ctx || !avg_ms || nrep < 1 || bytes < 16) return fail(ELA_ERR_BADARG, "bad argument");
    if (variant < 0 || variant >= copy_variants()) return fail(ELA_ERR_BADARG, "unknown copy variant");
    bytes = bytes / 16 * 16;
    HIPCHK(hipSetDevice(ctx->device));
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, bytes));
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return fail(ELA_ERR_OOM, "probe allocation failed");
    }
    hipStream_t s;
    HIPCHK(hipStreamCreate(&s));
    HIPCHK(hipMemsetAsync(a, 0, bytes, s));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    launch_copy_variant(variant, a, b, (long long)bytes, s);
    HIPCHK(hipEventRecord(e0, s));
    for (int r = 0; r < nrep; ++r) launch_copy_variant(variant, a, b, (long long)bytes, s);
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / nrep;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    (void)hipFree(a);
    (void)hipFree(b);
    return ELA_OK;
}

int ela_probe_barrier(ela_ctx* ctx, int blocks, int rounds, double* us_per_barrier, double* us_per_launch) {
    if (!ctx || !us_per_barrier || !us_per_launch || blocks < 1 || blocks > 1024 || rounds < 1)
        return fail(ELA_ERR_BADARG, "bad argument");
    HIPCHK(hipSetDevice(ctx->device));
    unsigned* cnt = nullptr;
    int* err = nullptr;
    HIPCHK(hipMalloc(&cnt, 256));
    err = (int*)((char*)cnt + 128);
    hipStream_t s;
    HIPCHK(hipStreamCreate(&s));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto timed = [&](int r, int nlaunch, float& ms) -> int {
        HIPCHK(hipMemsetAsync(cnt, 0, 256, s));
        launch_barrier_probe(blocks, cnt, err, 0, s);  // first launch of the kernel
        HIPCHK(hipEventRecord(e0, s));
        for (int q = 0; q < nlaunch; ++q) launch_barrier_probe(blocks, cnt, err, r, s);
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        return ELA_OK;
    };
    float mb = 0, ml = 0;
    int rc = timed(rounds, 1, mb);
    int bad = 0;
    if (!rc) rc = hipMemcpy(&bad, err, sizeof bad, hipMemcpyDeviceToHost) == hipSuccess ? ELA_OK : fail(ELA_ERR_HIP, "hipMemcpy failed");
    if (!rc) rc = timed(0, 1000, ml);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    (void)hipFree(cnt);
    if (rc) return rc;
    if (bad) return fail(ELA_ERR_STATE, "barrier probe timed out: the blocks were not all resident");
    *us_per_barrier = 1e3 * mb / rounds;
    *us_per_launch = 1e3 * ml / 1000;
    return ELA_OK;
}

}  // extern "C"
