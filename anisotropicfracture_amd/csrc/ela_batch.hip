// ela_batch.hip -- many independent box problems solved together on one GPU.
//
// A member (ela_rve) is a cube of n cells per side on [0,1]^3: a cubic
// Voronoi polycrystal (grain table, optionally times a per-node degradation
// factor), its own boundary-condition types and its own multigrid hierarchy,
// resident on the device from ela_rve_create to ela_rve_destroy.  Every
// kernel of a cycle is ONE launch over the boxes of all members at a level:
// a launch carries an array of box descriptors (BoxDesc) and a block -> box
// map, and each 256-thread block works on 256 consecutive nodes of its box.
//
// Levels are aligned at the coarsest grid: a member of depth D occupies the
// global levels G-D .. G-1 (G = deepest member), so every member's coarsest
// level is global level G-1 and one F-cycle schedule serves all of them.  In
// the schedule, the F-cycle step at global level g (a V-cycle from g) runs on
// the members whose finest level is at or above g; with the members ordered
// by their finest global level, those are a prefix of every level's list.
//
// Per member the arithmetic is that of the single-box solver (ela_solver.hip):
// FP64 outer iteration r = b - A u, u += cycle(r) with the right-hand side
// scaled by a power of two, F-cycles with nu1 = nu2 = 2 calls of 2 weighted
// Jacobi sweeps (omega = 2/3), full weighting of residuals, trilinear
// prolongation, grain-form stiffness on every level (the coarse levels' grain
// ids and factors injected, as a grain-form single box coarsens, so every
// level runs the grain kernels and, when the whole level is in flight, the
// marching tiles) and, on the coarsest level, a dense inverse (up to
// kMaxBatchDense unknowns) or BiCGStab to 1e-4 (larger coarsest levels: sides
// with a large odd factor).
// A member's results depend only on its own data (no cross-member reductions;
// a member's BiCGStab scalars come from its own blocks, and a member that
// stops iterating stays frozen while others go on), and it leaves the active
// set the cycle it converges, so a solve gives the same bits alone and in any
// mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ela.h"
#include "ela_grain.hpp"
#include "ela_internal.hpp"
#include "ela_kernels.hpp"

namespace ela {
namespace {

constexpr int TB = 256;                // nodes per block
constexpr int kMaxBatchDense = 1536;   // unknowns of a dense coarsest level (n_c <= 7)
constexpr double kBottomTol = 1e-4;    // BiCGStab coarsest solve: |r|inf <= 1e-4 |b|inf
constexpr int kBottomMaxIter = 200;    //   or this many iterations
constexpr int kBottomCheck = 4;        // iterations between the host's looks at the members' states
enum { BV_R = 0, BV_RH, BV_P, BV_V, BV_S, BV_T, BV_PH, BV_SH, BV_N };  // BV_PH, BV_SH: D^-1 p, D^-1 s
// Per-member BiCGStab scalars (FP64, on the device).
enum { BS_RHO = 0, BS_RHO1, BS_ALPHA, BS_OMEGA, BS_BETA, BS_TOL, BS_IT, BS_STATE, BS_N = 8 };
constexpr int kNC = kGrainNC;          // cubic stiffness: major/minor symmetric storage
enum { F_RES = 0, F_RC, F_COR, F_TMP, F_HOLD, F_N };

// One box of one launch.  Pointers of the precision that a kernel does not use may be null.
struct BoxDesc {
    Lay L;            // compact: px = nx, pl = nx * ny, cs = nodes
    int tile0;        // first block of this box in launches over its list
    int nodes;
    int nunk;         // coarsest level: unknowns of the dense inverse
    int uniform;      // SetUniform: drop the grad C term (callers use 0)
    // Marching launches over grain-form boxes: interior tiles of MTX x MTY
    // nodes marching kz planes in z, plus the boundary shell in blocks of TB.
    int mt0, ntx, nty, ntz;  // first marching block of the box, tiles per direction
    int kz, st0, nsh, ngr;   // planes per tile, first shell block, shell blocks, grains
    int fp0, nfp;            // first x-face patch block of the box, its patches (face_patch)
    Inv<double> i64;
    Inv<float> i32;
    double h[3];
    BCTab bc;
    // stiffness in grain form on every level: s[o] * tab[gid[o]] (the coarse levels' ids and factors are
    // injected from the next finer level's, as a grain-form single box coarsens)
    const unsigned short* gid;
    const double* tab64;
    const float* tab32;
    const double* sc64;  // per-node factor or null
    const float* sc32;
    double* d64;  // Jacobi diagonal in the cycle precision (one of d64 / d32)
    float* d32;   // FP32: its inverse
    // members without factors: the level's per-grain interior Jacobi diagonal (3 values a grain, the cycle
    // precision, FP32 the inverse), which the marching tiles read instead of d64 / d32; else null
    void* dgt;
    void* f[F_N];  // cycle fields, cycle precision
    double* u64;   // finest level: solution, right-hand side, residual / test output
    double* b64;
    double* r64;
    double* M;      // coarsest level: A^-1, nunk x nunk row-major, dense unknown order
    double* norm;   // member scalars on the device: |r|inf of the last residual, rhs scale
    double* scale;
    const double* E;  // set-up: identity of size nunk (columns = unit fields)
    double* W;        // set-up: [A | I] workspace, nunk x 2 nunk
    double* bs;       // coarsest level solved by BiCGStab: its scalars (BS_*) and vectors (BV_*)
    void* bv[BV_N];
};

template <typename T>
struct P;
template <>
struct P<double> {
    __device__ static const double* tab(const BoxDesc& d) { return d.tab64; }
    __device__ static const double* sc(const BoxDesc& d) { return d.sc64; }
    __device__ static const double* dg(const BoxDesc& d) { return d.d64; }
    __device__ static const double* dgt(const BoxDesc& d) { return (const double*)d.dgt; }
    __device__ static const Inv<double>& inv(const BoxDesc& d) { return d.i64; }
};
template <>
struct P<float> {
    __device__ static const float* tab(const BoxDesc& d) { return d.tab32; }
    __device__ static const float* sc(const BoxDesc& d) { return d.sc32; }
    __device__ static const float* dg(const BoxDesc& d) { return d.d32; }
    __device__ static const float* dgt(const BoxDesc& d) { return (const float*)d.dgt; }
    __device__ static const Inv<float>& inv(const BoxDesc& d) { return d.i32; }
};

// Node of this thread in its box (compact layout: offset = dense index); blk
// is the block's index in the list (launches over a sub-range add an offset).
__device__ __forceinline__ bool box_node(const BoxDesc& d, const Lay& L, int blk, int& i, int& j, int& k, long long& o) {
    const int q = (blk - d.tile0) * TB + (int)threadIdx.x;
    if (q >= d.nodes) return false;
    const int nxy = L.nx * L.ny;
    k = q / nxy;
    const int r = q - k * nxy;
    j = r / L.nx;
    i = r - j * L.nx;
    o = q;
    return true;
}

// f = A u at (i,j,k) in precision T (cubic, grain form).
template <typename T>
__device__ __forceinline__ void box_apply(const BoxDesc& d, const Lay& L, const T* __restrict__ u, int i, int j, int k,
                                          T f[3]) {
    node_apply_acc<T, K_MM, 3>(L, GrainAcc<T, kNC>{d.gid, P<T>::tab(d), P<T>::sc(d)}, u, i, j, k, d.bc, P<T>::inv(d),
                               d.uniform != 0, f);
}

// One weighted-Jacobi sweep out = in + omega (rhs - A in) / D; in < 0: out = omega rhs / D.
template <typename T>
__global__ __launch_bounds__(TB) void kb_smooth(const BoxDesc* __restrict__ D, const int* __restrict__ bmap, int blk0,
                                                int rhs, int in, int out, T omega) {
    const int blk = blk0 + blockIdx.x;
    const BoxDesc& d = D[bmap[blk]];
    const Lay L = d.L;
    int i, j, k;
    long long o;
    if (!box_node(d, L, blk, i, j, k, o)) return;
    const T* b = (const T*)d.f[rhs];
    const T* dg = P<T>::dg(d);
    T* xo = (T*)d.f[out];
    if (in < 0) {
#pragma unroll
        for (int p = 0; p < 3; ++p) xo[p * L.cs + o] = jacobi_first(omega, b[p * L.cs + o], dg[p * L.cs + o]);
        return;
    }
    const T* xi = (const T*)d.f[in];
    T f[3];
    box_apply<T>(d, L, xi, i, j, k, f);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const long long q = p * L.cs + o;
        xo[q] = jacobi_update(xi[q], omega, b[q] - f[p], dg[q]);
    }
}

// r = b - A x (cycle precision).
template <typename T>
__global__ __launch_bounds__(TB) void kb_resid(const BoxDesc* __restrict__ D, const int* __restrict__ bmap, int blk0,
                                               int x, int b, int r) {
    const int blk = blk0 + blockIdx.x;
    const BoxDesc& d = D[bmap[blk]];
    const Lay L = d.L;
    int i, j, k;
    long long o;
    if (!box_node(d, L, blk, i, j, k, o)) return;
    const T* bb = (const T*)d.f[b];
    T* rr = (T*)d.f[r];
    T f[3];
    box_apply<T>(d, L, (const T*)d.f[x], i, j, k, f);
#pragma unroll
    for (int p = 0; p < 3; ++p) rr[p * L.cs + o] = bb[p * L.cs + o] - f[p];
}

// Block max of a non-negative value (NaN propagates) into partials[slot].
__device__ __forceinline__ void block_max_at(double* partials, int slot, double v) {
    __shared__ double wmax[TB / 64];
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o > v || o != o) ? o : v;
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmax[0];
        for (int w = 1; w < TB / 64; ++w) m = (wmax[w] > m || wmax[w] != wmax[w]) ? wmax[w] : m;
        partials[slot] = m;
    }
}

// Block max of a non-negative value (NaN propagates) into partials[blockIdx.x].
__device__ __forceinline__ void block_max(double* partials, double v) {
    __shared__ double wmax[TB / 64];
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o > v || o != o) ? o : v;
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmax[0];
        for (int w = 1; w < TB / 64; ++w) m = (wmax[w] > m || wmax[w] != wmax[w]) ? wmax[w] : m;
        partials[blockIdx.x] = m;
    }
}

// FP64 solution residual of the finest level: r = b - A u; f[res] = r * scale
// (the next cycle's right-hand side) and the block maxima of |r|.  Every lane
// reaches the block reduction.
template <typename T>
__global__ __launch_bounds__(TB) void kb_resid64(const BoxDesc* __restrict__ D, const int* __restrict__ bmap, int res,
                                                 double* __restrict__ partials) {
    const BoxDesc& d = D[bmap[blockIdx.x]];
    const Lay L = d.L;
    int i, j, k;
    long long o;
    double m = 0.0;
    if (box_node(d, L, blockIdx.x, i, j, k, o)) {
        double f[3];
        box_apply<double>(d, L, d.u64, i, j, k, f);
        const double s = *d.scale;
        T* rs = (T*)d.f[res];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const long long q = p * L.cs + o;
            const double v = d.b64[q] - f[p];
            rs[q] = (T)(v * s);
            const double a = fabs(v);
            m = (a > m || a != a) ? a : m;
        }
    }
    block_max(partials, m);
}

// |r|inf of each box from its blocks' partials (fixed order): *norm and out[box].
__global__ __launch_bounds__(64) void kb_fold(const BoxDesc* __restrict__ D, const double* __restrict__ partials,
                                              double* __restrict__ out) {
    const BoxDesc& d = D[blockIdx.x];
    const int nt = (d.nodes + TB - 1) / TB;
    double m = 0.0;
    for (int q = threadIdx.x; q < nt; q += 64) {
        const double v = partials[d.tile0 + q];
        m = (v > m || v != v) ? v : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = (o > m || o != o) ? o : m;
    }
    if (threadIdx.x == 0) {
        *d.norm = m;
        out[blockIdx.x] = m;
    }
}

// scale = 2^-e with 2^e the power of two at or above |r|inf (exact in any precision).
__global__ void kb_scale(const BoxDesc* __restrict__ D, int nbox) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const double r = *D[b].norm;
    int e = 0;
    if (r > 0.0 && r < 1e300) frexp(r, &e);
    *D[b].scale = ldexp(1.0, -e);
}

// Full weighting of f[src] (fine) into f[dst] (coarse); launched over the coarse list.
template <typename T>
__global__ __launch_bounds__(TB) void kb_restrict(const BoxDesc* __restrict__ Dc, const BoxDesc* __restrict__ Df,
                                                  const int* __restrict__ bmap, int src, int dst) {
    const int box = bmap[blockIdx.x];
    const BoxDesc& c = Dc[box];
    const BoxDesc& f = Df[box];
    const Lay Lc = c.L, Lf = f.L;
    int I, J, K;
    long long oc;
    if (!box_node(c, Lc, blockIdx.x, I, J, K, oc)) return;
    const long long px = Lf.px, pl = Lf.pl, of = 2LL * I + px * (2LL * J) + pl * (2LL * K);
    const T* F = (const T*)f.f[src] + of;
    T* out = (T*)c.f[dst];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const T* Fp = F + p * Lf.cs;
        out[p * Lc.cs + oc] =
            restrict_node<T, 3>(Lc, I, J, K, [Fp, px, pl](int a, int b, int c) { return Fp[a + px * b + pl * c]; });
    }
}

// f[dst] (fine) (+)= P f[src] (coarse); launched over the fine list.
template <typename T>
__global__ __launch_bounds__(TB) void kb_prolong(const BoxDesc* __restrict__ Df, const BoxDesc* __restrict__ Dc,
                                                 const int* __restrict__ bmap, int src, int dst, int add) {
    const int box = bmap[blockIdx.x];
    const BoxDesc& f = Df[box];
    const BoxDesc& c = Dc[box];
    const Lay Lf = f.L, Lc = c.L;
    int i, j, k;
    long long o;
    if (!box_node(f, Lf, blockIdx.x, i, j, k, o)) return;
    const long long oc = (long long)(i >> 1) + Lc.px * (j >> 1) + Lc.pl * (k >> 1);
    const T* C = (const T*)c.f[src];
    T* out = (T*)f.f[dst];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const T v = prolong_node<T>(Lc, i, j, k, C + p * Lc.cs + oc);
        T* y = out + p * Lf.cs + o;
        *y = add ? (*y + v) : v;
    }
}

// The same with one thread per coarse node (I, J, K) of the coarse level's
// block list (bmap): its up to 8 fine nodes (2I + a, 2J + b, 2K + c), each
// prolong_node's expression from the thread's coarse corner (the bits of
// kb_prolong), so each coarse value is loaded once for 8 fine nodes.
template <typename T>
__global__ __launch_bounds__(TB) void kb_prolong_oct(const BoxDesc* __restrict__ Df, const BoxDesc* __restrict__ Dc,
                                                     const int* __restrict__ bmap, int src, int dst, int add) {
    const int box = bmap[blockIdx.x];
    const BoxDesc& f = Df[box];
    const BoxDesc& c = Dc[box];
    const Lay Lf = f.L, Lc = c.L;
    int I, J, K;
    long long oc;
    if (!box_node(c, Lc, blockIdx.x, I, J, K, oc)) return;
    const T* C = (const T*)c.f[src];
    T* out = (T*)f.f[dst];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const T* cp = C + p * Lc.cs + oc;
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
                    if (i >= Lf.nx || j >= Lf.ny || k >= Lf.nz) continue;
                    const T v = prolong_node<T>(Lc, i, j, k, cp);
                    T* y = out + p * Lf.cs + i + Lf.px * j + Lf.pl * k;
                    *y = add ? (*y + v) : v;
                }
    }
}

// f[y] += f[x]
template <typename T>
__global__ __launch_bounds__(TB) void kb_axpy(const BoxDesc* __restrict__ D, const int* __restrict__ bmap, int x, int y) {
    const BoxDesc& d = D[bmap[blockIdx.x]];
    const int q = ((int)blockIdx.x - d.tile0) * TB + (int)threadIdx.x;
    if (q >= d.nodes) return;
    const T* xs = (const T*)d.f[x];
    T* ys = (T*)d.f[y];
#pragma unroll
    for (int p = 0; p < 3; ++p) ys[p * d.L.cs + q] += xs[p * d.L.cs + q];
}

// u64 += f[src] / scale (the refinement update, FP64).
template <typename T>
__global__ __launch_bounds__(TB) void kb_update(const BoxDesc* __restrict__ D, const int* __restrict__ bmap, int src) {
    const BoxDesc& d = D[bmap[blockIdx.x]];
    const int q = ((int)blockIdx.x - d.tile0) * TB + (int)threadIdx.x;
    if (q >= d.nodes) return;
    const T* c = (const T*)d.f[src];
    const double s = 1.0 / *d.scale;  // a power of two: exact
#pragma unroll
    for (int p = 0; p < 3; ++p) d.u64[p * d.L.cs + q] += (double)c[p * d.L.cs + q] * s;
}

// Coarsest level: f[dst] = A^-1 f[src].  Grid (row blocks, boxes); each block
// gathers the right-hand side into LDS and each wavefront computes whole rows.
template <typename T>
__global__ __launch_bounds__(TB) void kb_bottom(const BoxDesc* __restrict__ D, int src, int dst) {
    __shared__ double sb[kMaxBatchDense];
    const BoxDesc& d = D[blockIdx.y];
    const int n = d.nunk;
    const int rows_per_block = 4 * (TB / 64);
    const int r0 = blockIdx.x * rows_per_block;
    if (r0 >= n) return;
    const T* b = (const T*)d.f[src];
    for (int q = threadIdx.x; q < n; q += TB) sb[q] = (double)b[q];  // compact layout: dense unknown order
    __syncthreads();
    T* x = (T*)d.f[dst];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int r = r0 + w; r < min(n, r0 + rows_per_block); r += TB / 64) {
        const double* row = d.M + (long long)r * n;
        double acc = 0.0;
        for (int q = lane; q < n; q += 64) acc += row[q] * sb[q];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) x[r] = (T)acc;
    }
}

// ---------------------------------------------------------- BiCGStab coarsest solve
// The coarsest levels beyond the dense inverse solve A x = b by BiCGStab (the
// single-box solver's bottom, the reference's bottom solver: Linear.H:23-29
// with MLMG's bottom tolerance 1e-4), right-preconditioned by the level's
// Jacobi diagonal D: A is applied to D^-1 p and D^-1 s and x gathers those, as
// the reference's operators scale them through their normalize hook
// (Operator.cpp:164-195).  Vectors in the cycle precision, dot products and
// norms in FP64.  Launches go over the list of such boxes, five
// per iteration.  Each block writes three partials (two sums and a maximum);
// the next launch that needs the scalar they define folds them again in
// every block of the member (its own blocks' partials, in a fixed order, so
// every block gets the same bits) and applies the scalar step.  The scalars
// (BS_*) are double-buffered: a launch reads set ph and the member's first
// block writes set ph ^ 1, so no block reads what another block of the same
// launch writes; the two partial buffers alternate the same way.  A member's
// state is 1 while it iterates, 2 when it ends after x += alpha p, 0 when
// done; the vector work skips members whose state says so, so a member stops
// where the single-box loop breaks and stays frozen while others go on.

__device__ __forceinline__ double amax(double m, double v) {
    const double a = fabs(v);
    return (a > m || a != a) ? a : m;
}
__device__ __forceinline__ double nmax(double a, double b) { return (b > a || b != b) ? b : a; }

// Sums of a and b and maximum of c over the block into partials[3 blk + 0..2].
__device__ __forceinline__ void block_sum2_max(double* partials, int blk, double a, double b, double c) {
    __shared__ double wred[3][TB / 64];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c = nmax(c, __shfl_xor(c, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        wred[0][threadIdx.x >> 6] = a;
        wred[1][threadIdx.x >> 6] = b;
        wred[2][threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = wred[0][0], sb = wred[1][0], m = wred[2][0];
        for (int w = 1; w < TB / 64; ++w) {
            sa += wred[0][w];
            sb += wred[1][w];
            m = nmax(m, wred[2][w]);
        }
        partials[3 * blk] = sa;
        partials[3 * blk + 1] = sb;
        partials[3 * blk + 2] = m;
    }
}

// Fold of a member's partials (sums of slots 0, 1, max of slot 2) into q on
// every thread of the block; the same order in every block.
__device__ __forceinline__ void fold_member(const BoxDesc& d, const double* __restrict__ P, double q[3]) {
    __shared__ double fr[3][TB / 64];
    const int nt = (d.nodes + TB - 1) / TB;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int t = threadIdx.x; t < nt; t += TB) {
        const double* pp = P + 3LL * (d.tile0 + t);
        a += pp[0];
        b += pp[1];
        c = nmax(c, pp[2]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c = nmax(c, __shfl_xor(c, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        fr[0][threadIdx.x >> 6] = a;
        fr[1][threadIdx.x >> 6] = b;
        fr[2][threadIdx.x >> 6] = c;
    }
    __syncthreads();
    q[0] = fr[0][0];
    q[1] = fr[1][0];
    q[2] = fr[2][0];
    for (int w = 1; w < TB / 64; ++w) {
        q[0] += fr[0][w];
        q[1] += fr[1][w];
        q[2] = nmax(q[2], fr[2][w]);
    }
    __syncthreads();  // fr is reused by a later fold of the same kernel
}

// The member's scalars of set ph (every thread holds a copy), and the write of
// set ph ^ 1 by the member's first block (state also to state[box]).
struct BScal {
    double v[BS_N];
};
__device__ __forceinline__ BScal scal_load(const BoxDesc& d, int ph) {
    BScal S;
#pragma unroll
    for (int k = 0; k < BS_N; ++k) S.v[k] = d.bs[ph * BS_N + k];
    return S;
}
__device__ __forceinline__ void scal_store(const BoxDesc& d, int blk, int ph, const BScal& S, int* state, int box) {
    if (blk != d.tile0 || threadIdx.x != 0) return;
#pragma unroll
    for (int k = 0; k < BS_N; ++k) d.bs[(ph ^ 1) * BS_N + k] = S.v[k];
    if (state) state[box] = (int)S.v[BS_STATE];
}

// z = D^-1 y at element e of the level's layout (FP32 levels hold the inverse diagonal).
template <typename T>
__device__ __forceinline__ T dinv_apply(const BoxDesc& d, long long e, T y) {
    if constexpr (sizeof(T) == 4) return y * P<T>::dg(d)[e];
    else return y / P<T>::dg(d)[e];
}

// x = f[dst] = 0, p = v = 0, r = rh = f[src]; partials rh.r and |r|inf into P.
template <typename T>
__global__ __launch_bounds__(TB) void kbb_init(const BoxDesc* __restrict__ D, const int* __restrict__ bmap, int src,
                                               int dst, double* __restrict__ P) {
    const BoxDesc& d = D[bmap[blockIdx.x]];
    const int q = ((int)blockIdx.x - d.tile0) * TB + (int)threadIdx.x;
    double s0 = 0.0, m = 0.0;
    if (q < d.nodes) {
        const T* b = (const T*)d.f[src];
        T* x = (T*)d.f[dst];
        T *r = (T*)d.bv[BV_R], *rh = (T*)d.bv[BV_RH], *p = (T*)d.bv[BV_P], *v = (T*)d.bv[BV_V];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long e = c * d.L.cs + q;
            const T bv = b[e];
            x[e] = p[e] = v[e] = (T)0;
            r[e] = rh[e] = bv;
            s0 += (double)bv * (double)bv;
            m = amax(m, (double)bv);
        }
    }
    block_sum2_max(P, blockIdx.x, s0, 0.0, m);
}

// Iteration start.  first: rho = rh.r and the tolerance from |b|inf (a zero
// right-hand side gives x = 0); else rho = rh.r of the last update, stop on
// |r| small, the iteration limit or a breakdown of rho, beta.  Then
// p = r + beta (p - omega v) and p^ = D^-1 p.
template <typename T>
__global__ __launch_bounds__(TB) void kbb_p(const BoxDesc* __restrict__ D, const int* __restrict__ bmap,
                                            const double* __restrict__ P, int ph, int first, int* state) {
    const int box = bmap[blockIdx.x];
    const BoxDesc& d = D[box];
    BScal S = scal_load(d, ph);
    if (!first && S.v[BS_STATE] != 1.0) {
        scal_store(d, blockIdx.x, ph, S, state, box);
        return;
    }
    double q[3];
    fold_member(d, P, q);
    if (first) {
        S.v[BS_RHO] = q[0];
        S.v[BS_RHO1] = S.v[BS_ALPHA] = S.v[BS_OMEGA] = 1.0;
        S.v[BS_BETA] = 0.0;
        S.v[BS_IT] = 0.0;
        S.v[BS_TOL] = kBottomTol * q[2];
        S.v[BS_STATE] = (q[2] > 0.0 && q[0] != 0.0 && isfinite(q[0])) ? 1.0 : 0.0;
    } else {
        S.v[BS_RHO1] = S.v[BS_RHO];
        S.v[BS_RHO] = q[0];
        S.v[BS_IT] += 1.0;
        if (q[2] <= S.v[BS_TOL] || !isfinite(q[2]) || S.v[BS_IT] >= (double)kBottomMaxIter || !(q[0] != 0.0) ||
            !isfinite(q[0]))
            S.v[BS_STATE] = 0.0;
        else
            S.v[BS_BETA] = (S.v[BS_RHO] / S.v[BS_RHO1]) * (S.v[BS_ALPHA] / S.v[BS_OMEGA]);
    }
    scal_store(d, blockIdx.x, ph, S, state, box);
    if (S.v[BS_STATE] != 1.0) return;
    const int qn = ((int)blockIdx.x - d.tile0) * TB + (int)threadIdx.x;
    if (qn >= d.nodes) return;
    const T beta = (T)S.v[BS_BETA], omega = (T)S.v[BS_OMEGA];
    const T *r = (const T*)d.bv[BV_R], *v = (const T*)d.bv[BV_V];
    T* p = (T*)d.bv[BV_P];
    T* ph_ = (T*)d.bv[BV_PH];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const long long e = c * d.L.cs + qn;
        const T pe = r[e] + beta * (p[e] - omega * v[e]);
        p[e] = pe;
        ph_[e] = dinv_apply<T>(d, e, pe);
    }
}

// v = A p^, partial rh.v into Pout.
template <typename T>
__global__ __launch_bounds__(TB) void kbb_av(const BoxDesc* __restrict__ D, const int* __restrict__ bmap,
                                             double* __restrict__ Pout, int ph) {
    const int box = bmap[blockIdx.x];
    const BoxDesc& d = D[box];
    const BScal S = scal_load(d, ph);
    scal_store(d, blockIdx.x, ph, S, nullptr, box);
    if (S.v[BS_STATE] != 1.0) return;  // uniform over the block
    const Lay L = d.L;
    int i, j, k;
    long long o;
    double s0 = 0.0;
    if (box_node(d, L, blockIdx.x, i, j, k, o)) {
        const T* in = (const T*)d.bv[BV_PH];
        const T* rh = (const T*)d.bv[BV_RH];
        T* out = (T*)d.bv[BV_V];
        T f[3];
        box_apply<T>(d, L, in, i, j, k, f);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long e = c * L.cs + o;
            out[e] = f[c];
            s0 += (double)rh[e] * (double)f[c];
        }
    }
    block_sum2_max(Pout, blockIdx.x, s0, 0.0, 0.0);
}

// alpha = rho / rh.v (a breakdown ends the member without an update);
// s = r - alpha v and s^ = D^-1 s, partial |s|inf into Pout.
template <typename T>
__global__ __launch_bounds__(TB) void kbb_s(const BoxDesc* __restrict__ D, const int* __restrict__ bmap,
                                            const double* __restrict__ Pin, double* __restrict__ Pout, int ph) {
    const int box = bmap[blockIdx.x];
    const BoxDesc& d = D[box];
    BScal S = scal_load(d, ph);
    if (S.v[BS_STATE] != 1.0) {
        scal_store(d, blockIdx.x, ph, S, nullptr, box);
        return;
    }
    double q[3];
    fold_member(d, Pin, q);
    if (!(q[0] != 0.0) || !isfinite(q[0])) S.v[BS_STATE] = 0.0;
    else S.v[BS_ALPHA] = S.v[BS_RHO] / q[0];
    scal_store(d, blockIdx.x, ph, S, nullptr, box);
    if (S.v[BS_STATE] != 1.0) return;
    const int qn = ((int)blockIdx.x - d.tile0) * TB + (int)threadIdx.x;
    double m = 0.0;
    if (qn < d.nodes) {
        const T alpha = (T)S.v[BS_ALPHA];
        const T *r = (const T*)d.bv[BV_R], *v = (const T*)d.bv[BV_V];
        T* sv = (T*)d.bv[BV_S];
        T* sh_ = (T*)d.bv[BV_SH];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long e = c * d.L.cs + qn;
            const T x = r[e] - alpha * v[e];
            sv[e] = x;
            sh_[e] = dinv_apply<T>(d, e, x);
            m = amax(m, (double)x);
        }
    }
    block_sum2_max(Pout, blockIdx.x, 0.0, 0.0, m);
}

// |s| small: the member ends after x += alpha p^ (state 2); else t = A s^,
// partials t.t and t.s into Pout.
template <typename T>
__global__ __launch_bounds__(TB) void kbb_at(const BoxDesc* __restrict__ D, const int* __restrict__ bmap,
                                             const double* __restrict__ Pin, double* __restrict__ Pout, int ph) {
    const int box = bmap[blockIdx.x];
    const BoxDesc& d = D[box];
    BScal S = scal_load(d, ph);
    if (S.v[BS_STATE] != 1.0) {
        scal_store(d, blockIdx.x, ph, S, nullptr, box);
        return;
    }
    double q[3];
    fold_member(d, Pin, q);
    if (q[2] <= S.v[BS_TOL]) {
        S.v[BS_STATE] = 2.0;
        S.v[BS_IT] += 1.0;
    }
    scal_store(d, blockIdx.x, ph, S, nullptr, box);
    if (S.v[BS_STATE] != 1.0) return;
    const Lay L = d.L;
    int i, j, k;
    long long o;
    double s0 = 0.0, s1 = 0.0;
    if (box_node(d, L, blockIdx.x, i, j, k, o)) {
        const T* in = (const T*)d.bv[BV_SH];
        const T* sv = (const T*)d.bv[BV_S];
        T* out = (T*)d.bv[BV_T];
        T f[3];
        box_apply<T>(d, L, in, i, j, k, f);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long e = c * L.cs + o;
            out[e] = f[c];
            s0 += (double)f[c] * (double)f[c];
            s1 += (double)f[c] * (double)sv[e];
        }
    }
    block_sum2_max(Pout, blockIdx.x, s0, s1, 0.0);
}

// State 2: x += alpha p^, done.  State 1: omega = t.s / t.t (a breakdown: x +=
// alpha p^, done), x += alpha p^ + omega s^, r = s - omega t, partials rh.r
// and |r|inf into Pout.
template <typename T>
__global__ __launch_bounds__(TB) void kbb_x(const BoxDesc* __restrict__ D, const int* __restrict__ bmap, int dst,
                                            const double* __restrict__ Pin, double* __restrict__ Pout, int ph) {
    const int box = bmap[blockIdx.x];
    const BoxDesc& d = D[box];
    BScal S = scal_load(d, ph);
    const double st = S.v[BS_STATE];
    if (st != 1.0 && st != 2.0) {
        scal_store(d, blockIdx.x, ph, S, nullptr, box);
        return;
    }
    bool only_p = st == 2.0;
    if (!only_p) {
        double q[3];
        fold_member(d, Pin, q);
        const double om = q[0] > 0.0 ? q[1] / q[0] : 0.0;
        if (!(om != 0.0) || !isfinite(om)) {
            only_p = true;
            S.v[BS_IT] += 1.0;
        } else {
            S.v[BS_OMEGA] = om;
        }
    }
    if (only_p) S.v[BS_STATE] = 0.0;
    scal_store(d, blockIdx.x, ph, S, nullptr, box);
    const int qn = ((int)blockIdx.x - d.tile0) * TB + (int)threadIdx.x;
    const T alpha = (T)S.v[BS_ALPHA], omega = (T)S.v[BS_OMEGA];
    T* x = (T*)d.f[dst];
    const T* p = (const T*)d.bv[BV_PH];
    if (only_p) {
        if (qn < d.nodes)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const long long e = c * d.L.cs + qn;
                x[e] += alpha * p[e];
            }
        return;
    }
    double s0 = 0.0, m = 0.0;
    if (qn < d.nodes) {
        const T *sv = (const T*)d.bv[BV_S], *t = (const T*)d.bv[BV_T], *rh = (const T*)d.bv[BV_RH];
        const T* sh_ = (const T*)d.bv[BV_SH];
        T* r = (T*)d.bv[BV_R];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long e = c * d.L.cs + qn;
            x[e] += alpha * p[e] + omega * sh_[e];
            const T rr = sv[e] - omega * t[e];
            r[e] = rr;
            s0 += (double)rh[e] * (double)rr;
            m = amax(m, (double)rr);
        }
    }
    block_sum2_max(Pout, blockIdx.x, s0, 0.0, m);
}

// f[dst] = A f[src] (cycle precision), and the FP64 finest-level r64 = A u64.
template <typename T>
__global__ __launch_bounds__(TB) void kb_apply(const BoxDesc* __restrict__ D, const int* __restrict__ bmap, int src,
                                               int dst) {
    const BoxDesc& d = D[bmap[blockIdx.x]];
    const Lay L = d.L;
    int i, j, k;
    long long o;
    if (!box_node(d, L, blockIdx.x, i, j, k, o)) return;
    T f[3];
    box_apply<T>(d, L, (const T*)d.f[src], i, j, k, f);
    T* out = (T*)d.f[dst];
#pragma unroll
    for (int p = 0; p < 3; ++p) out[p * L.cs + o] = f[p];
}

__global__ __launch_bounds__(TB) void kb_apply64(const BoxDesc* __restrict__ D, const int* __restrict__ bmap) {
    const BoxDesc& d = D[bmap[blockIdx.x]];
    const Lay L = d.L;
    int i, j, k;
    long long o;
    if (!box_node(d, L, blockIdx.x, i, j, k, o)) return;
    double f[3];
    box_apply<double>(d, L, d.u64, i, j, k, f);
#pragma unroll
    for (int p = 0; p < 3; ++p) d.r64[p * L.cs + o] = f[p];
}

// f[dst] = (T) u64 and r64 = (double) f[src] (test conversions on the finest level).
template <typename T>
__global__ __launch_bounds__(TB) void kb_cvt(const BoxDesc* __restrict__ D, const int* __restrict__ bmap, int dst,
                                             int src) {
    const BoxDesc& d = D[bmap[blockIdx.x]];
    const int q = ((int)blockIdx.x - d.tile0) * TB + (int)threadIdx.x;
    if (q >= d.nodes) return;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const long long e = p * d.L.cs + q;
        if (dst >= 0) ((T*)d.f[dst])[e] = (T)d.u64[e];
        if (src >= 0) d.r64[e] = (double)((const T*)d.f[src])[e];
    }
}

// ------------------------------------------------------------------ marching kernels (grain form)
// Operands of a marching / shell launch over a list: field indices (ela_grain.hpp
// MODE 0/1), or the member's u64 / b64 / f[out] / scale (MODE 2), and the
// partials of MODE 2 (marching blocks first, then the shell blocks).
struct MOps {
    int rhs, in, out;
    double omega;
    double* partials;
    int mtot;  // marching blocks of the launch
    int xcd;   // XCD-contiguous tile order (else hardware order)
    int gdiag = 0;  // Jacobi sweeps: the tiles take the interior diagonal per grain (BoxDesc::dgt)
};

template <typename T, int MODE>
__device__ __forceinline__ GrainBox<T> grain_of(const BoxDesc& d, const MOps& op) {
    GrainBox<T> g;
    g.L = d.L;
    g.iv = P<T>::inv(d);
    g.bc = &d.bc;
    g.uniform = d.uniform;
    g.gid = d.gid;
    g.tab = P<T>::tab(d);
    g.sc = P<T>::sc(d);
    g.ngr = d.ngr;
    g.ntx = d.ntx;
    g.nty = d.nty;
    g.kz = d.kz;
    g.xcd = op.xcd;
    g.ry = 1;
    g.pf = 1;
    g.nt = false;
    g.xface = true;  // the x faces run in face patches (kf_grain)
    g.dgt = MODE == 0 && op.gdiag ? P<T>::dgt(d) : nullptr;  // a sweep of the cycle precision
    g.in = MODE == 2 ? (const T*)d.u64 : (const T*)d.f[op.in];
    g.rhs = MODE == 2 ? (const T*)d.b64 : (const T*)d.f[op.rhs];
    g.out = d.f[op.out];
    g.out2 = nullptr;
    g.dg = P<T>::dg(d);
    g.scale = d.scale;
    g.omega = (T)op.omega;
    return g;
}

// Shell block b of a grain-form launch, reading the descriptor in place (a
// local copy of the operands would be indexed per lane and spill to scratch).
template <typename T, int MODE, typename TO>
__device__ __forceinline__ void shell_block(const BoxDesc* __restrict__ D, const int* __restrict__ bmap, int b,
                                            const MOps& op) {
    const BoxDesc& d = D[bmap[b]];
    const Lay L = d.L;
    int q = (b - d.st0) * TB + (int)threadIdx.x;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int A = nx * ny, Bf = nx * (nz - 2), Cf = (ny - 2) * (nz - 2);
    const bool in = q < 2 * A + 2 * Bf;  // the x faces (2 Cf nodes) run in face patches
    int i = 0, j = 0, k = 0;
    if (in) {
        if (q < 2 * A) {
            k = q < A ? 0 : nz - 1;
            const int r = q % A;
            j = r / nx;
            i = r - j * nx;
        } else if ((q -= 2 * A) < 2 * Bf) {
            j = q < Bf ? 0 : ny - 1;
            const int r = q % Bf;
            k = 1 + r / nx;
            i = r - (k - 1) * nx;
        } else {
            q -= 2 * Bf;
            i = q < Cf ? 0 : nx - 1;
            const int r = q % Cf;
            k = 1 + r / (ny - 2);
            j = 1 + r - (k - 1) * (ny - 2);
        }
    }
    double m = 0.0;
    if (in) {
        const T* u = MODE == 2 ? (const T*)d.u64 : (const T*)d.f[op.in];
        const T* rhs = MODE == 2 ? (const T*)d.b64 : (const T*)d.f[op.rhs];
        T f[3];
        node_apply_acc<T, K_MM, 3>(L, GrainAcc<T, kNC>{d.gid, P<T>::tab(d), P<T>::sc(d)}, u, i, j, k, d.bc,
                                   P<T>::inv(d), d.uniform != 0, f);
        const long long o = (long long)i + L.px * j + L.pl * k;
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const long long e = p * L.cs + o;
            if (MODE == 0) {
                // P<float>::dg holds the inverse diagonal: jacobi_update multiplies in FP32 and divides in FP64
                ((T*)d.f[op.out])[e] = jacobi_update(u[e], (T)op.omega, rhs[e] - f[p], P<T>::dg(d)[e]);
            } else if (MODE == 1) {
                ((T*)d.f[op.out])[e] = rhs[e] - f[p];
            } else {
                const double v = (double)rhs[e] - (double)f[p];
                ((TO*)d.f[op.out])[e] = (TO)(v * *d.scale);
                const double a = fabs(v);
                m = (a > m || a != a) ? a : m;
            }
        }
    }
    if (MODE == 2) block_max(op.partials + op.mtot, m);
}

// One launch per grain-form level: blocks [0, nsp) compute the shells' z and
// y faces (the last nsp - ns are padding), blocks [nsp, nsp + nfp) the x
// faces in LDS patches (face_patch: a face's nodes lie a row apart, so the
// patch stages the face and its neighbour column once; the last nfp - nf are
// padding), the others march the tiles.  The shells run beside the tiles
// instead of after them; nsp and nfp are multiples of 8, so a tile's hardware
// block lands on the XCD that xcd_block assigns it.
template <typename T, int MODE, typename TO, bool SC>
__global__ __launch_bounds__(MTX* MTY) void kf_grain(const BoxDesc* __restrict__ D, const int* __restrict__ bmap_m,
                                                     const int* __restrict__ bmap_s, const int* __restrict__ bmap_f,
                                                     int ns, int nsp, int nf, int nfp, MOps op) {
    const int b = (int)blockIdx.x;
    if (b < nsp) {
        if (b < ns) shell_block<T, MODE, TO>(D, bmap_s, b, op);
        return;
    }
    __shared__ MarchLds<T, SC> sh;
    if (b < nsp + nfp) {
        double m = 0.0;
        const int pf = b - nsp;
        if (pf < nf) {
            const BoxDesc& d = D[bmap_f[pf]];
            m = face_patch<T, MODE, TO>(grain_of<T, MODE>(d, op), pf - d.fp0, sh.su);
        }
        if (MODE == 2) block_max(op.partials + op.mtot, m);
        return;
    }
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    const int nm = (int)gridDim.x - nsp - nfp, bm = b - nsp - nfp;
    const int lb = op.xcd ? xcd_block(bm, nm) : bm;
    const BoxDesc& d = D[bmap_m[lb]];
    const double m = march_tile<T, MODE, TO, SC>(grain_of<T, MODE>(d, op), lb - d.mt0, sh, (T*)dyn_lds);
    if (MODE == 2) block_max_at(op.partials, lb, m);
}

// |r|inf of each box from the partials of its marching, shell and x-face
// patch blocks (the shells' slots follow the tiles', the patches' the padded
// shells').
__global__ __launch_bounds__(64) void kb_fold_m(const BoxDesc* __restrict__ D, const double* __restrict__ partials,
                                                int mtot, int nsp, double* __restrict__ out) {
    const BoxDesc& d = D[blockIdx.x];
    const int nm = d.ntx * d.nty * d.ntz;
    double m = 0.0;
    for (int q = threadIdx.x; q < nm + d.nsh + d.nfp; q += 64) {
        const double v = q < nm ? partials[d.mt0 + q]
                         : q < nm + d.nsh ? partials[mtot + d.st0 + q - nm]
                                          : partials[mtot + nsp + d.fp0 + q - nm - d.nsh];
        m = (v > m || v != v) ? v : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = (o > m || o != o) ? o : m;
    }
    if (threadIdx.x == 0) {
        *d.norm = m;
        out[blockIdx.x] = m;
    }
}

// Coarse level of a member in grain form: each coarse node takes the grain id
// (and factor) of the coincident fine node, as a grain-form single box
// coarsens (Elastic.cpp:759-803 averages per-node stiffness instead).
__global__ __launch_bounds__(TB) void kb_inject(const BoxDesc* __restrict__ Dc, const BoxDesc* __restrict__ Df,
                                               const int* __restrict__ bmap) {
    const int box = bmap[blockIdx.x];
    const BoxDesc& c = Dc[box];
    const BoxDesc& f = Df[box];
    const Lay Lc = c.L, Lf = f.L;
    int I, J, K;
    long long oc;
    if (!box_node(c, Lc, blockIdx.x, I, J, K, oc)) return;
    const long long of = 2LL * I + Lf.px * (2LL * J) + Lf.pl * (2LL * K);
    ((unsigned short*)c.gid)[oc] = f.gid[of];
    if (c.sc64) ((double*)c.sc64)[oc] = f.sc64[of];
    if (c.sc32) ((float*)c.sc32)[oc] = f.sc32[of];
}

// Per-grain interior Jacobi diagonal of each box of the list that keeps one
// (BoxDesc::dgt; grain_diag, as the single box's tables), in the cycle
// precision.  Grid (boxes), one thread a grain.
__global__ __launch_bounds__(256) void kb_grain_diag(const BoxDesc* __restrict__ D, int mixed) {
    const BoxDesc& d = D[blockIdx.x];
    if (!d.dgt) return;
    for (int gi = threadIdx.x; gi < d.ngr; gi += blockDim.x)
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const double dg = grain_diag(d.tab64 + (long long)kNC * gi, d.h, p);
            if (mixed) store_jacobi_diag((float*)d.dgt + 3 * gi + p, dg);  // the inverse
            else store_jacobi_diag((double*)d.dgt + 3 * gi + p, dg);
        }
}

// Jacobi diagonal (Elastic.cpp:264-352) from the FP64 stiffness, stored in the cycle precision.
__global__ __launch_bounds__(TB) void kb_diag(const BoxDesc* __restrict__ D, const int* __restrict__ bmap, int* flag) {
    const BoxDesc& d = D[bmap[blockIdx.x]];
    const Lay L = d.L;
    int i, j, k;
    long long o;
    if (!box_node(d, L, blockIdx.x, i, j, k, o)) return;
    double dg[3];
    node_diag_acc<K_MM, 3>(L, GrainAcc<double, kNC>{d.gid, d.tab64, d.sc64}, i, j, k, d.bc, d.h, dg);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        if (d.d64) d.d64[p * L.cs + o] = dg[p];
        if (d.d32) store_jacobi_diag(d.d32 + p * L.cs + o, dg[p]);  // the inverse
        if (!(dg[p] == dg[p]) || dg[p] == 0.0) *flag = 1;
    }
}

// Identity of size n (row-major): the unit fields of a coarsest layout.
__global__ void kb_identity(double* E, int n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (long long)n * n) E[t] = (t / n == t % n) ? 1.0 : 0.0;
}

// Dense coarsest operator by probing: block (column q, node block) of box y
// applies A to the unit field e_q and writes W[r][q] = (A e_q)_r, plus the
// identity in W's right half.  Grid (maxnunk * nb, boxes).
__global__ __launch_bounds__(TB) void kb_probe(const BoxDesc* __restrict__ D, int nb) {
    const BoxDesc& d = D[blockIdx.y];
    const int q = blockIdx.x / nb, t = (blockIdx.x % nb) * TB + threadIdx.x;
    if (q >= d.nunk || t >= d.nodes) return;
    const Lay L = d.L;
    const int nxy = L.nx * L.ny;
    const int k = t / nxy, j = (t - k * nxy) / L.nx, i = t - k * nxy - j * L.nx;
    double f[3];
    box_apply<double>(d, L, d.E + (long long)q * d.nunk, i, j, k, f);
    const long long w = 2LL * d.nunk;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const long long r = (long long)p * d.nodes + t;
        d.W[r * w + q] = f[p];
        if (r == q) d.W[r * w + d.nunk + r] = 1.0;
    }
}

__global__ __launch_bounds__(1024) void kb_gj_pivot(const BoxDesc* __restrict__ D, int c, int* flag) {
    __shared__ double sval[1024];
    __shared__ int sidx[1024];
    const BoxDesc& d = D[blockIdx.x];
    if (c >= d.nunk) return;
    gj_pivot_block(d.W, d.nunk, c, flag, sval, sidx);
}

__global__ __launch_bounds__(256) void kb_gj_eliminate(const BoxDesc* __restrict__ D, int c) {
    const BoxDesc& d = D[blockIdx.y];
    if (c >= d.nunk || (int)blockIdx.x >= d.nunk) return;
    gj_eliminate_row(d.W, d.nunk, c, blockIdx.x);
}

// Each member's Gauss-Jordan on its [A | I] in one launch a column
// (gj_column: the pivot/eliminate pairs' pivots and bits), its scratch after
// W: pivots (n doubles), two permutations (2n ints), two candidate columns (2n
// doubles).  Grid (maxnunk rows, boxes); block r updates row r of its member.
__global__ __launch_bounds__(256) void kb_gj_col(const BoxDesc* __restrict__ D, int c, int* flag) {
    __shared__ double sval[256];
    __shared__ int sidx[256];
    const BoxDesc& d = D[blockIdx.y];
    const int n = d.nunk;
    if (c >= n || (int)blockIdx.x >= n) return;
    double* dpiv = d.W + 2LL * n * n;
    int* perm = (int*)(dpiv + n);
    double* cand = dpiv + 2 * (size_t)n;
    gj_column(d.W, n, c, c ? perm + (size_t)((c - 1) & 1) * n : nullptr, perm + (size_t)(c & 1) * n, dpiv,
              c ? cand + (size_t)((c - 1) & 1) * n : nullptr, cand + (size_t)(c & 1) * n, flag, blockIdx.x, sval, sidx);
}

// After kb_gj_col's last column: row i of A^-1 (the right half of W row
// perm[i], the last pivot row over its pivot) into M.
__global__ __launch_bounds__(256) void kb_gj_final(const BoxDesc* __restrict__ D) {
    const BoxDesc& d = D[blockIdx.y];
    const int n = d.nunk, i = blockIdx.x;
    if (i >= n) return;
    const double* dpiv = d.W + 2LL * n * n;
    const int* perm = (const int*)(dpiv + n) + (size_t)((n - 1) & 1) * n;
    const double* src = d.W + (long long)perm[i] * 2 * n + n;
    const double dl = dpiv[n - 1];
    for (int q = threadIdx.x; q < n; q += blockDim.x) d.M[(long long)i * n + q] = i == n - 1 ? src[q] / dl : src[q];
}

// M = right half of W.  Grid (maxnunk rows, boxes).
__global__ __launch_bounds__(256) void kb_gj_extract(const BoxDesc* __restrict__ D) {
    const BoxDesc& d = D[blockIdx.y];
    const int r = blockIdx.x, n = d.nunk;
    if (r >= n) return;
    for (int q = threadIdx.x; q < n; q += blockDim.x) d.M[(long long)r * n + q] = d.W[(long long)r * 2 * n + n + q];
}

// Per-member model of the finest level: nearest Voronoi seed (exact int64
// distances, ties to the lowest index) and, with a crack, the degradation
// factor s = a eta^2 + b eta + c with eta = (1 - tanh(dist / width)) / 2, dist
// the distance to the half-plane {x : nrm.(x-p) = 0, tv.(x-p) <= 0}.
struct Crack {
    double p[3], nrm[3], tv[3];
    double width, a, b, c;
};
__global__ __launch_bounds__(TB) void kb_model(Lay L, double h, int nseeds, const long long* __restrict__ seeds, int qs,
                                               unsigned short* __restrict__ gid, int with_crack, Crack cr,
                                               double* __restrict__ sc64, float* __restrict__ sc32) {
    const int t = blockIdx.x * TB + threadIdx.x;
    const int nodes = L.nx * L.ny * L.nz;
    if (t >= nodes) return;
    const int nxy = L.nx * L.ny;
    const int k = t / nxy, j = (t - k * nxy) / L.nx, i = t - k * nxy - j * L.nx;
    if (gid) {
        const long long x = (long long)i * qs, y = (long long)j * qs, z = (long long)k * qs;
        long long best = -1;
        int bi = 0;
        for (int s = 0; s < nseeds; ++s) {
            const long long dx = x - seeds[3 * s], dy = y - seeds[3 * s + 1], dz = z - seeds[3 * s + 2];
            const long long dd = dx * dx + dy * dy + dz * dz;
            if (best < 0 || dd < best) { best = dd; bi = s; }
        }
        gid[t] = (unsigned short)bi;
    }
    if (with_crack && sc64) {
        const double X[3] = {i * h, j * h, k * h};
        double dn = 0, dt = 0;
        for (int a = 0; a < 3; ++a) {
            dn += cr.nrm[a] * (X[a] - cr.p[a]);
            dt += cr.tv[a] * (X[a] - cr.p[a]);
        }
        const double dist = sqrt(dn * dn + (dt > 0 ? dt * dt : 0.0));
        const double eta = 0.5 * (1.0 - tanh(dist / cr.width));
        const double s = (cr.a * eta + cr.b) * eta + cr.c;
        sc64[t] = s;
        if (sc32) sc32[t] = (float)s;
    }
}

// Right-hand side of an affine displacement u = G x on the Displacement rows
// of boundary nodes (BC::Operator::Elastic::Init with a field value), 0 on
// every other row (interior rows carry no body force, traction rows 0).
__global__ __launch_bounds__(TB) void kb_rhs_affine(Lay L, double h, BCTab bc, double g00, double g01, double g02,
                                                    double g10, double g11, double g12, double g20, double g21,
                                                    double g22, double* __restrict__ b) {
    const int t = blockIdx.x * TB + threadIdx.x;
    const int nodes = L.nx * L.ny * L.nz;
    if (t >= nodes) return;
    const int nxy = L.nx * L.ny;
    const int k = t / nxy, j = (t - k * nxy) / L.nx, i = t - k * nxy - j * L.nx;
    const int face = face_of<3>(i, j, k, L.nx - 1, L.ny - 1, L.nz - 1);
    const double x = i * h, y = j * h, z = k * h;
    const double v[3] = {g00 * x + g01 * y + g02 * z, g10 * x + g11 * y + g12 * z, g20 * x + g21 * y + g22 * z};
    for (int p = 0; p < 3; ++p) b[p * L.cs + t] = (face >= 0 && bc.t[face][p] == BC_DISP) ? v[p] : 0.0;
}

}  // namespace
}  // namespace ela

using namespace ela;

// ====================================================================== host side
namespace {

struct RLev {
    Lay L{};
    int n = 0;
    int nodes = 0;
    double h = 1;
    unsigned short* gid = nullptr;  // levels >= 1: grain ids (and factors) injected from the finer level
    double* sc64 = nullptr;
    float* sc32 = nullptr;
    double* d64 = nullptr;
    float* d32 = nullptr;
    void* dgt = nullptr;  // members without factors: per-grain interior Jacobi diagonal (BoxDesc::dgt)
    void* f[F_N] = {};
};

enum { ST_IDLE = 0, ST_QUEUED = 1, ST_ACTIVE = 2, ST_DONE = 3 };

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct ela_rve {
    ela_batch* B = nullptr;
    int n = 0, depth = 0, max_grains = 0, ngrains = 0, nunk = 0;
    int uniform = 0;
    bool factor = false, model = false, prepared = false;
    std::vector<RLev> lev;
    BCTab bc{};
    unsigned short* gid = nullptr;
    double *tab64 = nullptr, *sc64 = nullptr;
    float *tab32 = nullptr, *sc32 = nullptr;
    long long* seeds = nullptr;  // device copies of the seeds and rotations (model evaluation)
    double* rot = nullptr;
    double *u64 = nullptr, *b64 = nullptr, *r64 = nullptr, *M = nullptr, *scal = nullptr;
    bool bicg = false;  // coarsest level too large for the dense inverse: BiCGStab
    double* bs = nullptr;
    void* bv[BV_N] = {};
    void* block = nullptr;
    size_t bytes = 0;
    // solve state
    int state = ST_IDLE, status = ELA_OK, cycles = 0, max_cycles = 0;
    double tol_rel = 0, tol_abs = 0, res0 = 0, res = 0, target = 0, best = 0;
    int stall = 0;
    long long seq = 0;
    double t_submit = 0, t_done = 0;
};

struct ela_server;
struct ela_batch {
    ela_ctx* ctx = nullptr;
    ela_server* server = nullptr;  // an online server running on this batch
    int mode = ELA_SOLVE_MIXED;
    size_t tsize = 4;  // bytes of the cycle precision
    hipStream_t stream = nullptr;
    long long seq = 0;
    std::vector<ela_rve*> members;  // every live member
    std::vector<ela_rve*> active;   // in the schedule, ordered by (finest global level, seq)
    std::vector<ela_rve*> queued;   // submitted, join at the next step
    bool dirty = true;
    // launch lists of the active set
    int G = 0;                               // global levels
    std::vector<int> cnt;                    // cnt[g]: members whose finest global level <= g
    std::vector<std::vector<int>> tiles;     // tiles[g][m]: first block of the m-th box of level g (size cnt+1)
    std::vector<size_t> off_desc, off_bmap;  // device offsets of level lists (bytes)
    size_t off_top_desc = 0, off_top_bmap = 0;
    std::vector<int> top_tiles;
    // marching launches over the grain-form boxes of each level and of the top list
    struct March {
        size_t off_bm = 0, off_bs = 0, off_bf = 0;  // device offsets of the marching / shell / face block maps
        int nm = 0, ns = 0, nf = 0, maxng = 0;      // marching, shell and x-face patch blocks, largest grain count
        bool scaled = false;            // some box of the launch carries a per-node factor
        bool fits = true;               // every box's fields within 32-bit offsets (grain_offsets_fit)
    };
    std::vector<March> mlev;
    // mpre[g][c]: marching, shell and face-patch blocks and nodes of the first c boxes of level g's list
    // (the numbering runs box by box, so a prefix of the list launches a prefix of each part)
    std::vector<std::vector<std::array<long long, 4>>> mpre;
    March mtop;
    // coarsest boxes solved by BiCGStab (in the order of the coarsest list)
    size_t off_bdesc = 0, off_bbmap = 0;
    std::vector<int> btiles;  // first block of each such box (size nbicg + 1)
    std::vector<int> bpre;    // bpre[c]: such boxes among the first c boxes of the coarsest list
    int* dbstate = nullptr;   // their states after a fold (device, and pinned host)
    int* hbstate = nullptr;
    size_t bstate_cap = 0;
    bool march = true;
    int kz = 16;
    bool xcd = false;  // XCD-contiguous tile order of the marching launches
    // ELA_OPT_GRAPH: the F-cycle's launches captured into a HIP graph, re-captured
    // when the launch lists change (list_epoch), replayed otherwise
    bool graph_on = false;
    bool gj_one = true;  // ELA_OPT_GJ: the members' coarsest Gauss-Jordan in one launch a column
    bool gdiag = true;   // ELA_OPT_GRAIN_DIAG: the tiles' interior Jacobi diagonal per grain
    bool prolong_oct = true;  // prolongation with one thread per coarse node (ELA_OPT_PROLONG)
    long long list_epoch = 0;
    hipGraphExec_t gexec = nullptr;
    long long gexec_epoch = -1;
    size_t gexec_tsize = 0;
    int graph_captures = 0, graph_replays = 0;
    char* dlists = nullptr;
    size_t dlists_cap = 0;
    std::vector<char> hlists;
    // scratch
    double* partials = nullptr;
    size_t partials_cap = 0;
    double* dnorms = nullptr;
    double* hnorms = nullptr;  // pinned
    size_t norms_cap = 0;
    int* dflag = nullptr;
    // timing of the smoother launches (HIP events on the batch stream)
    bool timing = false;
    std::vector<hipEvent_t> ev;
    std::vector<std::pair<int, long long>> ev_launch;  // event pair index, nodes of the launch
    double ms_smooth = 0;
    long long smooth_nodes = 0;
    int n_smooth = 0;
};

namespace {

int fail(int code, const std::string& msg) { return set_error(code, msg); }

// Levels of a cube of n cells: coarsened while n is even and n / 2 >= 2.
int depth_of(int n) {
    int d = 1;
    while (n % 2 == 0 && n / 2 >= 2) { n /= 2; ++d; }
    return d;
}
int coarsest_of(int n) {
    while (n % 2 == 0 && n / 2 >= 2) n /= 2;
    return n;
}

size_t up256(size_t b) { return (b + 255) / 256 * 256; }

// Byte layout of a member's single allocation; with base == nullptr only sizes.
size_t layout_member(ela_rve* r, char* base) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + off : nullptr;
        off += up256(bytes);
        return p;
    };
    const size_t ts = r->B->tsize;
    const bool mixed = r->B->mode == ELA_SOLVE_MIXED;
    for (int l = 0; l < r->depth; ++l) {
        RLev& v = r->lev[l];
        const size_t N = (size_t)v.nodes;
        if (l == 0) {
            r->gid = (unsigned short*)take(N * 2);
            r->tab64 = (double*)take((size_t)r->max_grains * kNC * 8);
            if (mixed) r->tab32 = (float*)take((size_t)r->max_grains * kNC * 4);
            r->seeds = (long long*)take((size_t)r->max_grains * 3 * 8);
            r->rot = (double*)take((size_t)r->max_grains * 9 * 8);
            if (r->factor) {
                r->sc64 = (double*)take(N * 8);
                if (mixed) r->sc32 = (float*)take(N * 4);
            }
            r->u64 = (double*)take(3 * N * 8);
            r->b64 = (double*)take(3 * N * 8);
            r->r64 = (double*)take(3 * N * 8);
        } else {
            v.gid = (unsigned short*)take(N * 2);
            if (r->factor) {
                v.sc64 = (double*)take(N * 8);
                if (mixed) v.sc32 = (float*)take(N * 4);
            }
        }
        if (mixed) v.d32 = (float*)take(3 * N * 4);
        else v.d64 = (double*)take(3 * N * 8);
        v.dgt = r->factor ? nullptr : take((size_t)r->max_grains * 3 * ts);
        for (int q = 0; q < F_N; ++q) v.f[q] = take(3 * N * ts);
    }
    r->M = (double*)take((size_t)r->nunk * r->nunk * 8);
    r->scal = (double*)take(2 * 8);
    if (r->bicg) {
        const size_t N = (size_t)r->lev[r->depth - 1].nodes;
        for (int q = 0; q < BV_N; ++q) r->bv[q] = take(3 * N * ts);
        r->bs = (double*)take(2 * BS_N * 8);  // two sets (double-buffered)
    }
    return off;
}

BoxDesc desc_of(const ela_rve* r, int l) {
    const RLev& v = r->lev[l];
    BoxDesc d;
    std::memset(&d, 0, sizeof d);
    d.L = v.L;
    d.nodes = v.nodes;
    const double h3[3] = {v.h, v.h, v.h};
    d.i64 = make_inv<double>(h3);
    d.i32 = make_inv<float>(h3);
    d.h[0] = d.h[1] = d.h[2] = v.h;
    d.bc = r->bc;
    d.uniform = r->uniform;
    d.ngr = r->ngrains;
    d.tab64 = r->tab64;
    d.tab32 = r->tab32;
    if (l == 0) {
        d.gid = r->gid;
        d.sc64 = r->sc64;
        d.sc32 = r->sc32;
        d.u64 = r->u64;
        d.b64 = r->b64;
        d.r64 = r->r64;
    } else {
        d.gid = v.gid;
        d.sc64 = v.sc64;
        d.sc32 = v.sc32;
    }
    d.d64 = v.d64;
    d.d32 = v.d32;
    d.dgt = v.dgt;
    for (int q = 0; q < F_N; ++q) d.f[q] = v.f[q];
    if (l == r->depth - 1) {
        d.nunk = r->nunk;
        d.M = r->M;
        d.bs = r->bs;
        for (int q = 0; q < BV_N; ++q) d.bv[q] = r->bv[q];
    }
    d.norm = r->scal;
    d.scale = r->scal + 1;
    return d;
}

int ensure_device(ela_batch* B, void** p, size_t* cap, size_t bytes) {
    if (*cap >= bytes) return ELA_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    const size_t nb = std::max(bytes, (size_t)4096) * 3 / 2;
    ELA_HIPCHK(hipMalloc(p, nb));
    *cap = nb;
    return ELA_OK;
}

// A list of boxes with the block -> box map, packed into B->hlists.
struct ListBuild {
    std::vector<BoxDesc> descs;
    std::vector<int> tile;  // prefix of blocks (size descs + 1)
};

void list_push(ListBuild& lb, BoxDesc d) {
    if (lb.tile.empty()) lb.tile.push_back(0);
    d.tile0 = lb.tile.back();
    lb.descs.push_back(d);
    lb.tile.push_back(d.tile0 + (d.nodes + TB - 1) / TB);
}

void pack_list(ela_batch* B, const ListBuild& lb, size_t* off_desc, size_t* off_bmap) {
    auto align = [&]() { B->hlists.resize(up256(B->hlists.size())); };
    align();
    *off_desc = B->hlists.size();
    const size_t nd = lb.descs.size() * sizeof(BoxDesc);
    B->hlists.resize(*off_desc + nd);
    if (nd) std::memcpy(B->hlists.data() + *off_desc, lb.descs.data(), nd);
    align();
    *off_bmap = B->hlists.size();
    const int nblk = lb.tile.empty() ? 0 : lb.tile.back();
    B->hlists.resize(*off_bmap + (size_t)nblk * sizeof(int));
    int* bm = (int*)(B->hlists.data() + *off_bmap);
    for (size_t b = 0; b + 1 < lb.tile.size(); ++b)
        for (int t = lb.tile[b]; t < lb.tile[b + 1]; ++t) bm[t] = (int)b;
}

// Marching numbering of the grain-form boxes [a, b) of a list: interior tiles,
// shell blocks (z and y faces) and x-face patches per box, and the block ->
// box maps of the three parts of the launch.
void march_number(ListBuild& lb, int a, int b, int kz, std::vector<int>& bm, std::vector<int>& bs,
                  std::vector<int>& bf, ela_batch::March& mi) {
    mi = ela_batch::March{};
    bm.clear();
    bs.clear();
    bf.clear();
    for (int x = a; x < b; ++x) {
        BoxDesc& d = lb.descs[x];
        const Lay& L = d.L;
        d.kz = kz;
        grain_tiles(L, kz, d.ntx, d.nty, d.ntz, d.nsh, /*xface=*/true);
        d.mt0 = mi.nm;
        mi.nm += d.ntx * d.nty * d.ntz;
        d.st0 = mi.ns;
        mi.ns += d.nsh;
        d.nfp = grain_face_patches(L);
        d.fp0 = mi.nf;
        mi.nf += d.nfp;
        bf.insert(bf.end(), (size_t)d.nfp, x);
        mi.maxng = std::max(mi.maxng, d.ngr);
        mi.scaled = mi.scaled || d.sc64 != nullptr;
        mi.fits = mi.fits && grain_offsets_fit(L);
        bm.insert(bm.end(), (size_t)(d.ntx * d.ntz * d.nty), x);
        bs.insert(bs.end(), (size_t)d.nsh, x);
    }
}

size_t pack_ints(ela_batch* B, const std::vector<int>& v) {
    B->hlists.resize(up256(B->hlists.size()));
    const size_t off = B->hlists.size();
    B->hlists.resize(off + v.size() * sizeof(int));
    if (!v.empty()) std::memcpy(B->hlists.data() + off, v.data(), v.size() * sizeof(int));
    return off;
}

// Rebuild the level lists of the active set (members ordered by their finest global level).
int build_lists(ela_batch* B) {
    std::stable_sort(B->active.begin(), B->active.end(), [](const ela_rve* a, const ela_rve* b) {
        return a->depth != b->depth ? a->depth > b->depth : a->seq < b->seq;
    });
    int G = 0;
    for (auto* r : B->active) G = std::max(G, r->depth);
    B->G = G;
    B->cnt.assign(G, 0);
    B->tiles.assign(G, {});
    B->off_desc.assign(G, 0);
    B->off_bmap.assign(G, 0);
    B->hlists.clear();
    B->mlev.assign(G, ela_batch::March{});
    B->mpre.assign(G, {});
    std::vector<int> bm, bs, bf;
    for (int g = 0; g < G; ++g) {
        ListBuild lb;
        for (auto* r : B->active) {
            const int t = G - r->depth;
            if (t <= g) list_push(lb, desc_of(r, g - t));
        }
        B->cnt[g] = (int)lb.descs.size();
        B->tiles[g] = lb.tile.empty() ? std::vector<int>{0} : lb.tile;
        march_number(lb, 0, B->cnt[g], B->kz, bm, bs, bf, B->mlev[g]);
        long long nodes = 0;
        for (int c = 0; c < B->cnt[g]; ++c) {
            B->mpre[g].push_back({lb.descs[c].mt0, lb.descs[c].st0, lb.descs[c].fp0, nodes});
            nodes += lb.descs[c].nodes;
        }
        B->mpre[g].push_back({B->mlev[g].nm, B->mlev[g].ns, B->mlev[g].nf, nodes});
        pack_list(B, lb, &B->off_desc[g], &B->off_bmap[g]);
        B->mlev[g].off_bm = pack_ints(B, bm);
        B->mlev[g].off_bs = pack_ints(B, bs);
        B->mlev[g].off_bf = pack_ints(B, bf);
    }
    ListBuild bl;
    B->bpre.assign(1, 0);
    if (G > 0) {
        const BoxDesc* cd = (const BoxDesc*)(B->hlists.data() + B->off_desc[G - 1]);
        for (int m = 0; m < B->cnt[G - 1]; ++m) {
            if (cd[m].bs) list_push(bl, cd[m]);
            B->bpre.push_back((int)bl.descs.size());
        }
    }
    B->btiles = bl.tile.empty() ? std::vector<int>{0} : bl.tile;
    pack_list(B, bl, &B->off_bdesc, &B->off_bbmap);
    ListBuild top;
    for (auto* r : B->active) list_push(top, desc_of(r, 0));
    B->top_tiles = top.tile.empty() ? std::vector<int>{0} : top.tile;
    march_number(top, 0, (int)top.descs.size(), B->kz, bm, bs, bf, B->mtop);
    pack_list(B, top, &B->off_top_desc, &B->off_top_bmap);
    B->mtop.off_bm = pack_ints(B, bm);
    B->mtop.off_bs = pack_ints(B, bs);
    B->mtop.off_bf = pack_ints(B, bf);
    if (int rc = ensure_device(B, (void**)&B->dlists, &B->dlists_cap, B->hlists.size())) return rc;
    if (!B->hlists.empty())
        ELA_HIPCHK(hipMemcpyAsync(B->dlists, B->hlists.data(), B->hlists.size(), hipMemcpyHostToDevice, B->stream));
    const size_t nblk_top = std::max({(size_t)B->top_tiles.back(),
                                      (size_t)B->mtop.nm + (B->mtop.ns + 7) / 8 * 8 + (B->mtop.nf + 7) / 8 * 8,
                                      6 * (size_t)B->btiles.back()});
    if (int rc = ensure_device(B, (void**)&B->partials, &B->partials_cap, nblk_top * sizeof(double))) return rc;
    const size_t nbc = B->btiles.size() - 1;
    if (nbc && B->bstate_cap < nbc) {
        if (B->dbstate) (void)hipFree(B->dbstate);
        if (B->hbstate) (void)hipHostFree(B->hbstate);
        B->dbstate = nullptr;
        B->hbstate = nullptr;
        B->bstate_cap = 0;
        const size_t cap = std::max<size_t>(nbc * 2, 64);
        ELA_HIPCHK(hipMalloc((void**)&B->dbstate, cap * sizeof(int)));
        ELA_HIPCHK(hipHostMalloc((void**)&B->hbstate, cap * sizeof(int), hipHostMallocDefault));
        B->bstate_cap = cap;
    }
    if (B->norms_cap < B->active.size()) {
        if (B->dnorms) (void)hipFree(B->dnorms);
        if (B->hnorms) (void)hipHostFree(B->hnorms);
        B->dnorms = nullptr;
        B->hnorms = nullptr;
        B->norms_cap = 0;
        const size_t cap = std::max<size_t>(B->active.size() * 2, 64);
        ELA_HIPCHK(hipMalloc(&B->dnorms, cap * sizeof(double)));
        ELA_HIPCHK(hipHostMalloc((void**)&B->hnorms, cap * sizeof(double), hipHostMallocDefault));
        B->norms_cap = cap;
    }
    B->dirty = false;
    ++B->list_epoch;
    return ELA_OK;
}

// ------------------------------------------------------------------ launch helpers
const BoxDesc* LD(const ela_batch* B, int g) { return (const BoxDesc*)(B->dlists + B->off_desc[g]); }
const int* LM(const ela_batch* B, int g) { return (const int*)(B->dlists + B->off_bmap[g]); }
const BoxDesc* TD(const ela_batch* B) { return (const BoxDesc*)(B->dlists + B->off_top_desc); }
const int* TM(const ela_batch* B) { return (const int*)(B->dlists + B->off_top_bmap); }
int nblocks(const ela_batch* B, int g, int count) { return B->tiles[g][count]; }
long long nodes_of(const ela_batch* B, int g, int count) {
    const BoxDesc* d = (const BoxDesc*)(B->hlists.data() + B->off_desc[g]);
    long long s = 0;
    for (int m = 0; m < count; ++m) s += d[m].nodes;
    return s;
}

bool use_march(const ela_batch* B, const ela_batch::March& mi) {
    return B->march && mi.nm > 0 && mi.maxng <= kMarchMaxGrains && mi.fits;
}

// The marching launch of the first count boxes of level g.
ela_batch::March march_prefix(const ela_batch* B, int g, int count) {
    ela_batch::March mi = B->mlev[g];
    const auto& p = B->mpre[g][count];
    mi.nm = (int)p[0];
    mi.ns = (int)p[1];
    mi.nf = (int)p[2];
    return mi;
}

// A launch over the first count boxes of level g marches when the whole level
// is in flight or the boxes hold kMarchPrefixNodes nodes or more: over fewer,
// the tiles' per-block table staging costs more than the node kernels' reads
// (measured on MI355X: the serving stream at 20 requests/s, p50 267 -> 339 ms
// with every prefix marched).
constexpr long long kMarchPrefixNodes = 1LL << 20;
bool march_on(const ela_batch* B, int g, int count) {
    if (!use_march(B, B->mlev[g])) return false;
    return count == B->cnt[g] || B->mpre[g][count][3] >= kMarchPrefixNodes;
}

template <typename T, int MODE, typename TO>
void march_launch(ela_batch* B, const BoxDesc* D, const ela_batch::March& mi, const MOps& op) {
    const size_t lds = (size_t)mi.maxng * (kNC + 3) * sizeof(T);  // the records, then the per-grain diagonal
    const int nsp = (mi.ns + 7) / 8 * 8, nfp = (mi.nf + 7) / 8 * 8;
    const int* bm = (const int*)(B->dlists + mi.off_bm);
    const int* bs = (const int*)(B->dlists + mi.off_bs);
    const int* bf = (const int*)(B->dlists + mi.off_bf);
    if (mi.scaled)
        hipLaunchKernelGGL((kf_grain<T, MODE, TO, true>), dim3(nsp + nfp + mi.nm), dim3(MTX * MTY), lds, B->stream, D,
                           bm, bs, bf, mi.ns, nsp, mi.nf, nfp, op);
    else
        hipLaunchKernelGGL((kf_grain<T, MODE, TO, false>), dim3(nsp + nfp + mi.nm), dim3(MTX * MTY), lds, B->stream,
                           D, bm, bs, bf, mi.ns, nsp, mi.nf, nfp, op);
}

template <typename T>
void smooth(ela_batch* B, int g, int count, int rhs, int in, int out) {
    const int nb = nblocks(B, g, count);
    if (!nb) return;
    int e = -1;
    if (B->timing && B->ev_launch.size() * 2 + 2 <= B->ev.size()) {
        e = (int)B->ev_launch.size() * 2;
        (void)hipEventRecord(B->ev[e], B->stream);
    }
    if (in >= 0 && march_on(B, g, count))
        march_launch<T, 0, T>(B, LD(B, g), march_prefix(B, g, count), MOps{rhs, in, out, 2.0 / 3.0, nullptr, 0, B->xcd, B->gdiag ? 1 : 0});
    else
        hipLaunchKernelGGL((kb_smooth<T>), dim3(nb), dim3(TB), 0, B->stream, LD(B, g), LM(B, g), 0, rhs, in, out,
                           (T)(2.0 / 3.0));
    if (e >= 0) {
        (void)hipEventRecord(B->ev[e + 1], B->stream);
        B->ev_launch.push_back({e, nodes_of(B, g, count)});
    }
}

// nu calls of Fsmooth (2 sweeps each); the first call of a pre-smoothing starts from zero.
template <typename T>
void fsmooth(ela_batch* B, int g, int count, int rhs, bool zero_start) {
    for (int c = 0; c < 2; ++c) {
        smooth<T>(B, g, count, rhs, (zero_start && c == 0) ? -1 : F_COR, F_TMP);
        smooth<T>(B, g, count, rhs, F_TMP, F_COR);
    }
}

template <typename T>
void resid(ela_batch* B, int g, int count, int x, int b, int r) {
    const int nb = nblocks(B, g, count);
    if (!nb) return;
    if (march_on(B, g, count))
        march_launch<T, 1, T>(B, LD(B, g), march_prefix(B, g, count), MOps{b, x, r, 0.0, nullptr, 0, B->xcd});
    else
        hipLaunchKernelGGL((kb_resid<T>), dim3(nb), dim3(TB), 0, B->stream, LD(B, g), LM(B, g), 0, x, b, r);
}

template <typename T>
void restrict_to(ela_batch* B, int g, int count, int src, int dst) {  // level g -> g + 1
    const int nb = nblocks(B, g + 1, count);
    if (nb)
        hipLaunchKernelGGL((kb_restrict<T>), dim3(nb), dim3(TB), 0, B->stream, LD(B, g + 1), LD(B, g), LM(B, g + 1),
                           src, dst);
}

// Prolongation onto level g: one thread per coarse node (kb_prolong_oct) over
// the coarse list's blocks of the same members (measured on MI355X, 512 RVEs
// batched: mixed 350 -> 364 solves/s, FP64 216 -> 220), or one per fine node
// (ELA_OPT_PROLONG 0).
template <typename T>
void prolong_to(ela_batch* B, int g, int count, int src, int dst, int add) {  // level g + 1 -> g
    const int nbc = nblocks(B, g + 1, count);
    if (B->prolong_oct && nbc) {
        hipLaunchKernelGGL((kb_prolong_oct<T>), dim3(nbc), dim3(TB), 0, B->stream, LD(B, g), LD(B, g + 1),
                           LM(B, g + 1), src, dst, add);
        return;
    }
    const int nb = nblocks(B, g, count);
    if (nb)
        hipLaunchKernelGGL((kb_prolong<T>), dim3(nb), dim3(TB), 0, B->stream, LD(B, g), LD(B, g + 1), LM(B, g), src,
                           dst, add);
}

template <typename T>
void axpy(ela_batch* B, int g, int count, int x, int y) {
    const int nb = nblocks(B, g, count);
    if (nb) hipLaunchKernelGGL((kb_axpy<T>), dim3(nb), dim3(TB), 0, B->stream, LD(B, g), LM(B, g), x, y);
}

const BoxDesc* BD(const ela_batch* B) { return (const BoxDesc*)(B->dlists + B->off_bdesc); }
const int* BM(const ela_batch* B) { return (const int*)(B->dlists + B->off_bbmap); }

// BiCGStab on the first nbc boxes of the BiCGStab list: f[dst] = A^-1 f[src]
// to 1e-4, five launches an iteration.  Every kBottomCheck iterations the host
// reads the states the iteration's first launch wrote and stops once no member
// iterates; launches past a member's end leave it unchanged.
template <typename T>
void bicg_bottom(ela_batch* B, int nbc, int src, int dst) {
    const int nb = B->btiles[nbc];
    const BoxDesc* D = BD(B);
    const int* bm = BM(B);
    hipStream_t st = B->stream;
    double* Pa = B->partials;
    double* Pb = B->partials + 3LL * B->btiles.back();
    int ph = 0;
    hipLaunchKernelGGL((kbb_init<T>), dim3(nb), dim3(TB), 0, st, D, bm, src, dst, Pa);
    for (int it = 0; it < kBottomMaxIter; ++it) {
        hipLaunchKernelGGL((kbb_p<T>), dim3(nb), dim3(TB), 0, st, D, bm, (const double*)Pa, ph, it == 0 ? 1 : 0,
                           B->dbstate);
        ph ^= 1;
        if (it % kBottomCheck == 0) {
            (void)hipMemcpyAsync(B->hbstate, B->dbstate, nbc * sizeof(int), hipMemcpyDeviceToHost, st);
            (void)hipStreamSynchronize(st);
            bool any = false;
            for (int m = 0; m < nbc && !any; ++m) any = B->hbstate[m] != 0;
            if (!any) break;
        }
        hipLaunchKernelGGL((kbb_av<T>), dim3(nb), dim3(TB), 0, st, D, bm, Pb, ph);
        ph ^= 1;
        hipLaunchKernelGGL((kbb_s<T>), dim3(nb), dim3(TB), 0, st, D, bm, (const double*)Pb, Pa, ph);
        ph ^= 1;
        hipLaunchKernelGGL((kbb_at<T>), dim3(nb), dim3(TB), 0, st, D, bm, (const double*)Pa, Pb, ph);
        ph ^= 1;
        hipLaunchKernelGGL((kbb_x<T>), dim3(nb), dim3(TB), 0, st, D, bm, dst, (const double*)Pb, Pa, ph);
        ph ^= 1;
    }
}

template <typename T>
void bottom(ela_batch* B, int count, int src, int dst) {
    const int g = B->G - 1;
    if (!count) return;
    const BoxDesc* hd = (const BoxDesc*)(B->hlists.data() + B->off_desc[g]);
    int maxn = 0;
    for (int m = 0; m < count; ++m) maxn = std::max(maxn, hd[m].nunk);
    const int rpb = 4 * (TB / 64);
    if (maxn > 0)
        hipLaunchKernelGGL((kb_bottom<T>), dim3((maxn + rpb - 1) / rpb, count), dim3(TB), 0, B->stream, LD(B, g), src,
                           dst);
    const int nbc = B->bpre[count];
    if (nbc > 0) bicg_bottom<T>(B, nbc, src, dst);
}

// V-cycle from global level s on the members whose finest level is <= s; the
// right-hand side on level s is f[RC], on coarser levels f[RES].
template <typename T>
void vcycle(ela_batch* B, int s) {
    const int G = B->G, n = B->cnt[s];
    for (int g = s; g < G - 1; ++g) {
        const int rhs = g == s ? F_RC : F_RES;
        fsmooth<T>(B, g, n, rhs, true);
        resid<T>(B, g, n, F_COR, rhs, F_TMP);
        restrict_to<T>(B, g, n, F_TMP, F_RES);
    }
    bottom<T>(B, n, s == G - 1 ? F_RC : F_RES, F_COR);
    for (int g = G - 2; g >= s; --g) {
        const int rhs = g == s ? F_RC : F_RES;
        prolong_to<T>(B, g, n, F_COR, F_COR, 1);
        fsmooth<T>(B, g, n, rhs, false);
    }
}

// F-cycle on the whole active set: restrict the right-hand side to every
// level, solve the coarsest, then per level upward: hold = P cor_c,
// rc = res - A hold, V-cycle on rc, cor += hold.
template <typename T>
void fcycle(ela_batch* B) {
    const int G = B->G;
    for (int g = 0; g + 1 < G; ++g) restrict_to<T>(B, g, B->cnt[g], F_RES, F_RES);
    bottom<T>(B, B->cnt[G - 1], F_RES, F_COR);
    for (int g = G - 2; g >= 0; --g) {
        const int n = B->cnt[g];
        if (!n) continue;
        prolong_to<T>(B, g, n, F_COR, F_HOLD, 0);
        resid<T>(B, g, n, F_HOLD, F_RES, F_RC);
        vcycle<T>(B, g);
        axpy<T>(B, g, n, F_HOLD, F_COR);
    }
}

template <typename T>
void resid64_top(ela_batch* B, const BoxDesc* D, const int* bmap, int nbox, int nblk) {
    if (!nbox) return;
    hipLaunchKernelGGL((kb_resid64<T>), dim3(nblk), dim3(TB), 0, B->stream, D, bmap, F_RES, B->partials);
    hipLaunchKernelGGL(kb_fold, dim3(nbox), dim3(64), 0, B->stream, D, B->partials, B->dnorms);
}

int read_norms(ela_batch* B, int nbox) {
    ELA_HIPCHK(hipMemcpyAsync(B->hnorms, B->dnorms, nbox * sizeof(double), hipMemcpyDeviceToHost, B->stream));
    ELA_HIPCHK(hipStreamSynchronize(B->stream));
    return ELA_OK;
}

// Admit queued members: initial residual (u as uploaded, zero by default) with scale 1.
int admit(ela_batch* B) {
    if (B->queued.empty()) return ELA_OK;
    ListBuild jl;
    for (auto* r : B->queued) list_push(jl, desc_of(r, 0));
    std::vector<char> save;
    save.swap(B->hlists);
    size_t od = 0, om = 0;
    pack_list(B, jl, &od, &om);
    std::vector<char> jbytes;
    jbytes.swap(B->hlists);
    B->hlists.swap(save);
    // The join lists go to a separate device buffer region at the end of dlists.
    const size_t base = up256(B->hlists.size());
    if (int rc = ensure_device(B, (void**)&B->dlists, &B->dlists_cap, base + jbytes.size())) return rc;
    B->dirty = true;  // dlists may have been reallocated: the active lists are rebuilt anyway
    ELA_HIPCHK(hipMemcpyAsync(B->dlists + base, jbytes.data(), jbytes.size(), hipMemcpyHostToDevice, B->stream));
    const size_t nblk = (size_t)jl.tile.back();
    if (int rc = ensure_device(B, (void**)&B->partials, &B->partials_cap, nblk * sizeof(double))) return rc;
    if (B->norms_cap < B->queued.size()) {
        if (B->dnorms) (void)hipFree(B->dnorms);
        if (B->hnorms) (void)hipHostFree(B->hnorms);
        const size_t cap = std::max<size_t>(B->queued.size() * 2, 64);
        ELA_HIPCHK(hipMalloc(&B->dnorms, cap * sizeof(double)));
        ELA_HIPCHK(hipHostMalloc((void**)&B->hnorms, cap * sizeof(double), hipHostMallocDefault));
        B->norms_cap = cap;
    }
    const double one[2] = {0.0, 1.0};
    for (auto* r : B->queued) ELA_HIPCHK(hipMemcpyAsync(r->scal, one, sizeof one, hipMemcpyHostToDevice, B->stream));
    const BoxDesc* D = (const BoxDesc*)(B->dlists + base + od);
    const int* bm = (const int*)(B->dlists + base + om);
    if (B->mode == ELA_SOLVE_MIXED) resid64_top<float>(B, D, bm, (int)jl.descs.size(), (int)nblk);
    else resid64_top<double>(B, D, bm, (int)jl.descs.size(), (int)nblk);
    if (int rc = read_norms(B, (int)jl.descs.size())) return rc;
    for (size_t m = 0; m < B->queued.size(); ++m) {
        ela_rve* r = B->queued[m];
        const double r0 = B->hnorms[m];
        r->res0 = r->res = r->best = r0;
        r->cycles = 0;
        r->stall = 0;
        r->target = std::max(r->tol_abs, std::max(r->tol_rel, 1e-16) * r0);
        if (!(r0 == r0)) {
            r->status = fail(ELA_ERR_NAN, "initial residual is not finite");
            r->state = ST_DONE;
            r->t_done = now_s();
        } else if (r0 <= r->target) {
            r->status = ELA_OK;
            r->state = ST_DONE;
            r->t_done = now_s();
        } else {
            r->state = ST_ACTIVE;
            B->active.push_back(r);
        }
    }
    B->queued.clear();
    return ELA_OK;
}

// The F-cycle as a HIP graph (ELA_OPT_GRAPH): captured from the stream when
// the launch lists changed since the last capture, else replayed.  Kernel
// nodes only (the cycle has no memsets or copies).  Cycles with a BiCGStab
// coarsest level read the members' states on the host, and timed cycles
// record events, so those run on the stream.
template <typename T>
int fcycle_graph(ela_batch* B) {
    if (!B->gexec || B->gexec_epoch != B->list_epoch || B->gexec_tsize != sizeof(T)) {
        if (B->gexec) {
            ELA_HIPCHK(hipStreamSynchronize(B->stream));  // its last replay has finished
            (void)hipGraphExecDestroy(B->gexec);
            B->gexec = nullptr;
        }
        hipGraph_t g = nullptr;
        ELA_HIPCHK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
        fcycle<T>(B);
        const hipError_t e = hipStreamEndCapture(B->stream, &g);
        if (e != hipSuccess) return fail(ELA_ERR_HIP, std::string("batch cycle capture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&B->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            B->gexec = nullptr;
            return fail(ELA_ERR_HIP, std::string("batch cycle graph: ") + hipGetErrorString(ei));
        }
        B->gexec_epoch = B->list_epoch;
        B->gexec_tsize = sizeof(T);
        ++B->graph_captures;
    } else {
        ++B->graph_replays;
    }
    ELA_HIPCHK(hipGraphLaunch(B->gexec, B->stream));
    return ELA_OK;
}

// One F-cycle on every active member, then the FP64 update and residual;
// members that converge, fail or run out of cycles leave the active set.
template <typename T>
int step_t(ela_batch* B) {
    if (B->graph_on && !B->timing && B->bpre.back() == 0) {
        if (int rc = fcycle_graph<T>(B)) return rc;
    } else {
        fcycle<T>(B);
    }
    const int nbox = (int)B->active.size();
    const int nblk = B->top_tiles.back();
    hipLaunchKernelGGL((kb_update<T>), dim3(nblk), dim3(TB), 0, B->stream, TD(B), TM(B), F_COR);
    hipLaunchKernelGGL(kb_scale, dim3((nbox + 63) / 64), dim3(64), 0, B->stream, TD(B), nbox);
    if (use_march(B, B->mtop)) {
        march_launch<double, 2, T>(B, TD(B), B->mtop, MOps{0, 0, F_RES, 0.0, B->partials, B->mtop.nm, B->xcd});
        hipLaunchKernelGGL(kb_fold_m, dim3(nbox), dim3(64), 0, B->stream, TD(B), B->partials, B->mtop.nm,
                           (B->mtop.ns + 7) / 8 * 8, B->dnorms);
    } else {
        resid64_top<T>(B, TD(B), TM(B), nbox, nblk);
    }
    return read_norms(B, nbox);
}

int step(ela_batch* B) {
    if (int rc = admit(B)) return rc;
    if (B->active.empty()) return ELA_OK;
    if (B->dirty)
        if (int rc = build_lists(B)) return rc;
    const int rc = B->mode == ELA_SOLVE_MIXED ? step_t<float>(B) : step_t<double>(B);
    if (rc) return rc;
    const double t = now_s();
    std::vector<ela_rve*> keep;
    for (size_t m = 0; m < B->active.size(); ++m) {
        ela_rve* r = B->active[m];
        const double rn = B->hnorms[m];
        r->res = rn;
        ++r->cycles;
        bool done = false;
        if (!(rn == rn)) {
            r->status = ELA_ERR_NAN;
            done = true;
        } else if (rn <= r->target) {
            r->status = ELA_OK;
            done = true;
        } else {
            if (rn < 0.5 * r->best) {
                r->best = rn;
                r->stall = 0;
            } else if (++r->stall >= 12) {
                r->status = ELA_ERR_NONCONVERGED;
                done = true;
            }
            if (r->cycles >= r->max_cycles) {
                r->status = ELA_ERR_NONCONVERGED;
                done = true;
            }
        }
        if (done) {
            r->state = ST_DONE;
            r->t_done = t;
            B->dirty = true;
        } else {
            keep.push_back(r);
        }
    }
    B->active.swap(keep);
    return ELA_OK;
}

int check_rve(const ela_rve* r) {
    if (!r) return fail(ELA_ERR_BADARG, "null member");
    return ELA_OK;
}

// Batched set-up of the members: coarse grain ids and factors, diagonals, coarsest inverses.
int prepare(ela_batch* B, int count, ela_rve* const* rves) {
    std::vector<ela_rve*> set(rves, rves + count);
    for (auto* r : set) {
        if (int rc = check_rve(r)) return rc;
        if (r->B != B) return fail(ELA_ERR_BADARG, "member belongs to another batch");
        if (!r->model) return fail(ELA_ERR_STATE, "model not set (SetModel before prepareForSolve)");
        if (r->state == ST_QUEUED || r->state == ST_ACTIVE) return fail(ELA_ERR_STATE, "member is being solved");
    }
    if (set.empty()) return ELA_OK;
    std::stable_sort(set.begin(), set.end(), [](const ela_rve* a, const ela_rve* b) { return a->depth > b->depth; });
    const int G = set.front()->depth;
    // Per global level lists (the same alignment as the solve).
    std::vector<ListBuild> L(G);
    for (int g = 0; g < G; ++g)
        for (auto* r : set) {
            const int t = G - r->depth;
            if (t <= g) list_push(L[g], desc_of(r, g - t));
        }
    // Set-up workspace: identities per coarsest size and [A | I] per member.
    std::vector<int> sizes;
    for (auto* r : set)
        if (r->nunk > 0) sizes.push_back(r->nunk);
    std::sort(sizes.begin(), sizes.end());
    sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
    size_t ws = 0;
    for (int s : sizes) ws += up256((size_t)s * s * 8);
    // [A | I], then the one-launch-a-column scratch: pivots (n doubles), two permutations (2n ints) and two
    // candidate columns (2n doubles)
    auto wsize = [](int n) { return up256(((size_t)n * 2 * n + 4 * (size_t)n) * 8); };
    for (auto* r : set) ws += wsize(r->nunk);
    char* wbase = nullptr;
    ELA_HIPCHK(hipMalloc((void**)&wbase, std::max<size_t>(ws, 256)));
    ELA_HIPCHK(hipMemsetAsync(wbase, 0, std::max<size_t>(ws, 256), B->stream));
    size_t woff = 0;
    std::vector<double*> eye(sizes.size());
    for (size_t q = 0; q < sizes.size(); ++q) {
        eye[q] = (double*)(wbase + woff);
        woff += up256((size_t)sizes[q] * sizes[q] * 8);
        const long long nn = (long long)sizes[q] * sizes[q];
        hipLaunchKernelGGL(kb_identity, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, B->stream, eye[q], sizes[q]);
    }
    for (auto& d : L[G - 1].descs) {
        if (d.nunk == 0) continue;
        const int q = (int)(std::lower_bound(sizes.begin(), sizes.end(), d.nunk) - sizes.begin());
        d.E = eye[q];
        d.W = (double*)(wbase + woff);
        woff += wsize(d.nunk);
    }
    // Pack and upload the lists.
    std::vector<char> save;
    save.swap(B->hlists);
    std::vector<size_t> od(G), om(G);
    for (int g = 0; g < G; ++g) pack_list(B, L[g], &od[g], &om[g]);
    char* dl = nullptr;
    int rc = ELA_OK;
    if (hipMalloc((void**)&dl, std::max<size_t>(B->hlists.size(), 256)) != hipSuccess) rc = fail(ELA_ERR_OOM, "hipMalloc");
    if (!rc)
        (void)hipMemcpyAsync(dl, B->hlists.data(), B->hlists.size(), hipMemcpyHostToDevice, B->stream);
    B->hlists.swap(save);
    if (!rc) {
        ELA_HIPCHK(hipMemsetAsync(B->dflag, 0, sizeof(int), B->stream));
        auto D = [&](int g) { return (const BoxDesc*)(dl + od[g]); };
        auto M = [&](int g) { return (const int*)(dl + om[g]); };
        auto nb = [&](int g, int c) { return L[g].tile[c]; };
        // coarse grain ids and factors (level by level, each member's next level from its previous one)
        for (int g = 0; g + 1 < G; ++g) {
            const int c = (int)L[g].descs.size();
            hipLaunchKernelGGL(kb_inject, dim3(nb(g + 1, c)), dim3(TB), 0, B->stream, D(g + 1), D(g), M(g + 1));
        }
        for (int g = 0; g < G; ++g) {
            const int c = (int)L[g].descs.size();
            hipLaunchKernelGGL(kb_diag, dim3(nb(g, c)), dim3(TB), 0, B->stream, D(g), M(g), B->dflag);
            hipLaunchKernelGGL(kb_grain_diag, dim3(c), dim3(64), 0, B->stream, D(g), B->mode == ELA_SOLVE_MIXED ? 1 : 0);
        }
        // coarsest inverse: probe A, Gauss-Jordan on [A | I], keep A^-1
        const int nbox = (int)L[G - 1].descs.size();
        int maxn = 0, maxnodes = 0;
        for (auto& d : L[G - 1].descs) {
            maxn = std::max(maxn, d.nunk);
            maxnodes = std::max(maxnodes, d.nodes);
        }
        const int nbk = (maxnodes + TB - 1) / TB;
        if (maxn > 0) {  // (members with a BiCGStab coarsest level have nunk = 0 and are skipped)
            hipLaunchKernelGGL(kb_probe, dim3(maxn * nbk, nbox), dim3(TB), 0, B->stream, D(G - 1), nbk);
            if (B->gj_one) {
                for (int c = 0; c < maxn; ++c)
                    hipLaunchKernelGGL(kb_gj_col, dim3(maxn, nbox), dim3(256), 0, B->stream, D(G - 1), c, B->dflag);
                hipLaunchKernelGGL(kb_gj_final, dim3(maxn, nbox), dim3(256), 0, B->stream, D(G - 1));
            } else {
                for (int c = 0; c < maxn; ++c) {
                    hipLaunchKernelGGL(kb_gj_pivot, dim3(nbox), dim3(1024), 0, B->stream, D(G - 1), c, B->dflag);
                    hipLaunchKernelGGL(kb_gj_eliminate, dim3(maxn, nbox), dim3(256), 0, B->stream, D(G - 1), c);
                }
                hipLaunchKernelGGL(kb_gj_extract, dim3(maxn, nbox), dim3(256), 0, B->stream, D(G - 1));
            }
        }
        int flag = 0;
        (void)hipMemcpyAsync(&flag, B->dflag, sizeof flag, hipMemcpyDeviceToHost, B->stream);
        const hipError_t e = hipStreamSynchronize(B->stream);
        if (e != hipSuccess) rc = fail(ELA_ERR_HIP, std::string("prepare: ") + hipGetErrorString(e));
        else if (flag) rc = fail(ELA_ERR_NAN, "zero or non-finite diagonal or singular coarsest operator");
    }
    if (dl) (void)hipFree(dl);
    (void)hipFree(wbase);
    if (!rc)
        for (auto* r : set) r->prepared = true;
    return rc;
}

// ------------------------------------------------------------------ serving core
// One request with its arrays owned by the loop (an online submit copies them).
struct OwnedRequest {
    ela_request q{};
    std::vector<int64_t> seeds;
    std::vector<double> R;
    long long ticket = 0;
    double t_submit = 0;
};

// A member for one request: created against the HBM budget, its model and right-hand side evaluated.
int make_member(ela_batch* B, const ela_request& q, ela_rve** out) {
    *out = nullptr;
    ela_rve* r = nullptr;
    int rc = ela_rve_create(B, q.n, std::max(1, q.nseeds), q.crack ? 1 : 0, &r);
    if (rc) return rc;
    rc = ela_rve_set_voronoi_cubic(r, q.nseeds, q.seeds, q.q, q.R, q.C11, q.C12, q.C44);
    if (!rc && q.crack)
        rc = ela_rve_set_crack(r, q.crack_p, q.crack_nrm, q.crack_tv, q.crack_width, q.crack_g[0], q.crack_g[1],
                               q.crack_g[2]);
    if (!rc) rc = ela_rve_set_rhs_affine(r, q.G);
    if (rc) {
        ela_rve_destroy(r);
        return rc;
    }
    *out = r;
    return ELA_OK;
}

// Set-up of an admitted group.  A member whose set-up fails (a non-finite or zero
// diagonal, a singular coarsest operator) fails alone: the group is then set up
// member by member, which gives each member the bits it gets in any group.
// status[m] receives each member's result; the return value is a fatal error only.
int prepare_group(ela_batch* B, const std::vector<ela_rve*>& group, std::vector<int>& status) {
    status.assign(group.size(), ELA_OK);
    const int rc = prepare(B, (int)group.size(), group.data());
    if (rc == ELA_OK) return ELA_OK;
    if (rc != ELA_ERR_NAN) return rc;
    for (size_t m = 0; m < group.size(); ++m) {
        ela_rve* one = group[m];
        const int r1 = prepare(B, 1, &one);
        if (r1 != ELA_OK && r1 != ELA_ERR_NAN) return r1;
        status[m] = r1;
    }
    return ELA_OK;
}

// The continuous-batching loop of ela_serve and of the online server: requests
// waiting for admission, members in flight, and the time breakdown.
struct ServeCore {
    ela_batch* B;
    double tol_rel, tol_abs;
    int max_cycles, max_active;
    double t0;
    std::deque<std::pair<ela_request, long long>> waiting;  // request, caller id
    std::vector<std::pair<ela_rve*, long long>> live;
    std::vector<std::pair<long long, double*>> uout;          // caller id -> u_out
    size_t live_bytes = 0;
    long long done = 0;
    ela_serve_summary sm{};
    ServeCore(ela_batch* b, double tr, double ta, int mc, int ma)
        : B(b), tol_rel(tr), tol_abs(ta), max_cycles(mc), max_active(ma), t0(now_s()) {}
    bool busy() const { return !B->active.empty() || !B->queued.empty(); }
    static ela_result ended(int status, double t) {
        ela_result o{};
        o.status = status;
        o.admitted = o.finished = t;
        return o;
    }
    // Admit waiting requests first come, first served while memory and max_active allow.
    template <class Sink>
    int admit(Sink&& sink) {
        const double ta0 = now_s();
        std::vector<ela_rve*> group;
        std::vector<std::pair<ela_request, long long>> gq;
        while (!waiting.empty() && (int)live.size() + (int)group.size() < max_active) {
            const auto& w = waiting.front();
            ela_rve* r = nullptr;
            const int crc = make_member(B, w.first, &r);
            if (crc == ELA_ERR_OOM && (!live.empty() || !group.empty())) break;  // wait for memory
            if (crc) {  // a request that cannot be served (bad size or model, or larger than the whole budget)
                sink(w.second, ended(crc, now_s() - t0));
                ++done;
                waiting.pop_front();
                continue;
            }
            group.push_back(r);
            gq.push_back(w);
            live_bytes += r->bytes;
            waiting.pop_front();
        }
        sm.t_admit += now_s() - ta0;
        if (group.empty()) return ELA_OK;
        const double tp0 = now_s();
        std::vector<int> st;
        if (int rc = prepare_group(B, group, st)) return rc;
        sm.t_prepare += now_s() - tp0;
        ++sm.groups;
        sm.admitted += (int)group.size();
        const double ta = now_s() - t0;
        for (size_t m = 0; m < group.size(); ++m) {
            const ela_request& q = gq[m].first;
            if (st[m] != ELA_OK) {
                live_bytes -= group[m]->bytes;
                ela_rve_destroy(group[m]);
                ela_result o = ended(st[m], ta);
                sink(gq[m].second, o);
                ++done;
                continue;
            }
            if (int rc = ela_batch_submit(B, group[m], tol_rel, tol_abs, q.max_cycles > 0 ? q.max_cycles : max_cycles))
                return rc;
            live.push_back({group[m], gq[m].second});
            if (q.u_out) uout.push_back({gq[m].second, q.u_out});
            admitted_at.push_back({gq[m].second, ta});
        }
        sm.peak_active = std::max(sm.peak_active, (int)live.size());
        sm.peak_bytes = std::max(sm.peak_bytes, live_bytes);
        return ELA_OK;
    }
    std::vector<std::pair<long long, double>> admitted_at;
    template <class V>
    static typename V::value_type::second_type take(V& v, long long id, typename V::value_type::second_type dflt) {
        for (size_t q = 0; q < v.size(); ++q)
            if (v[q].first == id) {
                auto x = v[q].second;
                v[q] = v.back();
                v.pop_back();
                return x;
            }
        return dflt;
    }
    // One F-cycle over every member in flight; finished members are reported, copied out and freed.
    template <class Sink>
    int step_retire(Sink&& sink) {
        const double ts0 = now_s();
        if (int rc = step(B)) return rc;
        sm.t_step += now_s() - ts0;
        ++sm.steps;
        const double tr0 = now_s();
        std::vector<std::pair<ela_rve*, long long>> keep;
        for (auto& pr : live) {
            ela_rve* r = pr.first;
            if (r->state != ST_DONE) {
                keep.push_back(pr);
                continue;
            }
            ela_result o{};
            o.admitted = take(admitted_at, pr.second, -1.0);
            o.finished = r->t_done - t0;
            o.cycles = r->cycles;
            o.status = r->status;
            o.res0 = r->res0;
            o.res = r->res;
            if (double* u = take(uout, pr.second, (double*)nullptr)) ela_rve_download(r, 0, u);
            live_bytes -= r->bytes;
            ela_rve_destroy(r);
            ++done;
            sink(pr.second, o);
        }
        live.swap(keep);
        sm.t_retire += now_s() - tr0;
        return ELA_OK;
    }
    // Members still in flight are dropped (reported by the caller as not served).
    void drop() {
        for (auto& pr : live) ela_rve_destroy(pr.first);
        live.clear();
        B->active.clear();
        B->queued.clear();
        B->dirty = true;
    }
};

}  // namespace

// The online server: a host thread runs ServeCore over the requests submitted so far.
struct ela_server {
    ela_batch* B = nullptr;
    std::unique_ptr<ServeCore> core;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    bool stopping = false, drain = true;
    std::deque<OwnedRequest> inbox;                  // submitted, not yet taken by the loop
    std::deque<OwnedRequest> held;                   // taken: their arrays stay alive until done
    std::unordered_map<long long, ela_result> results;
    long long next_ticket = 1;
    int fatal = ELA_OK;
    std::string fatal_msg;
};

namespace {

void server_loop(ela_server* sv) {
    if (hipSetDevice(sv->B->ctx->device) != hipSuccess) {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->fatal = ELA_ERR_HIP;
        sv->fatal_msg = "hipSetDevice failed in the serving thread";
        return;
    }
    ServeCore& core = *sv->core;
    auto sink = [sv](long long ticket, const ela_result& r) {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->results[ticket] = r;
        for (auto it = sv->held.begin(); it != sv->held.end(); ++it)
            if (it->ticket == ticket) {
                sv->held.erase(it);
                break;
            }
    };
    for (;;) {
        {
            std::unique_lock<std::mutex> g(sv->mu);
            if (core.waiting.empty() && !core.busy() && sv->inbox.empty() && !sv->stopping)
                sv->cv.wait(g, [sv] { return !sv->inbox.empty() || sv->stopping; });
            while (!sv->inbox.empty()) {
                sv->held.push_back(std::move(sv->inbox.front()));
                sv->inbox.pop_front();
                OwnedRequest& o = sv->held.back();
                o.q.seeds = o.seeds.data();
                o.q.R = o.R.data();
                core.waiting.push_back({o.q, o.ticket});
            }
            if (sv->stopping && (!sv->drain || (core.waiting.empty() && !core.busy()))) break;
        }
        int rc = core.admit(sink);
        if (!rc && core.busy()) rc = core.step_retire(sink);
        if (rc) {
            std::lock_guard<std::mutex> g(sv->mu);
            sv->fatal = rc;
            sv->fatal_msg = ela_last_error();
            break;
        }
    }
    // requests still waiting or in flight are not served
    std::vector<long long> left;
    for (auto& w : core.waiting) left.push_back(w.second);
    for (auto& pr : core.live) left.push_back(pr.second);
    core.drop();
    std::lock_guard<std::mutex> g(sv->mu);
    for (long long t : left) {
        ela_result o{};
        o.status = -1;
        o.admitted = o.finished = -1.0;
        sv->results[t] = o;
    }
    for (auto& o : sv->inbox) {
        ela_result r{};
        r.status = -1;
        r.admitted = r.finished = -1.0;
        sv->results[o.ticket] = r;
    }
    sv->inbox.clear();
    sv->held.clear();
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int ela_batch_create(ela_ctx* ctx, int mode, ela_batch** out) {
    if (!ctx || !out) return fail(ELA_ERR_BADARG, "null argument");
    *out = nullptr;
    if (mode != ELA_SOLVE_FP64 && mode != ELA_SOLVE_MIXED) return fail(ELA_ERR_BADARG, "mode must be FP64 or MIXED");
    ELA_HIPCHK(hipSetDevice(ctx->device));
    ela_batch* B = new ela_batch();
    B->ctx = ctx;
    B->mode = mode;
    B->tsize = mode == ELA_SOLVE_MIXED ? 4 : 8;
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&B->dflag, sizeof(int)) != hipSuccess) {
        if (B->stream) (void)hipStreamDestroy(B->stream);
        delete B;
        return fail(ELA_ERR_HIP, "stream / scratch allocation failed");
    }
    *out = B;
    return ELA_OK;
}

int ela_batch_destroy(ela_batch* B) {
    if (!B) return ELA_OK;
    (void)hipStreamSynchronize(B->stream);
    std::vector<ela_rve*> ms = B->members;
    for (auto* r : ms) ela_rve_destroy(r);
    for (auto e : B->ev) (void)hipEventDestroy(e);
    if (B->dlists) (void)hipFree(B->dlists);
    if (B->partials) (void)hipFree(B->partials);
    if (B->dnorms) (void)hipFree(B->dnorms);
    if (B->hnorms) (void)hipHostFree(B->hnorms);
    if (B->dflag) (void)hipFree(B->dflag);
    if (B->dbstate) (void)hipFree(B->dbstate);
    if (B->hbstate) (void)hipHostFree(B->hbstate);
    if (B->gexec) (void)hipGraphExecDestroy(B->gexec);  // the stream was synchronized above
    (void)hipStreamDestroy(B->stream);
    delete B;
    return ELA_OK;
}

size_t ela_rve_bytes(int n, int mode, int max_grains, int with_factor) {
    if (n < 4 || max_grains < 1) return 0;
    ela_batch Bt;
    Bt.mode = mode;
    Bt.tsize = mode == ELA_SOLVE_MIXED ? 4 : 8;
    ela_rve r;
    r.B = &Bt;
    r.n = n;
    r.depth = depth_of(n);
    r.max_grains = max_grains;
    r.factor = with_factor != 0;
    const int c = coarsest_of(n);
    const int nunk = 3 * (c + 1) * (c + 1) * (c + 1);
    r.bicg = nunk > kMaxBatchDense;
    r.nunk = r.bicg ? 0 : nunk;
    r.lev.resize(r.depth);
    int m = n;
    for (int l = 0; l < r.depth; ++l, m /= 2) r.lev[l].nodes = (m + 1) * (m + 1) * (m + 1);
    return layout_member(&r, nullptr);
}

int ela_rve_create(ela_batch* B, int n, int max_grains, int with_factor, ela_rve** out) {
    if (!B || !out) return fail(ELA_ERR_BADARG, "null argument");
    *out = nullptr;
    if (n < 4 || n > 1024) return fail(ELA_ERR_BADARG, "member side must be 4..1024 cells");
    if (max_grains < 1 || max_grains > 65536) return fail(ELA_ERR_BADARG, "max_grains must be 1..65536");
    const int c = coarsest_of(n);
    const int nunk = 3 * (c + 1) * (c + 1) * (c + 1);
    ela_rve* r = new ela_rve();
    r->B = B;
    r->n = n;
    r->depth = depth_of(n);
    r->max_grains = max_grains;
    r->factor = with_factor != 0;
    r->bicg = nunk > kMaxBatchDense;
    r->nunk = r->bicg ? 0 : nunk;
    r->lev.resize(r->depth);
    int m = n;
    for (int l = 0; l < r->depth; ++l, m /= 2) {
        RLev& v = r->lev[l];
        v.n = m;
        v.h = 1.0 / m;
        v.L.dim = 3;
        v.L.nx = v.L.ny = v.L.nz = m + 1;
        v.L.px = m + 1;
        v.L.pl = (long long)(m + 1) * (m + 1);
        v.L.cs = v.L.pl * (m + 1);
        v.nodes = (int)v.L.cs;
    }
    r->bytes = layout_member(r, nullptr);
    if (int rc = budget_reserve(B->ctx, r->bytes)) {
        delete r;
        return rc;
    }
    if (hipMalloc(&r->block, r->bytes) != hipSuccess) {
        budget_release(B->ctx, r->bytes);
        delete r;
        return fail(ELA_ERR_OOM, "hipMalloc of a member failed");
    }
    layout_member(r, (char*)r->block);
    if (hipMemsetAsync(r->block, 0, r->bytes, B->stream) != hipSuccess) {
        (void)hipFree(r->block);
        budget_release(B->ctx, r->bytes);
        delete r;
        return fail(ELA_ERR_HIP, "hipMemsetAsync of a member failed");
    }
    if (r->factor) {  // factor 1 until a crack is set
        Crack one{};
        one.width = 1.0;
        one.c = 1.0;
        const RLev& v = r->lev[0];
        hipLaunchKernelGGL(kb_model, dim3((v.nodes + TB - 1) / TB), dim3(TB), 0, B->stream, v.L, v.h, 0,
                           (const long long*)nullptr, 1, (unsigned short*)nullptr, 1, one, r->sc64, r->sc32);
    }
    r->seq = B->seq++;
    B->members.push_back(r);
    *out = r;
    return ELA_OK;
}

int ela_rve_destroy(ela_rve* r) {
    if (!r) return ELA_OK;
    ela_batch* B = r->B;
    (void)hipStreamSynchronize(B->stream);
    auto drop = [r](std::vector<ela_rve*>& v) { v.erase(std::remove(v.begin(), v.end(), r), v.end()); };
    if (std::find(B->active.begin(), B->active.end(), r) != B->active.end()) B->dirty = true;
    drop(B->active);
    drop(B->queued);
    drop(B->members);
    (void)hipFree(r->block);
    budget_release(B->ctx, r->bytes);
    delete r;
    return ELA_OK;
}

size_t ela_rve_bytes_in_use(const ela_rve* r) { return r ? r->bytes : 0; }

int ela_rve_levels(const ela_rve* r) { return r ? r->depth : 0; }

int ela_rve_set_bc_types(ela_rve* r, const int* types) {
    if (int rc = check_rve(r)) return rc;
    if (!types) return fail(ELA_ERR_BADARG, "null types");
    for (int f = 0; f < 26; ++f)
        for (int c = 0; c < 3; ++c) {
            const int t = types[3 * f + c];
            if (t < 0 || t > 3) return fail(ELA_ERR_BADARG, "unknown BC type");
            r->bc.t[f][c] = (unsigned char)t;
        }
    r->prepared = false;
    return ELA_OK;
}

int ela_rve_set_uniform(ela_rve* r, int uniform) {
    if (int rc = check_rve(r)) return rc;
    r->uniform = uniform != 0;
    r->prepared = false;
    return ELA_OK;
}

int ela_rve_set_voronoi_cubic(ela_rve* r, int nseeds, const int64_t* seeds, int q, const double* R, double C11,
                              double C12, double C44) {
    if (int rc = check_rve(r)) return rc;
    if (nseeds < 1 || nseeds > r->max_grains || !seeds || !R || q < 1)
        return fail(ELA_ERR_BADARG, "need 1..max_grains seeds, rotations and q >= 1");
    ela_batch* B = r->B;
    ELA_HIPCHK(hipMemcpyAsync(r->rot, R, 9 * (size_t)nseeds * sizeof(double), hipMemcpyHostToDevice, B->stream));
    ELA_HIPCHK(
        hipMemcpyAsync(r->seeds, seeds, 3 * (size_t)nseeds * sizeof(long long), hipMemcpyHostToDevice, B->stream));
    launch_cubic_table(nseeds, r->rot, C11, C12, C44, r->tab64, B->stream);
    if (r->tab32) launch_convert<double, float>((long long)nseeds * kNC, r->tab64, r->tab32, B->stream);
    const RLev& v = r->lev[0];
    hipLaunchKernelGGL(kb_model, dim3((v.nodes + TB - 1) / TB), dim3(TB), 0, B->stream, v.L, v.h, nseeds,
                       (const long long*)r->seeds, q, r->gid, 0, Crack{}, (double*)nullptr, (float*)nullptr);
    // The host arrays are only borrowed for this call.
    const hipError_t e = hipStreamSynchronize(B->stream);
    if (e != hipSuccess) return fail(ELA_ERR_HIP, std::string("set_voronoi_cubic: ") + hipGetErrorString(e));
    r->ngrains = nseeds;
    r->model = true;
    r->prepared = false;
    return ELA_OK;
}

int ela_rve_set_crack(ela_rve* r, const double p[3], const double nrm[3], const double tv[3], double width, double a,
                      double b, double c) {
    if (int rc = check_rve(r)) return rc;
    if (!r->factor) return fail(ELA_ERR_STATE, "member was created without a degradation factor field");
    if (!p || !nrm || !tv || !(width > 0)) return fail(ELA_ERR_BADARG, "need p, normal, in-plane direction, width > 0");
    Crack cr;
    for (int q = 0; q < 3; ++q) {
        cr.p[q] = p[q];
        cr.nrm[q] = nrm[q];
        cr.tv[q] = tv[q];
    }
    cr.width = width;
    cr.a = a;
    cr.b = b;
    cr.c = c;
    const RLev& v = r->lev[0];
    hipLaunchKernelGGL(kb_model, dim3((v.nodes + TB - 1) / TB), dim3(TB), 0, r->B->stream, v.L, v.h, 0,
                       (const long long*)nullptr, 1, (unsigned short*)nullptr, 1, cr, r->sc64, r->sc32);
    ELA_HIPCHK(hipGetLastError());
    r->prepared = false;
    return ELA_OK;
}

int ela_rve_set_rhs_affine(ela_rve* r, const double G[9]) {
    if (int rc = check_rve(r)) return rc;
    if (!G) return fail(ELA_ERR_BADARG, "null gradient");
    const RLev& v = r->lev[0];
    hipLaunchKernelGGL(kb_rhs_affine, dim3((v.nodes + TB - 1) / TB), dim3(TB), 0, r->B->stream, v.L, v.h, r->bc, G[0],
                       G[1], G[2], G[3], G[4], G[5], G[6], G[7], G[8], r->b64);
    ELA_HIPCHK(hipGetLastError());
    return ELA_OK;
}

// field: 0 = u, 1 = b, 2 = r (test output / residual), 3 = degradation factor (1 comp)
int ela_rve_upload(ela_rve* r, int field, const double* host) {
    if (int rc = check_rve(r)) return rc;
    if (!host || field < 0 || field > 1) return fail(ELA_ERR_BADARG, "upload: field must be 0 (u) or 1 (b)");
    if (r->state == ST_QUEUED || r->state == ST_ACTIVE) return fail(ELA_ERR_STATE, "member is being solved");
    double* dst = field == 0 ? r->u64 : r->b64;
    ELA_HIPCHK(hipMemcpyAsync(dst, host, 3 * (size_t)r->lev[0].nodes * 8, hipMemcpyHostToDevice, r->B->stream));
    ELA_HIPCHK(hipStreamSynchronize(r->B->stream));
    return ELA_OK;
}

int ela_rve_download(const ela_rve* r, int field, double* host) {
    if (int rc = check_rve(r)) return rc;
    if (!host || field < 0 || field > 3) return fail(ELA_ERR_BADARG, "download: field must be 0..3");
    const size_t N = (size_t)r->lev[0].nodes;
    if (field == 3) {
        if (!r->sc64) return fail(ELA_ERR_STATE, "member has no degradation factor");
        ELA_HIPCHK(hipMemcpyAsync(host, r->sc64, N * 8, hipMemcpyDeviceToHost, r->B->stream));
    } else {
        const double* src = field == 0 ? r->u64 : (field == 1 ? r->b64 : r->r64);
        ELA_HIPCHK(hipMemcpyAsync(host, src, 3 * N * 8, hipMemcpyDeviceToHost, r->B->stream));
    }
    ELA_HIPCHK(hipStreamSynchronize(r->B->stream));
    return ELA_OK;
}

int ela_rve_grain_ids(const ela_rve* r, uint16_t* host) {
    if (int rc = check_rve(r)) return rc;
    if (!host) return fail(ELA_ERR_BADARG, "null output");
    ELA_HIPCHK(hipMemcpyAsync(host, r->gid, (size_t)r->lev[0].nodes * 2, hipMemcpyDeviceToHost, r->B->stream));
    ELA_HIPCHK(hipStreamSynchronize(r->B->stream));
    return ELA_OK;
}

int ela_batch_prepare(ela_batch* B, int count, ela_rve* const* rves) {
    if (!B || count < 0 || (count && !rves)) return fail(ELA_ERR_BADARG, "bad arguments");
    return prepare(B, count, rves);
}

// f = A u on the finest level of each member: prec ELA_F64 from u into r;
// ELA_F32: u rounded to FP32, the FP32 operator, result into r.
int ela_batch_fapply(ela_batch* B, int count, ela_rve* const* rves, int prec) {
    if (!B || count < 1 || !rves) return fail(ELA_ERR_BADARG, "bad arguments");
    if (prec == ELA_F32 && B->mode != ELA_SOLVE_MIXED) return fail(ELA_ERR_STATE, "no FP32 fields in an FP64 batch");
    ListBuild lb;
    for (int m = 0; m < count; ++m) {
        if (int rc = check_rve(rves[m])) return rc;
        if (!rves[m]->model) return fail(ELA_ERR_STATE, "model not set");
        list_push(lb, desc_of(rves[m], 0));
    }
    std::vector<char> save;
    save.swap(B->hlists);
    size_t od = 0, om = 0;
    pack_list(B, lb, &od, &om);
    char* dl = nullptr;
    ELA_HIPCHK(hipMalloc((void**)&dl, B->hlists.size()));
    (void)hipMemcpyAsync(dl, B->hlists.data(), B->hlists.size(), hipMemcpyHostToDevice, B->stream);
    B->hlists.swap(save);
    const BoxDesc* D = (const BoxDesc*)(dl + od);
    const int* bm = (const int*)(dl + om);
    const int nb = lb.tile.back();
    if (prec == ELA_F64) {
        hipLaunchKernelGGL(kb_apply64, dim3(nb), dim3(TB), 0, B->stream, D, bm);
    } else {
        hipLaunchKernelGGL((kb_cvt<float>), dim3(nb), dim3(TB), 0, B->stream, D, bm, F_TMP, -1);
        hipLaunchKernelGGL((kb_apply<float>), dim3(nb), dim3(TB), 0, B->stream, D, bm, F_TMP, F_HOLD);
        hipLaunchKernelGGL((kb_cvt<float>), dim3(nb), dim3(TB), 0, B->stream, D, bm, -1, F_HOLD);
    }
    const hipError_t e = hipStreamSynchronize(B->stream);
    (void)hipFree(dl);
    if (e != hipSuccess) return fail(ELA_ERR_HIP, std::string("batch fapply: ") + hipGetErrorString(e));
    return ELA_OK;
}

int ela_batch_submit(ela_batch* B, ela_rve* r, double tol_rel, double tol_abs, int max_cycles) {
    if (!B || !r || r->B != B) return fail(ELA_ERR_BADARG, "bad arguments");
    if (!r->prepared) return fail(ELA_ERR_STATE, "member not prepared (ela_batch_prepare)");
    if (r->state == ST_QUEUED || r->state == ST_ACTIVE) return fail(ELA_ERR_STATE, "member already submitted");
    if (max_cycles < 1) return fail(ELA_ERR_BADARG, "max_cycles must be >= 1");
    r->tol_rel = tol_rel;
    r->tol_abs = tol_abs;
    r->max_cycles = max_cycles;
    r->state = ST_QUEUED;
    r->status = ELA_OK;
    r->seq = B->seq++;
    r->t_submit = now_s();
    B->queued.push_back(r);
    return ELA_OK;
}

int ela_batch_step(ela_batch* B, int* n_active) {
    if (!B) return fail(ELA_ERR_BADARG, "null batch");
    const int rc = step(B);
    if (n_active) *n_active = (int)(B->active.size() + B->queued.size());
    return rc;
}

int ela_batch_solve(ela_batch* B, int count, ela_rve* const* rves, double tol_rel, double tol_abs, int max_cycles) {
    if (!B || count < 0 || (count && !rves)) return fail(ELA_ERR_BADARG, "bad arguments");
    for (int m = 0; m < count; ++m)
        if (int rc = ela_batch_submit(B, rves[m], tol_rel, tol_abs, max_cycles)) return rc;
    for (;;) {
        if (int rc = step(B)) return rc;
        if (B->active.empty() && B->queued.empty()) break;
    }
    int worst = ELA_OK;
    for (int m = 0; m < count; ++m)
        if (rves[m]->status != ELA_OK && worst == ELA_OK) worst = rves[m]->status;
    if (worst == ELA_ERR_NONCONVERGED) return fail(worst, "a member did not converge");
    if (worst == ELA_ERR_NAN) return fail(worst, "a member's residual became non-finite");
    return worst;
}

// state: 0 idle, 1 queued, 2 active, 3 done (status: ELA_OK, ELA_ERR_NONCONVERGED or ELA_ERR_NAN).
int ela_rve_status(const ela_rve* r, int* state, int* status, ela_stats* st) {
    if (int rc = check_rve(r)) return rc;
    if (state) *state = r->state;
    if (status) *status = r->status;
    if (st) {
        st->iters = r->cycles;
        st->converged = (r->state == ST_DONE && r->status == ELA_OK) ? 1 : 0;
        st->mg_levels = r->depth;
        st->mode = r->B->mode;
        st->res0 = r->res0;
        st->res = r->res;
        st->seconds = r->state == ST_DONE ? r->t_done - r->t_submit : 0.0;
        st->setup_seconds = 0;
    }
    return ELA_OK;
}

int ela_batch_set_option(ela_batch* B, int option, int value) {
    if (!B) return fail(ELA_ERR_BADARG, "null batch");
    if (!B->active.empty() || !B->queued.empty()) return fail(ELA_ERR_STATE, "batch is busy");
    if (option == ELA_OPT_MARCH) {
        B->march = value != 0;
    } else if (option == ELA_OPT_KZ) {
        if (value < 1 || value > 1024) return fail(ELA_ERR_BADARG, "planes per marching tile must be 1..1024");
        B->kz = value;
    } else if (option == ELA_OPT_XCD) {
        B->xcd = value != 0;
    } else if (option == ELA_OPT_GRAPH) {
        B->graph_on = value != 0;
    } else if (option == ELA_OPT_GJ) {
        B->gj_one = value != 0;
    } else if (option == ELA_OPT_GRAIN_DIAG) {
        B->gdiag = value != 0;
    } else if (option == ELA_OPT_PROLONG) {
        B->prolong_oct = value != 0;
    } else {
        return fail(ELA_ERR_BADARG, "unknown option");
    }
    B->dirty = true;
    return ELA_OK;
}

int ela_batch_graph_stats(const ela_batch* B, int* captures, int* replays) {
    if (!B) return fail(ELA_ERR_BADARG, "null batch");
    if (captures) *captures = B->graph_captures;
    if (replays) *replays = B->graph_replays;
    return ELA_OK;
}

int ela_batch_autotune(ela_batch* B, int count, ela_rve* const* rves, int nrep, double* ms) {
    if (!B || count < 1 || !rves || nrep < 1) return fail(ELA_ERR_BADARG, "bad arguments");
    if (!B->active.empty() || !B->queued.empty()) return fail(ELA_ERR_STATE, "batch is busy");
    for (int m = 0; m < count; ++m) {
        if (int rc = check_rve(rves[m])) return rc;
        if (!rves[m]->prepared) return fail(ELA_ERR_STATE, "member not prepared");
        if (rves[m]->ngrains > kMarchMaxGrains) return fail(ELA_ERR_STATE, "too many grains for the marching kernels");
    }
    const int kzs[3] = {8, 16, 32};
    double best = 1e300;
    int bkz = B->kz, bx = B->xcd ? 1 : 0;
    hipEvent_t e0, e1;
    ELA_HIPCHK(hipEventCreate(&e0));
    ELA_HIPCHK(hipEventCreate(&e1));
    char* dl = nullptr;
    size_t dl_cap = 0;
    int rc = ELA_OK;
    for (int a = 0; a < 3 && !rc; ++a) {
        // the finest levels of the members, numbered for this tile height
        ListBuild lb;
        for (int m = 0; m < count; ++m) list_push(lb, desc_of(rves[m], 0));
        std::vector<int> bm, bs, bf;
        ela_batch::March mi;
        march_number(lb, 0, count, kzs[a], bm, bs, bf, mi);
        std::vector<char> save;
        save.swap(B->hlists);
        size_t od = 0, om = 0;
        pack_list(B, lb, &od, &om);
        mi.off_bm = pack_ints(B, bm);
        mi.off_bs = pack_ints(B, bs);
        mi.off_bf = pack_ints(B, bf);
        std::vector<char> bytes;
        bytes.swap(B->hlists);
        B->hlists.swap(save);
        if ((rc = ensure_device(B, (void**)&dl, &dl_cap, bytes.size()))) break;
        if (hipMemcpyAsync(dl, bytes.data(), bytes.size(), hipMemcpyHostToDevice, B->stream) != hipSuccess) {
            rc = fail(ELA_ERR_HIP, "autotune upload");
            break;
        }
        char* keep = B->dlists;
        B->dlists = dl;  // march_launch reads the block maps through B->dlists
        const BoxDesc* D = (const BoxDesc*)(dl + od);
        for (int x = 0; x < 2; ++x) {
            const MOps op{F_RES, F_COR, F_TMP, 2.0 / 3.0, nullptr, 0, x};
            auto sweep = [&]() {
                if (B->mode == ELA_SOLVE_MIXED) march_launch<float, 0, float>(B, D, mi, op);
                else march_launch<double, 0, double>(B, D, mi, op);
            };
            sweep();
            (void)hipEventRecord(e0, B->stream);
            for (int r = 0; r < nrep; ++r) sweep();
            (void)hipEventRecord(e1, B->stream);
            (void)hipEventSynchronize(e1);
            float t = 0;
            (void)hipEventElapsedTime(&t, e0, e1);
            const double avg = (double)t / nrep;
            if (ms) ms[2 * a + x] = avg;
            if (avg < best) {
                best = avg;
                bkz = kzs[a];
                bx = x;
            }
        }
        B->dlists = keep;
        (void)hipStreamSynchronize(B->stream);
    }
    if (dl) (void)hipFree(dl);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    B->kz = bkz;
    B->xcd = bx != 0;
    B->dirty = true;
    return ELA_OK;
}

int ela_batch_set_timing(ela_batch* B, int on) {
    if (!B) return fail(ELA_ERR_BADARG, "null batch");
    B->timing = on != 0;
    B->ev_launch.clear();
    if (B->timing && B->ev.empty()) {
        B->ev.resize(16384);
        for (auto& e : B->ev) ELA_HIPCHK(hipEventCreate(&e));
    }
    return ELA_OK;
}

// Smoother launches since timing was switched on: total ms, total nodes swept, launch count.
int ela_batch_smoother_times(ela_batch* B, double* ms, long long* nodes, int* launches) {
    if (!B) return fail(ELA_ERR_BADARG, "null batch");
    ELA_HIPCHK(hipStreamSynchronize(B->stream));
    double t = 0;
    long long nn = 0;
    for (auto& pr : B->ev_launch) {
        float x = 0;
        ELA_HIPCHK(hipEventElapsedTime(&x, B->ev[pr.first], B->ev[pr.first + 1]));
        t += x;
        nn += pr.second;
    }
    if (ms) *ms = t;
    if (nodes) *nodes = nn;
    if (launches) *launches = (int)B->ev_launch.size();
    return ELA_OK;
}

size_t ela_batch_bytes_in_use(const ela_batch* B) {
    if (!B) return 0;
    size_t s = 0;
    for (auto* r : B->members) s += r->bytes;
    return s;
}

// Request loop with continuous batching.  Requests (sorted by arrival time)
// are admitted first-come first-served while the context's HBM budget and
// max_active allow: a member is created (its hierarchy stays resident until
// the request finishes), its model and right-hand side are evaluated on the
// device, the admitted group is prepared together and joins the running set
// at the next cycle.  Every step is one F-cycle over all members in flight;
// finished members are copied out (if the request asks) and freed, which
// lets waiting requests in.  Times are seconds from the start of the call.
int ela_serve(ela_batch* B, int nreq, const ela_request* req, double tol_rel, double tol_abs, int max_cycles,
              int max_active, double max_seconds, ela_result* out, ela_serve_summary* summary) {
    if (!B || nreq < 0 || (nreq && (!req || !out)) || max_active < 1 || max_cycles < 1)
        return fail(ELA_ERR_BADARG, "bad arguments");
    for (int i = 1; i < nreq; ++i)
        if (req[i].arrival < req[i - 1].arrival) return fail(ELA_ERR_BADARG, "requests must be sorted by arrival");
    if (!B->active.empty() || !B->queued.empty() || B->server) return fail(ELA_ERR_STATE, "batch is busy");
    for (int i = 0; i < nreq; ++i) {
        out[i] = ela_result{};
        out[i].status = -1;
        out[i].admitted = out[i].finished = -1.0;
    }
    ServeCore core(B, tol_rel, tol_abs, max_cycles, max_active);
    int next = 0, rc = ELA_OK;
    auto sink = [&](long long i, const ela_result& r) { out[i] = r; };
    while (core.done < nreq) {
        const double t = now_s() - core.t0;
        if (max_seconds > 0 && t > max_seconds) break;
        while (next < nreq && req[next].arrival <= t) {
            core.waiting.push_back({req[next], next});
            ++next;
        }
        if ((rc = core.admit(sink))) break;
        if (core.busy()) {
            if ((rc = core.step_retire(sink))) break;
        } else if (core.waiting.empty() && next < nreq) {
            const double ti0 = now_s();
            const double wait = req[next].arrival - (now_s() - core.t0);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
            core.sm.t_idle += now_s() - ti0;
        }
    }
    if (summary) *summary = core.sm;
    core.drop();  // leave nothing behind: requests still in flight at the time limit are dropped
    return rc;
}

// ------------------------------------------------------------------ online serving
int ela_server_start(ela_batch* B, double tol_rel, double tol_abs, int max_cycles, int max_active, ela_server** out) {
    if (!B || !out || max_cycles < 1 || max_active < 1) return fail(ELA_ERR_BADARG, "bad arguments");
    *out = nullptr;
    if (!B->active.empty() || !B->queued.empty() || B->server) return fail(ELA_ERR_STATE, "batch is busy");
    ela_server* sv = new ela_server();
    sv->B = B;
    sv->core.reset(new ServeCore(B, tol_rel, tol_abs, max_cycles, max_active));
    B->server = sv;
    sv->th = std::thread([sv] { server_loop(sv); });
    *out = sv;
    return ELA_OK;
}

int ela_server_submit(ela_server* sv, const ela_request* req, int64_t* ticket) {
    if (!sv || !req || !ticket) return fail(ELA_ERR_BADARG, "null argument");
    if (req->nseeds < 1 || !req->seeds || !req->R) return fail(ELA_ERR_BADARG, "a request needs seeds and rotations");
    std::lock_guard<std::mutex> g(sv->mu);
    if (sv->stopping) return fail(ELA_ERR_STATE, "server is stopping");
    OwnedRequest o;
    o.q = *req;
    o.seeds.assign(req->seeds, req->seeds + 3 * (size_t)req->nseeds);
    o.R.assign(req->R, req->R + 9 * (size_t)req->nseeds);
    o.ticket = sv->next_ticket++;
    o.t_submit = now_s() - sv->core->t0;
    *ticket = o.ticket;
    sv->inbox.push_back(std::move(o));
    sv->cv.notify_all();
    return ELA_OK;
}

int ela_server_poll(ela_server* sv, int64_t ticket, int* done, ela_result* out) {
    if (!sv || !done) return fail(ELA_ERR_BADARG, "null argument");
    std::lock_guard<std::mutex> g(sv->mu);
    *done = 0;
    auto it = sv->results.find(ticket);
    if (it == sv->results.end()) {
        if (ticket < 1 || ticket >= sv->next_ticket) return fail(ELA_ERR_BADARG, "unknown ticket");
        if (sv->fatal) return fail(sv->fatal, "the serving loop stopped: " + sv->fatal_msg);
        return ELA_OK;
    }
    *done = 1;
    if (out) *out = it->second;
    sv->results.erase(it);
    return ELA_OK;
}

int ela_server_stop(ela_server* sv, int drain) {
    if (!sv) return ELA_OK;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->stopping = true;
        sv->drain = drain != 0;
        sv->cv.notify_all();
    }
    if (sv->th.joinable()) sv->th.join();
    const int rc = sv->fatal;
    const std::string msg = sv->fatal_msg;
    sv->B->server = nullptr;
    delete sv;
    return rc ? fail(rc, "the serving loop stopped: " + msg) : ELA_OK;
}

}  // extern "C"
