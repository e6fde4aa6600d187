// ela_kernels.hip -- gfx950 kernels of the nodal elasticity multigrid and
// their launchers.  The operator math lives in ela_device.hpp.
//
// Launch shapes: node kernels run one thread per node on 64 x 4 thread
// blocks (one wavefront per x row of 64 nodes, so loads along x coalesce),
// grid (ceil(nx/64), ceil(ny/4), nz).  Every kernel checks its node indices
// against the level's node counts before touching memory.
#include "ela_kernels.hpp"

#include <hip/hip_runtime.h>

#include "ela_grain.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <initializer_list>
#include <type_traits>

namespace ela {

namespace {

constexpr int BX = 64, BY = 4;

__device__ __forceinline__ bool node_index(const Lay& L, int& i, int& j, int& k) {
    i = blockIdx.x * BX + threadIdx.x;
    j = blockIdx.y * BY + threadIdx.y;
    k = blockIdx.z;
    return i < L.nx && j < L.ny && k < L.nz;
}

inline dim3 node_grid(const Lay& L) { return dim3((L.nx + BX - 1) / BX, (L.ny + BY - 1) / BY, L.nz); }

// XCD-slab numbering of the stencil kernels on 3D levels.  Hardware block b
// runs on XCD b % 8; XCD x owns the rows [x*rows, (x+1)*rows).  Each plane of
// the slab, nx * rows nodes in row order, is cut into blocks of 256
// consecutive nodes, walked plane by plane: no thread idles on a 64-wide x
// tile past the row end (levels have 2^l + 1 nodes a row, so 64-wide tiles
// would leave a tile column holding one node).  rows == 0 marks the plain
// block grid (node_index).
struct Slab {
    int rows, ntx;
    // node-list order instead (NodeList): thread q of the launch takes node list[q]
    const unsigned long long* list = nullptr;
    long long nlist = 0;
    // 3D: regroup each block's nodes, those with every bit of the mask gsm first
    const unsigned char* gsm = nullptr;
};
inline Slab slab_of(const Lay& L) {
    Slab S;
    S.rows = (L.ny + 7) / 8;
    S.ntx = (int)(((long long)L.nx * S.rows + BX * BY - 1) / (BX * BY));  // blocks per slab plane
    return S;
}
inline dim3 slab_grid(const Lay& L) {
    const Slab S = slab_of(L);
    return dim3((unsigned)(8LL * L.nz * S.ntx));
}
__device__ __forceinline__ bool slab_index(const Lay& L, const Slab S, int& i, int& j, int& k) {
    const int b = blockIdx.x, x = b & 7, r = b >> 3;
    const int bl = r % S.ntx;  // 256 consecutive nodes of a slab plane, planes in turn
    k = r / S.ntx;
    const unsigned t = (unsigned)bl * (BX * BY) + threadIdx.x + BX * threadIdx.y;
    const unsigned jl = t / (unsigned)L.nx;
    i = (int)(t - jl * (unsigned)L.nx);
    j = x * S.rows + (int)jl;
    return (int)jl < S.rows && j < L.ny && k < L.nz;
}

// Node of thread q in a node-list launch (1D grid of BX x BY blocks).
__device__ __forceinline__ bool list_index(const Slab& S, int& i, int& j, int& k) {
    const long long q = (long long)blockIdx.x * (BX * BY) + threadIdx.x + BX * threadIdx.y;
    if (q >= S.nlist) return false;
    const unsigned long long v = S.list[q];
    i = (int)(v & 0x1FFFFFu);
    j = (int)((v >> 21) & 0x1FFFFFu);
    k = (int)(v >> 42);
    return true;
}
// The block's nodes (slab or grid order) regrouped: the interior nodes with
// all three mask bits (constant C around them) take the first threads, the
// others follow, each group in row order.  Whole wavefronts then skip the
// derivative of C while the block keeps its place in the launch order.
// Every thread of the block must call it (it synchronizes).
__device__ __forceinline__ bool group_index(const Lay& L, const Slab& S, int& i, int& j, int& k) {
    const bool valid = S.rows ? slab_index(L, S, i, j, k) : node_index(L, i, j, k);
    const bool in = valid && i > 0 && i < L.nx - 1 && j > 0 && j < L.ny - 1 && k > 0 && k < L.nz - 1;
    const bool pure = in && (S.gsm[(long long)i + L.px * j + L.pl * k] & 7u) == 7u;
    const unsigned long long bp = __ballot(pure), bo = __ballot(valid && !pure);
    __shared__ int wp[BY], wo[BY];
    __shared__ unsigned long long slot[BX * BY];
    const int lane = threadIdx.x, w = threadIdx.y;  // a wavefront is one 64-wide row of the block
    if (lane == 0) {
        wp[w] = __popcll(bp);
        wo[w] = __popcll(bo);
    }
    __syncthreads();
    int np = 0, no = 0, pp = 0, po = 0;
    for (int ww = 0; ww < BY; ++ww) {
        if (ww < w) {
            pp += wp[ww];
            po += wo[ww];
        }
        np += wp[ww];
        no += wo[ww];
    }
    const unsigned long long below = (1ull << lane) - 1;
    const unsigned long long v = (unsigned long long)i | ((unsigned long long)j << 21) | ((unsigned long long)k << 42);
    if (pure) slot[pp + __popcll(bp & below)] = v;
    else if (valid) slot[np + po + __popcll(bo & below)] = v;
    __syncthreads();
    const int tid = lane + BX * w;
    if (tid >= np + no) return false;
    const unsigned long long u = slot[tid];
    i = (int)(u & 0x1FFFFFu);
    j = (int)((u >> 21) & 0x1FFFFFu);
    k = (int)(u >> 42);
    return true;
}

__device__ __forceinline__ bool pick_index(const Lay& L, const Slab& S, int& i, int& j, int& k) {
    if (S.list) return list_index(S, i, j, k);
    if (S.gsm) return group_index(L, S, i, j, k);
    return S.rows ? slab_index(L, S, i, j, k) : node_index(L, i, j, k);
}

// Block max of a non-negative value (NaN propagates: a NaN compares as
// "greater"), written to partials[linear block id]; k_reduce_partials then
// folds the partials.  No atomics, so the result is deterministic.
__device__ __forceinline__ void block_max_store(double* partials, double v) {
    __shared__ double wmax[16];
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o > v || o != o) ? o : v;
    }
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int nw = (blockDim.x * blockDim.y + 63) / 64;
    if ((tid & 63) == 0) wmax[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double m = wmax[0];
        for (int w = 1; w < nw; ++w) m = (wmax[w] > m || wmax[w] != wmax[w]) ? wmax[w] : m;
        partials[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = m;
    }
}

__global__ __launch_bounds__(1024) void k_reduce_partials(const double* __restrict__ partials, int n,
                                                          unsigned long long* norm) {
    __shared__ double sm[1024];
    double m = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const double v = partials[q];
        m = (v > m || v != v) ? v : m;
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double v = sm[threadIdx.x + s];
            sm[threadIdx.x] = (v > sm[threadIdx.x] || v != v) ? v : sm[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *norm = (unsigned long long)__double_as_longlong(sm[0]);
}

template <typename T, int K, int DIM, typename Acc>
__global__ __launch_bounds__(256) void k_fapply(Lay L, Acc A, const T* __restrict__ u, T* __restrict__ f, BCTab bc,
                                                Inv<T> iv, int uniform, Slab S) {
    int i, j, k;
    if (!pick_index(L, S, i, j, k)) return;
    T r[DIM];
    node_apply_acc<T, K, DIM>(L, A, u, i, j, k, bc, iv, uniform != 0, r);
    const long long o = (long long)i + L.px * j + L.pl * k;
#pragma unroll
    for (int p = 0; p < DIM; ++p) f[p * L.cs + o] = r[p];
}

// r = b - A x; optional outputs: r in T, r * scale in TO (the next cycle's
// right-hand side), and the block maxima of |r|.  Every lane reaches the
// block reduction (out-of-range lanes contribute 0).
template <typename T, typename TO, int K, int DIM, typename Acc>
__global__ __launch_bounds__(256) void k_residual(Lay L, Acc A, const T* __restrict__ x,
                                                  const T* __restrict__ b, T* __restrict__ r, TO* __restrict__ rs,
                                                  const double* __restrict__ scale, double* __restrict__ partials,
                                                  BCTab bc, Inv<T> iv, int uniform, Slab S) {
    int i, j, k;
    const bool in = pick_index(L, S, i, j, k);
    double m = 0.0;
    if (in) {
        T f[DIM];
        node_apply_acc<T, K, DIM>(L, A, x, i, j, k, bc, iv, uniform != 0, f);
        const long long o = (long long)i + L.px * j + L.pl * k;
        const double s = rs ? *scale : 1.0;
#pragma unroll
        for (int p = 0; p < DIM; ++p) {
            const T v = b[p * L.cs + o] - f[p];
            if (r) r[p * L.cs + o] = v;
            if (rs) rs[p * L.cs + o] = (TO)((double)v * s);
            const double a = fabs((double)v);
            m = (a > m || a != a) ? a : m;
        }
    }
    if (partials) block_max_store(partials, m);
}

// One weighted-Jacobi sweep x_out = x_in + omega (b - A x_in) / D.
template <typename T, int K, int DIM, typename Acc>
__global__ __launch_bounds__(256) void k_jacobi(Lay L, Acc A, const T* __restrict__ dg, const T* __restrict__ xin,
                                                const T* __restrict__ b, T* __restrict__ xout, BCTab bc, Inv<T> iv,
                                                int uniform, T omega, Slab S) {
    int i, j, k;
    if (!pick_index(L, S, i, j, k)) return;
    T f[DIM];
    node_apply_acc<T, K, DIM>(L, A, xin, i, j, k, bc, iv, uniform != 0, f);
    const long long o = (long long)i + L.px * j + L.pl * k;
    const bool interior = i > 0 && i < L.nx - 1 && j > 0 && j < L.ny - 1 && (DIM == 2 || (k > 0 && k < L.nz - 1));
#pragma unroll
    for (int p = 0; p < DIM; ++p) {
        const long long q = p * L.cs + o;
        xout[q] = jacobi_update(xin[q], omega, b[q] - f[p], jdiag(A, dg, q, o, p, interior));
    }
}

// The first Jacobi sweep from x = 0: x = omega b / D.
template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_jacobi0(Lay L, const T* __restrict__ dg, const T* __restrict__ b,
                                                 T* __restrict__ xout, T omega) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const long long o = (long long)i + L.px * j + L.pl * k;
#pragma unroll
    for (int p = 0; p < DIM; ++p) {
        const long long q = p * L.cs + o;
        xout[q] = jacobi_first(omega, b[q], dg[q]);
    }
}

// The same on a grain-form level: an interior node's diagonal from its
// grain's entry of dgt (3 values a grain, the value the stored field holds
// there: the same bits), the boundary nodes' from dg.
template <typename T>
__global__ __launch_bounds__(256) void k_jacobi0_grain(Lay L, const unsigned short* __restrict__ gid,
                                                       const T* __restrict__ dgt, const T* __restrict__ dg,
                                                       const T* __restrict__ b, T* __restrict__ xout, T omega) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const long long o = (long long)i + L.px * j + L.pl * k;
    const bool interior = i > 0 && i < L.nx - 1 && j > 0 && j < L.ny - 1 && k > 0 && k < L.nz - 1;
    const int g3 = interior ? 3 * (int)gid[o] : 0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const long long q = p * L.cs + o;
        xout[q] = jacobi_first(omega, b[q], interior ? dgt[g3 + p] : dg[q]);
    }
}

template <int K, int DIM>
__global__ __launch_bounds__(256) void k_diag(Lay L, const double* __restrict__ C, double* __restrict__ d64,
                                              float* __restrict__ d32, BCTab bc, double hx, double hy, double hz,
                                              unsigned long long* nan_flag) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const double h[3] = {hx, hy, hz};
    double dg[DIM];
    node_diag<K, DIM>(L, C, i, j, k, bc, h, dg);
    const long long o = (long long)i + L.px * j + L.pl * k;
#pragma unroll
    for (int p = 0; p < DIM; ++p) {
        if (d64) d64[p * L.cs + o] = dg[p];
        if (d32) store_jacobi_diag(d32 + p * L.cs + o, dg[p]);  // the inverse
        if (!(dg[p] == dg[p]) || dg[p] == 0.0) atomicMax(nan_flag, 1ull);
    }
}

// Nodal full weighting from the fine level (Lf) onto the coarse level (Lc), nc components.
template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_restrict(Lay Lc, Lay Lf, int nc, const T* __restrict__ fine,
                                                  T* __restrict__ crse) {
    int I, J, K;
    if (!node_index(Lc, I, J, K)) return;
    const long long oc = (long long)I + Lc.px * J + Lc.pl * K;
    const long long px = Lf.px, pl = Lf.pl;
    for (int n = 0; n < nc; ++n) {
        const T* F = fine + n * Lf.cs + 2LL * I + px * (2LL * J) + pl * (2LL * K);
        crse[n * Lc.cs + oc] =
            restrict_node<T, DIM>(Lc, I, J, K, [F, px, pl](int a, int b, int c) { return F[a + px * b + pl * c]; });
    }
}

// Trilinear nodal prolongation: fine (+)= P crse.
template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_prolong(Lay Lf, Lay Lc, int nc, const T* __restrict__ crse,
                                                 T* __restrict__ fine, int add) {
    int i, j, k;
    if (!node_index(Lf, i, j, k)) return;
    const long long oc = (long long)(i >> 1) + Lc.px * (j >> 1) + Lc.pl * (k >> 1);
    const long long of = (long long)i + Lf.px * j + Lf.pl * k;
    for (int n = 0; n < nc; ++n) {
        const T v = prolong_node<T>(Lc, i, j, k, crse + n * Lc.cs + oc);
        T* fp = fine + n * Lf.cs + of;
        *fp = add ? (*fp + v) : v;
    }
}

// The same in 3D with one thread per coarse node (I, J, K): the up to 8 fine nodes
// (2I + a, 2J + b, 2K + c), a, b, c in {0, 1}, read only the coarse nodes
// (I..I+1, J..J+1, K..K+1), so each coarse value is loaded once per thread
// for 8 fine nodes, and each fine row pair moves in one access.  Each fine
// node is prolong_node's expression (its loads repeat the thread's and are
// merged), so the bits are the per-node kernel's.
template <typename T>
__global__ __launch_bounds__(256) void k_prolong_oct(Lay Lf, Lay Lc, int nc, const T* __restrict__ crse,
                                                     T* __restrict__ fine, int add) {
    using T2 = std::conditional_t<sizeof(T) == 8, double2, float2>;
    const unsigned t = blockIdx.x * 256u + threadIdx.x;  // coarse node of plane K = blockIdx.y
    if (t >= (unsigned)Lc.nx * (unsigned)Lc.ny) return;
    const int I = (int)(t % (unsigned)Lc.nx), J = (int)(t / (unsigned)Lc.nx), K = (int)blockIdx.y;
    const int i = 2 * I, j0 = 2 * J, k0 = 2 * K;
    const bool two = i + 1 < Lf.nx;
    const long long oc = (long long)I + Lc.px * J + Lc.pl * K;
    for (int n = 0; n < nc; ++n) {
        const T* c = crse + n * Lc.cs + oc;
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj) {
                const int j = j0 + dj, k = k0 + dk;
                if (j >= Lf.ny || k >= Lf.nz) continue;
                T* fp = fine + n * Lf.cs + (long long)i + Lf.px * j + Lf.pl * k;
                T v0 = prolong_node<T>(Lc, i, j, k, c);
                if (two) {
                    T v1 = prolong_node<T>(Lc, i + 1, j, k, c);
                    if (add) {
                        const T2 f = *(const T2*)fp;
                        v0 = f.x + v0;
                        v1 = f.y + v1;
                    }
                    T2 o;
                    o.x = v0;
                    o.y = v1;
                    *(T2*)fp = o;
                } else {
                    *fp = add ? (*fp + v0) : v0;
                }
            }
    }
}

// y (+)= a * x over nc components of a level (elementwise on valid nodes).
template <typename TX, typename TY>
__global__ __launch_bounds__(256) void k_axpy(Lay L, int nc, const TX* __restrict__ x, TY* __restrict__ y,
                                              const double* __restrict__ ascale, double a, int assign) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const long long o = (long long)i + L.px * j + L.pl * k;
    const double s = ascale ? a / *ascale : a;
    for (int n = 0; n < nc; ++n) {
        const long long q = n * L.cs + o;
        const double v = (double)x[q] * s;
        y[q] = assign ? (TY)v : (TY)((double)y[q] + v);
    }
}

// out = (double) (T) (cor + hold) over nc components on valid nodes: the last
// correction of an F-cycle (k_axpy's cor += hold) written straight into the
// FP64 preconditioned vector, with k_axpy's rounding.
template <typename T>
__global__ __launch_bounds__(256) void k_add_to64(Lay L, int nc, const T* __restrict__ hold,
                                                  const T* __restrict__ cor, double* __restrict__ out) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const long long o = (long long)i + L.px * j + L.pl * k;
    for (int n = 0; n < nc; ++n) {
        const long long q = n * L.cs + o;
        out[q] = (double)(T)((double)cor[q] + (double)hold[q]);
    }
}

// Operator<Node>::normalize (Operator.cpp:164-195): the reference's lambda
// divides x(i,j,k) -- component 0 -- by diag(i,j,k) -- component 0 -- once
// per component of the operator; the other components are left as they are.
// The diagonal is the FP64 one, rounded to the precision of x.
template <typename T>
__global__ __launch_bounds__(256) void k_normalize(Lay L, int ncomp, const double* __restrict__ dg, T* __restrict__ x) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const long long o = (long long)i + L.px * j + L.pl * k;
    const T d = (T)dg[o];
    T v = x[o];
    for (int n = 0; n < ncomp; ++n) v /= d;
    x[o] = v;
}

template <typename TX, typename TY>
__global__ __launch_bounds__(256) void k_convert(long long n, const TX* __restrict__ x, TY* __restrict__ y) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) y[q] = (TY)x[q];
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = 0ull;
}

// scale = 2^-e with 2^e the power of two at or above |r|inf (exact in any precision).
__global__ void k_scale_from_norm(const unsigned long long* norm, double* scale) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double r = __longlong_as_double((long long)*norm);
        int e = 0;
        if (r > 0.0 && r < 1e300) frexp(r, &e);
        *scale = ldexp(1.0, -e);
    }
}

// ---------------------------------------------------------------- dense coarsest level
// Dense (row-major) operator of the coarsest level: A[r][q] = (A e_q)[r],
// unknowns numbered comp * nodes + dense node index.  One thread per (node, column q).
// Copy column q of the coarsest operator (F = A e_q in the level layout) into
// the dense row-major matrix M (leading dimension ld), unknowns numbered
// comp * nodes + dense node index.
__global__ __launch_bounds__(256) void k_assemble_col(Lay L, int nodes, const double* __restrict__ F, int q,
                                                      long long ld, double* __restrict__ M, int dim) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nodes) return;
    const int i = t % L.nx, j = (t / L.nx) % L.ny, k = t / (L.nx * L.ny);
    const long long o = (long long)i + L.px * j + L.pl * k;
    for (int p = 0; p < dim; ++p) M[(long long)(p * nodes + t) * ld + q] = F[p * L.cs + o];
}

// Unit vector e_q in the level layout (E zeroed beforehand except the previous unit entry).
__global__ void k_unit(Lay L, int nodes, double* __restrict__ E, int qprev, int q) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    auto at = [&](int qq) -> double& {
        const int p = qq / nodes, t = qq % nodes;
        const int i = t % L.nx, j = (t / L.nx) % L.ny, k = t / (L.nx * L.ny);
        return E[p * L.cs + i + L.px * j + L.pl * k];
    };
    if (qprev >= 0) at(qprev) = 0.0;
    at(q) = 1.0;
}

// Every column of the coarsest operator in one launch: E holds nunk unit
// fields (E + q * fstride is e_q in the level layout, set by k_unit_all) and
// thread (node t, column q = blockIdx.y) writes A[p * nodes + t][q] = (A e_q)
// at node t, component p, into M (leading dimension ld) -- the per-node
// operator of the plane kernels, so the same bits as probing column by column.
template <int K, int DIM>
__global__ __launch_bounds__(256) void k_probe_all(Lay L, int nodes, const double* __restrict__ C,
                                                   const double* __restrict__ E, long long fstride, BCTab bc,
                                                   Inv<double> iv, int uniform, double* __restrict__ M, long long ld) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x, q = blockIdx.y;
    if (t >= nodes) return;
    const int i = t % L.nx, j = (t / L.nx) % L.ny, k = t / (L.nx * L.ny);
    double f[DIM];
    node_apply<double, K, DIM>(L, C, E + (long long)q * fstride, i, j, k, bc, iv, uniform != 0, f);
#pragma unroll
    for (int p = 0; p < DIM; ++p) M[(long long)(p * nodes + t) * ld + q] = f[p];
}

// E + q * fstride = e_q for every unknown q (E zeroed beforehand).
__global__ void k_unit_all(Lay L, int nodes, int dim, double* __restrict__ E, long long fstride) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nodes * dim) return;
    const int p = q / nodes, t = q % nodes;
    const int i = t % L.nx, j = (t / L.nx) % L.ny, k = t / (L.nx * L.ny);
    E[(long long)q * fstride + p * L.cs + i + L.px * j + L.pl * k] = 1.0;
}

// Right half of [A | I]: M[r][n + r] = 1.
__global__ void k_identity_right(double* __restrict__ M, int n) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) M[(long long)r * (2 * n) + n + r] = 1.0;
}

// x = A^-1 b on the coarsest level (b, x in the level layout, precision T).
// Every block gathers b into LDS (dense unknown order); each wavefront then
// computes whole rows of A^-1 b with coalesced row reads and a shuffle sum.
template <typename T>
__global__ __launch_bounds__(256) void k_bottom(Lay L, int nodes, int dim, const double* __restrict__ M,
                                                const T* __restrict__ b, T* __restrict__ x) {
    __shared__ double sb[3 * 9 * 9 * 9];
    const int nunk = nodes * dim;
    for (int q = threadIdx.x; q < nunk; q += blockDim.x) {
        const int p = q / nodes, t = q % nodes;
        const int i = t % L.nx, j = (t / L.nx) % L.ny, k = t / (L.nx * L.ny);
        sb[q] = (double)b[p * L.cs + i + L.px * j + L.pl * k];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int r = wave; r < nunk; r += nwaves) {
        const double* row = M + (long long)r * (2 * nunk);  // A^-1 in the left half
        double acc = 0.0;
        for (int q = lane; q < nunk; q += 64) acc += row[q] * sb[q];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) {
            const int p = r / nodes, t = r % nodes;
            const int i = t % L.nx, j = (t / L.nx) % L.ny, k = t / (L.nx * L.ny);
            x[p * L.cs + i + L.px * j + L.pl * k] = (T)acc;
        }
    }
}

// Gauss-Jordan on the dense coarsest operator, one column per launch pair
// (the step bodies are gj_pivot_block / gj_eliminate_row).
__global__ __launch_bounds__(1024) void k_gj_pivot(double* __restrict__ M, int n, int c, int* flag) {
    __shared__ double sval[1024];
    __shared__ int sidx[1024];
    gj_pivot_block(M, n, c, flag, sval, sidx);
}

__global__ __launch_bounds__(256) void k_gj_eliminate(double* __restrict__ M, int n, int c) {
    gj_eliminate_row(M, n, c, blockIdx.x);
}

// Gauss-Jordan of [A | I] (n rows of 2n), one launch a column with the pivot
// search in every block: the pivots of the pivot/eliminate launch pairs (the
// largest |M[r][c]| over the rows not yet pivots, ties to the first in the
// swapped order) and the same operations on every element they keep, so the
// same bits in half the launches.  The row swaps are a permutation (perm_in ->
// perm_out, written by block 0; perm_in null: the identity); the pivot row is
// divided by its pivot where it is read and, in the matrix, by its owner one
// column later (dpiv[c] keeps each pivot); the left half is updated right of
// the pivot column only (the columns left of it are never read again).  So no
// element that one block writes is read by another in the same launch.  Block
// r updates matrix row r.
__global__ __launch_bounds__(256) void k_gj_col(double* __restrict__ M, int n, int c, const int* __restrict__ perm_in,
                                                int* __restrict__ perm_out, double* __restrict__ dpiv,
                                                const double* __restrict__ cand_in, double* __restrict__ cand_out,
                                                int* flag) {
    __shared__ double sval[256];
    __shared__ int sidx[256];
    gj_column(M, n, c, perm_in, perm_out, dpiv, cand_in, cand_out, flag, blockIdx.x, sval, sidx);
}

// After k_gj_col's last column: row i of A^-1 is the right half of matrix row
// perm[i] (the last pivot row over its pivot); block i writes it into the left
// half of row i (no block reads a left half).
__global__ __launch_bounds__(256) void k_gj_final(double* __restrict__ M, int n, const int* __restrict__ perm,
                                                  const double* __restrict__ dpiv) {
    const long long w = 2LL * n;
    const int i = blockIdx.x;
    const double* src = M + (long long)perm[i] * w + n;
    double* dst = M + (long long)i * w;
    const double dl = dpiv[n - 1];
    for (int q = threadIdx.x; q < n; q += 256) dst[q] = i == n - 1 ? src[q] / dl : src[q];
}

// After the per-pivot launches: row r of A^-1 (right half) into the left half.
__global__ __launch_bounds__(256) void k_gj_left(double* __restrict__ M, int n) {
    const long long w = 2LL * n;
    for (int q = threadIdx.x; q < n; q += 256) M[blockIdx.x * w + q] = M[blockIdx.x * w + n + q];
}


// ---------------------------------------------------------------- stiffness evaluation
// Per-grain cubic stiffness (Linear::Cubic::Define, Cubic.H:30-59): the
// unrotated Ctmp, rotated R_pi R_sk Ctmp_ijkl R_qj R_tl, written slot by slot
// in p,q,s,t order so the last tuple of each symmetry class sets the value.
// One thread per grain.
__global__ void k_cubic_table(int ng, const double* __restrict__ R, double C11, double C12, double C44,
                              double* __restrict__ table) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ng) return;
    const double* r = R + 9 * gi;
    double* out = table + 21 * gi;
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q)
            for (int s = 0; s < 3; ++s)
                for (int t = 0; t < 3; ++t) {
                    double acc = 0.0;
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j)
                            for (int k = 0; k < 3; ++k)
                                for (int l = 0; l < 3; ++l) {
                                    double ct;
                                    if (i == j && j == k && k == l) ct = C11;
                                    else if (i == k && j == l) ct = C44;
                                    else if (i == j && k == l) ct = C12;
                                    else ct = 0.0;
                                    acc += r[3 * p + i] * r[3 * s + k] * ct * r[3 * q + j] * r[3 * t + l];
                                }
                    out[mm_slot<3>(p, q, s, t)] = acc;
                }
}

// Nearest Voronoi seed per node by exact int64 distance (ties -> lowest
// index), then the grain's 21 values into the stiffness planes.
__global__ __launch_bounds__(256) void k_voronoi_cubic(Lay L, int nseeds, const long long* __restrict__ seeds, int q,
                                                       const double* __restrict__ table, double* __restrict__ C,
                                                       unsigned short* __restrict__ gid) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const long long x = (long long)i * q, y = (long long)j * q, z = (long long)k * q;
    long long best = -1;
    int bi = 0;
    for (int s = 0; s < nseeds; ++s) {
        const long long dx = x - seeds[3 * s], dy = y - seeds[3 * s + 1], dz = z - seeds[3 * s + 2];
        const long long d = dx * dx + dy * dy + dz * dz;
        if (best < 0 || d < best) { best = d; bi = s; }
    }
    const long long o = (long long)i + L.px * j + L.pl * k;
    for (int c = 0; c < 21; ++c) C[c * L.cs + o] = table[21 * bi + c];
    if (gid) gid[o] = (unsigned short)bi;
}

// ---------------------------------------------------------------- model-field assembly
// Node value of an order-parameter field eta (component c) at node (i, j, k)
// of level L, from one of the sources of EtaSrc: node values, or cell values
// averaged onto the node the way the reference's callers do it.
template <int DIM>
__device__ __forceinline__ double eta_at(const Lay& L, const EtaSrc& e, int c, int i, int j, int k) {
    if (e.avg == ETA_NODES) return e.p[c * e.cs + (long long)i + (long long)L.nx * (j + (long long)L.ny * k)];
    // cells -1..n per direction (one ghost layer), x fastest
    auto cv = [&](int ci, int cj, int ck) {
        const long long kk = DIM == 3 ? (long long)(ck + 1) : 0;
        return e.p[c * e.cs + (long long)(ci + 1) + (long long)e.gx * ((cj + 1) + (long long)e.gy * kk)];
    };
    if (e.avg == ETA_XY4)  // PhaseFieldMicrostructure.cpp:537: the 4 cells of the x-y plane, also in 3D
        return 0.25 * (cv(i, j, k) + cv(i, j - 1, k) + cv(i - 1, j, k) + cv(i - 1, j - 1, k));
    // ETA_CELLS: every cell around the node, times 1/(2+2) or 1/(2+2+4)
    // (PolymerDegradation.cpp:817-827, in its order of summation)
    const double mul = 1.0 / (DIM == 3 ? 8.0 : 4.0);
    double t = cv(i, j, k) + cv(i - 1, j, k);
    t = t + cv(i, j - 1, k) + cv(i - 1, j - 1, k);
    if (DIM == 3) t = t + cv(i, j, k - 1) + cv(i - 1, j, k - 1) + cv(i, j - 1, k - 1) + cv(i - 1, j - 1, k - 1);
    return mul * t;
}

// Checks of DegradeModulus (Degradable/Isotropic.H:74-104) on the node
// values of components [c0, c1): bit 0 a NaN, bit 1 an infinite eta in
// component 0 (the reference's "Infs" test looks at eta[0] only), bit 2 a
// value outside [0, 1] (an infinite component 1 lands here, as in the reference).
template <int DIM>
__global__ __launch_bounds__(256) void k_eta_check(Lay L, EtaSrc e, int c0, int c1, int* __restrict__ flags) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    int f = 0;
    for (int c = c0; c < c1; ++c) {
        const double v = eta_at<DIM>(L, e, c, i, j, k);
        if (v != v) f |= 1;
        else if (c == 0 && (v == INFINITY || v == -INFINITY)) f |= 2;
        else if (v < 0.0 || v > 1.0) f |= 4;
    }
    if (f) atomicOr(flags, f);
}

// lambda = lambda0 (1 - eta_a), mu = mu0 (1 - eta_b) (DegradeModulus; a == b
// for its one-eta forms).
template <int DIM>
__global__ __launch_bounds__(256) void k_degrade_iso(Lay L, EtaSrc e, int ca, int cb, double lambda0, double mu0,
                                                     double* __restrict__ C) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const long long o = (long long)i + L.px * j + L.pl * k;
    const double ea = eta_at<DIM>(L, e, ca, i, j, k);
    const double eb = cb == ca ? ea : eta_at<DIM>(L, e, cb, i, j, k);
    C[o] = lambda0 * (1.0 - ea);
    C[L.cs + o] = mu0 * (1.0 - eb);
}

// Affine::Cubic::Combine (Affine/Cubic.H:70-94) at every node: the weights
// eta_n / sum(eta) of the node-averaged order parameters, each record and F0
// accumulated in model order from zero.  NC stiffness values, DIM x DIM F0.
template <int DIM, int NC>
__global__ __launch_bounds__(256) void k_combine(Lay L, EtaSrc e, int nph, const double* __restrict__ ddw,
                                                 const double* __restrict__ F0s, double* __restrict__ C,
                                                 double* __restrict__ F0out, int* __restrict__ flags) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const long long o = (long long)i + L.px * j + L.pl * k;
    double etasum = 0.;
    for (int n = 0; n < nph; ++n) etasum += eta_at<DIM>(L, e, n, i, j, k);
    double acc[NC], facc[DIM * DIM];
#pragma unroll
    for (int q = 0; q < NC; ++q) acc[q] = 0.0;
#pragma unroll
    for (int q = 0; q < DIM * DIM; ++q) facc[q] = 0.0;
    for (int n = 0; n < nph; ++n) {
        const double w = eta_at<DIM>(L, e, n, i, j, k) / etasum;
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[q] += ddw[n * NC + q] * w;
        if (F0s) {
#pragma unroll
            for (int a = 0; a < DIM; ++a)
#pragma unroll
                for (int b = 0; b < DIM; ++b) facc[a * DIM + b] += F0s[n * 9 + 3 * a + b] * w;
        }
    }
    bool bad = false;
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        C[q * L.cs + o] = acc[q];
        bad |= !(acc[q] == acc[q]);
    }
    if (F0out) {
#pragma unroll
        for (int q = 0; q < DIM * DIM; ++q) F0out[q * L.cs + o] = facc[q];
    }
    if (bad) atomicOr(flags, 1);
}

template <typename T>
__global__ __launch_bounds__(256) void k_pack(Lay L, int nc, const double* __restrict__ dense, T* __restrict__ dev,
                                              int to_dev) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const long long o = (long long)i + L.px * j + L.pl * k;
    const long long d = (long long)i + (long long)L.nx * (j + (long long)L.ny * k);
    const long long nn = (long long)L.nx * L.ny * L.nz;
    for (int n = 0; n < nc; ++n) {
        if (to_dev) dev[n * L.cs + o] = (T)dense[n * nn + d];
        else ((double*)dense)[n * nn + d] = (double)dev[n * L.cs + o];
    }
}

// Plain streaming copy (16 B per lane per step), the calibration probe of the
// HBM byte counters.
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ src, double2* __restrict__ dst, long long n) {
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
        dst[q] = src[q];
}

// The copy with U 16-B loads in flight per lane (each of a wavefront's U
// accesses is one coalesced 1 KiB run), all issued before the stores, and
// optionally non-temporal stores (nothing reads the destination again).
typedef double dvec2 __attribute__((ext_vector_type(2)));
template <int U, bool NT>
__global__ __launch_bounds__(256) void k_copy_u(const dvec2* __restrict__ src, dvec2* __restrict__ dst, long long n) {
    const long long stride = (long long)gridDim.x * 256 * U;
    for (long long base = (long long)blockIdx.x * 256 * U + threadIdx.x; base < n; base += stride) {
        dvec2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long q = base + (long long)u * 256;
            if (q < n) v[u] = src[q];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long q = base + (long long)u * 256;
            if (q < n) {
                if (NT) __builtin_nontemporal_store(v[u], dst + q);
                else dst[q] = v[u];
            }
        }
    }
}

// ---------------------------------------------------------------- flat vector kernels (Krylov)
// Vectors are whole level fields (dim * cstride doubles, padding kept at 0).
// FG blocks of FB threads: 16 waves per CU.  Each lane moves V consecutive
// elements per step (16-B FP64 accesses; V = 4 in the conversions, whose
// FP32 stores are then 16 B too) when every vector is 16-B aligned (all
// fields of the library are), else one.
constexpr int FB = 256, FG = 4096, CH = 8;

template <int V, typename T>
__device__ __forceinline__ void ldv(const T* __restrict__ p, long long q, T (&o)[V]) {
    if constexpr (V == 1) {
        o[0] = p[q];
    } else if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int e = 0; e < V; e += 2) {
            const double2 d = *(const double2*)(p + q + e);
            o[e] = d.x;
            o[e + 1] = d.y;
        }
    } else if constexpr (V == 2) {
        const float2 f = *(const float2*)(p + q);
        o[0] = f.x;
        o[1] = f.y;
    } else {
        static_assert(V == 4, "FP32 accesses of 1, 2 or 4 elements");
        const float4 f = *(const float4*)(p + q);
        o[0] = f.x;
        o[1] = f.y;
        o[2] = f.z;
        o[3] = f.w;
    }
}
template <int V, typename T>
__device__ __forceinline__ void stv(T* __restrict__ p, long long q, const T (&o)[V]) {
    if constexpr (V == 1) {
        p[q] = o[0];
    } else if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int e = 0; e < V; e += 2) *(double2*)(p + q + e) = make_double2(o[e], o[e + 1]);
    } else if constexpr (V == 2) {
        *(float2*)(p + q) = make_float2(o[0], o[1]);
    } else {
        static_assert(V == 4, "FP32 accesses of 1, 2 or 4 elements");
        *(float4*)(p + q) = make_float4(o[0], o[1], o[2], o[3]);
    }
}

// Steps of a flat kernel: lane (block, tid) takes elements [q, q + V) for
// q = (block * FB + tid) * V + r * FG * FB * V while q + V <= n; the last
// n % V elements go to lane 0 of block 0 after its steps, one at a time.
#define ELA_FLAT_LOOP(V, n, q) \
    for (long long q = ((long long)blockIdx.x * FB + threadIdx.x) * (V); q + (V) <= (n); q += (long long)FG * FB * (V))
#define ELA_FLAT_TAIL(V, n, q)                                        \
    if ((V) > 1 && blockIdx.x == 0 && threadIdx.x == 0)                \
        for (long long q = (n) - (n) % (V); q < (n); ++q)

// The steps of k_mdot with NV (compile-time) list vectors, so that every
// load of a step is issued before the first product.
template <typename TB, int V, int NV>
__device__ __forceinline__ void mdot_steps(const double* __restrict__ x, const VecListT<TB>& Y, int v0, long long n,
                                           int self, double (&acc)[CH], double& accs) {
    auto step = [&](long long q, auto vw) {
        constexpr int W = decltype(vw)::value;
        double xq[W];
        TB yq[NV > 0 ? NV : 1][W];
        ldv<W>(x, q, xq);
#pragma unroll
        for (int v = 0; v < NV; ++v) ldv<W>(Y.v[v0 + v], q, yq[v]);
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int e = 0; e < W; ++e) acc[v] += xq[e] * (double)yq[v][e];
        if (self >= 0) {
#pragma unroll
            for (int e = 0; e < W; ++e) accs += xq[e] * xq[e];
        }
    };
    ELA_FLAT_LOOP(V, n, q) step(q, std::integral_constant<int, V>{});
    ELA_FLAT_TAIL(V, n, q) step(q, std::integral_constant<int, 1>{});
}

// Partial dot products of x with up to CH vectors: partials[(v0 + v) * FG + block].
// TB: the list's precision (products and sums in FP64).  self >= 0: also
// (x, x) into partials[self * FG + block], summed as a list entry equal to x is.
template <typename TB, int V>
__global__ __launch_bounds__(FB) void k_mdot(const double* __restrict__ x, VecListT<TB> Y, int v0, int nv, long long n,
                                             double* __restrict__ partials, int self) {
    double acc[CH], accs = 0.0;
#pragma unroll
    for (int v = 0; v < CH; ++v) acc[v] = 0.0;
    switch (nv) {
        case 0: mdot_steps<TB, V, 0>(x, Y, v0, n, self, acc, accs); break;
        case 1: mdot_steps<TB, V, 1>(x, Y, v0, n, self, acc, accs); break;
        case 2: mdot_steps<TB, V, 2>(x, Y, v0, n, self, acc, accs); break;
        case 3: mdot_steps<TB, V, 3>(x, Y, v0, n, self, acc, accs); break;
        case 4: mdot_steps<TB, V, 4>(x, Y, v0, n, self, acc, accs); break;
        case 5: mdot_steps<TB, V, 5>(x, Y, v0, n, self, acc, accs); break;
        case 6: mdot_steps<TB, V, 6>(x, Y, v0, n, self, acc, accs); break;
        case 7: mdot_steps<TB, V, 7>(x, Y, v0, n, self, acc, accs); break;
        default: mdot_steps<TB, V, CH>(x, Y, v0, n, self, acc, accs); break;
    }
    __shared__ double red[FB / 64][CH];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < CH; ++v) {
        double a = acc[v];
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        if (lane == 0) red[w][v] = a;
    }
    __shared__ double reds[FB / 64];
    if (self >= 0) {
        double a = accs;
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        if (lane == 0) reds[w] = a;
    }
    __syncthreads();
    if (threadIdx.x < CH && (int)threadIdx.x < nv) {
        double a = 0.0;
        for (int ww = 0; ww < FB / 64; ++ww) a += red[ww][threadIdx.x];
        partials[(long long)(v0 + threadIdx.x) * FG + blockIdx.x] = a;
    }
    if (self >= 0 && threadIdx.x == CH) {
        double a = 0.0;
        for (int ww = 0; ww < FB / 64; ++ww) a += reds[ww];
        partials[(long long)self * FG + blockIdx.x] = a;
    }
}

// out[v] = sum of the FG partials of vector v, in a fixed order.
__global__ __launch_bounds__(256) void k_mdot_final(const double* __restrict__ partials, double* __restrict__ out) {
    __shared__ double sm[256];
    double a = 0.0;
    for (int q = threadIdx.x; q < FG; q += 256) a += partials[(long long)blockIdx.x * FG + q];
    sm[threadIdx.x] = a;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

// y += sign * sum_v coef[v] mul[v] X_v; NORM: also the block partials of |y|^2
// (the new y), reduced as k_mdot reduces (x, x): partials[block].  OUTF: the
// new y goes to o1 and o2 (may be null) as (float) (osc y) instead of into y.
template <typename TB, int V, bool NORM, bool OUTF = false>
__global__ __launch_bounds__(FB) void k_maxpy(double* __restrict__ y, VecListT<TB> X, int nv,
                                              const double* __restrict__ coef, CoefMul mul, double sign, long long n,
                                              double* __restrict__ partials, double osc, float* __restrict__ o1,
                                              float* __restrict__ o2) {
    double acc = 0.0;
    auto step = [&](long long q, auto vw) {
        constexpr int W = decltype(vw)::value;
        double yq[W], a[W];
        ldv<W>(y, q, yq);
#pragma unroll
        for (int e = 0; e < W; ++e) a[e] = 0.0;
        // CH vectors at a time, every load of a chunk issued before its products
        for (int v0 = 0; v0 < nv; v0 += CH) {
            TB xq[CH][W];
#pragma unroll
            for (int v = 0; v < CH; ++v)
                if (v0 + v < nv) ldv<W>(X.v[v0 + v], q, xq[v]);
#pragma unroll
            for (int v = 0; v < CH; ++v)
                if (v0 + v < nv) {
                    const double c = coef[v0 + v] * mul.f[v0 + v];
#pragma unroll
                    for (int e = 0; e < W; ++e) a[e] += c * (double)xq[v][e];
                }
        }
#pragma unroll
        for (int e = 0; e < W; ++e) {
            yq[e] = yq[e] + sign * a[e];
            if (NORM) acc += yq[e] * yq[e];
        }
        if constexpr (OUTF) {
            float f[W];
#pragma unroll
            for (int e = 0; e < W; ++e) f[e] = (float)(osc * yq[e]);
            stv<W>(o1, q, f);
            if (o2) stv<W>(o2, q, f);
        } else {
            stv<W>(y, q, yq);
        }
    };
    ELA_FLAT_LOOP(V, n, q) step(q, std::integral_constant<int, V>{});
    ELA_FLAT_TAIL(V, n, q) step(q, std::integral_constant<int, 1>{});
    if constexpr (NORM) {
        __shared__ double red[FB / 64];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) red[w] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0.0;
            for (int ww = 0; ww < FB / 64; ++ww) a += red[ww];
            partials[blockIdx.x] = a;
        }
    }
}

// y = a x, and yo = (TO) y when yo is non-null (the next MG cycle's
// right-hand side), flat.  y may be x.
template <typename TO, int V>
__global__ __launch_bounds__(FB) void k_scale_flat_to(double* y, const double* x, double a, TO* __restrict__ yo,
                                                      long long n) {
    auto step = [&](long long q, auto vw) {
        constexpr int W = decltype(vw)::value;
        double v[W];
        TO o[W];
        ldv<W>(x, q, v);
#pragma unroll
        for (int e = 0; e < W; ++e) {
            v[e] = a * v[e];
            o[e] = (TO)v[e];
        }
        stv<W>(y, q, v);
        if (yo) stv<W>(yo, q, o);
    };
    ELA_FLAT_LOOP(V, n, q) step(q, std::integral_constant<int, V>{});
    ELA_FLAT_TAIL(V, n, q) step(q, std::integral_constant<int, 1>{});
}

// v = (float) (a x) and r = v (r may be null), flat (an FP32 Krylov vector and
// the next MG cycle's right-hand side).
template <int V>
__global__ __launch_bounds__(FB) void k_scale_to2f(const double* __restrict__ x, double a, float* __restrict__ v,
                                                   float* __restrict__ r, long long n) {
    auto step = [&](long long q, auto vw) {
        constexpr int W = decltype(vw)::value;
        double xq[W];
        float f[W];
        ldv<W>(x, q, xq);
#pragma unroll
        for (int e = 0; e < W; ++e) f[e] = (float)(a * xq[e]);
        stv<W>(v, q, f);
        if (r) stv<W>(r, q, f);
    };
    ELA_FLAT_LOOP(V, n, q) step(q, std::integral_constant<int, V>{});
    ELA_FLAT_TAIL(V, n, q) step(q, std::integral_constant<int, 1>{});
}

#undef ELA_FLAT_LOOP
#undef ELA_FLAT_TAIL

// Whether the multi-element accesses apply: every pointer is 16-B aligned.
bool vec_ok(std::initializer_list<const void*> ps) {
    for (const void* p : ps)
        if (p && ((uintptr_t)p & 15)) return false;
    return true;
}
template <typename TB>
bool vec_ok(const void* x, const VecListT<TB>& Y, int nv) {
    if (!vec_ok({x})) return false;
    for (int v = 0; v < nv; ++v)
        if ((uintptr_t)Y.v[v] & 15) return false;
    return true;
}

}  // namespace

void launch_mdot(const double* x, const VecList& Y, int nv, long long n, double* partials, double* out,
                 hipStream_t s) {
    const bool p2 = vec_ok(x, Y, nv);
    for (int v0 = 0; v0 < nv; v0 += CH) {
        const int c = std::min(CH, nv - v0);
        if (p2) hipLaunchKernelGGL((k_mdot<double, 2>), dim3(FG), dim3(FB), 0, s, x, Y, v0, c, n, partials, -1);
        else hipLaunchKernelGGL((k_mdot<double, 1>), dim3(FG), dim3(FB), 0, s, x, Y, v0, c, n, partials, -1);
    }
    hipLaunchKernelGGL(k_mdot_final, dim3(nv), dim3(256), 0, s, partials, out);
}

void launch_mdot_self(const double* x, const VecList32& Y, int nv, long long n, double* partials, double* out,
                      hipStream_t s) {
    const bool p2 = vec_ok(x, Y, nv);
    for (int v0 = 0; v0 < nv || v0 == 0; v0 += CH) {
        const int c = std::max(0, std::min(CH, nv - v0)), self = v0 == 0 ? nv : -1;
        if (p2) hipLaunchKernelGGL((k_mdot<float, 2>), dim3(FG), dim3(FB), 0, s, x, Y, v0, c, n, partials, self);
        else hipLaunchKernelGGL((k_mdot<float, 1>), dim3(FG), dim3(FB), 0, s, x, Y, v0, c, n, partials, self);
    }
    hipLaunchKernelGGL(k_mdot_final, dim3(nv + 1), dim3(256), 0, s, partials, out);
}

namespace {
template <typename TB, bool NORM, bool OUTF = false>
void maxpy_any(double* y, const VecListT<TB>& X, int nv, const double* coef, const CoefMul& mul, double sign,
               long long n, double* partials, double* norm2, hipStream_t s, double osc = 1.0, float* o1 = nullptr,
               float* o2 = nullptr) {
    if (vec_ok(y, X, nv) && vec_ok({o1, o2}))
        hipLaunchKernelGGL((k_maxpy<TB, 2, NORM, OUTF>), dim3(FG), dim3(FB), 0, s, y, X, nv, coef, mul, sign, n,
                           partials, osc, o1, o2);
    else
        hipLaunchKernelGGL((k_maxpy<TB, 1, NORM, OUTF>), dim3(FG), dim3(FB), 0, s, y, X, nv, coef, mul, sign, n,
                           partials, osc, o1, o2);
    if (NORM) hipLaunchKernelGGL(k_mdot_final, dim3(1), dim3(256), 0, s, partials, norm2);
}
}  // namespace

void launch_maxpy(double* y, const VecList& X, int nv, const double* coef, double sign, long long n, hipStream_t s) {
    maxpy_any<double, false>(y, X, nv, coef, coef_ones(), sign, n, nullptr, nullptr, s);
}

void launch_maxpy_norm(double* y, const VecList& X, int nv, const double* coef, double sign, long long n,
                       double* partials, double* norm2, hipStream_t s) {
    maxpy_any<double, true>(y, X, nv, coef, coef_ones(), sign, n, partials, norm2, s);
}

void launch_maxpy32(double* y, const VecList32& X, int nv, const double* coef, const CoefMul& mul, double sign,
                    long long n, hipStream_t s) {
    maxpy_any<float, false>(y, X, nv, coef, mul, sign, n, nullptr, nullptr, s);
}

void launch_maxpy_norm32(double* y, const VecList32& X, int nv, const double* coef, const CoefMul& mul,
                         double sign, long long n, double* partials, double* norm2, hipStream_t s) {
    maxpy_any<float, true>(y, X, nv, coef, mul, sign, n, partials, norm2, s);
}

void launch_maxpy_norm32_to2f(const double* y, const VecList32& X, int nv, const double* coef, const CoefMul& mul,
                              double sign, long long n, double osc, float* o1, float* o2, double* partials,
                              double* norm2, hipStream_t s) {
    maxpy_any<float, true, true>(const_cast<double*>(y), X, nv, coef, mul, sign, n, partials, norm2, s, osc, o1, o2);
}

void launch_scale_flat(double* y, const double* x, double a, long long n, hipStream_t s) {
    launch_scale_flat_to<double>(y, x, a, nullptr, n, s);
}

void launch_scale_to2f(const double* x, double a, float* v, float* r, long long n, hipStream_t s) {
    if (vec_ok({x, v, r})) hipLaunchKernelGGL(k_scale_to2f<4>, dim3(FG), dim3(FB), 0, s, x, a, v, r, n);
    else hipLaunchKernelGGL(k_scale_to2f<1>, dim3(FG), dim3(FB), 0, s, x, a, v, r, n);
}

template <typename TO>
void launch_scale_flat_to(double* y, const double* x, double a, TO* yo, long long n, hipStream_t s) {
    if (vec_ok({y, x, yo})) hipLaunchKernelGGL((k_scale_flat_to<TO, 4>), dim3(FG), dim3(FB), 0, s, y, x, a, yo, n);
    else hipLaunchKernelGGL((k_scale_flat_to<TO, 1>), dim3(FG), dim3(FB), 0, s, y, x, a, yo, n);
}
template void launch_scale_flat_to<float>(double*, const double*, double, float*, long long, hipStream_t);
template void launch_scale_flat_to<double>(double*, const double*, double, double*, long long, hipStream_t);

long long mdot_partials(int nv) { return (long long)nv * FG; }

void launch_copy(const void* src, void* dst, long long bytes, hipStream_t s) {
    const long long n = bytes / 16;
    hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, s, (const double2*)src, (double2*)dst, n);
}

// Copy variants of the bandwidth probe: 0 is launch_copy (8192 blocks, one
// 16-B access per lane and step); 1..6 use U = 4 or 8 accesses in flight per
// lane, non-temporal stores, and 2048, 8192 or 32768 blocks.
constexpr int kCopyVariants = 7;
int copy_variants() { return kCopyVariants; }
void launch_copy_variant(int var, const void* src, void* dst, long long bytes, hipStream_t s) {
    const long long n = bytes / 16;
    if (var <= 0) {
        launch_copy(src, dst, bytes, s);
        return;
    }
    const int blocks[3] = {2048, 8192, 32768};
    const int b = blocks[(var - 1) % 3];
    if (var <= 3)
        hipLaunchKernelGGL((k_copy_u<4, true>), dim3(b), dim3(256), 0, s, (const dvec2*)src, (dvec2*)dst, n);
    else
        hipLaunchKernelGGL((k_copy_u<8, true>), dim3(b), dim3(256), 0, s, (const dvec2*)src, (dvec2*)dst, n);
}

// Device-wide barrier probe: each of the launch's blocks takes part in
// `rounds` barriers through one device-scope atomic counter; thread 0 spins
// until every block has arrived, at most spin_max polls a barrier (then it
// raises *err and every block leaves at its next poll or round), so every
// wave ends whatever the blocks' residency.  rounds 0: an empty launch.
__global__ __launch_bounds__(256) void k_barrier_probe(unsigned* cnt, int* err, int rounds, long long spin_max) {
    const unsigned G = gridDim.x;
    for (int r = 1; r <= rounds; ++r) {
        if (threadIdx.x == 0) {
            __threadfence();
            atomicAdd(cnt, 1u);
            long long s = 0;
            while (__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < (unsigned)r * G) {
                if (++s > spin_max) atomicExch(err, 1);
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            }
        }
        __syncthreads();
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    }
}
void launch_barrier_probe(int blocks, unsigned* cnt, int* err, int rounds, hipStream_t s) {
    hipLaunchKernelGGL(k_barrier_probe, dim3(blocks), dim3(256), 0, s, cnt, err, rounds, 1LL << 22);
}

// ------------------------------------------------------------------ launchers
#define ELA_DISPATCH(KIND, DIMV, ...)                                              \
    do {                                                                           \
        if ((DIMV) == 3) {                                                         \
            if ((KIND) == K_ISO) { constexpr int KK = K_ISO, DD = 3; __VA_ARGS__; } \
            else if ((KIND) == K_MM) { constexpr int KK = K_MM, DD = 3; __VA_ARGS__; } \
            else { constexpr int KK = K_DIAG, DD = 3; __VA_ARGS__; }              \
        } else {                                                                   \
            if ((KIND) == K_ISO) { constexpr int KK = K_ISO, DD = 2; __VA_ARGS__; } \
            else if ((KIND) == K_MM) { constexpr int KK = K_MM, DD = 2; __VA_ARGS__; } \
            else { constexpr int KK = K_DIAG, DD = 2; __VA_ARGS__; }              \
        }                                                                          \
    } while (0)

template <typename T>
Inv<T> make_inv(const double h[3]) {
    Inv<T> iv;
    for (int d = 0; d < 3; ++d) {
        iv.ih[d] = (T)(1.0 / h[d]);
        iv.ih2[d] = (T)(0.5 / h[d]);
        iv.ihh[d] = (T)(1.0 / (h[d] * h[d]));
    }
    iv.imix[0] = (T)(0.25 / (h[0] * h[1]));
    iv.imix[1] = (T)(0.25 / (h[0] * h[2]));
    iv.imix[2] = (T)(0.25 / (h[1] * h[2]));
    return iv;
}

// Launch order of the one-thread-per-node plane kernels: a node list (every
// node of the level once, in the list's order), the XCD-slab order (3D), or
// the plain block grid.
template <typename T>
void plane_launch(const Lay& L, const PlaneOpt& po, Slab& S, dim3& grid) {
    if (po.list) {
        S = Slab{0, 0};
        S.list = po.list;
        S.nlist = po.nlist;
        grid = dim3((unsigned)((po.nlist + BX * BY - 1) / (BX * BY)));
        return;
    }
    const bool sl = L.dim == 3;
    S = sl ? slab_of(L) : Slab{0, 0};
    grid = sl ? slab_grid(L) : node_grid(L);
    if (po.group && po.sm && L.dim == 3) S.gsm = po.sm;
}

template <typename T>
void launch_fapply(const Lay& L, int kind, const T* C, const T* u, T* f, const BCTab& bc, const double h[3],
                   int uniform, hipStream_t s, const PlaneOpt& po) {
    const Inv<T> iv = make_inv<T>(h);
    const unsigned char* sm = po.sm;
    Slab S;
    dim3 grid;
    plane_launch<T>(L, po, S, grid);
    if (sm)
        ELA_DISPATCH(kind, L.dim,
                     hipLaunchKernelGGL((k_fapply<T, KK, DD, MaskedPlaneAcc<T>>), grid, dim3(BX, BY), 0, s, L,
                                        (MaskedPlaneAcc<T>{C, L.cs, sm}), u, f, bc, iv, uniform, S));
    else
        ELA_DISPATCH(kind, L.dim,
                     hipLaunchKernelGGL((k_fapply<T, KK, DD, PlaneAcc<T>>), grid, dim3(BX, BY), 0, s, L,
                                        (PlaneAcc<T>{C, L.cs}), u, f, bc, iv, uniform, S));
}

template <typename T, typename TO>
void launch_residual(const Lay& L, int kind, const T* C, const T* x, const T* b, T* r, TO* rs, const double* scale,
                     double* partials, unsigned long long* norm, const BCTab& bc, const double h[3], int uniform,
                     hipStream_t s, const PlaneOpt& po) {
    const Inv<T> iv = make_inv<T>(h);
    const unsigned char* sm = po.sm;
    Slab S;
    dim3 grid;
    plane_launch<T>(L, po, S, grid);
    if (sm)
        ELA_DISPATCH(kind, L.dim,
                     hipLaunchKernelGGL((k_residual<T, TO, KK, DD, MaskedPlaneAcc<T>>), grid, dim3(BX, BY), 0, s, L,
                                        (MaskedPlaneAcc<T>{C, L.cs, sm}), x, b, r, rs, scale,
                                        norm ? partials : nullptr, bc, iv, uniform, S));
    else
        ELA_DISPATCH(kind, L.dim,
                     hipLaunchKernelGGL((k_residual<T, TO, KK, DD, PlaneAcc<T>>), grid, dim3(BX, BY), 0, s, L,
                                        (PlaneAcc<T>{C, L.cs}), x, b, r, rs, scale, norm ? partials : nullptr, bc,
                                        iv, uniform, S));
    if (norm) {
        const dim3 g = grid;
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, s, partials, (int)(g.x * g.y * g.z), norm);
    }
}

long long residual_partials(const Lay& L) {
    const dim3 g = node_grid(L), h = slab_grid(L);
    return std::max((long long)g.x * g.y * g.z, (long long)h.x);
}


template <typename T>
void launch_jacobi(const Lay& L, int kind, const T* C, const T* dg, const T* xin, const T* b, T* xout,
                   const BCTab& bc, const double h[3], int uniform, double omega, hipStream_t s,
                   const PlaneOpt& po) {
    const Inv<T> iv = make_inv<T>(h);
    const unsigned char* sm = po.sm;
    Slab S;
    dim3 grid;
    plane_launch<T>(L, po, S, grid);
    if (sm)
        ELA_DISPATCH(kind, L.dim,
                     hipLaunchKernelGGL((k_jacobi<T, KK, DD, MaskedPlaneAcc<T>>), grid, dim3(BX, BY), 0, s, L,
                                        (MaskedPlaneAcc<T>{C, L.cs, sm}), dg, xin, b, xout, bc, iv, uniform,
                                        (T)omega, S));
    else
        ELA_DISPATCH(kind, L.dim,
                     hipLaunchKernelGGL((k_jacobi<T, KK, DD, PlaneAcc<T>>), grid, dim3(BX, BY), 0, s, L,
                                        (PlaneAcc<T>{C, L.cs}), dg, xin, b, xout, bc, iv, uniform, (T)omega, S));
}

// Same-direction mask of a plane-form level (PlaneAcc::same_dir): bit d of an
// interior node is set when all nc stiffness planes hold bitwise-equal values
// at its two d-neighbours; boundary nodes get 0.  A wavefront (64 nodes along
// x) skips the d-derivative of C only when all its interior nodes have bit d,
// so stats[0] counts the wavefronts with an interior node and stats[1] the
// (wavefront, d) pairs that skip.
template <int DIM>
__global__ __launch_bounds__(256) void k_same_mask(Lay L, int nc, const double* __restrict__ C,
                                                   unsigned char* __restrict__ sm, unsigned long long* stats) {
    int i, j, k;
    const bool valid = node_index(L, i, j, k);
    const bool in = valid && i > 0 && i < L.nx - 1 && j > 0 && j < L.ny - 1 && (DIM == 2 || (k > 0 && k < L.nz - 1));
    const long long o = (long long)i + L.px * j + L.pl * k;
    unsigned m = 0;
    if (in) {
        const long long st[3] = {1, L.px, L.pl};
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            bool eq = true;
            for (int q = 0; q < nc && eq; ++q) eq = C[q * L.cs + o + st[d]] == C[q * L.cs + o - st[d]];
            m |= (eq ? 1u : 0u) << d;
        }
    }
    if (valid) sm[o] = (unsigned char)m;
    const unsigned long long wi = __ballot(in);
    unsigned long long skip = 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) skip += (wi != 0 && __ballot(in && ((m >> d) & 1)) == wi) ? 1 : 0;
    if (threadIdx.x == 0 && wi != 0) {  // one lane per wavefront (a wavefront is one 64-wide row)
        atomicAdd(stats, 1ull);
        atomicAdd(stats + 1, skip);
    }
}

void launch_same_mask(const Lay& L, int nc, const double* C, unsigned char* sm, unsigned long long* stats,
                      hipStream_t s) {
    if (L.dim == 3) hipLaunchKernelGGL(k_same_mask<3>, node_grid(L), dim3(BX, BY), 0, s, L, nc, C, sm, stats);
    else hipLaunchKernelGGL(k_same_mask<2>, node_grid(L), dim3(BX, BY), 0, s, L, nc, C, sm, stats);
}

// Node list of a plane level (PlaneOpt::list): the interior nodes with every
// direction bit of the mask (C constant across them: no derivative of C to
// evaluate) first, then every other node, each part in block order (64 x 4
// patches of node_grid, row-major inside), so the wavefronts of the first part
// all skip.  Entries pack (i, j, k) in 21 bits each.
template <int DIM>
__device__ __forceinline__ bool node_pure(const Lay& L, const unsigned char* sm, int i, int j, int k) {
    const bool in = i > 0 && i < L.nx - 1 && j > 0 && j < L.ny - 1 && (DIM == 2 || (k > 0 && k < L.nz - 1));
    const unsigned full = (1u << DIM) - 1;
    return in && (sm[(long long)i + L.px * j + L.pl * k] & full) == full;
}

__device__ __forceinline__ long long grid_block() {
    return blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
}

// cnt[2 b], cnt[2 b + 1]: the block's nodes of the first and of the second part.
template <int DIM>
__global__ __launch_bounds__(256) void k_list_count(Lay L, const unsigned char* __restrict__ sm,
                                                    int* __restrict__ cnt) {
    int i, j, k;
    const bool valid = node_index(L, i, j, k);
    const bool pure = valid && node_pure<DIM>(L, sm, i, j, k);
    const unsigned long long bp = __ballot(pure), bv = __ballot(valid);
    __shared__ int wp[BY], wv[BY];
    if (threadIdx.x == 0) {  // a wavefront is one 64-wide row of the block
        wp[threadIdx.y] = __popcll(bp);
        wv[threadIdx.y] = __popcll(bv);
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        int np = 0, nv = 0;
        for (int w = 0; w < BY; ++w) {
            np += wp[w];
            nv += wv[w];
        }
        const long long b = grid_block();
        cnt[2 * b] = np;
        cnt[2 * b + 1] = nv - np;
    }
}

// Exclusive offsets of every block in both parts (one block of 1024 threads,
// each scanning a contiguous run of blocks); totals = {first part, second part}.
__global__ __launch_bounds__(1024) void k_list_scan(int nb, const int* __restrict__ cnt, int* __restrict__ off,
                                                    long long* __restrict__ totals) {
    __shared__ long long sp[1024], sr[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024, b0 = min(nb, t * per), b1 = min(nb, b0 + per);
    long long p = 0, r = 0;
    for (int b = b0; b < b1; ++b) {
        p += cnt[2 * b];
        r += cnt[2 * b + 1];
    }
    sp[t] = p;
    sr[t] = r;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // inclusive scan of the runs' sums
        const long long ap = t >= d ? sp[t - d] : 0, ar = t >= d ? sr[t - d] : 0;
        __syncthreads();
        sp[t] += ap;
        sr[t] += ar;
        __syncthreads();
    }
    long long ep = sp[t] - p, er = sr[t] - r;
    for (int b = b0; b < b1; ++b) {
        off[2 * b] = (int)ep;
        off[2 * b + 1] = (int)er;
        ep += cnt[2 * b];
        er += cnt[2 * b + 1];
    }
    if (t == 1023) {
        totals[0] = sp[1023];
        totals[1] = sr[1023];
    }
}

template <int DIM>
__global__ __launch_bounds__(256) void k_list_scatter(Lay L, const unsigned char* __restrict__ sm,
                                                      const int* __restrict__ off, const long long* __restrict__ totals,
                                                      unsigned long long* __restrict__ list) {
    int i, j, k;
    const bool valid = node_index(L, i, j, k);
    const bool pure = valid && node_pure<DIM>(L, sm, i, j, k);
    const unsigned long long bp = __ballot(pure), br = __ballot(valid && !pure);
    const unsigned long long below = (1ull << threadIdx.x) - 1;  // lanes of this row before this one
    __shared__ int wp[BY], wr[BY];
    if (threadIdx.x == 0) {
        wp[threadIdx.y] = __popcll(bp);
        wr[threadIdx.y] = __popcll(br);
    }
    __syncthreads();
    if (!valid) return;
    int op = 0, orr = 0;
    for (int w = 0; w < (int)threadIdx.y; ++w) {
        op += wp[w];
        orr += wr[w];
    }
    const long long b = grid_block();
    const unsigned long long v =
        (unsigned long long)i | ((unsigned long long)j << 21) | ((unsigned long long)k << 42);
    if (pure) list[off[2 * b] + op + __popcll(bp & below)] = v;
    else list[totals[0] + off[2 * b + 1] + orr + __popcll(br & below)] = v;
}

long long node_list_blocks(const Lay& L) {
    const dim3 g = node_grid(L);
    return (long long)g.x * g.y * g.z;
}

void launch_node_list(const Lay& L, const unsigned char* sm, int* cnt, long long* totals, unsigned long long* list,
                      hipStream_t s) {
    const dim3 g = node_grid(L);
    const int nb = (int)node_list_blocks(L);
    int* off = cnt + 2LL * nb;
    if (L.dim == 3) {
        hipLaunchKernelGGL(k_list_count<3>, g, dim3(BX, BY), 0, s, L, sm, cnt);
        hipLaunchKernelGGL(k_list_scan, dim3(1), dim3(1024), 0, s, nb, cnt, off, totals);
        hipLaunchKernelGGL(k_list_scatter<3>, g, dim3(BX, BY), 0, s, L, sm, off, totals, list);
    } else {
        hipLaunchKernelGGL(k_list_count<2>, g, dim3(BX, BY), 0, s, L, sm, cnt);
        hipLaunchKernelGGL(k_list_scan, dim3(1), dim3(1024), 0, s, nb, cnt, off, totals);
        hipLaunchKernelGGL(k_list_scatter<2>, g, dim3(BX, BY), 0, s, L, sm, off, totals, list);
    }
}

template <typename T>
void launch_jacobi0(const Lay& L, const T* dg, const T* b, T* xout, double omega, hipStream_t s,
                    const unsigned short* gid, const T* dgt) {
    if (L.dim == 3 && gid && dgt)
        hipLaunchKernelGGL((k_jacobi0_grain<T>), node_grid(L), dim3(BX, BY), 0, s, L, gid, dgt, dg, b, xout,
                           (T)omega);
    else if (L.dim == 3) hipLaunchKernelGGL((k_jacobi0<T, 3>), node_grid(L), dim3(BX, BY), 0, s, L, dg, b, xout, (T)omega);
    else hipLaunchKernelGGL((k_jacobi0<T, 2>), node_grid(L), dim3(BX, BY), 0, s, L, dg, b, xout, (T)omega);
}

void launch_diag(const Lay& L, int kind, const double* C, double* d64, float* d32, const BCTab& bc,
                 const double h[3], unsigned long long* flag, hipStream_t s) {
    ELA_DISPATCH(kind, L.dim,
                 hipLaunchKernelGGL((k_diag<KK, DD>), node_grid(L), dim3(BX, BY), 0, s, L, C, d64, d32, bc, h[0],
                                    h[1], h[2], flag));
}

// Injection from the finer level: crse(I, J, K) = fine(2I, 2J, 2K), nc
// components (coarse grain ids and stiffness of a grain-form hierarchy).
template <typename T>
__global__ __launch_bounds__(256) void k_inject_level(Lay Lc, Lay Lf, int nc, const T* __restrict__ fine,
                                                      T* __restrict__ crse) {
    int I, J, K;
    if (!node_index(Lc, I, J, K)) return;
    const long long oc = (long long)I + Lc.px * J + Lc.pl * K;
    const long long of = 2LL * I + Lf.px * (2LL * J) + Lf.pl * (Lc.dim == 3 ? 2LL * K : 0LL);
    for (int n = 0; n < nc; ++n) crse[n * Lc.cs + oc] = fine[n * Lf.cs + of];
}

template <typename T>
void launch_inject_level(const Lay& Lc, const Lay& Lf, int nc, const T* fine, T* crse, hipStream_t s) {
    hipLaunchKernelGGL((k_inject_level<T>), node_grid(Lc), dim3(BX, BY), 0, s, Lc, Lf, nc, fine, crse);
}
template void launch_inject_level<double>(const Lay&, const Lay&, int, const double*, double*, hipStream_t);
template void launch_inject_level<unsigned short>(const Lay&, const Lay&, int, const unsigned short*,
                                                  unsigned short*, hipStream_t);

template <typename T>
void launch_restrict(const Lay& Lc, const Lay& Lf, int nc, const T* fine, T* crse, hipStream_t s) {
    if (Lc.dim == 3) hipLaunchKernelGGL((k_restrict<T, 3>), node_grid(Lc), dim3(BX, BY), 0, s, Lc, Lf, nc, fine, crse);
    else hipLaunchKernelGGL((k_restrict<T, 2>), node_grid(Lc), dim3(BX, BY), 0, s, Lc, Lf, nc, fine, crse);
}

template <typename T>
void launch_prolong(const Lay& Lf, const Lay& Lc, int nc, const T* crse, T* fine, int add, hipStream_t s) {
    // The octant kernel: fine levels of 129+ nodes a row (measured, FP32: a
    // 257^3 level takes 193 us with a thread per fine node, 159 us with a
    // thread per fine x-pair, 106 us with a thread per coarse node; 129^3:
    // 26, 23, 17 us; at 65^3 and below the per-node kernel's shorter
    // dependency chain is as fast or faster), pair-aligned fine accesses (even
    // pitches, a 2-element-aligned base).
    const bool pairs = Lf.dim == 3 && Lf.nx >= 129 && Lf.px % 2 == 0 && Lf.pl % 2 == 0 && Lf.cs % 2 == 0 &&
                       (uintptr_t)fine % (2 * sizeof(T)) == 0;
    if (pairs) {
        const unsigned nt = (unsigned)Lc.nx * (unsigned)Lc.ny;
        hipLaunchKernelGGL((k_prolong_oct<T>), dim3((nt + 255) / 256, Lc.nz), dim3(256), 0, s, Lf, Lc, nc, crse,
                           fine, add);
    } else if (Lf.dim == 3) {
        hipLaunchKernelGGL((k_prolong<T, 3>), node_grid(Lf), dim3(BX, BY), 0, s, Lf, Lc, nc, crse, fine, add);
    } else {
        hipLaunchKernelGGL((k_prolong<T, 2>), node_grid(Lf), dim3(BX, BY), 0, s, Lf, Lc, nc, crse, fine, add);
    }
}

template <typename TX, typename TY>
void launch_axpy(const Lay& L, int nc, const TX* x, TY* y, const double* ascale, double a, int assign, hipStream_t s) {
    hipLaunchKernelGGL((k_axpy<TX, TY>), node_grid(L), dim3(BX, BY), 0, s, L, nc, x, y, ascale, a, assign);
}

template <typename T>
void launch_add_to64(const Lay& L, int nc, const T* hold, const T* cor, double* out, hipStream_t s) {
    hipLaunchKernelGGL((k_add_to64<T>), node_grid(L), dim3(BX, BY), 0, s, L, nc, hold, cor, out);
}
template void launch_add_to64<float>(const Lay&, int, const float*, const float*, double*, hipStream_t);
template void launch_add_to64<double>(const Lay&, int, const double*, const double*, double*, hipStream_t);

template <typename T>
void launch_normalize(const Lay& L, int ncomp, const double* dg, T* x, hipStream_t s) {
    hipLaunchKernelGGL((k_normalize<T>), node_grid(L), dim3(BX, BY), 0, s, L, ncomp, dg, x);
}
template void launch_normalize<double>(const Lay&, int, const double*, double*, hipStream_t);
template void launch_normalize<float>(const Lay&, int, const double*, float*, hipStream_t);

template <typename TX, typename TY>
void launch_convert(long long n, const TX* x, TY* y, hipStream_t s) {
    const int bs = 256;
    hipLaunchKernelGGL((k_convert<TX, TY>), dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, n, x, y);
}

void launch_zero_u64(unsigned long long* p, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}

void launch_scale_from_norm(const unsigned long long* norm, double* scale, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_from_norm, dim3(1), dim3(64), 0, s, norm, scale);
}

void launch_assemble_col(const Lay& L, int nodes, const double* F, int q, long long ld, double* M, int dim,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_assemble_col, dim3((nodes + 255) / 256), dim3(256), 0, s, L, nodes, F, q, ld, M, dim);
}

void launch_probe_all(const Lay& L, int kind, int nodes, const double* C, double* E, long long fstride,
                      const BCTab& bc, const double h[3], int uniform, double* M, long long ld, hipStream_t s) {
    const int n = nodes * L.dim;
    hipLaunchKernelGGL(k_unit_all, dim3((n + 255) / 256), dim3(256), 0, s, L, nodes, L.dim, E, fstride);
    const Inv<double> iv = make_inv<double>(h);
    const dim3 grid((nodes + 255) / 256, n);
    ELA_DISPATCH(kind, L.dim,
                 hipLaunchKernelGGL((k_probe_all<KK, DD>), grid, dim3(256), 0, s, L, nodes, C, E, fstride, bc, iv,
                                    uniform, M, ld));
}

void launch_unit(const Lay& L, int nodes, double* E, int qprev, int q, hipStream_t s) {
    hipLaunchKernelGGL(k_unit, dim3(1), dim3(64), 0, s, L, nodes, E, qprev, q);
}

void launch_identity_right(double* M, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_identity_right, dim3((n + 255) / 256), dim3(256), 0, s, M, n);
}

void launch_gauss_jordan(double* M, int n, int* flag, hipStream_t s) {
    for (int c = 0; c < n; ++c) {
        hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(1024), 0, s, M, n, c, flag);
        hipLaunchKernelGGL(k_gj_eliminate, dim3(n), dim3(256), 0, s, M, n, c);
    }
    hipLaunchKernelGGL(k_gj_left, dim3(n), dim3(256), 0, s, M, n);
}

void launch_gauss_jordan_col(double* M, int n, int* flag, int* perm, double* dpiv, double* cand, hipStream_t s) {
    for (int c = 0; c < n; ++c)
        hipLaunchKernelGGL(k_gj_col, dim3(n), dim3(256), 0, s, M, n, c, c ? perm + (size_t)((c - 1) & 1) * n : nullptr,
                           perm + (size_t)(c & 1) * n, dpiv, c ? cand + (size_t)((c - 1) & 1) * n : nullptr,
                           cand + (size_t)(c & 1) * n, flag);
    hipLaunchKernelGGL(k_gj_final, dim3(n), dim3(256), 0, s, M, n, perm + (size_t)((n - 1) & 1) * n, dpiv);
}

template <typename T>
void launch_bottom(const Lay& L, int nodes, int dim, const double* M, const T* b, T* x, hipStream_t s) {
    const int nunk = nodes * dim;
    const int blocks = std::min(64, (nunk + 3) / 4);  // 4 wavefronts per block
    hipLaunchKernelGGL((k_bottom<T>), dim3(blocks), dim3(256), 0, s, L, nodes, dim, M, b, x);
}

void launch_cubic_table(int ng, const double* R, double C11, double C12, double C44, double* table, hipStream_t s) {
    hipLaunchKernelGGL(k_cubic_table, dim3((ng + 63) / 64), dim3(64), 0, s, ng, R, C11, C12, C44, table);
}

void launch_voronoi_cubic(const Lay& L, int nseeds, const long long* seeds, int q, const double* table, double* C,
                          unsigned short* gid, hipStream_t s) {
    hipLaunchKernelGGL(k_voronoi_cubic, node_grid(L), dim3(BX, BY), 0, s, L, nseeds, seeds, q, table, C, gid);
}

// ---------------------------------------------------------------- BiCGStab bottom solve
// Flat reductions over n elements of one precision: out[0] = (a, b), out[1] =
// (c, d), out[2] = max |e| (each term skipped when its inputs are null), in FP64,
// deterministic (block partials, then one block folds them in block order).
template <typename T>
__global__ __launch_bounds__(256) void k_bicg_reduce(long long n, const T* __restrict__ a, const T* __restrict__ b,
                                                     const T* __restrict__ c, const T* __restrict__ d,
                                                     const T* __restrict__ e, double* __restrict__ partials) {
    __shared__ double s0[256], s1[256], s2[256];
    double x0 = 0, x1 = 0, x2 = 0;
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
        if (a) x0 += (double)a[q] * (double)b[q];
        if (c) x1 += (double)c[q] * (double)d[q];
        if (e) {
            const double v = fabs((double)e[q]);
            x2 = (v > x2 || v != v) ? v : x2;
        }
    }
    s0[threadIdx.x] = x0;
    s1[threadIdx.x] = x1;
    s2[threadIdx.x] = x2;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            s0[threadIdx.x] += s0[threadIdx.x + st];
            s1[threadIdx.x] += s1[threadIdx.x + st];
            const double o = s2[threadIdx.x + st];
            if (o > s2[threadIdx.x] || o != o) s2[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x] = s0[0];
        partials[3 * blockIdx.x + 1] = s1[0];
        partials[3 * blockIdx.x + 2] = s2[0];
    }
}

__global__ __launch_bounds__(256) void k_bicg_fold(const double* __restrict__ partials, int nb, double* __restrict__ out) {
    __shared__ double s0[256], s1[256], s2[256];
    double x0 = 0, x1 = 0, x2 = 0;
    for (int q = threadIdx.x; q < nb; q += 256) {
        x0 += partials[3 * q];
        x1 += partials[3 * q + 1];
        const double v = partials[3 * q + 2];
        x2 = (v > x2 || v != v) ? v : x2;
    }
    s0[threadIdx.x] = x0;
    s1[threadIdx.x] = x1;
    s2[threadIdx.x] = x2;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            s0[threadIdx.x] += s0[threadIdx.x + st];
            s1[threadIdx.x] += s1[threadIdx.x + st];
            const double o = s2[threadIdx.x + st];
            if (o > s2[threadIdx.x] || o != o) s2[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s0[0];
        out[1] = s1[0];
        out[2] = s2[0];
    }
}

// The vector updates of one BiCGStab step (coefficients from the host):
// op 0: p = r + beta (p - omega v);  op 1: s = r - alpha v;
// op 2: x += alpha p + omega s, r = s - omega t;  op 3: x += alpha p.
// BiCGStab vector steps, right-preconditioned by the Jacobi diagonal dg of the
// level (inv: dg holds the inverse, FP32 levels): op 0 p = r + beta (p -
// omega v) and p^ = D^-1 p; op 1 s = r - alpha v and s^ = D^-1 s; op 2 x +=
// alpha p^ + omega s^, r = s - omega t; op 3 x += alpha p^.  Elements of the
// layout's padding get p^ = s^ = 0.
template <typename T>
__device__ __forceinline__ T bicg_precond(const Lay& L, long long q, const T* __restrict__ dg, int inv, T y) {
    const long long w = q % L.cs;
    const long long k = w / L.pl, rr = w - k * L.pl, j = rr / L.px, i = rr - j * L.px;
    if (i >= L.nx || j >= L.ny || k >= L.nz) return T(0);
    return inv ? (T)(y * dg[q]) : (T)(y / dg[q]);
}
template <typename T>
__global__ __launch_bounds__(256) void k_bicg_update(long long n, int op, double alpha, double beta, double omega,
                                                     T* __restrict__ x, T* __restrict__ r, T* __restrict__ p,
                                                     const T* __restrict__ v, T* __restrict__ sv,
                                                     const T* __restrict__ t, const T* __restrict__ dg, int inv,
                                                     T* __restrict__ ph, T* __restrict__ sh, Lay L) {
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
        if (op == 0) {
            const T pq = (T)((double)r[q] + beta * ((double)p[q] - omega * (double)v[q]));
            p[q] = pq;
            ph[q] = bicg_precond(L, q, dg, inv, pq);
        } else if (op == 1) {
            const T sq = (T)((double)r[q] - alpha * (double)v[q]);
            sv[q] = sq;
            sh[q] = bicg_precond(L, q, dg, inv, sq);
        } else if (op == 2) {
            x[q] = (T)((double)x[q] + alpha * (double)ph[q] + omega * (double)sh[q]);
            r[q] = (T)((double)sv[q] - omega * (double)t[q]);
        } else {
            x[q] = (T)((double)x[q] + alpha * (double)ph[q]);
        }
    }
}

constexpr int kBicgBlocks = 512;
long long bicg_partials() { return 3LL * kBicgBlocks; }

template <typename T>
void launch_bicg_reduce(long long n, const T* a, const T* b, const T* c, const T* d, const T* e, double* partials,
                        double* out, hipStream_t s) {
    const int nb = (int)std::min<long long>(kBicgBlocks, (n + 255) / 256);
    hipLaunchKernelGGL(k_bicg_reduce<T>, dim3(nb), dim3(256), 0, s, n, a, b, c, d, e, partials);
    hipLaunchKernelGGL(k_bicg_fold, dim3(1), dim3(256), 0, s, (const double*)partials, nb, out);
}

template <typename T>
void launch_bicg_update(long long n, int op, double alpha, double beta, double omega, T* x, T* r, T* p, const T* v,
                        T* sv, const T* t, const T* dg, T* ph, T* sh, const Lay& L, hipStream_t s) {
    const int nb = (int)std::min<long long>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_bicg_update<T>, dim3(nb), dim3(256), 0, s, n, op, alpha, beta, omega, x, r, p, v, sv, t, dg,
                       sizeof(T) == 4 ? 1 : 0, ph, sh, L);
}

template void launch_bicg_reduce<double>(long long, const double*, const double*, const double*, const double*,
                                         const double*, double*, double*, hipStream_t);
template void launch_bicg_reduce<float>(long long, const float*, const float*, const float*, const float*,
                                        const float*, double*, double*, hipStream_t);
template void launch_bicg_update<double>(long long, int, double, double, double, double*, double*, double*,
                                         const double*, double*, const double*, const double*, double*, double*,
                                         const Lay&, hipStream_t);
template void launch_bicg_update<float>(long long, int, double, double, double, float*, float*, float*,
                                        const float*, float*, const float*, const float*, float*, float*, const Lay&,
                                        hipStream_t);

void launch_eta_check(const Lay& L, const EtaSrc& e, int c0, int c1, int* flags, hipStream_t s) {
    if (L.dim == 3) hipLaunchKernelGGL(k_eta_check<3>, node_grid(L), dim3(BX, BY), 0, s, L, e, c0, c1, flags);
    else hipLaunchKernelGGL(k_eta_check<2>, node_grid(L), dim3(BX, BY), 0, s, L, e, c0, c1, flags);
}
void launch_degrade_iso(const Lay& L, const EtaSrc& e, int ca, int cb, double lambda0, double mu0, double* C,
                        hipStream_t s) {
    if (L.dim == 3)
        hipLaunchKernelGGL(k_degrade_iso<3>, node_grid(L), dim3(BX, BY), 0, s, L, e, ca, cb, lambda0, mu0, C);
    else hipLaunchKernelGGL(k_degrade_iso<2>, node_grid(L), dim3(BX, BY), 0, s, L, e, ca, cb, lambda0, mu0, C);
}
void launch_combine(const Lay& L, const EtaSrc& e, int nph, const double* ddw, const double* F0s, double* C,
                    double* F0out, int* flags, hipStream_t s) {
    if (L.dim == 3)
        hipLaunchKernelGGL((k_combine<3, 21>), node_grid(L), dim3(BX, BY), 0, s, L, e, nph, ddw, F0s, C, F0out, flags);
    else
        hipLaunchKernelGGL((k_combine<2, 6>), node_grid(L), dim3(BX, BY), 0, s, L, e, nph, ddw, F0s, C, F0out, flags);
}

template <typename T>
void launch_pack(const Lay& L, int nc, const double* dense, T* dev, int to_dev, hipStream_t s) {
    hipLaunchKernelGGL((k_pack<T>), node_grid(L), dim3(BX, BY), 0, s, L, nc, dense, dev, to_dev);
}

// Grain-table variants (3D major/minor levels): the stiffness of node o is tab[gid[o]].
template <typename T>
void launch_fapply_tab(const Lay& L, const TabDesc<T>& t, const T* u, T* f, const BCTab& bc, const double h[3],
                       int uniform, hipStream_t s) {
    const Inv<T> iv = make_inv<T>(h);
    const Slab S = slab_of(L);
    const dim3 grid = slab_grid(L);
    using A = TabAcc<T, 21>;
    hipLaunchKernelGGL((k_fapply<T, K_MM, 3, A>), grid, dim3(BX, BY), 0, s, L, (A{t.gid, t.tab, t.dgt}), u, f, bc, iv,
                       uniform, S);
}

template <typename T>
void launch_jacobi_tab(const Lay& L, const TabDesc<T>& t, const T* dg, const T* xin, const T* b, T* xout,
                       const BCTab& bc, const double h[3], int uniform, double omega, hipStream_t s) {
    const Inv<T> iv = make_inv<T>(h);
    const Slab S = slab_of(L);
    const dim3 grid = slab_grid(L);
    using A = TabAcc<T, 21>;
    hipLaunchKernelGGL((k_jacobi<T, K_MM, 3, A>), grid, dim3(BX, BY), 0, s, L, (A{t.gid, t.tab, t.dgt}), dg, xin, b,
                       xout, bc, iv, uniform, (T)omega, S);
}

template <typename T, typename TO>
void launch_residual_tab(const Lay& L, const TabDesc<T>& t, const T* x, const T* b, T* r, TO* rs,
                         const double* scale, double* partials, unsigned long long* norm, const BCTab& bc,
                         const double h[3], int uniform, hipStream_t s) {
    const Inv<T> iv = make_inv<T>(h);
    const Slab S = slab_of(L);
    const dim3 grid = slab_grid(L);
    using A = TabAcc<T, 21>;
    hipLaunchKernelGGL((k_residual<T, TO, K_MM, 3, A>), grid, dim3(BX, BY), 0, s, L, (A{t.gid, t.tab, t.dgt}), x, b,
                       r, rs, scale, norm ? partials : nullptr, bc, iv, uniform, S);
    if (norm) hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, s, partials, (int)(grid.x * grid.y * grid.z), norm);
}

// Post-processing (Elastic.cpp:399-616): Strain = sym(grad u), Stress =
// C : grad u (DW of the linear models), Energy = W(grad u) + sigma : eps / 2
// with W = tr(grad u^T sigma) / 2, as the reference adds them (which is
// sigma : eps for the symmetric sigma of these models).  what: 0 strain,
// 1 stress, 2 energy; voigt: (00,11,22,12,20,01) in 3D, (00,11,01) in 2D,
// otherwise all DIM*DIM entries row-major.
namespace {
template <int K, int DIM>
__global__ __launch_bounds__(256) void k_post(Lay L, const double* __restrict__ C, const double* __restrict__ u,
                                              double* __restrict__ out, Inv<double> iv, int what, int voigt) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const long long o = (long long)i + L.px * j + L.pl * k;
    double g[DIM][DIM], e[DIM][DIM], s[DIM][DIM] = {};
    node_grad<double, DIM>(L, u, i, j, k, iv, g);
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
        for (int b = 0; b < DIM; ++b) e[a][b] = 0.5 * (g[a][b] + g[b][a]);
    if (what != 0) stress<double, K, DIM>(PlaneRec<double>{C + o, L.cs}, g, s);
    if (what == 2) {
        double w = 0.0, se = 0.0;
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int b = 0; b < DIM; ++b) w += g[a][b] * s[a][b];
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int b = 0; b < DIM; ++b) se += 0.5 * s[a][b] * e[a][b];
        out[o] = 0.5 * w + se;
        return;
    }
    double m[DIM][DIM];
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
        for (int b = 0; b < DIM; ++b) m[a][b] = what == 0 ? e[a][b] : s[a][b];
    if (voigt) {
        if (DIM == 3) {
            const int vi[6] = {0, 1, 2, 1, 2, 0}, vj[6] = {0, 1, 2, 2, 0, 1};
#pragma unroll
            for (int c = 0; c < 6; ++c) out[c * L.cs + o] = m[vi[c]][vj[c]];
        } else {
            const int vi[3] = {0, 1, 0}, vj[3] = {0, 1, 1};
#pragma unroll
            for (int c = 0; c < 3; ++c) out[c * L.cs + o] = m[vi[c]][vj[c]];
        }
    } else {
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int b = 0; b < DIM; ++b) out[(a * DIM + b) * L.cs + o] = m[a][b];
    }
}
}  // namespace

void launch_post(const Lay& L, int kind, const double* C, const double* u, double* out, const double h[3], int what,
                 int voigt, hipStream_t s) {
    const Inv<double> iv = make_inv<double>(h);
    ELA_DISPATCH(kind, L.dim,
                 hipLaunchKernelGGL((k_post<KK, DD>), node_grid(L), dim3(BX, BY), 0, s, L, C, u, out, iv, what,
                                    voigt));
}

// Newton<T>::prepareForSolve (src/Solver/Nonlocal/Newton.H:30-140), models with
// kinematic variable grad u: dw = DDW : (grad u - F0) per node (one-sided
// gradient on the walls), stored as dim*dim planes (row-major).
template <int K, int DIM>
__global__ __launch_bounds__(256) void k_newton_dw(Lay L, const double* __restrict__ C, const double* __restrict__ F0,
                                                  const double* __restrict__ u, double* __restrict__ dw,
                                                  Inv<double> iv, double* __restrict__ w, double wfac) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const long long o = (long long)i + L.px * j + L.pl * k;
    double g[DIM][DIM], s[DIM][DIM] = {};
    node_grad<double, DIM>(L, u, i, j, k, iv, g);
    if (F0) {
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int b = 0; b < DIM; ++b) g[a][b] -= F0[(a * DIM + b) * L.cs + o];
    }
    stress<double, K, DIM>(PlaneRec<double>{C + o, L.cs}, g, s);
    if (dw) {
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int b = 0; b < DIM; ++b) dw[(a * DIM + b) * L.cs + o] = s[a][b];
    }
    if (w) {
        // W = trace((wfac g)^T (DDW g)) (Linear/Isotropic.H:31-34, Affine/Isotropic.H:31-34):
        // column sums of the product, then the trace, in index order
        double tr = 0.0;
#pragma unroll
        for (int b = 0; b < DIM; ++b) {
            double col = 0.0;
#pragma unroll
            for (int a = 0; a < DIM; ++a) col += (wfac * g[a][b]) * s[a][b];
            tr += col;
        }
        w[o] = tr;
    }
}

// rhs = b - BC row(u, grad u, dw) on the boundary, b - div dw (central
// differences of dw, in the reference's order of operations) inside.
template <int DIM>
__global__ __launch_bounds__(256) void k_newton_rhs(Lay L, const double* __restrict__ u, const double* __restrict__ b,
                                                   const double* __restrict__ dw, double* __restrict__ rhs, BCTab bc,
                                                   Inv<double> iv, double hx, double hy, double hz) {
    int i, j, k;
    if (!node_index(L, i, j, k)) return;
    const long long o = (long long)i + L.px * j + L.pl * k;
    const int ex = L.nx - 1, ey = L.ny - 1, ez = L.nz - 1;
    const bool wall = i == 0 || i == ex || j == 0 || j == ey || (DIM == 3 && (k == 0 || k == ez));
    if (wall) {
        double g[DIM][DIM], s[DIM][DIM], uc[DIM], f[DIM];
        node_grad<double, DIM>(L, u, i, j, k, iv, g);
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            uc[a] = u[a * L.cs + o];
#pragma unroll
            for (int c = 0; c < DIM; ++c) s[a][c] = dw[(a * DIM + c) * L.cs + o];
        }
        bc_row<double, DIM>(bc, face_of<DIM>(i, j, k, ex, ey, ez), uc, g, s, f);
#pragma unroll
        for (int a = 0; a < DIM; ++a) rhs[a * L.cs + o] = b[a * L.cs + o] - f[a];
        return;
    }
    const double h[3] = {hx, hy, hz};
    const long long st[3] = {1, L.px, L.pl};
#pragma unroll
    for (int p = 0; p < DIM; ++p) {
        double div = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double* plane = dw + (p * DIM + d) * L.cs + o;
            div += (plane[st[d]] - plane[-st[d]]) / 2. / h[d];
        }
        rhs[p * L.cs + o] = b[p * L.cs + o] - div;
    }
}

void launch_newton_w_dw(const Lay& L, int kind, const double* C, const double* F0, const double* u, double wfac,
                        double* w, double* dw, const double h[3], hipStream_t s) {
    const Inv<double> iv = make_inv<double>(h);
    ELA_DISPATCH(kind, L.dim,
                 hipLaunchKernelGGL((k_newton_dw<KK, DD>), node_grid(L), dim3(BX, BY), 0, s, L, C, F0, u, dw, iv, w,
                                    wfac));
}

void launch_newton_rhs(const Lay& L, int kind, const double* C, const double* F0, const double* u, const double* b,
                       double* dw, double* rhs, const BCTab& bc, const double h[3], hipStream_t s) {
    const Inv<double> iv = make_inv<double>(h);
    ELA_DISPATCH(kind, L.dim,
                 hipLaunchKernelGGL((k_newton_dw<KK, DD>), node_grid(L), dim3(BX, BY), 0, s, L, C, F0, u, dw, iv,
                                    (double*)nullptr, 0.0));
    if (L.dim == 3)
        hipLaunchKernelGGL((k_newton_rhs<3>), node_grid(L), dim3(BX, BY), 0, s, L, u, b, dw, rhs, bc, iv, h[0], h[1],
                           h[2]);
    else
        hipLaunchKernelGGL((k_newton_rhs<2>), node_grid(L), dim3(BX, BY), 0, s, L, u, b, dw, rhs, bc, iv, h[0], h[1],
                           1.0);
}

// Grain-form z-marching stencil on one box (ela_grain.hpp), one launch: the
// first nshp blocks compute the shell (boundary rows, latency-bound), the
// others march the interior tiles, so the shell overlaps the tiles instead of
// running after them.  nshp is a multiple of 8, so a tile's hardware block
// lands on the XCD that xcd_block assigns it.
namespace {
// Occupancy floor of the sweep: FP32 with one row per thread keeps 5 waves per
// SIMD (93 VGPRs); two rows per thread need about 155 VGPRs (two stencils in
// flight), so 3 waves (at 4 the compiler spills).
template <typename T, int RY, int PF>
constexpr int grain_waves() { return sizeof(T) == 4 ? (RY == 1 ? (PF == 1 ? 5 : 4) : 3) : 1; }
template <typename T, int MODE, typename TO, bool SC, int RY, int PF>
__global__ __launch_bounds__(MTX* MTY) __attribute__((amdgpu_waves_per_eu(grain_waves<T, RY, PF>()))) void k_grain_sweep(GrainBox<T> g, int nshp, int nfp, double* partials) {
    __shared__ MarchLds<T, SC, RY> sh;
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    const int b = (int)blockIdx.x;
    double m;
    if (b < nshp) {
        m = shell_node<T, MODE, TO>(g, b * (MTX * MTY) + (int)threadIdx.x);
    } else if (b < nshp + nfp) {
        m = b - nshp < grain_face_patches(g.L) ? face_patch<T, MODE, TO>(g, b - nshp, sh.su) : 0.0;
    } else {
        const int nm = (int)gridDim.x - nshp - nfp, bm = b - nshp - nfp;
        const int t = g.xcd ? xcd_block(bm, nm) : bm;
        m = march_tile<T, MODE, TO, SC, RY, PF>(g, t, sh, (T*)dyn_lds);
    }
    if (MODE == 2 && partials) block_max_store(partials, m);  // a max: the slot order does not matter
}
}  // namespace

static_assert(3 * (FJ + 2) * (FK + 2) * 2 <= MRING * 3 * MPL, "a face patch fits the marching LDS ring");
static_assert(FJ * FK <= MTX * MTY, "one face node per thread");

long long grain_partials(const Lay& L, int kz) {
    int ntx, nty, ntz, nsh;
    grain_tiles(L, kz, ntx, nty, ntz, nsh);
    return (long long)ntx * nty * ntz + (nsh + 7) / 8 * 8 + (grain_face_patches(L) + 7) / 8 * 8;
}

// MODE 2: the residual's max norm from the launch's block maxima.
template <int MODE>
void launch_grain_norm(int nblocks, double* partials, unsigned long long* norm, hipStream_t s) {
    if (MODE == 2 && norm) hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, s, partials, nblocks, norm);
}

template <typename T, int MODE, typename TO>
void launch_grain(GrainBox<T> g, int kz, double* partials, unsigned long long* norm, hipStream_t s) {
    int ntx, nty, ntz, nsh;
    if (g.ry != 2 || sizeof(T) != 4) g.ry = 1;  // two rows per thread: FP32 (the FP64 ring would need 64 KB)
    grain_tiles(g.L, kz, ntx, nty, ntz, nsh, g.xface, g.ry);
    g.ntx = ntx;
    g.nty = nty;
    g.kz = kz;
    const int nm = ntx * nty * ntz, nshp = (nsh + 7) / 8 * 8;
    const int nfp = g.xface ? (grain_face_patches(g.L) + 7) / 8 * 8 : 0;
    const size_t lds = (size_t)g.ngr * (kGrainNC + (g.dgt ? 3 : 0)) * sizeof(T);
    const dim3 grid(nshp + nfp + nm), block(MTX * MTY);
    constexpr int R2 = sizeof(T) == 4 ? 2 : 1;
    auto go = [&](auto sc, auto ry, auto pf) {
        hipLaunchKernelGGL((k_grain_sweep<T, MODE, TO, decltype(sc)::value, decltype(ry)::value, decltype(pf)::value>),
                           grid, block, lds, s, g, nshp, nfp, partials);
    };
    using B0 = std::false_type;
    using B1 = std::true_type;
    using P1 = std::integral_constant<int, 1>;
    using P2 = std::integral_constant<int, 2>;
    using RR = std::integral_constant<int, R2>;
    // two planes ahead: FP32 tiles of one row per thread (4 waves a SIMD, 104-124
    // VGPRs); FP64 (about 200 VGPRs) and two-row tiles (spills) keep one
    if constexpr (sizeof(T) == 4) {
        if (g.pf == 2 && g.ry != 2) {
            if (g.sc) go(B1{}, P1{}, P2{});
            else go(B0{}, P1{}, P2{});
            return launch_grain_norm<MODE>(nm + nshp + nfp, partials, norm, s);
        }
    }
    if (g.sc && g.ry == 2) go(B1{}, RR{}, P1{});
    else if (g.sc) go(B1{}, P1{}, P1{});
    else if (g.ry == 2) go(B0{}, RR{}, P1{});
    else go(B0{}, P1{}, P1{});
    launch_grain_norm<MODE>(nm + nshp + nfp, partials, norm, s);
}
template void launch_grain<float, 0, float>(GrainBox<float>, int, double*, unsigned long long*, hipStream_t);
template void launch_grain<double, 0, double>(GrainBox<double>, int, double*, unsigned long long*, hipStream_t);
template void launch_grain<float, 1, float>(GrainBox<float>, int, double*, unsigned long long*, hipStream_t);
template void launch_grain<double, 1, double>(GrainBox<double>, int, double*, unsigned long long*, hipStream_t);
template void launch_grain<double, 2, float>(GrainBox<double>, int, double*, unsigned long long*, hipStream_t);
template void launch_grain<double, 2, double>(GrainBox<double>, int, double*, unsigned long long*, hipStream_t);
template void launch_grain<float, 3, float>(GrainBox<float>, int, double*, unsigned long long*, hipStream_t);
template void launch_grain<float, 4, float>(GrainBox<float>, int, double*, unsigned long long*, hipStream_t);
template void launch_grain<double, 4, double>(GrainBox<double>, int, double*, unsigned long long*, hipStream_t);
template void launch_grain<double, 3, double>(GrainBox<double>, int, double*, unsigned long long*, hipStream_t);

// The z-marching sweep of a box whose stiffness is per-node isotropic planes
// (the reference's per-node model field, src/Operator/Elastic.cpp:151-260):
// the grain-form launch layout (shell blocks, x-face patches, XCD-ordered
// tiles), with each node's (lambda, mu) read from its planes instead of a
// grain table.  Measured on MI355X (257^3 nodes, every node its own record):
// FP64 Fapply 0.62 -> 0.40 ms, FP32 0.31 -> 0.22 ms against the node kernels.
// Major/minor planes stay on the node kernels: their derivative of C reads
// 6 x 18 neighbour values a node, and through the tiles the FP64 apply was
// no faster (2.45 against 2.50 ms; 1.97 against 1.60 ms with the mask).
namespace {
// Occupancy floor: FP32 4 waves a SIMD (3 with two planes ahead), FP64 3.
template <typename T, int PF>
constexpr int plane_waves() { return sizeof(T) == 4 ? (PF == 1 ? 4 : 3) : 3; }
template <typename T, int MODE, typename TO, int PF>
__global__ __launch_bounds__(MTX* MTY) __attribute__((amdgpu_waves_per_eu(plane_waves<T, PF>()))) void k_plane_sweep(GrainBox<T> g, int nshp, int nfp, double* partials) {
    __shared__ MarchLds<T, false, 1> sh;
    const int b = (int)blockIdx.x;
    double m;
    if (b < nshp) {
        m = shell_node<T, MODE, TO, kSrcPlanes, K_ISO>(g, b * (MTX * MTY) + (int)threadIdx.x);
    } else if (b < nshp + nfp) {
        m = b - nshp < grain_face_patches(g.L) ? face_patch<T, MODE, TO, kSrcPlanes, K_ISO>(g, b - nshp, sh.su) : 0.0;
    } else {
        const int nm = (int)gridDim.x - nshp - nfp, bm = b - nshp - nfp;
        const int t = g.xcd ? xcd_block(bm, nm) : bm;
        m = march_tile<T, MODE, TO, false, 1, PF, kSrcPlanes, K_ISO>(g, t, sh, nullptr);
    }
    if (MODE == 2 && partials) block_max_store(partials, m);
}
}  // namespace

template <typename T, int MODE, typename TO>
void launch_plane(GrainBox<T> g, int kz, double* partials, unsigned long long* norm, hipStream_t s) {
    int ntx, nty, ntz, nsh;
    g.ry = 1;
    grain_tiles(g.L, kz, ntx, nty, ntz, nsh, g.xface, 1);
    g.ntx = ntx;
    g.nty = nty;
    g.kz = kz;
    const int nm = ntx * nty * ntz, nshp = (nsh + 7) / 8 * 8;
    const int nfp = g.xface ? (grain_face_patches(g.L) + 7) / 8 * 8 : 0;
    const dim3 grid(nshp + nfp + nm), block(MTX * MTY);
    bool done = false;
    if constexpr (sizeof(T) == 4) {  // two planes ahead: FP32 (FP64 keeps one)
        if (g.pf == 2) {
            hipLaunchKernelGGL((k_plane_sweep<T, MODE, TO, 2>), grid, block, 0, s, g, nshp, nfp, partials);
            done = true;
        }
    }
    if (!done) hipLaunchKernelGGL((k_plane_sweep<T, MODE, TO, 1>), grid, block, 0, s, g, nshp, nfp, partials);
    launch_grain_norm<MODE>(nm + nshp + nfp, partials, norm, s);
}
template void launch_plane<float, 0, float>(GrainBox<float>, int, double*, unsigned long long*, hipStream_t);
template void launch_plane<double, 0, double>(GrainBox<double>, int, double*, unsigned long long*, hipStream_t);
template void launch_plane<float, 1, float>(GrainBox<float>, int, double*, unsigned long long*, hipStream_t);
template void launch_plane<double, 1, double>(GrainBox<double>, int, double*, unsigned long long*, hipStream_t);
template void launch_plane<double, 2, float>(GrainBox<double>, int, double*, unsigned long long*, hipStream_t);
template void launch_plane<double, 2, double>(GrainBox<double>, int, double*, unsigned long long*, hipStream_t);
template void launch_plane<float, 3, float>(GrainBox<float>, int, double*, unsigned long long*, hipStream_t);
template void launch_plane<double, 3, double>(GrainBox<double>, int, double*, unsigned long long*, hipStream_t);

// Per-grain interior diagonal sum_b C_{p b p b} (-2 / h_b^2) (Elastic.cpp:311-321 on an interior node).
__global__ void k_grain_diag(int ng, const double* __restrict__ tab, double hx, double hy, double hz,
                             double* __restrict__ d64, float* __restrict__ d32) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ng) return;
    const double h[3] = {hx, hy, hz};
    for (int p = 0; p < 3; ++p) {
        const double acc = grain_diag(tab + 21 * gi, h, p);
        if (d64) d64[3 * gi + p] = acc;
        if (d32) store_jacobi_diag(d32 + 3 * gi + p, acc);  // the inverse
    }
}

void launch_grain_diag(int ng, const double* tab, const double h[3], double* d64, float* d32, hipStream_t s) {
    hipLaunchKernelGGL(k_grain_diag, dim3((ng + 63) / 64), dim3(64), 0, s, ng, tab, h[0], h[1], h[2], d64, d32);
}


template void launch_fapply_tab<double>(const Lay&, const TabDesc<double>&, const double*, double*, const BCTab&,
                                        const double*, int, hipStream_t);
template void launch_fapply_tab<float>(const Lay&, const TabDesc<float>&, const float*, float*, const BCTab&,
                                       const double*, int, hipStream_t);
template void launch_jacobi_tab<float>(const Lay&, const TabDesc<float>&, const float*, const float*, const float*,
                                       float*, const BCTab&, const double*, int, double, hipStream_t);
template void launch_jacobi_tab<double>(const Lay&, const TabDesc<double>&, const double*, const double*,
                                        const double*, double*, const BCTab&, const double*, int, double, hipStream_t);
template void launch_residual_tab<double, float>(const Lay&, const TabDesc<double>&, const double*, const double*,
                                                 double*, float*, const double*, double*, unsigned long long*,
                                                 const BCTab&, const double*, int, hipStream_t);
template void launch_residual_tab<double, double>(const Lay&, const TabDesc<double>&, const double*, const double*,
                                                  double*, double*, const double*, double*, unsigned long long*,
                                                  const BCTab&, const double*, int, hipStream_t);
template void launch_residual_tab<float, float>(const Lay&, const TabDesc<float>&, const float*, const float*,
                                                float*, float*, const double*, double*, unsigned long long*,
                                                const BCTab&, const double*, int, hipStream_t);

// explicit instantiations
template void launch_fapply<double>(const Lay&, int, const double*, const double*, double*, const BCTab&,
                                    const double*, int, hipStream_t, const PlaneOpt&);
template void launch_fapply<float>(const Lay&, int, const float*, const float*, float*, const BCTab&, const double*,
                                   int, hipStream_t, const PlaneOpt&);
template void launch_residual<double, double>(const Lay&, int, const double*, const double*, const double*, double*,
                                              double*, const double*, double*, unsigned long long*, const BCTab&,
                                              const double*, int, hipStream_t, const PlaneOpt&);
template void launch_residual<double, float>(const Lay&, int, const double*, const double*, const double*, double*,
                                             float*, const double*, double*, unsigned long long*, const BCTab&, const double*,
                                             int, hipStream_t, const PlaneOpt&);
template void launch_residual<float, float>(const Lay&, int, const float*, const float*, const float*, float*,
                                            float*, const double*, double*, unsigned long long*, const BCTab&, const double*,
                                            int, hipStream_t, const PlaneOpt&);
template void launch_jacobi<double>(const Lay&, int, const double*, const double*, const double*, const double*,
                                    double*, const BCTab&, const double*, int, double, hipStream_t, const PlaneOpt&);
template void launch_jacobi<float>(const Lay&, int, const float*, const float*, const float*, const float*, float*,
                                   const BCTab&, const double*, int, double, hipStream_t, const PlaneOpt&);
template void launch_jacobi0<double>(const Lay&, const double*, const double*, double*, double, hipStream_t,
                                     const unsigned short*, const double*);
template void launch_jacobi0<float>(const Lay&, const float*, const float*, float*, double, hipStream_t,
                                    const unsigned short*, const float*);
template void launch_restrict<double>(const Lay&, const Lay&, int, const double*, double*, hipStream_t);
template void launch_restrict<float>(const Lay&, const Lay&, int, const float*, float*, hipStream_t);
template void launch_prolong<double>(const Lay&, const Lay&, int, const double*, double*, int, hipStream_t);
template void launch_prolong<float>(const Lay&, const Lay&, int, const float*, float*, int, hipStream_t);
template void launch_axpy<double, double>(const Lay&, int, const double*, double*, const double*, double, int,
                                          hipStream_t);
template void launch_axpy<float, double>(const Lay&, int, const float*, double*, const double*, double, int,
                                         hipStream_t);
template void launch_axpy<double, float>(const Lay&, int, const double*, float*, const double*, double, int,
                                         hipStream_t);
template void launch_axpy<float, float>(const Lay&, int, const float*, float*, const double*, double, int,
                                        hipStream_t);
template void launch_convert<double, float>(long long, const double*, float*, hipStream_t);
template void launch_convert<float, double>(long long, const float*, double*, hipStream_t);
template void launch_bottom<double>(const Lay&, int, int, const double*, const double*, double*, hipStream_t);
template void launch_bottom<float>(const Lay&, int, int, const double*, const float*, float*, hipStream_t);
template void launch_pack<double>(const Lay&, int, const double*, double*, int, hipStream_t);
template void launch_pack<float>(const Lay&, int, const double*, float*, int, hipStream_t);

}  // namespace ela

namespace ela {
template Inv<double> make_inv<double>(const double*);
template Inv<float> make_inv<float>(const double*);
}  // namespace ela
