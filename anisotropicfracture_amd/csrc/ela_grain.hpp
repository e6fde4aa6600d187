// ela_grain.hpp -- LDS z-marching stencil for boxes whose stiffness is in
// grain form (a 16-bit grain id per node, one record of 21 values per grain,
// optionally times a per-node factor).
//
// The interior of a box is cut into MTX x MTY tiles of (i, j) that march kz
// planes in k.  A ring of three planes of u (tile + one-node halo), of the
// grain ids and of the factors lives in LDS with the grain table; each plane
// of u is read from memory once per tile column instead of once per
// neighbouring row of every wavefront.  The stencil is interior_apply on an
// LDS context, so every node gets the same bits as from the global-memory
// kernels.  Boundary nodes (the box's shell) go through the global-memory
// boundary rows, one thread per shell node.
#pragma once
#include "ela_device.hpp"

namespace ela {

constexpr int MTX = 32, MTY = 8, MRW = MTY + 2, MCW = MTX + 2, MPL = MRW * MCW;
constexpr int kGrainNC = 21;
constexpr int kMarchMaxGrains = 256;  // the table is staged in LDS up to this many grains

// SC: the launch carries per-node factors (LDS ss holds them, 1 for a box
// without); else the records are read as they are.  The values are the same,
// since 1 x = x exactly.
template <typename T, bool SC>
struct LdsCtx {
    const T* su;               // [3 slots][3 comps][MRW][MCW]
    const unsigned short* sg;  // [3 slots][MRW][MCW]
    const T* ss;               // [3 slots][MRW][MCW] (SC only)
    const T* stab;             // grain table
    int sl[3];                 // ring slots of planes k-1, k, k+1
    int r, c;                  // this node's row / column in the tile arrays
    __device__ __forceinline__ int at(int dx, int dy, int dz) const {
        return (sl[dz + 1] * MRW + r + dy) * MCW + c + dx;
    }
    __device__ __forceinline__ T u(int p, int dx, int dy, int dz) const {
        return su[((sl[dz + 1] * 3 + p) * MRW + r + dy) * MCW + c + dx];
    }
    __device__ __forceinline__ auto rec(int dx, int dy, int dz) const {
        const int q = at(dx, dy, dz);
        if constexpr (SC) return ScaledRec<T>{stab + (int)sg[q] * kGrainNC, ss[q]};
        else return TabRec<T>{stab + (int)sg[q] * kGrainNC};
    }
    __device__ __forceinline__ bool same(int d) const {
        const int qp = at(d == 0, d == 1, d == 2), qm = at(-(d == 0), -(d == 1), -(d == 2));
        if constexpr (SC) return sg[qp] == sg[qm] && ss[qp] == ss[qm];
        else return sg[qp] == sg[qm];
    }
};

// Operands of one grain-form box.  MODE 0: Jacobi sweep out = in + omega (rhs -
// A in) / dg; MODE 1: out = rhs - A in; MODE 2: v = rhs - A in in FP64, out
// (precision TO, may be null) = v * (*scale), out2 (FP64, may be null) = v,
// block maxima of |v|; MODE 3: out = A in.
template <typename T>
struct GrainBox {
    Lay L;
    Inv<T> iv;
    const BCTab* bc;
    int uniform;
    const unsigned short* gid;
    const T* tab;
    const T* sc;
    int ngr;
    int ntx, nty, kz;
    int xcd;  // XCD-contiguous tile order
    const T* in;
    const T* rhs;
    void* out;
    double* out2;
    const T* dg;
    const double* scale;
    T omega;
};

template <typename T, int MODE, typename TO>
__device__ __forceinline__ double grain_epilogue(const GrainBox<T>& g, long long o, const T xc[3], const T f[3]) {
    const long long cs = g.L.cs;
    double m = 0.0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const long long q = p * cs + o;
        if (MODE == 0) {
            ((T*)g.out)[q] = xc[p] + g.omega * (g.rhs[q] - f[p]) / g.dg[q];
        } else if (MODE == 1) {
            ((T*)g.out)[q] = g.rhs[q] - f[p];
        } else if (MODE == 2) {
            const double v = (double)g.rhs[q] - (double)f[p];
            if (g.out) ((TO*)g.out)[q] = (TO)(v * *g.scale);
            if (g.out2) g.out2[q] = v;
            const double a = fabs(v);
            m = (a > m || a != a) ? a : m;
        } else {
            ((T*)g.out)[q] = f[p];
        }
    }
    return m;
}

// LDS arrays of one marching block.
template <typename T>
struct MarchLds {
    T su[3 * 3 * MPL];
    unsigned short sg[3 * MPL];
    T ss[3 * MPL];
};

// Tile t of box g (256 threads); returns the thread's max |v| (MODE 2).
// Every thread of the block calls this (it synchronizes).
template <typename T, int MODE, typename TO, bool SC>
__device__ __forceinline__ double march_tile(const GrainBox<T>& g, int t, MarchLds<T>& sh, T* stab) {
    static_assert(MODE != 2 || sizeof(T) == 8, "the solution residual is FP64");
    const Lay& L = g.L;
    const int tid = threadIdx.x;
    const int bx = t % g.ntx, by = (t / g.ntx) % g.nty, bz = t / (g.ntx * g.nty);
    const int i0 = 1 + bx * MTX, j0 = 1 + by * MTY;  // first interior node of the tile
    const int tx = tid % MTX, ty = tid / MTX;
    const int i = i0 + tx, j = j0 + ty;
    const bool active = i <= L.nx - 2 && j <= L.ny - 2;
    const int kb = 1 + bz * g.kz, ke = min(kb + g.kz, L.nz - 1);
    for (int q = tid; q < g.ngr * kGrainNC; q += MTX * MTY) stab[q] = g.tab[q];

    constexpr int PU = (3 * MPL + MTX * MTY - 1) / (MTX * MTY), PG = (MPL + MTX * MTY - 1) / (MTX * MTY);
    T ru[PU], rs[PG];
    unsigned short rg[PG];
    auto load = [&](int kk) {
#pragma unroll
        for (int e = 0; e < PU; ++e) {
            const int q = tid + e * MTX * MTY;
            T v = 0;
            if (q < 3 * MPL) {
                const int p = q / MPL, rem = q - p * MPL, rr = rem / MCW, cc = rem - rr * MCW;
                const int ii = i0 - 1 + cc, jj = j0 - 1 + rr;
                if (ii < L.nx && jj < L.ny) v = g.in[p * L.cs + ii + L.px * jj + L.pl * kk];
            }
            ru[e] = v;
        }
#pragma unroll
        for (int e = 0; e < PG; ++e) {
            const int q = tid + e * MTX * MTY;
            unsigned short gv = 0;
            T sv = 1;
            if (q < MPL) {
                const int rr = q / MCW, cc = q - rr * MCW;
                const int ii = i0 - 1 + cc, jj = j0 - 1 + rr;
                if (ii < L.nx && jj < L.ny) {
                    const long long o = ii + L.px * jj + L.pl * kk;
                    gv = g.gid[o];
                    // a batched launch takes SC when any of its boxes has a factor;
                    // the boxes without one read s = 1
                    if (SC && g.sc) sv = g.sc[o];
                }
            }
            rg[e] = gv;
            rs[e] = sv;
        }
    };
    auto store = [&](int slot) {
#pragma unroll
        for (int e = 0; e < PU; ++e) {
            const int q = tid + e * MTX * MTY;
            if (q < 3 * MPL) sh.su[slot * 3 * MPL + q] = ru[e];
        }
#pragma unroll
        for (int e = 0; e < PG; ++e) {
            const int q = tid + e * MTX * MTY;
            if (q < MPL) {
                sh.sg[slot * MPL + q] = rg[e];
                if (SC) sh.ss[slot * MPL + q] = rs[e];
            }
        }
    };
    load(kb - 1);
    store((kb - 1) % 3);
    load(kb);
    store(kb % 3);
    load(kb + 1);
    store((kb + 1) % 3);
    __syncthreads();
    double m = 0.0;
    for (int k = kb; k < ke; ++k) {
        const bool more = k + 1 < ke;
        if (more) load(k + 2);
        if (active) {
            LdsCtx<T, SC> S{sh.su, sh.sg, sh.ss, stab, {(k - 1) % 3, k % 3, (k + 1) % 3}, ty + 1, tx + 1};
            T f[3], xc[3];
            interior_apply<T, K_MM, 3>(S, g.iv, g.uniform != 0, f);
#pragma unroll
            for (int p = 0; p < 3; ++p) xc[p] = S.u(p, 0, 0, 0);
            const double mm = grain_epilogue<T, MODE, TO>(g, (long long)i + L.px * j + L.pl * k, xc, f);
            m = (mm > m || mm != mm) ? mm : m;
        }
        __syncthreads();  // every read of plane k-1's slot is done
        if (more) store((k + 2) % 3);
        __syncthreads();
    }
    return m;
}

// Shell node q of box g (both z faces, then the y faces and the x faces
// between them); returns false past the shell.  Returns the node's |v| max in m.
template <typename T, int MODE, typename TO>
__device__ __forceinline__ double shell_node(const GrainBox<T>& g, int q) {
    const Lay& L = g.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int A = nx * ny, Bf = nx * (nz - 2), Cf = (ny - 2) * (nz - 2);
    if (q >= 2 * A + 2 * Bf + 2 * Cf) return 0.0;
    int i, j, k;
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
    T f[3], xc[3];
    node_apply_acc<T, K_MM, 3>(L, GrainAcc<T, kGrainNC>{g.gid, g.tab, g.sc}, g.in, i, j, k, *g.bc, g.iv,
                               g.uniform != 0, f);
    const long long o = (long long)i + L.px * j + L.pl * k;
#pragma unroll
    for (int p = 0; p < 3; ++p) xc[p] = g.in[p * L.cs + o];
    return grain_epilogue<T, MODE, TO>(g, o, xc, f);
}

// XCD-contiguous block numbering: hardware block b of a launch of n blocks
// runs on XCD b % 8; give each XCD a contiguous range of logical blocks, so
// neighbouring tiles (which share halo lines) run under the same L2.
__device__ __forceinline__ int xcd_block(int b, int n) {
    const int per = (n + 7) / 8;
    const int x = b % 8, r = b / 8;
    const int full = n % 8 == 0 ? 8 : n % 8;  // XCDs that receive `per` blocks
    if (x < full) return x * per + r;
    return full * per + (x - full) * (per - 1) + r;
}

// Tiles and shell blocks of a box (host and device).
__host__ __device__ inline void grain_tiles(const Lay& L, int kz, int& ntx, int& nty, int& ntz, int& nsh) {
    ntx = (L.nx - 2 + MTX - 1) / MTX;
    nty = (L.ny - 2 + MTY - 1) / MTY;
    ntz = (L.nz - 2 + kz - 1) / kz;
    const long long shell = (long long)L.nx * L.ny * L.nz - (long long)(L.nx - 2) * (L.ny - 2) * (L.nz - 2);
    nsh = (int)((shell + 255) / 256);
}

}  // namespace ela
