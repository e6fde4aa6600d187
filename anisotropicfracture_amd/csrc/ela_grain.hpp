// ela_grain.hpp -- LDS z-marching stencil for boxes whose stiffness is in
// grain form (a 16-bit grain id per node, one record of 21 values per grain,
// optionally times a per-node factor), or in per-node isotropic planes (the
// reference's per-node model field of (lambda, mu), read once per node).
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

constexpr int MTX = 32, MTY = 8, MCW = MTX + 2;
// A tile is MTX x (MTY RY) nodes: each of the MTX x MTY threads takes RY
// adjacent rows (j), whose stencils share two thirds of their LDS reads.
template <int RY>
struct MTile {
    static constexpr int TY = MTY * RY, RW = TY + 2, PL = RW * MCW;
    static constexpr int HALO = 2 * MCW + 2 * TY;                       // nodes of a plane's one-node halo ring
    static constexpr int NH = (3 * HALO + MTX * MTY - 1) / (MTX * MTY);  // halo loads per thread and plane
};
constexpr int MRW = MTile<1>::RW, MPL = MTile<1>::PL, MHALO = MTile<1>::HALO;
static_assert(MTile<1>::NH == 1 && MTile<2>::NH == 2, "halo loads per thread");
constexpr int kGrainNC = 21;
constexpr int kMarchMaxGrains = 256;  // the table is staged in LDS up to this many grains

// First i of a box's tile columns.  Where rows start on a multiple of MTX
// elements (the single-box layout), tiles start at i = 0, so a tile row of the
// interior streams (rhs, diagonal, output) is whole cache lines; node 0 of the
// first tile is a shell node and is skipped.  Compact layouts start at i = 1.
__host__ __device__ inline int grain_x0(const Lay& L) { return L.px % MTX == 0 ? 0 : 1; }

// SC: the launch carries per-node factors (LDS ss holds them, 1 for a box
// without); else the records are read as they are.  The values are the same,
// since 1 x = x exactly.
template <typename T, bool SC, int PL = MPL>
struct LdsCtx {
    const T* su;               // [slots][3 comps][rows][MCW] (PL = rows * MCW per component plane)
    const unsigned short* sg;  // [slots][rows][MCW]
    const T* ss;               // [slots][rows][MCW] (SC only)
    const T* stab;             // grain table
    int pb[3];                 // slot * MPL of planes k-1, k, k+1 (uniform over the block)
    int tb;                    // this node's row * MCW + column in the tile arrays
    // Index = uniform plane base + per-thread base + a compile-time constant,
    // so each access is one LDS read with an immediate offset.
    __device__ __forceinline__ int at(int dx, int dy, int dz) const { return pb[dz + 1] + tb + (dy * MCW + dx); }
    __device__ __forceinline__ T u(int p, int dx, int dy, int dz) const {
        return su[3 * pb[dz + 1] + tb + (p * PL + dy * MCW + dx)];
    }
    __device__ __forceinline__ auto rec(int dx, int dy, int dz) const {
        const int q = at(dx, dy, dz);
        if constexpr (SC) return ScaledRec<T>{stab + (int)sg[q] * kGrainNC, ss[q]};
        else return TabRec<T>{stab + (int)sg[q] * kGrainNC};
    }
    __device__ __forceinline__ auto center() const { return rec(0, 0, 0); }
    __device__ __forceinline__ bool same(int d) const {
        const int qp = at(d == 0, d == 1, d == 2), qm = at(-(d == 0), -(d == 1), -(d == 2));
        if constexpr (SC) return sg[qp] == sg[qm] && ss[qp] == ss[qm];
        else return sg[qp] == sg[qm];
    }
};

// A node's record in per-node planes, addressed by 32-bit byte offsets from
// the planes' base (uniform over the block): each slot is one load at the
// node's offset from a uniform slot base, so the record costs no 64-bit
// address registers (plane_offsets_fit holds for every box marched).
template <typename T>
struct PlaneRecB {
    const char* base;
    unsigned off, csb;  // the node's byte offset; bytes between two slots' planes
    __device__ __forceinline__ T operator[](int q) const {
        return *(const T*)((base + (size_t)q * csb) + off);
    }
};

// The same on a box whose stiffness is per-node planes: u from the LDS ring,
// the node's own record from registers (cr, loaded ahead of the stencil), a
// neighbour's from the planes at its offset (a cache hit after the
// neighbour's own read), and the mask byte m of the node where the level has
// a mask (else m = 0: every derivative of C is computed).
template <typename T, int PL = MPL>
struct PlaneLdsCtx {
    const T* su;
    int pb[3];
    int tb;
    const char* base;
    unsigned off, csb;
    int px, pl;
    unsigned m;
    const T* cr;
    __device__ __forceinline__ TabRec<T> center() const { return TabRec<T>{cr}; }
    __device__ __forceinline__ T u(int p, int dx, int dy, int dz) const {
        return su[3 * pb[dz + 1] + tb + (p * PL + dy * MCW + dx)];
    }
    __device__ __forceinline__ PlaneRecB<T> rec(int dx, int dy, int dz) const {
        return PlaneRecB<T>{base, off + (unsigned)((dx + dy * px + dz * pl) * (int)sizeof(T)), csb};
    }
    __device__ __forceinline__ bool same(int d) const { return (m >> d) & 1u; }
};

// Stiffness source of a marching box: grain form (ids, table, factors) or
// per-node planes, with the model kind K of the planes (grain records are
// always major/minor).
constexpr int kSrcGrain = 0, kSrcPlanes = 1;

// Operands of one grain-form box.  MODE 0: Jacobi sweep out = in + omega (rhs -
// A in) / dg; MODE 1: out = rhs - A in; MODE 2: v = rhs - A in in FP64, out
// (precision TO, may be null) = v * (*scale), out2 (FP64, may be null) = v,
// block maxima of |v|; MODE 3: out = A in; MODE 4: the Jacobi sweep of MODE 0
// followed by k_add_to64's step: out2 (FP64) = (T) (x + hold) with x the
// sweep's value, which is not stored (a cycle's last sweep, whose correction
// goes straight into an FP64 Krylov vector).
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
    const T* C;                // plane form: stiffness planes (component stride L.cs)
    const unsigned char* sm;   // plane form: same-direction mask (bit d: skip the d-derivative of C) or null
    int ntx, nty, kz;
    int ry;   // rows per thread of the marching tiles (1 or 2)
    int pf;   // planes loaded ahead by the marching tiles (1 or 2; march_tile's PF)
    int xcd;  // XCD-contiguous tile order
    bool nt;     // rhs, diagonal and Jacobi output non-temporal (see epilogue_load)
    bool xface;  // x faces in face blocks (face_patch); the shell then skips them
    const T* in;
    const T* rhs;
    void* out;
    double* out2;
    const T* dg;
    // Grain form without factors: the level's per-grain interior Jacobi
    // diagonal (3 values a grain, as the smoothers hold it: FP32 the inverse),
    // or null.  An interior node's diagonal is sum_b C_{p b p b} (-2 / h_b^2)
    // of its own record (the first- and mixed-derivative stencils weigh the
    // node itself by 0), so the tiles, whose nodes are all interior, take it
    // from LDS beside the table instead of streaming dg; the shell and face
    // nodes read dg.
    const T* dgt;
    const T* hold;  // MODE 4
    const double* scale;
    T omega;
};

template <typename T, int SRC>
__device__ __forceinline__ auto box_acc(const GrainBox<T>& g) {
    if constexpr (SRC == kSrcGrain) return GrainAcc<T, kGrainNC>{g.gid, g.tab, g.sc};
    else return PlaneAcc<T>{g.C, g.L.cs};
}

// The epilogue's operands of node o: rhs (MODES 0-2) and the diagonal (MODES 0
// and 4, unless ldg is false: the caller takes it from the grain's record).
// With g.nt (levels far larger than the caches) nothing reads them again
// before they are evicted, so they are loaded non-temporally: they then leave
// more of the L2 to the lines of u and of the grain ids that neighbouring
// tiles re-read as halos (measured: 0.299 -> 0.291 ms a 257^3 FP32 sweep).
// The Jacobi output is stored the same way.  A 129^3 level's fields stay in
// the caches between its sweeps, and the hint made its sweep slower (59 -> 67
// us), so smaller levels load normally.
template <typename T, int MODE>
__device__ __forceinline__ void epilogue_load(const GrainBox<T>& g, long long o, T rb[3], T db[3], bool ldg = true) {
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const long long q = p * g.L.cs + o;
        if (g.nt) {
            rb[p] = MODE != 3 ? __builtin_nontemporal_load(g.rhs + q) : T(0);
            db[p] = (MODE == 0 || MODE == 4) && ldg ? __builtin_nontemporal_load(g.dg + q) : T(1);
        } else {
            rb[p] = MODE != 3 ? g.rhs[q] : T(0);
            db[p] = (MODE == 0 || MODE == 4) && ldg ? g.dg[q] : T(1);
        }
    }
}

// Output of node o from its stencil value f, its current value xc and the
// preloaded operands rb, db of epilogue_load.
template <typename T, int MODE, typename TO>
__device__ __forceinline__ double grain_epilogue(const GrainBox<T>& g, long long o, const T xc[3], const T f[3],
                                                 const T rb[3], const T db[3]) {
    const long long cs = g.L.cs;
    double m = 0.0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const long long q = p * cs + o;
        if (MODE == 0) {
            const T x = jacobi_update(xc[p], g.omega, rb[p] - f[p], db[p]);
            if (g.nt) __builtin_nontemporal_store(x, (T*)g.out + q);
            else ((T*)g.out)[q] = x;
        } else if (MODE == 4) {
            const T x = jacobi_update(xc[p], g.omega, rb[p] - f[p], db[p]);
            const double v = (double)(T)((double)x + (double)g.hold[q]);
            if (g.nt) __builtin_nontemporal_store(v, g.out2 + q);
            else g.out2[q] = v;
        } else if (MODE == 1) {
            ((T*)g.out)[q] = rb[p] - f[p];
        } else if (MODE == 2) {
            const double v = (double)rb[p] - (double)f[p];
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

// LDS arrays of one marching block: a ring of MRING planes (three are read
// while the fourth is written, so a plane needs one barrier, not two).  The
// factor planes exist only in launches that carry factors (SC).
constexpr int MRING = 4;
template <typename T, bool SC, int RY = 1>
struct MarchLds {
    static constexpr int PL = MTile<RY>::PL;
    T su[MRING * 3 * PL];
    unsigned short sg[MRING * PL];
    T ss[SC ? MRING * PL : 1];
};

// Tile t of box g (256 threads); returns the thread's max |v| (MODE 2).
// Every thread of the block calls this (it synchronizes).  SRC: the box's
// stiffness is in grain form (the ids, and factors with SC, ride in the LDS
// ring, the table sits in stab) or in per-node planes of model kind K (no
// ring for them: the stencil reads each node's record from its planes).
template <typename T, int MODE, typename TO, bool SC, int RY = 1, int PF = 1, int SRC = kSrcGrain, int K = K_MM>
__device__ __forceinline__ double march_tile(const GrainBox<T>& g, int t, MarchLds<T, SC, RY>& sh, T* stab) {
    static_assert(MODE != 2 || sizeof(T) == 8, "the solution residual is FP64");
    static_assert(SRC == kSrcPlanes || K == K_MM, "grain records are major/minor");
    constexpr bool GR = SRC == kSrcGrain;
    using TL = MTile<RY>;
    constexpr int PL = TL::PL, HALO = TL::HALO, NH = TL::NH, RW = TL::RW;
    const Lay& L = g.L;
    const int tid = threadIdx.x;
    const int bx = t % g.ntx, by = (t / g.ntx) % g.nty, bz = t / (g.ntx * g.nty);
    const int i0 = grain_x0(L) + bx * MTX, j0 = 1 + by * TL::TY;  // first node of the tile
    const int tx = tid % MTX, ty = tid / MTX;
    const int i = i0 + tx;
    const int kb = 1 + bz * g.kz, ke = min(kb + g.kz, L.nz - 1);
    // the interior diagonal from the grains' records (GrainBox::dgt), staged
    // after the table; a box with factors reads dg
    const bool tdg = GR && (MODE == 0 || MODE == 4) && g.dgt != nullptr && !(SC && g.sc != nullptr);
    const T* sdg = GR ? stab + g.ngr * kGrainNC : nullptr;
    if constexpr (GR) {
        for (int q = tid; q < g.ngr * kGrainNC; q += MTX * MTY) stab[q] = g.tab[q];
        if (tdg)
            for (int q = tid; q < g.ngr * 3; q += MTX * MTY) stab[g.ngr * kGrainNC + q] = g.dgt[q];
    }

    // Plane loads.  Every thread reads its own RY nodes (i, j0 + RY ty + y) of
    // the tile for the three components, at 32-bit offsets from the plane's
    // base; the halo ring's 3 HALO values (node, component) are spread over the
    // threads, NH per thread, and the loads of component 0 also read the node's
    // grain id (and factor).  Offsets fit 32 bits: grain_offsets_fit holds for
    // every box marched.
    int tb[RY];
    bool act[RY], iok[RY];
    unsigned ie[RY];
#pragma unroll
    for (int y = 0; y < RY; ++y) {
        const int r = ty * RY + y, j = j0 + r;
        tb[y] = (r + 1) * MCW + tx + 1;  // this node's position in a tile plane
        act[y] = i >= 1 && i <= L.nx - 2 && j <= L.ny - 2;
        iok[y] = i < L.nx && j < L.ny;
        ie[y] = (unsigned)(i + L.px * j);
    }
    unsigned he[NH];
    int hpos[NH];
    bool hok[NH], gok[NH];
#pragma unroll
    for (int h = 0; h < NH; ++h) {
        const int e = tid + h * (MTX * MTY);
        const int hp = e / HALO, hh = e - hp * HALO;
        const int hr = hh < MCW ? 0 : hh < 2 * MCW ? RW - 1 : 1 + ((hh - 2 * MCW) >> 1);
        const int hc = hh < MCW ? hh : hh < 2 * MCW ? hh - MCW : (((hh - 2 * MCW) & 1) ? MCW - 1 : 0);
        hok[h] = e < 3 * HALO && i0 - 1 + hc >= 0 && i0 - 1 + hc < L.nx && j0 - 1 + hr < L.ny;
        gok[h] = hok[h] && e < HALO;
        he[h] = (unsigned)(hp * L.cs + (i0 - 1 + hc) + L.px * (j0 - 1 + hr));
        hpos[h] = hp * PL + hr * MCW + hc;
    }
    const bool scl = SC && g.sc != nullptr;  // a batched launch takes SC when any of its boxes has
                                             // a factor; the boxes without one read s = 1
    // One plane's loads in registers: the tile's nodes (ru, grain ids rg,
    // factors rs) and this thread's halo values (rh, rgh, rsh).
    struct Regs {
        T ru[RY][3], rh[NH], rs[RY], rsh[NH];
        unsigned short rg[RY], rgh[NH];
    };
    auto at = [](const void* base, unsigned bytes) { return (const char*)base + bytes; };
    auto load = [&](int kk, Regs& R) {
        const T* up = g.in + L.pl * kk;
        const unsigned short* gp = GR ? g.gid + L.pl * kk : nullptr;
#pragma unroll
        for (int y = 0; y < RY; ++y)
#pragma unroll
            for (int p = 0; p < 3; ++p)
                R.ru[y][p] = iok[y] ? *(const T*)at(up + p * L.cs, ie[y] * (unsigned)sizeof(T)) : T(0);
#pragma unroll
        for (int h = 0; h < NH; ++h) R.rh[h] = hok[h] ? *(const T*)at(up, he[h] * (unsigned)sizeof(T)) : T(0);
        if constexpr (GR) {
#pragma unroll
            for (int y = 0; y < RY; ++y) R.rg[y] = iok[y] ? *(const unsigned short*)at(gp, ie[y] * 2u) : (unsigned short)0;
#pragma unroll
            for (int h = 0; h < NH; ++h) R.rgh[h] = gok[h] ? *(const unsigned short*)at(gp, he[h] * 2u) : (unsigned short)0;
        }
        if constexpr (SC) {
            const T* sp = g.sc + L.pl * kk;
#pragma unroll
            for (int y = 0; y < RY; ++y) R.rs[y] = iok[y] && scl ? *(const T*)at(sp, ie[y] * (unsigned)sizeof(T)) : T(1);
#pragma unroll
            for (int h = 0; h < NH; ++h) R.rsh[h] = gok[h] && scl ? *(const T*)at(sp, he[h] * (unsigned)sizeof(T)) : T(1);
        } else {
#pragma unroll
            for (int y = 0; y < RY; ++y) R.rs[y] = 1;
#pragma unroll
            for (int h = 0; h < NH; ++h) R.rsh[h] = 1;
        }
    };
    auto store = [&](int slot, const Regs& R) {
        T* su = sh.su + slot * 3 * PL;
#pragma unroll
        for (int y = 0; y < RY; ++y) {
#pragma unroll
            for (int p = 0; p < 3; ++p) su[p * PL + tb[y]] = R.ru[y][p];
            if constexpr (GR) sh.sg[slot * PL + tb[y]] = R.rg[y];
            if constexpr (SC) sh.ss[slot * PL + tb[y]] = R.rs[y];
        }
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            const int e = tid + h * (MTX * MTY);
            if (e < 3 * HALO) su[hpos[h]] = R.rh[h];
            if (GR && e < HALO) {  // component 0: the node's grain id (and factor)
                sh.sg[slot * PL + hpos[h]] = R.rgh[h];
                if constexpr (SC) sh.ss[slot * PL + hpos[h]] = R.rsh[h];
            }
        }
    };
    // ring slots of planes k-1, k, k+1 and the free slot s3, which receives
    // plane k+2 while plane k is computed.  PF 1: plane k+2 is loaded into
    // registers (A) at the start of iteration k and the epilogue's operands
    // of plane k there too.  PF 2: plane k+3 is loaded (B) and the epilogue's
    // operands of plane k+1, so each load has a whole plane's computation
    // more to arrive; A then holds plane k+2 from the previous iteration.
    int s0 = 0, s1 = 1, s2 = 2, s3 = 3;
    Regs A, B;
    load(kb - 1, A);
    store(s0, A);
    load(kb, A);
    store(s1, A);
    load(kb + 1, A);
    store(s2, A);
    // the epilogue's operands of the current plane (separate scalars per row:
    // arrays indexed by the row cost the one-row kernel registers), and of the
    // next plane (PF 2)
    T rb0[3], db0[3], rb1[3], db1[3], nb0[3], nd0[3], nb1[3], nd1[3];
    const long long ob0 = (long long)i + L.px * (j0 + ty * RY), ob1 = ob0 + L.px;
    // plane form: the node's own record (and mask byte), loaded with the
    // epilogue's operands so their latency overlaps the stencil
    static_assert(GR || RY == 1, "plane-form tiles take one row per thread");
    constexpr int NCR = GR ? 1 : NComp<3, K>::v;
    T cr[NCR], ncr[NCR];
    unsigned mk = 0, nmk = 0;
    const unsigned csb = (unsigned)L.cs * (unsigned)sizeof(T);
    auto rec_load = [&](long long o, T* r, unsigned& mm) {
        if constexpr (!GR) {
            const unsigned off = (unsigned)o * (unsigned)sizeof(T);
#pragma unroll
            for (int q = 0; q < NCR; ++q) r[q] = *(const T*)(((const char*)g.C + (size_t)q * csb) + off);
            mm = g.sm ? (unsigned)g.sm[o] : 0u;
        }
    };
    if constexpr (PF == 2) {
        if (kb + 1 < ke) load(kb + 2, A);
        if (act[0]) {
            epilogue_load<T, MODE>(g, ob0 + L.pl * kb, rb0, db0, !tdg);
            rec_load(ob0 + L.pl * kb, cr, mk);
        }
        if constexpr (RY == 2)
            if (act[RY - 1]) epilogue_load<T, MODE>(g, ob1 + L.pl * kb, rb1, db1, !tdg);
    }
    __syncthreads();
    double m = 0.0;
    for (int k = kb; k < ke; ++k) {
        const bool more = k + 1 < ke;
        const long long o0 = ob0 + L.pl * k, o1 = ob1 + L.pl * k;
        if constexpr (PF == 1) {
            // the epilogue's loads are issued before the stencils, so their
            // latency overlaps them, as the next plane's loads do
            if (act[0]) {
                epilogue_load<T, MODE>(g, o0, rb0, db0, !tdg);
                rec_load(o0, cr, mk);
            }
            if constexpr (RY == 2)
                if (act[RY - 1]) epilogue_load<T, MODE>(g, o1, rb1, db1, !tdg);
            if (more) load(k + 2, A);
        } else {
            if (k + 2 < ke) load(k + 3, B);
            if (more) {
                if (act[0]) {
                    epilogue_load<T, MODE>(g, o0 + L.pl, nb0, nd0, !tdg);
                    rec_load(o0 + L.pl, ncr, nmk);
                }
                if constexpr (RY == 2)
                    if (act[RY - 1]) epilogue_load<T, MODE>(g, o1 + L.pl, nb1, nd1, !tdg);
            }
        }
        auto node = [&](int y, long long o, const T* rb, const T* db) {
            T f[3], xc[3], r3[3] = {rb[0], rb[1], rb[2]}, d3[3] = {db[0], db[1], db[2]};
            if constexpr (GR) {
                LdsCtx<T, SC, PL> S{sh.su, sh.sg, sh.ss, stab, {s0 * PL, s1 * PL, s2 * PL}, tb[y]};
                interior_apply<T, K_MM, 3>(S, g.iv, g.uniform != 0, f);
            } else {
                PlaneLdsCtx<T, PL> S{sh.su, {s0 * PL, s1 * PL, s2 * PL}, tb[y], (const char*)g.C,
                                     (unsigned)o * (unsigned)sizeof(T), csb, (int)L.px, (int)L.pl, mk, cr};
                interior_apply<T, K, 3>(S, g.iv, g.uniform != 0, f);
            }
#pragma unroll
            for (int p = 0; p < 3; ++p) xc[p] = sh.su[3 * (s1 * PL) + tb[y] + p * PL];
            if (tdg) {
                const int gc = (int)sh.sg[s1 * PL + tb[y]] * 3;
#pragma unroll
                for (int p = 0; p < 3; ++p) d3[p] = sdg[gc + p];
            }
            const double mm = grain_epilogue<T, MODE, TO>(g, o, xc, f, r3, d3);
            m = (mm > m || mm != mm) ? mm : m;
        };
        if (act[0]) node(0, o0, rb0, db0);
        if constexpr (RY == 2)
            if (act[RY - 1]) node(RY - 1, o1, rb1, db1);
        // s3 held plane k-2, whose last reads were in the previous plane,
        // before the previous barrier
        if (more) store(s3, A);
        __syncthreads();  // plane k+2 is written; every read of plane k-1 is done
        const int t0 = s0;
        s0 = s1;
        s1 = s2;
        s2 = s3;
        s3 = t0;
        if constexpr (PF == 2) {
            A = B;
            if constexpr (!GR) {
#pragma unroll
                for (int q = 0; q < NCR; ++q) cr[q] = ncr[q];
                mk = nmk;
            }
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                rb0[p] = nb0[p];
                db0[p] = nd0[p];
                if constexpr (RY == 2) {
                    rb1[p] = nb1[p];
                    db1[p] = nd1[p];
                }
            }
        }
    }
    return m;
}

// Shell node q of box g (both z faces, then the y faces and the x faces
// between them); returns false past the shell.  Returns the node's |v| max in m.
template <typename T, int MODE, typename TO, int SRC = kSrcGrain, int K = K_MM>
__device__ __forceinline__ double shell_node(const GrainBox<T>& g, int q) {
    const Lay& L = g.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int A = nx * ny, Bf = nx * (nz - 2), Cf = (ny - 2) * (nz - 2);
    if (q >= 2 * A + 2 * Bf + (g.xface ? 0 : 2 * Cf)) return 0.0;
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
    node_apply_acc<T, K, 3>(L, box_acc<T, SRC>(g), g.in, i, j, k, *g.bc, g.iv,
                               g.uniform != 0, f);
    const long long o = (long long)i + L.px * j + L.pl * k;
#pragma unroll
    for (int p = 0; p < 3; ++p) xc[p] = g.in[p * L.cs + o];
    T rb[3], db[3];
    epilogue_load<T, MODE>(g, o, rb, db);
    return grain_epilogue<T, MODE, TO>(g, o, xc, f, rb, db);
}

// Face blocks (GrainBox::xface): block p takes a FJ x FK patch of an x face
// (i = 0 for p < half the patches, else nx - 1) over j, k in [1, n - 2].  The
// face's nodes lie a row apart, so every load of a wavefront there touches 64
// lines; the block first stages the patch's two columns (the face and its
// neighbour) with their one-node (j, k) ring in LDS, each line once, then
// computes node_apply_acc's boundary row from LDS with node_grad's
// expressions (one-sided x, central y and z differences), so the bits are the
// shell's.  scratch: 3 * (FJ + 2) * (FK + 2) * 2 values of T.
constexpr int FJ = 16, FK = 16;
__host__ __device__ inline int grain_face_patches(const Lay& L) {
    return 2 * ((L.ny - 2 + FJ - 1) / FJ) * ((L.nz - 2 + FK - 1) / FK);
}
template <typename T, int MODE, typename TO, int SRC = kSrcGrain, int K = K_MM>
__device__ __forceinline__ double face_patch(const GrainBox<T>& g, int p, T* su) {
    const Lay& L = g.L;
    const int pj = (L.ny - 2 + FJ - 1) / FJ, pk = (L.nz - 2 + FK - 1) / FK;
    const int side = p / (pj * pk), r = p - side * (pj * pk);
    const int i = side ? L.nx - 1 : 0, sx = side ? -1 : 1;
    const int j0 = 1 + (r % pj) * FJ, k0 = 1 + (r / pj) * FK;
    constexpr int RJ = FJ + 2, RK = FK + 2, NS = RJ * RK * 2;
    const int tid = threadIdx.x;
    for (int e = tid; e < 3 * RJ * RK; e += MTX * MTY) {
        const int c = e / (RJ * RK), q = e - c * (RJ * RK), kk = q / RJ, jj = q - kk * RJ;
        const int j = j0 - 1 + jj, k = k0 - 1 + kk;
        if (j < L.ny && k < L.nz) {
            const T* up = g.in + c * L.cs + (long long)i + L.px * j + L.pl * k;
            su[c * NS + q * 2] = up[0];
            su[c * NS + q * 2 + 1] = up[sx];
        }
    }
    __syncthreads();
    const int jj = tid % FJ, kk = tid / FJ;
    const int j = j0 + jj, k = k0 + kk;
    if (kk >= FK || j > L.ny - 2 || k > L.nz - 2) return 0.0;
    const long long o = (long long)i + L.px * j + L.pl * k;
    T rb[3], db[3];
    epilogue_load<T, MODE>(g, o, rb, db);
    auto U = [&](int c, int dj, int dk, int col) { return su[c * NS + ((kk + 1 + dk) * RJ + (jj + 1 + dj)) * 2 + col]; };
    T gr[3][3], s[3][3], uc[3], f[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const T u0 = U(c, 0, 0, 0);
        uc[c] = u0;
        gr[c][0] = side ? (u0 - U(c, 0, 0, 1)) * g.iv.ih[0] : (U(c, 0, 0, 1) - u0) * g.iv.ih[0];
        gr[c][1] = (U(c, 1, 0, 0) - U(c, -1, 0, 0)) * g.iv.ih2[1];
        gr[c][2] = (U(c, 0, 1, 0) - U(c, 0, -1, 0)) * g.iv.ih2[2];
    }
    stress<T, K, 3>(box_acc<T, SRC>(g).rec(o), gr, s);
    bc_row<T, 3>(*g.bc, face_of<3>(i, j, k, L.nx - 1, L.ny - 1, L.nz - 1), uc, gr, s, f);
    return grain_epilogue<T, MODE, TO>(g, o, uc, f, rb, db);
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

// The marching kernels address a box's fields with 32-bit byte offsets from a
// plane's base (three components of FP64 at most), and per-node stiffness
// planes (nc planes of es bytes) from the planes' base.
__host__ __device__ inline bool grain_offsets_fit(const Lay& L) { return 3LL * L.cs * 8 < (1LL << 32); }
__host__ __device__ inline bool plane_offsets_fit(const Lay& L, int nc, int es) {
    return (long long)nc * L.cs * es < (1LL << 32);
}

// Tiles and shell blocks of a box (host and device); ry rows per thread.
__host__ __device__ inline void grain_tiles(const Lay& L, int kz, int& ntx, int& nty, int& ntz, int& nsh,
                                            bool xface = false, int ry = 1) {
    ntx = (L.nx - 1 - grain_x0(L) + MTX - 1) / MTX;
    nty = (L.ny - 2 + MTY * ry - 1) / (MTY * ry);
    ntz = (L.nz - 2 + kz - 1) / kz;
    const long long shell = (long long)L.nx * L.ny * L.nz - (long long)(L.nx - 2) * (L.ny - 2) * (L.nz - 2) -
                            (xface ? 2LL * (L.ny - 2) * (L.nz - 2) : 0LL);
    nsh = (int)((shell + 255) / 256);
}

}  // namespace ela
