// ela_device.hpp -- per-node math of the nodal elastic operator for gfx950.
//
// The discrete operator is the reference's (src/Operator/Elastic.cpp:151-260):
// on interior nodes f = C : grad grad u + sum_d (d_d C : grad u) e_d with
// central differences; on domain-boundary nodes the boundary row of
// BC::Operator::Elastic (src/BC/Operator/Elastic.H:222-303), with one-sided
// first differences on the boundary (src/Numeric/Stencil.H:16-102).
//
// Second and mixed differences are evaluated differences-first,
// ((u+ - u) + (u- - u)) / h^2 and ((u++ - u+-) - (u-+ - u--)) / (4 h_a h_b):
// the same stencil as u+ - 2u + u- but without the cancellation of two
// O(|u|) terms, which keeps the FP64 residual floor near eps * |h grad u|
// rather than eps * |u|.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ela {

enum { K_ISO = 0, K_MM = 1, K_DIAG = 2 };
enum { BC_DISP = 0, BC_TRAC = 1, BC_PER = 2, BC_NEU = 3 };

// One MG level: a single box of nodes, component-major, x fastest.
struct Lay {
    int nx, ny, nz;   // node counts (nz = 1 in 2D)
    int dim;
    long long px;     // pitch between j rows
    long long pl;     // pitch between k planes
    long long cs;     // component stride
};

// Boundary-condition types per face class and component (faces numbered as
// the reference's BC::Operator::Elastic::Face enum).
struct BCTab {
    unsigned char t[26][3];
};

// Inverse spacings used by the stencils.
template <typename T>
struct Inv {
    T ih[3];     // 1 / h
    T ih2[3];    // 0.5 / h       (central first difference)
    T ihh[3];    // 1 / h^2
    T imix[3];   // 0.25 / (h_a h_b) for the pairs (x,y), (x,z), (y,z)
};

// Stiffness accessors.  rec(o) is a view of node o's stiffness record, indexed
// by storage slot.  PlaneAcc: ncomp planes in the level layout (the
// reference's per-node model field).  TabAcc: a per-node grain id and one
// record per grain (nc values, contiguous); two nodes of the same grain have
// bitwise-identical records, so their central difference is exactly zero.
// GrainAcc: a grain table times an optional per-node factor s (a degraded
// polycrystal, s = g(eta)); the stiffness of node o is s[o] * tab[gid[o]].
template <typename T>
struct PlaneRec {
    const T* p;
    long long cs;
    __device__ __forceinline__ T operator[](int q) const { return p[q * cs]; }
};
template <typename T>
struct TabRec {
    const T* p;
    __device__ __forceinline__ T operator[](int q) const { return p[q]; }
};
template <typename T>
struct ScaledRec {
    const T* p;
    T s;
    __device__ __forceinline__ T operator[](int q) const { return s * p[q]; }
};

// same_dir(o, d, a, b): the stiffness records of nodes a = o + e_d and
// b = o - e_d are bitwise equal, so the d-derivative of C at o is exactly 0
// and its term may be skipped.  PlaneAcc never skips; MaskedPlaneAcc reads a
// per-node mask (bit d) built from the level's FP64 planes (equal FP64
// records stay equal in FP32).
template <typename T>
struct PlaneAcc {
    const T* C;
    long long cs;
    __device__ __forceinline__ PlaneRec<T> rec(long long o) const { return PlaneRec<T>{C + o, cs}; }
    __device__ __forceinline__ bool same_dir(long long, int, long long, long long) const { return false; }
};
template <typename T>
struct MaskedPlaneAcc {
    const T* C;
    long long cs;
    const unsigned char* sm;
    __device__ __forceinline__ PlaneRec<T> rec(long long o) const { return PlaneRec<T>{C + o, cs}; }
    __device__ __forceinline__ bool same_dir(long long o, int d, long long, long long) const { return (sm[o] >> d) & 1; }
};
template <typename T, int NC>
struct TabAcc {
    const unsigned short* gid;
    const T* tab;
    const T* dgt;  // per-grain interior diagonal (3 values; FP32: the inverse), may be null
    __device__ __forceinline__ TabRec<T> rec(long long o) const { return TabRec<T>{tab + (long long)gid[o] * NC}; }
    __device__ __forceinline__ bool same_dir(long long, int, long long a, long long b) const { return gid[a] == gid[b]; }
};
template <typename T, int NC>
struct GrainAcc {
    const unsigned short* gid;
    const T* tab;
    const T* s;  // per-node factor, may be null (factor 1)
    __device__ __forceinline__ ScaledRec<T> rec(long long o) const {
        return ScaledRec<T>{tab + (long long)gid[o] * NC, s ? s[o] : T(1)};
    }
    __device__ __forceinline__ bool same_dir(long long, int, long long a, long long b) const {
        return gid[a] == gid[b] && (!s || s[a] == s[b]);
    }
};

// Weighted Jacobi update x + omega r / D (Operator.cpp:119-157).  FP64 divides
// by the diagonal, as the reference does; the FP32 smoothers hold the inverse
// diagonal, 1/D rounded once from the FP64 value, and multiply (a division is
// a dozen instructions per component in the smoother's inner loop).
__device__ __forceinline__ double jacobi_update(double x, double omega, double r, double d) {
    return x + omega * r / d;
}
__device__ __forceinline__ float jacobi_update(float x, float omega, float r, float dinv) {
    return x + omega * r * dinv;
}
// The first sweep from x = 0: omega b / D.
__device__ __forceinline__ double jacobi_first(double omega, double b, double d) { return omega * b / d; }
__device__ __forceinline__ float jacobi_first(float omega, float b, float dinv) { return omega * b * dinv; }
// The value the smoothers of precision T store for diagonal d.
__device__ __forceinline__ void store_jacobi_diag(double* p, double d) { *p = d; }
__device__ __forceinline__ void store_jacobi_diag(float* p, double d) { *p = (float)(1.0 / d); }

// Jacobi diagonal of a node as the smoothers hold it (the inverse in FP32):
// the stored field, or the per-grain interior value.
template <typename T, typename Acc>
__device__ __forceinline__ T jdiag(const Acc&, const T* dg, long long q, long long, int, bool) {
    return dg[q];
}
template <typename T, int NC>
__device__ __forceinline__ T jdiag(const TabAcc<T, NC>& A, const T* dg, long long q, long long o, int p,
                                   bool interior) {
    return (interior && A.dgt) ? A.dgt[(long long)A.gid[o] * 3 + p] : dg[q];
}

template <int DIM, int K>
struct NComp;
template <int DIM> struct NComp<DIM, K_ISO> { static constexpr int v = 2; };
template <> struct NComp<3, K_MM> { static constexpr int v = 21; };
template <> struct NComp<2, K_MM> { static constexpr int v = 6; };
template <int DIM> struct NComp<DIM, K_DIAG> { static constexpr int v = DIM * DIM; };

// Voigt index of a symmetric pair, order (00,01,02,11,12,22) in 3D and (00,01,11) in 2D.
template <int DIM>
__host__ __device__ constexpr int voigt(int i, int j) {
    return DIM == 3 ? ((i <= j) ? (i == 0 ? j : (i == 1 ? 2 + j : 5)) : (j == 0 ? i : (j == 1 ? 2 + i : 5)))
                    : (i + j);
}

// Slot of C_ijkl in the upper-triangular Voigt storage.
template <int DIM>
__host__ __device__ constexpr int mm_slot(int i, int j, int k, int l) {
    const int nv = DIM == 3 ? 6 : 3;
    const int a = voigt<DIM>(i, j), b = voigt<DIM>(k, l);
    const int I = a < b ? a : b, J = a < b ? b : a;
    return I * nv - I * (I - 1) / 2 + (J - I);
}

// Face classification with the reference's precedence (corners, then the YZ,
// ZX and XY edges, then faces; BC/Operator/Elastic.H:239-278).  Returns -1 for
// interior nodes.
template <int DIM>
__host__ __device__ __forceinline__ int face_of(int i, int j, int k, int hx, int hy, int hz) {
    const bool xl = i == 0, xh = i == hx, yl = j == 0, yh = j == hy;
    if (DIM == 2) {
        if (xl && yl) return 4;
        if (xl && yh) return 5;
        if (xh && yl) return 6;
        if (xh && yh) return 7;
        if (xl) return 0;
        if (xh) return 2;
        if (yl) return 1;
        if (yh) return 3;
        return -1;
    }
    const bool zl = k == 0, zh = k == hz;
    if ((xl || xh) && (yl || yh) && (zl || zh)) return 18 + (xh ? 4 : 0) + (yh ? 2 : 0) + (zh ? 1 : 0);
    if ((yl || yh) && (zl || zh)) return 6 + (yh ? 2 : 0) + (zh ? 1 : 0);
    if ((zl || zh) && (xl || xh)) return 10 + (zh ? 2 : 0) + (xh ? 1 : 0);
    if ((xl || xh) && (yl || yh)) return 14 + (xh ? 2 : 0) + (yh ? 1 : 0);
    if (xl) return 0;
    if (xh) return 3;
    if (yl) return 1;
    if (yh) return 4;
    if (zl) return 2;
    if (zh) return 5;
    return -1;
}

// Outward normal of a face class; the reference's truncated 1/sqrt(3) and
// 1/sqrt(2) (BC/Operator/Elastic.H:215-216) on corners and edges.
template <typename T, int DIM>
__device__ __forceinline__ void normal_of(int face, T n[3]) {
    const T s3 = T(0.57735026919), s2 = T(0.70710678118);
    n[0] = n[1] = n[2] = T(0);
    if (DIM == 2) {
        switch (face) {
            case 0: n[0] = -1; break;
            case 1: n[1] = -1; break;
            case 2: n[0] = 1; break;
            case 3: n[1] = 1; break;
            default: {
                const int c = face - 4;  // XLO_YLO, XLO_YHI, XHI_YLO, XHI_YHI
                n[0] = (c & 2) ? s2 : -s2;
                n[1] = (c & 1) ? s2 : -s2;
            }
        }
        return;
    }
    if (face < 6) {
        n[face % 3] = face < 3 ? T(-1) : T(1);
    } else if (face < 10) {  // Y?_Z?
        const int c = face - 6;
        n[1] = (c & 2) ? s2 : -s2;
        n[2] = (c & 1) ? s2 : -s2;
    } else if (face < 14) {  // Z?_X?
        const int c = face - 10;
        n[2] = (c & 2) ? s2 : -s2;
        n[0] = (c & 1) ? s2 : -s2;
    } else if (face < 18) {  // X?_Y?
        const int c = face - 14;
        n[0] = (c & 2) ? s2 : -s2;
        n[1] = (c & 1) ? s2 : -s2;
    } else {
        const int c = face - 18;
        n[0] = (c & 4) ? s3 : -s3;
        n[1] = (c & 2) ? s3 : -s3;
        n[2] = (c & 1) ? s3 : -s3;
    }
}

// sigma_ij = C_ijkl g_kl at one node; c is the node's stiffness record.
template <typename T, int K, int DIM, typename Rec>
__device__ __forceinline__ void stress(const Rec& c, const T g[DIM][DIM], T s[DIM][DIM]) {
    if (K == K_ISO) {
        const T lam = c[0], mu = c[1];
        T tr = 0;
#pragma unroll
        for (int a = 0; a < DIM; ++a) tr += g[a][a];
        // explicit fused products: every kernel that inlines this rounds the
        // same way (a free a*b + c*d may be contracted either way)
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int b = 0; b < DIM; ++b)
                s[a][b] = a == b ? fma(mu, g[a][b] + g[b][a], lam * tr) : mu * (g[a][b] + g[b][a]);
    } else if (K == K_MM) {
        T gs[DIM][DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k)
#pragma unroll
            for (int l = 0; l < DIM; ++l) gs[k][l] = (k == l) ? g[k][k] : g[k][l] + g[l][k];
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int b = a; b < DIM; ++b) {
                T acc = 0;
#pragma unroll
                for (int k = 0; k < DIM; ++k)
#pragma unroll
                    for (int l = k; l < DIM; ++l) acc += c[mm_slot<DIM>(a, b, k, l)] * gs[k][l];
                s[a][b] = acc;
                s[b][a] = acc;
            }
    } else {
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int b = 0; b < DIM; ++b) s[a][b] = c[a * DIM + b] * g[a][b];
    }
}

// Boundary row (BC/Operator/Elastic.H:286-303).
template <typename T, int DIM>
__device__ __forceinline__ void bc_row(const BCTab& bc, int face, const T u[DIM], const T g[DIM][DIM],
                                       const T s[DIM][DIM], T f[DIM]) {
    T n[3];
    normal_of<T, DIM>(face, n);
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
        const int t = bc.t[face][a];
        T v = 0;
        if (t == BC_DISP) {
            v = u[a];
        } else if (t == BC_TRAC) {
#pragma unroll
            for (int b = 0; b < DIM; ++b) v += s[a][b] * n[b];
        } else if (t == BC_NEU) {
#pragma unroll
            for (int b = 0; b < DIM; ++b) v += g[a][b] * n[b];
        }
        f[a] = v;
    }
}

// Interior row of A (Elastic.cpp:151-260): f = C : grad grad u + sum_d (d_d C :
// grad u) e_d with central differences, on a stencil context S that provides
//   S.u(p, dx, dy, dz)  component p of u at the neighbour (dx, dy, dz),
//   S.rec(dx, dy, dz)   the stiffness record at +-e_d,
//   S.center()          the node's own record,
//   S.same(d)           whether the records at +e_d and -e_d are bitwise equal
//                       (their central difference is then exactly zero and is
//                       skipped).  Every
// context evaluates the same expressions in the same order, so a node gets
// the same bits whether u comes from global memory or from an LDS tile.
template <typename T, int K, int DIM, typename Ctx>
__device__ __forceinline__ void interior_apply(const Ctx& S, const Inv<T>& iv, bool uniform, T f[DIM]) {
    T g[DIM][DIM], gg[DIM][DIM][DIM];  // gg[p][a][b] = u_{p,ab}
#pragma unroll
    for (int p = 0; p < DIM; ++p) {
        const T c0 = S.u(p, 0, 0, 0);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const T vp = S.u(p, d == 0, d == 1, d == 2), vm = S.u(p, -(d == 0), -(d == 1), -(d == 2));
            g[p][d] = (vp - vm) * iv.ih2[d];
            gg[p][d][d] = ((vp - c0) + (vm - c0)) * iv.ihh[d];
        }
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int b = a + 1; b < DIM; ++b) {
                const int ax = a == 0, ay = a == 1, az = a == 2, bx = b == 0, by = b == 1, bz = b == 2;
                const T pp = S.u(p, ax + bx, ay + by, az + bz), pm = S.u(p, ax - bx, ay - by, az - bz);
                const T mp = S.u(p, -ax + bx, -ay + by, -az + bz), mm = S.u(p, -ax - bx, -ay - by, -az - bz);
                const T m = ((pp - pm) - (mp - mm)) * iv.imix[a + b - 1];
                gg[p][a][b] = m;
                gg[p][b][a] = m;
            }
    }
    const auto c = S.center();
    if (K == K_ISO) {
        const T lam = c[0], mu = c[1];
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            T lap = 0, gd = 0;
#pragma unroll
            for (int b = 0; b < DIM; ++b) {
                lap += gg[a][b][b];
                gd += gg[b][a][b];
            }
            f[a] = fma(mu, lap, (mu + lam) * gd);  // fused explicitly, as in stress
        }
        if (!uniform) {
            T tr = 0;
#pragma unroll
            for (int a = 0; a < DIM; ++a) tr += g[a][a];
            T dl[DIM], dm[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                dl[d] = dm[d] = T(0);
                if (!S.same(d)) {
                    const auto cp = S.rec(d == 0, d == 1, d == 2);
                    const auto cm = S.rec(-(d == 0), -(d == 1), -(d == 2));
                    dl[d] = (cp[0] - cm[0]) * iv.ih2[d];
                    dm[d] = (cp[1] - cm[1]) * iv.ih2[d];
                }
            }
#pragma unroll
            for (int a = 0; a < DIM; ++a) {
                T add = dl[a] * tr;
#pragma unroll
                for (int d = 0; d < DIM; ++d) add = fma(dm[d], g[a][d] + g[d][a], add);
                f[a] += add;
            }
        }
    } else if (K == K_MM) {
        // sum over b, k, l of C_abkl u_k,lb.  C_abkl = C_ablk (minor symmetry), so the pair (k, l), k < l,
        // takes one product with u_k,lb + u_l,kb: per (a, b) DIM + DIM (DIM - 1) / 2 products instead of DIM^2
        constexpr int NP = DIM * (DIM - 1) / 2;
        T sp[NP][DIM];
#pragma unroll
        for (int b = 0; b < DIM; ++b) {
            int q = 0;
#pragma unroll
            for (int kk = 0; kk < DIM; ++kk)
#pragma unroll
                for (int l = kk + 1; l < DIM; ++l) sp[q++][b] = gg[kk][l][b] + gg[l][kk][b];
        }
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            T acc = 0;
#pragma unroll
            for (int b = 0; b < DIM; ++b) {
#pragma unroll
                for (int kk = 0; kk < DIM; ++kk) acc += c[mm_slot<DIM>(a, b, kk, kk)] * gg[kk][kk][b];
                int q = 0;
#pragma unroll
                for (int kk = 0; kk < DIM; ++kk)
#pragma unroll
                    for (int l = kk + 1; l < DIM; ++l) acc += c[mm_slot<DIM>(a, b, kk, l)] * sp[q++][b];
            }
            f[a] = acc;
        }
        if (!uniform) {
            T gs[DIM][DIM];
#pragma unroll
            for (int kk = 0; kk < DIM; ++kk)
#pragma unroll
                for (int l = 0; l < DIM; ++l) gs[kk][l] = (kk == l) ? g[kk][kk] : g[kk][l] + g[l][kk];
#pragma unroll
            for (int a = 0; a < DIM; ++a) {
                T add = 0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    if (S.same(d)) continue;
                    const auto cp = S.rec(d == 0, d == 1, d == 2);
                    const auto cm = S.rec(-(d == 0), -(d == 1), -(d == 2));
                    T acc = 0;
#pragma unroll
                    for (int kk = 0; kk < DIM; ++kk)
#pragma unroll
                        for (int l = kk; l < DIM; ++l) {
                            const int q = mm_slot<DIM>(a, d, kk, l);
                            acc += (cp[q] - cm[q]) * gs[kk][l];
                        }
                    add += acc * iv.ih2[d];
                }
                f[a] += add;
            }
        }
    } else {
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            T acc = 0;
#pragma unroll
            for (int b = 0; b < DIM; ++b) acc += c[a * DIM + b] * gg[a][b][b];
            f[a] = acc;
        }
        if (!uniform) {
#pragma unroll
            for (int a = 0; a < DIM; ++a) {
                T add = 0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    if (S.same(d)) continue;
                    const int q = a * DIM + d;
                    add += (S.rec(d == 0, d == 1, d == 2)[q] - S.rec(-(d == 0), -(d == 1), -(d == 2))[q]) * iv.ih2[d] *
                           g[a][d];
                }
                f[a] += add;
            }
        }
    }
}

// Stencil context on global memory: u in the level layout, stiffness through an accessor.
template <typename T, typename Acc>
struct GlobalCtx {
    const T* __restrict__ up;
    long long cs, o, sx, sy, sz;
    const Acc& A;
    __device__ __forceinline__ long long at(int dx, int dy, int dz) const { return o + dx * sx + dy * sy + dz * sz; }
    __device__ __forceinline__ T u(int p, int dx, int dy, int dz) const { return up[p * cs + at(dx, dy, dz)]; }
    __device__ __forceinline__ auto rec(int dx, int dy, int dz) const { return A.rec(at(dx, dy, dz)); }
    __device__ __forceinline__ auto center() const { return rec(0, 0, 0); }
    __device__ __forceinline__ bool same(int d) const {
        return A.same_dir(o, d, at(d == 0, d == 1, d == 2), at(-(d == 0), -(d == 1), -(d == 2)));
    }
};

// grad u at node (i,j,k), g[p][d] = du_p/dx_d: central differences inside,
// one-sided on domain walls (Numeric::GetStencil with the domain box,
// src/Numeric/Stencil.H:16-102).
template <typename T, int DIM>
__device__ __forceinline__ void node_grad(const Lay& L, const T* __restrict__ u, int i, int j, int k,
                                          const Inv<T>& iv, T g[DIM][DIM]) {
    const long long o = (long long)i + L.px * j + L.pl * k;
    const long long st[3] = {1, L.px, L.pl};
    const int idx[3] = {i, j, k};
    const int top[3] = {L.nx - 1, L.ny - 1, L.nz - 1};
#pragma unroll
    for (int p = 0; p < DIM; ++p) {
        const T* up = u + p * L.cs + o;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            if (idx[d] == 0) g[p][d] = (up[st[d]] - up[0]) * iv.ih[d];
            else if (idx[d] == top[d]) g[p][d] = (up[0] - up[-st[d]]) * iv.ih[d];
            else g[p][d] = (up[st[d]] - up[-st[d]]) * iv.ih2[d];
        }
    }
}

// f = A u at node (i,j,k).  C points at plane 0 of the stiffness, u at comp 0.
template <typename T, int K, int DIM, typename Acc>
__device__ __forceinline__ void node_apply_acc(const Lay& L, const Acc& A, const T* __restrict__ u, int i, int j,
                                               int k, const BCTab& bc, const Inv<T>& iv, bool uniform, T f[DIM]) {
    const long long o = (long long)i + L.px * j + L.pl * k;
    const int hx = L.nx - 1, hy = L.ny - 1, hz = L.nz - 1;
    const bool bnd = (i == 0 || i == hx || j == 0 || j == hy || (DIM == 3 && (k == 0 || k == hz)));
    const long long st[3] = {1, L.px, L.pl};
    const int idx[3] = {i, j, k};
    const int top[3] = {hx, hy, hz};
    if (bnd) {
        T uc[DIM], g[DIM][DIM], s[DIM][DIM];
        node_grad<T, DIM>(L, u, i, j, k, iv, g);
#pragma unroll
        for (int p = 0; p < DIM; ++p) uc[p] = u[p * L.cs + o];
        stress<T, K, DIM>(A.rec(o), g, s);
        bc_row<T, DIM>(bc, face_of<DIM>(i, j, k, hx, hy, hz), uc, g, s, f);
        return;
    }
    const GlobalCtx<T, Acc> S{u, L.cs, o, st[0], st[1], st[2], A};
    interior_apply<T, K, DIM>(S, iv, uniform, f);
}

template <typename T, int K, int DIM>
__device__ __forceinline__ void node_apply(const Lay& L, const T* __restrict__ C, const T* __restrict__ u, int i,
                                           int j, int k, const BCTab& bc, const Inv<T>& iv, bool uniform, T f[DIM]) {
    node_apply_acc<T, K, DIM>(L, PlaneAcc<T>{C, L.cs}, u, i, j, k, bc, iv, uniform, f);
}

// Diagonal of A at node (i,j,k) (Elastic.cpp:285-349), FP64.
template <int K, int DIM, typename Acc>
__device__ __forceinline__ void node_diag_acc(const Lay& L, const Acc& A, int i, int j, int k, const BCTab& bc,
                                              const double h[3], double dg[DIM]) {
    const long long o = (long long)i + L.px * j + L.pl * k;
    const int hx = L.nx - 1, hy = L.ny - 1, hz = L.nz - 1;
    const bool mn[3] = {i == 0, j == 0, DIM == 3 && k == 0};
    const bool mx[3] = {i == hx, j == hy, DIM == 3 && k == hz};
    const bool bnd = mn[0] || mx[0] || mn[1] || mx[1] || mn[2] || mx[2];
    const auto c = A.rec(o);
#pragma unroll
    for (int p = 0; p < DIM; ++p) {
        if (bnd) {
            double g[DIM][DIM], s[DIM][DIM], uu[DIM], f[DIM];
#pragma unroll
            for (int q = 0; q < DIM; ++q) {
                uu[q] = (q == p) ? 1.0 : 0.0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    const double e = (p == q) ? 1.0 : 0.0;
                    g[q][d] = ((!mx[d] ? 0.0 : e) - (!mn[d] ? 0.0 : e)) / ((mn[d] || mx[d] ? 1.0 : 2.0) * h[d]);
                }
            }
            stress<double, K, DIM>(c, g, s);
            bc_row<double, DIM>(bc, face_of<DIM>(i, j, k, hx, hy, hz), uu, g, s, f);
            dg[p] = f[p];
        } else {
            // f_p = sum_j C_{p j p j} (-2 / h_j^2): the only entries of C : grad grad e_p at the node.
            double acc = 0;
#pragma unroll
            for (int b = 0; b < DIM; ++b) {
                double cpp;
                if (K == K_ISO) cpp = c[1] + (b == p ? c[1] + c[0] : 0.0);
                else if (K == K_MM) cpp = c[mm_slot<DIM>(p, b, p, b)];
                else cpp = c[p * DIM + b];
                acc += cpp * (-2.0 / h[b] / h[b]);
            }
            dg[p] = acc;
        }
    }
}

// The interior diagonal of component p from a major/minor record r alone
// (node_diag_acc's interior branch, the same operations): a grain's, per level.
__device__ __forceinline__ double grain_diag(const double* r, const double h[3], int p) {
    double acc = 0;
#pragma unroll
    for (int b = 0; b < 3; ++b) acc += r[mm_slot<3>(p, b, p, b)] * (-2.0 / h[b] / h[b]);
    return acc;
}

template <int K, int DIM>
__device__ __forceinline__ void node_diag(const Lay& L, const double* __restrict__ C, int i, int j, int k,
                                          const BCTab& bc, const double h[3], double dg[DIM]) {
    node_diag_acc<K, DIM>(L, PlaneAcc<double>{C, L.cs}, i, j, k, bc, h, dg);
}

// Nodal full weighting onto coarse node (I,J,K) of the coarse level Lc
// (Operator.cpp:486-543; the same stencil coarsens the stiffness,
// Elastic.cpp:759-803, and refluxes the fine residual, Operator.cpp:779-822).
// F(a, b, c) returns the fine value at fine node (2I+a, 2J+b, 2K+c).  The
// branches follow the coarse level's domain boundary; in 2D the k direction
// counts as "on the boundary" (the reference's degenerate z box), which
// selects the 2D branches.
template <typename T, int DIM, typename Get>
__device__ __forceinline__ T restrict_node(const Lay& Lc, int I, int J, int K, const Get& F) {
    const bool bI = I == 0 || I == Lc.nx - 1, bJ = J == 0 || J == Lc.ny - 1;
    const bool bK = DIM == 2 ? true : (K == 0 || K == Lc.nz - 1);
    if (bI && bJ && bK) return F(0, 0, 0);
    if (bJ && bK) return T(0.25) * F(-1, 0, 0) + T(0.5) * F(0, 0, 0) + T(0.25) * F(1, 0, 0);
    if (bK && bI) return T(0.25) * F(0, -1, 0) + T(0.5) * F(0, 0, 0) + T(0.25) * F(0, 1, 0);
    if (bI && bJ) return T(0.25) * F(0, 0, -1) + T(0.5) * F(0, 0, 0) + T(0.25) * F(0, 0, 1);
    if (bI || bJ || bK) {
        // face: (1 2 1) x (1 2 1) / 16 over the two tangential directions u, w
        const int ux = bI ? 0 : 1, uy = bI ? 1 : 0;              // u: y on an x face, else x
        const int wy = bK ? 1 : 0, wz = bK ? 0 : 1;              // w: y on a z face, else z
        auto G = [&](int a, int c) { return F(a * ux, a * uy + c * wy, c * wz); };
        return (G(-1, -1) + T(2) * G(0, -1) + G(1, -1) + T(2) * G(-1, 0) + T(4) * G(0, 0) + T(2) * G(1, 0) +
                G(-1, 1) + T(2) * G(0, 1) + G(1, 1)) /
               T(16);
    }
    T corners = 0, edges = 0, faces = 0;
#pragma unroll
    for (int a = -1; a <= 1; a += 2)
#pragma unroll
        for (int bb = -1; bb <= 1; bb += 2)
#pragma unroll
            for (int c = -1; c <= 1; c += 2) corners += F(a, bb, c);
#pragma unroll
    for (int a = -1; a <= 1; a += 2)
#pragma unroll
        for (int bb = -1; bb <= 1; bb += 2) edges += F(0, a, bb) + F(a, 0, bb) + F(a, bb, 0);
#pragma unroll
    for (int a = -1; a <= 1; a += 2) faces += F(a, 0, 0) + F(0, a, 0) + F(0, 0, a);
    return corners / T(64) + edges / T(32) + faces / T(16) + F(0, 0, 0) / T(8);
}

// Trilinear nodal prolongation (Operator.cpp:589-618): the coarse-grid
// interpolant at fine node (i,j,k); c points at coarse node (i/2, j/2, k/2).
template <typename T>
__device__ __forceinline__ T prolong_node(const Lay& Lc, int i, int j, int k, const T* __restrict__ c) {
    const bool ei = (i & 1) == 0, ej = (j & 1) == 0, ek = (k & 1) == 0;
    const long long sx = 1, sy = Lc.px, sz = Lc.pl;
    if (ei && ej && ek) return c[0];
    if (ej && ek) return T(0.5) * (c[0] + c[sx]);
    if (ek && ei) return T(0.5) * (c[0] + c[sy]);
    if (ei && ej) return T(0.5) * (c[0] + c[sz]);
    if (ei) return T(0.25) * (c[0] + c[sy] + c[sz] + c[sy + sz]);
    if (ej) return T(0.25) * (c[0] + c[sz] + c[sx] + c[sx + sz]);
    if (ek) return T(0.25) * (c[0] + c[sx] + c[sy] + c[sx + sy]);
    return T(0.125) * (c[0] + c[sx] + c[sy] + c[sz] + c[sy + sz] + c[sx + sz] + c[sx + sy] + c[sx + sy + sz]);
}

// Gauss-Jordan step for pivot column c on M = [A | A^-1-in-progress] (n x 2n,
// row-major), executed by one block: find the pivot row (largest |M[r][c]|,
// r >= c, lowest index on ties), swap it into row c and scale row c.  sval /
// sidx are blockDim.x entries of LDS.  Sets *flag on a zero pivot.
__device__ __forceinline__ void gj_pivot_block(double* __restrict__ M, int n, int c, int* flag, double* sval,
                                               int* sidx) {
    const int tid = threadIdx.x, nt = blockDim.x, w = 2 * n;
    double best = -1.0;
    int bi = c;
    for (int r = c + tid; r < n; r += nt) {
        const double a = fabs(M[(long long)r * w + c]);
        if (a > best) { best = a; bi = r; }
    }
    sval[tid] = best;
    sidx[tid] = bi;
    __syncthreads();
    for (int st = nt / 2; st > 0; st >>= 1) {
        if (tid < st && (sval[tid + st] > sval[tid] || (sval[tid + st] == sval[tid] && sidx[tid + st] < sidx[tid]))) {
            sval[tid] = sval[tid + st];
            sidx[tid] = sidx[tid + st];
        }
        __syncthreads();
    }
    const int p = sidx[0];
    if (tid == 0 && !(sval[0] > 0.0)) *flag = 1;
    const double d = M[(long long)p * w + c];
    __syncthreads();
    for (int q = tid; q < w; q += nt) {
        const double vp = M[(long long)p * w + q];
        if (p != c) M[(long long)p * w + q] = M[(long long)c * w + q];
        M[(long long)c * w + q] = vp / d;
    }
}

// Row r != c of the same step: M[r][q] -= M[r][c] * M[c][q], by one block.
// The multiplier is read before any thread of the block writes column c.
__device__ __forceinline__ void gj_eliminate_row(double* __restrict__ M, int n, int c, int r) {
    const int w = 2 * n;
    if (r == c) return;
    const double m = M[(long long)r * w + c];
    __syncthreads();
    if (m == 0.0) return;
    for (int q = threadIdx.x; q < w; q += blockDim.x) M[(long long)r * w + q] -= m * M[(long long)c * w + q];
}

// Column c of Gauss-Jordan on [A | I] (n rows of 2n) by the block that owns
// matrix row r, with the pivot search in every block: the pivots of the
// pivot/eliminate step pair (the largest |M[r][c]| over the rows not yet
// pivots, ties to the first in the swapped order) and the same operations on
// every element they keep, so the same bits.  The row swaps are a permutation
// (perm_in -> perm_out, written by row 0's block; perm_in null: the
// identity); the pivot row is divided by its pivot where it is read and, in
// the matrix, by its owner one column later (dpiv[c] keeps each pivot); the
// left half is updated right of the pivot column only (the columns left of it
// are never read again), so no element one block writes is read by another
// in the same step.  cand_in holds |M[row][c]| by matrix row (null: read
// column c itself), and each row's owner writes |M[r][c+1]| after its update
// into cand_out, so the next column's search reads n contiguous values
// instead of a strided column.  256 threads; sval, sidx: 256 entries of LDS.
__device__ __forceinline__ void gj_column(double* __restrict__ M, int n, int c, const int* __restrict__ perm_in,
                                          int* __restrict__ perm_out, double* __restrict__ dpiv,
                                          const double* __restrict__ cand_in, double* __restrict__ cand_out,
                                          int* flag, int r, double* sval, int* sidx) {
    const int tid = threadIdx.x;
    const long long w = 2LL * n;
    auto pin = [&](int i) { return perm_in ? perm_in[i] : i; };
    double best = -1.0;
    int bi = c;
    for (int i = c + tid; i < n; i += 256) {
        const int pr = pin(i);
        const double a = cand_in ? cand_in[pr] : fabs(M[(long long)pr * w + c]);
        if (a > best) { best = a; bi = i; }
    }
    sval[tid] = best;
    sidx[tid] = bi;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st && (sval[tid + st] > sval[tid] || (sval[tid + st] == sval[tid] && sidx[tid + st] < sidx[tid]))) {
            sval[tid] = sval[tid + st];
            sidx[tid] = sidx[tid + st];
        }
        __syncthreads();
    }
    const int pi = sidx[0];
    const int prow = pin(pi);
    const double d = M[(long long)prow * w + c];
    if (r == 0) {
        for (int i = tid; i < n; i += 256) perm_out[i] = i == c ? prow : (i == pi ? pin(c) : pin(i));
        if (tid == 0) {
            dpiv[c] = d;
            if (!(sval[0] > 0.0)) *flag = 1;
        }
    }
    if (r == prow) return;
    double* Mr = M + (long long)r * w;
    const double* P = M + (long long)prow * w;
    const int nl = n - 1 - c, nq = nl + n;  // columns kept: c+1 .. n-1 and n .. 2n-1
    if (c > 0 && r == pin(c - 1)) {  // last column's pivot row: its division, then this column's elimination
        const double dp = dpiv[c - 1];
        const double m = Mr[c] / dp;
        for (int t = tid; t < nq; t += 256) {
            const int q = t < nl ? c + 1 + t : n + (t - nl);
            const double s = Mr[q] / dp;
            Mr[q] = m == 0.0 ? s : s - m * (P[q] / d);
        }
    } else {
        const double m = Mr[c];
        if (m != 0.0)
            for (int t = tid; t < nq; t += 256) {
                const int q = t < nl ? c + 1 + t : n + (t - nl);
                Mr[q] -= m * (P[q] / d);
            }
    }
    // thread 0 updated column c+1 of this row itself (t = 0)
    if (tid == 0 && cand_out && nl > 0) cand_out[r] = fabs(Mr[c + 1]);
}

}  // namespace ela
