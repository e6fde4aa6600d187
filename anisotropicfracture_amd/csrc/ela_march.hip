// ela_march.hip -- LDS-staged, z-marching stencil kernels for 3D levels.
//
// The operator (f = A u) and the two sweeps built on it (weighted Jacobi and
// the residual) are split into
//   * an interior kernel: a 64 x TY block owns an (x, y) tile of interior
//     nodes and marches up a chunk of z planes.  u sits in a 3-plane LDS ring
//     with a one-node x/y halo (the next plane is prefetched into registers
//     while the current one is computed); the stiffness of the node's own
//     column is carried in registers from plane to plane (its z-1 record is
//     the previous step's centre); x/y neighbours' stiffness records are read
//     through the vector L1, where the neighbouring threads of the same step
//     have just brought them;
//   * a boundary kernel over the six faces (one thread per boundary node),
//     so interior wavefronts never diverge on boundary rows.
// Blocks are numbered XCD-major: consecutive logical tiles land on the same
// XCD (hardware dispatch is round-robin over the 8 XCDs), so the halo lines
// two neighbouring tiles share stay in one L2.
#include <hip/hip_runtime.h>

#include "ela_march.hpp"

namespace ela {

namespace {

constexpr int TX = 64;

template <typename T, typename TO>
struct OpArgs {
    const T* C;
    const T* u;
    T* out;              // fapply: f; jacobi: x_out; residual: r (may be null)
    TO* outs;            // residual: r * (*scale) (may be null)
    const T* b;
    const T* dg;
    const double* scale;
    double* partials;    // residual: per-block max |r| (may be null)
    T omega;
};

// Epilogue of one node: mode 0 f = A u, 1 Jacobi, 2 residual.
template <int MODE, typename T, typename TO>
__device__ __forceinline__ void epilogue(const OpArgs<T, TO>& a, long long cs, long long o, const T xc[3],
                                         const T f[3], double& m, double s) {
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const long long q = p * cs + o;
        if (MODE == 0) {
            a.out[q] = f[p];
        } else if (MODE == 1) {
            a.out[q] = jacobi_update(xc[p], a.omega, a.b[q] - f[p], a.dg[q]);
        } else {
            const T v = a.b[q] - f[p];
            if (a.out) a.out[q] = v;
            if (a.outs) a.outs[q] = (TO)((double)v * s);
            const double av = fabs((double)v);
            m = (av > m || av != av) ? av : m;
        }
    }
}

__device__ __forceinline__ void block_max_store2(double* partials, long long slot, double v) {
    __shared__ double wmax[16];
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o > v || o != o) ? o : v;
    }
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int nw = (blockDim.x * blockDim.y + 63) / 64;
    __syncthreads();
    if ((tid & 63) == 0) wmax[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double mm = wmax[0];
        for (int w = 1; w < nw; ++w) mm = (wmax[w] > mm || wmax[w] != wmax[w]) ? wmax[w] : mm;
        partials[slot] = mm;
    }
}

// XCD-major block numbering: hardware block id b runs on XCD b % 8; give the
// blocks of one XCD a contiguous range of logical ids.
__device__ __forceinline__ int logical_block(int b, int nb) {
    const int per = (nb + 7) / 8;
    const int xcd = b % 8, r = b / 8;
    const int full = nb % 8 == 0 ? 8 : nb % 8;  // XCDs that receive `per` blocks
    if (xcd < full) return xcd * per + r;
    return full * per + (xcd - full) * (per - 1) + r;
}

template <typename T, int K, int TY, int MODE, typename TO>
__global__ __launch_bounds__(TX* TY) void k_march(Lay L, OpArgs<T, TO> a, Inv<T> iv, int uniform, int kz, int ntx,
                                                  int nty, int ntz) {
    constexpr int NC = NComp<3, K>::v;
    constexpr int RW = TY + 2, CW = TX + 2;
    __shared__ T su[3][3][RW][CW];  // [ring slot][comp][row][col]

    const int nb = ntx * nty * ntz;
    const int lb = logical_block(blockIdx.x, nb);
    const int bx = lb % ntx, by = (lb / ntx) % nty, bz = lb / (ntx * nty);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = tx + TX * ty;
    const int i0 = 1 + bx * TX, j0 = 1 + by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const bool active = i <= L.nx - 2 && j <= L.ny - 2;
    const int kb = 1 + bz * kz;
    const int ke = min(kb + kz, L.nz - 1);  // exclusive; interior planes 1..nz-2
    const long long cs = L.cs, px = L.px, pl = L.pl;
    const double s = (MODE == 2 && a.outs) ? *a.scale : 1.0;
    double m = 0.0;

    // cooperative plane load of u (tile + halo) for plane kk
    constexpr int PL = 3 * RW * CW;
    constexpr int PER = (PL + TX * TY - 1) / (TX * TY);
    auto load_plane = [&](int kk, T reg[PER]) {
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int q = tid + r * TX * TY;
            T v = 0;
            if (q < PL) {
                const int p = q / (RW * CW), rem = q % (RW * CW);
                const int rr = rem / CW, cc = rem % CW;
                const int ii = i0 - 1 + cc, jj = j0 - 1 + rr;
                if (ii < L.nx && jj < L.ny && kk < L.nz) v = a.u[p * cs + ii + px * jj + pl * kk];
            }
            reg[r] = v;
        }
    };
    auto store_plane = [&](int slot, const T reg[PER]) {
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int q = tid + r * TX * TY;
            if (q < PL) {
                const int p = q / (RW * CW), rem = q % (RW * CW);
                (&su[slot][p][0][0])[rem] = reg[r];
            }
        }
    };

    T pre[PER];
    load_plane(kb - 1, pre);
    store_plane((kb - 1) % 3, pre);
    load_plane(kb, pre);
    store_plane(kb % 3, pre);
    load_plane(kb + 1, pre);
    store_plane((kb + 1) % 3, pre);

    const long long oc0 = (long long)i + px * j;
    T Cm[NC];  // stiffness of (i, j, k-1)
    if (active) {
#pragma unroll
        for (int q = 0; q < NC; ++q) Cm[q] = a.C[q * cs + oc0 + pl * (kb - 1)];
    }
    __syncthreads();

    for (int k = kb; k < ke; ++k) {
        // prefetch plane k+2 (consumed next step)
        const bool more = k + 1 < ke;
        if (more) load_plane(k + 2, pre);
        const int sm = (k - 1) % 3, s0 = k % 3, sp = (k + 1) % 3;
        const long long o = oc0 + pl * k;
        T Cc[NC], Cp[NC];
        if (active) {
#pragma unroll
            for (int q = 0; q < NC; ++q) Cc[q] = a.C[q * cs + o];
#pragma unroll
            for (int q = 0; q < NC; ++q) Cp[q] = a.C[q * cs + o + pl];

            T g[3][3], gg[3][3][3], xc[3];
            const int r = ty + 1, c = tx + 1;
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                const T c0 = su[s0][p][r][c];
                xc[p] = c0;
                const T xp = su[s0][p][r][c + 1], xm = su[s0][p][r][c - 1];
                const T yp = su[s0][p][r + 1][c], ym = su[s0][p][r - 1][c];
                const T zp = su[sp][p][r][c], zm = su[sm][p][r][c];
                g[p][0] = (xp - xm) * iv.ih2[0];
                g[p][1] = (yp - ym) * iv.ih2[1];
                g[p][2] = (zp - zm) * iv.ih2[2];
                gg[p][0][0] = ((xp - c0) + (xm - c0)) * iv.ihh[0];
                gg[p][1][1] = ((yp - c0) + (ym - c0)) * iv.ihh[1];
                gg[p][2][2] = ((zp - c0) + (zm - c0)) * iv.ihh[2];
                const T mxy = ((su[s0][p][r + 1][c + 1] - su[s0][p][r - 1][c + 1]) -
                               (su[s0][p][r + 1][c - 1] - su[s0][p][r - 1][c - 1])) * iv.imix[0];
                const T mxz = ((su[sp][p][r][c + 1] - su[sm][p][r][c + 1]) - (su[sp][p][r][c - 1] - su[sm][p][r][c - 1])) *
                              iv.imix[1];
                const T myz = ((su[sp][p][r + 1][c] - su[sm][p][r + 1][c]) - (su[sp][p][r - 1][c] - su[sm][p][r - 1][c])) *
                              iv.imix[2];
                gg[p][0][1] = gg[p][1][0] = mxy;
                gg[p][0][2] = gg[p][2][0] = mxz;
                gg[p][1][2] = gg[p][2][1] = myz;
            }
            T f[3];
            const long long st[3] = {1, px, pl};
            if (K == K_ISO) {
                const T lam = Cc[0], mu = Cc[1];
#pragma unroll
                for (int aa = 0; aa < 3; ++aa) {
                    T lap = 0, gd = 0;
#pragma unroll
                    for (int bb = 0; bb < 3; ++bb) {
                        lap += gg[aa][bb][bb];
                        gd += gg[bb][aa][bb];
                    }
                    f[aa] = mu * lap + (mu + lam) * gd;
                }
                if (!uniform) {
                    const T tr = g[0][0] + g[1][1] + g[2][2];
                    T dl[3], dm[3];
#pragma unroll
                    for (int d = 0; d < 2; ++d) {
                        dl[d] = (a.C[o + st[d]] - a.C[o - st[d]]) * iv.ih2[d];
                        dm[d] = (a.C[cs + o + st[d]] - a.C[cs + o - st[d]]) * iv.ih2[d];
                    }
                    dl[2] = (Cp[0] - Cm[0]) * iv.ih2[2];
                    dm[2] = (Cp[1] - Cm[1]) * iv.ih2[2];
#pragma unroll
                    for (int aa = 0; aa < 3; ++aa) {
                        T add = dl[aa] * tr;
#pragma unroll
                        for (int d = 0; d < 3; ++d) add += dm[d] * (g[aa][d] + g[d][aa]);
                        f[aa] += add;
                    }
                }
            } else if (K == K_MM) {
#pragma unroll
                for (int aa = 0; aa < 3; ++aa) {
                    T acc = 0;
#pragma unroll
                    for (int bb = 0; bb < 3; ++bb)
#pragma unroll
                        for (int kk = 0; kk < 3; ++kk)
#pragma unroll
                            for (int l = 0; l < 3; ++l) acc += Cc[mm_slot<3>(aa, bb, kk, l)] * gg[kk][l][bb];
                    f[aa] = acc;
                }
                if (!uniform) {
                    T gs[3][3];
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk)
#pragma unroll
                        for (int l = 0; l < 3; ++l) gs[kk][l] = (kk == l) ? g[kk][kk] : g[kk][l] + g[l][kk];
#pragma unroll
                    for (int aa = 0; aa < 3; ++aa) {
                        T add = 0;
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            T acc = 0;
#pragma unroll
                            for (int kk = 0; kk < 3; ++kk)
#pragma unroll
                                for (int l = kk; l < 3; ++l) {
                                    const int sl = mm_slot<3>(aa, d, kk, l);
                                    T dc;
                                    if (d == 2) dc = Cp[sl] - Cm[sl];
                                    else dc = a.C[sl * cs + o + st[d]] - a.C[sl * cs + o - st[d]];
                                    acc += dc * gs[kk][l];
                                }
                            add += acc * iv.ih2[d];
                        }
                        f[aa] += add;
                    }
                }
            } else {
#pragma unroll
                for (int aa = 0; aa < 3; ++aa) {
                    T acc = 0;
#pragma unroll
                    for (int bb = 0; bb < 3; ++bb) acc += Cc[aa * 3 + bb] * gg[aa][bb][bb];
                    f[aa] = acc;
                }
                if (!uniform) {
#pragma unroll
                    for (int aa = 0; aa < 3; ++aa) {
                        T add = 0;
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            const int sl = aa * 3 + d;
                            T dc;
                            if (d == 2) dc = Cp[sl] - Cm[sl];
                            else dc = a.C[sl * cs + o + st[d]] - a.C[sl * cs + o - st[d]];
                            add += dc * iv.ih2[d] * g[aa][d];
                        }
                        f[aa] += add;
                    }
                }
            }
            epilogue<MODE>(a, cs, o, xc, f, m, s);
#pragma unroll
            for (int q = 0; q < NC; ++q) Cm[q] = Cc[q];
        }
        __syncthreads();  // every read of the ring slot (k-1) % 3 is done
        if (more) store_plane((k + 2) % 3, pre);
        __syncthreads();
    }
    if (MODE == 2 && a.partials) block_max_store2(a.partials, blockIdx.x, m);
}

// Boundary nodes: the six faces, one thread each.  Count: 2 nx ny (z faces)
// + 2 nx (nz-2) (y faces) + 2 (ny-2)(nz-2) (x faces).
template <typename T, int K, int MODE, typename TO>
__global__ __launch_bounds__(256) void k_boundary(Lay L, OpArgs<T, TO> a, BCTab bc, Inv<T> iv, int uniform,
                                                  long long nbnd, long long slot0) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (t < nbnd) {
        const long long nx = L.nx, ny = L.ny, nz = L.nz;
        const long long fz = nx * ny, fy = nx * (nz - 2);
        int i, j, k;
        long long q = t;
        if (q < 2 * fz) {
            k = q < fz ? 0 : (int)nz - 1;
            q %= fz;
            i = (int)(q % nx);
            j = (int)(q / nx);
        } else if ((q -= 2 * fz) < 2 * fy) {
            j = q < fy ? 0 : (int)ny - 1;
            q %= fy;
            i = (int)(q % nx);
            k = 1 + (int)(q / nx);
        } else {
            q -= 2 * fy;
            const long long fx = (ny - 2) * (nz - 2);
            i = q < fx ? 0 : (int)nx - 1;
            q %= fx;
            j = 1 + (int)(q % (ny - 2));
            k = 1 + (int)(q / (ny - 2));
        }
        T f[3];
        node_apply<T, K, 3>(L, a.C, a.u, i, j, k, bc, iv, uniform != 0, f);
        const long long o = (long long)i + L.px * j + L.pl * k;
        T xc[3];
#pragma unroll
        for (int p = 0; p < 3; ++p) xc[p] = a.u[p * L.cs + o];
        const double s = (MODE == 2 && a.outs) ? *a.scale : 1.0;
        epilogue<MODE>(a, L.cs, o, xc, f, m, s);
    }
    if (MODE == 2 && a.partials) block_max_store2(a.partials, slot0 + blockIdx.x, m);
}

template <typename T, int K, int MODE, typename TO>
void launch_impl(const MarchPlan& P, const Lay& L, const OpArgs<T, TO>& a, const BCTab& bc, const double h[3],
                 int uniform, hipStream_t s) {
    const Inv<T> iv = make_inv<T>(h);
    const int nb = P.ntx * P.nty * P.ntz;
    if (P.ty == 4)
        hipLaunchKernelGGL((k_march<T, K, 4, MODE, TO>), dim3(nb), dim3(TX, 4), 0, s, L, a, iv, uniform, P.kz, P.ntx,
                           P.nty, P.ntz);
    else
        hipLaunchKernelGGL((k_march<T, K, 2, MODE, TO>), dim3(nb), dim3(TX, 2), 0, s, L, a, iv, uniform, P.kz, P.ntx,
                           P.nty, P.ntz);
    const long long nbnd = march_boundary_nodes(L);
    hipLaunchKernelGGL((k_boundary<T, K, MODE, TO>), dim3((unsigned)((nbnd + 255) / 256)), dim3(256), 0, s, L, a, bc,
                       iv, uniform, nbnd, (long long)nb);
}

template <typename T, int MODE, typename TO>
void dispatch(const MarchPlan& P, const Lay& L, int kind, const OpArgs<T, TO>& a, const BCTab& bc,
              const double h[3], int uniform, hipStream_t s) {
    if (kind == K_ISO) launch_impl<T, K_ISO, MODE, TO>(P, L, a, bc, h, uniform, s);
    else if (kind == K_MM) launch_impl<T, K_MM, MODE, TO>(P, L, a, bc, h, uniform, s);
    else launch_impl<T, K_DIAG, MODE, TO>(P, L, a, bc, h, uniform, s);
}

__global__ __launch_bounds__(1024) void k_reduce_partials2(const double* __restrict__ partials, int n,
                                                           unsigned long long* norm) {
    __shared__ double sm[1024];
    double m = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const double v = partials[q];
        m = (v > m || v != v) ? v : m;
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            const double v = sm[threadIdx.x + st];
            sm[threadIdx.x] = (v > sm[threadIdx.x] || v != v) ? v : sm[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *norm = (unsigned long long)__double_as_longlong(sm[0]);
}

}  // namespace

long long march_boundary_nodes(const Lay& L) {
    const long long nx = L.nx, ny = L.ny, nz = L.nz;
    return 2 * nx * ny + 2 * nx * (nz - 2) + 2 * (ny - 2) * (nz - 2);
}

MarchPlan march_plan(const Lay& L, int kz, int ty) {
    MarchPlan P;
    P.ty = ty == 2 ? 2 : 4;
    P.kz = kz > 0 ? kz : 32;
    P.ntx = (L.nx - 2 + TX - 1) / TX;
    P.nty = (L.ny - 2 + P.ty - 1) / P.ty;
    P.ntz = (L.nz - 2 + P.kz - 1) / P.kz;
    P.ok = L.dim == 3 && L.nx >= 18 && L.ny >= 6 && L.nz >= 4;
    return P;
}

long long march_partials(const MarchPlan& P, const Lay& L) {
    return (long long)P.ntx * P.nty * P.ntz + (march_boundary_nodes(L) + 255) / 256;
}

template <typename T>
void march_fapply(const MarchPlan& P, const Lay& L, int kind, const T* C, const T* u, T* f, const BCTab& bc,
                  const double h[3], int uniform, hipStream_t s) {
    OpArgs<T, T> a{C, u, f, nullptr, nullptr, nullptr, nullptr, nullptr, T(0)};
    dispatch<T, 0, T>(P, L, kind, a, bc, h, uniform, s);
}

template <typename T>
void march_jacobi(const MarchPlan& P, const Lay& L, int kind, const T* C, const T* dg, const T* xin, const T* b,
                  T* xout, const BCTab& bc, const double h[3], int uniform, double omega, hipStream_t s) {
    OpArgs<T, T> a{C, xin, xout, nullptr, b, dg, nullptr, nullptr, (T)omega};
    dispatch<T, 1, T>(P, L, kind, a, bc, h, uniform, s);
}

template <typename T, typename TO>
void march_residual(const MarchPlan& P, const Lay& L, int kind, const T* C, const T* x, const T* b, T* r, TO* rs,
                    const double* scale, double* partials, unsigned long long* norm, const BCTab& bc,
                    const double h[3], int uniform, hipStream_t s) {
    OpArgs<T, TO> a{C, x, r, rs, b, nullptr, scale, norm ? partials : nullptr, T(0)};
    dispatch<T, 2, TO>(P, L, kind, a, bc, h, uniform, s);
    if (norm)
        hipLaunchKernelGGL(k_reduce_partials2, dim3(1), dim3(1024), 0, s, partials, (int)march_partials(P, L), norm);
}

template void march_fapply<double>(const MarchPlan&, const Lay&, int, const double*, const double*, double*,
                                   const BCTab&, const double*, int, hipStream_t);
template void march_fapply<float>(const MarchPlan&, const Lay&, int, const float*, const float*, float*, const BCTab&,
                                  const double*, int, hipStream_t);
template void march_jacobi<double>(const MarchPlan&, const Lay&, int, const double*, const double*, const double*,
                                   const double*, double*, const BCTab&, const double*, int, double, hipStream_t);
template void march_jacobi<float>(const MarchPlan&, const Lay&, int, const float*, const float*, const float*,
                                  const float*, float*, const BCTab&, const double*, int, double, hipStream_t);
template void march_residual<double, double>(const MarchPlan&, const Lay&, int, const double*, const double*,
                                             const double*, double*, double*, const double*, double*,
                                             unsigned long long*, const BCTab&, const double*, int, hipStream_t);
template void march_residual<double, float>(const MarchPlan&, const Lay&, int, const double*, const double*,
                                            const double*, double*, float*, const double*, double*,
                                            unsigned long long*, const BCTab&, const double*, int, hipStream_t);
template void march_residual<float, float>(const MarchPlan&, const Lay&, int, const float*, const float*,
                                           const float*, float*, float*, const double*, double*,
                                           unsigned long long*, const BCTab&, const double*, int, hipStream_t);

}  // namespace ela
