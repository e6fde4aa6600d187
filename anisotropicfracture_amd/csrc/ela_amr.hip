// ela_amr.hip -- transfers between AMR levels of the nodal hierarchy: the
// reflux of the fine residual into the coarse one (Operator<Node>::reflux,
// src/Operator/Operator.cpp:728-839) and the average-down of the fine
// solution (averageDownSolutionRHS, src/Operator/Operator.cpp:626-641).
//
// An AMR level's refined region is one box of nodes with its own layout (the
// fine box); its position is the fine-level node index of its first node
// (even: fine boxes are refinements of coarse cell boxes).
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/ela.h"
#include "ela_internal.hpp"
#include "ela_kernels.hpp"

namespace ela {
namespace {

constexpr int RX = 64, RY = 4;

// Coarse node (I,J,K) = (ox/2 + a, oy/2 + b, oz/2 + c) of the coarsened fine
// box takes the full weighting of the fine residual; fine values outside the
// box are its ghost nodes, zero.
// The coarse field may be a window of the coarse level: its element of
// coarse node (I,J,K) is at (I - cw.x) + px (J - cw.y) + pl (K - cw.z), while
// Lc's node counts (the coarse domain's) decide the boundary branches.
struct Win {
    int x, y, z;
};
template <typename T, int DIM>
__global__ __launch_bounds__(RX* RY) void k_reflux(Lay Lc, Lay Lf, int ox, int oy, int oz, int nc,
                                                   const T* __restrict__ fine, T* __restrict__ crse, Win cw) {
    const int cnx = (Lf.nx - 1) / 2 + 1, cny = (Lf.ny - 1) / 2 + 1, cnz = DIM == 3 ? (Lf.nz - 1) / 2 + 1 : 1;
    const int a = blockIdx.x * RX + threadIdx.x, b = blockIdx.y * RY + threadIdx.y, c = blockIdx.z;
    if (a >= cnx || b >= cny || c >= cnz) return;
    const int I = ox / 2 + a, J = oy / 2 + b, K = oz / 2 + c;
    const long long px = Lf.px, pl = Lf.pl;
    const int fx = 2 * a, fy = 2 * b, fz = 2 * c, nx = Lf.nx, ny = Lf.ny, nz = Lf.nz;
    for (int n = 0; n < nc; ++n) {
        const T* F = fine + n * Lf.cs;
        auto get = [=](int di, int dj, int dk) -> T {
            const int x = fx + di, y = fy + dj, z = fz + dk;
            if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return T(0);
            return F[x + px * y + pl * z];
        };
        crse[n * Lc.cs + (I - cw.x) + Lc.px * (J - cw.y) + Lc.pl * (K - cw.z)] = restrict_node<T, DIM>(Lc, I, J, K, get);
    }
}

// Nodal average-down of the solution: coarse nodes of the coarsened fine box
// take the coincident fine value (amrex::average_down on nodal data).
template <typename T, int DIM>
__global__ __launch_bounds__(RX* RY) void k_average_down(Lay Lc, Lay Lf, int ox, int oy, int oz, int nc,
                                                         const T* __restrict__ fine, T* __restrict__ crse, Win cw) {
    const int cnx = (Lf.nx - 1) / 2 + 1, cny = (Lf.ny - 1) / 2 + 1, cnz = DIM == 3 ? (Lf.nz - 1) / 2 + 1 : 1;
    const int a = blockIdx.x * RX + threadIdx.x, b = blockIdx.y * RY + threadIdx.y, c = blockIdx.z;
    if (a >= cnx || b >= cny || c >= cnz) return;
    const int I = ox / 2 + a, J = oy / 2 + b, K = oz / 2 + c;
    for (int n = 0; n < nc; ++n)
        crse[n * Lc.cs + (I - cw.x) + Lc.px * (J - cw.y) + Lc.pl * (K - cw.z)] =
            fine[n * Lf.cs + 2 * a + Lf.px * (2 * b) + Lf.pl * (2 * c)];
}

Lay lay_of(const ela_level_layout& l) {
    Lay L;
    L.dim = l.dim;
    L.nx = l.nodes[0];
    L.ny = l.nodes[1];
    L.nz = l.dim == 3 ? l.nodes[2] : 1;
    L.px = l.pitch;
    L.pl = l.plane;
    L.cs = l.cstride;
    return L;
}

int check_pair(const ela_level_layout* c, const ela_level_layout* f, const int* o, int ncomp, int prec) {
    if (!c || !f || !o) return set_error(ELA_ERR_BADARG, "null argument");
    if (c->dim != f->dim || (c->dim != 2 && c->dim != 3)) return set_error(ELA_ERR_BADARG, "dimension mismatch");
    if (ncomp < 1 || (prec != ELA_F64 && prec != ELA_F32)) return set_error(ELA_ERR_BADARG, "bad ncomp or precision");
    for (int d = 0; d < c->dim; ++d) {
        if (o[d] < 0 || o[d] % 2) return set_error(ELA_ERR_BADARG, "fine box origin must be even and >= 0");
        if (f->nodes[d] < 1 || f->nodes[d] % 2 == 0)
            return set_error(ELA_ERR_BADARG, "fine box must span an even number of cells");
        if (o[d] + f->nodes[d] - 1 > 2 * (c->nodes[d] - 1))
            return set_error(ELA_ERR_BADARG, "fine box extends past the refined coarse domain");
        if (c->pitch < c->nodes[0] || f->pitch < f->nodes[0]) return set_error(ELA_ERR_BADARG, "bad pitch");
    }
    return ELA_OK;
}

// A window of the coarse level (win: its layout, wo: the coarse node index of
// its first node) must hold every coarse node under the fine box.
int check_window(const ela_level_layout* c, const ela_level_layout* win, const int* wo, const ela_level_layout* f,
                 const int* o) {
    if (!win || !wo) return set_error(ELA_ERR_BADARG, "null window");
    if (win->dim != c->dim) return set_error(ELA_ERR_BADARG, "dimension mismatch");
    for (int d = 0; d < c->dim; ++d) {
        const int lo = o[d] / 2, hi = (o[d] + f->nodes[d] - 1) / 2;
        if (wo[d] < 0 || lo < wo[d] || hi > wo[d] + win->nodes[d] - 1)
            return set_error(ELA_ERR_BADARG, "the coarse window does not cover the coarsened fine box");
    }
    if (win->pitch < win->nodes[0] || (win->dim == 3 && win->plane < win->pitch * win->nodes[1]))
        return set_error(ELA_ERR_BADARG, "bad window pitch");
    return ELA_OK;
}

template <template <typename, int> class Kern>
int launch_pair(const ela_level_layout* crse, void* crse_f, const ela_level_layout* fine, const void* fine_f,
                const int* o, int ncomp, int prec, void* stream, const ela_level_layout* win = nullptr,
                const int* wo = nullptr) {
    if (int rc = check_pair(crse, fine, o, ncomp, prec)) return rc;
    if (win)
        if (int rc = check_window(crse, win, wo, fine, o)) return rc;
    if (!crse_f || !fine_f) return set_error(ELA_ERR_BADARG, "null field");
    Lay Lc = lay_of(*crse);
    const Lay Lf = lay_of(*fine);
    Win cw{0, 0, 0};
    if (win) {  // node counts of the domain, strides of the window
        Lc.px = win->pitch;
        Lc.pl = win->plane;
        Lc.cs = win->cstride;
        cw = Win{wo[0], wo[1], crse->dim == 3 ? wo[2] : 0};
    }
    const int d3 = crse->dim == 3;
    const int cnx = (Lf.nx - 1) / 2 + 1, cny = (Lf.ny - 1) / 2 + 1, cnz = d3 ? (Lf.nz - 1) / 2 + 1 : 1;
    const dim3 grid((cnx + RX - 1) / RX, (cny + RY - 1) / RY, cnz), block(RX, RY);
    hipStream_t s = (hipStream_t)stream;
    const int oz = d3 ? o[2] : 0;
    if (prec == ELA_F64) {
        if (d3) hipLaunchKernelGGL((Kern<double, 3>::k), grid, block, 0, s, Lc, Lf, o[0], o[1], oz, ncomp,
                                   (const double*)fine_f, (double*)crse_f, cw);
        else hipLaunchKernelGGL((Kern<double, 2>::k), grid, block, 0, s, Lc, Lf, o[0], o[1], oz, ncomp,
                                (const double*)fine_f, (double*)crse_f, cw);
    } else {
        if (d3) hipLaunchKernelGGL((Kern<float, 3>::k), grid, block, 0, s, Lc, Lf, o[0], o[1], oz, ncomp,
                                   (const float*)fine_f, (float*)crse_f, cw);
        else hipLaunchKernelGGL((Kern<float, 2>::k), grid, block, 0, s, Lc, Lf, o[0], o[1], oz, ncomp,
                                (const float*)fine_f, (float*)crse_f, cw);
    }
    ELA_HIPCHK(hipGetLastError());
    if (!stream) ELA_HIPCHK(hipStreamSynchronize(nullptr));
    return ELA_OK;
}

template <typename T, int DIM>
struct Reflux {
    static constexpr auto k = k_reflux<T, DIM>;
};
template <typename T, int DIM>
struct AverageDown {
    static constexpr auto k = k_average_down<T, DIM>;
};

}  // namespace
}  // namespace ela

using namespace ela;

extern "C" {

int ela_reflux(const ela_level_layout* crse, void* crse_res, const ela_level_layout* fine, const void* fine_res,
               const int fine_origin[3], int ncomp, int prec, void* stream) {
    return launch_pair<Reflux>(crse, crse_res, fine, fine_res, fine_origin, ncomp, prec, stream);
}

int ela_reflux_window(const ela_level_layout* crse, const ela_level_layout* crse_win, const int crse_origin[3],
                      void* crse_res, const ela_level_layout* fine, const void* fine_res, const int fine_origin[3],
                      int ncomp, int prec, void* stream) {
    return launch_pair<Reflux>(crse, crse_res, fine, fine_res, fine_origin, ncomp, prec, stream, crse_win, crse_origin);
}

int ela_average_down_window(const ela_level_layout* crse, const ela_level_layout* crse_win, const int crse_origin[3],
                            void* crse_sol, const ela_level_layout* fine, const void* fine_sol,
                            const int fine_origin[3], int ncomp, int prec, void* stream) {
    return launch_pair<AverageDown>(crse, crse_sol, fine, fine_sol, fine_origin, ncomp, prec, stream, crse_win,
                                    crse_origin);
}

int ela_average_down(const ela_level_layout* crse, void* crse_sol, const ela_level_layout* fine,
                     const void* fine_sol, const int fine_origin[3], int ncomp, int prec, void* stream) {
    return launch_pair<AverageDown>(crse, crse_sol, fine, fine_sol, fine_origin, ncomp, prec, stream);
}

}  // extern "C"
