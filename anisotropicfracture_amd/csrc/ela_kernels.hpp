// ela_kernels.hpp -- launchers of the gfx950 kernels (ela_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "ela_device.hpp"
#include "ela_grain.hpp"

namespace ela {

template <typename T>
Inv<T> make_inv(const double h[3]);

// Options of the per-node-plane stencil launchers: sm, the same-direction mask
// (launch_same_mask) or null; list, the level's nodes in the order to launch
// them (launch_node_list; nlist entries, every node once) or null.
struct PlaneOpt {
    const unsigned char* sm = nullptr;
    const unsigned long long* list = nullptr;
    long long nlist = 0;
    bool group = false;  // 3D, with sm: constant-C nodes take each block's first wavefronts
};

template <typename T>
void launch_fapply(const Lay& L, int kind, const T* C, const T* u, T* f, const BCTab& bc, const double h[3],
                   int uniform, hipStream_t s, const PlaneOpt& po = PlaneOpt{});

// r = b - A x; r (T), rs = r * (*scale) (TO), max|r| into *norm (bit pattern of the double,
// via residual_partials(L) doubles of scratch in partials); any output may be null.
template <typename T, typename TO>
void launch_residual(const Lay& L, int kind, const T* C, const T* x, const T* b, T* r, TO* rs, const double* scale,
                     double* partials, unsigned long long* norm, const BCTab& bc, const double h[3], int uniform,
                     hipStream_t s, const PlaneOpt& po = PlaneOpt{});
long long residual_partials(const Lay& L);

template <typename T>
void launch_jacobi(const Lay& L, int kind, const T* C, const T* dg, const T* xin, const T* b, T* xout,
                   const BCTab& bc, const double h[3], int uniform, double omega, hipStream_t s,
                   const PlaneOpt& po = PlaneOpt{});

// sm (one byte per node): bit d set where the stiffness records at the two
// d-neighbours of an interior node are bitwise equal (the plane launchers
// above then skip that d-derivative of C, whose terms are exactly zero).
// stats[0] += wavefronts with an interior node, stats[1] += (wavefront,
// direction) pairs whose interior nodes all have the bit (the work skipped).
void launch_same_mask(const Lay& L, int nc, const double* C, unsigned char* sm, unsigned long long* stats,
                      hipStream_t s);
// Node list of a plane level from its mask (PlaneOpt::list): the interior
// nodes with every direction bit first, then the others; list holds one
// entry per node, totals = {first part, second part} (device), cnt
// 4 * node_list_blocks(L) ints of scratch.
long long node_list_blocks(const Lay& L);
void launch_node_list(const Lay& L, const unsigned char* sm, int* cnt, long long* totals, unsigned long long* list,
                      hipStream_t s);

// Barrier probe (ela_probe_barrier): `rounds` device-wide barriers among
// `blocks` blocks of 256 threads, bounded spin; *err set when they time out.
void launch_barrier_probe(int blocks, unsigned* cnt, int* err, int rounds, hipStream_t s);

template <typename T>
void launch_jacobi0(const Lay& L, const T* dg, const T* b, T* xout, double omega, hipStream_t s,
                    const unsigned short* gid = nullptr, const T* dgt = nullptr);

void launch_diag(const Lay& L, int kind, const double* C, double* d64, float* d32, const BCTab& bc,
                 const double h[3], unsigned long long* flag, hipStream_t s);

template <typename T>
void launch_restrict(const Lay& Lc, const Lay& Lf, int nc, const T* fine, T* crse, hipStream_t s);
// crse(I, J, K) = fine(2I, 2J, 2K) for nc components.
template <typename T>
void launch_inject_level(const Lay& Lc, const Lay& Lf, int nc, const T* fine, T* crse, hipStream_t s);

template <typename T>
void launch_prolong(const Lay& Lf, const Lay& Lc, int nc, const T* crse, T* fine, int add, hipStream_t s);

// y = a*x/(*ascale) (assign) or y += a*x/(*ascale); ascale may be null (then y (+)= a*x).
template <typename TX, typename TY>
void launch_axpy(const Lay& L, int nc, const TX* x, TY* y, const double* ascale, double a, int assign,
                 hipStream_t s);

template <typename TX, typename TY>
void launch_convert(long long n, const TX* x, TY* y, hipStream_t s);

// x0 /= diag0, ncomp times (component 0 only, as the reference's normalize does).
template <typename T>
void launch_normalize(const Lay& L, int ncomp, const double* dg, T* x, hipStream_t s);

void launch_zero_u64(unsigned long long* p, int n, hipStream_t s);
void launch_scale_from_norm(const unsigned long long* norm, double* scale, hipStream_t s);

void launch_assemble_col(const Lay& L, int nodes, const double* F, int q, long long ld, double* M, int dim,
                         hipStream_t s);
void launch_unit(const Lay& L, int nodes, double* E, int qprev, int q, hipStream_t s);
// All columns of the coarsest operator into M (rows p * nodes + t, leading
// dimension ld) in two launches; E: nodes * dim zeroed fields of fstride
// doubles (scratch).
void launch_probe_all(const Lay& L, int kind, int nodes, const double* C, double* E, long long fstride,
                      const BCTab& bc, const double h[3], int uniform, double* M, long long ld, hipStream_t s);
void launch_identity_right(double* M, int n, hipStream_t s);
void launch_gauss_jordan(double* M, int n, int* flag, hipStream_t s);
// Gauss-Jordan of [A | I] in one launch a column (k_gj_col): the same pivots
// and bits as launch_gauss_jordan's two launches a column.  Scratch: perm 2n
// ints, dpiv n doubles, cand 2n doubles.  Both leave row r of A^-1 in the
// left half of row r.
void launch_gauss_jordan_col(double* M, int n, int* flag, int* perm, double* dpiv, double* cand, hipStream_t s);

template <typename T>
void launch_bottom(const Lay& L, int nodes, int dim, const double* M, const T* b, T* x, hipStream_t s);

void launch_cubic_table(int ng, const double* R, double C11, double C12, double C44, double* table, hipStream_t s);
void launch_voronoi_cubic(const Lay& L, int nseeds, const long long* seeds, int q, const double* table, double* C,
                          unsigned short* gid, hipStream_t s);

// Grain-table stiffness (3D major/minor): per-node uint16 grain id, 21 values per grain, and the
// per-grain interior diagonal (3 values) for the Jacobi sweep.
template <typename T>
struct TabDesc {
    const unsigned short* gid;
    const T* tab;
    const T* dgt;
};
template <typename T>
void launch_fapply_tab(const Lay& L, const TabDesc<T>& t, const T* u, T* f, const BCTab& bc, const double h[3],
                       int uniform, hipStream_t s);
template <typename T>
void launch_jacobi_tab(const Lay& L, const TabDesc<T>& t, const T* dg, const T* xin, const T* b, T* xout,
                       const BCTab& bc, const double h[3], int uniform, double omega, hipStream_t s);
template <typename T, typename TO>
void launch_residual_tab(const Lay& L, const TabDesc<T>& t, const T* x, const T* b, T* r, TO* rs,
                         const double* scale, double* partials, unsigned long long* norm, const BCTab& bc,
                         const double h[3], int uniform, hipStream_t s);
void launch_grain_diag(int ng, const double* tab, const double h[3], double* d64, float* d32, hipStream_t s);

// Newton<T>::prepareForSolve: dw = DDW : (grad u - F0) into dw (dim*dim planes; F0 the same, or null),
// then rhs = b - div dw inside, b - BC row(u, grad u, dw) on the boundary.  FP64.
void launch_newton_rhs(const Lay& L, int kind, const double* C, const double* F0, const double* u, const double* b,
                       double* dw, double* rhs, const BCTab& bc, const double h[3], hipStream_t s);

// Strain / Stress / Energy of u (what 0 / 1 / 2) into out (6 or 9 / 6 or 9 / 1 planes in 3D).
void launch_post(const Lay& L, int kind, const double* C, const double* u, double* out, const double h[3], int what,
                 int voigt, hipStream_t s);

// Grain-form z-marching stencil on one box (modes of GrainBox, ela_grain.hpp);
// MODE 2 writes nm + nsh block maxima into partials (grain_partials(L, kz)
// doubles) and folds them into *norm when norm is non-null.
template <typename T, int MODE, typename TO>
void launch_grain(GrainBox<T> g, int kz, double* partials, unsigned long long* norm, hipStream_t s);
long long grain_partials(const Lay& L, int kz);
// The same z-marching on a box whose stiffness is per-node isotropic planes
// (g.C: lambda, mu; optional same-direction mask g.sm): modes 0-3 of
// GrainBox, tiles of one row per thread and g.pf planes loaded ahead.
template <typename T, int MODE, typename TO>
void launch_plane(GrainBox<T> g, int kz, double* partials, unsigned long long* norm, hipStream_t s);
// Order-parameter fields for the model assembly: node values (dense over the
// level's nodes), or cell values with one ghost layer ((n+2) cells per
// direction, x fastest) averaged onto the nodes over every cell around a node
// (ETA_CELLS) or over the 4 cells of its x-y plane (ETA_XY4).  cs: elements
// between components.
enum { ETA_NODES = 0, ETA_CELLS = 1, ETA_XY4 = 2 };
struct EtaSrc {
    const double* p;
    int avg;
    int gx, gy;  // cells per row / rows per plane, ghosts included (cell sources)
    long long cs;
};
void launch_eta_check(const Lay& L, const EtaSrc& e, int c0, int c1, int* flags, hipStream_t s);
void launch_degrade_iso(const Lay& L, const EtaSrc& e, int ca, int cb, double lambda0, double mu0, double* C,
                        hipStream_t s);
void launch_combine(const Lay& L, const EtaSrc& e, int nph, const double* ddw, const double* F0s, double* C,
                    double* F0out, int* flags, hipStream_t s);
// Newton::W / DW (Newton.H:233-333): dw = DDW : (grad u - F0) (dim*dim planes),
// W = trace((wfac (grad u - F0))^T dw); either output may be null.
void launch_newton_w_dw(const Lay& L, int kind, const double* C, const double* F0, const double* u, double wfac,
                        double* w, double* dw, const double h[3], hipStream_t s);

// BiCGStab bottom solve (flat over whole fields, zero padding included):
// out[0] = (a, b), out[1] = (c, d), out[2] = max |e| in FP64 (null inputs skip a
// term; partials: bicg_partials() doubles), and the step updates
// (op 0: p = r + beta (p - omega v); 1: s = r - alpha v;
//  2: x += alpha p + omega s, r = s - omega t; 3: x += alpha p).
template <typename T>
void launch_bicg_reduce(long long n, const T* a, const T* b, const T* c, const T* d, const T* e, double* partials,
                        double* out, hipStream_t s);
template <typename T>
void launch_bicg_update(long long n, int op, double alpha, double beta, double omega, T* x, T* r, T* p, const T* v,
                        T* sv, const T* t, const T* dg, T* ph, T* sh, const Lay& L, hipStream_t s);
long long bicg_partials();

// Flat FP64 vector kernels of the Krylov accelerator (whole fields incl. zero padding).
constexpr int kMaxKrylov = 33;
// FGMRES re-orthogonalizes w when |w'|^2 <= kReorth2 |w|^2 after a classical
// Gram-Schmidt pass (|w'| < 0.1 |w|).  Measured on the 256^3 polycrystal with
// FGMRES(6): the bound 1/2 re-orthogonalized nearly every step at the same
// iteration count and final residual, 10 % slower.
constexpr double kReorth2 = 0.01;
template <typename TB>
struct VecListT {
    const TB* v[kMaxKrylov];
};
using VecList = VecListT<double>;
using VecList32 = VecListT<float>;
// Per-vector multipliers of an update's coefficients (coef[v] * f[v]): an FP32
// Krylov basis may hold vector v scaled by s_v, whose coefficient is then
// taken times 1 / s_v^2.  A multiplier of 1 leaves the coefficient's bits.
struct CoefMul {
    double f[kMaxKrylov];
};
inline CoefMul coef_ones() {
    CoefMul c;
    for (double& x : c.f) x = 1.0;
    return c;
}
// out[v] = (x, Y_v) for v < nv (deterministic two-pass reduction; partials: mdot_partials(nv) doubles).
void launch_mdot(const double* x, const VecList& Y, int nv, long long n, double* partials, double* out,
                 hipStream_t s);
// out[v] = (x, Y_v) for v < nv with an FP32 list (FP64 sums), and out[nv] = (x, x)
// (partials: mdot_partials(nv + 1) doubles).
void launch_mdot_self(const double* x, const VecList32& Y, int nv, long long n, double* partials, double* out,
                      hipStream_t s);
long long mdot_partials(int nv);
// y += sign * sum_v coef[v] X_v, coef on the device.
void launch_maxpy(double* y, const VecList& X, int nv, const double* coef, double sign, long long n, hipStream_t s);
// launch_maxpy, and |y|^2 of the result into *norm2 (device), summed in the
// order launch_mdot sums (y, y) (partials: mdot_partials(1) doubles).
void launch_maxpy_norm(double* y, const VecList& X, int nv, const double* coef, double sign, long long n,
                       double* partials, double* norm2, hipStream_t s);
// The same with an FP32 list (FP64 arithmetic), the coefficients times mul.
void launch_maxpy32(double* y, const VecList32& X, int nv, const double* coef, const CoefMul& mul, double sign,
                    long long n, hipStream_t s);
void launch_maxpy_norm32(double* y, const VecList32& X, int nv, const double* coef, const CoefMul& mul,
                         double sign, long long n, double* partials, double* norm2, hipStream_t s);
// launch_maxpy_norm32 whose FP64 result r is not stored back into y: o1 and o2
// (o2 may be null) receive (float) (osc r), osc a power of two (exact scaling),
// and *norm2 = |r|^2 (the next FP32 Krylov vector and the MG cycle's
// right-hand side in one pass, the normalization left to the caller).
void launch_maxpy_norm32_to2f(const double* y, const VecList32& X, int nv, const double* coef, const CoefMul& mul,
                              double sign, long long n, double osc, float* o1, float* o2, double* partials,
                              double* norm2, hipStream_t s);
// y = a x
void launch_scale_flat(double* y, const double* x, double a, long long n, hipStream_t s);
// y = a x and yo = (TO) y, flat over n elements (yo: a field of the same layout)
template <typename TO>
void launch_scale_flat_to(double* y, const double* x, double a, TO* yo, long long n, hipStream_t s);
// v = (float) (a x), r = v, flat over n elements
void launch_scale_to2f(const double* x, double a, float* v, float* r, long long n, hipStream_t s);
// out = (double) (T) (cor + hold) on the valid nodes of a level (nc components)
template <typename T>
void launch_add_to64(const Lay& L, int nc, const T* hold, const T* cor, double* out, hipStream_t s);

// Streaming copy of bytes (multiple of 16) from src to dst.
void launch_copy(const void* src, void* dst, long long bytes, hipStream_t s);
// Variant var in [0, copy_variants()) of the copy (0: launch_copy; others more bytes in flight).
int copy_variants();
void launch_copy_variant(int var, const void* src, void* dst, long long bytes, hipStream_t s);

// Dense host-order array (nc planes of nx*ny*nz, already on the device) <-> level layout.
template <typename T>
void launch_pack(const Lay& L, int nc, const double* dense, T* dev, int to_dev, hipStream_t s);

}  // namespace ela
