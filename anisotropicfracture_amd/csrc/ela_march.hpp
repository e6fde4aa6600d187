// ela_march.hpp -- LDS-staged z-marching stencil kernels (3D levels).
#pragma once
#include <hip/hip_runtime.h>

#include "ela_kernels.hpp"

namespace ela {

struct MarchPlan {
    int ty = 4;   // rows per block (block = 64 x ty threads)
    int kz = 32;  // z planes marched per block
    int ntx = 0, nty = 0, ntz = 0;
    bool ok = false;  // level large enough for the marching kernels
};

MarchPlan march_plan(const Lay& L, int kz, int ty);
long long march_boundary_nodes(const Lay& L);
long long march_partials(const MarchPlan& P, const Lay& L);

template <typename T>
void march_fapply(const MarchPlan& P, const Lay& L, int kind, const T* C, const T* u, T* f, const BCTab& bc,
                  const double h[3], int uniform, hipStream_t s);
template <typename T>
void march_jacobi(const MarchPlan& P, const Lay& L, int kind, const T* C, const T* dg, const T* xin, const T* b,
                  T* xout, const BCTab& bc, const double h[3], int uniform, double omega, hipStream_t s);
template <typename T, typename TO>
void march_residual(const MarchPlan& P, const Lay& L, int kind, const T* C, const T* x, const T* b, T* r, TO* rs,
                    const double* scale, double* partials, unsigned long long* norm, const BCTab& bc,
                    const double h[3], int uniform, hipStream_t s);

}  // namespace ela
