// ela_hooks.hpp -- what the composite (AMR) solver (ela_composite.hip) uses of
// a single-box hierarchy (ela_solver.hip): the level-0 correction fields, one
// multigrid cycle on them, and a shared stream.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/ela.h"
#include "ela_device.hpp"

namespace ela {

// Layout of level 0 and its current right-hand side / correction fields of the
// correction cycle in precision prec (ELA_F64 or ELA_F32).  The cycle swaps
// these buffers: read res right before filling it and cor right after a cycle.
int hier_level0(ela_hier* h, Lay* L);
void* hier_res0(ela_hier* h, int prec);
void* hier_cor0(ela_hier* h, int prec);

// One multigrid cycle (the hierarchy's cycle kind) from a zero correction:
// level-0 res -> level-0 cor, on the hierarchy's stream.  Needs ela_prepare
// and a dense coarsest-level inverse.
int hier_cycle(ela_hier* h, int prec);

// Run the hierarchy's kernels on stream s (owned by the caller) from now on.
int hier_share_stream(ela_hier* h, hipStream_t s);
hipStream_t hier_stream(const ela_hier* h);

// Level-0 stiffness planes (FP64, level-0 layout) for the caller to fill on
// the device, then hier_model_set() to mark the model as set.
double* hier_C0(ela_hier* h);
void hier_model_set(ela_hier* h);

}  // namespace ela
