// ela_internal.hpp -- host-side pieces shared by the single-box solver
// (ela_solver.hip) and the batched engine (ela_batch.hip): the device
// context with its HBM budget, and the thread-local error message.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <string>
#include <unordered_map>

#include "../../include/ela.h"

struct ela_ctx {
    int device = 0;
    size_t budget = 0;
    size_t in_use = 0;
    std::mutex mu;
    std::unordered_map<void*, size_t> raw;  // ela_dev_alloc'd buffers and their sizes
};

namespace ela {

// Record msg as the thread's last error and return code.
int set_error(int code, const std::string& msg);

// Reserve / release bytes against the context's budget (budget 0 = unlimited).
int budget_reserve(ela_ctx* ctx, size_t bytes);
void budget_release(ela_ctx* ctx, size_t bytes);

}  // namespace ela

#define ELA_HIPCHK(expr)                                                                             \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return ::ela::set_error(ELA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
