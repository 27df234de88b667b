// Error reporting shared by the translation units of the C ABI library.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../../include/heliocubed.h"

namespace hc
{
// Message of the calling thread's last failed C-ABI call.
std::string& lastError();

inline int fail(const std::string& msg)
{
    lastError() = msg;
    return -1;
}

// Fourth-order level (defined with the fourth-order kernels): stage
// preparation JU -> USph (ghost 5) and the box-table RHS over the level.
struct Level4;
Level4* level4Create(const hc_shell& shell, const hc_options& opt, int boxSize);
void level4Destroy(Level4* L);
void level4Prepare(Level4& L, const double* U0, const double* K, double a_dt, const double* inner, double* US,
                   hipStream_t s);
void level4Rhs(Level4& L, const double* US, double* K, double* U, double b_dt, hipStream_t s);
} // namespace hc

#define HC_THROW(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
