// Error reporting shared by the translation units of the C ABI library.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace hc
{
// Message of the calling thread's last failed C-ABI call.
std::string& lastError();

inline int fail(const std::string& msg)
{
    lastError() = msg;
    return -1;
}
} // namespace hc

#define HC_THROW(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
