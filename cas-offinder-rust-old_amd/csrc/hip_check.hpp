// HIP error check for libcof host code: HIP failures become cof::Error with
// COF_E_HIP (the reference panics on DriverError, search.rs:1616-1619).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "cof_internal.hpp"

#define HIPCHECK(expr)                                                                                  \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::cof::Error(e_ == hipErrorOutOfMemory ? COF_E_NOMEM : COF_E_HIP,                     \
                               std::string(#expr) + ": " + hipGetErrorString(e_));                      \
    } while (0)
