// fvm_api_util.hpp -- error reporting and configuration checks shared by the
// single-problem (fvm_api.cpp) and batch (fvm_batch_api.cpp) entry points.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/rpfvm.h"

namespace fvm {
namespace api {

// Records the message returned by rpfvm_last_error (per host thread); returns code.
int fail(int code, const char* fmt, ...);

// Checks one step configuration (0 or a negative RPFVM_ERR_* code with the message).
int validate(const rpfvm_config* c);

constexpr int kFrontPad = 12;  // makes column GC of every row 128-byte aligned

}  // namespace api
}  // namespace fvm

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::fvm::api::fail(RPFVM_ERR_HIP, "%s failed: %s", #expr,                   \
                                    hipGetErrorString(e_));                                  \
    } while (0)
