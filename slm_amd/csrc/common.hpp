// common.hpp -- error reporting and launch helpers shared by the HIP sources
// of libslm_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/slm_hip.h"

// Records a printf-style message for slm_last_error() and returns code.
int slm_fail(int code, const char* fmt, ...);
// Validates the grid and ABI fields of an SlmParams (SLM_OK or an error).
int slm_check_params(const SlmParams* p);
// Checks the last launch, synchronises the stream when SLM_FLAG_SYNC is set.
int slm_finish(void* stream, uint32_t flags);

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return slm_fail(SLM_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
