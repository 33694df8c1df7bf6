// common.hpp -- error reporting and launch helpers shared by the HIP sources
// of libslm_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include "../../include/slm_hip.h"

// Records a printf-style message for slm_last_error() and returns code.
int slm_fail(int code, const char* fmt, ...);
// Validates the grid and ABI fields of an SlmParams (SLM_OK or an error).
int slm_check_params(const SlmParams* p);
// Checks the last launch, synchronises the stream when SLM_FLAG_SYNC is set.
int slm_finish(void* stream, uint32_t flags);

// TRAJECTORY_RESOLUTION is a power of two in [1, 2^31]: approximate's division
// is then an exact scaling (approx_r<true>), and kernels specialised on it
// issue no division there.
inline bool slm_res_is_pow2(const SlmParams* p) {
    const float r = p->trajectory_resolution;
    if (!(r >= 1.0f && r <= 2147483648.0f)) return false;
    const uint32_t u = (uint32_t)r;
    return (float)u == r && (u & (u - 1u)) == 0u;
}

// The lanes of a wave that hold one key act through one atomic (neighbouring
// items mostly share a key, so per-lane atomics queue on the same address):
// per distinct key other than `none`, the lowest such lane adds the group's
// size to ctr[key], and the group's lanes take consecutive positions after
// the old value, in lane order.  Returns the lane's position (0 for `none`).
// Every lane of the wave calls it.
__device__ __forceinline__ uint32_t wave_group_add(uint32_t key, uint32_t none, uint32_t* __restrict__ ctr) {
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t todo = __ballot(key != none);
    uint32_t at = 0;
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t lk = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
        const uint64_t same = __ballot(key == lk) & todo;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&ctr[lk], (uint32_t)__popcll(same));
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
        if ((same >> lane) & 1ull) at = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        todo &= ~same;
    }
    return at;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return slm_fail(SLM_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Grid of a persist_seeds kernel (256-thread blocks): every resident block
// (occupancy x CUs), or fewer when there are fewer 64-seed chunks than waves.
template <typename K>
inline dim3 persist_grid(K kernel, uint32_t n_items) {
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (!cus[dev] && hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus[dev] = 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const uint32_t want = (uint32_t)std::max(cus[dev], 1) * (uint32_t)per_cu;
    const uint32_t chunks = (n_items + 63u) / 64u;
    const uint32_t need = std::max(1u, (chunks + 3u) / 4u);
    return dim3(std::min(want, need));
}

