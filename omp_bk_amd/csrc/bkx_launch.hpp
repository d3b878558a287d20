// bkx_launch.hpp -- host-side helpers of the C ABI: status codes, order
// dispatch and launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <utility>

#include "bkx.h"

namespace bkx {

// Run-time p -> compile-time P in [lo, hi]: calls f(std::integral_constant<int, P>)
// (the role of bk::with_order, include/bk_common.h:143-178).
template <int lo, int hi, typename F>
int with_order(int p, F&& f)
{
    if constexpr (lo > hi) {
        (void)p;
        (void)f;
        return BKX_ERROR_ORDER;
    } else {
        if (p == lo) {
            return f(std::integral_constant<int, lo>{});
        }
        return with_order<lo + 1, hi>(p, std::forward<F>(f));
    }
}

inline int launch_status()
{
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? BKX_SUCCESS : BKX_ERROR_LAUNCH;
}

// Grid size for nelmt elements, units elements per workgroup; 0 when the grid
// would exceed the 2^31 - 1 workgroups of one dimension.
inline unsigned grid_for(size_t nelmt, size_t per_block)
{
    const size_t g = (nelmt + per_block - 1) / per_block;
    return g > 0x7fffffffULL ? 0u : unsigned(g);
}

// Workgroups of `Kernel` (block threads, static LDS only) that are resident on
// the current device at once: compute units times the occupancy per unit.
// Computed once per device; 0 if the runtime cannot tell.  Persistent
// kernels are launched with at most this many workgroups, so that every
// workgroup runs from the start and none waits for another to finish.
constexpr int kMaxDevices = 64;
template <auto Kernel>
unsigned resident_blocks(int block)
{
    static std::atomic<int> cache[kMaxDevices];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        return 0;
    }
    int n = cache[dev].load(std::memory_order_relaxed);
    if (n == 0) {
        int per_cu = 0;
        int cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, Kernel, block, 0) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            return 0;
        }
        n = per_cu * cus;
        cache[dev].store(n, std::memory_order_relaxed);
    }
    return unsigned(n);
}

}  // namespace bkx
