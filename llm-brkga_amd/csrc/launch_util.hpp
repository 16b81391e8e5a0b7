// launch_util.hpp - small helpers shared by the kernel launchers.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

namespace brkga {

// Dynamic LDS above 64 KiB must be allowed per kernel.
template <typename F>
inline void ensure_lds(F func, int bytes)
{
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(func), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  bytes);
}

// Plain launch, or with start / stop events taken from the kernel's own
// dispatch (hipExtLaunchKernel) when either is given.
template <typename K, typename... Args>
inline void launch_maybe_timed(K kernel, dim3 grid, dim3 block, int lds, hipStream_t s, hipEvent_t start,
                               hipEvent_t stop, Args... args)
{
    if (start || stop)
        hipExtLaunchKernelGGL(kernel, grid, block, (uint32_t)lds, s, start, stop, 0u, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
}

}  // namespace brkga
