// hip_compat.hpp - the one place that includes the HIP runtime.  Kernel launches and the
// dynamic LDS declaration go through two macros so that the same sources can also be built
// against the CPU emulation layer of the test suite (tests/emu/), where an out-of-bounds
// access shows up under AddressSanitizer instead of as a GPU fault.
#pragma once

#ifndef IPL_CPU_EMULATION
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
// launch of a kernel that uses no __syncthreads
#define IPL_LAUNCH(kernel, grid, block, shmem, stream, ...) \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, stream, __VA_ARGS__)
// launch of a kernel whose work-items synchronise inside a workgroup
#define IPL_LAUNCH_SYNC(kernel, grid, block, shmem, stream, ...) \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, stream, __VA_ARGS__)
#define IPL_DYNAMIC_LDS(type, name)                  \
    extern __shared__ unsigned char ipl_lds_raw_[]; \
    type* name = reinterpret_cast<type*>(ipl_lds_raw_)
#endif
