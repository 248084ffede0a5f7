// hip_emu.hpp - CPU emulation of the subset of HIP the library uses (test infrastructure).
//
// Force-included (-include) when the library sources are compiled with g++ and
// -DIPL_CPU_EMULATION.  Kernels become host functions; IPL_LAUNCH runs every work-item of
// every workgroup in order, IPL_LAUNCH_SYNC runs a workgroup's work-items as fibers that
// yield at __syncthreads.  Device memory is host memory, so an out-of-bounds
// index is caught by AddressSanitizer on the CPU instead of faulting a GPU.
#pragma once
#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <math.h>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __builtin_amdgcn_readfirstlane(x) (x)  // the values it is given are uniform over a wavefront
#define __builtin_assume(x) ((void)0)
#define __builtin_amdgcn_sched_barrier(x) ((void)0)  // an instruction-scheduling fence only
#define __constant__
#define __shared__ static
#define __launch_bounds__(...)
#define __align__(n) alignas(n)

using std::isfinite;
using std::sqrt;
using std::min;
using std::max;

struct ipl_emu_idx3 { unsigned x = 0, y = 0, z = 0; };
struct int2 { int x, y; };
struct alignas(16) uint4 { unsigned x, y, z, w; };
inline int2 make_int2(int a, int b) { return int2{a, b}; }
extern thread_local ipl_emu_idx3 threadIdx, blockIdx, blockDim, gridDim;
extern thread_local std::barrier<>* ipl_emu_barrier;
extern thread_local unsigned char* ipl_emu_lds;

namespace ipl_emu {
bool in_fiber();
void fiber_barrier();
}  // namespace ipl_emu
// a barrier of the workgroup: the work-item's fiber yields to the scheduler (a no-op outside a
// synchronising launch)
inline void __syncthreads() {
    if (ipl_emu::in_fiber()) ipl_emu::fiber_barrier();
}

// Wavefront vote of a 64-work-item workgroup (the library's kernels that vote are one wavefront
// each and vote from uniform control flow): every work-item posts its bit in one barrier phase and
// reads the word in the next; the word is cleared in the phase before, after every work-item has
// read the previous vote.
namespace ipl_emu { extern thread_local unsigned long long ballot_word; }
inline unsigned long long __ballot(int p) {
    __syncthreads();
    if (threadIdx.x == 0) ipl_emu::ballot_word = 0;
    __syncthreads();
    if (p) ipl_emu::ballot_word |= 1ull << threadIdx.x;
    __syncthreads();
    return ipl_emu::ballot_word;
}
inline unsigned long long __lanemask_lt() { return (1ull << (threadIdx.x & 63)) - 1; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }

template <typename T> inline T atomicAdd(T* p, T v) {
    std::atomic_ref<T> r(*p);
    T old = r.load();
    while (!r.compare_exchange_weak(old, old + v)) {}
    return old;
}
inline int atomicOr(int* p, int v) { return std::atomic_ref<int>(*p).fetch_or(v); }

// IEEE binary16 with round-to-nearest-even
struct __half { uint16_t bits; };
inline float __half2float(__half h) {
    const uint32_t s = (h.bits & 0x8000u) << 16, e = (h.bits >> 10) & 0x1f, m = h.bits & 0x3ff;
    uint32_t f;
    if (e == 0) {
        if (m == 0) f = s;
        else {  // subnormal
            int ee = -1; uint32_t mm = m;
            do { ++ee; mm <<= 1; } while (!(mm & 0x400));
            f = s | ((uint32_t)(127 - 15 - ee) << 23) | ((mm & 0x3ff) << 13);
        }
    } else if (e == 31) f = s | 0x7f800000u | (m << 13);
    else f = s | ((e - 15 + 127) << 23) | (m << 13);
    float out; std::memcpy(&out, &f, 4); return out;
}
inline __half __float2half_rn(float x) {
    uint32_t f; std::memcpy(&f, &x, 4);
    const uint32_t s = (f >> 16) & 0x8000u;
    const int e = (int)((f >> 23) & 0xff) - 127 + 15;
    uint32_t m = f & 0x7fffffu;
    __half h;
    if (((f >> 23) & 0xff) == 0xff) { h.bits = (uint16_t)(s | 0x7c00u | (m ? 0x200u : 0)); return h; }
    if (e >= 31) { h.bits = (uint16_t)(s | 0x7c00u); return h; }
    if (e <= 0) {
        if (e < -10) { h.bits = (uint16_t)s; return h; }
        m |= 0x800000u;
        const int shift = 14 - e;
        uint32_t hm = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (hm & 1))) ++hm;
        h.bits = (uint16_t)(s | hm); return h;
    }
    uint32_t hm = m >> 13;
    const uint32_t rem = m & 0x1fffu;
    uint32_t he = (uint32_t)e;
    if (rem > 0x1000u || (rem == 0x1000u && (hm & 1))) { if (++hm == 0x400u) { hm = 0; ++he; } }
    if (he >= 31) { h.bits = (uint16_t)(s | 0x7c00u); return h; }
    h.bits = (uint16_t)(s | (he << 10) | hm); return h;
}

// ---- device allocations: every access through the kernels' element helper is checked against
// the allocation that holds its row base (an access that lands in another live allocation is
// not seen by AddressSanitizer, whose red zones only border each block)
namespace ipl_emu {
void register_alloc(const void* p, size_t n);
void unregister_alloc(const void* p);
void check_access(const void* base, const void* addr, size_t n);
}  // namespace ipl_emu

// ---- runtime API subset ------------------------------------------------------------------
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 };
typedef void* hipStream_t;
typedef void* hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
#define hipStreamNonBlocking 1
inline const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "success" : "emulated error"; }
namespace ipl_emu { void count_alloc(); }
inline hipError_t hipMalloc(void** p, size_t n) {
    ipl_emu::count_alloc();
    *p = std::malloc(n);
    if (!*p) return hipErrorOutOfMemory;
    ipl_emu::register_alloc(*p, n);
    return hipSuccess;
}
template <typename T> inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
inline hipError_t hipFree(void* p) {
    ipl_emu::unregister_alloc(p);
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipMallocAsync(void** p, size_t n, hipStream_t) { return hipMalloc(p, n); }
inline hipError_t hipFreeAsync(void* p, hipStream_t) { return hipFree(p); }
typedef void* hipMemPool_t;
enum hipMemPoolAttr { hipMemPoolAttrReleaseThreshold = 4 };
inline hipError_t hipDeviceGetDefaultMemPool(hipMemPool_t* m, int) { *m = nullptr; return hipSuccess; }
inline hipError_t hipMemPoolSetAttribute(hipMemPool_t, hipMemPoolAttr, void*) { return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = (void*)1; return hipSuccess; }
#define hipEventDisableTiming 2
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = (void*)1; return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }  // work runs in call order
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0; return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
// a file-scope __device__ array is a host array here: its address is the symbol's
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return hipSuccess; }
extern "C" long long ipl_emu_alloc_count();
// graphs are not emulated: capture reports an error, so the engines launch every step directly
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1, hipStreamCaptureModeRelaxed = 2 };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorInvalidValue; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = nullptr; return hipErrorInvalidValue; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t, void*, void*, size_t) { *e = nullptr; return hipErrorInvalidValue; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorInvalidValue; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }

namespace ipl_emu {
void launch(unsigned grid, unsigned block, const std::function<void()>& body);
void launch_sync(unsigned grid, unsigned block, size_t lds, const std::function<void()>& body);
}  // namespace ipl_emu

#define IPL_LAUNCH(kernel, grid, block, shmem, stream, ...) \
    ipl_emu::launch((unsigned)(grid), (unsigned)(block), [&]() { kernel(__VA_ARGS__); })
#define IPL_LAUNCH_SYNC(kernel, grid, block, shmem, stream, ...) \
    ipl_emu::launch_sync((unsigned)(grid), (unsigned)(block), (size_t)(shmem), [&]() { kernel(__VA_ARGS__); })
#define IPL_DYNAMIC_LDS(type, name) type* name = reinterpret_cast<type*>(ipl_emu_lds)
