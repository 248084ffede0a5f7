// reduce.hip - deterministic FP64 reductions on the device: a fixed grid of partial sums,
// each a fixed-order shared-memory tree, then one block over the partials.  Used for the
// volume-weighted conservation sums of output() (basic.cu:9-13) and the NaN check the
// drivers do at save intervals (LockExchange.cu:259-261).
#include "hip_compat.hpp"

#include <mutex>

#include "engine.hpp"

namespace ipl {

namespace {

constexpr int kRedBlock = 256;
constexpr int kRedGrid = 1024;

// Scratch of the allocation-free sums (ipl_global_sums_noalloc): the partial sums of the fixed
// reduction grid for up to three fields, then the three results.  Every GPU the library runs on
// holds one copy until the process ends; calls that overlap on one GPU share it.
constexpr size_t kNoAllocScratchDoubles = 3 * kRedGrid + 3;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

template <typename T>
__global__ void k_partial_wsum(const T* a0, const T* a1, const T* a2, const T* wx, const T* wy, const T* wz,
                               int64_t n, double* partial) {
    __shared__ double sh[3][kRedBlock];
    double acc[3] = {0, 0, 0};
    for (int64_t i = blockIdx.x * (int64_t)kRedBlock + threadIdx.x; i < n; i += (int64_t)kRedGrid * kRedBlock) {
        const double vol = (double)wx[i] * (double)wy[i] * (double)wz[i];
        if (a0) acc[0] += (double)a0[i] * vol;
        if (a1) acc[1] += (double)a1[i] * vol;
        if (a2) acc[2] += (double)a2[i] * vol;
    }
    for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = kRedBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) partial[q * kRedGrid + blockIdx.x] = sh[q][0];
}

__global__ void k_final_sum(const double* partial, int nq, double scale, double* out) {
    __shared__ double sh[kRedBlock];
    for (int q = 0; q < nq; ++q) {
        double acc = 0;
        for (int i = threadIdx.x; i < kRedGrid; i += kRedBlock) acc += partial[q * kRedGrid + i];
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int w = kRedBlock / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = sh[0] * scale;
        __syncthreads();
    }
}

template <typename T>
__global__ void k_nonfinite(const T* a0, const T* a1, const T* a2, const T* a3, int64_t n, int* flag) {
    int bad = 0;
    for (int64_t i = blockIdx.x * (int64_t)kRedBlock + threadIdx.x; i < n; i += (int64_t)kRedGrid * kRedBlock) {
        if (a0 && !isfinite((double)a0[i])) bad = 1;
        if (a1 && !isfinite((double)a1[i])) bad = 1;
        if (a2 && !isfinite((double)a2[i])) bad = 1;
        if (a3 && !isfinite((double)a3[i])) bad = 1;
    }
    if (bad) atomicOr(flag, 1);
}

__global__ void k_zero_int(int* p) { *p = 0; }

}  // namespace

Status noalloc_scratch(double** out, size_t* capacity) {
    // the device address of this GPU's copy of `scratch`, looked up once per device with that
    // device current (the array's name is a host-side stand-in, not a device address)
    static std::mutex m;
    static double* ptr[64] = {};
    int dev = 0;
    IPL_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return Status::err(IPL_ERR_HIP, "device index beyond 63");
    std::lock_guard<std::mutex> lock(m);
    if (!ptr[dev]) {
        void* p = nullptr;
        IPL_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[dev] = static_cast<double*>(p);
    }
    *out = ptr[dev];
    *capacity = kNoAllocScratchDoubles;
    return Status::ok();
}

size_t weighted_sums_scratch_doubles() { return 3 * (size_t)kRedGrid + 3; }

template <typename T>
Status device_weighted_sums(const T* const* fields, int nfields, const T* wx, const T* wy, const T* wz, int64_t n,
                            double inv_vol, double* out, hipStream_t st, double* scratch_buf) {
    double* buf = scratch_buf;  // the caller's scratch (every slot is written before it is read), or our own
    if (!buf) IPL_HIP_TRY(hipMallocAsync((void**)&buf, (3 * kRedGrid + 3) * sizeof(double), st));
    IPL_LAUNCH_SYNC((k_partial_wsum<T>), kRedGrid, kRedBlock, 0, st, nfields > 0 ? fields[0] : nullptr, nfields > 1 ? fields[1] : nullptr,
                                                        nfields > 2 ? fields[2] : nullptr, wx, wy, wz, n, buf);
    IPL_LAUNCH_SYNC(k_final_sum, 1, kRedBlock, 0, st, buf, 3, inv_vol, buf + 3 * kRedGrid);
    double h[3];
    IPL_HIP_TRY(hipMemcpyAsync(h, buf + 3 * kRedGrid, sizeof(h), hipMemcpyDeviceToHost, st));
    if (!scratch_buf) IPL_HIP_TRY(hipFreeAsync(buf, st));
    IPL_HIP_TRY(hipStreamSynchronize(st));
    for (int q = 0; q < nfields && q < 3; ++q) out[q] = h[q];
    return Status::ok();
}

template <typename T>
Status device_all_finite(const T* const* fields, int nfields, int64_t n, bool* finite, hipStream_t st) {
    int* flag = nullptr;
    IPL_HIP_TRY(hipMallocAsync((void**)&flag, sizeof(int), st));
    IPL_LAUNCH(k_zero_int, 1, 1, 0, st, flag);
    IPL_LAUNCH((k_nonfinite<T>), kRedGrid, kRedBlock, 0, st, nfields > 0 ? fields[0] : nullptr, nfields > 1 ? fields[1] : nullptr,
                                                     nfields > 2 ? fields[2] : nullptr, nfields > 3 ? fields[3] : nullptr,
                                                     n, flag);
    int h = 0;
    IPL_HIP_TRY(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    IPL_HIP_TRY(hipFreeAsync(flag, st));
    IPL_HIP_TRY(hipStreamSynchronize(st));
    *finite = h == 0;
    return Status::ok();
}

template Status device_weighted_sums<double>(const double* const*, int, const double*, const double*, const double*,
                                             int64_t, double, double*, hipStream_t, double*);
template Status device_weighted_sums<float>(const float* const*, int, const float*, const float*, const float*,
                                            int64_t, double, double*, hipStream_t, double*);
template Status device_all_finite<double>(const double* const*, int, int64_t, bool*, hipStream_t);
template Status device_all_finite<float>(const float* const*, int, int64_t, bool*, hipStream_t);

}  // namespace ipl
