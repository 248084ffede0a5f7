// engine.hpp - the per-context engine interface behind the C ABI.
#pragma once

#include "hip_compat.hpp"

#include <cstdint>
#include <memory>
#include <string>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/ipl.h"
#include "graph.hpp"
#include "plain.hpp"

namespace ipl {

// Status with a message; every HIP error is turned into one of these.
struct Status {
    int code = IPL_OK;
    std::string msg;
    static Status ok() { return {}; }
    static Status err(int c, std::string m) { return {c, std::move(m)}; }
    explicit operator bool() const { return code == IPL_OK; }
};

#define IPL_HIP_TRY(expr)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ::ipl::Status::err(e_ == hipErrorOutOfMemory ? IPL_ERR_OUT_OF_MEMORY : IPL_ERR_HIP, \
                                      std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

// Device allocation that frees itself.  Once bound to the owner's stream, memory comes from the
// device's stream-ordered pool (hipMallocAsync) and returns to it with hipFreeAsync on that
// stream: releasing a context waits for nothing but its own stream, where hipFree would wait for
// every stream of the device (a server releases finished jobs while the others step).  The pool
// keeps freed memory for the next allocation instead of unmapping it.
class DeviceArena {
public:
    ~DeviceArena() { release(); }
    void bind(hipStream_t st) { stream_ = st; bound_ = true; }
    Status alloc(void** p, size_t bytes) {
        *p = nullptr;
        if (bytes == 0) bytes = 256;
        if (bound_) {
            keep_pool_memory();
            IPL_HIP_TRY(hipMallocAsync(p, bytes, stream_));
            IPL_HIP_TRY(hipStreamSynchronize(stream_));
        } else {
            IPL_HIP_TRY(hipMalloc(p, bytes));
        }
        ptrs_.push_back({*p, bound_});
        total_ += bytes;
        return Status::ok();
    }
    template <typename T> Status alloc_n(T** p, size_t n) { return alloc((void**)p, n * sizeof(T)); }
    void release() {  // each buffer the way it was allocated
        for (const auto& [p, pooled] : ptrs_) (void)(pooled ? hipFreeAsync(p, stream_) : hipFree(p));
        ptrs_.clear();
        total_ = 0;
    }
    size_t total() const { return total_; }

private:
    // release threshold of the current device's default pool: keep what is freed
    static void keep_pool_memory() {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return;
        static std::once_flag once[64];
        if (dev < 0 || dev >= 64) return;
        std::call_once(once[dev], [dev] {
            hipMemPool_t pool;
            if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
                uint64_t keep = UINT64_MAX;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            }
        });
    }
    std::vector<std::pair<void*, bool>> ptrs_;  // (buffer, from the stream-ordered pool)
    size_t total_ = 0;
    hipStream_t stream_ = nullptr;
    bool bound_ = false;
};

// Host-side initial state expanded to all cells (calc + ghosts) in FP64.
struct HostState {
    std::vector<double> p, u, v, w, rho, phi, S, tau, taus;
    std::vector<double> f, g;  // [k][Ntot]
};
Status make_host_state(const Graph& g, const ipl_config& cfg, const ipl_init* init, HostState& hs);

// The template arguments, block and LDS size of a fused step launch: contexts with equal keys can
// share one batched launch (fusedk::k_fused_step).
struct BatchKey {
    int storage = 0, Q = 0, Qs = 0, npts = 0, s = 0, flux = 0, ext = 0;
    unsigned block = 0;
    size_t lds = 0;
    auto tie() const { return std::make_tuple(storage, Q, Qs, npts, s, flux, ext, block, lds); }
    bool operator==(const BatchKey& o) const { return tie() == o.tie(); }
    bool operator<(const BatchKey& o) const { return tie() < o.tie(); }
};

// Workgroup of the fused step (fused.hip): `lines` line slots of `slot` work-items (whole
// wavefronts) per workgroup of `items` work-items.
struct WorkgroupShape {
    int lines = 1, slot = 64, items = 64;
};
WorkgroupShape workgroup_shape(int ns, int requested, size_t line_lds_bytes);
WorkgroupShape batch_workgroup_shape(int ns, size_t line_lds_bytes);

// A fused-path context whose steps can run inside a batched launch.
class BatchMember {
public:
    virtual ~BatchMember() = default;
    virtual BatchKey batch_key() const = 0;
    virtual int64_t batch_lines() const = 0;     // workgroups of one step
    virtual size_t batch_job_bytes() const = 0;  // size of its fusedk::BatchJob
    // writes its fusedk::BatchJob for launches that start from the current buffer roles, or from
    // the swapped ones (after an odd number of steps not yet applied by batch_advance)
    virtual void batch_fill(void* job, int first_block, bool swapped) const = 0;
    // the buffer roles after nsteps batched steps
    virtual void batch_advance(int nsteps) = 0;
};

// One batched step launch of the jobs (device memory, njobs fusedk::BatchJob of key's types).
Status fused_batch_launch(const BatchKey& key, const void* jobs, int njobs, int64_t blocks, int step, hipStream_t st);

class Engine {
public:
    virtual ~Engine() = default;
    virtual Status step(int nsteps) = 0;
    // step() without the final host check: launches only (a batch launches every member first)
    virtual Status launch_steps(int nsteps) { return step(nsteps); }
    // host-side check after the launched steps completed (the IB box bound)
    virtual Status post_step() { return Status::ok(); }
    // replay steps from a captured HIP graph where the engine supports it
    virtual void set_graph(bool) {}
    // capture the step graphs ahead of the first step (no device work)
    virtual Status prepare() { return Status::ok(); }
    // the context shares the device with others that step at the same time (a batch member):
    // work that a context alone overlaps on a second stream stays on its own stream
    virtual Status set_shared(bool) { return Status::ok(); }
    // number of graph captures so far
    virtual int graph_captures() const { return 0; }
    virtual Status stage(int s) = 0;
    virtual Status read(int field, double* host, size_t count) = 0;
    virtual Status write(int field, const double* host, size_t count) = 0;
    // scratch: device buffer of at least weighted_sums_scratch_doubles() doubles for the partial
    // sums, or null for the engine's own
    virtual Status global_sums(double out[3], double* scratch) = 0;
    virtual Status check_finite(bool* finite) = 0;
    // Runs nsteps steps timing every launch with HIP events on the engine's stream:
    // out[0] = mean duration of the dominant kernel's launch (ms), out[1] = mean per-step time of
    // the other kernels (ms), out[2] = summed time of all timed launches (ms).
    virtual Status profile(int nsteps, double out[3]) = 0;
    virtual int64_t device_bytes() const = 0;
    virtual int path() const = 0;
    // immersed boundary (fused path only)
    virtual Status ib_attach(const ipl_ib_config&) {
        return Status::err(IPL_ERR_UNSUPPORTED, "the immersed boundary runs on the fused path");
    }
    virtual Status ib_body_state(double*) { return Status::err(IPL_ERR_ARGUMENT, "no body attached"); }
    virtual Status ib_read_markers(int, double*, size_t) { return Status::err(IPL_ERR_ARGUMENT, "no body attached"); }
    virtual int64_t ib_marker_count() const { return 0; }
    virtual Status ib_stage(int) { return Status::err(IPL_ERR_UNSUPPORTED, "the immersed boundary runs on the fused path"); }
    // the engine as a member of batched launches, or null when its step has more than the fused
    // kernel (an immersed body, LES, an open boundary, the 3D phase field's scalar pass)
    virtual BatchMember* batch_member() { return nullptr; }
    virtual Status change_dt(double) { return Status::err(IPL_ERR_UNSUPPORTED, "time-step changes run on the fused path"); }
    // vals: (p, u, v, w) per listed cell
    virtual Status set_open_boundary(const int32_t*, int64_t, const double*) {
        return Status::err(IPL_ERR_UNSUPPORTED, "open boundaries run on the fused path");
    }
    hipStream_t stream = nullptr;
};

Params make_params(const ipl_config& cfg, const Graph& g);

std::unique_ptr<Engine> make_plain_engine(const Graph& g, const ipl_config& cfg, const HostState& hs,
                                          hipStream_t st, Status& status);
std::unique_ptr<Engine> make_fused_engine(const Graph& g, const ipl_config& cfg, const ipl_init* init,
                                          hipStream_t st, Status& status);
// Whether the fused path can run this configuration; fills why when it cannot.
bool fused_supported(const Graph& g, const ipl_config& cfg, std::string& why);

// Deterministic FP64 reductions on the device (fixed block partition, fixed order).  scratch:
// weighted_sums_scratch_doubles() doubles for the partial sums, or null to allocate them.
template <typename T>
Status device_weighted_sums(const T* const* fields, int nfields, const T* wx, const T* wy, const T* wz,
                            int64_t n, double inv_vol, double* out, hipStream_t st, double* scratch);
size_t weighted_sums_scratch_doubles();
// The device address and capacity (doubles) of the current GPU's file-scope scratch array.
Status noalloc_scratch(double** out, size_t* capacity);
template <typename T>
Status device_all_finite(const T* const* fields, int nfields, int64_t n, bool* finite, hipStream_t st);

}  // namespace ipl
