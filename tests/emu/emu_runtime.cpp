// emu_runtime.cpp - work-item scheduling and the device-allocation registry of the CPU
// emulation (test infrastructure).
//
// A workgroup whose kernel synchronises runs its work-items as fibers on one host thread: each
// has its own stack, __syncthreads switches back to the scheduler, and one sweep over the
// work-items resumes every one that waits, so each sweep is one barrier phase.  AddressSanitizer
// is told about every stack switch.
#include "hip_emu.hpp"

#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <ucontext.h>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define IPL_EMU_ASAN 1
#else
#define IPL_EMU_ASAN 0
#endif

thread_local ipl_emu_idx3 threadIdx, blockIdx, blockDim, gridDim;
thread_local std::barrier<>* ipl_emu_barrier = nullptr;
thread_local unsigned char* ipl_emu_lds = nullptr;
namespace ipl_emu { thread_local unsigned long long ballot_word = 0; }

namespace ipl_emu {

namespace {
std::map<uintptr_t, size_t>& allocs() {
    static std::map<uintptr_t, size_t> m;
    return m;
}
std::shared_mutex& allocs_mutex() {
    static std::shared_mutex m;
    return m;
}

// ---- fibers ------------------------------------------------------------------------------
constexpr size_t kStack = 128 * 1024;  // AddressSanitizer clears the shadow of a whole stack per switch

struct Fiber {
    ucontext_t ctx;
    std::unique_ptr<char[]> stack;
    bool done = false;
};

struct Sched {
    ucontext_t main;
    std::vector<Fiber> fibers;
    const std::function<void()>* body = nullptr;
    unsigned cur = 0;
    bool active = false;
#if IPL_EMU_ASAN
    const void* main_bottom = nullptr;
    size_t main_size = 0;
#endif
};
thread_local Sched g_sched;

// switch from the running fiber back to the scheduler
void to_main() {
    Sched& s = g_sched;
#if IPL_EMU_ASAN
    void* fake = nullptr;
    __sanitizer_start_switch_fiber(&fake, s.main_bottom, s.main_size);
#endif
    swapcontext(&s.fibers[s.cur].ctx, &s.main);
#if IPL_EMU_ASAN
    __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
}

void fiber_entry() {
    Sched& s = g_sched;
#if IPL_EMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &s.main_bottom, &s.main_size);
#endif
    (*s.body)();
    s.fibers[s.cur].done = true;
#if IPL_EMU_ASAN
    __sanitizer_start_switch_fiber(nullptr, s.main_bottom, s.main_size);  // this fiber ends
#endif
    setcontext(&s.main);
}

// resume fiber t until it waits at a barrier or ends
void resume(unsigned t) {
    Sched& s = g_sched;
    s.cur = t;
    threadIdx.x = t;
#if IPL_EMU_ASAN
    void* fake = nullptr;
    __sanitizer_start_switch_fiber(&fake, s.fibers[t].stack.get(), kStack);
#endif
    swapcontext(&s.main, &s.fibers[t].ctx);
#if IPL_EMU_ASAN
    __sanitizer_finish_switch_fiber(fake, &s.main_bottom, &s.main_size);
#endif
}
}  // namespace

bool in_fiber() { return g_sched.active; }

std::atomic<long long> g_allocs{0};
void count_alloc() { ++g_allocs; }
void fiber_barrier() { to_main(); }

void register_alloc(const void* p, size_t n) {
    std::unique_lock<std::shared_mutex> l(allocs_mutex());
    allocs()[(uintptr_t)p] = n;
}

void unregister_alloc(const void* p) {
    if (!p) return;
    std::unique_lock<std::shared_mutex> l(allocs_mutex());
    allocs().erase((uintptr_t)p);
}

// aborts when [addr, addr + n) is not inside the live allocation that holds base
void check_access(const void* base, const void* addr, size_t n) {
    const uintptr_t b = (uintptr_t)base, a = (uintptr_t)addr;
    std::shared_lock<std::shared_mutex> l(allocs_mutex());
    auto& m = allocs();
    auto it = m.upper_bound(b);
    if (it == m.begin()) {
        std::fprintf(stderr, "[emu] access through %p: not a device allocation\n", base);
        std::abort();
    }
    --it;
    const uintptr_t lo = it->first, hi = it->first + it->second;
    if (b >= hi) {
        std::fprintf(stderr, "[emu] access through %p: not a device allocation\n", base);
        std::abort();
    }
    if (a < lo || a + n > hi) {
        std::fprintf(stderr, "[emu] out of bounds: %zu bytes at offset %lld of an allocation of %zu bytes (row base at offset %lld)\n",
                     n, (long long)(a - lo), (size_t)it->second, (long long)(b - lo));
        std::abort();
    }
}

void launch(unsigned grid, unsigned block, const std::function<void()>& body) {
    blockDim.x = block; gridDim.x = grid;
    for (unsigned b = 0; b < grid; ++b)
        for (unsigned t = 0; t < block; ++t) {
            blockIdx.x = b; threadIdx.x = t;
            body();
        }
}

// IPL_EMU_REVERSE=1 runs workgroups in reverse order: results that change expose two
// workgroups writing one location (an order the GPU does not guarantee)
static bool reverse_order() {
    const char* e = std::getenv("IPL_EMU_REVERSE");
    return e && e[0] == '1';
}

void launch_sync(unsigned grid, unsigned block, size_t lds, const std::function<void()>& body) {
    Sched& s = g_sched;
    if (s.active) std::abort();  // no launch from inside a kernel
    blockDim.x = block; gridDim.x = grid;
    // exact size: an LDS overrun is a heap overflow under AddressSanitizer
    std::vector<unsigned char> mem(lds);
    ipl_emu_lds = lds ? mem.data() : nullptr;
    if (s.fibers.size() < block) s.fibers.resize(block);
    for (unsigned t = 0; t < block; ++t)
        if (!s.fibers[t].stack) s.fibers[t].stack.reset(new char[kStack]);
    s.body = &body;
    s.active = true;
    const bool rev = reverse_order();
    for (unsigned bb = 0; bb < grid; ++bb) {
        blockIdx.x = rev ? grid - 1 - bb : bb;
        for (unsigned t = 0; t < block; ++t) {
            Fiber& f = s.fibers[t];
            getcontext(&f.ctx);
            f.ctx.uc_stack.ss_sp = f.stack.get();
            f.ctx.uc_stack.ss_size = kStack;
            f.ctx.uc_link = nullptr;
            makecontext(&f.ctx, fiber_entry, 0);
            f.done = false;
        }
        // one sweep per barrier phase, until every work-item has returned
        for (unsigned left = block; left > 0;) {
            left = 0;
            for (unsigned t = 0; t < block; ++t) {
                if (s.fibers[t].done) continue;
                resume(t);
                left += !s.fibers[t].done;
            }
        }
    }
    s.active = false;
    s.body = nullptr;
    ipl_emu_lds = nullptr;
}

}  // namespace ipl_emu

// number of device allocations made so far (hipMalloc and hipMallocAsync)
extern "C" long long ipl_emu_alloc_count() { return ipl_emu::g_allocs.load(); }
