// emu_runtime.cpp - work-item scheduling of the CPU emulation (test infrastructure).
#include "hip_emu.hpp"

thread_local ipl_emu_idx3 threadIdx, blockIdx, blockDim, gridDim;
thread_local std::barrier<>* ipl_emu_barrier = nullptr;
thread_local unsigned char* ipl_emu_lds = nullptr;

namespace ipl_emu {

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
    std::barrier<> bar((std::ptrdiff_t)block);
    // exact size: an LDS overrun is a heap overflow under AddressSanitizer
    std::vector<unsigned char> mem(lds);
    std::vector<std::thread> pool;
    pool.reserve(block);
    for (unsigned t = 0; t < block; ++t)
        pool.emplace_back([&, t]() {
            blockDim.x = block; gridDim.x = grid; threadIdx.x = t;
            ipl_emu_barrier = &bar;
            ipl_emu_lds = lds ? mem.data() : nullptr;
            const bool rev = reverse_order();
            for (unsigned bb = 0; bb < grid; ++bb) {
                blockIdx.x = rev ? grid - 1 - bb : bb;
                body();
                bar.arrive_and_wait();  // the next workgroup reuses the LDS buffer
            }
            ipl_emu_barrier = nullptr;
        });
    for (auto& th : pool) th.join();
}

}  // namespace ipl_emu
