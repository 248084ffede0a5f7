// batch.hpp - batched step launches of many fused-path contexts (see batch.cpp).
#pragma once

#include "hip_compat.hpp"

#include <vector>

#include "engine.hpp"

namespace ipl {

class BatchRunner {
public:
    ~BatchRunner();
    // Advances members[i] by counts[i] steps (counts >= 1) with batched launches on the runner's
    // stream, and waits for them.  The members' own streams must be idle.
    Status run(const std::vector<BatchMember*>& members, const std::vector<int>& counts);
    int graph_captures() const { return captures_; }
    int64_t launches() const { return launches_; }        // k_fused_step launches (graph replays included)
    int64_t member_steps() const { return member_steps_; }  // member-steps those launches advanced
    void set_graph(bool on) { use_graph_ = on; }

private:
    struct Group {
        BatchKey key;
        int njobs = 0;
        int64_t blocks = 0;
        size_t offset = 0;  // bytes into the job table
    };
    struct Cached {
        std::vector<Group> layout;
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
    };
    static constexpr size_t kMaxGraphs = 8;  // captured layouts kept (least recently used dropped)
    Status init();
    Status phase(const std::vector<BatchMember*>& members, int nsteps);
    Status launch_layout(int step);
    Status graph_for_layout(hipGraphExec_t* exec);
    static bool same_layout(const std::vector<Group>& a, const std::vector<Group>& b);
    void drop_graph();

    hipStream_t stream_ = nullptr;
    DeviceArena mem_;
    char* jobs_dev_ = nullptr;
    std::vector<char> host_;  // the job table as uploaded (alive until its copy has run)
    size_t cap_ = 0;
    std::vector<Group> layout_;
    std::vector<Cached> graphs_;
    uint64_t clock_ = 0;
    int captures_ = 0;
    int64_t launches_ = 0, member_steps_ = 0;
#ifdef IPL_CPU_EMULATION
    bool use_graph_ = false;
#else
    bool use_graph_ = true;
#endif
};

}  // namespace ipl
