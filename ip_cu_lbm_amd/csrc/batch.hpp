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
    // streams, and waits for them.  The members' own streams must be idle.
    Status run(const std::vector<BatchMember*>& members, const std::vector<int>& counts);
    int graph_captures() const { return captures_; }
    int64_t launches() const { return launches_; }        // k_fused_step launches (graph replays included)
    int64_t member_steps() const { return member_steps_; }  // member-steps those launches advanced
    void set_graph(bool on) { use_graph_ = on; }

    // One run of steps of one launch group: njobs jobs (table entries at `offset` bytes), whose
    // lines are `blocks` workgroups of one launch.
    struct Slice {
        BatchKey key;  // key.lds: the largest LDS request of the slice's jobs
        int njobs = 0;
        int64_t blocks = 0;
        size_t offset = 0;
        int nsteps = 0;
    };

    static constexpr int kGraphSteps = 20;    // steps per captured graph (even: roles return)
    static constexpr int kMaxStreams = 8;     // launch groups run concurrently on these
    static constexpr size_t kMaxGraphs = 64;  // captured slice shapes kept between runs

private:
    struct Cached {
        Slice slice;  // (nsteps unused)
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
    };
    Status init(size_t nstreams);
    Status launch_slice(const Slice& s, int first_step, int nsteps, hipStream_t st);
    Status graph_for(const Slice& s, hipStream_t st, hipGraphExec_t* exec);
    static bool same_slice(const Slice& a, const Slice& b);
    void drop_graphs();

    hipStream_t stream_ = nullptr;  // table uploads
    hipEvent_t uploaded_ = nullptr;
    std::vector<hipStream_t> streams_;
    DeviceArena mem_;
    char* jobs_dev_ = nullptr;
    std::vector<char> host_;  // the job table as uploaded (alive until the run has ended)
    size_t cap_ = 0;
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
