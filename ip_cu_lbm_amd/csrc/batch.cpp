// batch.cpp - batched step launches of many fused-path contexts (ipl_batch_step).
//
// Contexts whose step is the fused kernel alone are grouped by launch key (storage, lattices,
// stencil, stream axis, workgroup size); jobs of one group advance together, one launch of
// k_fused_step per step covering the lines of all of them.  Members step by different counts: a
// group runs as slices of decreasing membership (the members with steps left), in order of
// increasing count.  Every slice's job table is built on the host and uploaded in one copy per
// run; the groups then run concurrently, each on a stream of its own, and the run waits for all.
// A slice of at least kGraphSteps steps replays a captured graph of kGraphSteps launches; a graph
// holds its slice's launch geometry and table offset and reads the table when it runs, so it
// serves every later slice of the same shape (the kMaxGraphs most recently used are kept).
#include "hip_compat.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "batch.hpp"

namespace ipl {

bool BatchRunner::same_slice(const Slice& a, const Slice& b) {
    return a.key == b.key && a.key.lds == b.key.lds && a.njobs == b.njobs && a.blocks == b.blocks && a.offset == b.offset;
}

BatchRunner::~BatchRunner() {
    for (hipStream_t s : streams_) (void)hipStreamSynchronize(s);
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graphs();
    if (stream_) {
        mem_.release();
        (void)hipStreamSynchronize(stream_);
    }
    for (hipStream_t s : streams_) (void)hipStreamDestroy(s);
    if (uploaded_) (void)hipEventDestroy(uploaded_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

Status BatchRunner::init(size_t nstreams) {
    if (!stream_) {
        IPL_HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        IPL_HIP_TRY(hipEventCreateWithFlags(&uploaded_, hipEventDisableTiming));
        mem_.bind(stream_);
    }
    while (streams_.size() < std::min(nstreams, (size_t)kMaxStreams)) {
        hipStream_t s = nullptr;
        IPL_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        streams_.push_back(s);
    }
    return Status::ok();
}

Status BatchRunner::launch_slice(const Slice& s, int first_step, int nsteps, hipStream_t st) {
    for (int k = 0; k < nsteps; ++k) {
        Status r = fused_batch_launch(s.key, jobs_dev_ + s.offset, s.njobs, s.blocks, first_step + k, st);
        if (!r) return r;
    }
    return Status::ok();
}

Status BatchRunner::graph_for(const Slice& s, hipStream_t st, hipGraphExec_t* exec) {
    for (Cached& c : graphs_)
        if (same_slice(c.slice, s)) {
            c.used = ++clock_;
            *exec = c.exec;
            return Status::ok();
        }
    IPL_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    Status r = launch_slice(s, 0, kGraphSteps, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (!r) { if (g) (void)hipGraphDestroy(g); return r; }
    if (e != hipSuccess) return Status::err(IPL_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) return Status::err(IPL_ERR_HIP, std::string("batch graph instantiate: ") + hipGetErrorString(e2));
    ++captures_;
    graphs_.push_back(Cached{s, x, ++clock_});
    *exec = x;
    return Status::ok();
}

Status BatchRunner::run(const std::vector<BatchMember*>& members, const std::vector<int>& counts) {
    if (members.empty()) return Status::ok();
    // launch groups in key order, members in the caller's order within a group (jobs whose lines
    // round to the same workgroup size share a launch whatever their lengths: the kernel masks
    // work-items past its job's line; the launch takes the largest LDS request of its jobs)
    std::map<BatchKey, std::vector<size_t>> groups;
    for (size_t i = 0; i < members.size(); ++i) {
        BatchKey k = members[i]->batch_key();
        k.lds = 0;
        groups[k].push_back(i);
    }
    // the slices of every group and their job tables, in one host buffer
    std::vector<std::vector<Slice>> plan;
    std::vector<char> host;
    for (auto& [key, idx] : groups) {
        std::vector<int> levels;
        for (size_t i : idx) levels.push_back(counts[i]);
        std::sort(levels.begin(), levels.end());
        levels.erase(std::unique(levels.begin(), levels.end()), levels.end());
        std::vector<Slice> slices;
        int done = 0;
        for (int lvl : levels) {
            Slice s;
            s.key = key;
            s.offset = host.size();
            s.nsteps = lvl - done;
            int64_t first = 0;
            for (size_t i : idx) {
                if (counts[i] < lvl) continue;
                BatchMember* m = members[i];
                if (first + m->batch_lines() > INT32_MAX) return Status::err(IPL_ERR_UNSUPPORTED, "a batched launch of more than 2^31 lines");
                const size_t jb = m->batch_job_bytes();
                host.resize(host.size() + jb);
                // the roles at the slice's start: swapped after an odd number of earlier steps
                m->batch_fill(host.data() + host.size() - jb, (int)first, (done & 1) != 0);
                first += m->batch_lines();
                s.key.lds = std::max(s.key.lds, m->batch_key().lds);
                ++s.njobs;
            }
            s.blocks = first;
            slices.push_back(s);
            done = lvl;
        }
        plan.push_back(std::move(slices));
    }
    Status r = init(plan.size());
    if (!r) return r;
    // the previous run has ended (it waits for its streams), so the table, its host copy and the
    // graphs are idle: trim the graphs to the most recently used
    if (graphs_.size() > kMaxGraphs) {
        std::sort(graphs_.begin(), graphs_.end(), [](const Cached& a, const Cached& b) { return a.used > b.used; });
        for (size_t i = kMaxGraphs; i < graphs_.size(); ++i) (void)hipGraphExecDestroy(graphs_[i].exec);
        graphs_.resize(kMaxGraphs);
    }
    host_.swap(host);
    if (host_.size() > cap_) {  // a larger table: the captured graphs held the old one
        drop_graphs();
        if (!(r = mem_.alloc_n(&jobs_dev_, host_.size() * 2))) return r;
        cap_ = host_.size() * 2;
    }
    IPL_HIP_TRY(hipMemcpyAsync(jobs_dev_, host_.data(), host_.size(), hipMemcpyHostToDevice, stream_));
    IPL_HIP_TRY(hipEventRecord(uploaded_, stream_));
    for (size_t g = 0; g < plan.size(); ++g) {
        hipStream_t st = streams_[g % streams_.size()];
        IPL_HIP_TRY(hipStreamWaitEvent(st, uploaded_, 0));
        for (const Slice& s : plan[g]) {
            int left = s.nsteps;
            if (use_graph_ && left >= kGraphSteps) {
                hipGraphExec_t exec = nullptr;
                if (!(r = graph_for(s, st, &exec))) return r;
                for (; left >= kGraphSteps; left -= kGraphSteps) IPL_HIP_TRY(hipGraphLaunch(exec, st));
            }
            // the rest, launched directly (the step parity continues from the replays)
            if (!(r = launch_slice(s, s.nsteps - left, left, st))) return r;
            launches_ += s.nsteps;
            member_steps_ += (int64_t)s.nsteps * s.njobs;
        }
    }
    for (hipStream_t st : streams_) IPL_HIP_TRY(hipStreamSynchronize(st));
    for (size_t i = 0; i < members.size(); ++i) members[i]->batch_advance(counts[i]);
    return Status::ok();
}

void BatchRunner::drop_graphs() {
    for (Cached& c : graphs_) (void)hipGraphExecDestroy(c.exec);
    graphs_.clear();
}

}  // namespace ipl
