// batch.cpp - batched step launches of many fused-path contexts (ipl_batch_step).
//
// Contexts whose step is the fused kernel alone are grouped by launch key (storage, lattices,
// stencil, stream axis, workgroup size); each group advances with one launch of k_fused_step
// per step, whose workgroups cover the lines of every job of the group.  The job table is uploaded
// once per phase (a run of steps with one set of members), and the launches of two steps (the
// buffer roles swap every step) are captured in a HIP graph that is replayed.  A graph holds the
// launch geometry of its layout (per group: key, job count, workgroups, table offset) and reads
// the table when it runs, so it serves every later phase of the same layout; the last few
// layouts' graphs are kept.  Members step by different counts: the phases run in order of
// increasing count, each with the members that still have steps left.
#include "hip_compat.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "batch.hpp"

namespace ipl {

bool BatchRunner::same_layout(const std::vector<Group>& a, const std::vector<Group>& b) {
    return a.size() == b.size() && std::equal(a.begin(), a.end(), b.begin(), [](const Group& x, const Group& y) {
               return x.key == y.key && x.key.lds == y.key.lds && x.njobs == y.njobs && x.blocks == y.blocks && x.offset == y.offset;
           });
}

BatchRunner::~BatchRunner() {
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        drop_graph();
        mem_.release();
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

Status BatchRunner::init() {
    if (stream_) return Status::ok();
    IPL_HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    mem_.bind(stream_);
    return Status::ok();
}

Status BatchRunner::launch_layout(int step) {
    for (const Group& g : layout_) {
        Status r = fused_batch_launch(g.key, jobs_dev_ + g.offset, g.njobs, g.blocks, step, stream_);
        if (!r) return r;
    }
    return Status::ok();
}

Status BatchRunner::phase(const std::vector<BatchMember*>& members, int nsteps) {
    // groups in key order, members in the caller's order within a group
    // (jobs whose lines round to the same workgroup size share a launch whatever their lengths:
    // the kernel masks work-items past its job's line; the launch takes the largest LDS request)
    std::map<BatchKey, std::vector<BatchMember*>> groups;
    for (BatchMember* m : members) {
        BatchKey k = m->batch_key();
        k.lds = 0;
        groups[k].push_back(m);
    }
    std::vector<Group> layout;
    std::vector<char> host;
    for (auto& [key, ms] : groups) {
        Group g;
        g.key = key;
        for (BatchMember* m : ms) g.key.lds = std::max(g.key.lds, m->batch_key().lds);
        g.njobs = (int)ms.size();
        g.offset = host.size();
        const size_t jb = ms[0]->batch_job_bytes();
        host.resize(host.size() + jb * ms.size());
        int64_t first = 0;
        for (size_t i = 0; i < ms.size(); ++i) {
            if (first + ms[i]->batch_lines() > INT32_MAX) return Status::err(IPL_ERR_UNSUPPORTED, "a batched launch of more than 2^31 lines");
            ms[i]->batch_fill(host.data() + g.offset + i * jb, (int)first);
            first += ms[i]->batch_lines();
        }
        g.blocks = first;
        layout.push_back(g);
    }
    // the previous phase's launches have read the job table (and its host copy) before both change
    IPL_HIP_TRY(hipStreamSynchronize(stream_));
    host_.swap(host);
    if (host_.size() > cap_) {  // a larger job table: the captured launches held the old one
        Status r = mem_.alloc_n(&jobs_dev_, host_.size() * 2);
        if (!r) return r;
        cap_ = host_.size() * 2;
        drop_graph();
    }
    IPL_HIP_TRY(hipMemcpyAsync(jobs_dev_, host_.data(), host_.size(), hipMemcpyHostToDevice, stream_));
    layout_ = layout;
    int left = nsteps;
    if (use_graph_ && left >= 2) {
        hipGraphExec_t exec = nullptr;
        Status r = graph_for_layout(&exec);
        if (!r) return r;
        for (; left >= 2; left -= 2) IPL_HIP_TRY(hipGraphLaunch(exec, stream_));
    }
    for (int s = 0; s < left; ++s) {  // an odd remainder, or graphs off (step parity continues)
        Status r = launch_layout((nsteps - left + s) & 1);
        if (!r) return r;
    }
    for (BatchMember* m : members) m->batch_advance(nsteps);
    launches_ += (int64_t)nsteps * (int64_t)layout_.size();
    member_steps_ += (int64_t)nsteps * (int64_t)members.size();
    return Status::ok();
}

Status BatchRunner::run(const std::vector<BatchMember*>& members, const std::vector<int>& counts) {
    if (members.empty()) return Status::ok();
    Status r = init();
    if (!r) return r;
    std::vector<int> levels(counts);
    std::sort(levels.begin(), levels.end());
    levels.erase(std::unique(levels.begin(), levels.end()), levels.end());
    int done = 0;
    for (int lvl : levels) {
        if (lvl <= done) continue;
        std::vector<BatchMember*> active;
        for (size_t i = 0; i < members.size(); ++i)
            if (counts[i] >= lvl) active.push_back(members[i]);
        if (!(r = phase(active, lvl - done))) return r;
        done = lvl;
    }
    IPL_HIP_TRY(hipStreamSynchronize(stream_));
    return Status::ok();
}

Status BatchRunner::graph_for_layout(hipGraphExec_t* exec) {
    for (Cached& c : graphs_)
        if (same_layout(c.layout, layout_)) {
            c.used = ++clock_;
            *exec = c.exec;
            return Status::ok();
        }
    IPL_HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    Status r = launch_layout(0);
    if (r) r = launch_layout(1);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(stream_, &g);
    if (!r) { if (g) (void)hipGraphDestroy(g); return r; }
    if (e != hipSuccess) return Status::err(IPL_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) return Status::err(IPL_ERR_HIP, std::string("batch graph instantiate: ") + hipGetErrorString(e2));
    ++captures_;
    if (graphs_.size() >= kMaxGraphs) {  // an evicted graph may still run on the stream: wait first
        auto lru = std::min_element(graphs_.begin(), graphs_.end(), [](const Cached& a, const Cached& b) { return a.used < b.used; });
        IPL_HIP_TRY(hipStreamSynchronize(stream_));
        (void)hipGraphExecDestroy(lru->exec);
        graphs_.erase(lru);
    }
    graphs_.push_back(Cached{layout_, x, ++clock_});
    *exec = x;
    return Status::ok();
}

void BatchRunner::drop_graph() {
    for (Cached& c : graphs_) (void)hipGraphExecDestroy(c.exec);
    graphs_.clear();
}

}  // namespace ipl
