// api.cpp - the extern "C" boundary (include/ipl.h).
#include "hip_compat.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <fstream>
#include <string>
#include <vector>

#include "batch.hpp"
#include "engine.hpp"
#include "ibm.hpp"

struct ipl_ctx {
    ipl_config cfg{};
    ipl::Graph graph;
    std::vector<double> axis_pos[3], axis_wid[3];  // per axis index, for the file writers
    std::unique_ptr<ipl::Engine> engine;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string last_error;
    std::string path_reason;  // why the automatic choice took the plain path ("" on the fused path)
    int device = 0;
};

namespace {

int fail(ipl_ctx* ctx, const ipl::Status& s) {
    if (ctx) ctx->last_error = s.msg;
    return s.code;
}

thread_local std::string g_create_error;

}  // namespace

namespace {

// The initial fields' checks of the set-up (make_host_state): finite p, u, v, w; positive tau, tau_s.
ipl::Status check_initial_fields(const ipl::Graph& hdr, const ipl_config& cfg, const ipl_init* init) {
    const int64_t N = hdr.N;
    auto at = [&](const double* a, int64_t i, double dflt) { return a ? a[i] : dflt; };
    for (int64_t i = 0; i < N; ++i) {
        const double tau = at(init ? init->tau : nullptr, i, cfg.tau), taus = at(init ? init->taus : nullptr, i, cfg.taus);
        if (!(tau > 0) || !(taus > 0)) return ipl::Status::err(IPL_ERR_INVALID_CONFIG, "tau and taus must be positive");
    }
    const double* v[4] = {init ? init->p : nullptr, init ? init->u : nullptr, init ? init->v : nullptr, init ? init->w : nullptr};
    for (const double* a : v)
        if (a)
            for (int64_t i = 0; i < N; ++i)
                if (!std::isfinite(a[i])) return ipl::Status::err(IPL_ERR_NONFINITE, "non-finite initial field");
    return ipl::Status::ok();
}

bool uniform_rates(const ipl::Graph& hdr, const ipl_init* init) {
    for (const double* a : {init ? init->tau : nullptr, init ? init->taus : nullptr})
        if (a)
            for (int64_t i = 1; i < hdr.N; ++i)
                if (a[i] != a[0]) return false;
    return true;
}

}  // namespace

extern "C" {

const char* ipl_error_string(int code) {
    switch (code) {
    case IPL_OK: return "ok";
    case IPL_ERR_INVALID_CONFIG: return "invalid configuration";
    case IPL_ERR_HIP: return "HIP runtime error";
    case IPL_ERR_OUT_OF_MEMORY: return "out of device memory";
    case IPL_ERR_NONFINITE: return "non-finite value in the macroscopic fields";
    case IPL_ERR_UNSUPPORTED: return "configuration not supported by the requested path";
    case IPL_ERR_ARGUMENT: return "bad argument";
    case IPL_ERR_BUDGET: return "job does not fit the memory budget";
    default: return "unknown error";
    }
}

const char* ipl_path_reason(ipl_ctx* ctx) { return ctx ? ctx->path_reason.c_str() : ""; }

const char* ipl_last_error_message(ipl_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : g_create_error.c_str();
}

int ipl_create(const ipl_config* cfg, const ipl_init* init, ipl_ctx** out) {
    if (!cfg || !out) return IPL_ERR_ARGUMENT;
    *out = nullptr;
    g_create_error.clear();
    ipl_ctx* ctx = new (std::nothrow) ipl_ctx();
    if (!ctx) return IPL_ERR_OUT_OF_MEMORY;
    ctx->cfg = *cfg;
    ctx->cfg.solid = nullptr;
    for (int a = 0; a < 3; ++a) { ctx->cfg.pos[a] = nullptr; ctx->cfg.width[a] = nullptr; }
    // the fused path needs only the header of the graph, its cells' coordinates and one line's
    // stencil weights; the full tables (neighbours, ghosts, slip remap, per-cell weights) are
    // built only for the plain path
    std::string err;
    ipl::Graph hdr;
    if (!ipl::graph_header(*cfg, hdr, err)) {
        g_create_error = err;
        delete ctx;
        return IPL_ERR_INVALID_CONFIG;
    }
    std::string why;
    const int req = cfg->path;
    bool fused_ok = req != IPL_PATH_PLAIN && ipl::fused_supported(hdr, *cfg, why);
    {
        const ipl::Status v = check_initial_fields(hdr, *cfg, init);
        if (!v) { g_create_error = v.msg; delete ctx; return v.code; }
    }
    if (fused_ok && !cfg->les && !uniform_rates(hdr, init)) {
        fused_ok = false;
        why = "the fused path needs uniform tau and taus";
    }
    if (req == IPL_PATH_FUSED && !fused_ok) {
        g_create_error = "fused path: " + why;
        delete ctx;
        return IPL_ERR_UNSUPPORTED;
    }
    auto bail = [&](const ipl::Status& s) {
        g_create_error = s.msg;
        ctx->engine.reset();  // its memory returns to the pool on the stream, destroyed after it
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
        if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
        delete ctx;
        return s.code;
    };
    {
        hipError_t e = hipGetDevice(&ctx->device);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
        if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
        if (e != hipSuccess) return bail(ipl::Status::err(IPL_ERR_HIP, std::string("HIP setup: ") + hipGetErrorString(e)));
    }
    for (int a = 0; a < 3; ++a) {
        const int na = a == 0 ? cfg->nx : (a == 1 ? cfg->ny : cfg->nz);
        ctx->axis_pos[a].assign(cfg->pos[a], cfg->pos[a] + na);
        ctx->axis_wid[a].assign(cfg->width[a], cfg->width[a] + na);
    }
    // engines are built from the caller's config: its geometry pointers are valid for this call
    ipl::Status st;
    if (fused_ok) {
        if (!ipl::build_box_graph(*cfg, ctx->graph, err)) return bail(ipl::Status::err(IPL_ERR_INVALID_CONFIG, err));
        ctx->engine = ipl::make_fused_engine(ctx->graph, *cfg, init, ctx->stream, st);
        if (!ctx->engine && req == IPL_PATH_FUSED) return bail(st);
        if (!ctx->engine) why = st.msg;
    }
    if (!ctx->engine) {
        ctx->graph = ipl::Graph();
        if (!ipl::build_graph(*cfg, ctx->graph, err)) return bail(ipl::Status::err(IPL_ERR_INVALID_CONFIG, err));
        ipl::HostState hs;
        st = ipl::make_host_state(ctx->graph, *cfg, init, hs);
        if (!st) return bail(st);
        ctx->path_reason = req == IPL_PATH_PLAIN ? "the plain path was requested" : "fused path: " + why;
        ctx->engine = ipl::make_plain_engine(ctx->graph, *cfg, hs, ctx->stream, st);
    }
    if (!ctx->engine) return bail(st ? ipl::Status::err(IPL_ERR_UNSUPPORTED, "no engine") : st);
    *out = ctx;
    return IPL_OK;
}

int ipl_step(ipl_ctx* ctx, int32_t nsteps) {
    if (!ctx || nsteps < 0) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->step(nsteps);
    if (!s) return fail(ctx, s);
    if (ctx->cfg.check_finite) {
        bool finite = true;
        s = ctx->engine->check_finite(&finite);
        if (!s) return fail(ctx, s);
        if (!finite) return fail(ctx, ipl::Status::err(IPL_ERR_NONFINITE, "NaN or Inf in p/u/v/w"));
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ctx, ipl::Status::err(IPL_ERR_HIP, hipGetErrorString(e)));
    return IPL_OK;
}

int ipl_step_timed(ipl_ctx* ctx, int32_t nsteps, double* ms) {
    if (!ctx || nsteps < 0 || !ms) return IPL_ERR_ARGUMENT;
    hipError_t e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e != hipSuccess) return fail(ctx, ipl::Status::err(IPL_ERR_HIP, hipGetErrorString(e)));
    ipl::Status s = ctx->engine->step(nsteps);
    if (!s) return fail(ctx, s);
    e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
    if (e != hipSuccess) return fail(ctx, ipl::Status::err(IPL_ERR_HIP, hipGetErrorString(e)));
    *ms = t;
    return IPL_OK;
}

int ipl_profile_steps(ipl_ctx* ctx, int32_t nsteps, double out[3]) {
    if (!ctx || nsteps <= 0 || !out) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->profile(nsteps, out);
    return s ? IPL_OK : fail(ctx, s);
}

int ipl_stage(ipl_ctx* ctx, int32_t stage) {
    if (!ctx) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->stage(stage);
    if (!s) return fail(ctx, s);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ctx, ipl::Status::err(IPL_ERR_HIP, hipGetErrorString(e)));
    return IPL_OK;
}

int ipl_read(ipl_ctx* ctx, int32_t field, double* host, size_t count) {
    if (!ctx || !host) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->read(field, host, count);
    return s ? IPL_OK : fail(ctx, s);
}

int ipl_write(ipl_ctx* ctx, int32_t field, const double* host, size_t count) {
    if (!ctx || !host) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->write(field, host, count);
    return s ? IPL_OK : fail(ctx, s);
}

int ipl_global_sums(ipl_ctx* ctx, double out[3]) {
    if (!ctx || !out) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->global_sums(out, nullptr);
    return s ? IPL_OK : fail(ctx, s);
}

int ipl_global_sums_noalloc(ipl_ctx* ctx, double out[3]) {
    if (!ctx || !out) return IPL_ERR_ARGUMENT;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != ctx->device)
        return fail(ctx, ipl::Status::err(IPL_ERR_ARGUMENT, "the context's device must be current"));
    double* scratch = nullptr;
    size_t capacity = 0;
    ipl::Status s = ipl::noalloc_scratch(&scratch, &capacity);
    if (!s) return fail(ctx, s);
    if (ipl::weighted_sums_scratch_doubles() > capacity)
        return fail(ctx, ipl::Status::err(IPL_ERR_ARGUMENT, "the reduction needs more scratch than the library holds"));
    s = ctx->engine->global_sums(out, scratch);
    return s ? IPL_OK : fail(ctx, s);
}

int ipl_info(ipl_ctx* ctx, int64_t* n_calc, int64_t* n_ghost, int64_t* device_bytes, int32_t* path) {
    if (!ctx) return IPL_ERR_ARGUMENT;
    if (n_calc) *n_calc = ctx->graph.N;
    if (n_ghost) *n_ghost = ctx->graph.G;
    if (device_bytes) *device_bytes = ctx->engine->device_bytes();
    if (path) *path = ctx->engine->path();
    return IPL_OK;
}

int ipl_destroy(ipl_ctx* ctx) {
    if (!ctx) return IPL_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->engine.reset();  // frees on ctx->stream (stream-ordered), before the stream goes
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return IPL_OK;
}

int ipl_ib_attach(ipl_ctx* ctx, const ipl_ib_config* ib) {
    if (!ctx || !ib) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->ib_attach(*ib);
    return s ? IPL_OK : fail(ctx, s);
}

int ipl_ib_stage(ipl_ctx* ctx, int32_t stage) {
    if (!ctx) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->ib_stage(stage);
    if (!s) return fail(ctx, s);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ctx, ipl::Status::err(IPL_ERR_HIP, hipGetErrorString(e)));
    s = ctx->engine->post_step();
    return s ? IPL_OK : fail(ctx, s);
}

int ipl_ib_body_state(ipl_ctx* ctx, double out[19]) {
    if (!ctx || !out) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->ib_body_state(out);
    return s ? IPL_OK : fail(ctx, s);
}

int ipl_ib_read_markers(ipl_ctx* ctx, int32_t which, double* out, size_t count) {
    if (!ctx || !out) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->ib_read_markers(which, out, count);
    return s ? IPL_OK : fail(ctx, s);
}

}  // extern "C"

struct ipl_batch {
    std::vector<ipl_ctx*> members;
    std::string last_error;
    int launch = IPL_BATCH_LAUNCH_BATCHED;
    int64_t max_lines = 0;  // members with more lines step alone (0: no limit)
    int64_t own_steps = 0;  // member-steps launched on the members' own streams

    ipl::BatchMember* batched_member(size_t i) const {
        if (launch != IPL_BATCH_LAUNCH_BATCHED) return nullptr;
        ipl::BatchMember* m = members[i]->engine->batch_member();
        return m && (max_lines == 0 || m->batch_lines() <= max_lines) ? m : nullptr;
    }
    ipl::BatchRunner runner;
};

extern "C" {

int ipl_batch_create(ipl_batch** out) {
    if (!out) return IPL_ERR_ARGUMENT;
    *out = new (std::nothrow) ipl_batch();
    return *out ? IPL_OK : IPL_ERR_OUT_OF_MEMORY;
}

int ipl_batch_add(ipl_batch* b, ipl_ctx* ctx) {
    if (!b || !ctx) return IPL_ERR_ARGUMENT;
    for (ipl_ctx* m : b->members)
        if (m == ctx) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->set_shared(true);
    if (!s) return fail(ctx, s);
    b->members.push_back(ctx);
    return IPL_OK;
}

int ipl_batch_remove(ipl_batch* b, ipl_ctx* ctx) {
    if (!b || !ctx) return IPL_ERR_ARGUMENT;
    for (size_t i = 0; i < b->members.size(); ++i)
        if (b->members[i] == ctx) {
            b->members.erase(b->members.begin() + (long)i);
            ipl::Status s = ctx->engine->set_shared(false);
            return s ? IPL_OK : fail(ctx, s);
        }
    return IPL_ERR_ARGUMENT;
}

int ipl_batch_step(ipl_batch* b, const int32_t* nsteps, int32_t all) {
    if (!b || !nsteps) return IPL_ERR_ARGUMENT;
    const size_t n = b->members.size();
    for (size_t i = 0; i < n; ++i)
        if ((all ? nsteps[0] : nsteps[i]) < 0) return IPL_ERR_ARGUMENT;
    // members whose step is the fused kernel alone join the batched launches; the others launch
    // their own steps, each on its own stream, first
    std::vector<ipl::BatchMember*> batched;
    std::vector<int> counts;
    for (size_t i = 0; i < n; ++i) {
        const int k = all ? nsteps[0] : nsteps[i];
        if (k == 0) continue;
        ipl::BatchMember* m = b->batched_member(i);
        if (m) {
            batched.push_back(m);
            counts.push_back(k);
            continue;
        }
        ipl::Status s = b->members[i]->engine->launch_steps(k);
        if (!s) { b->last_error = s.msg; return s.code; }
        b->own_steps += k;
    }
    if (!batched.empty()) {
        for (size_t i = 0; i < n; ++i)  // their own streams are idle before the batched launches
            if (b->batched_member(i) && (all ? nsteps[0] : nsteps[i]) > 0) {
                hipError_t e = hipStreamSynchronize(b->members[i]->stream);
                if (e != hipSuccess) { b->last_error = hipGetErrorString(e); return IPL_ERR_HIP; }
            }
        ipl::Status s = b->runner.run(batched, counts);
        if (!s) { b->last_error = s.msg; return s.code; }
    }
    for (size_t i = 0; i < n; ++i) {
        ipl_ctx* c = b->members[i];
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) { b->last_error = hipGetErrorString(e); return IPL_ERR_HIP; }
        ipl::Status s = c->engine->post_step();
        if (!s) { b->last_error = s.msg; return s.code; }
    }
    return IPL_OK;
}

int ipl_batch_set_launch(ipl_batch* b, int32_t launch) {
    if (!b || (launch != IPL_BATCH_LAUNCH_PER_MEMBER && launch != IPL_BATCH_LAUNCH_BATCHED)) return IPL_ERR_ARGUMENT;
    b->launch = launch;
    return IPL_OK;
}

int ipl_batch_set_max_lines(ipl_batch* b, int64_t max_lines) {
    if (!b || max_lines < 0) return IPL_ERR_ARGUMENT;
    b->max_lines = max_lines;
    return IPL_OK;
}

int ipl_batch_counters(ipl_batch* b, int64_t* out) {
    if (!b || !out) return IPL_ERR_ARGUMENT;
    out[0] = b->runner.launches();
    out[1] = b->runner.member_steps();
    out[2] = b->own_steps;
    out[3] = b->runner.graph_captures();
    return IPL_OK;
}

int ipl_batch_destroy(ipl_batch* b) {
    delete b;
    return IPL_OK;
}

const char* ipl_batch_last_error_message(ipl_batch* b) { return b ? b->last_error.c_str() : ""; }

int ipl_set_graph(ipl_ctx* ctx, int32_t on) {
    if (!ctx) return IPL_ERR_ARGUMENT;
    ctx->engine->set_graph(on != 0);
    return IPL_OK;
}

int ipl_prepare(ipl_ctx* ctx) {
    if (!ctx) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->prepare();
    return s ? IPL_OK : fail(ctx, s);
}

int ipl_set_shared(ipl_ctx* ctx, int32_t on) {
    if (!ctx) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->set_shared(on != 0);
    return s ? IPL_OK : fail(ctx, s);
}

int ipl_graph_captures(ipl_ctx* ctx, int32_t* count) {
    if (!ctx || !count) return IPL_ERR_ARGUMENT;
    *count = ctx->engine->graph_captures();
    return IPL_OK;
}

int ipl_estimate_device_bytes(const ipl_config* cfg, int64_t* bytes) {
    if (!cfg || !bytes) return IPL_ERR_ARGUMENT;
    // the path ipl_create would take: the fused path's checks on the graph's header (no tables)
    ipl::Graph hdr;
    std::string why;
    if (!ipl::graph_header(*cfg, hdr, why)) {
        g_create_error = why;
        return IPL_ERR_INVALID_CONFIG;
    }
    const bool fused = cfg->path != IPL_PATH_PLAIN && ipl::fused_supported(hdr, *cfg, why);
    const int64_t n = hdr.N;
    const int Q = cfg->lattice;
    const int Qs = cfg->scalar == IPL_SCALAR_NONE ? 0 : (cfg->scalar_lattice ? cfg->scalar_lattice : Q);
    const int64_t s = cfg->storage == IPL_F64 ? 8 : (cfg->storage == IPL_F32 ? 4 : 2);
    const int64_t m = cfg->storage == IPL_F64 ? 8 : 4;
    const int64_t per_axis = 4 * (int64_t)(cfg->nx + cfg->ny + cfg->nz) * 8;
    if (fused) {
        // fused: ping-pong populations, force rows, the phase field's scalar-field and history rows,
        // FP64 cell volumes and the FP64 readout staging rows
        int64_t b = 2 * (int64_t)(Q + Qs) * s * n + 6 * m * n + 8 * n + 7 * 8 * n;
        if (cfg->scalar == IPL_SCALAR_PHASE_FIELD) b += 6 * m * n;
        if (cfg->les) b += 6 * m * n;  // velocity, viscosity, tau and tau_s rows
        *bytes = b + per_axis + (1 << 20);
    } else {
        // plain: five population sets per field, neighbour table, ghosts (~n^(2/3) of n), fields
        const int64_t nt = n + 6 * (int64_t)std::cbrt((double)n) * (int64_t)std::cbrt((double)n);
        *bytes = 5 * (int64_t)(Q + Qs) * s * nt + 4 * (int64_t)Q * nt + 20 * m * nt + per_axis + (1 << 20);
    }
    return IPL_OK;
}

int ipl_estimate_ib_bytes(const ipl_config* cfg, const ipl_ib_config* ib, int64_t* bytes) {
    if (!cfg || !ib || !bytes || ib->n_markers < 0 || (ib->n_markers > 0 && !ib->marker_pos)) return IPL_ERR_ARGUMENT;
    const int n[3] = {cfg->nx, cfg->ny, cfg->nz};
    const int dim = cfg->lattice == 9 ? 2 : 3;
    *bytes = ipl::IbSolver::estimate_bytes(*ib, n, cfg->periodic, dim, cfg->dz);
    return IPL_OK;
}

int ipl_set_open_boundary(ipl_ctx* ctx, const int32_t* cells, int64_t n, const double* p, const double* u,
                          const double* v, const double* w) {
    if (!ctx || n < 0 || (n > 0 && (!cells || !p || !u || !v || !w))) return IPL_ERR_ARGUMENT;
    std::vector<double> vals((size_t)4 * n);
    for (int64_t i = 0; i < n; ++i) {
        vals[4 * i] = p[i]; vals[4 * i + 1] = u[i]; vals[4 * i + 2] = v[i]; vals[4 * i + 3] = w[i];
    }
    ipl::Status s = ctx->engine->set_open_boundary(cells, n, vals.data());
    return s ? IPL_OK : fail(ctx, s);
}

namespace {

// the fields of the reference's output() rows, canonical order: p, S, phi, rho, u, v, w, Fx, Fy, Fz
ipl::Status read_output_fields(ipl_ctx* ctx, std::vector<double> (&v)[10]) {
    static const int fields[10] = {IPL_FIELD_P, IPL_FIELD_S, IPL_FIELD_PHI, IPL_FIELD_RHO, IPL_FIELD_U,
                                   IPL_FIELD_V, IPL_FIELD_W, IPL_FIELD_FX, IPL_FIELD_FY, IPL_FIELD_FZ};
    const size_t N = (size_t)ctx->graph.N;
    for (int q = 0; q < 10; ++q) {
        v[q].assign(N, 0.0);
        ipl::Status s = ctx->engine->read(fields[q], v[q].data(), N);
        if (!s) return s;
    }
    return ipl::Status::ok();
}

}  // namespace

int ipl_write_fields_csv(ipl_ctx* ctx, const char* path) {
    if (!ctx || !path) return IPL_ERR_ARGUMENT;
    std::vector<double> v[10];
    ipl::Status s = read_output_fields(ctx, v);
    if (!s) return fail(ctx, s);
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(ctx, ipl::Status::err(IPL_ERR_ARGUMENT, std::string("cannot open ") + path));
    const ipl::Graph& g = ctx->graph;
    std::fprintf(fp, "num_calc,%d\n", (int)g.N);
    std::fprintf(fp, "x       ,y       ,z       ,delx    ,dely    ,pressure,salinity ,phi     ,rho     ,u       ,v       ,w       ,Fx      ,Fy      ,Fz      \n");
    // cell positions and widths from the axes (a calc cell's are those of its indices)
    auto pos = [&](int64_t i, int a) { return ctx->axis_pos[a][(size_t)g.cell_ijk[3 * i + a]]; };
    auto wid = [&](int64_t i, int a) { return ctx->axis_wid[a][(size_t)g.cell_ijk[3 * i + a]]; };
    for (int64_t i = 0; i < g.N; ++i)
        std::fprintf(fp, "%f,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f,%f\n", pos(i, 0), pos(i, 1), pos(i, 2),
                     wid(i, 0), wid(i, 1), v[0][i], v[1][i], v[2][i], v[3][i], v[4][i], v[5][i], v[6][i], v[7][i],
                     v[8][i], v[9][i]);
    const bool ok = std::fclose(fp) == 0;
    return ok ? IPL_OK : fail(ctx, ipl::Status::err(IPL_ERR_ARGUMENT, std::string("write failed: ") + path));
}

int ipl_write_markers_csv(ipl_ctx* ctx, const char* path) {
    if (!ctx || !path) return IPL_ERR_ARGUMENT;
    double body[19];
    ipl::Status s = ctx->engine->ib_body_state(body);
    if (!s) return fail(ctx, s);
    const int64_t M = ctx->engine->ib_marker_count();
    std::vector<double> pos(3 * M), vel(3 * M), gw(3 * M);
    if (!(s = ctx->engine->ib_read_markers(IPL_IB_POS, pos.data(), pos.size()))) return fail(ctx, s);
    if (!(s = ctx->engine->ib_read_markers(IPL_IB_VEL, vel.data(), vel.size()))) return fail(ctx, s);
    if (!(s = ctx->engine->ib_read_markers(IPL_IB_FORCE, gw.data(), gw.size()))) return fail(ctx, s);
    std::ofstream file(path);
    if (!file) return fail(ctx, ipl::Status::err(IPL_ERR_ARGUMENT, std::string("cannot open ") + path));
    file << "x,y,z,velx,vely,velz,ax,ay,az\n";
    for (int64_t i = 0; i < M; ++i)
        file << pos[i * 3] << "," << pos[i * 3 + 1] << "," << pos[i * 3 + 2] << "," << vel[i * 3] << "," << vel[i * 3 + 1]
             << "," << vel[i * 3 + 2] << "," << gw[i * 3] << "," << gw[i * 3 + 1] << "," << gw[i * 3 + 2] << "\n";
    file.close();
    return file ? IPL_OK : fail(ctx, ipl::Status::err(IPL_ERR_ARGUMENT, std::string("write failed: ") + path));
}

int ipl_write_cd_csv(const char* path, const double* c_d, const double* time, int64_t n) {
    if (!path || n < 0 || (n > 0 && (!c_d || !time))) return IPL_ERR_ARGUMENT;
    std::ofstream file(path);
    if (!file) return IPL_ERR_ARGUMENT;
    file << "C_D,C_time\n";
    for (int64_t i = 0; i < n; ++i) file << c_d[i] << "," << time[i] << "\n";  // the stream's default format
    file.close();
    return file ? IPL_OK : IPL_ERR_ARGUMENT;
}

int ipl_write_vtu(ipl_ctx* ctx, const char* path) {
    if (!ctx || !path) return IPL_ERR_ARGUMENT;
    std::vector<double> v[10];
    ipl::Status s = read_output_fields(ctx, v);
    if (!s) return fail(ctx, s);
    const ipl::Graph& g = ctx->graph;
    const int64_t n[3] = {ctx->cfg.nx, ctx->cfg.ny, ctx->cfg.nz};
    // corner coordinates per axis: the faces between cells
    std::vector<double> corner[3];
    for (int a = 0; a < 3; ++a) {
        corner[a].resize((size_t)n[a] + 1);
        corner[a][0] = ctx->axis_pos[a][0] - ctx->axis_wid[a][0] / 2.0;
        for (int64_t k = 1; k <= n[a]; ++k) corner[a][k] = ctx->axis_pos[a][k - 1] + ctx->axis_wid[a][k - 1] / 2.0;
    }
    auto cid = [&](int64_t i, int64_t j, int64_t l) { return (i * (n[1] + 1) + j) * (n[2] + 1) + l; };
    std::vector<int64_t> pid((size_t)((n[0] + 1) * (n[1] + 1) * (n[2] + 1)), -1);
    std::vector<int64_t> conn((size_t)8 * g.N);
    std::vector<int64_t> pts;
    for (int64_t c = 0; c < g.N; ++c) {
        const int64_t i = g.cell_ijk[3 * c], j = g.cell_ijk[3 * c + 1], l = g.cell_ijk[3 * c + 2];
        // VTK_HEXAHEDRON order: bottom face (z-) counter-clockwise from (x-, y-), then the top face
        const int64_t corners[8] = {cid(i, j, l), cid(i + 1, j, l), cid(i + 1, j + 1, l), cid(i, j + 1, l),
                                    cid(i, j, l + 1), cid(i + 1, j, l + 1), cid(i + 1, j + 1, l + 1), cid(i, j + 1, l + 1)};
        for (int q = 0; q < 8; ++q) {
            int64_t& p = pid[(size_t)corners[q]];
            if (p < 0) { p = (int64_t)pts.size(); pts.push_back(corners[q]); }
            conn[(size_t)8 * c + q] = p;
        }
    }
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(ctx, ipl::Status::err(IPL_ERR_ARGUMENT, std::string("cannot open ") + path));
    std::fprintf(fp, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    std::fprintf(fp, "<UnstructuredGrid>\n<Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n", (long long)pts.size(), (long long)g.N);
    std::fprintf(fp, "<Points>\n<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (int64_t q : pts) {
        const int64_t l = q % (n[2] + 1), j = (q / (n[2] + 1)) % (n[1] + 1), i = q / ((n[1] + 1) * (n[2] + 1));
        std::fprintf(fp, "%.17g %.17g %.17g\n", corner[0][i], corner[1][j], corner[2][l]);
    }
    std::fprintf(fp, "</DataArray>\n</Points>\n<Cells>\n<DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n");
    for (int64_t c = 0; c < g.N; ++c) {
        for (int q = 0; q < 8; ++q) std::fprintf(fp, q ? " %lld" : "%lld", (long long)conn[(size_t)8 * c + q]);
        std::fprintf(fp, "\n");
    }
    std::fprintf(fp, "</DataArray>\n<DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
    for (int64_t c = 0; c < g.N; ++c) std::fprintf(fp, "%lld\n", (long long)(8 * (c + 1)));
    std::fprintf(fp, "</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
    for (int64_t c = 0; c < g.N; ++c) std::fprintf(fp, "12\n");
    std::fprintf(fp, "</DataArray>\n</Cells>\n<CellData>\n");
    const char* names[4] = {"salinity", "phi", "rho", "pressure"};
    const int slot[4] = {1, 2, 3, 0};
    for (int q = 0; q < 4; ++q) {
        std::fprintf(fp, "<DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n", names[q]);
        for (int64_t c = 0; c < g.N; ++c) std::fprintf(fp, "%.17g\n", v[slot[q]][c]);
        std::fprintf(fp, "</DataArray>\n");
    }
    const char* vnames[2] = {"velocity", "force"};
    for (int q = 0; q < 2; ++q) {
        std::fprintf(fp, "<DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"3\" format=\"ascii\">\n", vnames[q]);
        for (int64_t c = 0; c < g.N; ++c)
            std::fprintf(fp, "%.17g %.17g %.17g\n", v[4 + 3 * q][c], v[5 + 3 * q][c], v[6 + 3 * q][c]);
        std::fprintf(fp, "</DataArray>\n");
    }
    std::fprintf(fp, "</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
    const bool ok = std::fclose(fp) == 0;
    return ok ? IPL_OK : fail(ctx, ipl::Status::err(IPL_ERR_ARGUMENT, std::string("write failed: ") + path));
}

int ipl_change_dt(ipl_ctx* ctx, double alpha) {
    if (!ctx) return IPL_ERR_ARGUMENT;
    ipl::Status s = ctx->engine->change_dt(alpha);
    if (!s) return fail(ctx, s);
    ctx->cfg.dt /= alpha;
    return IPL_OK;
}

int ipl_hydrostatic_pressure(const ipl_config* cfg, int32_t boussinesq, double center_z, double* p_out, int64_t count) {
    if (!cfg || !p_out) return IPL_ERR_ARGUMENT;
    ipl::Graph g;
    std::string err;
    if (!ipl::build_graph(*cfg, g, err)) {
        g_create_error = err;
        return IPL_ERR_INVALID_CONFIG;
    }
    if (count != g.N) return IPL_ERR_ARGUMENT;
    const double grav = 9.81, c2 = g.c * g.c;
    for (int64_t i = 0; i < g.N; ++i) {
        const double z = g.pos[2][i];
        p_out[i] = boussinesq ? 3.0 * grav * (cfg->nz / 2.0 * cfg->dz - z) / c2 / 1000.0 : 3.0 * grav * (center_z - z) / c2;
    }
    return IPL_OK;
}

int ipl_wall_cells(const ipl_config* cfg, int32_t face, int32_t* out, int64_t* count) {
    if (!cfg || !count || face < 1 || face > 6) return IPL_ERR_ARGUMENT;
    ipl::Graph g;
    std::string err;
    if (!ipl::build_graph(*cfg, g, err)) {
        g_create_error = err;
        return IPL_ERR_INVALID_CONFIG;
    }
    // directions of set_walln (set_bound.cu:11-13); the rest direction never has a ghost
    static const int d9[7] = {0, 1, 0, 3, 0, 2, 4}, d3[7] = {0, 1, 2, 3, 4, 5, 6};
    const int dir = g.Q == 9 ? d9[face] : d3[face];
    int64_t m = 0;
    for (int64_t i = 0; i < g.N; ++i)
        if (g.neib[(size_t)dir * g.Ntot + i] >= g.N) {
            if (out) {
                if (m >= *count) return IPL_ERR_ARGUMENT;
                out[m] = (int32_t)i;
            }
            ++m;
        }
    *count = m;
    return IPL_OK;
}

int ipl_build_tables(const ipl_config* cfg, int32_t which, void* out, int64_t* count) {
    if (!cfg || !count) return IPL_ERR_ARGUMENT;
    ipl::Graph g;
    std::string err;
    if (!ipl::build_graph(*cfg, g, err)) {
        g_create_error = err;
        return IPL_ERR_INVALID_CONFIG;
    }
    const std::vector<int32_t>* t = nullptr;
    if (which == 0) t = &g.neib;
    else if (which == 1) t = &g.nextB;
    else if (which == 2) t = &g.nextK;
    if (which == 3) {
        if (!out) { *count = (int64_t)g.weights.size(); return IPL_OK; }
        if (*count != (int64_t)g.weights.size()) return IPL_ERR_ARGUMENT;
        std::memcpy(out, g.weights.data(), g.weights.size() * sizeof(double));
        return IPL_OK;
    }
    if (!t) return IPL_ERR_ARGUMENT;
    if (!out) { *count = (int64_t)t->size(); return IPL_OK; }
    if (*count != (int64_t)t->size()) return IPL_ERR_ARGUMENT;
    std::memcpy(out, t->data(), t->size() * sizeof(int32_t));
    return IPL_OK;
}

}  // extern "C"
