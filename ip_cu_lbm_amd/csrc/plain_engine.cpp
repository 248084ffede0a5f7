// plain_engine.cpp - host side of the plain path: buffers, uploads, the stage loop, readout.
#include "hip_compat.hpp"

#include <cmath>
#include <cstring>

#include "engine.hpp"

namespace ipl {

Params make_params(const ipl_config& cfg, const Graph& g) {
    Params P{};
    P.dz = g.dz; P.dt = g.dt; P.c = g.c;
    for (int i = 0; i < 27; ++i) P.rates[i] = cfg.s_rates[i];
    P.rho_L = cfg.rho_L; P.rho_H = cfg.rho_H; P.beta_s = cfg.beta_s; P.gravity = cfg.gravity;
    P.sigma = cfg.sigma; P.pf_thickness = cfg.pf_thickness; P.pf_tau = cfg.pf_tau;
    P.boussinesq = cfg.boussinesq; P.collision = cfg.collision; P.scalar = cfg.scalar;
    P.wall_bc = cfg.wall_bc;
    P.flux_mode = g.flux_mode; P.flux_axis = g.flux_axis < 0 ? 0 : g.flux_axis; P.npts = g.npts;
    for (int a = 0; a < 3; ++a) { P.periodic[a] = cfg.periodic[a]; P.period[a] = g.period[a]; }
    return P;
}

// Initial state: the reference's driver fill (LockExchange.cu:84-139) and the ghost copies of
// set_neibghor_wall (basic.cu:221-226), computed once in FP64.
Status make_host_state(const Graph& g, const ipl_config& cfg, const ipl_init* init, HostState& hs) {
    const int64_t N = g.N, Nt = g.Ntot;
    auto take = [&](const double* src, double dflt, std::vector<double>& v) {
        v.assign(Nt, 0.0);
        for (int64_t i = 0; i < N; ++i) v[i] = src ? src[i] : dflt;
    };
    take(init ? init->p : nullptr, 0.0, hs.p);
    take(init ? init->u : nullptr, 0.0, hs.u);
    take(init ? init->v : nullptr, 0.0, hs.v);
    take(init ? init->w : nullptr, 0.0, hs.w);
    take(init ? init->S : nullptr, 0.0, hs.S);
    take(init ? init->phi : nullptr, 0.0, hs.phi);
    take(init ? init->tau : nullptr, cfg.tau, hs.tau);
    take(init ? init->taus : nullptr, cfg.taus, hs.taus);
    for (int64_t i = 0; i < N; ++i) {
        if (!(hs.tau[i] > 0) || !(hs.taus[i] > 0))
            return Status::err(IPL_ERR_INVALID_CONFIG, "tau and taus must be positive");
        if (!std::isfinite(hs.p[i]) || !std::isfinite(hs.u[i]) || !std::isfinite(hs.v[i]) || !std::isfinite(hs.w[i]))
            return Status::err(IPL_ERR_NONFINITE, "non-finite initial field");
    }
    hs.rho.assign(Nt, 0.0);
    for (int64_t i = 0; i < N; ++i) hs.rho[i] = cfg.rho_L + hs.phi[i] * (cfg.rho_H - cfg.rho_L) + hs.S[i] * cfg.beta_s;
    for (int64_t gi = N; gi < Nt; ++gi) {
        const int64_t o = g.neib[gi];
        hs.p[gi] = hs.p[o]; hs.rho[gi] = hs.rho[o]; hs.phi[gi] = hs.phi[o];
        hs.tau[gi] = hs.tau[o]; hs.taus[gi] = hs.taus[o];
    }
    const int Q = g.Q;
    const double c = g.c, c2 = c * c;
    hs.f.assign((size_t)Q * Nt, 0.0);
    for (int64_t i = 0; i < N; ++i) {
        const double u2 = hs.u[i] * hs.u[i] + hs.v[i] * hs.v[i] + hs.w[i] * hs.w[i];
        for (int k = 0; k < Q; ++k) {
            const double t = (g.dirs.cx[k] * c * hs.u[i] + g.dirs.cy[k] * c * hs.v[i] + g.dirs.cz[k] * c * hs.w[i]) / c2;
            hs.f[(size_t)k * Nt + i] = g.dirs.w[k] * (hs.p[i] + 3.0 * t + 4.5 * t * t - 1.5 * u2 / c2);
        }
    }
    if (g.Qs) {
        const int Qs = g.Qs;
        const bool linear = Qs == 7;
        const double cs2inv = linear ? 4.0 : 3.0;
        hs.g.assign((size_t)Qs * Nt, 0.0);
        for (int64_t i = 0; i < N; ++i) {
            const double s = cfg.scalar == IPL_SCALAR_SALINITY ? hs.S[i] : hs.phi[i];
            const double u2 = hs.u[i] * hs.u[i] + hs.v[i] * hs.v[i] + hs.w[i] * hs.w[i];
            for (int k = 0; k < Qs; ++k) {
                const double t = (g.sdirs.cx[k] * c * hs.u[i] + g.sdirs.cy[k] * c * hs.v[i] + g.sdirs.cz[k] * c * hs.w[i]) / c2;
                double e = 1.0 + cs2inv * t;
                if (!linear) e += 4.5 * t * t - 1.5 * u2 / c2;
                hs.g[(size_t)k * Nt + i] = g.sdirs.w[k] * s * e;
            }
        }
    }
    return Status::ok();
}

namespace {

template <typename T>
Status upload(T* dst, const std::vector<double>& src, size_t n, hipStream_t st) {
    std::vector<T> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (T)src[i];
    IPL_HIP_TRY(hipMemcpyAsync(dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice, st));
    IPL_HIP_TRY(hipStreamSynchronize(st));
    return Status::ok();
}

template <typename T>
class PlainEngine final : public Engine {
public:
    PlainEngine(const Graph& g, const ipl_config& cfg) : g_(g), cfg_(cfg) { P_ = make_params(cfg, g); }

    Status init(const HostState& hs, hipStream_t st) {
        stream = st;
        mem_.bind(st);
        const int64_t N = g_.N, Nt = g_.Ntot, G = g_.G;
        const int Q = g_.Q, Qs = g_.Qs;
        PlainView<T>& s = v_;
        std::memset(&s, 0, sizeof(s));
        s.N = N; s.Ntot = Nt;
        Status r;
        T** popf[5] = {&s.f, &s.ftmp, &s.fout, &s.feq, &s.Fk};
        for (T** p : popf) if (!(r = mem_.alloc_n(p, (size_t)Q * Nt))) return r;
        if (Qs) {
            T** popg[4] = {&s.g, &s.gtmp, &s.gout, &s.geq};
            for (T** p : popg) if (!(r = mem_.alloc_n(p, (size_t)Qs * Nt))) return r;
        }
        T** fields[] = {&s.p, &s.u, &s.v, &s.w, &s.rho, &s.phi, &s.S, &s.tau, &s.taus, &s.Fx, &s.Fy, &s.Fz,
                        &s.phiold, &s.uold, &s.vold, &s.wold};
        for (T** p : fields) if (!(r = mem_.alloc_n(p, (size_t)Nt))) return r;
        for (int a = 0; a < 3; ++a) {
            if (!(r = mem_.alloc_n(&s.pos[a], (size_t)Nt))) return r;
            if (!(r = mem_.alloc_n(&s.del[a], (size_t)Nt))) return r;
            if (!(r = mem_.alloc_n(&s.rdel[a], (size_t)Nt))) return r;
        }
        if (!(r = mem_.alloc_n(&s.W, g_.weights.size()))) return r;
        int32_t *nb, *nB, *nK, *snB = nullptr, *snK = nullptr;
        if (!(r = mem_.alloc_n(&nb, g_.neib.size()))) return r;
        if (!(r = mem_.alloc_n(&nB, g_.nextB.size()))) return r;
        if (!(r = mem_.alloc_n(&nK, g_.nextK.size()))) return r;
        if (Qs) {
            if (!(r = mem_.alloc_n(&snB, g_.snextB.size()))) return r;
            if (!(r = mem_.alloc_n(&snK, g_.snextK.size()))) return r;
        }
        IPL_HIP_TRY(hipMemcpyAsync(nb, g_.neib.data(), g_.neib.size() * 4, hipMemcpyHostToDevice, st));
        if (G) {
            IPL_HIP_TRY(hipMemcpyAsync(nB, g_.nextB.data(), g_.nextB.size() * 4, hipMemcpyHostToDevice, st));
            IPL_HIP_TRY(hipMemcpyAsync(nK, g_.nextK.data(), g_.nextK.size() * 4, hipMemcpyHostToDevice, st));
            if (Qs) {
                IPL_HIP_TRY(hipMemcpyAsync(snB, g_.snextB.data(), g_.snextB.size() * 4, hipMemcpyHostToDevice, st));
                IPL_HIP_TRY(hipMemcpyAsync(snK, g_.snextK.data(), g_.snextK.size() * 4, hipMemcpyHostToDevice, st));
            }
        }
        s.neib = nb; s.nextB = nB; s.nextK = nK; s.snextB = snB; s.snextK = snK;
        // zero everything first (ghost scratch entries start at zero, basic.cu:216-219)
        for (T** p : popf) IPL_HIP_TRY(hipMemsetAsync(*p, 0, (size_t)Q * Nt * sizeof(T), st));
        if (Qs) {
            T* pg[4] = {s.g, s.gtmp, s.gout, s.geq};
            for (T* p : pg) IPL_HIP_TRY(hipMemsetAsync(p, 0, (size_t)Qs * Nt * sizeof(T), st));
        }
        for (T** p : fields) IPL_HIP_TRY(hipMemsetAsync(*p, 0, (size_t)Nt * sizeof(T), st));
        if (!(r = upload(s.f, hs.f, (size_t)Q * Nt, st))) return r;
        if (Qs && !(r = upload(s.g, hs.g, (size_t)Qs * Nt, st))) return r;
        const std::vector<double>* hf[] = {&hs.p, &hs.u, &hs.v, &hs.w, &hs.rho, &hs.phi, &hs.S, &hs.tau, &hs.taus};
        T* df[] = {s.p, s.u, s.v, s.w, s.rho, s.phi, s.S, s.tau, s.taus};
        for (int q = 0; q < 9; ++q) if (!(r = upload(df[q], *hf[q], (size_t)Nt, st))) return r;
        // the phase-field history starts equal to the initial fields (LockExchange.cu:187)
        if (!(r = upload(s.phiold, hs.phi, (size_t)Nt, st))) return r;
        if (!(r = upload(s.uold, hs.u, (size_t)Nt, st))) return r;
        if (!(r = upload(s.vold, hs.v, (size_t)Nt, st))) return r;
        if (!(r = upload(s.wold, hs.w, (size_t)Nt, st))) return r;
        for (int a = 0; a < 3; ++a) {
            std::vector<double> rd(Nt);
            for (int64_t i = 0; i < Nt; ++i) rd[i] = g_.dz / g_.del[a][i];
            if (!(r = upload(s.pos[a], g_.pos[a], (size_t)Nt, st))) return r;
            if (!(r = upload(s.del[a], g_.del[a], (size_t)Nt, st))) return r;
            if (!(r = upload(s.rdel[a], rd, (size_t)Nt, st))) return r;
        }
        if (!(r = upload(s.W, g_.weights, g_.weights.size(), st))) return r;
        return Status::ok();
    }

    Status stage(int s) override {
        const int rc = plain_stage<T>(v_, P_, g_.Q, g_.Qs, s, stream);
        if (rc != IPL_OK) return Status::err(rc, "stage not available for this lattice");
        IPL_HIP_TRY(hipGetLastError());
        return Status::ok();
    }

    Status step(int nsteps) override {
        for (int n = 0; n < nsteps; ++n)
            for (int s = IPL_STAGE_EQU_F; s <= IPL_STAGE_RESET_F; ++s) {
                Status r = stage(s);
                if (!r) return r;
            }
        return Status::ok();
    }

    // the plain path has no single dominant kernel: out[0] is the whole step
    Status profile(int nsteps, double out[3]) override {
        std::vector<hipEvent_t> ev((size_t)nsteps + 1, nullptr);
        for (auto& e : ev) IPL_HIP_TRY(hipEventCreate(&e));
        Status r = Status::ok();
        IPL_HIP_TRY(hipEventRecord(ev[0], stream));
        for (int n = 0; n < nsteps && r; ++n) {
            r = step(1);
            if (r) IPL_HIP_TRY(hipEventRecord(ev[n + 1], stream));
        }
        double tot = 0;
        if (r) {
            IPL_HIP_TRY(hipEventSynchronize(ev[nsteps]));
            for (int n = 0; n < nsteps; ++n) {
                float m = 0;
                IPL_HIP_TRY(hipEventElapsedTime(&m, ev[n], ev[n + 1]));
                tot += m;
            }
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        if (!r) return r;
        out[0] = tot / nsteps;
        out[1] = 0;
        out[2] = tot;
        return Status::ok();
    }

    T* field_ptr(int field) const {
        switch (field) {
        case IPL_FIELD_P: return v_.p; case IPL_FIELD_U: return v_.u; case IPL_FIELD_V: return v_.v;
        case IPL_FIELD_W: return v_.w; case IPL_FIELD_RHO: return v_.rho; case IPL_FIELD_S: return v_.S;
        case IPL_FIELD_PHI: return v_.phi; case IPL_FIELD_FX: return v_.Fx; case IPL_FIELD_FY: return v_.Fy;
        case IPL_FIELD_FZ: return v_.Fz; case IPL_FIELD_TAU: return v_.tau; case IPL_FIELD_TAUS: return v_.taus;
        default: return nullptr;
        }
    }

    Status read(int field, double* host, size_t count) override {
        const int64_t N = g_.N, Nt = g_.Ntot;
        if (field == IPL_FIELD_F || field == IPL_FIELD_G) {
            const int q = field == IPL_FIELD_F ? g_.Q : g_.Qs;
            T* src = field == IPL_FIELD_F ? v_.f : v_.g;
            if (!q || count != (size_t)q * N) return Status::err(IPL_ERR_ARGUMENT, "population read size mismatch");
            std::vector<T> tmp((size_t)q * N);
            for (int k = 0; k < q; ++k)
                IPL_HIP_TRY(hipMemcpyAsync(tmp.data() + (size_t)k * N, src + (size_t)k * Nt, N * sizeof(T),
                                           hipMemcpyDeviceToHost, stream));
            IPL_HIP_TRY(hipStreamSynchronize(stream));
            for (size_t i = 0; i < tmp.size(); ++i) host[i] = (double)tmp[i];
            return Status::ok();
        }
        T* src = field_ptr(field);
        if (!src || count != (size_t)N) return Status::err(IPL_ERR_ARGUMENT, "field id or size mismatch");
        std::vector<T> tmp(N);
        IPL_HIP_TRY(hipMemcpyAsync(tmp.data(), src, N * sizeof(T), hipMemcpyDeviceToHost, stream));
        IPL_HIP_TRY(hipStreamSynchronize(stream));
        for (int64_t i = 0; i < N; ++i) host[i] = (double)tmp[i];
        return Status::ok();
    }

    Status write(int field, const double* host, size_t count) override {
        const int64_t N = g_.N, Nt = g_.Ntot;
        if (field == IPL_FIELD_F || field == IPL_FIELD_G) {
            const int q = field == IPL_FIELD_F ? g_.Q : g_.Qs;
            T* dst = field == IPL_FIELD_F ? v_.f : v_.g;
            if (!q || count != (size_t)q * N) return Status::err(IPL_ERR_ARGUMENT, "population write size mismatch");
            std::vector<T> tmp(count);
            for (size_t i = 0; i < count; ++i) tmp[i] = (T)host[i];
            for (int k = 0; k < q; ++k)
                IPL_HIP_TRY(hipMemcpyAsync(dst + (size_t)k * Nt, tmp.data() + (size_t)k * N, N * sizeof(T),
                                           hipMemcpyHostToDevice, stream));
            IPL_HIP_TRY(hipStreamSynchronize(stream));
            return Status::ok();
        }
        T* dst = field_ptr(field);
        if (!dst || count != (size_t)N) return Status::err(IPL_ERR_ARGUMENT, "field id or size mismatch");
        std::vector<T> tmp(N);
        for (int64_t i = 0; i < N; ++i) tmp[i] = (T)host[i];
        IPL_HIP_TRY(hipMemcpyAsync(dst, tmp.data(), N * sizeof(T), hipMemcpyHostToDevice, stream));
        IPL_HIP_TRY(hipStreamSynchronize(stream));
        return Status::ok();
    }

    Status global_sums(double out[3], double* scratch) override {
        const T* f[3] = {v_.p, v_.S, v_.phi};
        const double dz = g_.dz;
        return device_weighted_sums<T>(f, 3, v_.del[0], v_.del[1], v_.del[2], g_.N, 1.0 / (dz * dz * dz), out, stream,
                                       scratch);
    }

    Status check_finite(bool* finite) override {
        const T* f[4] = {v_.p, v_.u, v_.v, v_.w};
        return device_all_finite<T>(f, 4, g_.N, finite, stream);
    }

    int64_t device_bytes() const override { return (int64_t)mem_.total(); }
    int path() const override { return IPL_PATH_PLAIN; }

private:
    const Graph& g_;
    ipl_config cfg_;
    Params P_;
    PlainView<T> v_;
    DeviceArena mem_;
};

}  // namespace

std::unique_ptr<Engine> make_plain_engine(const Graph& g, const ipl_config& cfg, const HostState& hs,
                                          hipStream_t st, Status& status) {
    if (cfg.storage == IPL_F64) {
        auto e = std::make_unique<PlainEngine<double>>(g, cfg);
        status = e->init(hs, st);
        if (!status) return nullptr;
        return e;
    }
    if (cfg.storage == IPL_F32) {
        auto e = std::make_unique<PlainEngine<float>>(g, cfg);
        status = e->init(hs, st);
        if (!status) return nullptr;
        return e;
    }
    status = Status::err(IPL_ERR_UNSUPPORTED, "the plain path stores FP64 or FP32 populations");
    return nullptr;
}

}  // namespace ipl
