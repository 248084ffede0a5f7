// graph.cpp - cell graph, ghost cells, slip remap tables and stencil weights (host, FP64).
#include "graph.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace ipl {

int DirTable::find(int x, int y, int z) const {
    for (int j = 0; j < Q; ++j)
        if (cx[j] == x && cy[j] == y && cz[j] == z) return j;
    return -1;
}

DirTable make_dirs(int Q) {
    static const int d2[9][3] = {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {-1, 0, 0}, {0, 0, -1},
                                 {1, 0, 1}, {-1, 0, 1}, {-1, 0, -1}, {1, 0, -1}};
    static const int d3[27][3] = {
        {0, 0, 0},  {1, 0, 0},  {0, 1, 0},  {-1, 0, 0}, {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
        {1, 1, 0},  {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, 0, 1},  {1, 0, -1}, {-1, 0, 1},
        {-1, 0, -1}, {0, 1, 1},  {0, 1, -1}, {0, -1, 1}, {0, -1, -1}, {1, 1, 1},  {1, 1, -1},
        {1, -1, 1}, {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
    DirTable t;
    t.Q = Q;
    for (int k = 0; k < Q; ++k) {
        const int* d = Q == 9 ? d2[k] : d3[k];
        t.cx[k] = d[0]; t.cy[k] = d[1]; t.cz[k] = d[2];
        const int n2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (Q == 9) t.w[k] = n2 == 0 ? 4.0 / 9 : (n2 == 1 ? 1.0 / 9 : 1.0 / 36);
        else if (Q == 27) t.w[k] = n2 == 0 ? 8.0 / 27 : (n2 == 1 ? 2.0 / 27 : (n2 == 2 ? 1.0 / 54 : 1.0 / 216));
        else if (Q == 19) t.w[k] = n2 == 0 ? 1.0 / 3 : (n2 == 1 ? 1.0 / 18 : 1.0 / 36);
        else t.w[k] = n2 == 0 ? 0.25 : 0.125;
    }
    for (int k = 0; k < Q; ++k) t.opp[k] = t.find(-t.cx[k], -t.cy[k], -t.cz[k]);
    return t;
}

void lagrange_weights(double* coeff, const double* x, double xa, int n) {
    for (int i = 0; i < n; ++i) {
        coeff[i] = 1.0;
        for (int j = 0; j < i; ++j) coeff[i] *= (xa - x[j]) / (x[i] - x[j]);
        for (int j = i + 1; j < n; ++j) coeff[i] *= (xa - x[j]) / (x[i] - x[j]);
    }
}

namespace {

// Slip remap of one ghost (set_bound2D/3D, set_bound.cu:25-147), expressed as the set of
// reflected axes: a component of the incoming direction is reflected when the owner's
// neighbour on the side the population comes from is missing; in 3D the xz / yz diagonal
// neighbour can restore or add a reflection of both components (set_bound.cu:477-501).  The
// source cell is the owner shifted by minus the unreflected components and the direction is
// the incoming one with the reflected components reversed; with nothing reflected the link
// falls back to bounce-back.
void slip_tables(const Graph& g, const DirTable& d, std::vector<int32_t>& nB, std::vector<int32_t>& nK) {
    const int Q = d.Q, QN = g.Q;
    const int64_t N = g.N, Ntot = g.Ntot;
    nB.assign((size_t)Q * g.G, 0);
    nK.assign((size_t)Q * g.G, 0);
    const DirTable& f = g.dirs;
    auto nbr = [&](int64_t o, int x, int y, int z) -> int64_t {
        const int dd = f.find(x, y, z);
        return dd < 0 ? -1 : g.neib[(size_t)dd * Ntot + o];
    };
    for (int64_t gi = 0; gi < g.G; ++gi) {
        const int64_t o = g.neib[(size_t)0 * Ntot + N + gi];
        for (int k = 1; k < Q; ++k) {
            const int c[3] = {d.cx[k], d.cy[k], d.cz[k]};
            int r[3] = {c[0], c[1], c[2]};
            const int kk = f.find(-c[0], -c[1], -c[2]);
            const bool src_missing = g.neib[(size_t)kk * Ntot + o] >= N;
            bool reflected_any = false;
            if (src_missing) {
                for (int a = 0; a < 3; ++a) {
                    if (c[a] == 0) continue;
                    int e[3] = {0, 0, 0};
                    e[a] = -c[a];
                    const int64_t t = nbr(o, e[0], e[1], e[2]);
                    if (t >= N) r[a] = 0;
                }
                if (g.dim == 3) {
                    const int pairs[2][2] = {{0, 2}, {1, 2}};
                    for (int p = 0; p < 2; ++p) {  // restore when the diagonal source side is open
                        const int a = pairs[p][0], b = pairs[p][1];
                        if (c[a] == 0 || c[b] == 0) continue;
                        int e[3] = {0, 0, 0};
                        e[a] = -c[a]; e[b] = -c[b];
                        const int64_t t = nbr(o, e[0], e[1], e[2]);
                        if (t >= 0 && t < N) { r[a] = c[a]; r[b] = c[b]; }
                    }
                    for (int p = 0; p < 2; ++p) {  // reflect both when it is closed
                        const int a = pairs[p][0], b = pairs[p][1];
                        if (r[a] == 0 || r[b] == 0) continue;
                        int e[3] = {0, 0, 0};
                        e[a] = -r[a]; e[b] = -r[b];
                        const int64_t t = nbr(o, e[0], e[1], e[2]);
                        if (t >= N) { r[a] = 0; r[b] = 0; }
                    }
                }
                for (int a = 0; a < 3; ++a) if (c[a] != 0 && r[a] == 0) reflected_any = true;
            }
            int b = 0, kn = d.opp[k];
            if (reflected_any) {
                int bv[3], kv[3];
                for (int a = 0; a < 3; ++a) {
                    const bool refl = c[a] != 0 && r[a] == 0;
                    bv[a] = refl ? 0 : -c[a];
                    kv[a] = refl ? -c[a] : c[a];
                }
                b = f.find(bv[0], bv[1], bv[2]);
                kn = d.find(kv[0], kv[1], kv[2]);
            }
            nB[(size_t)k * g.G + gi] = b;
            nK[(size_t)k * g.G + gi] = kn;
        }
    }
}

// Stencil positions of the slab reconstruction: ghost positions are pushed out by dz steps
// (set_posl_x / set_posl_y, adv.cu:15-28) and periodic neighbours are unwrapped.
void fix_positions(const Graph& g, const ipl_config& cfg, int axis, double* x, int n, int s) {
    if (cfg.periodic[axis]) {
        const double Lp = g.period[axis];
        const int c = n / 2;
        for (int j = c + 1; j < n; ++j) while (s * (x[j] - x[j - 1]) <= 0) x[j] += s * Lp;
        for (int j = c - 1; j >= 0; --j) while (s * (x[j + 1] - x[j]) <= 0) x[j] -= s * Lp;
    }
    const double d = g.dz;
    const int c = n / 2;
    for (int j = c - 1; j >= 0; --j) x[j] = s > 0 ? std::min(x[j], x[j + 1] - d) : std::max(x[j], x[j + 1] + d);
    for (int j = c + 1; j < n; ++j) x[j] = s > 0 ? std::max(x[j], x[j - 1] + d) : std::min(x[j], x[j - 1] - d);
}

}  // namespace

bool graph_header(const ipl_config& cfg, Graph& g, std::string& err) {
    const int Q = cfg.lattice;
    if (Q != 9 && Q != 19 && Q != 27) { err = "lattice must be 9, 19 or 27"; return false; }
    g.Q = Q;
    g.dim = Q == 9 ? 2 : 3;
    g.dirs = make_dirs(Q);
    g.n[0] = cfg.nx; g.n[1] = cfg.ny; g.n[2] = cfg.nz;
    if (cfg.nx < 1 || cfg.ny < 1 || cfg.nz < 1) { err = "grid sizes must be positive"; return false; }
    if (g.dim == 2 && cfg.ny != 1) { err = "D2Q9 needs ny == 1"; return false; }
    if (!(cfg.dz > 0) || !(cfg.dt > 0)) { err = "dz and dt must be positive"; return false; }
    g.dz = cfg.dz; g.dt = cfg.dt; g.c = cfg.dz / cfg.dt;
    for (int a = 0; a < 3; ++a) {
        if (!cfg.pos[a] || !cfg.width[a]) { err = "missing axis geometry"; return false; }
        if (cfg.periodic[a]) {
            double L = 0;
            for (int i = 0; i < g.n[a]; ++i) L += cfg.width[a][i];
            g.period[a] = L;
        }
    }
    int fa = cfg.flux_axes;
    if (fa == IPL_FLUX_REFERENCE) fa = g.dim == 2 ? IPL_FLUX_X : (IPL_FLUX_X | IPL_FLUX_Y);
    if (fa == IPL_FLUX_NONE) g.flux_mode = FLUX_MODE_NONE;
    else if (fa == IPL_FLUX_X || fa == IPL_FLUX_Y || fa == IPL_FLUX_Z) {
        g.flux_mode = FLUX_MODE_ONE_AXIS;
        g.flux_axis = fa == IPL_FLUX_X ? 0 : (fa == IPL_FLUX_Y ? 1 : 2);
        if (g.dim == 2 && g.flux_axis == 1) { err = "no y flux axis in 2D"; return false; }
    } else if (fa == (IPL_FLUX_X | IPL_FLUX_Y) && g.dim == 3) {
        g.flux_mode = FLUX_MODE_CORNER;
    } else { err = "unsupported flux axes"; return false; }
    g.npts = cfg.stencil == 3 ? 3 : (cfg.stencil == 5 ? 5 : 0);
    if (g.npts == 0) { err = "stencil must be 3 or 5"; return false; }
    // the corner directions' tensor stencil is 5 x 5 (adv.cu:100-145), and the one-axis
    // directions of the same grid share its per-axis weights
    if (g.flux_mode == FLUX_MODE_CORNER && g.npts != 5) { err = "the two-axis corner scheme needs the 5-point stencil"; return false; }
    bool flux_axis_on[3] = {false, false, false};
    if (g.flux_mode == FLUX_MODE_ONE_AXIS) flux_axis_on[g.flux_axis] = true;
    if (g.flux_mode == FLUX_MODE_CORNER) flux_axis_on[0] = flux_axis_on[1] = true;
    for (int a = 0; a < 3; ++a) {
        if (!flux_axis_on[a]) continue;
        for (int i = 0; i < g.n[a]; ++i)
            if (!(cfg.width[a][i] >= cfg.dz * (1 - 1e-12))) {
                err = "cell width below the lattice spacing on a flux axis";
                return false;
            }
        if (cfg.periodic[a] && g.n[a] < g.npts) { err = "periodic flux axis shorter than the stencil"; return false; }
    }
    g.Qs = 0;
    if (cfg.scalar != IPL_SCALAR_NONE) {
        const int Qs = cfg.scalar_lattice ? cfg.scalar_lattice : Q;
        if (!(Qs == Q || (Qs == 7 && g.dim == 3))) { err = "scalar lattice must equal the flow lattice or be 7 in 3D"; return false; }
        if (cfg.scalar == IPL_SCALAR_PHASE_FIELD && Qs != Q) { err = "the phase field uses the flow lattice"; return false; }
        if (g.flux_mode == FLUX_MODE_CORNER && Qs == 7) { err = "a D3Q7 scalar needs a single flux axis"; return false; }
        g.Qs = Qs;
    }
    int64_t N = 0;
    const size_t ncell = (size_t)cfg.nx * cfg.ny * cfg.nz;
    for (size_t q = 0; q < ncell; ++q) N += !(cfg.solid && cfg.solid[q]);
    if (N == 0) { err = "no calc cells"; return false; }
    if (N > (int64_t)0x3fffffff) { err = "too many cells for 32-bit cell indices"; return false; }
    g.N = N;
    return true;
}

bool build_graph(const ipl_config& cfg, Graph& g, std::string& err) {
    if (!graph_header(cfg, g, err)) return false;
    const int Q = g.Q;

    // calc cells in lnum order
    const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
    const size_t ncell = (size_t)nx * ny * nz;
    std::vector<int64_t> lnum(ncell, -1);
    int64_t N = 0;
    for (size_t q = 0; q < ncell; ++q)
        if (!(cfg.solid && cfg.solid[q])) lnum[q] = N++;
    g.cell_ijk.resize((size_t)3 * N);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int l = 0; l < nz; ++l) {
                const int64_t id = lnum[((size_t)i * ny + j) * nz + l];
                if (id < 0) continue;
                g.cell_ijk[3 * id] = i; g.cell_ijk[3 * id + 1] = j; g.cell_ijk[3 * id + 2] = l;
            }
    // which calc cells get a ghost, and their ghost index
    const DirTable& d = g.dirs;
    std::vector<int32_t> nb((size_t)N * Q);  // N < 2^30 (checked above)
    std::vector<int64_t> ghost_of(N, -1);
    int64_t G = 0;
    for (int64_t id = 0; id < N; ++id) {
        bool missing = false;
        for (int k = 0; k < Q; ++k) {
            int t[3] = {g.cell_ijk[3 * id] + d.cx[k], g.cell_ijk[3 * id + 1] + d.cy[k], g.cell_ijk[3 * id + 2] + d.cz[k]};
            bool ok = true;
            for (int a = 0; a < 3; ++a)
                if (t[a] < 0 || t[a] >= g.n[a]) {
                    if (cfg.periodic[a]) t[a] = (t[a] + g.n[a]) % g.n[a];
                    else ok = false;
                }
            const int32_t v = ok ? (int32_t)lnum[((size_t)t[0] * ny + t[1]) * nz + t[2]] : -1;
            nb[(size_t)id * Q + k] = v;
            if (v < 0) missing = true;
        }
        if (missing) ghost_of[id] = N + G++;
    }
    g.G = G;
    g.Ntot = N + G;
    const int64_t Ntot = g.Ntot;
    g.neib.assign((size_t)Q * Ntot, 0);
    for (int64_t id = 0; id < N; ++id)
        for (int k = 0; k < Q; ++k) {
            int64_t v = nb[(size_t)id * Q + k];
            if (v < 0) v = ghost_of[id];
            g.neib[(size_t)k * Ntot + id] = (int32_t)v;
        }
    for (int64_t id = 0; id < N; ++id) {
        const int64_t gh = ghost_of[id];
        if (gh < 0) continue;
        g.neib[gh] = (int32_t)id;  // k = 0: owner
        for (int k = 1; k < Q; ++k) g.neib[(size_t)k * Ntot + gh] = (int32_t)gh;
        for (int k = 1; k < Q; ++k)
            if (nb[(size_t)id * Q + k] < 0) g.neib[(size_t)d.opp[k] * Ntot + gh] = (int32_t)id;
    }
    // geometry (ghosts copy their owner)
    for (int a = 0; a < 3; ++a) { g.pos[a].resize(Ntot); g.del[a].resize(Ntot); }
    for (int64_t id = 0; id < Ntot; ++id) {
        const int64_t src = id < N ? id : g.neib[id];
        for (int a = 0; a < 3; ++a) {
            const int ix = g.cell_ijk[3 * src + a];
            g.pos[a][id] = cfg.pos[a][ix];
            g.del[a][id] = cfg.width[a][ix];
        }
    }
    // slip tables
    slip_tables(g, g.dirs, g.nextB, g.nextK);
    if (g.Qs) {
        const int Qs = g.Qs;
        g.sdirs = make_dirs(Qs);
        slip_tables(g, g.sdirs, g.snextB, g.snextK);
    }
    // stencil weights per flux axis and sign
    const int np = g.npts;
    g.weights.assign((size_t)3 * 2 * np * N, 0.0);
    for (int a = 0; a < 3; ++a) {
        const bool flux_axis_on = (g.flux_mode == FLUX_MODE_ONE_AXIS && g.flux_axis == a) ||
                                  (g.flux_mode == FLUX_MODE_CORNER && a < 2);
        if (!flux_axis_on) continue;
        g.axis_has_weights[a] = true;
        for (int si = 0; si < 2; ++si) {
            const int s = si == 0 ? 1 : -1;
            const int dp = d.find(a == 0 ? s : 0, a == 1 ? s : 0, a == 2 ? s : 0);
            const int dm = d.opp[dp];
            for (int64_t id = 0; id < N; ++id) {
                const int64_t cp = g.neib[(size_t)dp * Ntot + id], cm = g.neib[(size_t)dm * Ntot + id];
                double x[5], coeff[5];
                if (np == 5) {
                    const int64_t cpp = g.neib[(size_t)dp * Ntot + cp], cmm = g.neib[(size_t)dm * Ntot + cm];
                    x[0] = g.pos[a][cmm]; x[1] = g.pos[a][cm]; x[2] = g.pos[a][id]; x[3] = g.pos[a][cp]; x[4] = g.pos[a][cpp];
                } else {
                    x[0] = g.pos[a][cm]; x[1] = g.pos[a][id]; x[2] = g.pos[a][cp];
                }
                fix_positions(g, cfg, a, x, np, s);
                const double xa = g.pos[a][id] + s * (g.del[a][id] - g.dz) / 2.0;
                lagrange_weights(coeff, x, xa, np);
                for (int j = 0; j < np; ++j) g.weights[((size_t)(a * 2 + si) * np + j) * N + id] = coeff[j];
            }
        }
    }
    return true;
}

bool build_box_graph(const ipl_config& cfg, Graph& g, std::string& err) {
    if (cfg.solid) { err = "a cell mask needs the full graph"; return false; }
    if (!graph_header(cfg, g, err)) return false;
    const int n[3] = {cfg.nx, cfg.ny, cfg.nz};
    g.cell_ijk.resize((size_t)3 * g.N);
    int64_t id = 0;
    for (int i = 0; i < n[0]; ++i)
        for (int j = 0; j < n[1]; ++j)
            for (int l = 0; l < n[2]; ++l, ++id) {
                g.cell_ijk[3 * id] = i; g.cell_ijk[3 * id + 1] = j; g.cell_ijk[3 * id + 2] = l;
            }
    // a cell has a ghost when a link leaves the box across a wall: along an axis the lattice has
    // components on, that is not periodic, at index 0 or n-1 (every lattice here holds the unit
    // vectors of its axes, and every link with a component along such an axis crosses the wall)
    int64_t interior = 1;
    for (int a = 0; a < 3; ++a) {
        const bool lattice_axis = !(g.dim == 2 && a == 1);
        interior *= (!lattice_axis || cfg.periodic[a]) ? n[a] : std::max(0, n[a] - 2);
    }
    g.G = g.N - interior;
    g.Ntot = g.N + g.G;
    return true;
}

bool axis_line_weights(const ipl_config& cfg, int axis, std::vector<double>& w, std::string& err) {
    // one line along the axis: the other axes reduced to their first cell
    ipl_config c = cfg;
    c.solid = nullptr;
    int* dims[3] = {&c.nx, &c.ny, &c.nz};
    for (int a = 0; a < 3; ++a)
        if (a != axis) *dims[a] = 1;
    Graph g;
    if (!build_graph(c, g, err)) return false;
    if (!g.axis_has_weights[axis]) { err = "no flux along the axis"; return false; }
    const int np = g.npts, na = g.n[axis];
    w.assign((size_t)2 * np * na, 0.0);
    for (int si = 0; si < 2; ++si)
        for (int j = 0; j < np; ++j)
            for (int i = 0; i < na; ++i)  // cell i of the line has canonical index i
                w[((size_t)si * np + j) * na + i] = g.weights[((size_t)(axis * 2 + si) * np + j) * g.N + i];
    return true;
}

}  // namespace ipl
