// fused.hip - the fused path: one kernel per step does what the reference spreads over
// equ_f, Force, col_f_MRT, IP_process (ten launches), col_g_reg / col_PF, IP_process(g),
// update_scalar and update_rho (LockExchange.cu:220-278).
//
// Layout.  Cells are grouped in "lines" along the stream axis s (the flux axis, or z when no
// axis is stretched); line L holds cells L*ns .. L*ns+ns-1 and every population array is
// [k][cell].  One workgroup owns one whole line, so the slab reconstruction along s
// (adv.cu:76-99) needs no halo: the collided populations of the line are exchanged through
// LDS, the slab values F are formed with per-s-index Lagrange weights (FP64-precomputed on the
// host), and the flux-updated value of cell n,
//     U[n,k] = P[n,k] + (F[n - sigma_k s, k] r_n - F[n,k] r_n),   r_n = dz / width_s(n),
// is exactly the reference's set_tmp value T[m,k] of m = n - sigma_k s (adv.cu:241), which
// propagation then pulls from m + c_k (propagation.cu:9).  The kernel pushes it there instead:
// the destination lies on a neighbouring line at the same s index, so the 64 lanes of a
// wavefront store 64 contiguous values.
//
// Boundaries (box domains).  A push whose destination leaves the box is the reference's ghost
// redirect resolved at the sender: bounce-back writes it to the source cell m with the opposite
// direction (set_wall_propagation bc=1, set_bound.cu:515-518); slip reflects the destination
// and the direction across the walls it crossed, which is what the nextB/nextK tables of
// set_bound2D/3D encode for a box.  At the downstream wall of the stream axis the reference
// reads the ghost's mirrored copy (set_wall_f_inv, adv.cu:188-232); that value
//     T_mir[n,k] = P[n,mk] + (F[n,k] r_n - F[n,mk] r_n),   mk = k mirrored along s,
// is pushed by the wall cell itself.  Stencil points beyond a wall take the wall cell's value
// (the ghost copies its owner, adv.cu:3-13) with positions extrapolated by dz, which the host
// folds into the weights.
//
// The macroscopic fields of step n are recomputed from f^n, g^n and the force of step n-1,
// which is all update_rho needs (collision.cu:238-257); a small pass before the step writes
// rho (and phi) for the gradients of Force and col_PF when they are not uniform.
#include "hip_compat.hpp"

#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "engine.hpp"
#include "lattice.hpp"

namespace ipl {

namespace {

struct FGeom {
    int n[3];
    int periodic[3];
    int s, o1, o2;
    int ns;
    int64_t lines, ncell;
    int wall_bc;
    int flux;        // 1 = interpolated flux along s
    int need_grad;   // gradients of rho / phi are needed (Force or col_PF)
    int scalar;
    float inv_scale; // FP16 storage: stored value = f / w_k * scale
    float scale;
};

template <typename MT>
struct FTables {
    const MT* W;       // [sign(0:+,1:-)][j][ns]
    const MT* rdel;    // [ns]
    const MT* pos[3];  // per axis index
    const MT* wid[3];
    MT period[3];      // sum of widths along periodic axes
};

template <typename ST, typename MT>
struct FBufs {
    const ST* f_src; ST* f_dst;
    const ST* g_src; ST* g_dst;
    const MT* F_old; MT* F_new;   // [3][ncell]
    const MT* rho_n;              // [ncell] rho of step n (gradient pass), or null
    const MT* phi_n;              // [ncell] phi of step n, or null
    MT *phiold, *uold, *vold, *wold;
};

// ---- storage conversion ------------------------------------------------------------------
template <typename ST, typename MT> struct Store;
template <typename MT> struct Store<double, MT> {
    __device__ static MT load(const double* p, double, float) { return (MT)*p; }
    __device__ static void store(double* p, MT v, double, float) { *p = (double)v; }
};
template <typename MT> struct Store<float, MT> {
    __device__ static MT load(const float* p, double, float) { return (MT)*p; }
    __device__ static void store(float* p, MT v, double, float) { *p = (float)v; }
};
// FP16: populations are already deviations from the rest state (sum f = p ~ 0, collision.cu:11);
// they are stored divided by their lattice weight and scaled by a power of two.
template <typename MT> struct Store<__half, MT> {
    __device__ static MT load(const __half* p, double w, float inv_scale) {
        return (MT)(__half2float(*p) * ((float)w * inv_scale));
    }
    __device__ static void store(__half* p, MT v, double w, float scale) {
        *p = __float2half_rn((float)v * (scale / (float)w));
    }
};

__device__ inline int wrap_or_clamp(int x, int n, int periodic) {
    if (x < 0) return periodic ? x + n : 0;
    if (x >= n) return periodic ? x - n : n - 1;
    return x;
}

__device__ inline int64_t cell_index(const FGeom& G, const int c[3]) {
    return ((int64_t)c[G.o1] * G.n[G.o2] + c[G.o2]) * G.ns + c[G.s];
}

template <int Q> __device__ inline int mirror_dir(int k, int axis) {
    return axis == 0 ? mirror<Q>(k, 0) : (axis == 1 ? mirror<Q>(k, 1) : mirror<Q>(k, 2));
}

template <typename ST, typename MT, int Q>
__device__ inline void push_value(const FGeom& G, ST* dst, int k, const int m[3], MT value) {
    const Dir ck = Lattice<Q>::c(k);
    const int cv[3] = {ck.x, ck.y, ck.z};
    int d[3];
    int out = 0;  // bit mask of axes left
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int x = m[a] + cv[a];
        if (x < 0 || x >= G.n[a]) {
            if (G.periodic[a]) x = x < 0 ? x + G.n[a] : x - G.n[a];
            else out |= 1 << a;
        }
        d[a] = x;
    }
    if (!out) {
        Store<ST, MT>::store(dst + (int64_t)k * G.ncell + cell_index(G, d), value, Lattice<Q>::w(k), G.scale);
        return;
    }
    if (G.wall_bc == IPL_BC_BOUNCE_BACK) {
        const int kb = opposite<Q>(k);
        Store<ST, MT>::store(dst + (int64_t)kb * G.ncell + cell_index(G, m), value, Lattice<Q>::w(kb), G.scale);
        return;
    }
    // slip: reflect the destination and the direction across every wall crossed
    int r[3] = {cv[0], cv[1], cv[2]};
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (out & (1 << a)) { d[a] = d[a] < 0 ? 0 : G.n[a] - 1; r[a] = -r[a]; }
    const int kr = find_dir<Q>(r[0], r[1], r[2]);
    Store<ST, MT>::store(dst + (int64_t)kr * G.ncell + cell_index(G, d), value, Lattice<Q>::w(kr), G.scale);
}

// Slab flux, flux update and push of one population set (flow or scalar) of a line.
// P: collided populations of this thread's cell.  lds: [Q][ns] scratch of the workgroup.
template <typename ST, typename MT, int Q, int NPTS>
__device__ inline void transport(const FGeom& G, const FTables<MT>& tb, ST* dst, const MT (&P)[Q], MT* lds,
                                 int i, bool active, const int cc[3]) {
    const int s = G.s, ns = G.ns;
    if (!G.flux) {
        if (active) {
#pragma unroll
            for (int k = 0; k < Q; ++k) push_value<ST, MT, Q>(G, dst, k, cc, P[k]);
        }
        return;
    }
    // 1. collided values of the s-directions into LDS
    if (active) {
#pragma unroll
        for (int k = 0; k < Q; ++k)
            if (comp<Q>(k, s) != 0) lds[k * ns + i] = P[k];
    }
    __syncthreads();
    // 2. slab values F (adv.cu:76-87); stencil points beyond a wall repeat the wall cell
    MT Fv[Q];
    const int per = G.periodic[s];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        Fv[k] = 0;
        const int sg = comp<Q>(k, s);
        if (sg == 0 || !active) continue;
        const MT* Wk = tb.W + (sg > 0 ? 0 : NPTS) * ns;
        MT acc = 0;
#pragma unroll
        for (int j = 0; j < NPTS; ++j) {
            const int off = (j - NPTS / 2) * sg;
            const int ii = wrap_or_clamp(i + off, ns, per);
            const MT term = Wk[j * ns + i] * lds[k * ns + ii];
            acc = j == 0 ? term : acc + term;
        }
        Fv[k] = acc;
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int k = 0; k < Q; ++k)
            if (comp<Q>(k, s) != 0) lds[k * ns + i] = Fv[k];
    }
    __syncthreads();
    if (!active) return;
    const MT rn = tb.rdel[i];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int sg = comp<Q>(k, s);
        if (sg == 0) {  // no component along s: set_tmp returns P exactly (adv.cu:64-65,241)
            push_value<ST, MT, Q>(G, dst, k, cc, P[k]);
            continue;
        }
        // U of this cell, transported value of m = n - sg s
        const int im = i - sg;
        if (im >= 0 && im < ns) {
            const MT Um = P[k] + (lds[k * ns + im] * rn - Fv[k] * rn);
            int m[3] = {cc[0], cc[1], cc[2]};
            m[s] = im;
            push_value<ST, MT, Q>(G, dst, k, m, Um);
        } else if (per) {
            const int imw = im < 0 ? im + ns : im - ns;
            const MT Um = P[k] + (lds[k * ns + imw] * rn - Fv[k] * rn);
            int m[3] = {cc[0], cc[1], cc[2]};
            m[s] = imw;
            push_value<ST, MT, Q>(G, dst, k, m, Um);
        }
        // downstream wall: the ghost's mirrored copy (adv.cu:196-199 along s)
        const int ip = i + sg;
        if (!per && (ip < 0 || ip >= ns)) {
            const int mk = mirror_dir<Q>(k, s);
            const MT Tm = P[mk] + (Fv[k] * rn - Fv[mk] * rn);
            push_value<ST, MT, Q>(G, dst, k, cc, Tm);
        }
    }
}

// MRT / SRT collision in moment space (collision.cu:115-165)
template <typename MT, int Q>
__device__ inline void collide_cell(const MT (&f)[Q], const MT (&feq)[Q], const MT (&fk)[Q], MT tau,
                                    const double* rates, int collision, MT (&P)[Q]) {
    MT d[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) d[k] = f[k] - feq[k] + fk[k] / MT(2.0);
    if (collision == IPL_COLL_SRT) {
#pragma unroll
        for (int k = 0; k < Q; ++k) P[k] = f[k] - d[k] / tau + fk[k];
        return;
    }
    const MT itau = MT(1) / tau;
    MT m[Q];
#pragma unroll
    for (int r = 0; r < Q; ++r) {
        MT acc = 0;
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const double mk = moment<Q>(r, k);
            if (mk != 0.0) acc += MT(mk) * d[k];
        }
        m[r] = acc * (is_stress<Q>(r) ? itau : MT(rates[r]));
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        MT acc = 0;
#pragma unroll
        for (int r = 0; r < Q; ++r) {
            const double mi = moment_inv<Q>(k, r);
            if (mi != 0.0) acc += MT(mi) * m[r];
        }
        P[k] = f[k] - acc + fk[k];
    }
}

struct FParams {
    double dz, dt, c;
    double rates[27];
    double rho_L, rho_H, beta_s, gravity, sigma, pf_thickness, pf_tau;
    double tau, taus;
    int boussinesq, collision;
};

// neighbour of cell cc along axis a at offset +-1: the wall cell itself beyond a wall (the
// ghost copies its owner, basic.cu:221-225), wrapped on a periodic axis
__device__ inline int64_t neighbour(const FGeom& G, const int cc[3], int a, int off) {
    int c[3] = {cc[0], cc[1], cc[2]};
    c[a] = wrap_or_clamp(cc[a] + off, G.n[a], G.periodic[a]);
    return cell_index(G, c);
}

// (q[+] - q[-]) / (pos[+] - pos[-] + dz/1000) along axis a (Force, collision.cu:66-69), the
// period added across a periodic seam
template <typename MT>
__device__ inline MT grad_axis(const FGeom& G, const FTables<MT>& tb, const MT* q, const int cc[3], int a, MT eps) {
    const int ip = wrap_or_clamp(cc[a] + 1, G.n[a], G.periodic[a]);
    const int im = wrap_or_clamp(cc[a] - 1, G.n[a], G.periodic[a]);
    MT d = tb.pos[a][ip] - tb.pos[a][im];
    if (G.periodic[a] && d <= MT(0)) d += tb.period[a];
    return (q[neighbour(G, cc, a, 1)] - q[neighbour(G, cc, a, -1)]) / (d + eps);
}

template <typename ST, typename MT, int Q, int QS, int NPTS>
__global__ void __launch_bounds__(1024) k_fused_step(FGeom G, FTables<MT> tb, FBufs<ST, MT> b, FParams P) {
    IPL_DYNAMIC_LDS(MT, lds_f);
    MT* lds_g = lds_f + Q * G.ns;
    const int i = threadIdx.x;
    const bool active = i < G.ns;
    const int64_t L = blockIdx.x;
    int cc[3];
    cc[G.s] = active ? i : 0;
    cc[G.o1] = (int)(L / G.n[G.o2]);
    cc[G.o2] = (int)(L % G.n[G.o2]);
    const int64_t n = L * G.ns + (active ? i : 0);
    const int64_t NC = G.ncell;
    const MT c = (MT)P.c, c2 = (MT)(P.c * P.c), dt = (MT)P.dt;

    MT f[Q], g[QS > 0 ? QS : 1];
    MT Pf[Q], Pg[QS > 0 ? QS : 1];
    if (active) {
#pragma unroll
        for (int k = 0; k < Q; ++k) f[k] = Store<ST, MT>::load(b.f_src + k * NC + n, Lattice<Q>::w(k), G.inv_scale);
        if constexpr (QS > 0) {
#pragma unroll
            for (int k = 0; k < QS; ++k) g[k] = Store<ST, MT>::load(b.g_src + k * NC + n, Lattice<QS>::w(k), G.inv_scale);
        }
        // ---- macroscopic fields of step n (update_rho, collision.cu:244-254) ----
        MT p = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const Dir d = Lattice<Q>::c(k);
            p += f[k];
            jx += MT(d.x) * c * f[k]; jy += MT(d.y) * c * f[k]; jz += MT(d.z) * c * f[k];
        }
        MT sc = 0;
        if constexpr (QS > 0) {
#pragma unroll
            for (int k = 0; k < QS; ++k) sc += g[k];
        }
        const MT S = G.scalar == IPL_SCALAR_SALINITY ? sc : MT(0);
        const MT phi = G.scalar == IPL_SCALAR_PHASE_FIELD ? sc : MT(0);
        const MT rho = MT(P.rho_L) + (MT(P.rho_H) - MT(P.rho_L)) * phi + S * MT(P.beta_s);
        const MT u = jx + b.F_old[n] * dt / (MT(2.0) * rho);
        const MT v = jy + b.F_old[NC + n] * dt / (MT(2.0) * rho);
        const MT w = jz + b.F_old[2 * NC + n] * dt / (MT(2.0) * rho);
        // ---- equ_f (collision.cu:3-14) ----
        const MT u2 = u * u + v * v + w * w;
        MT feq[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const Dir d = Lattice<Q>::c(k);
            const MT t = (MT(d.x) * c * u + MT(d.y) * c * v + MT(d.z) * c * w) / c2;
            feq[k] = MT(Lattice<Q>::w(k)) * (p + MT(3.0) * t + MT(4.5) * t * t - MT(1.5) * u2 / c2);
        }
        // ---- Force (collision.cu:57-112) ----
        constexpr bool d3 = Lattice<Q>::dim == 3;
        MT Fx = 0, Fy = 0, Fz = 0;
        MT gpx = 0, gpy = 0, gpz = 0;  // grad phi (col_PF reuses it)
        const MT eps = MT(P.dz) / MT(1000);
        if (G.need_grad) {
            gpx = grad_axis<MT>(G, tb, b.phi_n, cc, 0, eps);
            if (d3) gpy = grad_axis<MT>(G, tb, b.phi_n, cc, 1, eps);
            gpz = grad_axis<MT>(G, tb, b.phi_n, cc, 2, eps);
            // x term over delX^2, the term of directions 2/4 (z in 2D, y in 3D) over dz^2
            // (collision.cu:70-71)
            const int a2 = d3 ? 1 : 2;
            const MT wx = tb.wid[0][cc[0]], dzz = MT(P.dz);
            const MT lap = (b.phi_n[neighbour(G, cc, 0, 1)] - MT(2) * phi + b.phi_n[neighbour(G, cc, 0, -1)]) / (wx * wx) +
                           (b.phi_n[neighbour(G, cc, a2, 1)] - MT(2) * phi + b.phi_n[neighbour(G, cc, a2, -1)]) / (dzz * dzz);
            const MT beta = MT(12.0) * MT(P.sigma) / MT(P.pf_thickness);
            const MT kap = MT(3.0) * MT(P.sigma) * MT(P.pf_thickness) / MT(2.0);
            const MT st = MT(4) * beta * phi * (phi - MT(1.0)) * (phi - MT(0.5)) - kap * lap;
            Fx += st * gpx; Fy += st * gpy; Fz += st * gpz;
        }
        if (P.boussinesq) Fz += -S * MT(P.beta_s) * MT(P.gravity);
        else Fz += -rho * MT(P.gravity);
        if (G.need_grad) {
            const MT grx = grad_axis<MT>(G, tb, b.rho_n, cc, 0, eps);
            const MT gry = d3 ? grad_axis<MT>(G, tb, b.rho_n, cc, 1, eps) : MT(0);
            const MT grz = grad_axis<MT>(G, tb, b.rho_n, cc, 2, eps);
            Fx += -p * c2 / MT(3.0) * grx;
            Fy += -p * c2 / MT(3.0) * gry;
            Fz += -p * c2 / MT(3.0) * grz;
        }
        b.F_new[n] = Fx; b.F_new[NC + n] = Fy; b.F_new[2 * NC + n] = Fz;
        MT fk[Q];
        const MT den = rho * c2 / MT(3.0);
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const Dir d = Lattice<Q>::c(k);
            fk[k] = MT(Lattice<Q>::w(k)) * dt * (MT(d.x) * c * Fx + MT(d.y) * c * Fy + MT(d.z) * c * Fz) / den;
        }
        // ---- collision ----
        collide_cell<MT, Q>(f, feq, fk, MT(P.tau), P.rates, P.collision, Pf);
        // ---- scalar collision ----
        if constexpr (QS > 0) {
            if (G.scalar == IPL_SCALAR_SALINITY) {  // col_g_reg (collision.cu:167-188)
                constexpr bool linear = QS == 7;
                const MT cs2inv = linear ? MT(4.0) : MT(3.0);
                MT geq[QS];
                MT gx = 0, gy = 0, gz = 0;
#pragma unroll
                for (int k = 0; k < QS; ++k) {
                    const Dir d = Lattice<QS>::c(k);
                    const MT t = (MT(d.x) * c * u + MT(d.y) * c * v + MT(d.z) * c * w) / c2;
                    MT e = MT(1.0) + cs2inv * t;
                    if (!linear) e += MT(4.5) * t * t - MT(1.5) * u2 / c2;
                    geq[k] = MT(Lattice<QS>::w(k)) * S * e;
                    const MT gn = g[k] - geq[k];
                    gx += gn * MT(d.x) * c; gy += gn * MT(d.y) * c; gz += gn * MT(d.z) * c;
                }
                const MT om = MT(1.0) - MT(1.0) / MT(P.taus);
#pragma unroll
                for (int k = 0; k < QS; ++k) {
                    const Dir d = Lattice<QS>::c(k);
                    const MT gneq = MT(Lattice<QS>::w(k)) * cs2inv * (gx * MT(d.x) * c + gy * MT(d.y) * c + gz * MT(d.z) * c) / c2;
                    Pg[k] = geq[k] + om * gneq;
                }
            } else {  // col_PF (collision.cu:190-223)
                const MT nabu = (phi * u - b.phiold[n] * b.uold[n]) / dt;
                const MT nabv = (phi * v - b.phiold[n] * b.vold[n]) / dt;
                const MT nabw = (phi * w - b.phiold[n] * b.wold[n]) / dt;
                b.phiold[n] = phi; b.uold[n] = u; b.vold[n] = v; b.wold[n] = w;
                const MT nrm = sqrt(gpx * gpx + gpy * gpy + gpz * gpz + MT(1e-8));
                const MT src = (MT(1.0) - MT(4.0) * (phi - MT(0.5)) * (phi - MT(0.5))) / MT(P.pf_thickness);
                const MT tp = MT(P.pf_tau);
#pragma unroll
                for (int k = 0; k < QS; ++k) {
                    const Dir d = Lattice<QS>::c(k);
                    const MT cx = MT(d.x) * c, cy = MT(d.y) * c, cz = MT(d.z) * c;
                    const MT t = (cx * u + cy * v + cz * w) / c2;
                    const MT geq = MT(Lattice<QS>::w(k)) * phi * (MT(1.0) + MT(3.0) * t + MT(4.5) * t * t - MT(1.5) * u2 / c2);
                    const MT Fkp = dt * MT(Lattice<QS>::w(k)) *
                                   (src * (cx * gpx + cy * gpy + cz * gpz) / nrm + MT(3) * (cx * nabu + cy * nabv + cz * nabw) / c2);
                    Pg[k] = g[k] - (g[k] - geq - Fkp * (tp - MT(0.5))) / tp;
                }
            }
        }
    }
    // ---- interpolated propagation of f and g ----
    transport<ST, MT, Q, NPTS>(G, tb, b.f_dst, Pf, lds_f, i, active, cc);
    if constexpr (QS > 0) transport<ST, MT, QS, NPTS>(G, tb, b.g_dst, Pg, lds_g, i, active, cc);
}

// rho (and phi) of step n for the gradients of the next fused step
template <typename ST, typename MT, int QS>
__global__ void k_scalar_fields(FGeom G, const ST* g_src, MT* rho_n, MT* phi_n, FParams P) {
    const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= G.ncell) return;
    MT sc = 0;
    if constexpr (QS > 0) {
#pragma unroll
        for (int k = 0; k < QS; ++k) sc += Store<ST, MT>::load(g_src + k * G.ncell + n, Lattice<QS>::w(k), G.inv_scale);
    }
    const MT S = G.scalar == IPL_SCALAR_SALINITY ? sc : MT(0);
    const MT phi = G.scalar == IPL_SCALAR_PHASE_FIELD ? sc : MT(0);
    rho_n[n] = MT(P.rho_L) + (MT(P.rho_H) - MT(P.rho_L)) * phi + S * MT(P.beta_s);
    phi_n[n] = phi;
}

// readout: p, u, v, w, rho, S, phi of the current state into [7][ncell] doubles
template <typename ST, typename MT, int Q, int QS>
__global__ void k_readout(FGeom G, const ST* f_src, const ST* g_src, const MT* F_old, double* out, FParams P) {
    const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= G.ncell) return;
    const int64_t NC = G.ncell;
    const MT c = (MT)P.c;
    MT p = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const Dir d = Lattice<Q>::c(k);
        const MT fv = Store<ST, MT>::load(f_src + k * NC + n, Lattice<Q>::w(k), G.inv_scale);
        p += fv; jx += MT(d.x) * c * fv; jy += MT(d.y) * c * fv; jz += MT(d.z) * c * fv;
    }
    MT sc = 0;
    if constexpr (QS > 0) {
#pragma unroll
        for (int k = 0; k < QS; ++k) sc += Store<ST, MT>::load(g_src + k * NC + n, Lattice<QS>::w(k), G.inv_scale);
    }
    const MT S = G.scalar == IPL_SCALAR_SALINITY ? sc : MT(0);
    const MT phi = G.scalar == IPL_SCALAR_PHASE_FIELD ? sc : MT(0);
    const MT rho = MT(P.rho_L) + (MT(P.rho_H) - MT(P.rho_L)) * phi + S * MT(P.beta_s);
    out[0 * NC + n] = p;
    out[1 * NC + n] = jx + F_old[n] * MT(P.dt) / (MT(2.0) * rho);
    out[2 * NC + n] = jy + F_old[NC + n] * MT(P.dt) / (MT(2.0) * rho);
    out[3 * NC + n] = jz + F_old[2 * NC + n] * MT(P.dt) / (MT(2.0) * rho);
    out[4 * NC + n] = rho;
    out[5 * NC + n] = S;
    out[6 * NC + n] = phi;
}

template <typename ST, typename MT>
__global__ void k_convert_pops(FGeom G, const ST* src, double* dst, int q, int Q) {
    const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= G.ncell) return;
    double w = 0;
    if (Q == 9) w = Lattice<9>::w(q);
    else if (Q == 19) w = Lattice<19>::w(q);
    else if (Q == 27) w = Lattice<27>::w(q);
    else w = Lattice<7>::w(q);
    dst[n] = (double)Store<ST, MT>::load(src + (int64_t)q * G.ncell + n, w, G.inv_scale);
}

template <typename ST, typename MT>
__global__ void k_store_pops(FGeom G, const double* src, ST* dst, int q, int Q) {
    const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= G.ncell) return;
    double w = 0;
    if (Q == 9) w = Lattice<9>::w(q);
    else if (Q == 19) w = Lattice<19>::w(q);
    else if (Q == 27) w = Lattice<27>::w(q);
    else w = Lattice<7>::w(q);
    Store<ST, MT>::store(dst + (int64_t)q * G.ncell + n, (MT)src[n], w, G.scale);
}

template <typename MT>
__global__ void k_wsum(const double* fields, int nf, int64_t n, const double* vol, double* partial) {
    __shared__ double sh[3][256];
    double acc[3] = {0, 0, 0};
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += 1024ll * 256)
        for (int q = 0; q < nf; ++q) acc[q] += fields[q * n + i] * vol[i];
    for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) partial[q * 1024 + blockIdx.x] = sh[q][0];
}

__global__ void k_wsum_final(const double* partial, double* out) {
    __shared__ double sh[256];
    for (int q = 0; q < 3; ++q) {
        double acc = 0;
        for (int i = threadIdx.x; i < 1024; i += 256) acc += partial[q * 1024 + i];
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = sh[0];
        __syncthreads();
    }
}

template <typename ST, typename MT>
class FusedEngine final : public Engine {
public:
    FusedEngine(const Graph& g, const ipl_config& cfg) : gr_(g), cfg_(cfg) {}

    Status init(const HostState& hs, hipStream_t st) {
        stream = st;
        FGeom& G = geo_;
        std::memset(&G, 0, sizeof(G));
        for (int a = 0; a < 3; ++a) { G.n[a] = gr_.n[a]; G.periodic[a] = cfg_.periodic[a]; }
        G.flux = gr_.flux_mode == FLUX_MODE_ONE_AXIS;
        G.s = G.flux ? gr_.flux_axis : 2;
        G.o1 = G.s == 0 ? 1 : 0;
        G.o2 = G.s == 2 ? 1 : 2;
        G.ns = G.n[G.s];
        G.lines = (int64_t)G.n[G.o1] * G.n[G.o2];
        G.ncell = G.lines * G.ns;
        G.wall_bc = cfg_.wall_bc;
        G.scalar = cfg_.scalar;
        const bool rho_varies = (cfg_.scalar == IPL_SCALAR_SALINITY && cfg_.beta_s != 0.0) ||
                                (cfg_.scalar == IPL_SCALAR_PHASE_FIELD && cfg_.rho_H != cfg_.rho_L);
        G.need_grad = rho_varies || cfg_.sigma != 0.0 || cfg_.scalar == IPL_SCALAR_PHASE_FIELD;
        double scale = 1.0;
        if (std::is_same<ST, __half>::value) scale = cfg_.f16_scale > 0 ? cfg_.f16_scale : 8192.0;
        G.scale = (float)scale;
        G.inv_scale = (float)(1.0 / scale);
        P_ = FParams{};
        P_.dz = gr_.dz; P_.dt = gr_.dt; P_.c = gr_.c;
        for (int i = 0; i < 27; ++i) P_.rates[i] = cfg_.s_rates[i];
        P_.rho_L = cfg_.rho_L; P_.rho_H = cfg_.rho_H; P_.beta_s = cfg_.beta_s; P_.gravity = cfg_.gravity;
        P_.sigma = cfg_.sigma; P_.pf_thickness = cfg_.pf_thickness; P_.pf_tau = cfg_.pf_tau;
        P_.tau = hs.tau[0]; P_.taus = hs.taus[0];
        P_.boussinesq = cfg_.boussinesq; P_.collision = cfg_.collision;
        const int Q = gr_.Q, Qs = gr_.Qs;
        const int64_t NC = G.ncell, N = gr_.N;
        // internal <-> canonical permutation
        perm_.resize(NC);
        for (int64_t id = 0; id < N; ++id) {
            int c[3] = {gr_.cell_ijk[3 * id], gr_.cell_ijk[3 * id + 1], gr_.cell_ijk[3 * id + 2]};
            perm_[id] = ((int64_t)c[G.o1] * G.n[G.o2] + c[G.o2]) * G.ns + c[G.s];
        }
        Status r;
        // populations
        if (!(r = mem_.alloc_n(&fa_, (size_t)Q * NC))) return r;
        if (!(r = mem_.alloc_n(&fb_, (size_t)Q * NC))) return r;
        if (Qs) {
            if (!(r = mem_.alloc_n(&ga_, (size_t)Qs * NC))) return r;
            if (!(r = mem_.alloc_n(&gb_, (size_t)Qs * NC))) return r;
        }
        if (!(r = mem_.alloc_n(&Fa_, (size_t)3 * NC))) return r;
        if (!(r = mem_.alloc_n(&Fb_, (size_t)3 * NC))) return r;
        IPL_HIP_TRY(hipMemsetAsync(Fa_, 0, 3 * NC * sizeof(MT), st));
        IPL_HIP_TRY(hipMemsetAsync(Fb_, 0, 3 * NC * sizeof(MT), st));
        if (G.need_grad) {
            if (!(r = mem_.alloc_n(&rho_n_, (size_t)NC))) return r;
            if (!(r = mem_.alloc_n(&phi_n_, (size_t)NC))) return r;
        }
        if (cfg_.scalar == IPL_SCALAR_PHASE_FIELD) {
            MT** pf[4] = {&phiold_, &uold_, &vold_, &wold_};
            const std::vector<double>* src[4] = {&hs.phi, &hs.u, &hs.v, &hs.w};
            for (int q = 0; q < 4; ++q) {
                if (!(r = mem_.alloc_n(pf[q], (size_t)NC))) return r;
                std::vector<MT> t(NC);
                for (int64_t id = 0; id < N; ++id) t[perm_[id]] = (MT)(*src[q])[id];
                IPL_HIP_TRY(hipMemcpyAsync(*pf[q], t.data(), NC * sizeof(MT), hipMemcpyHostToDevice, st));
                IPL_HIP_TRY(hipStreamSynchronize(st));
            }
        }
        if (!(r = upload_pops(hs.f, Q, gr_.Ntot, fa_))) return r;
        if (Qs && !(r = upload_pops(hs.g, Qs, gr_.Ntot, ga_))) return r;
        // tables
        const int np = gr_.npts;
        std::vector<MT> W((size_t)2 * np * G.ns, MT(0)), rd(G.ns, MT(1));
        if (G.flux) {
            // weights of line 0: cells (o1=0, o2=0, s=i) in canonical order
            for (int i = 0; i < G.ns; ++i) {
                int c[3] = {0, 0, 0};
                c[G.s] = i;
                const int64_t id = ((int64_t)c[0] * gr_.n[1] + c[1]) * gr_.n[2] + c[2];
                for (int si = 0; si < 2; ++si)
                    for (int j = 0; j < np; ++j)
                        W[((size_t)si * np + j) * G.ns + i] = (MT)gr_.weights[((size_t)(G.s * 2 + si) * np + j) * N + id];
                rd[i] = (MT)(gr_.dz / cfg_widths_[G.s][i]);
            }
        }
        MT* dW; MT* drd;
        if (!(r = mem_.alloc_n(&dW, W.size()))) return r;
        if (!(r = mem_.alloc_n(&drd, rd.size()))) return r;
        IPL_HIP_TRY(hipMemcpyAsync(dW, W.data(), W.size() * sizeof(MT), hipMemcpyHostToDevice, st));
        IPL_HIP_TRY(hipMemcpyAsync(drd, rd.data(), rd.size() * sizeof(MT), hipMemcpyHostToDevice, st));
        tb_.W = dW; tb_.rdel = drd;
        std::vector<std::vector<MT>> axis_h(6);
        for (int a = 0; a < 3; ++a) {
            axis_h[a].resize(G.n[a]);
            axis_h[3 + a].resize(G.n[a]);
            for (int x = 0; x < G.n[a]; ++x) { axis_h[a][x] = (MT)cfg_pos_[a][x]; axis_h[3 + a][x] = (MT)cfg_widths_[a][x]; }
            MT *dp, *dw;
            if (!(r = mem_.alloc_n(&dp, G.n[a]))) return r;
            if (!(r = mem_.alloc_n(&dw, G.n[a]))) return r;
            IPL_HIP_TRY(hipMemcpyAsync(dp, axis_h[a].data(), G.n[a] * sizeof(MT), hipMemcpyHostToDevice, st));
            IPL_HIP_TRY(hipMemcpyAsync(dw, axis_h[3 + a].data(), G.n[a] * sizeof(MT), hipMemcpyHostToDevice, st));
            tb_.pos[a] = dp; tb_.wid[a] = dw;
            tb_.period[a] = (MT)gr_.period[a];
        }
        // per-cell volume (FP64) for the conservation sums, internal order
        std::vector<double> vol(NC);
        for (int64_t id = 0; id < N; ++id) {
            const int x = gr_.cell_ijk[3 * id], y = gr_.cell_ijk[3 * id + 1], z = gr_.cell_ijk[3 * id + 2];
            vol[perm_[id]] = cfg_widths_[0][x] * cfg_widths_[1][y] * cfg_widths_[2][z] / (gr_.dz * gr_.dz * gr_.dz);
        }
        if (!(r = mem_.alloc_n(&vol_, (size_t)NC))) return r;
        if (!(r = mem_.alloc_n(&out_, (size_t)7 * NC))) return r;
        if (!(r = mem_.alloc_n(&partial_, (size_t)3 * 1024 + 3))) return r;
        IPL_HIP_TRY(hipMemcpyAsync(vol_, vol.data(), NC * sizeof(double), hipMemcpyHostToDevice, st));
        IPL_HIP_TRY(hipStreamSynchronize(st));
        block_ = ((G.ns + 63) / 64) * 64;
        lds_bytes_ = (size_t)(Q + Qs) * G.ns * sizeof(MT);
        return Status::ok();
    }

    void set_axes(const double* const* pos, const double* const* wid) {
        for (int a = 0; a < 3; ++a) {
            cfg_pos_[a].assign(pos[a], pos[a] + gr_.n[a]);
            cfg_widths_[a].assign(wid[a], wid[a] + gr_.n[a]);
        }
    }

    Status upload_pops(const std::vector<double>& src, int q, int64_t Nt, ST* dst) {
        const int64_t NC = geo_.ncell, N = gr_.N;
        double* tmp;
        IPL_HIP_TRY(hipMalloc(&tmp, NC * sizeof(double)));
        std::vector<double> h(NC);
        for (int k = 0; k < q; ++k) {
            for (int64_t id = 0; id < N; ++id) h[perm_[id]] = src[(size_t)k * Nt + id];
            IPL_HIP_TRY(hipMemcpyAsync(tmp, h.data(), NC * sizeof(double), hipMemcpyHostToDevice, stream));
            IPL_LAUNCH((k_store_pops<ST, MT>), (NC + 255) / 256, 256, 0, stream, geo_, tmp, dst, k, q);
            IPL_HIP_TRY(hipStreamSynchronize(stream));
        }
        IPL_HIP_TRY(hipFree(tmp));
        return Status::ok();
    }

    template <int Q, int QS, int NPTS>
    Status launch_step() {
        FBufs<ST, MT> b{};
        b.f_src = fa_; b.f_dst = fb_; b.g_src = ga_; b.g_dst = gb_;
        b.F_old = Fa_; b.F_new = Fb_;
        b.rho_n = rho_n_; b.phi_n = phi_n_;
        b.phiold = phiold_; b.uold = uold_; b.vold = vold_; b.wold = wold_;
        if (geo_.need_grad)
            IPL_LAUNCH((k_scalar_fields<ST, MT, QS>), (geo_.ncell + 255) / 256, 256, 0, stream, geo_, ga_, rho_n_, phi_n_, P_);
        IPL_LAUNCH_SYNC((k_fused_step<ST, MT, Q, QS, NPTS>), (unsigned)geo_.lines, block_, lds_bytes_, stream,
                        geo_, tb_, b, P_);
        std::swap(fa_, fb_);
        std::swap(ga_, gb_);
        std::swap(Fa_, Fb_);
        return Status::ok();
    }

    template <int Q, int QS>
    Status step_q(int nsteps) {
        for (int s = 0; s < nsteps; ++s) {
            Status r = gr_.npts == 5 ? launch_step<Q, QS, 5>() : launch_step<Q, QS, 3>();
            if (!r) return r;
        }
        IPL_HIP_TRY(hipGetLastError());
        return Status::ok();
    }

    Status step(int nsteps) override {
        const int Q = gr_.Q, Qs = gr_.Qs;
#define IPL_FCASE(QQ, QSS) if (Q == QQ && Qs == QSS) return step_q<QQ, QSS>(nsteps);
        IPL_FCASE(9, 0) IPL_FCASE(9, 9) IPL_FCASE(19, 0) IPL_FCASE(19, 19) IPL_FCASE(19, 7)
        IPL_FCASE(27, 0) IPL_FCASE(27, 27) IPL_FCASE(27, 7)
#undef IPL_FCASE
        return Status::err(IPL_ERR_UNSUPPORTED, "lattice combination");
    }

    Status stage(int) override { return Status::err(IPL_ERR_UNSUPPORTED, "the fused path has no separate stages"); }

    template <int Q, int QS>
    void launch_readout() {
        IPL_LAUNCH((k_readout<ST, MT, Q, QS>), (geo_.ncell + 255) / 256, 256, 0, stream, geo_, fa_, ga_, Fa_, out_, P_);
    }

    Status readout() {
        const int Q = gr_.Q, Qs = gr_.Qs;
#define IPL_RCASE(QQ, QSS) if (Q == QQ && Qs == QSS) { launch_readout<QQ, QSS>(); IPL_HIP_TRY(hipGetLastError()); return Status::ok(); }
        IPL_RCASE(9, 0) IPL_RCASE(9, 9) IPL_RCASE(19, 0) IPL_RCASE(19, 19) IPL_RCASE(19, 7)
        IPL_RCASE(27, 0) IPL_RCASE(27, 27) IPL_RCASE(27, 7)
#undef IPL_RCASE
        return Status::err(IPL_ERR_UNSUPPORTED, "lattice combination");
    }

    Status read(int field, double* host, size_t count) override {
        const int64_t N = gr_.N, NC = geo_.ncell;
        std::vector<double> h(NC);
        if (field == IPL_FIELD_F || field == IPL_FIELD_G) {
            const int q = field == IPL_FIELD_F ? gr_.Q : gr_.Qs;
            if (!q || count != (size_t)q * N) return Status::err(IPL_ERR_ARGUMENT, "population read size mismatch");
            for (int k = 0; k < q; ++k) {
                IPL_LAUNCH((k_convert_pops<ST, MT>), (NC + 255) / 256, 256, 0, stream, geo_, field == IPL_FIELD_F ? fa_ : ga_, out_, k, q);
                IPL_HIP_TRY(hipMemcpyAsync(h.data(), out_, NC * sizeof(double), hipMemcpyDeviceToHost, stream));
                IPL_HIP_TRY(hipStreamSynchronize(stream));
                for (int64_t id = 0; id < N; ++id) host[(size_t)k * N + id] = h[perm_[id]];
            }
            return Status::ok();
        }
        if (count != (size_t)N) return Status::err(IPL_ERR_ARGUMENT, "field size mismatch");
        int slot = -1;
        switch (field) {
        case IPL_FIELD_P: slot = 0; break; case IPL_FIELD_U: slot = 1; break; case IPL_FIELD_V: slot = 2; break;
        case IPL_FIELD_W: slot = 3; break; case IPL_FIELD_RHO: slot = 4; break; case IPL_FIELD_S: slot = 5; break;
        case IPL_FIELD_PHI: slot = 6; break;
        default: break;
        }
        if (field == IPL_FIELD_FX || field == IPL_FIELD_FY || field == IPL_FIELD_FZ) {
            // the force of the last step (before resetF, collision.cu:258-264)
            std::vector<MT> t(NC);
            const int a = field - IPL_FIELD_FX;
            IPL_HIP_TRY(hipMemcpyAsync(t.data(), Fa_ + (size_t)a * NC, NC * sizeof(MT), hipMemcpyDeviceToHost, stream));
            IPL_HIP_TRY(hipStreamSynchronize(stream));
            for (int64_t id = 0; id < N; ++id) host[id] = (double)t[perm_[id]];
            return Status::ok();
        }
        if (field == IPL_FIELD_TAU || field == IPL_FIELD_TAUS) {
            for (int64_t id = 0; id < N; ++id) host[id] = field == IPL_FIELD_TAU ? P_.tau : P_.taus;
            return Status::ok();
        }
        if (slot < 0) return Status::err(IPL_ERR_ARGUMENT, "field id");
        Status r = readout();
        if (!r) return r;
        IPL_HIP_TRY(hipMemcpyAsync(h.data(), out_ + (size_t)slot * NC, NC * sizeof(double), hipMemcpyDeviceToHost, stream));
        IPL_HIP_TRY(hipStreamSynchronize(stream));
        for (int64_t id = 0; id < N; ++id) host[id] = h[perm_[id]];
        return Status::ok();
    }

    Status write(int field, const double* host, size_t count) override {
        const int64_t N = gr_.N, NC = geo_.ncell;
        if (field == IPL_FIELD_F || field == IPL_FIELD_G) {
            const int q = field == IPL_FIELD_F ? gr_.Q : gr_.Qs;
            if (!q || count != (size_t)q * N) return Status::err(IPL_ERR_ARGUMENT, "population write size mismatch");
            std::vector<double> h(NC, 0.0);
            for (int k = 0; k < q; ++k) {
                for (int64_t id = 0; id < N; ++id) h[perm_[id]] = host[(size_t)k * N + id];
                IPL_HIP_TRY(hipMemcpyAsync(out_, h.data(), NC * sizeof(double), hipMemcpyHostToDevice, stream));
                IPL_LAUNCH((k_store_pops<ST, MT>), (NC + 255) / 256, 256, 0, stream, geo_, out_, field == IPL_FIELD_F ? fa_ : ga_, k, q);
                IPL_HIP_TRY(hipStreamSynchronize(stream));
            }
            return Status::ok();
        }
        return Status::err(IPL_ERR_UNSUPPORTED, "the fused path writes populations only");
    }

    Status global_sums(double out3[3]) override {
        Status r = readout();
        if (!r) return r;
        // fields p (slot 0), S (slot 5), phi (slot 6) are not contiguous: sum each separately
        double res[3];
        const int slots[3] = {0, 5, 6};
        for (int q = 0; q < 3; ++q) {
            IPL_LAUNCH_SYNC((k_wsum<MT>), 1024, 256, 0, stream, out_ + (size_t)slots[q] * geo_.ncell, 1, geo_.ncell, vol_, partial_);
            IPL_LAUNCH_SYNC(k_wsum_final, 1, 256, 0, stream, partial_, partial_ + 3 * 1024);
            IPL_HIP_TRY(hipMemcpyAsync(&res[q], partial_ + 3 * 1024, sizeof(double), hipMemcpyDeviceToHost, stream));
            IPL_HIP_TRY(hipStreamSynchronize(stream));
        }
        for (int q = 0; q < 3; ++q) out3[q] = res[q];
        return Status::ok();
    }

    Status check_finite(bool* finite) override {
        Status r = readout();
        if (!r) return r;
        const double* f[4] = {out_, out_ + geo_.ncell, out_ + 2 * geo_.ncell, out_ + 3 * geo_.ncell};
        return device_all_finite<double>(f, 4, geo_.ncell, finite, stream);
    }

    int64_t device_bytes() const override { return (int64_t)mem_.total(); }
    int path() const override { return IPL_PATH_FUSED; }

    const FGeom& geom() const { return geo_; }
    size_t lds_bytes() const { return lds_bytes_; }

private:
    const Graph& gr_;
    ipl_config cfg_;
    std::vector<double> cfg_pos_[3], cfg_widths_[3];
    FGeom geo_;
    FParams P_;
    FTables<MT> tb_{};
    DeviceArena mem_;
    ST *fa_ = nullptr, *fb_ = nullptr, *ga_ = nullptr, *gb_ = nullptr;
    MT *Fa_ = nullptr, *Fb_ = nullptr, *rho_n_ = nullptr, *phi_n_ = nullptr;
    MT *phiold_ = nullptr, *uold_ = nullptr, *vold_ = nullptr, *wold_ = nullptr;
    double *vol_ = nullptr, *out_ = nullptr, *partial_ = nullptr;
    std::vector<int64_t> perm_;
    unsigned block_ = 64;
    size_t lds_bytes_ = 0;
};

}  // namespace

bool fused_supported(const Graph& g, const ipl_config& cfg, std::string& why) {
    if (cfg.solid) { why = "cell masks need the plain path"; return false; }
    if (g.flux_mode == FLUX_MODE_CORNER) { why = "the two-axis corner scheme runs on the plain path"; return false; }
    const int s = g.flux_mode == FLUX_MODE_ONE_AXIS ? g.flux_axis : 2;
    if (g.n[s] > 1024) { why = "stream-axis lines longer than 1024 cells"; return false; }
    if (g.flux_mode == FLUX_MODE_ONE_AXIS && !cfg.periodic[s] && g.n[s] < 3) { why = "stream axis too short"; return false; }
    const size_t mt = cfg.storage == IPL_F64 ? 8 : 4;
    const size_t lds = (size_t)(g.Q + g.Qs) * g.n[s] * mt;
    if (lds > 64 * 1024) { why = "line does not fit the LDS budget"; return false; }
    for (int a = 0; a < 3; ++a)
        if (g.n[a] > 1 && !cfg.periodic[a] && g.n[a] < 2) { why = "axis too short"; return false; }
    return true;
}

std::unique_ptr<Engine> make_fused_engine(const Graph& g, const ipl_config& cfg, const HostState& hs,
                                          hipStream_t st, Status& status) {
    for (int64_t i = 1; i < g.N; ++i)
        if (hs.tau[i] != hs.tau[0] || hs.taus[i] != hs.taus[0]) {
            status = Status::err(IPL_ERR_UNSUPPORTED, "the fused path needs uniform tau and taus");
            return nullptr;
        }
    auto make = [&](auto* tag) -> std::unique_ptr<Engine> {
        using E = std::remove_pointer_t<decltype(tag)>;
        auto e = std::make_unique<E>(g, cfg);
        e->set_axes(cfg.pos, cfg.width);
        status = e->init(hs, st);
        if (!status) return nullptr;
        return e;
    };
    if (cfg.storage == IPL_F64) return make((FusedEngine<double, double>*)nullptr);
    if (cfg.storage == IPL_F32) return make((FusedEngine<float, float>*)nullptr);
    if (cfg.storage == IPL_F16) {
        // FP16 storage is switched off: its first GPU run (D3Q19 + D3Q7 lock exchange) ended in
        // a device memory fault whose cause is not yet shown; the CPU build of this code is clean.
        status = Status::err(IPL_ERR_UNSUPPORTED, "FP16 storage is disabled pending a GPU fault analysis");
        return nullptr;
    }
    status = Status::err(IPL_ERR_ARGUMENT, "storage type");
    return nullptr;
}

}  // namespace ipl
