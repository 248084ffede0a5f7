// fused_kernels.hpp - device code of the fused path (included by the fused_q*.hip files, which
// instantiate it per lattice; the engine in fused.hip sees only fused_common.hpp).
//
// One kernel per step does what the reference spreads over equ_f, Force, col_f_MRT,
// IP_process (ten launches), col_g_reg / col_PF, IP_process(g), update_scalar and update_rho
// (LockExchange.cu:220-278).
//
// Layout.  Cells are grouped in lines along the stream axis S (the flux axis, or z when no
// axis is stretched); line L holds cells L*ns .. L*ns+ns-1 and every population array is
// [k][cell].  One workgroup owns one whole line, so the slab reconstruction along S
// (adv.cu:76-99) needs no halo: the collided populations of the line's S-directions are
// exchanged through LDS, the slab values F are formed with per-S-index Lagrange weights
// (FP64-precomputed on the host), and the flux-updated value of cell n,
//     U[n,k] = P[n,k] + (F[n - sigma_k S, k] r_n - F[n,k] r_n),   r_n = dz / width_S(n),
// is exactly the reference's set_tmp value T[m,k] of m = n - sigma_k S (adv.cu:241), which
// propagation then pulls from m + c_k (propagation.cu:9).  The kernel pushes it there instead:
// the destination lies on a neighbouring line at the same S index, so the lanes of a
// wavefront store contiguous values.
//
// Boundaries (box domains).  A push whose destination leaves the box is the reference's ghost
// redirect resolved at the sender: bounce-back writes it to the source cell m with the opposite
// direction (set_wall_propagation bc=1, set_bound.cu:515-518); slip reflects the destination
// and the direction across the walls it crossed, which is what the nextB/nextK tables of
// set_bound2D/3D encode for a box.  At the downstream wall of the stream axis the reference
// reads the ghost's mirrored copy (set_wall_f_inv, adv.cu:188-232); that value
//     T_mir[n,k] = P[n,mk] + (F[n,k] r_n - F[n,mk] r_n),   mk = k mirrored along S,
// is pushed by the wall cell itself.  Stencil points beyond a wall take the wall cell's value
// (the ghost copies its owner, adv.cu:3-13) with positions extrapolated by dz, which the host
// folds into the weights.
//
// Every direction property (components, mirror, opposite, reflections, LDS slot) is a
// compile-time constant of the direction and the stream axis, which are template arguments:
// the per-direction code is fully specialised and no table is indexed at run time.
//
// The macroscopic fields of step n are recomputed from f^n, g^n and the force of step n-1,
// which is all update_rho needs (collision.cu:238-257); a small pass before the step writes
// rho (and phi) for the gradients of Force and col_PF when they are not uniform.
#pragma once

#include <cmath>
#include <type_traits>
#include <utility>

#include "fused_common.hpp"
#include "lattice.hpp"

namespace ipl {
namespace fusedk {

constexpr int kMaxLine = 512;  // work-items per workgroup = cells per line, at most

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

// compile-time loop: f(std::integral_constant<int, I>) for I = 0 .. N-1
template <int N, int I = 0, typename F>
__device__ inline void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, I + 1>(f);
    }
}

// number of directions with a component along axis S before direction k (the LDS slot)
template <int Q, int S> constexpr int sdir_slot(int k) {
    int n = 0;
    for (int j = 0; j < k; ++j) n += comp<Q>(j, S) != 0;
    return n;
}
template <int Q, int S> constexpr int sdir_count() { return sdir_slot<Q, S>(Q); }

// direction k with the components of the axes in mask reversed
template <int Q> constexpr int reflect_dir(int k, int mask) {
    Dir c = Lattice<Q>::c(k);
    if (mask & 1) c.x = -c.x;
    if (mask & 2) c.y = -c.y;
    if (mask & 4) c.z = -c.z;
    return find_dir<Q>(c.x, c.y, c.z);
}

__device__ inline int wrap_or_clamp(int x, int n, int periodic) {
    if (x < 0) return periodic ? x + n : 0;
    if (x >= n) return periodic ? x - n : n - 1;
    return x;
}

template <int S> __device__ inline int64_t cell_index(const FGeom& G, const int c[3]) {
    constexpr int o1 = S == 0 ? 1 : 0, o2 = S == 2 ? 1 : 2;
    return ((int64_t)c[o1] * G.n[o2] + c[o2]) * G.ns + c[S];
}

// Writes the transported value of direction K whose source is cell m (destination m + c_K).
template <typename ST, typename MT, int Q, int K, int S>
__device__ inline void push(const FGeom& G, ST* dst, const int m[3], MT value) {
    constexpr Dir ck = Lattice<Q>::c(K);
    const int cv[3] = {ck.x, ck.y, ck.z};
    int d[3];
    int out = 0;  // mask of the axes the destination leaves
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int x = m[a] + cv[a];
        if (cv[a] != 0 && (x < 0 || x >= G.n[a])) {
            if (G.periodic[a]) x = x < 0 ? x + G.n[a] : x - G.n[a];
            else out |= 1 << a;
        }
        d[a] = x;
    }
    if (!out) {
        Store<ST, MT>::store(dst + (int64_t)K * G.ncell + cell_index<S>(G, d), value, Lattice<Q>::w(K), G.scale);
        return;
    }
    if (G.wall_bc == IPL_BC_BOUNCE_BACK) {
        constexpr int kb = opposite<Q>(K);
        Store<ST, MT>::store(dst + (int64_t)kb * G.ncell + cell_index<S>(G, m), value, Lattice<Q>::w(kb), G.scale);
        return;
    }
    // slip: reflect the destination and the direction across every wall crossed
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (out & (1 << a)) d[a] = d[a] < 0 ? 0 : G.n[a] - 1;
    constexpr int r1 = reflect_dir<Q>(K, 1), r2 = reflect_dir<Q>(K, 2), r3 = reflect_dir<Q>(K, 3);
    constexpr int r4 = reflect_dir<Q>(K, 4), r5 = reflect_dir<Q>(K, 5), r6 = reflect_dir<Q>(K, 6);
    constexpr int r7 = reflect_dir<Q>(K, 7);
    const int kr = out == 1 ? r1 : out == 2 ? r2 : out == 3 ? r3 : out == 4 ? r4 : out == 5 ? r5 : out == 6 ? r6 : r7;
    const double wr = Lattice<Q>::w(K);  // reflection keeps |c|, hence the weight
    Store<ST, MT>::store(dst + (int64_t)kr * G.ncell + cell_index<S>(G, d), value, wr, G.scale);
}

// Slab flux, flux update and push of one population set (flow or scalar) of a line.
// P: collided populations of this work-item's cell.  lds: [sdir slot][ns] of the workgroup.
template <typename ST, typename MT, int Q, int NPTS, int S>
__device__ inline void transport(const FGeom& G, const FTables<MT>& tb, ST* dst, const MT (&P)[Q], MT* lds,
                                 int i, bool active, const int cc[3]) {
    const int ns = G.ns;
    if (!G.flux) {
        if (active)
            static_for<Q>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                push<ST, MT, Q, k, S>(G, dst, cc, P[k]);
            });
        return;
    }
    // 1. collided values of the S-directions into LDS
    if (active)
        static_for<Q>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if constexpr (comp<Q>(k, S) != 0) lds[sdir_slot<Q, S>(k) * ns + i] = P[k];
        });
    __syncthreads();
    // 2. slab values F (adv.cu:76-87); stencil points beyond a wall repeat the wall cell
    MT Fv[Q];
    const int per = G.periodic[S];
    static_for<Q>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr int sg = comp<Q>(k, S);
        Fv[k] = MT(0);
        if constexpr (sg != 0) {
            if (active) {
                const MT* Wk = tb.W + (sg > 0 ? 0 : NPTS) * ns;
                constexpr int slot = sdir_slot<Q, S>(k);
                MT acc = 0;
#pragma unroll
                for (int j = 0; j < NPTS; ++j) {
                    const int ii = wrap_or_clamp(i + (j - NPTS / 2) * sg, ns, per);
                    const MT term = Wk[j * ns + i] * lds[slot * ns + ii];
                    acc = j == 0 ? term : acc + term;
                }
                Fv[k] = acc;
            }
        }
    });
    __syncthreads();
    if (active)
        static_for<Q>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if constexpr (comp<Q>(k, S) != 0) lds[sdir_slot<Q, S>(k) * ns + i] = Fv[k];
        });
    __syncthreads();
    if (!active) return;
    const MT rn = tb.rdel[i];
    static_for<Q>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr int sg = comp<Q>(k, S);
        if constexpr (sg == 0) {
            // no component along S: set_tmp returns P exactly (adv.cu:64-65,241)
            push<ST, MT, Q, k, S>(G, dst, cc, P[k]);
        } else {
            constexpr int slot = sdir_slot<Q, S>(k);
            // U of this cell is the transported value of m = n - sg S
            int im = i - sg;
            bool have = im >= 0 && im < ns;
            if (!have && per) { im = im < 0 ? im + ns : im - ns; have = true; }
            if (have) {
                const MT Um = P[k] + (lds[slot * ns + im] * rn - Fv[k] * rn);
                int m[3] = {cc[0], cc[1], cc[2]};
                m[S] = im;
                push<ST, MT, Q, k, S>(G, dst, m, Um);
            }
            // downstream wall: the ghost's mirrored copy (adv.cu:196-199 along S)
            const int ip = i + sg;
            if (!per && (ip < 0 || ip >= ns)) {
                constexpr int mk = mirror<Q>(k, S);
                const MT Tm = P[mk] + (Fv[k] * rn - Fv[mk] * rn);
                push<ST, MT, Q, k, S>(G, dst, cc, Tm);
            }
        }
    });
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

// neighbour of cell cc along axis a at offset +-1: the wall cell itself beyond a wall (the
// ghost copies its owner, basic.cu:221-225), wrapped on a periodic axis
template <int S>
__device__ inline int64_t neighbour(const FGeom& G, const int cc[3], int a, int off) {
    int c[3] = {cc[0], cc[1], cc[2]};
    c[a] = wrap_or_clamp(cc[a] + off, G.n[a], G.periodic[a]);
    return cell_index<S>(G, c);
}

// (q[+] - q[-]) / (pos[+] - pos[-] + dz/1000) along axis a (Force, collision.cu:66-69), the
// period added across a periodic seam
template <typename MT, int S>
__device__ inline MT grad_axis(const FGeom& G, const FTables<MT>& tb, const MT* q, const int cc[3], int a, MT eps) {
    const int ip = wrap_or_clamp(cc[a] + 1, G.n[a], G.periodic[a]);
    const int im = wrap_or_clamp(cc[a] - 1, G.n[a], G.periodic[a]);
    MT d = tb.pos[a][ip] - tb.pos[a][im];
    if (G.periodic[a] && d <= MT(0)) d += tb.period[a];
    return (q[neighbour<S>(G, cc, a, 1)] - q[neighbour<S>(G, cc, a, -1)]) / (d + eps);
}

template <typename ST, typename MT, int Q, int QS, int NPTS, int S>
__global__ void __launch_bounds__(kMaxLine) k_fused_step(FGeom G, FTables<MT> tb, FBufs<ST, MT> b, FParams P) {
    IPL_DYNAMIC_LDS(MT, lds_f);
    MT* lds_g = lds_f + sdir_count<Q, S>() * G.ns;
    constexpr int o1 = S == 0 ? 1 : 0, o2 = S == 2 ? 1 : 2;
    const int i = threadIdx.x;
    const bool active = i < G.ns;
    const int64_t L = blockIdx.x;
    int cc[3];
    cc[S] = active ? i : 0;
    cc[o1] = (int)(L / G.n[o2]);
    cc[o2] = (int)(L % G.n[o2]);
    const int64_t n = L * G.ns + (active ? i : 0);
    const int64_t NC = G.ncell;
    const MT c = (MT)P.c, c2 = (MT)(P.c * P.c), dt = (MT)P.dt;

    MT Pf[Q], Pg[QS > 0 ? QS : 1];
    if (active) {
        MT f[Q], g[QS > 0 ? QS : 1];
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
        const MT S_ = G.scalar == IPL_SCALAR_SALINITY ? sc : MT(0);
        const MT phi = G.scalar == IPL_SCALAR_PHASE_FIELD ? sc : MT(0);
        const MT rho = MT(P.rho_L) + (MT(P.rho_H) - MT(P.rho_L)) * phi + S_ * MT(P.beta_s);
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
            gpx = grad_axis<MT, S>(G, tb, b.phi_n, cc, 0, eps);
            if (d3) gpy = grad_axis<MT, S>(G, tb, b.phi_n, cc, 1, eps);
            gpz = grad_axis<MT, S>(G, tb, b.phi_n, cc, 2, eps);
            // x term over delX^2, the term of directions 2/4 (z in 2D, y in 3D) over dz^2
            // (collision.cu:70-71)
            constexpr int a2 = d3 ? 1 : 2;
            const MT wx = tb.wid[0][cc[0]], dzz = MT(P.dz);
            const MT lap = (b.phi_n[neighbour<S>(G, cc, 0, 1)] - MT(2) * phi + b.phi_n[neighbour<S>(G, cc, 0, -1)]) / (wx * wx) +
                           (b.phi_n[neighbour<S>(G, cc, a2, 1)] - MT(2) * phi + b.phi_n[neighbour<S>(G, cc, a2, -1)]) / (dzz * dzz);
            const MT beta = MT(12.0) * MT(P.sigma) / MT(P.pf_thickness);
            const MT kap = MT(3.0) * MT(P.sigma) * MT(P.pf_thickness) / MT(2.0);
            const MT st = MT(4) * beta * phi * (phi - MT(1.0)) * (phi - MT(0.5)) - kap * lap;
            Fx += st * gpx; Fy += st * gpy; Fz += st * gpz;
        }
        if (P.boussinesq) Fz += -S_ * MT(P.beta_s) * MT(P.gravity);
        else Fz += -rho * MT(P.gravity);
        if (G.need_grad) {
            const MT grx = grad_axis<MT, S>(G, tb, b.rho_n, cc, 0, eps);
            const MT gry = d3 ? grad_axis<MT, S>(G, tb, b.rho_n, cc, 1, eps) : MT(0);
            const MT grz = grad_axis<MT, S>(G, tb, b.rho_n, cc, 2, eps);
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
                    geq[k] = MT(Lattice<QS>::w(k)) * S_ * e;
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
    transport<ST, MT, Q, NPTS, S>(G, tb, b.f_dst, Pf, lds_f, i, active, cc);
    if constexpr (QS > 0) transport<ST, MT, QS, NPTS, S>(G, tb, b.g_dst, Pg, lds_g, i, active, cc);
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

// population row q <-> FP64 staging buffer
template <typename ST, typename MT, int Q>
__global__ void k_convert_pops(FGeom G, const ST* src, double* dst, int q) {
    const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= G.ncell) return;
    dst[n] = (double)Store<ST, MT>::load(src + (int64_t)q * G.ncell + n, Lattice<Q>::w(q), G.inv_scale);
}

template <typename ST, typename MT, int Q>
__global__ void k_store_pops(FGeom G, const double* src, ST* dst, int q) {
    const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= G.ncell) return;
    Store<ST, MT>::store(dst + (int64_t)q * G.ncell + n, (MT)src[n], Lattice<Q>::w(q), G.scale);
}

// ---- launchers (declared in fused_common.hpp, instantiated in fused_q*.hip) ---------------
template <typename ST, typename MT, int Q, int QS>
void fused_launch_step(const FGeom& G, const FTables<MT>& tb, const FBufs<ST, MT>& b, const FParams& P, int npts,
                       unsigned block, size_t lds, hipStream_t st) {
    const unsigned grid = (unsigned)G.lines;
#define IPL_FSTEP(NP, SS) IPL_LAUNCH_SYNC((k_fused_step<ST, MT, Q, QS, NP, SS>), grid, block, lds, st, G, tb, b, P)
    if constexpr (Lattice<Q>::dim == 2) {
        if (npts == 5) { if (G.s == 0) IPL_FSTEP(5, 0); else IPL_FSTEP(5, 2); }
        else { if (G.s == 0) IPL_FSTEP(3, 0); else IPL_FSTEP(3, 2); }
    } else {
        if (npts == 5) { if (G.s == 0) IPL_FSTEP(5, 0); else if (G.s == 1) IPL_FSTEP(5, 1); else IPL_FSTEP(5, 2); }
        else { if (G.s == 0) IPL_FSTEP(3, 0); else if (G.s == 1) IPL_FSTEP(3, 1); else IPL_FSTEP(3, 2); }
    }
#undef IPL_FSTEP
}

template <typename ST, typename MT, int Q, int QS>
size_t fused_lds_bytes(int s, int ns) {
    int nf = 0, ng = 0;
    for (int k = 0; k < Q; ++k) nf += comp<Q>(k, s) != 0;
    if constexpr (QS > 0)
        for (int k = 0; k < QS; ++k) ng += comp<QS>(k, s) != 0;
    return (size_t)(nf + ng) * ns * sizeof(MT);
}

template <typename ST, typename MT, int QS>
void fused_launch_scalar_fields(const FGeom& G, const ST* g_src, MT* rho_n, MT* phi_n, const FParams& P, hipStream_t st) {
    IPL_LAUNCH((k_scalar_fields<ST, MT, QS>), (G.ncell + 255) / 256, 256, 0, st, G, g_src, rho_n, phi_n, P);
}

template <typename ST, typename MT, int Q, int QS>
void fused_launch_readout(const FGeom& G, const ST* f_src, const ST* g_src, const MT* F_old, double* out,
                          const FParams& P, hipStream_t st) {
    IPL_LAUNCH((k_readout<ST, MT, Q, QS>), (G.ncell + 255) / 256, 256, 0, st, G, f_src, g_src, F_old, out, P);
}

template <typename ST, typename MT, int Q>
void fused_launch_convert(const FGeom& G, const ST* src, double* dst, int q, bool to_double, ST* wdst,
                          hipStream_t st) {
    if (to_double) IPL_LAUNCH((k_convert_pops<ST, MT, Q>), (G.ncell + 255) / 256, 256, 0, st, G, src, dst, q);
    else IPL_LAUNCH((k_store_pops<ST, MT, Q>), (G.ncell + 255) / 256, 256, 0, st, G, dst, wdst, q);
}

}  // namespace fusedk
}  // namespace ipl

// explicit instantiation of every launcher of one (storage, math, Q, QS) combination
#define IPL_FUSED_INSTANTIATE(ST, MT, Q, QS)                                                                \
    template void ipl::fusedk::fused_launch_step<ST, MT, Q, QS>(                                             \
        const FGeom&, const FTables<MT>&, const FBufs<ST, MT>&, const FParams&, int, unsigned, size_t, hipStream_t); \
    template size_t ipl::fusedk::fused_lds_bytes<ST, MT, Q, QS>(int, int);                                    \
    template void ipl::fusedk::fused_launch_readout<ST, MT, Q, QS>(const FGeom&, const ST*, const ST*, const MT*, \
                                                                    double*, const FParams&, hipStream_t);
#define IPL_FUSED_INSTANTIATE_SCALAR(ST, MT, QS)                                                           \
    template void ipl::fusedk::fused_launch_scalar_fields<ST, MT, QS>(const FGeom&, const ST*, MT*, MT*,    \
                                                                       const FParams&, hipStream_t);
#define IPL_FUSED_INSTANTIATE_POPS(ST, MT, Q)                                                               \
    template void ipl::fusedk::fused_launch_convert<ST, MT, Q>(const FGeom&, const ST*, double*, int, bool, ST*, \
                                                                hipStream_t);
