// plain.hip - the plain GPU path: one kernel per stage of the reference step, structure-of-
// arrays populations [k][cell] (coalesced across the 64 lanes of a wavefront), neighbour and
// ghost tables from graph.cpp, FP64 Lagrange weights precomputed at init.  It follows the
// reference's stage semantics exactly (including its ghost copies and scratch aliasing inside
// IP_process, adv.cu:245-257) so it is bitwise-comparable with the FP64 oracle up to
// reassociation, and it is the "plain FP64" baseline of the fused path.
#include "hip_compat.hpp"

#include "lattice.hpp"
#include "plain.hpp"

namespace ipl {

namespace {

constexpr int kBlock = 256;

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Reference direction tables of the 3D two-axis corner scheme (set_cpm / set_cxypm,
// basic.cu:122-153); unset entries are 0 (= the cell itself), as in the reference.  The kernels
// index them only at compile-time directions.
constexpr int kCp[27] = {0, 1, 2, 3, 4, 0, 0, 7, 8, 9, 10, 1, 1, 3, 3, 2, 2, 4, 4, 7, 7, 8, 8, 9, 9, 10, 10};
constexpr int kCm[27] = {0, 3, 4, 1, 2, 0, 0, 10, 9, 8, 7, 3, 3, 1, 1, 4, 4, 2, 2, 10, 10, 9, 9, 8, 8, 7, 7};
constexpr int kCip[27] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 8, 8, 7, 7, 10, 10, 9, 9};
constexpr int kCim[27] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 9, 9, 10, 10, 7, 7, 8, 8};
constexpr int kCxp[27] = {0, 1, 0, 3, 0, 0, 0, 1, 1, 3, 3, 1, 1, 3, 3, 0, 0, 0, 0, 1, 1, 1, 1, 3, 3, 3, 3};
constexpr int kCxm[27] = {0, 3, 0, 1, 0, 0, 0, 3, 3, 1, 1, 3, 3, 1, 1, 0, 0, 0, 0, 3, 3, 3, 3, 1, 1, 1, 1};
constexpr int kCyp[27] = {0, 0, 2, 0, 4, 0, 0, 2, 4, 2, 4, 0, 0, 0, 0, 2, 2, 4, 4, 2, 2, 4, 4, 2, 2, 4, 4};
constexpr int kCym[27] = {0, 0, 4, 0, 2, 0, 0, 4, 2, 4, 2, 0, 0, 0, 0, 4, 4, 2, 2, 4, 4, 2, 2, 4, 4, 2, 2};

template <typename T> __device__ inline void atomic_add(T* a, T v) { atomicAdd(a, v); }

// position difference a - b along an axis, adding the period across a periodic seam
template <typename T>
__device__ inline T dpos(const PlainView<T>& s, const Params& P, int axis, int a, int b) {
    T d = s.pos[axis][a] - s.pos[axis][b];
    if (P.periodic[axis] && d <= T(0)) d += T(P.period[axis]);
    return d;
}

// ---- equ_f (collision.cu:3-14) -----------------------------------------------------------
template <typename T, int Q>
__global__ void __launch_bounds__(kBlock) k_equ_f(PlainView<T> s, Params P) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const T c = P.c, c2 = P.c * P.c;
    const T u = s.u[i], v = s.v[i], w = s.w[i], p = s.p[i];
    const T u2 = u * u + v * v + w * w;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const Dir d = Lattice<Q>::c(k);
        const T t = (T(d.x) * c * u + T(d.y) * c * v + T(d.z) * c * w) / c2;
        s.feq[k * s.Ntot + i] = T(Lattice<Q>::w(k)) * (p + T(3.0) * t + T(4.5) * t * t - T(1.5) * u2 / c2);
    }
}

// ---- Force (collision.cu:57-112) ---------------------------------------------------------
template <typename T, int Q>
__global__ void __launch_bounds__(kBlock) k_force(PlainView<T> s, Params P) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    constexpr bool d3 = Lattice<Q>::dim == 3;
    constexpr int up = d3 ? 5 : 2, down = d3 ? 6 : 4;
    const int64_t Nt = s.Ntot;
    const int n1 = s.neib[1 * Nt + i], n3 = s.neib[3 * Nt + i], n2 = s.neib[2 * Nt + i], n4 = s.neib[4 * Nt + i];
    const int nu = s.neib[up * Nt + i], nd = s.neib[down * Nt + i];
    const T eps = T(P.dz) / T(1000);
    const T* ph = s.phi; const T* rh = s.rho;
    T gx = (ph[n1] - ph[n3]) / (dpos(s, P, 0, n1, n3) + eps);
    T gy = 0;
    if (d3) gy = (ph[n2] - ph[n4]) / (dpos(s, P, 1, n2, n4) + eps);
    T gz = (ph[nu] - ph[nd]) / (dpos(s, P, 2, nu, nd) + eps);
    const T wdx = s.del[0][i], dz = P.dz;
    const T lap = (ph[n1] - T(2) * ph[i] + ph[n3]) / (wdx * wdx) + (ph[n2] - T(2) * ph[i] + ph[n4]) / (dz * dz);
    const T beta = T(12.0) * T(P.sigma) / T(P.pf_thickness), kap = T(3.0) * T(P.sigma) * T(P.pf_thickness) / T(2.0);
    const T st = T(4) * beta * ph[i] * (ph[i] - T(1.0)) * (ph[i] - T(0.5)) - kap * lap;
    T Fx = s.Fx[i] + st * gx, Fy = s.Fy[i] + st * gy, Fz = s.Fz[i] + st * gz;
    if (P.boussinesq) Fz += -s.S[i] * T(P.beta_s) * T(P.gravity);
    else Fz += -rh[i] * T(P.gravity);
    gx = (rh[n1] - rh[n3]) / (dpos(s, P, 0, n1, n3) + eps);
    gy = 0;
    if (d3) gy = (rh[n2] - rh[n4]) / (dpos(s, P, 1, n2, n4) + eps);
    gz = (rh[nu] - rh[nd]) / (dpos(s, P, 2, nu, nd) + eps);
    const T c = P.c, c2 = P.c * P.c;
    Fx += -s.p[i] * c2 / T(3.0) * gx;
    Fy += -s.p[i] * c2 / T(3.0) * gy;
    Fz += -s.p[i] * c2 / T(3.0) * gz;
    s.Fx[i] = Fx; s.Fy[i] = Fy; s.Fz[i] = Fz;
    const T den = rh[i] * c2 / T(3.0);
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const Dir d = Lattice<Q>::c(k);
        s.Fk[k * Nt + i] = T(Lattice<Q>::w(k)) * T(P.dt) * (T(d.x) * c * Fx + T(d.y) * c * Fy + T(d.z) * c * Fz) / den;
    }
}

// ---- col_f_MRT / col_f_SRT (collision.cu:115-165) ----------------------------------------
// MRT in moment space: m = M d, m_j *= s_j (stress rows 1/tau of the cell), f -= M^-1 m.
// M and M^-1 are compile-time constants, so the transforms reduce to adds.
template <typename T, int Q>
__global__ void __launch_bounds__(kBlock) k_collide(PlainView<T> s, Params P) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const int64_t Nt = s.Ntot;
    T d[Q];  // (the forcing term is read again where it is added: fewer live registers)
#pragma unroll
    for (int k = 0; k < Q; ++k) d[k] = s.f[k * Nt + i] - s.feq[k * Nt + i] + s.Fk[k * Nt + i] / T(2.0);
    const T tau = s.tau[i];
    if (P.collision == IPL_COLL_SRT) {
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const T fn = s.f[k * Nt + i] - d[k] / tau + s.Fk[k * Nt + i];
            s.ftmp[k * Nt + i] = fn;
        }
    } else {
        const T itau = T(1) / tau;
        T m[Q];
        static_for<Q>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            T acc = 0;
            static_for<Q>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                constexpr double mk = moment<Q>(r, k);
                if constexpr (mk != 0.0) acc += T(mk) * d[k];
            });
            m[r] = acc * (is_stress<Q>(r) ? itau : T(P.rates[r]));
        });
        static_for<Q>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            T acc = 0;
            static_for<Q>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                constexpr double mi = moment_inv<Q>(k, r);
                if constexpr (mi != 0.0) acc += T(mi) * m[r];
            });
            s.ftmp[k * Nt + i] = s.f[k * Nt + i] - acc + s.Fk[k * Nt + i];
        });
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) { s.f[k * Nt + i] = 0; s.feq[k * Nt + i] = 0; }
}

// ---- equ_g (collision.cu:16-27): the scalar's equilibrium, written into its populations ----
// (linear with 1/cs^2 = 4 on D3Q7, as in col_g_reg below)
template <typename T, int QS>
__global__ void __launch_bounds__(kBlock) k_equ_g(PlainView<T> s, Params P, const T* __restrict__ sc) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const T c = P.c, c2 = P.c * P.c;
    constexpr bool linear = QS == 7;
    const T cs2inv = linear ? T(4.0) : T(3.0);
    const T u = s.u[i], v = s.v[i], w = s.w[i], S = sc[i];
    const T u2 = u * u + v * v + w * w;
#pragma unroll
    for (int k = 0; k < QS; ++k) {
        const Dir d = Lattice<QS>::c(k);
        const T t = (T(d.x) * c * u + T(d.y) * c * v + T(d.z) * c * w) / c2;
        T e = T(1.0) + cs2inv * t;
        if (!linear) e += T(4.5) * t * t - T(1.5) * u2 / c2;
        s.g[k * s.Ntot + i] = T(Lattice<QS>::w(k)) * S * e;
    }
}

// ---- col_g_reg (collision.cu:167-188) ----------------------------------------------------
template <typename T, int QS>
__global__ void __launch_bounds__(kBlock) k_col_g_reg(PlainView<T> s, Params P) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const int64_t Nt = s.Ntot;
    const T c = P.c, c2 = P.c * P.c;
    constexpr bool linear = QS == 7;
    const T cs2inv = linear ? T(4.0) : T(3.0);
    const T u = s.u[i], v = s.v[i], w = s.w[i], S = s.S[i];
    const T u2 = u * u + v * v + w * w;
    T geq[QS];
    T jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int k = 0; k < QS; ++k) {
        const Dir d = Lattice<QS>::c(k);
        const T t = (T(d.x) * c * u + T(d.y) * c * v + T(d.z) * c * w) / c2;
        T e = T(1.0) + cs2inv * t;
        if (!linear) e += T(4.5) * t * t - T(1.5) * u2 / c2;
        geq[k] = T(Lattice<QS>::w(k)) * S * e;
        const T gn = s.g[k * Nt + i] - geq[k];
        jx += gn * T(d.x) * c; jy += gn * T(d.y) * c; jz += gn * T(d.z) * c;
    }
    const T om = T(1.0) - T(1.0) / s.taus[i];
#pragma unroll
    for (int k = 0; k < QS; ++k) {
        const Dir d = Lattice<QS>::c(k);
        const T gneq = T(Lattice<QS>::w(k)) * cs2inv * (jx * T(d.x) * c + jy * T(d.y) * c + jz * T(d.z) * c) / c2;
        s.gtmp[k * Nt + i] = geq[k] + om * gneq;
        s.geq[k * Nt + i] = 0;
        s.g[k * Nt + i] = 0;
    }
}

// ---- col_PF (collision.cu:190-223) -------------------------------------------------------
template <typename T, int Q>
__global__ void __launch_bounds__(kBlock) k_col_pf(PlainView<T> s, Params P) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    constexpr bool d3 = Lattice<Q>::dim == 3;
    constexpr int up = d3 ? 5 : 2, down = d3 ? 6 : 4;
    const int64_t Nt = s.Ntot;
    const T eps = T(P.dz) / T(1000);
    const T* ph = s.phi;
    const int n1 = s.neib[1 * Nt + i], n3 = s.neib[3 * Nt + i], n2 = s.neib[2 * Nt + i], n4 = s.neib[4 * Nt + i];
    const int nu = s.neib[up * Nt + i], nd = s.neib[down * Nt + i];
    const T gx = (ph[n1] - ph[n3]) / (dpos(s, P, 0, n1, n3) + eps);
    T gy = 0;
    if (d3) gy = (ph[n2] - ph[n4]) / (dpos(s, P, 1, n2, n4) + eps);
    const T gz = (ph[nu] - ph[nd]) / (dpos(s, P, 2, nu, nd) + eps);
    const T u = s.u[i], v = s.v[i], w = s.w[i], phi = ph[i];
    const T nabu = (phi * u - s.phiold[i] * s.uold[i]) / T(P.dt);
    const T nabv = (phi * v - s.phiold[i] * s.vold[i]) / T(P.dt);
    const T nabw = (phi * w - s.phiold[i] * s.wold[i]) / T(P.dt);
    s.phiold[i] = phi; s.uold[i] = u; s.vold[i] = v; s.wold[i] = w;
    const T c = P.c, c2 = P.c * P.c;
    const T u2 = u * u + v * v + w * w;
    const T nrm = sqrt(gx * gx + gy * gy + gz * gz + T(1e-8));
    const T src = (T(1.0) - T(4.0) * (phi - T(0.5)) * (phi - T(0.5))) / T(P.pf_thickness);
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const Dir d = Lattice<Q>::c(k);
        const T cx = T(d.x) * c, cy = T(d.y) * c, cz = T(d.z) * c;
        const T t = (cx * u + cy * v + cz * w) / c2;
        const T geq = T(Lattice<Q>::w(k)) * phi * (T(1.0) + T(3.0) * t + T(4.5) * t * t - T(1.5) * u2 / c2);
        const T Fkp = T(P.dt) * T(Lattice<Q>::w(k)) *
                      (src * (cx * gx + cy * gy + cz * gz) / nrm + T(3) * (cx * nabu + cy * nabv + cz * nabw) / c2);
        const T g = s.g[k * Nt + i];
        s.gtmp[k * Nt + i] = g - (g - geq - Fkp * (T(P.pf_tau) - T(0.5))) / T(P.pf_tau);
        s.geq[k * Nt + i] = 0;
        s.g[k * Nt + i] = 0;
    }
}

// ---- update_scalar (collision.cu:225-236) ------------------------------------------------
template <typename T, int QS>
__global__ void __launch_bounds__(kBlock) k_update_scalar(PlainView<T> s, T* out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    T a = 0;
#pragma unroll
    for (int k = 0; k < QS; ++k) a += s.g[k * s.Ntot + i];
    out[i] = a;
}

// ---- update_rho (collision.cu:238-257) ---------------------------------------------------
template <typename T, int Q>
__global__ void __launch_bounds__(kBlock) k_update_rho(PlainView<T> s, Params P) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const T c = P.c;
    T p = 0, ux = 0, uy = 0, uz = 0;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const Dir d = Lattice<Q>::c(k);
        const T f = s.f[k * s.Ntot + i];
        p += f; ux += T(d.x) * c * f; uy += T(d.y) * c * f; uz += T(d.z) * c * f;
    }
    const T rho = T(P.rho_L) + (T(P.rho_H) - T(P.rho_L)) * s.phi[i] + s.S[i] * T(P.beta_s);
    s.p[i] = p;
    s.rho[i] = rho;
    s.u[i] = ux + s.Fx[i] * T(P.dt) / (T(2.0) * rho);
    s.v[i] = uy + s.Fy[i] * T(P.dt) / (T(2.0) * rho);
    s.w[i] = uz + s.Fz[i] * T(P.dt) / (T(2.0) * rho);
}

// ---- set_wall_rho (basic.cu:254-261) / resetF (collision.cu:258-264) ---------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_set_wall_field(PlainView<T> s, T* a) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x + s.N;
    if (g >= s.Ntot) return;
    a[g] = a[s.neib[g]];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_reset_force(PlainView<T> s) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    s.Fx[i] = 0; s.Fy[i] = 0; s.Fz[i] = 0;
}

// ======================= IP_process (adv.cu:245-257) =====================================
// Pops<T> names the four buffers of one transported field: f doubles as finy, feq as fin.
template <typename T> struct Pops { T *f, *ftmp, *fout, *fin; const int32_t *nextB, *nextK; };

// set_wall_f (adv.cu:3-13)
template <typename T, int QF>
__global__ void __launch_bounds__(kBlock) k_set_wall_f(PlainView<T> s, Pops<T> b) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x + s.N;
    if (g >= s.Ntot) return;
    const int o = s.neib[g];
#pragma unroll
    for (int k = 0; k < QF; ++k) b.ftmp[k * s.Ntot + g] = b.ftmp[k * s.Ntot + o];
}

// slab reconstruction along one axis for direction k (adv.cu:76-99)
template <typename T, int QN, int QF>
__device__ inline void flux_1d(const PlainView<T>& s, const Pops<T>& b, const Params& P, int64_t i, int k, int axis) {
    const int64_t Nt = s.Ntot;
    const int sg = comp<QF>(k, axis);
    const int dp = axis_dir<QN>(axis, sg), dm = axis_dir<QN>(axis, -sg);
    const int cp = s.neib[dp * Nt + i], cm = s.neib[dm * Nt + i];
    const int si = sg > 0 ? 0 : 1;
    const int np = P.npts;
    const T* W = s.W + (size_t)((axis * 2 + si) * np) * s.N;
    T F;
    if (np == 5) {
        const int cpp = s.neib[dp * Nt + cp], cmm = s.neib[dm * Nt + cm];
        F = W[0 * s.N + i] * b.ftmp[k * Nt + cmm] + W[1 * s.N + i] * b.ftmp[k * Nt + cm] +
            W[2 * s.N + i] * b.ftmp[k * Nt + i] + W[3 * s.N + i] * b.ftmp[k * Nt + cp] +
            W[4 * s.N + i] * b.ftmp[k * Nt + cpp];
    } else {
        F = W[0 * s.N + i] * b.ftmp[k * Nt + cm] + W[1 * s.N + i] * b.ftmp[k * Nt + i] + W[2 * s.N + i] * b.ftmp[k * Nt + cp];
    }
    b.fin[k * Nt + i] = F * s.rdel[axis][cp];
    b.fout[k * Nt + i] = F * s.rdel[axis][i];
}

// set_out of the 3D x-y scheme (adv.cu:50-147), grouped so that what directions share is read
// once per cell.  The one-axis directions of one axis and sign (x, xz, or y, yz) share their
// neighbours, 5-point weights and widths.  The corner directions of one x-y orientation (sx, sy)
// share their 5 x 5 stencil in the x-y plane: the xy direction and the two corner directions
// z = +-1 take the same direction tables except cip / cim, which are 0 for the xy direction, so
// that its anti-diagonal points are the cell itself (the rest neighbour of a cell is the cell).
// Each direction's arithmetic is the reference's num3 loop; its three scattered contributions
// are added with FP atomics as in the reference (at most two land on one destination, so their
// order cannot change a bit).  A scheduling fence after each direction group keeps the compiler
// from hoisting the next group's gathers over the current one: hoisted, the FP64 D3Q27 kernel
// held 256 VGPRs (one wave per SIMD) and the FP32 one 157; fenced, the falling-sphere grid's
// plain step runs 7 % (FP32) and 10 % (FP64) faster than with one stencil per direction.
template <typename T, int QN, int QF, int AX, int SG>
__device__ inline void corner_scheme_axis(const PlainView<T>& s, const Pops<T>& b, int64_t i) {
    constexpr int dp = axis_dir<QN>(AX, SG), dm = axis_dir<QN>(AX, -SG);
    const int64_t Nt = s.Ntot;
    const int cp = s.neib[dp * Nt + i], cm = s.neib[dm * Nt + i];
    const int cpp = s.neib[dp * Nt + cp], cmm = s.neib[dm * Nt + cm];
    const T* W = s.W + (size_t)((AX * 2 + (SG > 0 ? 0 : 1)) * 5) * s.N;
    const T w0 = W[0 * s.N + i], w1 = W[1 * s.N + i], w2 = W[2 * s.N + i], w3 = W[3 * s.N + i], w4 = W[4 * s.N + i];
    const T rp = s.rdel[AX][cp], ri = s.rdel[AX][i];
    static_for<QF>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (comp<QF>(k, AX) == SG && comp<QF>(k, 1 - AX) == 0) {
            const T F = w0 * b.ftmp[k * Nt + cmm] + w1 * b.ftmp[k * Nt + cm] + w2 * b.ftmp[k * Nt + i] +
                        w3 * b.ftmp[k * Nt + cp] + w4 * b.ftmp[k * Nt + cpp];
            b.fin[k * Nt + i] = F * rp;
            b.fout[k * Nt + i] = F * ri;
        }
    });
}

template <typename T>
__device__ inline void corner_dir(const PlainView<T>& s, const Pops<T>& b, const Params& P, int64_t i, int kk,
                                  const int (&idf)[25], const T (&cxw)[5], const T (&cyw)[5], T dXi, T dYi,
                                  T dX13, T dY13, T dX17, T dY17) {
    const int64_t Nt = s.Ntot;
    const T* ft = b.ftmp + (size_t)kk * Nt;
    T fx = 0, fy = 0, fxy = 0;
#pragma unroll
    for (int j = 0; j < 5; ++j) fx += cxw[j] * ft[idf[j + 10]];
#pragma unroll
    for (int j = 0; j < 5; ++j) fy += cyw[j] * ft[idf[2 + j * 5]];
#pragma unroll
    for (int a = 0; a < 5; ++a)
#pragma unroll
        for (int c = 0; c < 5; ++c) fxy += cxw[a] * cyw[c] * ft[idf[a + c * 5]];
    const T dz = P.dz;
    b.fout[kk * Nt + i] = fx * dz / dXi + fy * dz / dYi - fxy * (dz * dz) / (dXi * dYi);
    atomic_add(&b.fin[kk * Nt + i], fxy * (dz * dz) / (dX13 * dY17));
    atomic_add(&b.fin[kk * Nt + idf[11]], (fy - fxy * dz / dX17) * dz / dY17);   // the -x neighbour
    atomic_add(&b.f[kk * Nt + idf[7]], (fx - fxy * dz / dY13) * dz / dX13);      // the -y neighbour
}

template <typename T, int QF, int SX, int SY>
__device__ inline void corner_scheme_orientation(const PlainView<T>& s, const Pops<T>& b, const Params& P, int64_t i) {
    constexpr int kxy = find_dir<QF>(SX, SY, 0);
    constexpr int kz1 = QF == 27 ? find_dir<QF>(SX, SY, 1) : kxy, kz2 = QF == 27 ? find_dir<QF>(SX, SY, -1) : kxy;
    static_assert(kxy > 0 && kz1 > 0 && kz2 > 0, "an x-y orientation without its directions");
    static_assert(kCxp[kz1] == kCxp[kxy] && kCxm[kz1] == kCxm[kxy] && kCyp[kz1] == kCyp[kxy] &&
                  kCym[kz1] == kCym[kxy] && kCp[kz1] == kCp[kxy] && kCm[kz1] == kCm[kxy] &&
                  kCxp[kz2] == kCxp[kxy] && kCxm[kz2] == kCxm[kxy] && kCyp[kz2] == kCyp[kxy] &&
                  kCym[kz2] == kCym[kxy] && kCp[kz2] == kCp[kxy] && kCm[kz2] == kCm[kxy] &&
                  kCip[kz2] == kCip[kz1] && kCim[kz2] == kCim[kz1] && kCip[kxy] == 0 && kCim[kxy] == 0,
                  "directions of one orientation share their stencil");
    constexpr int xp = kCxp[kxy], xm = kCxm[kxy], yp = kCyp[kxy], ym = kCym[kxy], dpp = kCp[kxy], dmm = kCm[kxy];
    const int64_t Nt = s.Ntot;
    auto nb = [&](int cell, int d) { return s.neib[d * Nt + cell]; };
    int idf[25];
    idf[12] = (int)i;
    idf[13] = nb(i, xp); idf[14] = nb(idf[13], xp);
    idf[11] = nb(i, xm); idf[10] = nb(idf[11], xm);
    idf[17] = nb(i, yp); idf[22] = nb(idf[17], yp);
    idf[7] = nb(i, ym); idf[2] = nb(idf[7], ym);
    idf[18] = nb(i, dpp); idf[24] = nb(idf[18], dpp);
    idf[6] = nb(i, dmm); idf[0] = nb(idf[6], dmm);
    idf[19] = nb(idf[14], yp); idf[9] = nb(idf[14], ym);
    idf[15] = nb(idf[10], yp); idf[5] = nb(idf[10], ym);
    idf[23] = nb(idf[22], xp); idf[21] = nb(idf[22], xm);
    idf[3] = nb(idf[2], xp); idf[1] = nb(idf[2], xm);
    idf[8] = idf[4] = idf[16] = idf[20] = (int)i;
    const T* Wx = s.W + (size_t)((0 * 2 + (SX > 0 ? 0 : 1)) * 5) * s.N;
    const T* Wy = s.W + (size_t)((1 * 2 + (SY > 0 ? 0 : 1)) * 5) * s.N;
    T cxw[5], cyw[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) { cxw[j] = Wx[j * s.N + i]; cyw[j] = Wy[j * s.N + i]; }
    const T* dX = s.del[0]; const T* dY = s.del[1];
    const T dXi = dX[i], dYi = dY[i], dX13 = dX[idf[13]], dY13 = dY[idf[13]], dX17 = dX[idf[17]], dY17 = dY[idf[17]];
    corner_dir<T>(s, b, P, i, kxy, idf, cxw, cyw, dXi, dYi, dX13, dY13, dX17, dY17);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (QF == 27) {
        constexpr int ip = kCip[kz1], im = kCim[kz1];
        idf[8] = nb(i, ip); idf[4] = nb(idf[8], ip);
        idf[16] = nb(i, im); idf[20] = nb(idf[16], im);
        corner_dir<T>(s, b, P, i, kz1, idf, cxw, cyw, dXi, dYi, dX13, dY13, dX17, dY17);
        __builtin_amdgcn_sched_barrier(0);
        corner_dir<T>(s, b, P, i, kz2, idf, cxw, cyw, dXi, dYi, dX13, dY13, dX17, dY17);
    }
}

template <typename T, int QN, int QF>
__device__ inline void set_out_corner_scheme(const PlainView<T>& s, const Pops<T>& b, const Params& P, int64_t i) {
    const int64_t Nt = s.Ntot;
    static_for<QF>([&](auto kc) {  // no x or y component: no flux
        constexpr int k = decltype(kc)::value;
        if constexpr (comp<QF>(k, 0) == 0 && comp<QF>(k, 1) == 0) {
            const T v = b.ftmp[k * Nt + i];
            b.fout[k * Nt + i] = v;
            b.fin[k * Nt + i] = v;
        }
    });
    corner_scheme_axis<T, QN, QF, 0, 1>(s, b, i); __builtin_amdgcn_sched_barrier(0);
    corner_scheme_axis<T, QN, QF, 0, -1>(s, b, i); __builtin_amdgcn_sched_barrier(0);
    corner_scheme_axis<T, QN, QF, 1, 1>(s, b, i); __builtin_amdgcn_sched_barrier(0);
    corner_scheme_axis<T, QN, QF, 1, -1>(s, b, i); __builtin_amdgcn_sched_barrier(0);
    corner_scheme_orientation<T, QF, 1, 1>(s, b, P, i); __builtin_amdgcn_sched_barrier(0);
    corner_scheme_orientation<T, QF, 1, -1>(s, b, P, i); __builtin_amdgcn_sched_barrier(0);
    corner_scheme_orientation<T, QF, -1, 1>(s, b, P, i); __builtin_amdgcn_sched_barrier(0);
    corner_scheme_orientation<T, QF, -1, -1>(s, b, P, i);
}

// set_out (adv.cu:50-147)
template <typename T, int QN, int QF>
__global__ void __launch_bounds__(kBlock) k_set_out(PlainView<T> s, Pops<T> b, Params P) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const int64_t Nt = s.Ntot;
    if constexpr (QF == 19 || QF == 27) {  // the corner scheme is 3D and takes no D3Q7 scalar
        if (P.flux_mode == FLUX_MODE_CORNER) {
            set_out_corner_scheme<T, QN, QF>(s, b, P, i);
            return;
        }
    }
#pragma unroll
    for (int k = 0; k < QF; ++k) {
        const bool flux = P.flux_mode == FLUX_MODE_ONE_AXIS && comp<QF>(k, P.flux_axis) != 0;
        if (!flux) {
            const T v = b.ftmp[k * Nt + i];
            b.fout[k * Nt + i] = v;
            b.fin[k * Nt + i] = v;
        } else {
            flux_1d<T, QN, QF>(s, b, P, i, k, P.flux_axis);
        }
    }
}

// set_fin (adv.cu:148-175), corner scheme only; atomics where the reference races
template <typename T, int QF>
__global__ void __launch_bounds__(kBlock) k_set_fin(PlainView<T> s, Pops<T> b) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x + s.N;
    if (g >= s.Ntot) return;
    const int64_t Nt = s.Ntot;
    const int nk[8] = {9, 8, 10, 7, 7, 10, 8, 9};
    const int nkz[8] = {19, 20, 21, 22, 23, 24, 25, 26};
    const int nneib[8] = {1, 2, 1, 4, 3, 2, 3, 4};
    const int k3[12] = {7, 8, 9, 10, 19, 21, 23, 25, 20, 22, 24, 26};
    for (int k = 0; k < 4; ++k) {
        const int t0 = s.neib[nneib[k * 2] * Nt + g], t1 = s.neib[nneib[k * 2 + 1] * Nt + g];
        if (t0 != g) {  // the reference's (1 - neib/id) factor is 0 on self links
            const int a = nk[2 * (k3[k] - 7)];
            atomic_add(&b.fin[a * Nt + t0], b.fin[k3[k] * Nt + g]);
            if (QF == 27) {
                atomic_add(&b.fin[nkz[2 * (a - 7)] * Nt + t0], b.fin[k3[k + 4] * Nt + g]);
                atomic_add(&b.fin[nkz[2 * (a - 7) + 1] * Nt + t0], b.fin[k3[k + 8] * Nt + g]);
            }
        }
        if (t1 != g) {
            const int bb = nk[2 * (k3[k] - 7) + 1];
            atomic_add(&b.f[bb * Nt + t1], b.f[k3[k] * Nt + g]);
            if (QF == 27) {
                atomic_add(&b.f[nkz[2 * (bb - 7)] * Nt + t1], b.f[k3[k + 4] * Nt + g]);
                atomic_add(&b.f[nkz[2 * (bb - 7) + 1] * Nt + t1], b.f[k3[k + 8] * Nt + g]);
            }
        }
    }
}

// reset_f (adv.cu:176-187)
template <typename T, int QF>
__global__ void __launch_bounds__(kBlock) k_reset_f(PlainView<T> s, Pops<T> b) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
#pragma unroll
    for (int k = 0; k < QF; ++k) {
        const int64_t j = k * s.Ntot + i;
        b.fin[j] += b.f[j];
        b.f[j] = b.ftmp[j];
    }
}

// set_wall_f_inv (adv.cu:188-232): ghost <- owner, mirrored along the flux axis; the 3D
// corner scheme mirrors x and y and takes the xy / corner directions through nextB/nextK
template <typename T, int QN, int QF>
__global__ void __launch_bounds__(kBlock) k_set_wall_f_inv(PlainView<T> s, Pops<T> b, T* a, Params P) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x + s.N;
    if (g >= s.Ntot) return;
    const int64_t Nt = s.Ntot;
    const int o = s.neib[g];
    if (P.flux_mode == FLUX_MODE_CORNER) {
        const int kxy[19] = {0, 3, 4, 1, 2, 0, 0, 0, 0, 0, 0, 13, 14, 11, 12, 17, 18, 15, 16};
        const int kn1[15] = {0, 1, 2, 3, 4, 5, 6, 11, 12, 13, 14, 15, 16, 17, 18};
        for (int q = 0; q < 15; ++q) a[kn1[q] * Nt + g] = a[kxy[kn1[q]] * Nt + o];
        const int k3[12] = {7, 8, 9, 10, 19, 21, 23, 25, 20, 22, 24, 26};
        const int nkz[8] = {20, 19, 22, 21, 24, 23, 26, 25};
        const int64_t G = s.Ntot - s.N, wid = g - s.N;
        for (int q = 0; q < 4; ++q) {
            const int kq = opposite<QF>(k3[q]);
            const int nk = opposite<QF>(b.nextK[kq * G + wid]);
            const int neB = b.nextB[kq * G + wid];
            const int src = s.neib[neB * Nt + o];
            a[k3[q] * Nt + g] = a[nk * Nt + src];
            if (QF == 27) {
                a[k3[q + 4] * Nt + g] = a[nkz[2 * (nk - 7) + 1] * Nt + src];
                a[k3[q + 8] * Nt + g] = a[nkz[2 * (nk - 7)] * Nt + src];
            }
        }
    } else if (P.flux_mode == FLUX_MODE_ONE_AXIS) {
#pragma unroll
        for (int k = 1; k < QF; ++k) {
            const int ax = P.flux_axis;
            if (comp<QF>(k, ax) == 0) continue;  // never read by set_tmp
            a[k * Nt + g] = a[mirror<QF>(k, ax) * Nt + o];
        }
    }
}

// set_tmp (adv.cu:234-244)
template <typename T, int QN, int QF>
__global__ void __launch_bounds__(kBlock) k_set_tmp(PlainView<T> s, Pops<T> b, Params P) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const int64_t Nt = s.Ntot;
    static_for<QF>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr int dcorner = QF == 19 || QF == 27 ? kCp[k] : 0;
        int d = 0;
        if (P.flux_mode == FLUX_MODE_ONE_AXIS) {
            const int sg = comp<QF>(k, P.flux_axis);
            if (sg) d = axis_dir<QN>(P.flux_axis, sg);
        } else if (P.flux_mode == FLUX_MODE_CORNER) {
            d = dcorner;
        }
        const int n = d ? s.neib[d * Nt + i] : (int)i;
        b.ftmp[k * Nt + i] = b.f[k * Nt + n] + (b.fin[k * Nt + i] - b.fout[k * Nt + n]);
    });
}

// set_wall_propagation (set_bound.cu:149-168)
template <typename T, int QF>
__global__ void __launch_bounds__(kBlock) k_set_wall_propagation(PlainView<T> s, Pops<T> b, Params P) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x + s.N;
    if (g >= s.Ntot) return;
    const int64_t Nt = s.Ntot, G = s.Ntot - s.N, wid = g - s.N;
    const int o = s.neib[g];
#pragma unroll
    for (int k = 1; k < QF; ++k) {
        if (P.wall_bc == IPL_BC_BOUNCE_BACK) {
            b.ftmp[k * Nt + g] = b.ftmp[opposite<QF>(k) * Nt + o];
        } else {
            const int src = s.neib[b.nextB[k * G + wid] * Nt + o];
            b.ftmp[k * Nt + g] = b.ftmp[b.nextK[k * G + wid] * Nt + src];
        }
    }
}

// propagation (propagation.cu:2-12)
template <typename T, int QF>
__global__ void __launch_bounds__(kBlock) k_propagation(PlainView<T> s, Pops<T> b) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const int64_t Nt = s.Ntot;
#pragma unroll
    for (int k = 0; k < QF; ++k) b.f[k * Nt + i] = b.ftmp[k * Nt + s.neib[opposite<QF>(k) * Nt + i]];
}

// reset_wall_fin (basic.cu:274-284)
template <typename T, int QF>
__global__ void __launch_bounds__(kBlock) k_reset_wall_fin(PlainView<T> s, Pops<T> b) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x + s.N;
    if (g >= s.Ntot) return;
#pragma unroll
    for (int k = 0; k < QF; ++k) { b.fin[k * s.Ntot + g] = 0; b.f[k * s.Ntot + g] = 0; }
}

template <typename T, int QN, int QF>
void ip_process(const PlainView<T>& s, const Pops<T>& b, const Params& P, hipStream_t st) {
    const int64_t G = s.Ntot - s.N;
    const unsigned bn = blocks_for(s.N), bg = blocks_for(G > 0 ? G : 1);
    if (G > 0) IPL_LAUNCH((k_set_wall_f<T, QF>), bg, kBlock, 0, st, s, b);
    IPL_LAUNCH((k_set_out<T, QN, QF>), bn, kBlock, 0, st, s, b, P);
    if (G > 0 && P.flux_mode == FLUX_MODE_CORNER) IPL_LAUNCH((k_set_fin<T, QF>), bg, kBlock, 0, st, s, b);
    IPL_LAUNCH((k_reset_f<T, QF>), bn, kBlock, 0, st, s, b);
    if (G > 0) {
        IPL_LAUNCH((k_set_wall_f_inv<T, QN, QF>), bg, kBlock, 0, st, s, b, b.f, P);
        IPL_LAUNCH((k_set_wall_f_inv<T, QN, QF>), bg, kBlock, 0, st, s, b, b.fout, P);
    }
    IPL_LAUNCH((k_set_tmp<T, QN, QF>), bn, kBlock, 0, st, s, b, P);
    if (G > 0) IPL_LAUNCH((k_set_wall_propagation<T, QF>), bg, kBlock, 0, st, s, b, P);
    IPL_LAUNCH((k_propagation<T, QF>), bn, kBlock, 0, st, s, b);
    if (G > 0) IPL_LAUNCH((k_reset_wall_fin<T, QF>), bg, kBlock, 0, st, s, b);
}

template <typename T, int Q, int QS>
int run_stage(const PlainView<T>& s, const Params& P, int stage, hipStream_t st) {
    const unsigned bn = blocks_for(s.N);
    const int64_t G = s.Ntot - s.N;
    const unsigned bg = blocks_for(G > 0 ? G : 1);
    switch (stage) {
    case IPL_STAGE_EQU_F: IPL_LAUNCH((k_equ_f<T, Q>), bn, kBlock, 0, st, s, P); break;
    case IPL_STAGE_FORCE: IPL_LAUNCH((k_force<T, Q>), bn, kBlock, 0, st, s, P); break;
    case IPL_STAGE_COLLIDE: IPL_LAUNCH((k_collide<T, Q>), bn, kBlock, 0, st, s, P); break;
    case IPL_STAGE_IP_F: {
        Pops<T> b{s.f, s.ftmp, s.fout, s.feq, s.nextB, s.nextK};
        ip_process<T, Q, Q>(s, b, P, st);
        break;
    }
    case IPL_STAGE_SCALAR_COLLIDE:
        if constexpr (QS > 0) {
            if (P.scalar == IPL_SCALAR_SALINITY) IPL_LAUNCH((k_col_g_reg<T, QS>), bn, kBlock, 0, st, s, P);
            else if constexpr (QS == Q) IPL_LAUNCH((k_col_pf<T, Q>), bn, kBlock, 0, st, s, P);
        }
        break;
    case IPL_STAGE_IP_G:
        if constexpr (QS > 0) {
            Pops<T> b{s.g, s.gtmp, s.gout, s.geq, s.snextB, s.snextK};
            ip_process<T, Q, QS>(s, b, P, st);
        }
        break;
    case IPL_STAGE_UPDATE_SCALAR:
        if constexpr (QS > 0)
            IPL_LAUNCH((k_update_scalar<T, QS>), bn, kBlock, 0, st, s, P.scalar == IPL_SCALAR_SALINITY ? s.S : s.phi);
        break;
    case IPL_STAGE_UPDATE_RHO: IPL_LAUNCH((k_update_rho<T, Q>), bn, kBlock, 0, st, s, P); break;
    case IPL_STAGE_SET_WALL_RHO:
        if (G > 0) {
            IPL_LAUNCH((k_set_wall_field<T>), bg, kBlock, 0, st, s, s.rho);
            if (P.scalar == IPL_SCALAR_PHASE_FIELD) IPL_LAUNCH((k_set_wall_field<T>), bg, kBlock, 0, st, s, s.phi);
        }
        break;
    case IPL_STAGE_RESET_F: IPL_LAUNCH((k_reset_force<T>), bn, kBlock, 0, st, s); break;
    case IPL_STAGE_EQU_G:
        if constexpr (QS > 0) {
            IPL_LAUNCH((k_equ_g<T, QS>), bn, kBlock, 0, st, s, P, P.scalar == IPL_SCALAR_SALINITY ? s.S : s.phi);
            break;
        }
        return IPL_ERR_UNSUPPORTED;
    default: return IPL_ERR_ARGUMENT;
    }
    return IPL_OK;
}

}  // namespace

template <typename T>
int plain_stage(const PlainView<T>& s, const Params& P, int Q, int Qs, int stage, hipStream_t st) {
#define IPL_CASE(QQ, QSS) if (Q == QQ && Qs == QSS) return run_stage<T, QQ, QSS>(s, P, stage, st);
    IPL_CASE(9, 0) IPL_CASE(9, 9) IPL_CASE(19, 0) IPL_CASE(19, 19) IPL_CASE(19, 7)
    IPL_CASE(27, 0) IPL_CASE(27, 27) IPL_CASE(27, 7)
#undef IPL_CASE
    return IPL_ERR_UNSUPPORTED;
}

template int plain_stage<double>(const PlainView<double>&, const Params&, int, int, int, hipStream_t);
template int plain_stage<float>(const PlainView<float>&, const Params&, int, int, int, hipStream_t);

}  // namespace ipl
