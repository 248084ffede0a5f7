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

// minimum waves per SIMD the register allocator is asked to leave room for (0: no hint)
#ifndef IPL_FUSED_WAVES_PER_EU
#define IPL_FUSED_WAVES_PER_EU 0
#endif
#if IPL_FUSED_WAVES_PER_EU > 0 && !defined(IPL_CPU_EMULATION)
#define IPL_FUSED_OCCUPANCY __attribute__((amdgpu_waves_per_eu(IPL_FUSED_WAVES_PER_EU)))
#else
#define IPL_FUSED_OCCUPANCY
#endif

// ---- storage conversion ------------------------------------------------------------------
template <typename ST, typename MT> struct Store;
// (P: a pointer to the stored type in the generic or the global address space)
template <typename MT> struct Store<double, MT> {
    template <typename P> __device__ static MT load(P p, double, float) { return (MT)*p; }
    template <typename P> __device__ static void store(P p, MT v, double, float) { *p = (double)v; }
};
template <typename MT> struct Store<float, MT> {
    template <typename P> __device__ static MT load(P p, double, float) { return (MT)*p; }
    template <typename P> __device__ static void store(P p, MT v, double, float) { *p = (float)v; }
};
// FP16: populations are already deviations from the rest state (sum f = p ~ 0, collision.cu:11);
// they are stored divided by their lattice weight and scaled by a power of two.
template <typename MT> struct Store<__half, MT> {
    template <typename P> __device__ static MT load(P p, double w, float inv_scale) {
        return (MT)(__half2float(*p) * ((float)w * inv_scale));
    }
    template <typename P> __device__ static void store(P p, MT v, double w, float scale) {
        *p = __float2half_rn((float)v * (scale / (float)w));
    }
};

// The population rows a step streams: each cell's own rows read once, every row written once.
// With NT the accesses carry the non-temporal hint, for steps whose population arrays exceed the
// last-level cache (the next step would not find them there); otherwise plain accesses, which keep
// them cached for the next step.
template <bool NT, typename ST, typename MT, typename P> __device__ inline MT load_pop(P p, double w, float s) {
#ifndef IPL_CPU_EMULATION
    if constexpr (NT && !std::is_same<ST, __half>::value) return (MT)__builtin_nontemporal_load(p);
#endif
    return Store<ST, MT>::load(p, w, s);
}
template <bool NT, typename ST, typename MT, typename P> __device__ inline void store_pop(P p, MT v, double w, float s) {
#ifndef IPL_CPU_EMULATION
    if constexpr (NT && !std::is_same<ST, __half>::value) { __builtin_nontemporal_store((ST)v, p); return; }
#endif
    Store<ST, MT>::store(p, v, w, s);
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

// pointer to element idx of an array: a 32-bit byte offset from a base held in scalar registers,
// the form of the global instructions' scalar-base addressing (the host keeps every population and
// field array below 4 GiB, so idx counts rows as well: idx = row * ncell + cell)
// The pointer is typed global: the batched kernel loads its arrays' bases from the job table, and
// the compiler treats a pointer loaded from memory as generic (flat instructions, 64-bit vector
// addresses) unless the access itself names the global address space.
#if defined(__HIP_DEVICE_COMPILE__)
#define IPL_GLOBAL __attribute__((address_space(1)))
#else
#define IPL_GLOBAL
#endif
template <typename T> __device__ inline IPL_GLOBAL T* elem(T* row, uint32_t idx) {
    T* p = reinterpret_cast<T*>(reinterpret_cast<typename std::conditional<std::is_const<T>::value, const char, char>::type*>(row) +
                                (uint64_t)(idx * (uint32_t)sizeof(T)));
#ifdef IPL_CPU_EMULATION
    ipl_emu::check_access(row, p, sizeof(T));  // the CPU test build checks every access
#endif
    return (IPL_GLOBAL T*)p;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  Line L of the
// grid goes to the XCD that owns a contiguous block of lines, so the neighbour lines a line
// gathers from (and pushes to) are mostly processed by the same XCD at the same time.
constexpr int kXcds = 8;
__device__ inline int xcd_line(int b, int lines) {
    const int q = lines / kXcds, r = lines % kXcds;
    const int x = b % kXcds, slot = b / kXcds;
    return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

template <int S> struct LineAxes {
    static constexpr int o1 = S == 0 ? 1 : 0, o2 = S == 2 ? 1 : 2;
};

// The workgroup's line (c1, c2) and its eight neighbour lines, all wave-uniform: base<da, db>() is
// the first cell of line (c1+da, c2+db), wrapped on a periodic axis and the workgroup's own line
// where that line would leave the box (the ghost copies its owner, basic.cu:221-225); out1<da>() /
// out2<db>() flag the displacements that leave the box through a wall.  Named members only, no
// arrays: a selection between array elements becomes a load from private memory.
struct LineNb {
    uint32_t bmm, bm0, bmp, b0m, b00, b0p, bpm, bp0, bpp;
    bool o1m, o1p, o2m, o2p;
    int c1, c2;
    template <int DA, int DB> __device__ uint32_t base() const {
        if constexpr (DA < 0) return DB < 0 ? bmm : (DB == 0 ? bm0 : bmp);
        else if constexpr (DA == 0) return DB < 0 ? b0m : (DB == 0 ? b00 : b0p);
        else return DB < 0 ? bpm : (DB == 0 ? bp0 : bpp);
    }
    template <int DA> __device__ bool out1() const { return DA < 0 ? o1m : (DA > 0 ? o1p : false); }
    template <int DB> __device__ bool out2() const { return DB < 0 ? o2m : (DB > 0 ? o2p : false); }
};

template <int S>
__device__ inline LineNb line_neighbours(const FGeom& G, int c1, int c2) {
    constexpr int o1 = LineAxes<S>::o1, o2 = LineAxes<S>::o2;
    const int n1 = G.n[o1], n2 = G.n[o2];
    const uint32_t ns = (uint32_t)G.ns;
    LineNb nb;
    nb.c1 = c1; nb.c2 = c2;
    auto step = [&](int c, int d, int n, int per, bool& out) {
        int a = c + d;
        out = false;
        if (a < 0 || a >= n) {
            if (per) a = a < 0 ? a + n : a - n;
            else { a = c; out = true; }
        }
        return a;
    };
    const int a_m = step(c1, -1, n1, G.periodic[o1], nb.o1m), a_p = step(c1, 1, n1, G.periodic[o1], nb.o1p);
    const int b_m = step(c2, -1, n2, G.periodic[o2], nb.o2m), b_p = step(c2, 1, n2, G.periodic[o2], nb.o2p);
    auto lb = [&](int a, int b) { return (uint32_t)(a * n2 + b) * ns; };
    nb.bmm = lb(a_m, b_m); nb.bm0 = lb(a_m, c2); nb.bmp = lb(a_m, b_p);
    nb.b0m = lb(c1, b_m);  nb.b00 = lb(c1, c2);  nb.b0p = lb(c1, b_p);
    nb.bpm = lb(a_p, b_m); nb.bp0 = lb(a_p, c2); nb.bpp = lb(a_p, b_p);
    return nb;
}

// Writes the transported value of direction K whose source is the cell at S-index si of the
// workgroup's line; the destination is that cell + c_K.
template <typename ST, typename MT, int Q, int K, int S, bool NT>
__device__ inline void push(const FGeom& G, const LineNb& nb, ST* __restrict__ dst, int si, MT value) {
    constexpr int o1 = LineAxes<S>::o1, o2 = LineAxes<S>::o2;
    constexpr int ka = comp<Q>(K, o1), kb = comp<Q>(K, o2), ks = comp<Q>(K, S);
    const int ns = G.ns;
    const uint32_t NC = (uint32_t)G.ncell;
    // the destination line leaves the box: wave-uniform
    const bool out1 = nb.template out1<ka>(), out2 = nb.template out2<kb>();
    // destination S-index: per work-item
    int xs = si + ks;
    bool sout = false;
    if (ks != 0) {
        if (G.periodic[S]) xs = xs < 0 ? xs + ns : (xs >= ns ? xs - ns : xs);
        else sout = xs < 0 || xs >= ns;
    }
    constexpr double wk = Lattice<Q>::w(K);  // bounce-back and reflection keep |c|, hence the weight
    constexpr int kopp = opposite<Q>(K);
    const uint32_t own = nb.b00 + (uint32_t)si;
    if (!out1 && !out2) {
        const uint32_t lbase = nb.template base<ka, kb>();
        if (!sout) {
            store_pop<NT, ST, MT>(elem(dst, K * NC + lbase + (uint32_t)xs), value, wk, G.scale);
        } else if (G.wall_bc == IPL_BC_BOUNCE_BACK) {
            store_pop<NT, ST, MT>(elem(dst, kopp * NC + own), value, wk, G.scale);
        } else {
            constexpr int kr = reflect_dir<Q>(K, 1 << S);
            store_pop<NT, ST, MT>(elem(dst, kr * NC + lbase + (uint32_t)si), value, wk, G.scale);
        }
        return;
    }
    // boundary lines only
    if (G.wall_bc == IPL_BC_BOUNCE_BACK) {
        store_pop<NT, ST, MT>(elem(dst, kopp * NC + own), value, wk, G.scale);
        return;
    }
    // slip: reflect the destination and the direction across every wall crossed; the crossed
    // coordinate stays the source's, which lies on that wall
    // (recomputed from the coordinates: a selection between table members would move the table
    // to private memory)
    const int n1 = G.n[o1], n2 = G.n[o2];
    int y1 = nb.c1, y2 = nb.c2;
    if (!out1) { y1 += ka; y1 = y1 < 0 ? y1 + n1 : (y1 >= n1 ? y1 - n1 : y1); }
    if (!out2) { y2 += kb; y2 = y2 < 0 ? y2 + n2 : (y2 >= n2 ? y2 - n2 : y2); }
    const uint32_t lbase = (uint32_t)(y1 * n2 + y2) * (uint32_t)ns;
    const int ys = sout ? si : xs;
    const int out = (out1 ? 1 << o1 : 0) | (out2 ? 1 << o2 : 0) | (sout ? 1 << S : 0);
    constexpr int r1 = reflect_dir<Q>(K, 1), r2 = reflect_dir<Q>(K, 2), r3 = reflect_dir<Q>(K, 3);
    constexpr int r4 = reflect_dir<Q>(K, 4), r5 = reflect_dir<Q>(K, 5), r6 = reflect_dir<Q>(K, 6);
    constexpr int r7 = reflect_dir<Q>(K, 7);
    const int kr = out == 1 ? r1 : out == 2 ? r2 : out == 3 ? r3 : out == 4 ? r4 : out == 5 ? r5 : out == 6 ? r6 : r7;
    store_pop<NT, ST, MT>(elem(dst, (uint32_t)kr * NC + lbase + (uint32_t)ys), value, wk, G.scale);
}

// Slab flux, flux update and push of one population set (flow or scalar) of a line.
// P: collided populations of this work-item's cell.  lds: [ns][SL] of the workgroup, one row of
// the S-directions' values per cell (SL = their count + 1, an odd stride for the LDS banks), so a
// direction's slot is an immediate offset from the cell's row.
template <typename ST, typename MT, int Q, int NPTS, int S, bool FLUX, bool NT>
__device__ inline void transport(const FGeom& G, const LineNb& nb, const MT (&Wp)[NPTS], const MT (&Wm)[NPTS], MT rn,
                                 ST* __restrict__ dst, const MT (&P)[Q], MT* lds, int i, bool active) {
    const int ns = G.ns;
    constexpr int SL = sdir_count<Q, S>() + 1;
    if constexpr (!FLUX) {
        // uniform grid: exact streaming (propagation.cu:2-12).  Opposite S-directions (k, kb_) are
        // stored as a pair at this cell's S-index, so a wavefront writes whole lines: row k of the
        // line displaced by (ka, kb) receives P_k of the cell before (i - 1, through LDS), row kb_
        // of the opposite line P_kb of the cell after.  The first (last) cell of a walled line has
        // no cell before (after); its slot carries instead the wall redirect of its own P_kb (P_k):
        // bounce-back into its opposite row, slip into the row reflected across the S-wall
        // (set_wall_propagation, set_bound.cu:149-168).
        if (active)
            static_for<Q>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                if constexpr (comp<Q>(k, S) != 0) lds[i * SL + sdir_slot<Q, S>(k)] = P[k];
            });
        __syncthreads();
        if (!active) return;
        constexpr int o1 = LineAxes<S>::o1, o2 = LineAxes<S>::o2;
        const uint32_t NC = (uint32_t)G.ncell;
        const int per = G.periodic[S];
        static_for<Q>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int sg = comp<Q>(k, S);
            constexpr int ka = comp<Q>(k, o1), kb = comp<Q>(k, o2);
            constexpr double wk = Lattice<Q>::w(k);
            if constexpr (sg == 0) {
                if (!nb.template out1<ka>() && !nb.template out2<kb>())
                    store_pop<NT, ST, MT>(elem(dst, k * NC + nb.template base<ka, kb>() + (uint32_t)i), P[k], wk, G.scale);
                else
                    push<ST, MT, Q, k, S, NT>(G, nb, dst, i, P[k]);
            } else if constexpr (sg > 0) {
                constexpr int kb_ = opposite<Q>(k);
                constexpr int sk = sdir_slot<Q, S>(k), skb = sdir_slot<Q, S>(kb_);
                const bool lines_in = !nb.template out1<ka>() && !nb.template out2<kb>() &&
                                      !nb.template out1<-ka>() && !nb.template out2<-kb>();
                const bool last = !per && i == ns - 1, first = !per && i == 0;
                const int im = wrap_or_clamp(i - 1, ns, per), ip = wrap_or_clamp(i + 1, ns, per);
                if (lines_in) {
                    const MT Uk = lds[im * SL + sk], Ukb = lds[ip * SL + skb];
                    const uint32_t fwd = nb.template base<ka, kb>() + (uint32_t)i;
                    const uint32_t bwd = nb.template base<-ka, -kb>() + (uint32_t)i;
                    const uint32_t own = nb.b00 + (uint32_t)i;
                    uint32_t a1, a2;
                    if (G.wall_bc == IPL_BC_BOUNCE_BACK) {
                        a1 = k * NC + (first ? own : fwd);
                        a2 = kb_ * NC + (last ? own : bwd);
                    } else {
                        constexpr int rk = reflect_dir<Q>(k, 1 << S), rkb = reflect_dir<Q>(kb_, 1 << S);
                        a1 = first ? rkb * NC + bwd : k * NC + fwd;
                        a2 = last ? rk * NC + fwd : kb_ * NC + bwd;
                    }
                    store_pop<NT, ST, MT>(elem(dst, a1), first ? P[kb_] : Uk, wk, G.scale);
                    store_pop<NT, ST, MT>(elem(dst, a2), last ? P[k] : Ukb, wk, G.scale);
                } else {
                    push<ST, MT, Q, k, S, NT>(G, nb, dst, i, P[k]);
                    push<ST, MT, Q, kb_, S, NT>(G, nb, dst, i, P[kb_]);
                }
            }
        });
        return;
    } else {
    // 1. collided values of the S-directions into LDS
    if (active)
        static_for<Q>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if constexpr (comp<Q>(k, S) != 0) lds[i * SL + sdir_slot<Q, S>(k)] = P[k];
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
                constexpr int slot = sdir_slot<Q, S>(k);
                MT acc = 0;
#pragma unroll
                for (int j = 0; j < NPTS; ++j) {
                    const int ii = wrap_or_clamp(i + (j - NPTS / 2) * sg, ns, per);
                    const MT term = (sg > 0 ? Wp[j] : Wm[j]) * lds[ii * SL + slot];
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
            if constexpr (comp<Q>(k, S) != 0) lds[i * SL + sdir_slot<Q, S>(k)] = Fv[k];
        });
    __syncthreads();
    if (!active) return;
    constexpr int o1 = LineAxes<S>::o1, o2 = LineAxes<S>::o2;
    const uint32_t NC = (uint32_t)G.ncell;
    static_for<Q>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr int sg = comp<Q>(k, S);
        constexpr int ka = comp<Q>(k, o1), kb = comp<Q>(k, o2);
        constexpr double wk = Lattice<Q>::w(k);
        if constexpr (sg == 0) {
            // no component along S: set_tmp returns P exactly (adv.cu:64-65,241); the destination
            // keeps the S-index
            if (!nb.template out1<ka>() && !nb.template out2<kb>())
                store_pop<NT, ST, MT>(elem(dst, k * NC + nb.template base<ka, kb>() + (uint32_t)i), P[k], wk, G.scale);
            else
                push<ST, MT, Q, k, S, NT>(G, nb, dst, i, P[k]);
        } else if constexpr (sg > 0) {
            // the pair (k, kb_) of opposite S-directions.  U of this cell is the transported value
            // of m = n - sg S, which lands back on this cell's S-index: Uk comes from m = i - 1,
            // Ukb from m = i + 1.  At the downstream S-wall of a direction the wall cell also
            // pushes the ghost's mirrored copy T (adv.cu:196-199 along S): the last cell for k, the
            // first for kb_.  The first cell has no Uk and the last no Ukb, so each work-item
            // stores exactly two values: (Uk or T of kb_) and (Ukb or T of k).
            constexpr int kb_ = opposite<Q>(k);
            constexpr int mk = mirror<Q>(k, S), mkb = mirror<Q>(kb_, S);
            constexpr int sk = sdir_slot<Q, S>(k), skb = sdir_slot<Q, S>(kb_);
            const int im = wrap_or_clamp(i - 1, ns, per), ip = wrap_or_clamp(i + 1, ns, per);
            const MT Uk = P[k] + (lds[im * SL + sk] * rn - Fv[k] * rn);
            const MT Ukb = P[kb_] + (lds[ip * SL + skb] * rn - Fv[kb_] * rn);
            const bool first = !per && i == 0, last = !per && i == ns - 1;
            const bool lines_in = !nb.template out1<ka>() && !nb.template out2<kb>() &&
                                  !nb.template out1<-ka>() && !nb.template out2<-kb>();
            if (lines_in) {
                const MT Tkb = P[mkb] + (Fv[kb_] * rn - Fv[mkb] * rn);
                const MT Tk = P[mk] + (Fv[k] * rn - Fv[mk] * rn);
                const uint32_t fwd = nb.template base<ka, kb>() + (uint32_t)i;
                const uint32_t bwd = nb.template base<-ka, -kb>() + (uint32_t)i;
                uint32_t a1, a2;
                if (G.wall_bc == IPL_BC_BOUNCE_BACK) {
                    // T of kb_ at the first cell returns as k into that cell; T of k at the last as kb_
                    a1 = k * NC + (first ? nb.b00 + (uint32_t)i : fwd);
                    a2 = kb_ * NC + (last ? nb.b00 + (uint32_t)i : bwd);
                } else {
                    // slip: reflected across the S-wall, the line offset kept
                    constexpr int rkb = reflect_dir<Q>(kb_, 1 << S), rk = reflect_dir<Q>(k, 1 << S);
                    a1 = first ? rkb * NC + bwd : k * NC + fwd;
                    a2 = last ? rk * NC + fwd : kb_ * NC + bwd;
                }
                store_pop<NT, ST, MT>(elem(dst, a1), first ? Tkb : Uk, wk, G.scale);
                store_pop<NT, ST, MT>(elem(dst, a2), last ? Tk : Ukb, wk, G.scale);
            } else {
                // boundary lines: the general redirect of push()
                if (!first) push<ST, MT, Q, k, S, NT>(G, nb, dst, im, Uk);
                if (!last) push<ST, MT, Q, kb_, S, NT>(G, nb, dst, ip, Ukb);
                if (last) push<ST, MT, Q, k, S, NT>(G, nb, dst, i, P[mk] + (Fv[k] * rn - Fv[mk] * rn));
                if (first) push<ST, MT, Q, kb_, S, NT>(G, nb, dst, i, P[mkb] + (Fv[kb_] * rn - Fv[mkb] * rn));
            }
        }
    });
    }
}

// A direction that represents its opposite pair (the rest direction stands alone).
template <int Q> constexpr bool is_rep(int k) { return k != 0 && k < opposite<Q>(k); }

// +1 when row r of the moment basis is even under c -> -c, -1 when it is odd, 0 otherwise
template <int Q> constexpr int row_parity(int r) {
    bool even = true, odd = true;
    for (int k = 0; k < Q; ++k) {
        const double a = moment<Q>(r, k), b = moment<Q>(r, opposite<Q>(k));
        even = even && a == b;
        odd = odd && a == -b;
    }
    return even ? 1 : (odd ? -1 : 0);
}

// first column among {0} and the pair representatives that contributes to row r
template <int Q> constexpr int first_term_col(int r) {
    for (int k = 0; k < Q; ++k)
        if ((k == 0 || is_rep<Q>(k)) && moment<Q>(r, k) != 0.0 && (k != 0 || row_parity<Q>(r) > 0)) return k;
    return Q;
}
// first row that contributes to column k with the given parity
template <int Q> constexpr int first_term_row(int k, int parity) {
    for (int r = 0; r < Q; ++r)
        if (row_parity<Q>(r) == parity && moment<Q>(r, k) != 0.0) return r;
    return Q;
}

template <typename MT> __device__ inline MT scaled(double c, MT x) {
    return c == 1.0 ? x : (c == -1.0 ? -x : MT(c) * x);
}

// DX x + DY y + DZ z over the nonzero components only
template <int DX, int DY, int DZ, typename MT> __device__ inline MT dot3(MT x, MT y, MT z) {
    if constexpr (DX != 0) {
        MT acc = scaled<MT>(DX, x);
        if constexpr (DY != 0) acc += scaled<MT>(DY, y);
        if constexpr (DZ != 0) acc += scaled<MT>(DZ, z);
        return acc;
    } else if constexpr (DY != 0) {
        MT acc = scaled<MT>(DY, y);
        if constexpr (DZ != 0) acc += scaled<MT>(DZ, z);
        return acc;
    } else if constexpr (DZ != 0) {
        return scaled<MT>(DZ, z);
    } else {
        return MT(0);
    }
}

// Equilibrium (equ_f, collision.cu:3-14), forcing term Fk (collision.cu:100-110) and the MRT /
// SRT collision (collision.cu:115-165) of one cell, written for opposite pairs (k, kb):
// feq_k = B + A and feq_kb = B - A with B = w (p + 4.5 t^2 - 1.5 u^2/c^2), A = 3 w t, and
// Fk_kb = -Fk_k, so the pair sums s = d_k + d_kb = f_k + f_kb - 2B and differences
// a = d_k - d_kb = f_k - f_kb - 2A + Fk_k of d = f - feq + Fk/2 are formed directly.  The basis
// rows are even or odd under c -> -c, so even moments take the sums and odd moments the
// differences, and M^-1 = M^T diag(1/|row|^2) (orthogonal rows, norms folded into the rates on
// the host) gives f*_k = f_k - (E + O) + Fk_k, f*_kb = f_kb - (E - O) - Fk_k with E, O the even
// and odd parts of column k.  f is replaced by f*.
template <typename MT, int Q>
__device__ inline void collide_pairs(MT (&f)[Q], MT p, MT u, MT v, MT w, MT u2c, MT Fx, MT Fy, MT Fz, MT fkc, MT ic,
                                     const FCoef<MT>& P, bool les, MT itau_c) {
    MT sv[Q], av[Q], fkv[Q];
    sv[0] = f[0] - MT(Lattice<Q>::w(0)) * (p - MT(1.5) * u2c);
    static_for<Q>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (is_rep<Q>(k)) {
            constexpr int kb = opposite<Q>(k);
            constexpr Dir d = Lattice<Q>::c(k);
            constexpr double wk = Lattice<Q>::w(k);
            const MT t = dot3<d.x, d.y, d.z>(u, v, w) * ic;
            const MT B = MT(wk) * (p + MT(4.5) * t * t - MT(1.5) * u2c);
            const MT A = MT(3.0 * wk) * t;
            const MT eF = dot3<d.x, d.y, d.z>(Fx, Fy, Fz);
            fkv[k] = MT(wk) * fkc * eF;
            sv[k] = (f[k] + f[kb]) - MT(2) * B;
            av[k] = (f[k] - f[kb]) - MT(2) * A + fkv[k];
        }
    });
    if (P.collision == IPL_COLL_SRT) {
        const MT itau = les ? itau_c : P.itau;
        f[0] = f[0] - sv[0] * itau;
        static_for<Q>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if constexpr (is_rep<Q>(k)) {
                constexpr int kb = opposite<Q>(k);
                const MT dk = MT(0.5) * (sv[k] + av[k]), dkb = MT(0.5) * (sv[k] - av[k]);
                f[k] = f[k] - dk * itau + fkv[k];
                f[kb] = f[kb] - dkb * itau - fkv[k];
            }
        });
        return;
    }
    MT m[Q];
    static_for<Q>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        constexpr int par = row_parity<Q>(r);
        static_assert(par != 0, "moment rows must be even or odd");
        MT acc = 0;
        static_for<Q>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if constexpr (k == 0 || is_rep<Q>(k)) {
                constexpr double mk = moment<Q>(r, k);
                if constexpr (mk != 0.0 && (k != 0 || par > 0)) {
                    const MT term = scaled<MT>(mk, par > 0 ? sv[k] : av[k]);
                    if constexpr (k == first_term_col<Q>(r)) acc = term;
                    else acc += term;
                }
            }
        });
        if constexpr (is_stress<Q>(r)) m[r] = acc * (les ? itau_c * P.inv_n2[r] : P.rnorm[r]);
        else m[r] = acc * P.rnorm[r];
    });
    static_for<Q>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (k == 0 || is_rep<Q>(k)) {
            MT E = 0, O = 0;
            static_for<Q>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                constexpr double mk = moment<Q>(r, k);
                constexpr int par = row_parity<Q>(r);
                if constexpr (mk != 0.0) {
                    const MT term = scaled<MT>(mk, m[r]);
                    if constexpr (par > 0) {
                        if constexpr (r == first_term_row<Q>(k, 1)) E = term; else E += term;
                    } else {
                        if constexpr (r == first_term_row<Q>(k, -1)) O = term; else O += term;
                    }
                }
            });
            if constexpr (k == 0) {
                f[0] = f[0] - E;
            } else {
                constexpr int kb = opposite<Q>(k);
                f[k] = f[k] - (E + O) + fkv[k];
                f[kb] = f[kb] - (E - O) - fkv[k];
            }
        }
    });
}

// index of the neighbour of the work-item's cell along axis A at offset +-1: the wall cell
// itself beyond a wall (the ghost copies its owner, basic.cu:221-225), wrapped on a periodic axis
template <int S, int A, int OFF>
__device__ inline uint32_t nb_index(const FGeom& G, const LineNb& nb, int i) {
    constexpr int o1 = LineAxes<S>::o1;
    if constexpr (A == S) return nb.b00 + (uint32_t)wrap_or_clamp(i + OFF, G.ns, G.periodic[S]);
    else if constexpr (A == o1) return nb.template base<OFF, 0>() + (uint32_t)i;
    else return nb.template base<0, OFF>() + (uint32_t)i;
}

// wall_function (collision.cu:28-55) for one wall: u_tau of Spalding's law by 15 Newton steps
// from the cell's speed and its distance y_dist to the wall, then the shear on the two
// tangential components (vt1, vt2; vn is the normal one) into Ft1, Ft2
template <typename MT>
__device__ inline void spalding_shear(MT vt1, MT vt2, MT vn, MT y_dist, MT wid_x, MT wid_y, MT dz, MT rho, MT nu,
                                      MT& Ft1, MT& Ft2) {
    const MT volume = wid_x * wid_y * dz;
    const MT abs_u = sqrt(vt1 * vt1 + vt2 * vt2 + vn * vn);
    const MT eta = abs_u * MT(1e-7) + MT(1e-20);
    const MT kap = MT(0.41), Beta = MT(5.5);
    const MT area = volume / (y_dist * MT(2.0));
    const MT ekb = exp(-kap * Beta);
    MT utau = abs_u * MT(0.11);
    for (int i = 0; i < 15; ++i) {
        const MT tmp = kap * abs_u / (utau + eta);
        const MT et = exp(tmp);
        const MT eps = -((abs_u / (utau + eta) - y_dist * utau / nu +
                          ekb * (et - MT(1.0) - tmp - tmp * tmp / MT(2.0) - tmp * tmp * tmp / MT(6.0))) /
                         (-y_dist / nu - abs_u / (utau * utau + eta) +
                          ekb * tmp / (utau + eta) * (-et + MT(1.0) + tmp + tmp * tmp / MT(2.0)) + eta));
        utau += eps;
    }
    const MT shear = utau * utau * rho * (area / volume) / (abs_u + eta);
    Ft1 += -vt1 * shear;
    Ft2 += -vt2 * shear;
}

// the driver's six wall-function calls (LockExchange.cu:214-219) for the cell at index ii of the
// line (c1, c2): faces +x, -x, +z, -z, +y, -y, each on the cells next to a (non-periodic) wall
template <typename MT, int S, bool D3>
__device__ inline void wall_shear(const FGeom& G, const FTables<MT>& tb, const FCoef<MT>& P, int ii, int c1, int c2,
                                  MT u, MT v, MT w, MT rho, MT& Fx, MT& Fy, MT& Fz) {
    constexpr int o1 = LineAxes<S>::o1;
    int c[3];
    c[S] = ii;
    c[o1] = c1;
    c[LineAxes<S>::o2] = c2;
    const MT wx = *elem(tb.wid[0], (uint32_t)c[0]), wy = *elem(tb.wid[1], (uint32_t)c[1]);
    const MT dz = P.dz;
    if (!G.periodic[0]) {
        const int hits = (c[0] == G.n[0] - 1) + (c[0] == 0);
        for (int h = 0; h < hits; ++h) spalding_shear<MT>(v, w, u, wx / MT(2.0), wx, wy, dz, rho, P.wall_nu, Fy, Fz);
    }
    if (!G.periodic[2]) {
        const int hits = (c[2] == G.n[2] - 1) + (c[2] == 0);
        for (int h = 0; h < hits; ++h) spalding_shear<MT>(u, v, w, dz / MT(2.0), wx, wy, dz, rho, P.wall_nu, Fx, Fy);
    }
    if (D3 && !G.periodic[1]) {
        const int hits = (c[1] == G.n[1] - 1) + (c[1] == 0);
        for (int h = 0; h < hits; ++h) spalding_shear<MT>(u, w, v, wy / MT(2.0), wx, wy, dz, rho, P.wall_nu, Fx, Fz);
    }
}

// (q[+] - q[-]) / (pos[+] - pos[-] + dz/1000) along axis A (Force, collision.cu:66-69); the
// reciprocal, with the period added across a periodic seam, is a host table
template <typename MT, int S, int A>
__device__ inline MT grad_axis(const FGeom& G, const LineNb& nb, const FTables<MT>& tb, const MT* __restrict__ q,
                               int c1, int c2, int i) {
    constexpr int o1 = LineAxes<S>::o1;
    const int x = A == S ? i : (A == o1 ? c1 : c2);
    return (*elem(q, nb_index<S, A, 1>(G, nb, i)) - *elem(q, nb_index<S, A, -1>(G, nb, i))) * *elem(tb.invd[A], (uint32_t)x);
}

// rho of step n of a salinity cell from the sum of its g (update_rho, collision.cu:244-254)
template <typename MT>
__device__ inline MT rho_from_sum(const FCoef<MT>& P, MT sc) { return P.rho_L + sc * P.beta_s; }

// rho of a phase-field cell from its phi (update_rho; the salinity term is zero)
template <typename MT>
__device__ inline MT rho_from_phi(const FCoef<MT>& P, MT phi) { return P.rho_L + P.drho * phi; }

// the sum of the g rows of cell m (S or phi), in the order every other pass sums them
template <typename ST, typename MT, int QS>
__device__ inline MT scalar_from_g(const FGeom& G, const ST* __restrict__ gsrc, uint32_t m) {
    const uint32_t NC = (uint32_t)G.ncell;
    MT sc = Store<ST, MT>::load(elem(gsrc, m), Lattice<QS>::w(0), G.inv_scale);
#pragma unroll
    for (int k = 1; k < QS; ++k) sc += Store<ST, MT>::load(elem(gsrc, k * NC + m), Lattice<QS>::w(k), G.inv_scale);
    return sc;
}

// rho of a salinity cell m from its g rows
template <typename ST, typename MT, int QS>
__device__ inline MT rho_from_g(const FGeom& G, const FCoef<MT>& P, const ST* __restrict__ gsrc, uint32_t m) {
    return rho_from_sum(P, scalar_from_g<ST, MT, QS>(G, gsrc, m));
}

// phi of the two neighbours of the work-item's cell along axis A (phase field, no scalar-field
// pass): along the line from the workgroup's LDS row of phi, across lines from the neighbour
// cells' g rows
template <typename ST, typename MT, int QS, int S, int A>
__device__ inline void phi_pair_inline(const FGeom& G, const LineNb& nb, const ST* __restrict__ gsrc,
                                       const MT* lds_s, int i, MT& hi, MT& lo) {
    if constexpr (A == S) {
        hi = lds_s[wrap_or_clamp(i + 1, G.ns, G.periodic[S])];
        lo = lds_s[wrap_or_clamp(i - 1, G.ns, G.periodic[S])];
    } else {
        hi = scalar_from_g<ST, MT, QS>(G, gsrc, nb_index<S, A, 1>(G, nb, i));
        lo = scalar_from_g<ST, MT, QS>(G, gsrc, nb_index<S, A, -1>(G, nb, i));
    }
}

// grad rho along axis A without the scalar-field pass: along the line from the workgroup's LDS
// row of rho, across lines from the neighbour cells' g rows (their lines are processed by the
// neighbouring workgroups of the same XCD at about the same time, so the rows come from L2)
template <typename ST, typename MT, int QS, int S, int A>
__device__ inline MT grad_rho_inline(const FGeom& G, const LineNb& nb, const FTables<MT>& tb, const FCoef<MT>& P,
                                     const ST* __restrict__ gsrc, const MT* lds_r, int c1, int c2, int i) {
    constexpr int o1 = LineAxes<S>::o1;
    const int x = A == S ? i : (A == o1 ? c1 : c2);
    MT hi, lo;
    if constexpr (A == S) {
        hi = lds_r[wrap_or_clamp(i + 1, G.ns, G.periodic[S])];
        lo = lds_r[wrap_or_clamp(i - 1, G.ns, G.periodic[S])];
    } else {
        hi = rho_from_g<ST, MT, QS>(G, P, gsrc, nb_index<S, A, 1>(G, nb, i));
        lo = rho_from_g<ST, MT, QS>(G, P, gsrc, nb_index<S, A, -1>(G, nb, i));
    }
    return (hi - lo) * *elem(tb.invd[A], (uint32_t)x);
}

// One line of the step: work-item i of the line's slot owns cell i of line L.
// EXT: the wall function and LES are compiled in (they add registers to every step otherwise)
// line_ok: the work-item's slot holds a line of the job (else L is another line of the workgroup,
// whose cells the work-item reads but never writes, and it only takes part in the barriers).
template <typename ST, typename MT, int Q, int QS, int NPTS, int S, bool FLUX, bool EXT, bool NT>
__device__ __forceinline__ void fused_line(const FGeom& G, const FTables<MT>& tb, const FBufs<ST, MT>& b, const FCoef<MT>& P,
                                           int L, int i, bool line_ok, MT* lds_f) {
    MT* lds_g = lds_f + (sdir_count<Q, S>() + 1) * G.ns;
    constexpr int o1 = LineAxes<S>::o1, o2 = LineAxes<S>::o2;
    const bool active = line_ok && i < G.ns;
    const int n2 = G.n[o2];
    const int c1 = L / n2, c2 = L - (L / n2) * n2;
    const int ii = active ? i : 0;
    const uint32_t n = (uint32_t)L * (uint32_t)G.ns + (uint32_t)ii;
    const uint32_t NC = (uint32_t)G.ncell;
    const ST* __restrict__ fsrc = b.f_src;
    const ST* __restrict__ gsrc = b.g_src;
    const MT c = P.c, ic = P.inv_c, ic2 = P.inv_c2;
    const LineNb nb = line_neighbours<S>(G, c1, c2);

    MT Pf[Q], Pg[QS > 0 ? QS : 1];
    MT g[QS > 0 ? QS : 1];
    MT* lds_r = lds_g;  // [ns] rho (inline_rho) or phi (inline_phi) of the line, after the g rows
    if constexpr (QS > 0) lds_r += (sdir_count<QS, S>() + 1) * G.ns;
    if constexpr (QS > 0) {
        if (active) {
#pragma unroll
            for (int k = 0; k < QS; ++k) g[k] = load_pop<NT, ST, MT>(elem(gsrc, k * NC + n), Lattice<QS>::w(k), G.inv_scale);
        }
        if (G.inline_rho || G.inline_phi) {  // uniform over the launch: every work-item reaches the barrier
            if (active) {
                MT sc = g[0];
#pragma unroll
                for (int k = 1; k < QS; ++k) sc += g[k];
                lds_r[i] = G.inline_phi ? sc : rho_from_sum(P, sc);
            }
            __syncthreads();
        }
    }
    if (active) {
#pragma unroll
        for (int k = 0; k < Q; ++k) Pf[k] = load_pop<NT, ST, MT>(elem(fsrc, k * NC + n), Lattice<Q>::w(k), G.inv_scale);
        // ---- macroscopic fields of step n (update_rho, collision.cu:244-254) ----
        MT p = Pf[0], ex = 0, ey = 0, ez = 0;
        static_for<Q - 1>([&](auto kc) {
            constexpr int k = decltype(kc)::value + 1;
            constexpr Dir d = Lattice<Q>::c(k);
            p += Pf[k];
            if constexpr (d.x > 0) ex += Pf[k]; else if constexpr (d.x < 0) ex -= Pf[k];
            if constexpr (d.y > 0) ey += Pf[k]; else if constexpr (d.y < 0) ey -= Pf[k];
            if constexpr (d.z > 0) ez += Pf[k]; else if constexpr (d.z < 0) ez -= Pf[k];
        });
        MT sc = 0;
        if constexpr (QS > 0) {
            sc = g[0];
#pragma unroll
            for (int k = 1; k < QS; ++k) sc += g[k];
        }
        const MT S_ = G.scalar == IPL_SCALAR_SALINITY ? sc : MT(0);
        const MT phi = G.scalar == IPL_SCALAR_PHASE_FIELD ? sc : MT(0);
        const MT rho = P.rho_L + P.drho * phi + S_ * P.beta_s;
        const MT irho = MT(1) / rho;
        // u = sum e c f + F_old dt / (2 rho); after an immersed-boundary sub-step u = sum e c f
        // (update_velIBM: no force term).  Only the loads are conditional, so the code stays one path.
        MT F0 = MT(0), F1 = MT(0), F2 = MT(0);
        if (G.half_force && !G.zero_force) {
            F0 = *elem(b.F_old, n);
            F1 = *elem(b.F_old, NC + n);
            F2 = *elem(b.F_old, 2 * NC + n);
        }
        const MT hr = P.half_dt * irho;
        const MT u = ex * P.uc + F0 * hr;
        const MT v = ey * P.uc + F1 * hr;
        const MT w = ez * P.uc + F2 * hr;
        // ---- Force (collision.cu:57-112) ----
        constexpr bool d3 = Lattice<Q>::dim == 3;
        MT Fx = 0, Fy = 0, Fz = 0;
        if constexpr (EXT) {
            if (G.wall_fn) wall_shear<MT, S, d3>(G, tb, P, ii, c1, c2, u, v, w, rho, Fx, Fy, Fz);
        }
        MT gpx = 0, gpy = 0, gpz = 0;  // grad phi (col_PF reuses it)
        MT grx = 0, grz = 0;           // grad rho of a 2D phase field (inline_phi)
        if constexpr (QS == 9 && !d3) {
            if (G.inline_phi) {
                // the same terms from the neighbours' phi formed here (x across lines, z along S
                // or across); rho of a neighbour is rho_L + (rho_H - rho_L) phi
                MT xh, xl, zh, zl;
                phi_pair_inline<ST, MT, QS, S, 0>(G, nb, gsrc, lds_r, ii, xh, xl);
                phi_pair_inline<ST, MT, QS, S, 2>(G, nb, gsrc, lds_r, ii, zh, zl);
                const MT idx = *elem(tb.invd[0], (uint32_t)(S == 0 ? ii : (o1 == 0 ? c1 : c2)));
                const MT idz = *elem(tb.invd[2], (uint32_t)(S == 2 ? ii : (o1 == 2 ? c1 : c2)));
                gpx = (xh - xl) * idx;
                gpz = (zh - zl) * idz;
                const MT lap = (xh - MT(2) * phi + xl) * *elem(tb.inv_wx2, (uint32_t)(S == 0 ? ii : (o1 == 0 ? c1 : c2))) +
                               (zh - MT(2) * phi + zl) * P.inv_dz2;
                const MT st = P.st_beta * phi * (phi - MT(1.0)) * (phi - MT(0.5)) - P.kap * lap;
                Fx += st * gpx; Fz += st * gpz;
                grx = (rho_from_phi(P, xh) - rho_from_phi(P, xl)) * idx;
                grz = (rho_from_phi(P, zh) - rho_from_phi(P, zl)) * idz;
            }
        }
        if (G.grad_phi && !G.inline_phi) {
            gpx = grad_axis<MT, S, 0>(G, nb, tb, b.phi_n, c1, c2, ii);
            if constexpr (d3) gpy = grad_axis<MT, S, 1>(G, nb, tb, b.phi_n, c1, c2, ii);
            gpz = grad_axis<MT, S, 2>(G, nb, tb, b.phi_n, c1, c2, ii);
            // x term over delX^2, the term of directions 2/4 (z in 2D, y in 3D) over dz^2
            // (collision.cu:70-71)
            constexpr int a2 = d3 ? 1 : 2;
            const MT lap = (*elem(b.phi_n, nb_index<S, 0, 1>(G, nb, ii)) - MT(2) * phi +
                            *elem(b.phi_n, nb_index<S, 0, -1>(G, nb, ii))) * *elem(tb.inv_wx2, (uint32_t)(S == 0 ? ii : (o1 == 0 ? c1 : c2))) +
                           (*elem(b.phi_n, nb_index<S, a2, 1>(G, nb, ii)) - MT(2) * phi +
                            *elem(b.phi_n, nb_index<S, a2, -1>(G, nb, ii))) * P.inv_dz2;
            const MT st = P.st_beta * phi * (phi - MT(1.0)) * (phi - MT(0.5)) - P.kap * lap;
            Fx += st * gpx; Fy += st * gpy; Fz += st * gpz;
        }
        if (P.boussinesq) Fz += -S_ * P.beta_s * P.gravity;
        else Fz += -rho * P.gravity;
        if (G.grad_rho && G.inline_phi) {
            const MT pp = -p * P.press;
            Fx += pp * grx;
            Fz += pp * grz;
        }
        if (G.grad_rho && !G.inline_phi) {
            const MT pp = -p * P.press;
            if constexpr (QS > 0) {
                if (G.inline_rho) {
                    Fx += pp * grad_rho_inline<ST, MT, QS, S, 0>(G, nb, tb, P, gsrc, lds_r, c1, c2, ii);
                    if constexpr (d3) Fy += pp * grad_rho_inline<ST, MT, QS, S, 1>(G, nb, tb, P, gsrc, lds_r, c1, c2, ii);
                    Fz += pp * grad_rho_inline<ST, MT, QS, S, 2>(G, nb, tb, P, gsrc, lds_r, c1, c2, ii);
                }
            }
            if (!G.inline_rho) {
                Fx += pp * grad_axis<MT, S, 0>(G, nb, tb, b.rho_n, c1, c2, ii);
                if constexpr (d3) Fy += pp * grad_axis<MT, S, 1>(G, nb, tb, b.rho_n, c1, c2, ii);
                Fz += pp * grad_axis<MT, S, 2>(G, nb, tb, b.rho_n, c1, c2, ii);
            }
        }
        if (!G.zero_force) { *elem(b.F_new, n) = Fx; *elem(b.F_new, NC + n) = Fy; *elem(b.F_new, 2 * NC + n) = Fz; }
        // ---- equ_f, forcing term and collision ----
        const MT u2c = (u * u + v * v + w * w) * ic2;
        const MT fkc = P.fk_coef * irho;
        MT itau_c = P.itau, om_c = P.om_s;  // per-cell rates when an LES pass sets tau, tau_s
        bool les = false;
        if constexpr (EXT) {
            les = G.les != 0;
            if (les) {
                itau_c = MT(1) / *elem(b.tau, n);
                if constexpr (QS > 0) om_c = MT(1) - MT(1) / *elem(b.taus, n);
            }
        }
        // ---- scalar collision (before the flow's: g and its collision are then no longer live
        // while the flow collides) ----
        if constexpr (QS > 0) {
            if (G.scalar == IPL_SCALAR_SALINITY) {  // col_g_reg (collision.cu:167-188)
                constexpr bool linear = QS == 7;
                const MT cs2inv = linear ? MT(4.0) : MT(3.0);
                MT geq[QS];
                MT jx = 0, jy = 0, jz = 0;
                static_for<QS>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    constexpr Dir d = Lattice<QS>::c(k);
                    const MT t = dot3<d.x, d.y, d.z>(u, v, w) * ic;
                    MT e = MT(1.0) + cs2inv * t;
                    if constexpr (!linear) e += MT(4.5) * t * t - MT(1.5) * u2c;
                    geq[k] = MT(Lattice<QS>::w(k)) * S_ * e;
                    const MT gn = g[k] - geq[k];
                    if constexpr (d.x > 0) jx += gn; else if constexpr (d.x < 0) jx -= gn;
                    if constexpr (d.y > 0) jy += gn; else if constexpr (d.y < 0) jy -= gn;
                    if constexpr (d.z > 0) jz += gn; else if constexpr (d.z < 0) jz -= gn;
                });
                const MT om = om_c;
                static_for<QS>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    constexpr Dir d = Lattice<QS>::c(k);
                    const MT gneq = MT(Lattice<QS>::w(k)) * cs2inv * dot3<d.x, d.y, d.z>(jx, jy, jz);
                    Pg[k] = geq[k] + om * gneq;
                });
            } else {  // col_PF (collision.cu:190-223)
                const MT idt = P.inv_dt;
                const MT nabu = (phi * u - *elem(b.phiold, n) * *elem(b.uold, n)) * idt;
                const MT nabv = (phi * v - *elem(b.phiold, n) * *elem(b.vold, n)) * idt;
                const MT nabw = (phi * w - *elem(b.phiold, n) * *elem(b.wold, n)) * idt;
                *elem(b.phiold, n) = phi; *elem(b.uold, n) = u; *elem(b.vold, n) = v; *elem(b.wold, n) = w;
                const MT inrm = MT(1) / sqrt(gpx * gpx + gpy * gpy + gpz * gpz + MT(1e-8));
                const MT src = (MT(1.0) - MT(4.0) * (phi - MT(0.5)) * (phi - MT(0.5))) * P.inv_pfW * c * inrm;
                const MT nc = MT(3) * c * ic2;
                const MT dtc = P.dt;
                const MT itp = P.pf_itau, fco = P.pf_fcoef;
                static_for<QS>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    constexpr Dir d = Lattice<QS>::c(k);
                    constexpr double wk = Lattice<QS>::w(k);
                    const MT t = dot3<d.x, d.y, d.z>(u, v, w) * ic;
                    const MT geq = MT(wk) * phi * (MT(1.0) + MT(3.0) * t + MT(4.5) * t * t - MT(1.5) * u2c);
                    const MT Fkp = dtc * MT(wk) *
                                   (src * dot3<d.x, d.y, d.z>(gpx, gpy, gpz) + nc * dot3<d.x, d.y, d.z>(nabu, nabv, nabw));
                    Pg[k] = g[k] - (g[k] - geq - Fkp * fco) * itp;
                });
            }
        }
        collide_pairs<MT, Q>(Pf, p, u, v, w, u2c, Fx, Fy, Fz, fkc, ic, P, les, itau_c);
    }
    // ---- interpolated propagation of f and g (one set of slab weights for both) ----
    MT Wp[NPTS], Wm[NPTS], rn = MT(0);
    if constexpr (FLUX) {
#pragma unroll
        for (int j = 0; j < NPTS; ++j) {
            Wp[j] = *elem(tb.W, (uint32_t)(j * G.ns + ii));
            Wm[j] = *elem(tb.W, (uint32_t)((NPTS + j) * G.ns + ii));
        }
        rn = *elem(tb.rdel, (uint32_t)ii);
    }
    // the scalar's populations first, whose collided values were formed first, then the flow's
    if constexpr (QS > 0) transport<ST, MT, QS, NPTS, S, FLUX, NT>(G, nb, Wp, Wm, rn, b.g_dst, Pg, lds_g, ii, active);
    transport<ST, MT, Q, NPTS, S, FLUX, NT>(G, nb, Wp, Wm, rn, b.f_dst, Pf, lds_f, ii, active);
}

// The fused step of one or many simulations in one launch.  The workgroups of job j are blocks
// jobs[j].first_block .. + groups - 1, in the job's XCD order; each holds lpw consecutive lines
// of that order in slots of tpl work-items (whole wavefronts), and its buffer roles at this launch
// are b[step & 1] (the roles swap every step).  Jobs of one launch share the
// template arguments; each keeps its own geometry, tables, buffers and coefficients.  A context
// stepped alone launches this kernel with a table of one job, so a job runs the same instruction
// stream alone and in any batch, and its results are bit-identical either way.
//
// The table is read through the constant address space: the kernel never writes it, and loads
// the compiler knows to be invariant are issued as early, wide scalar loads (as kernel arguments
// are) rather than one by one next to their uses.
#if defined(__HIP_DEVICE_COMPILE__)
#define IPL_CONST_AS __attribute__((address_space(4)))
#else
#define IPL_CONST_AS
#endif
template <typename T> __device__ inline T load_invariant(const IPL_CONST_AS T* p) {
    static_assert(sizeof(T) % 4 == 0, "copied as 32-bit words");
    T v;
    const IPL_CONST_AS uint32_t* src = reinterpret_cast<const IPL_CONST_AS uint32_t*>(p);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = src[i];
    return v;
}

template <typename ST, typename MT, int Q, int QS, int NPTS, int S, bool FLUX, bool EXT, bool NT>
__global__ void __launch_bounds__(kMaxLine) IPL_FUSED_OCCUPANCY
k_fused_step(const BatchJob<ST, MT>* __restrict__ jobs, int njobs, int step) {
    IPL_DYNAMIC_LDS(MT, lds_f);
    const IPL_CONST_AS BatchJob<ST, MT>* cj = (const IPL_CONST_AS BatchJob<ST, MT>*)jobs;
    const int bid = (int)blockIdx.x;
    int lo = 0, hi = njobs - 1;  // the last job whose first block is at or before this one (uniform)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cj[mid].first_block <= bid) lo = mid;
        else hi = mid - 1;
    }
    const FGeom G = load_invariant(&cj[lo].G);
    const FTables<MT> tb = load_invariant(&cj[lo].tb);
    const FBufs<ST, MT> b = load_invariant(&cj[lo].b[step & 1]);
    const FCoef<MT> P = load_invariant(&cj[lo].P);
    // tpl is a multiple of 64: a wavefront's work-items share their slot, so the line and all
    // that follows from it (its neighbour lines, its walls) stay in scalar registers
    const int slot = __builtin_amdgcn_readfirstlane((int)threadIdx.x / G.tpl);
    const int i = (int)threadIdx.x - slot * G.tpl;
    __builtin_assume(i >= 0 && i < kMaxLine);
    const int first = xcd_line(bid - cj[lo].first_block, G.groups) * G.lpw;
    const bool line_ok = slot < G.lpw && first + slot < (int)G.lines;
    fused_line<ST, MT, Q, QS, NPTS, S, FLUX, EXT, NT>(G, tb, b, P, line_ok ? first + slot : first, i, line_ok,
                                                  lds_f + (line_ok ? slot : 0) * G.lds_line);
}

// rho (and phi) of step n for the gradients of the next fused step; V consecutive cells per
// work-item with V-wide loads when the rows allow it (ncell a multiple of V).  The host takes
// V = 4 on large grids and fewer cells per work-item where that would leave the chip with too few
// workgroups to hide the load latency (the 1024 x 256 phase-field wave has 262,144 cells).
template <typename T, int V> struct alignas(V * sizeof(T)) VecN { T v[V]; };

template <typename ST, typename MT, int QS, int V>
__global__ void k_scalar_fields(FGeom G, const ST* g_src, MT* rho_n, MT* phi_n, FCoef<MT> P) {
    const int64_t n0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) * V;
    if (n0 >= G.ncell) return;
    const int64_t NC = G.ncell;
    MT sc[V];
#pragma unroll
    for (int j = 0; j < V; ++j) sc[j] = MT(0);
    const bool vec = (NC % V) == 0;
    if constexpr (QS > 0) {
#pragma unroll
        for (int k = 0; k < QS; ++k) {
            ST x[V];
            if (vec) {
                const VecN<ST, V> v = *reinterpret_cast<const VecN<ST, V>*>(g_src + k * NC + n0);
#pragma unroll
                for (int j = 0; j < V; ++j) x[j] = v.v[j];
            } else {
#pragma unroll
                for (int j = 0; j < V; ++j) x[j] = n0 + j < NC ? g_src[k * NC + n0 + j] : ST{};
            }
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const MT y = Store<ST, MT>::load(&x[j], Lattice<QS>::w(k), G.inv_scale);
                sc[j] = k == 0 ? y : sc[j] + y;
            }
        }
    }
    VecN<MT, V> r, ph;
#pragma unroll
    for (int j = 0; j < V; ++j) {
        const MT S = G.scalar == IPL_SCALAR_SALINITY ? sc[j] : MT(0);
        const MT phi = G.scalar == IPL_SCALAR_PHASE_FIELD ? sc[j] : MT(0);
        r.v[j] = P.rho_L + P.drho * phi + S * P.beta_s;
        ph.v[j] = phi;
    }
    if (vec) {
        if (G.grad_rho) *reinterpret_cast<VecN<MT, V>*>(rho_n + n0) = r;
        if (G.grad_phi) *reinterpret_cast<VecN<MT, V>*>(phi_n + n0) = ph;
    } else {
#pragma unroll
        for (int j = 0; j < V; ++j)
            if (n0 + j < NC) {
                if (G.grad_rho) rho_n[n0 + j] = r.v[j];
                if (G.grad_phi) phi_n[n0 + j] = ph.v[j];
            }
    }
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
    if (G.half_force) {
        out[1 * NC + n] = jx + F_old[n] * MT(P.dt) / (MT(2.0) * rho);
        out[2 * NC + n] = jy + F_old[NC + n] * MT(P.dt) / (MT(2.0) * rho);
        out[3 * NC + n] = jz + F_old[2 * NC + n] * MT(P.dt) / (MT(2.0) * rho);
    } else {  // update_velIBM: u = sum c e f (ibm.cu:356-360)
        out[1 * NC + n] = jx;
        out[2 * NC + n] = jy;
        out[3 * NC + n] = jz;
    }
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

// Initial state of a context (the drivers' fill, LockExchange.cu:133-139): f = feq(p, u) and
// g = geq(S or phi, u) of every cell, in FP64 with the operation order of the host fill and no
// fused multiply-adds, so every value equals the host computation's bit for bit before it is
// stored; the FP64 cell volumes; the phase field's history rows; the LES rate rows.  The fields
// come in canonical cell order, as the caller holds them: internal cell n is canonical cell
// ((x * ny) + y) * nz + z of its coordinates.
template <typename ST, typename MT, int Q, int QS>
__global__ void k_init_state(FGeom G, InitArgs A, ST* __restrict__ f, ST* __restrict__ g, double* __restrict__ vol,
                             MT* __restrict__ phiold, MT* __restrict__ uold, MT* __restrict__ vold, MT* __restrict__ wold,
                             MT* __restrict__ tau, MT* __restrict__ taus) {
#pragma clang fp contract(off)
    const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= G.ncell) return;
    const int s = G.s, o1 = G.o1, o2 = G.o2;
    const int64_t L = n / G.ns;
    int c[3];
    c[s] = (int)(n - L * G.ns);
    c[o1] = (int)(L / G.n[o2]);
    c[o2] = (int)(L - (L / G.n[o2]) * G.n[o2]);
    const int64_t id = ((int64_t)c[0] * G.n[1] + c[1]) * G.n[2] + c[2];
    double val[7];
#pragma unroll
    for (int r = 0; r < 7; ++r) val[r] = (A.present >> r) & 1u ? A.rows[(int64_t)r * G.ncell + id] : A.dflt[r];
    const double p = val[0], u = val[1], v = val[2], w = val[3], sc = val[4];
    const double cl = A.c, c2 = cl * cl;
    const double u2 = u * u + v * v + w * w;
    const uint32_t NC = (uint32_t)G.ncell, cell = (uint32_t)n;
    static_for<Q>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr Dir d = Lattice<Q>::c(k);
        const double t = ((double)d.x * cl * u + (double)d.y * cl * v + (double)d.z * cl * w) / c2;
        const double fk = Lattice<Q>::w(k) * (p + 3.0 * t + 4.5 * t * t - 1.5 * u2 / c2);
        Store<ST, MT>::store(elem(f, (uint32_t)k * NC + cell), (MT)fk, Lattice<Q>::w(k), G.scale);
    });
    if constexpr (QS > 0) {
        constexpr bool linear = QS == 7;
        const double cs2inv = linear ? 4.0 : 3.0;
        static_for<QS>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr Dir d = Lattice<QS>::c(k);
            const double t = ((double)d.x * cl * u + (double)d.y * cl * v + (double)d.z * cl * w) / c2;
            double e = 1.0 + cs2inv * t;
            if constexpr (!linear) e += 4.5 * t * t - 1.5 * u2 / c2;
            const double gk = Lattice<QS>::w(k) * sc * e;
            Store<ST, MT>::store(elem(g, (uint32_t)k * NC + cell), (MT)gk, Lattice<QS>::w(k), G.scale);
        });
    }
    *elem(vol, cell) = A.wid[0][c[0]] * A.wid[1][c[1]] * A.wid[2][c[2]] / (A.dz * A.dz * A.dz);
    if (phiold) {
        *elem(phiold, cell) = (MT)sc; *elem(uold, cell) = (MT)u; *elem(vold, cell) = (MT)v; *elem(wold, cell) = (MT)w;
    }
    if (tau) { *elem(tau, cell) = (MT)val[5]; *elem(taus, cell) = (MT)val[6]; }
}

// ---- launchers (declared in fused_common.hpp, instantiated in fused_q*.hip) ---------------
template <typename ST, typename MT, int Q, int QS>
void fused_launch_init(const FGeom& G, const InitArgs& A, ST* f, ST* g, double* vol, MT* phiold, MT* uold, MT* vold,
                       MT* wold, MT* tau, MT* taus, hipStream_t st) {
    IPL_LAUNCH((k_init_state<ST, MT, Q, QS>), (G.ncell + 255) / 256, 256, 0, st, G, A, f, g, vol, phiold, uold, vold,
               wold, tau, taus);
}

template <typename ST, typename MT, int Q, int QS>
void fused_launch_step(const BatchJob<ST, MT>* jobs, int njobs, int64_t blocks, int step, int npts, int s, bool flux,
                             bool ext, bool nt, unsigned block, size_t lds, hipStream_t st) {
    const unsigned grid = (unsigned)blocks;
    // ext: the wall function or LES is compiled in (it adds registers to every step otherwise);
    // nt: non-temporal population accesses (taken only without ext)
#define IPL_FSTEPB(NP, SS, FL)                                                                                        \
    do {                                                                                                              \
        if (ext) IPL_LAUNCH_SYNC((k_fused_step<ST, MT, Q, QS, NP, SS, FL, true, false>), grid, block, lds, st, jobs, njobs, step); \
        else if (nt) IPL_LAUNCH_SYNC((k_fused_step<ST, MT, Q, QS, NP, SS, FL, false, true>), grid, block, lds, st, jobs, njobs, step); \
        else IPL_LAUNCH_SYNC((k_fused_step<ST, MT, Q, QS, NP, SS, FL, false, false>), grid, block, lds, st, jobs, njobs, step); \
    } while (0)
    if (!flux) {
        if constexpr (Lattice<Q>::dim == 2) {
            if (s == 0) IPL_FSTEPB(5, 0, false); else IPL_FSTEPB(5, 2, false);
        } else {
            if (s == 0) IPL_FSTEPB(5, 0, false); else if (s == 1) IPL_FSTEPB(5, 1, false); else IPL_FSTEPB(5, 2, false);
        }
    } else if constexpr (Lattice<Q>::dim == 2) {
        if (npts == 5) { if (s == 0) IPL_FSTEPB(5, 0, true); else IPL_FSTEPB(5, 2, true); }
        else { if (s == 0) IPL_FSTEPB(3, 0, true); else IPL_FSTEPB(3, 2, true); }
    } else {
        if (npts == 5) { if (s == 0) IPL_FSTEPB(5, 0, true); else if (s == 1) IPL_FSTEPB(5, 1, true); else IPL_FSTEPB(5, 2, true); }
        else { if (s == 0) IPL_FSTEPB(3, 0, true); else if (s == 1) IPL_FSTEPB(3, 1, true); else IPL_FSTEPB(3, 2, true); }
    }
#undef IPL_FSTEPB
}

template <typename ST, typename MT, int Q, int QS>
size_t fused_lds_bytes(int s, int ns) {
    // one padded row of the S-directions' values per cell for each population set
    int nf = 1, ng = 0;
    for (int k = 0; k < Q; ++k) nf += comp<Q>(k, s) != 0;
    if constexpr (QS > 0) {
        ng = 2;  // + the line's rho (inline_rho) or phi (inline_phi)
        for (int k = 0; k < QS; ++k) ng += comp<QS>(k, s) != 0;
    }
    return (size_t)(nf + ng) * ns * sizeof(MT);
}

// ---- LES (basic.cu:285-339), after the step: velocities, then the model ---------------------
// k_les_fields: u of every cell as update_rho leaves it (sum e c f + F dt / (2 rho), with the
// step's force) and the molecular kinematic viscosity (mu_L + phi (mu_H - mu_L)) / rho
template <typename ST, typename MT, int Q, int QS>
__global__ void k_les_fields(FGeom G, const ST* __restrict__ f, const ST* __restrict__ g, const MT* __restrict__ F,
                             MT* __restrict__ vel, MT* __restrict__ nu, FCoef<MT> P) {
    const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= G.ncell) return;
    const uint32_t NC = (uint32_t)G.ncell, c = (uint32_t)n;
    MT ex = 0, ey = 0, ez = 0;
    static_for<Q>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr Dir d = Lattice<Q>::c(k);
        if constexpr (d.x != 0 || d.y != 0 || d.z != 0) {
            const MT fv = Store<ST, MT>::load(elem(f, (uint32_t)k * NC + c), Lattice<Q>::w(k), G.inv_scale);
            if constexpr (d.x > 0) ex += fv; else if constexpr (d.x < 0) ex -= fv;
            if constexpr (d.y > 0) ey += fv; else if constexpr (d.y < 0) ey -= fv;
            if constexpr (d.z > 0) ez += fv; else if constexpr (d.z < 0) ez -= fv;
        }
    });
    MT sc = 0;
    if constexpr (QS > 0) {
#pragma unroll
        for (int k = 0; k < QS; ++k) sc += Store<ST, MT>::load(elem(g, (uint32_t)k * NC + c), Lattice<QS>::w(k), G.inv_scale);
    }
    const MT S_ = G.scalar == IPL_SCALAR_SALINITY ? sc : MT(0);
    const MT phi = G.scalar == IPL_SCALAR_PHASE_FIELD ? sc : MT(0);
    const MT rho = P.rho_L + P.drho * phi + S_ * P.beta_s;
    const MT hr = P.dt / (MT(2.0) * rho);
    *elem(vel, c) = ex * P.c + *elem(F, c) * hr;
    *elem(vel, NC + c) = ey * P.c + *elem(F, NC + c) * hr;
    *elem(vel, 2 * NC + c) = ez * P.c + *elem(F, 2 * NC + c) * hr;
    *elem(nu, c) = (P.mu_L + phi * (P.mu_H - P.mu_L)) / rho;
}

// velocity component q of the neighbour along axis A at offset OFF: zero beyond a wall (a ghost
// has zero velocity, basic.cu:226)
template <int S, int A, int OFF, typename MT>
__device__ inline MT les_nb(const FGeom& G, const LineNb& nb, const MT* vel, uint32_t q, int i) {
    constexpr int o1 = LineAxes<S>::o1;
    bool wall;
    if constexpr (A == S) wall = (i + OFF < 0 || i + OFF >= G.ns) && !G.periodic[S];
    else if constexpr (A == o1) wall = OFF < 0 ? nb.o1m : nb.o1p;
    else wall = OFF < 0 ? nb.o2m : nb.o2p;
    return wall ? MT(0) : *elem(vel, q * (uint32_t)G.ncell + nb_index<S, A, OFF>(G, nb, i));
}

// k_les: the coherent-structure model per cell, one workgroup per line; gradients
// (v[+1] - v[-1]) / (pos[+1] - pos[-1] + dz/1000) with the reciprocal tables of Force.  In 2D the
// reference takes the z differences over the same neighbour twice, so they vanish.
template <typename MT, int Q, int S>
__global__ void k_les(FGeom G, FTables<MT> tb, const MT* __restrict__ vel, const MT* __restrict__ nu,
                      MT* __restrict__ tau, MT* __restrict__ taus, FCoef<MT> P) {
    constexpr bool d3 = Lattice<Q>::dim == 3;
    constexpr int o1 = LineAxes<S>::o1;
    const int i = threadIdx.x;
    if (i >= G.ns) return;
    const int L = (int)blockIdx.x;
    const int n2 = G.n[LineAxes<S>::o2];
    const int c1 = L / n2, c2 = L - (L / n2) * n2;
    const LineNb nb = line_neighbours<S>(G, c1, c2);
    const uint32_t n = (uint32_t)L * (uint32_t)G.ns + (uint32_t)i;
    int cc[3];
    cc[S] = i; cc[o1] = c1; cc[LineAxes<S>::o2] = c2;
    MT gr[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // gr[axis][component]
    auto axis_grad = [&](auto ac) {
        constexpr int A = decltype(ac)::value;
        const MT inv = *elem(tb.invd[A], (uint32_t)cc[A]);
#pragma unroll
        for (int q = 0; q < 3; ++q)
            gr[A][q] = (les_nb<S, A, 1>(G, nb, vel, (uint32_t)q, i) - les_nb<S, A, -1>(G, nb, vel, (uint32_t)q, i)) * inv;
    };
    axis_grad(std::integral_constant<int, 0>{});
    if constexpr (d3) {
        axis_grad(std::integral_constant<int, 1>{});
        axis_grad(std::integral_constant<int, 2>{});
    }
    const MT nxx = gr[0][0], nxy = gr[0][1], nxz = gr[0][2];
    const MT nyx = gr[1][0], nyy = gr[1][1], nyz = gr[1][2];
    const MT nzx = gr[2][0], nzy = gr[2][1], nzz = gr[2][2];
    const MT arr[9] = {nxx, (nxy + nyx) / MT(2), (nxz + nzx) / MT(2), (nyx + nxy) / MT(2), nyy,
                       (nyz + nzy) / MT(2), (nzx + nxz) / MT(2), (nzy + nyz) / MT(2), nzz};
    const MT Qi = MT(-0.5) * (nxx * nxx + nxy * nyx + nxz * nzx + nyx * nxy + nyy * nyy + nyz * nzy +
                              nzx * nxz + nzy * nyz + nzz * nzz);
    const MT sum = nxx + nxy + nxz + nyx + nyy + nyz + nzx + nzy + nzz;
    const MT E = MT(1) / (MT(2) * sum * sum + MT(1e-40));
    const MT Fcs = fmax(MT(-1), fmin(Qi / E, MT(1)));
    const MT af = fabs(Fcs);
    const MT C = MT(0.05) * af * sqrt(af);
    MT Ssum = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) Ssum += sqrt(MT(2) * arr[k] * arr[k]);
    const MT c2dt = P.c * P.c * P.dt;
    const MT dz2 = P.dz * P.dz;
    *elem(tau, n) = MT(3) * (*elem(nu, n) + C * Ssum * dz2) / c2dt + MT(0.5);
    *elem(taus, n) = MT(3) * (MT(1e-9) + C * Ssum * dz2 / MT(1000)) / c2dt + MT(0.5);
}

template <typename ST, typename MT, int Q, int QS>
void fused_launch_les(const FGeom& G, const FTables<MT>& tb, const ST* f, const ST* g, const MT* F, MT* vel, MT* nu,
                      MT* tau, MT* taus, const FCoef<MT>& P, unsigned block, hipStream_t st) {
    IPL_LAUNCH((k_les_fields<ST, MT, Q, QS>), (G.ncell + 255) / 256, 256, 0, st, G, f, g, F, vel, nu, P);
    if (G.s == 0) IPL_LAUNCH((k_les<MT, Q, 0>), (unsigned)G.lines, block, 0, st, G, tb, vel, nu, tau, taus, P);
    else if (G.s == 1) IPL_LAUNCH((k_les<MT, Q, 1>), (unsigned)G.lines, block, 0, st, G, tb, vel, nu, tau, taus, P);
    else IPL_LAUNCH((k_les<MT, Q, 2>), (unsigned)G.lines, block, 0, st, G, tb, vel, nu, tau, taus, P);
}

template <typename ST, typename MT, int QS>
void fused_launch_scalar_fields(const FGeom& G, const ST* g_src, MT* rho_n, MT* phi_n, const FCoef<MT>& P, hipStream_t st) {
    // at least 1024 workgroups of 256 work-items (four per CU) before cells are grouped
    const int64_t NC = G.ncell;
    if (NC >= (int64_t)4 * 256 * 1024) {
        const int64_t items = (NC + 3) / 4;
        IPL_LAUNCH((k_scalar_fields<ST, MT, QS, 4>), (items + 255) / 256, 256, 0, st, G, g_src, rho_n, phi_n, P);
    } else if (NC >= (int64_t)2 * 256 * 1024) {
        const int64_t items = (NC + 1) / 2;
        IPL_LAUNCH((k_scalar_fields<ST, MT, QS, 2>), (items + 255) / 256, 256, 0, st, G, g_src, rho_n, phi_n, P);
    } else {
        IPL_LAUNCH((k_scalar_fields<ST, MT, QS, 1>), (NC + 255) / 256, 256, 0, st, G, g_src, rho_n, phi_n, P);
    }
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
    template void ipl::fusedk::fused_launch_les<ST, MT, Q, QS>(const FGeom&, const FTables<MT>&, const ST*, const ST*, \
                                                              const MT*, MT*, MT*, MT*, MT*, const FCoef<MT>&,        \
                                                              unsigned, hipStream_t);                                 \
    template void ipl::fusedk::fused_launch_step<ST, MT, Q, QS>(const BatchJob<ST, MT>*, int, int64_t, int, int, int, \
                                                                       bool, bool, bool, unsigned, size_t, hipStream_t); \
    template size_t ipl::fusedk::fused_lds_bytes<ST, MT, Q, QS>(int, int);                                    \
    template void ipl::fusedk::fused_launch_readout<ST, MT, Q, QS>(const FGeom&, const ST*, const ST*, const MT*, \
                                                                    double*, const FParams&, hipStream_t);      \
    template void ipl::fusedk::fused_launch_init<ST, MT, Q, QS>(const FGeom&, const ipl::fusedk::InitArgs&, ST*, ST*, double*, \
                                                                 MT*, MT*, MT*, MT*, MT*, MT*, hipStream_t);
#define IPL_FUSED_INSTANTIATE_SCALAR(ST, MT, QS)                                                           \
    template void ipl::fusedk::fused_launch_scalar_fields<ST, MT, QS>(const FGeom&, const ST*, MT*, MT*,    \
                                                                       const FCoef<MT>&, hipStream_t);
#define IPL_FUSED_INSTANTIATE_POPS(ST, MT, Q)                                                               \
    template void ipl::fusedk::fused_launch_convert<ST, MT, Q>(const FGeom&, const ST*, double*, int, bool, ST*, \
                                                                hipStream_t);
