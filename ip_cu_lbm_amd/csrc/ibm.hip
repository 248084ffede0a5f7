// ibm.hip - immersed-boundary sub-step of the fused path (function/ibm.cu); see ibm.hpp.
#include "ibm.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "lattice.hpp"

namespace ipl {
namespace ibk {

// body state, FP64, one array on the device
enum BodySlot { B_POS = 0, B_VEL = 3, B_OMEGA = 6, B_QUAT = 9, B_QUATOLD = 13, B_R = 17, B_FB = 26, B_T = 29, B_SIZE = 32 };

struct IbArgs {
    int n[3], periodic[3];
    int s, o1, o2, ns;
    int dim, Qs, scalar;
    uint32_t ncell;
    double dz, dt, c, tau;
    double rho_L, drho, beta_s;
    int M, S, lo, hi, block3;   // slots per marker; node offsets lo..hi per axis; reference block table
    int method, kernel, moving;
    double spread, wscale, gwscale;
    double sina, inv_c, inv_c2, half_itau;  // stress method: sin of the split angle, 1/c, 1/c^2, 1/(2 tau)
    int spm, bpad;                          // smoothed-profile forcing; extra box cells it needs
    double spm_a, spm_b;
    int mds;                                // multi-direct forcing's later passes in marker space
    double kappa;                           // their velocity change per unit overlap force: c^2 dt spread / 3
    double rho_f, density, mass, inertia[3], gravity[3];
    int64_t vmax;
    const double* pos[3];
};

__device__ inline uint32_t cell_of(const IbArgs& A, const int c[3]) {
    return (uint32_t)(c[A.o1] * A.n[A.o2] + c[A.o2]) * (uint32_t)A.ns + (uint32_t)c[A.s];
}

// node displaced by d from base: false where the reference's neighbour is a ghost (outside a
// wall); pos receives its coordinates (a ghost sits on its owner; unwrapped across a seam)
__device__ inline bool node_at(const IbArgs& A, const int base[3], const int d[3], int c[3], double pos[3]) {
    bool ghost = false;
    int wrap[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a) {
        c[a] = base[a] + d[a];
        if (c[a] < 0 || c[a] >= A.n[a]) {
            if (A.periodic[a]) {
                wrap[a] = c[a] < 0 ? -1 : 1;
                c[a] -= wrap[a] * A.n[a];
            } else {
                ghost = true;
            }
        }
    }
    // a periodic axis unwraps by its period (sum of widths), held after the coordinates
    for (int a = 0; a < 3; ++a)
        pos[a] = ghost ? A.pos[a][base[a]] : A.pos[a][c[a]] + wrap[a] * A.pos[a][A.n[a]];
    return !ghost;
}

// displacement of slot s of the node block: i (x) outer, j, l (z) inner over lo..hi; the
// reference's table (idfIndex, ibm.cu:132-141) maps (x-1, z-1) of the 2D block to (x+1, z+1)
__device__ inline void slot_offset(const IbArgs& A, int s, int d[3]) {
    const int np = A.hi - A.lo + 1, npy = A.dim == 3 ? np : 1;
    d[0] = s / (npy * np) + A.lo;
    d[1] = A.dim == 3 ? (s / np) % npy + A.lo : 0;
    d[2] = s % np + A.lo;
    if (A.block3 && A.dim == 2 && d[0] == -1 && d[2] == -1) { d[0] = 1; d[2] = 1; }
}

__device__ inline double peskin4(double r) {
    if (r < 1.0) return (3.0 - 2.0 * r + sqrt(1.0 + 4.0 * r - 4.0 * r * r)) / 8.0;
    if (r <= 2.0) return (5.0 - 2.0 * r - sqrt(-7.0 + 12.0 * r - 4.0 * r * r)) / 8.0;
    return 0.0;
}

// Lagrange_interpolation (adv.cu:40-48) with n = 3
__device__ inline void lagrange3(double* coeff, const double* x, double xa) {
    for (int i = 0; i < 3; ++i) {
        coeff[i] = 1.0;
        for (int j = 0; j < 3; ++j)
            if (j != i) coeff[i] *= (xa - x[j]) / (x[i] - x[j]);
    }
}

// Node weights per axis (get_IBMGw2 ibm.cu:146-189 for the stress method, IB_directForcing
// ibm.cu:292-313 for direct forcing, the full Peskin support for the PESKIN4 kernel)
__device__ inline void axis_weights(const IbArgs& A, const int base[3], const double xw[3], double W[3][5]) {
    const int np = A.hi - A.lo + 1;
    const bool stress = A.method == IPL_IB_STRESS;
    const bool pk4 = !stress && A.kernel == IPL_IB_KERNEL_PESKIN4;
    for (int a = 0; a < 3; ++a) {
        if (A.dim == 2 && a == 1) {
            for (int t = 0; t < np; ++t) W[a][t] = t == -A.lo ? 1.0 : 0.0;
            continue;
        }
        double xs[5];
        bool ok[5];
        for (int t = 0; t < np; ++t) {
            int d[3] = {0, 0, 0}, c[3];
            d[a] = A.lo + t;
            double pp[3];
            ok[t] = node_at(A, base, d, c, pp);
            xs[t] = pp[a];
        }
        const bool peskin = pk4 || (stress && (a == 0 || (A.dim == 2 && a == 2)));
        if (peskin) {
            for (int t = 0; t < np; ++t) W[a][t] = peskin4(fabs(xw[a] - xs[t]) / A.dz) * (ok[t] ? 1.0 : 0.0);
            if (stress && A.dim == 2) {
                double sum = 0;
                for (int t = 0; t < np; ++t) sum += W[a][t];
                for (int t = 0; t < np; ++t) W[a][t] /= sum;
            }
        } else {
            lagrange3(W[a], xs, xw[a]);
        }
    }
}

__device__ inline double slot_weight(const IbArgs& A, const double W[3][5], int s) {
    const int np = A.hi - A.lo + 1, npy = A.dim == 3 ? np : 1;
    const int i = s / (npy * np), j = A.dim == 3 ? (s / np) % npy : -A.lo, l = s % np;
    return W[0][i] * W[1][j] * W[2][l];
}

template <typename ST> __device__ inline double ldpop(const ST* f, uint32_t idx) { return (double)f[idx]; }

// index of cell c (wrapped) in the box, -1 outside
__device__ inline int64_t box_index(const IbArgs& A, const int* box, const int c[3]) {
    int b[3];
    for (int a = 0; a < 3; ++a) {
        int x = c[a] - box[a];
        if (x < 0 && A.periodic[a]) x += A.n[a];
        if (x < 0 || x >= box[3 + a]) return -1;
        b[a] = x;
    }
    return ((int64_t)b[0] * box[4] + b[1]) * box[5] + b[2];
}

// ---- k_ib_rank: markers sorted by (nearest cell, index) and the marker run of every cell -------
// Keys are (nearest cell << 32 | marker index), all distinct.  A workgroup of 16 wavefronts takes
// 64 markers: it loads every key into LDS and finds the bounding box of the node blocks; lane l of
// wavefront p counts, over the p-th sixteenth of the keys, those below marker l's key (its rank)
// and those below the first and past the last key of its cell (the cell's run).  The sixteen
// partial counts are added in order; the first marker of a cell writes the cell's run (start,
// count), the apply kernel of the previous step having cleared the others.
constexpr int kIbRankThreads = 1024, kIbRankMarkers = 64, kIbRankParts = kIbRankThreads / kIbRankMarkers;

__global__ void __launch_bounds__(kIbRankThreads) k_ib_rank(IbArgs A, const int* __restrict__ node,
                                                             unsigned long long* __restrict__ sorted, int* __restrict__ box,
                                                             int2* __restrict__ range, int mpad) {
    IPL_DYNAMIC_LDS(unsigned long long, keys);
    __shared__ int work[6][kIbRankThreads];  // min / max of the node coordinates, then partial counts
    __shared__ int bx[6];
    const int t = threadIdx.x;
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (int i = t; i < mpad; i += kIbRankThreads) {
        if (i < A.M) {
            const int c[3] = {node[3 * i], node[3 * i + 1], node[3 * i + 2]};
            keys[i] = ((unsigned long long)cell_of(A, c) << 32) | (unsigned long long)i;
            for (int a = 0; a < 3; ++a) { lo[a] = min(lo[a], c[a]); hi[a] = max(hi[a], c[a]); }
        } else {
            keys[i] = ~0ull;  // above every key
        }
    }
    for (int a = 0; a < 3; ++a) { work[a][t] = lo[a]; work[3 + a][t] = hi[a]; }
    __syncthreads();
    for (int w = kIbRankThreads / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int a = 0; a < 3; ++a) {
                work[a][t] = min(work[a][t], work[a][t + w]);
                work[3 + a][t] = max(work[3 + a][t], work[3 + a][t + w]);
            }
        __syncthreads();
    }
    if (t == 0) {
        int64_t vol = 1;
        for (int a = 0; a < 3; ++a) {
            int l = work[a][0] + A.lo - A.bpad, h = work[3 + a][0] + A.hi + A.bpad;
            if (A.periodic[a]) {
                if (h - l + 1 >= A.n[a]) { l = 0; h = A.n[a] - 1; }
            } else {
                l = max(l, 0);
                h = min(h, A.n[a] - 1);
            }
            bx[3 + a] = h - l + 1;
            bx[a] = l < 0 ? l + A.n[a] : l;  // a periodic box across the low seam starts at its image
            vol *= bx[3 + a];
        }
        if (blockIdx.x == 0) {
            for (int a = 0; a < 6; ++a) box[a] = bx[a];
            box[6] = vol <= A.vmax ? 0 : 1;  // overflow flag (the host sizes vmax so that it stays 0)
        }
    }
    __syncthreads();
    const int l = t % kIbRankMarkers, part = t / kIbRankMarkers;
    const int m = blockIdx.x * kIbRankMarkers + l;
    const unsigned long long mine = m < A.M ? keys[m] : ~0ull;
    const unsigned long long first = mine & 0xffffffff00000000ull, past = first + (1ull << 32);
    int below = 0, lower = 0, upto = 0;
    const int chunk = mpad / kIbRankParts;  // mpad is a multiple of 2 * kIbRankParts
    for (int j = part * chunk; j < (part + 1) * chunk; j += 2) {
        const unsigned long long k0 = keys[j], k1 = keys[j + 1];
        below += (k0 < mine) + (k1 < mine);
        lower += (k0 < first) + (k1 < first);
        upto += (k0 < past) + (k1 < past);
    }
    __syncthreads();  // done with the box work area
    work[0][t] = below; work[1][t] = lower; work[2][t] = upto;
    __syncthreads();
    if (part != 0 || m >= A.M) return;
    for (int p = 1; p < kIbRankParts; ++p) {
        below += work[0][p * kIbRankMarkers + l];
        lower += work[1][p * kIbRankMarkers + l];
        upto += work[2][p * kIbRankMarkers + l];
    }
    sorted[below] = mine;
    if (below == lower) {  // first marker of its cell
        const int c[3] = {node[3 * m], node[3 * m + 1], node[3 * m + 2]};
        const int64_t b = box_index(A, bx, c);
        if (b >= 0 && b < A.vmax) range[b] = make_int2(lower, upto - lower);
    }
}

// ---- k_ib_ubase: update_rho's velocity (collision.cu:244-254) at every cell of the box -------
// (direct forcing interpolates it; computed once per cell instead of once per marker and slot)
template <typename ST, typename MT, int Q>
__global__ void k_ib_ubase(IbArgs A, const ST* __restrict__ f, const ST* __restrict__ g, const MT* __restrict__ F,
                           const int* __restrict__ box, double* __restrict__ Ubox) {
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t vol = (int64_t)box[3] * box[4] * box[5];
    if (b >= vol || b >= A.vmax) return;
    int c[3];
    c[2] = (int)(b % box[5]);
    c[1] = (int)((b / box[5]) % box[4]);
    c[0] = (int)(b / ((int64_t)box[4] * box[5]));
    for (int a = 0; a < 3; ++a) {
        c[a] += box[a];
        if (c[a] >= A.n[a]) c[a] -= A.n[a];
    }
    const uint32_t idx = cell_of(A, c);
    const double cl = A.c;
    double jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const Dir e = Lattice<Q>::c(k);
        const double fv = ldpop(f, (uint32_t)k * A.ncell + idx);
        jx += e.x * cl * fv; jy += e.y * cl * fv; jz += e.z * cl * fv;
    }
    double sc = 0;
    for (int k = 0; k < A.Qs; ++k) sc += ldpop(g, (uint32_t)k * A.ncell + idx);
    const double phi = A.scalar == IPL_SCALAR_PHASE_FIELD ? sc : 0.0, S = A.scalar == IPL_SCALAR_SALINITY ? sc : 0.0;
    const double rho = A.rho_L + A.drho * phi + S * A.beta_s;
    Ubox[3 * b + 0] = jx + (double)F[idx] * A.dt / (2.0 * rho);
    Ubox[3 * b + 1] = jy + (double)F[A.ncell + idx] * A.dt / (2.0 * rho);
    Ubox[3 * b + 2] = jz + (double)F[2 * (size_t)A.ncell + idx] * A.dt / (2.0 * rho);
}

// ---- k_ib_spm: smoothed-profile forcing of the body interior ----------------------------------
// u of the cell as update_rho leaves it (sum e c f + F dt / (2 rho), the step's force), the
// body-frame ellipsoid distance, the driver's profile (profile_s22: transition width fixed at
// 0.1), and f += w dt 3 (e c . F) / c^2 with F = profile (v_B - u) / dt
__device__ inline double spm_profile(double r, double Rada) {
    const double lim = 0.1;
    const double in_line = 1.0 - lim + (lim / 2) * (lim / 2), out_line = 1.0 + lim + (lim / 2) * (lim / 2);
    if (r < in_line) return 1.0;
    if (r <= out_line) return 0.5 * (sin(3.14159 * r / lim / Rada) + 1.0);
    return 0.0;
}

// profile_s2 (ibm.cu:50-58) of SPM_ellipse: 1 inside, 0 outside, a sine ramp of width lim between
__device__ inline double spm_profile_s2(double lim, double r) {
    const double in_line = 1.0 - lim + (lim / 2) * (lim / 2), out_line = 1.0 + lim + (lim / 2) * (lim / 2);
    if (r < in_line) return 1.0;
    if (r <= out_line) return 0.5 * (sin(3.14159 * r / lim) + 1.0);
    return 0.0;
}

// profile_s (ibm.cu:44-48) of SPM: 1 more than dz/2 inside the circle of radius R - dz/2, 0 more
// than dz/2 outside, a sine ramp between
__device__ inline double spm_profile_s(double dz, double R, double dist) {
    const double r = R - 0.5 * dz - dist;
    if (r > 0.5 * dz) return 1.0;
    if (r >= -0.5 * dz) return 0.5 * (sin(3.14159 * r / dz) + 1.0);
    return 0.0;
}

// Smoothed-profile forcing of the body interior, f += w dt 3 (e c . F) / c^2 with F = profile
// (v - u) / dt (u of update_rho): spm = 1 SPM_ellipse3D (SphereFall3D/SphereFall.cu:13-30) on the
// cells of the box, the ellipsoid (a, a, b) in the body frame R^T d; spm = 2 SPM_ellipse (ibm.cu:89-122)
// on every cell, the ellipse (a, b) over the body-frame x and z of R d (x and y on a 3D lattice:
// its "xy face"), v the body velocity; spm = 3 SPM (ibm.cu:68-87) on every cell, the circle of
// radius a about the body centre in the plane of the lattice's first and last axes (x-z in 2D,
// x-y in 3D), v the velocity of marker 0, F over dz.
template <typename ST, typename MT, int Q>
__global__ void k_ib_spm(IbArgs A, ST* __restrict__ f, const ST* __restrict__ g, const MT* __restrict__ F,
                         const int* __restrict__ box, const double* __restrict__ body, const double* __restrict__ velw) {
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int c[3];
    if (A.spm == 1) {
        const int64_t vol = (int64_t)box[3] * box[4] * box[5];
        if (b >= vol || b >= A.vmax) return;
        c[2] = (int)(b % box[5]);
        c[1] = (int)((b / box[5]) % box[4]);
        c[0] = (int)(b / ((int64_t)box[4] * box[5]));
        for (int a = 0; a < 3; ++a) {
            c[a] += box[a];
            if (c[a] >= A.n[a]) c[a] -= A.n[a];
        }
    } else {  // every cell, from its internal index
        if (b >= (int64_t)A.ncell) return;
        c[A.s] = (int)(b % A.ns);
        const int64_t line = b / A.ns;
        c[A.o2] = (int)(line % A.n[A.o2]);
        c[A.o1] = (int)(line / A.n[A.o2]);
    }
    const double* R = body + B_R;
    const double d[3] = {A.pos[0][c[0]] - body[B_POS], A.pos[1][c[1]] - body[B_POS + 1], A.pos[2][c[2]] - body[B_POS + 2]};
    double pr, vb[3];
    if (A.spm == 1) {
        const double X1 = R[0] * d[0] + R[3] * d[1] + R[6] * d[2];
        const double Y1 = R[1] * d[0] + R[4] * d[1] + R[7] * d[2];
        const double Z1 = R[2] * d[0] + R[5] * d[1] + R[8] * d[2];
        const double distance = (X1 / A.spm_a) * (X1 / A.spm_a) + (Y1 / A.spm_a) * (Y1 / A.spm_a) + (Z1 / A.spm_b) * (Z1 / A.spm_b);
        pr = spm_profile(distance, A.spm_a);
        for (int a = 0; a < 3; ++a) vb[a] = body[B_VEL + a];
    } else if (A.spm == 2) {
        const double X1 = R[0] * d[0] + R[1] * d[1] + R[2] * d[2];
        double Z1 = R[6] * d[0] + R[7] * d[1] + R[8] * d[2];
        if (A.dim == 3) Z1 = R[3] * d[0] + R[4] * d[1] + R[5] * d[2];  // the xy face
        const double distance = (X1 / A.spm_a) * (X1 / A.spm_a) + (Z1 / A.spm_b) * (Z1 / A.spm_b);
        pr = spm_profile_s2(A.dz / A.spm_a, distance);
        for (int a = 0; a < 3; ++a) vb[a] = body[B_VEL + a];
    } else {
        const int t = A.dim == 3 ? 1 : 2;
        pr = spm_profile_s(A.dz, A.spm_a, sqrt(d[0] * d[0] + d[t] * d[t]));
        for (int a = 0; a < 3; ++a) vb[a] = velw[a];
    }
    if (pr == 0.0) return;
    const uint32_t idx = cell_of(A, c);
    const double cl = A.c;
    double jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const Dir e = Lattice<Q>::c(k);
        const double fv = ldpop(f, (uint32_t)k * A.ncell + idx);
        jx += e.x * cl * fv; jy += e.y * cl * fv; jz += e.z * cl * fv;
    }
    double sc = 0;
    for (int k = 0; k < A.Qs; ++k) sc += ldpop(g, (uint32_t)k * A.ncell + idx);
    const double phi = A.scalar == IPL_SCALAR_PHASE_FIELD ? sc : 0.0, S = A.scalar == IPL_SCALAR_SALINITY ? sc : 0.0;
    const double rho = A.rho_L + A.drho * phi + S * A.beta_s;
    const double u[3] = {jx + (double)F[idx] * A.dt / (2.0 * rho), jy + (double)F[A.ncell + idx] * A.dt / (2.0 * rho),
                         jz + (double)F[2 * (size_t)A.ncell + idx] * A.dt / (2.0 * rho)};
    const double per = A.spm == 3 ? A.dz : A.dt;
    double fs[3];
    for (int a = 0; a < 3; ++a) fs[a] = pr * (vb[a] - u[a]) / per;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const Dir e = Lattice<Q>::c(k);
        ST& fv = f[(uint32_t)k * A.ncell + idx];
        fv = (ST)((double)fv + Lattice<Q>::w(k) * A.dt * 3.0 * (e.x * cl * fs[0] + e.y * cl * fs[1] + e.z * cl * fs[2]) / (cl * cl));
    }
}

// ---- k_ib_weights: direct forcing's node weights, computed with the ranking ---------------------
// One 64-lane workgroup per marker: the slot weights and box indices of its node block
// (IB_directForcing, ibm.cu:292-313, or the Peskin support; -2: a ghost node), the axis weights of
// the marker-space passes (a ghost node's zero on its own axis), and the marks of the box cells
// the block covers (the gather visits marked cells only).  The markers do not move between the
// ranking and the forcing passes, so every pass reads these.
constexpr int kIbAxisWeights = 15;  // per marker: 3 axes x 5 nodes (3-node blocks leave the last two 0)

__global__ void __launch_bounds__(64) k_ib_weights(IbArgs A, const int* __restrict__ node, const double* __restrict__ posw,
                                                   const int* __restrict__ box, double* __restrict__ wcache,
                                                   int* __restrict__ bcache, double* __restrict__ wax, int* __restrict__ mark) {
    __shared__ double xs[3][5], Wl[3][5];
    __shared__ int okl[3][5];
    const int m = blockIdx.x, t = threadIdx.x;
    const int base[3] = {node[3 * m], node[3 * m + 1], node[3 * m + 2]};
    const int np = A.hi - A.lo + 1;
    // lane 5a + i: node i of the block along axis a (its coordinate, and whether it is a ghost)
    const int a = t / 5, i = t % 5;
    const bool axis_lane = t < kIbAxisWeights && i < np && !(A.dim == 2 && a == 1);
    if (axis_lane) {
        int d[3] = {0, 0, 0}, c[3];
        double pp[3];
        d[a] = A.lo + i;
        okl[a][i] = node_at(A, base, d, c, pp);
        xs[a][i] = pp[a];
    }
    __syncthreads();
    // its weight (axis_weights' rules for direct forcing: the Peskin function, or 3-point Lagrange)
    if (t < kIbAxisWeights) {
        double w = 0.0;
        if (A.dim == 2 && a == 1) {
            w = i == -A.lo ? 1.0 : 0.0;
        } else if (i < np) {
            const double xw = posw[3 * m + a];
            if (A.kernel == IPL_IB_KERNEL_PESKIN4) {
                w = peskin4(fabs(xw - xs[a][i]) / A.dz) * (okl[a][i] ? 1.0 : 0.0);
            } else {
                w = 1.0;
                for (int j = 0; j < 3; ++j)
                    if (j != i) w *= (xw - xs[a][j]) / (xs[a][i] - xs[a][j]);
            }
        }
        Wl[a][i] = w;
        if (A.mds) {  // a ghost node's axis weight is zero
            const int x = base[a] + A.lo + i;
            const bool in = A.periodic[a] || (x >= 0 && x < A.n[a]);
            wax[(size_t)m * kIbAxisWeights + t] = i < np && in ? w : 0.0;
        }
    }
    __syncthreads();
    for (int j = 0; j < 2; ++j) {
        const int s = t + j * 64;
        if (s >= A.S) break;
        int d[3], cc[3];
        double pp[3];
        slot_offset(A, s, d);
        const bool ok = node_at(A, base, d, cc, pp);
        int64_t bx = -1;
        if (ok) {
            bx = box_index(A, box, cc);
            if (bx >= 0 && bx < A.vmax) mark[bx] = 1;
        }
        wcache[(size_t)m * A.S + s] = slot_weight(A, Wl, s);
        bcache[(size_t)m * A.S + s] = ok ? (int)bx : -2;
    }
}

// ---- k_ib_force: one 64-lane workgroup per marker, lanes over the slots of its node block ----
// Stress method (get_IBMGw2, ibm.cu:124-274): f interpolated at the marker, split into outer (fw1)
// and inner (fw2) populations by the normal; minus the traction of (fw1 - fw2) spread, marker force
// Gw from the traction of fw1.  Direct forcing (IB_directForcing, ibm.cu:276-344): node velocity
// (plus c^2 dt F_ib / 3 of earlier passes) interpolated, (U_w - u)/dt spread.  Slot sums are reduced in
// a fixed tree, so every run adds in the same order.
constexpr int kIbLanes = 64;

template <typename ST, typename MT, int Q>
__global__ void __launch_bounds__(kIbLanes) k_ib_force(IbArgs A, int it, const ST* __restrict__ f, const int* __restrict__ node,
                                                      const double* __restrict__ posw, const double* __restrict__ nB,
                                                      const double* __restrict__ velw, double* __restrict__ Gw,
                                                      double* __restrict__ contrib, const int* __restrict__ box,
                                                      const double* __restrict__ Fbox, const double* __restrict__ Ubox,
                                                      const double* __restrict__ wcache, const int* __restrict__ bcache,
                                                      int* __restrict__ mark, double* __restrict__ um,
                                                      double* __restrict__ gnext, double* __restrict__ gtot) {
    constexpr int kRows = Q > 20 ? Q : 20;  // the slot sums, then the 20 moment sums
    __shared__ double red[kIbLanes * kRows];
    __shared__ double share[3];
    const int m = blockIdx.x, t = threadIdx.x;
    const double U[3] = {velw[3 * m], velw[3 * m + 1], velw[3 * m + 2]};
    const double c = A.c, c2 = c * c;
    double* out = contrib + (size_t)m * A.S * 3;
    const int nr = A.method == IPL_IB_STRESS ? Q : 3;
    // the lane's slots (at most two: S <= 2 x 64): node weight, node valid (not a ghost) and its
    // cell (stress) or box index (direct forcing).  Direct forcing reads them from the ranking's
    // weights (k_ib_weights).
    double wsl[2] = {0.0, 0.0};
    bool oksl[2] = {false, false};
    int64_t bsl[2] = {-1, -1};
    uint32_t isl[2] = {0, 0};
    const bool direct = A.method == IPL_IB_DIRECT;
    if (direct) {
        for (int j = 0; j < 2; ++j) {
            const int s = t + j * kIbLanes;
            if (s >= A.S) break;
            const int bc = bcache[(size_t)m * A.S + s];
            wsl[j] = wcache[(size_t)m * A.S + s];
            oksl[j] = bc != -2;
            bsl[j] = bc;
        }
    } else {
        const int base[3] = {node[3 * m], node[3 * m + 1], node[3 * m + 2]};
        const double xw[3] = {posw[3 * m], posw[3 * m + 1], posw[3 * m + 2]};
        double W[3][5];
        axis_weights(A, base, xw, W);
        for (int j = 0; j < 2; ++j) {
            const int s = t + j * kIbLanes;
            if (s >= A.S) break;
            int d[3], cc[3];
            double pp[3];
            slot_offset(A, s, d);
            oksl[j] = node_at(A, base, d, cc, pp);
            wsl[j] = slot_weight(A, W, s);
            if (oksl[j]) {
                const int64_t bx = box_index(A, box, cc);
                isl[j] = cell_of(A, cc);
                if (bx >= 0 && bx < A.vmax) mark[bx] = 1;  // the gather visits marked cells only
            }
        }
    }
    // 1. per-lane partial sums over the lane's slots
    double part[Q > 3 ? Q : 3];
#pragma unroll
    for (int k = 0; k < (Q > 3 ? Q : 3); ++k) part[k] = 0.0;
    for (int j = 0; j < 2; ++j) {
        if (!oksl[j]) continue;
        const double wgt = wsl[j];
        if (!direct) {
            const uint32_t idx = isl[j];
#pragma unroll
            for (int k = 0; k < Q; ++k) part[k] += wgt * ldpop(f, (uint32_t)k * A.ncell + idx);
        } else {
            const int64_t b = bsl[j];
            if (b < 0 || b >= A.vmax) continue;  // cannot happen: the box holds every node block
            for (int a = 0; a < 3; ++a) {
                double u = Ubox[3 * b + a];
                if (it > 0) u += A.c * A.c * A.dt * Fbox[3 * b + a] / 3.0;  // what update_velIBM adds
                part[a] += wgt * u;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < (Q > 3 ? Q : 3); ++k)
        if (k < nr) red[k * kIbLanes + t] = part[k];
    __syncthreads();
    for (int w = kIbLanes / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < nr; ++k) red[k * kIbLanes + t] += red[k * kIbLanes + t + w];
        __syncthreads();
    }
    // 2. the marker force
    if (A.method == IPL_IB_STRESS) {
        // Lane k takes direction k: f at the marker split by the normal into the outer (fw1) and
        // inner (fw2) populations (ibm.cu:191-236), then its share of the moments sum_k g,
        // sum_k g e_a and sum_k g e_a e_b for g = fw1 - fw2 and g = fw1.
        double mom[20];
#pragma unroll
        for (int q = 0; q < 20; ++q) mom[q] = 0.0;
        if (t < Q) {
            int ex = 0, ey = 0, ez = 0, op = 0;
            double wk = 0.0;
#pragma unroll
            for (int k = 0; k < Q; ++k)
                if (t == k) {
                    const Dir e = Lattice<Q>::c(k);
                    ex = e.x; ey = e.y; ez = e.z;
                    op = opposite<Q>(k);
                    wk = Lattice<Q>::w(k);
                }
            const double fk = red[t * kIbLanes], fo = red[op * kIbLanes];
            const double nv[3] = {nB[3 * m], nB[3 * m + 1], nB[3 * m + 2]};
            // the reference's lattice velocities are c e (items[IDX_cx], Items::setc, basic.cu:79-96)
            const double ev[3] = {ex * A.c, ey * A.c, ez * A.c};
            const double cn = (nv[0] * ev[0] + nv[1] * ev[1] + nv[2] * ev[2]) * A.inv_c;
            const double lim = sqrt(ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2]) * A.inv_c * A.sina;
            const double bounce = fo + 6.0 * wk * (ev[0] * U[0] + ev[1] * U[1] + ev[2] * U[2]) * A.inv_c2;
            const double beta = fabs(cn) / (lim + 1e-10);
            const double mixed = (1.0 - beta) * fk + beta * bounce;
            const double fw1 = cn <= 0 ? fk : (cn <= lim ? mixed : bounce);
            const double fw2 = cn >= 0 ? fk : (cn >= -lim ? mixed : bounce);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const double g = h == 0 ? fw1 - fw2 : fw1;
                double* o = mom + 10 * h;
                o[0] = g;
                o[1] = g * ev[0]; o[2] = g * ev[1]; o[3] = g * ev[2];
                o[4] = g * ev[0] * ev[0]; o[5] = g * ev[1] * ev[1]; o[6] = g * ev[2] * ev[2];
                o[7] = g * ev[0] * ev[1]; o[8] = g * ev[0] * ev[2]; o[9] = g * ev[1] * ev[2];
            }
        }
        __syncthreads();  // every lane has read its slot sums
#pragma unroll
        for (int q = 0; q < 20; ++q) red[q * kIbLanes + t] = mom[q];
        __syncthreads();
        for (int w = kIbLanes / 2; w > 0; w >>= 1) {
            if (t < w)
#pragma unroll
                for (int q = 0; q < 20; ++q) red[q * kIbLanes + t] += red[q * kIbLanes + t + w];
            __syncthreads();
        }
        // traction sigma . n with sigma_ab = -sum_k [g_k delta_ab / (2 tau)
        //   + (1 - 1/(2 tau)) 3 (g_k (c e_a - U_a)(c e_b - U_b) - (g_k - w_k) U_a U_b) / c^2]
        // = -[G delta_ab / (2 tau) + (1 - 1/(2 tau)) 3 / c^2 (G_ab - G_a U_b - U_a G_b + U_a U_b)]
        // (sum_k w_k = 1), the same on every lane
        const double nv[3] = {nB[3 * m], nB[3 * m + 1], nB[3 * m + 2]};
        const double cf = (1.0 - A.half_itau) * 3.0 * A.inv_c2;
        double tr[2][3];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            double M[10];
#pragma unroll
            for (int q = 0; q < 10; ++q) M[q] = red[(10 * h + q) * kIbLanes];
            const double G2[3][3] = {{M[4], M[7], M[8]}, {M[7], M[5], M[9]}, {M[8], M[9], M[6]}};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                double acc = 0.0;
#pragma unroll
                for (int b2 = 0; b2 < 3; ++b2) {
                    const double sg = -((a == b2 ? M[0] * A.half_itau : 0.0) +
                                        cf * (G2[a][b2] - M[1 + a] * U[b2] - U[a] * M[1 + b2] + U[a] * U[b2]));
                    acc += sg * nv[b2];
                }
                tr[h][a] = acc;
            }
        }
        if (t == 0)
            for (int a = 0; a < 3; ++a) Gw[3 * m + a] = tr[1][a] * A.gwscale;
        const double gs0 = -tr[0][0], gs1 = -tr[0][1], gs2 = -tr[0][2];
        for (int j = 0; j < 2; ++j) {
            const int s = t + j * kIbLanes;
            if (s >= A.S) break;
            const bool ok = oksl[j];
            const double wgt = wsl[j];
            out[3 * s + 0] = ok ? gs0 * A.spread * wgt : 0.0;
            out[3 * s + 1] = ok ? gs1 * A.spread * wgt : 0.0;
            out[3 * s + 2] = ok ? gs2 * A.spread * wgt : 0.0;
        }
        return;
    }
    if (t == 0) {
        for (int a = 0; a < 3; ++a) {
            const double gwv = (U[a] - red[a * kIbLanes]) / A.dt;
            share[a] = gwv;
            Gw[3 * m + a] = (it == 0 ? 0.0 : Gw[3 * m + a]) + -gwv * A.wscale;
            if (A.mds) { um[3 * m + a] = red[a * kIbLanes]; gnext[3 * m + a] = gwv; gtot[3 * m + a] = gwv; }
        }
    }
    if (A.mds) return;  // the later passes run in marker space (k_ib_correct), the last writes the slots
    __syncthreads();
    // 3. contribution of every slot to the spread force
    const double g0 = share[0], g1 = share[1], g2 = share[2];
    for (int j = 0; j < 2; ++j) {
        const int s = t + j * kIbLanes;
        if (s >= A.S) break;
        const bool ok = oksl[j];
        const double wgt = wsl[j];
        out[3 * s + 0] = ok ? g0 * A.spread * wgt : 0.0;
        out[3 * s + 1] = ok ? g1 * A.spread * wgt : 0.0;
        out[3 * s + 2] = ok ? g2 * A.spread * wgt : 0.0;
    }
}

// ---- k_ib_correct: a later multi-direct forcing pass, in marker space ---------------------------
// The spread of the previous pass's marker forces g changes the node velocities by (c^2 dt / 3) dF
// (update_velIBM), so the velocity interpolated at marker m changes by
//   kappa sum_m' O_mm' g_m',  O_mm' = sum over the cells both node blocks reach of w_m w_m'.
// Node weights are products of axis weights (a ghost node's zero on its own axis), so O_mm' is the
// product over the axes of the overlaps of the two markers' axis weights.  One 64-lane workgroup
// per marker: the markers near it (listed with the ranking, k_ib_near) are taken one per lane in
// list order and the lane sums are added in a fixed tree, so every run adds in the same order.  The last pass writes every slot's share
// of the summed force of all passes (the spread of a sum is the sum of the spreads), which one
// gather then adds up per cell.
// Overlap of two markers' axis weights on axis a: node i of m (cell bm + lo + i) meets node j of
// k (cell bk + lo + j) where j = i + bm - bk, taking the nearest periodic image (a periodic axis
// of marker-space passes holds at least 9 cells, so no other image can meet).
// wk holds k's five axis weights in registers (loaded whole, so that the loads need not wait for
// the node offset), and the node pairs are selected with compile-time indices.
__device__ inline double axis_overlap(const IbArgs& A, int a, const double* wm, const double* wk, int bm, int bk) {
    int d0 = bm - bk;
    if (A.periodic[a]) {
        if (2 * d0 > A.n[a]) d0 -= A.n[a];
        else if (2 * d0 < -A.n[a]) d0 += A.n[a];
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j)
            if (j - i == d0) s += wm[i] * wk[j];
    return s;
}

// ---- k_ib_near: the markers whose node blocks can meet marker m's ------------------------------
// Nearest cells within the block span on every axis (nearest periodic image), in index order: one
// 64-lane workgroup per marker votes over 64 markers at a time and appends the hits in lane order.
// The first kIbNearMax go to m's row; the count is the full one (a longer list is not kept, and
// the marker-space pass then scans all markers for m itself).
constexpr int kIbNearMax = 256;

__device__ inline bool node_near(const IbArgs& A, const int bm[3], const int* __restrict__ node, int k) {
    const int span = A.hi - A.lo;
    bool near = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int d = node[3 * k + a] - bm[a];
        if (A.periodic[a]) {
            if (2 * d > A.n[a]) d -= A.n[a];
            else if (2 * d < -A.n[a]) d += A.n[a];
        }
        near = near && d <= span && d >= -span;
    }
    return near;
}

__global__ void __launch_bounds__(kIbLanes) k_ib_near(IbArgs A, const int* __restrict__ node, int* __restrict__ near,
                                                     int* __restrict__ near_count) {
    const int m = blockIdx.x, t = threadIdx.x;
    const int bm[3] = {node[3 * m], node[3 * m + 1], node[3 * m + 2]};
    int* row = near + (size_t)m * kIbNearMax;
    int cnt = 0;  // the same on every lane
    for (int base = 0; base < A.M; base += kIbLanes) {
        const int k = base + t;
        const bool hit = k < A.M && node_near(A, bm, node, k);
        const unsigned long long mask = __ballot(hit);
        const int pos = cnt + (int)__popcll(mask & __lanemask_lt());
        if (hit && pos < kIbNearMax) row[pos] = k;
        cnt += (int)__popcll(mask);
    }
    if (t == 0) near_count[m] = cnt;
}

__global__ void __launch_bounds__(kIbLanes) k_ib_correct(IbArgs A, int last, const int* __restrict__ node,
                                                        const int* __restrict__ near, const int* __restrict__ near_count,
                                                        const double* __restrict__ velw, const double* __restrict__ wax,
                                                        double* __restrict__ um, const double* __restrict__ gprev,
                                                        double* __restrict__ gnext, double* __restrict__ gtot,
                                                        double* __restrict__ Gw, const double* __restrict__ wcache,
                                                        const int* __restrict__ bcache, double* __restrict__ contrib) {
    __shared__ double red[3][kIbLanes];
    __shared__ double share[3];
    const int m = blockIdx.x, t = threadIdx.x;
    const int bm[3] = {node[3 * m], node[3 * m + 1], node[3 * m + 2]};
    double wm[3][5];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int i = 0; i < 5; ++i) wm[a][i] = wax[(size_t)m * kIbAxisWeights + 5 * a + i];
    double acc[3] = {0.0, 0.0, 0.0};
    // the overlap with marker k times k's force, added to the lane's share of the sum
    auto add = [&](int k) {
        double wk[kIbAxisWeights];
#pragma unroll
        for (int q = 0; q < kIbAxisWeights; ++q) wk[q] = wax[(size_t)k * kIbAxisWeights + q];
        double o = axis_overlap(A, 0, wm[0], wk, bm[0], node[3 * k]);
        o *= axis_overlap(A, 1, wm[1], wk + 5, bm[1], node[3 * k + 1]);
        o *= axis_overlap(A, 2, wm[2], wk + 10, bm[2], node[3 * k + 2]);
        acc[0] += o * gprev[3 * k]; acc[1] += o * gprev[3 * k + 1]; acc[2] += o * gprev[3 * k + 2];
    };
    // What does not depend on the list is read first, so that its latency overlaps the list's:
    // lane a < 3 holds component a of m's velocity, forces and Gw; the last pass's slot weights.
    const int cnt = near_count[m];
    const int* row = near + (size_t)m * kIbNearMax;
    const int first = row[t];  // t < 64 <= kIbNearMax: inside the row, used only when t < cnt
    const int a3 = t < 3 ? t : 0;
    const double velw_a = velw[3 * m + a3], um_a = um[3 * m + a3], gtot_a = gtot[3 * m + a3], Gw_a = Gw[3 * m + a3];
    double wsl[2] = {0.0, 0.0};
    bool oksl[2] = {false, false};
    if (last)
        for (int j = 0; j < 2; ++j) {
            const int s = t + j * kIbLanes;
            if (s >= A.S) break;
            oksl[j] = bcache[(size_t)m * A.S + s] != -2;
            wsl[j] = wcache[(size_t)m * A.S + s];
        }
    // lane l takes entries l, l + 64, ... of m's list of near markers (index order), or, where the
    // list was too long to keep, tests every marker l, l + 64, ... itself
    if (cnt <= kIbNearMax) {
        for (int e = t; e < cnt; e += kIbLanes) add(e == t ? first : row[e]);
    } else {
        for (int k = t; k < A.M; k += kIbLanes)
            if (node_near(A, bm, node, k)) add(k);
    }
    for (int a = 0; a < 3; ++a) red[a][t] = acc[a];
    __syncthreads();
    for (int w = kIbLanes / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int a = 0; a < 3; ++a) red[a][t] += red[a][t + w];
        __syncthreads();
    }
    if (t < 3) {  // component t
        const double u = um_a + A.kappa * red[t][0];
        const double gwv = (velw_a - u) / A.dt;
        um[3 * m + t] = u;
        gnext[3 * m + t] = gwv;
        const double gt = gtot_a + gwv;
        gtot[3 * m + t] = gt;
        Gw[3 * m + t] = Gw_a + -gwv * A.wscale;
        share[t] = gt;
    }
    if (!last) return;  // uniform
    __syncthreads();
    const double g0 = share[0], g1 = share[1], g2 = share[2];
    double* out = contrib + (size_t)m * A.S * 3;
    for (int j = 0; j < 2; ++j) {
        const int s = t + j * kIbLanes;
        if (s >= A.S) break;
        out[3 * s + 0] = oksl[j] ? g0 * A.spread * wsl[j] : 0.0;
        out[3 * s + 1] = oksl[j] ? g1 * A.spread * wsl[j] : 0.0;
        out[3 * s + 2] = oksl[j] ? g2 * A.spread * wsl[j] : 0.0;
    }
}

// ---- k_ib_gather: spread force of every cell of the box ----------------------------------------
// One wavefront per cell, lanes over the block slots (two each for the 125-slot support): cell c
// receives slot s of every marker whose nearest cell is c - d(s) (markers by index), and the
// lane sums are added in a fixed tree.  A cell that no marker block reaches leaves at once.
constexpr int kIbGatherLanes = 64;

__global__ void __launch_bounds__(kIbGatherLanes) k_ib_gather(IbArgs A, int it, const unsigned long long* __restrict__ sorted,
                                                               const int2* __restrict__ range, const double* __restrict__ contrib,
                                                               const int* __restrict__ box, double* __restrict__ Fbox,
                                                               const int* __restrict__ mark) {
    __shared__ double red[3][kIbGatherLanes];
    __shared__ int hit;
    const int64_t b = blockIdx.x;
    const int t = threadIdx.x;
    const int64_t vol = (int64_t)box[3] * box[4] * box[5];
    if (b >= vol || b >= A.vmax) return;  // whole workgroup: b is uniform
    if (!mark[b]) {  // no marker block covers this cell (marked by the first force pass)
        if (t == 0 && it == 0) { Fbox[3 * b] = 0.0; Fbox[3 * b + 1] = 0.0; Fbox[3 * b + 2] = 0.0; }
        return;
    }
    int c[3];
    c[2] = (int)(b % box[5]);
    c[1] = (int)((b / box[5]) % box[4]);
    c[0] = (int)(b / ((int64_t)box[4] * box[5]));
    for (int a = 0; a < 3; ++a) {
        c[a] += box[a];
        if (c[a] >= A.n[a]) c[a] -= A.n[a];
    }
    if (t == 0) hit = 0;
    __syncthreads();
    double F[3] = {0, 0, 0};
    bool any = false;
    for (int s = t; s < A.S; s += kIbGatherLanes) {
        int d[3], src[3];
        slot_offset(A, s, d);
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            src[a] = c[a] - d[a];
            if (src[a] < 0 || src[a] >= A.n[a]) {
                if (A.periodic[a]) src[a] += src[a] < 0 ? A.n[a] : -A.n[a];
                else ok = false;
            }
        }
        const int64_t bs = ok ? box_index(A, box, src) : -1;  // no marker is nearest to a cell outside the box
        if (bs >= 0 && bs < A.vmax) {
            const int2 r = range[bs];
            any = any || r.y > 0;
            for (int q = r.x; q < r.x + r.y; ++q) {
                const int m = (int)(sorted[q] & 0xffffffffull);
                const double* cm = contrib + ((size_t)m * A.S + s) * 3;
                F[0] += cm[0]; F[1] += cm[1]; F[2] += cm[2];
            }
        }
    }
    if (any) hit = 1;
    __syncthreads();
    if (!hit) {  // uniform: no contribution reaches this cell
        if (t == 0 && it == 0) { Fbox[3 * b] = 0.0; Fbox[3 * b + 1] = 0.0; Fbox[3 * b + 2] = 0.0; }
        return;
    }
    for (int a = 0; a < 3; ++a) red[a][t] = F[a];
    __syncthreads();
    for (int w = kIbGatherLanes / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int a = 0; a < 3; ++a) red[a][t] += red[a][t + w];
        __syncthreads();
    }
    if (t == 0)
        for (int a = 0; a < 3; ++a) Fbox[3 * b + a] = (it == 0 ? 0.0 : Fbox[3 * b + a]) + red[a][0];
}

// ---- k_ib_clear_runs: no marker run in any box cell, before a ranking writes its runs ---------
__global__ void k_ib_clear_runs(IbArgs A, int2* __restrict__ range) {
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b < A.vmax) range[b] = make_int2(0, 0);
}

// ---- k_ib_apply: update_velIBM (ibm.cu:346-365) on the box: f += w dt (c e . F) ---------------
// A moving body's marks are cleared for the next ranking of this set (a fixed body keeps its own).
template <typename ST, int Q>
__global__ void k_ib_apply(IbArgs A, ST* __restrict__ f, const int* __restrict__ box, const double* __restrict__ Fbox,
                           int* __restrict__ mark) {
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (A.moving && b < A.vmax) mark[b] = 0;
    const int64_t vol = (int64_t)box[3] * box[4] * box[5];
    if (b >= vol || b >= A.vmax) return;
    const double F[3] = {Fbox[3 * b], Fbox[3 * b + 1], Fbox[3 * b + 2]};
    if (F[0] == 0.0 && F[1] == 0.0 && F[2] == 0.0) return;
    int c[3];
    c[2] = (int)(b % box[5]);
    c[1] = (int)((b / box[5]) % box[4]);
    c[0] = (int)(b / ((int64_t)box[4] * box[5]));
    for (int a = 0; a < 3; ++a) {
        c[a] += box[a];
        if (c[a] >= A.n[a]) c[a] -= A.n[a];
    }
    const uint32_t idx = cell_of(A, c);
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const Dir e = Lattice<Q>::c(k);
        ST& fv = f[(uint32_t)k * A.ncell + idx];
        fv = (ST)((double)fv + Lattice<Q>::w(k) * A.dt * A.c * (e.x * F[0] + e.y * F[1] + e.z * F[2]));
    }
}

// rotation matrix of a quaternion (set_quaternionS, ibm.cu:16-28)
__device__ inline void rotation(const double* q, double* R) {
    const double a = q[0], b = q[1], c = q[2], d = q[3];
    R[0] = a * a - c * c - d * d + b * b; R[1] = 2 * (b * c - a * d);         R[2] = 2 * (b * d + a * c);
    R[3] = 2 * (b * c + a * d);         R[4] = a * a - d * d - b * b + c * c; R[5] = 2 * (c * d - a * b);
    R[6] = 2 * (b * d - a * c);         R[7] = 2 * (c * d + a * b);         R[8] = a * a - b * b - c * c + d * d;
}

// ---- k_ib_body: update_IBbody (ibm.cu:367-424) in one workgroup: the force and torque sums in a
// fixed tree, the body update on work-item 0
constexpr int kIbBodyThreads = 1024;

__device__ inline void ib_body_update(const IbArgs& A, double fx, double fy, double fz, double tx, double ty, double tz,
                                      double* body);

__global__ void __launch_bounds__(kIbBodyThreads) k_ib_body(IbArgs A, const double* __restrict__ Gw,
                                                            const double* __restrict__ posw, double* body) {
    __shared__ double red[6][kIbBodyThreads];
    const int t = threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const double* P = body + B_POS;
    for (int m = t; m < A.M; m += kIbBodyThreads) {
        const double* G = Gw + 3 * m;
        const double r[3] = {posw[3 * m] - P[0], posw[3 * m + 1] - P[1], posw[3 * m + 2] - P[2]};
        acc[0] += G[0]; acc[1] += G[1]; acc[2] += G[2];
        acc[3] += r[1] * G[2] - r[2] * G[1];
        acc[4] += r[2] * G[0] - r[0] * G[2];
        acc[5] += r[0] * G[1] - r[1] * G[0];
    }
    for (int q = 0; q < 6; ++q) red[q][t] = acc[q];
    __syncthreads();
    for (int w = kIbBodyThreads / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int q = 0; q < 6; ++q) red[q][t] += red[q][t + w];
        __syncthreads();
    }
    if (t == 0) ib_body_update(A, red[0][0], red[1][0], red[2][0], red[3][0], red[4][0], red[5][0], body);
}

// the Newton-Euler update of the body from its force and torque sums
__device__ inline void ib_body_update(const IbArgs& A, double fx, double fy, double fz, double tx, double ty, double tz,
                                      double* body) {
    const double sums[6] = {fx, fy, fz, tx, ty, tz};
    double* FB = body + B_FB;
    double* T = body + B_T;
    double FBold[3], Told[3], velBold[3], angold[3], Rold[9];
    rotation(body + B_QUATOLD, Rold);
    for (int a = 0; a < 3; ++a) {
        FBold[a] = FB[a]; Told[a] = T[a];
        velBold[a] = body[B_VEL + a]; angold[a] = body[B_OMEGA + a];
        FB[a] = sums[a];
        T[a] = sums[3 + a];
    }
    for (int a = 0; a < 3; ++a) FB[a] += (1.0 - A.rho_f / A.density) * A.mass * A.gravity[a];
    if (!A.moving) return;
    const double dt = A.dt;
    for (int a = 0; a < 3; ++a) {
        body[B_VEL + a] += dt * (3.0 * FB[a] - FBold[a]) / A.mass / 2.0;
        body[B_POS + a] += dt * (3.0 * body[B_VEL + a] - velBold[a]) / 2.0;
    }
    const double* R = body + B_R;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            body[B_OMEGA + i] += dt * (3.0 * R[j * 3 + i] * T[j] - Rold[j * 3 + i] * Told[j]) / A.inertia[i] / 2.0;
    double q[4], w[3], qo[4];
    for (int i = 0; i < 4; ++i) { q[i] = body[B_QUAT + i]; qo[i] = body[B_QUATOLD + i]; }
    for (int i = 0; i < 3; ++i) w[i] = body[B_OMEGA + i];
    double* Qn = body + B_QUAT;
    Qn[0] += dt * 3.0 * (-w[0] * q[1] - w[1] * q[2] - w[2] * q[3]) / 4.0;
    Qn[1] += dt * 3.0 * (w[0] * q[0] + w[2] * q[2] - w[1] * q[3]) / 4.0;
    Qn[2] += dt * 3.0 * (w[1] * q[0] - w[2] * q[1] + w[0] * q[3]) / 4.0;
    Qn[3] += dt * 3.0 * (w[2] * q[0] + w[1] * q[1] - w[0] * q[2]) / 4.0;
    Qn[0] -= dt * (-angold[0] * qo[1] - angold[1] * qo[2] - angold[2] * qo[3]) / 4.0;
    Qn[1] -= dt * (angold[0] * qo[0] + angold[2] * qo[2] - angold[1] * qo[3]) / 4.0;
    Qn[2] -= dt * (angold[1] * qo[0] - angold[2] * qo[1] + angold[0] * qo[3]) / 4.0;
    Qn[3] -= dt * (angold[2] * qo[0] + angold[1] * qo[1] - angold[0] * qo[2]) / 4.0;
    for (int i = 0; i < 4; ++i) body[B_QUATOLD + i] = q[i];
    rotation(Qn, body + B_R);
}

// ---- k_ib_points: update_IBpoint (ibm.cu:426-447) and search_IBlattice (ibm.cu:449-463) --------
template <int Q>
__global__ void k_ib_points(IbArgs A, const double* __restrict__ body, const double* __restrict__ oposw,
                            const double* __restrict__ onB, double* __restrict__ posw, double* __restrict__ nB,
                            double* __restrict__ velw, int* __restrict__ node) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= A.M) return;
    const double* R = body + B_R;
    const double* w = body + B_OMEGA;
    const double* op = oposw + 3 * m;
    const double* on = onB + 3 * m;
    const double wxo[3] = {w[1] * op[2] - w[2] * op[1], w[2] * op[0] - w[0] * op[2], w[0] * op[1] - w[1] * op[0]};
    double x[3];
    for (int i = 0; i < 3; ++i) {
        double p = body[B_POS + i], nv = 0.0;
        for (int j = 0; j < 3; ++j) {
            p += R[i * 3 + j] * op[j];
            nv += R[i * 3 + j] * on[j];
        }
        posw[3 * m + i] = p;
        nB[3 * m + i] = nv;
        velw[3 * m + i] = body[B_VEL + i] + R[i * 3 + 0] * wxo[0] + R[i * 3 + 1] * wxo[1] + R[i * 3 + 2] * wxo[2];
        x[i] = p;
    }
    // nearest cell among the current one and its lattice neighbours; a ghost sits on its owner
    // and never wins, a wrapped neighbour keeps its own (far) coordinates
    int best[3] = {node[3 * m], node[3 * m + 1], node[3 * m + 2]};
    const int cur[3] = {best[0], best[1], best[2]};
    auto d2 = [&](const int c[3]) {
        double s = 0;
        for (int a = 0; a < 3; ++a) s += (A.pos[a][c[a]] - x[a]) * (A.pos[a][c[a]] - x[a]);
        return s;
    };
    double bd = d2(cur);
#pragma unroll
    for (int k = 1; k < Q; ++k) {
        const Dir e = Lattice<Q>::c(k);
        int c[3] = {cur[0] + e.x, cur[1] + e.y, cur[2] + e.z};
        bool ok = true;
        for (int a = 0; a < 3; ++a)
            if (c[a] < 0 || c[a] >= A.n[a]) {
                if (A.periodic[a]) c[a] += c[a] < 0 ? A.n[a] : -A.n[a];
                else ok = false;
            }
        if (!ok) continue;
        const double dd = d2(c);
        if (dd < bd) { bd = dd; best[0] = c[0]; best[1] = c[1]; best[2] = c[2]; }
    }
    node[3 * m] = best[0]; node[3 * m + 1] = best[1]; node[3 * m + 2] = best[2];
}

}  // namespace ibk

namespace {

void rotation_host(const double* q, double* R) {
    const double a = q[0], b = q[1], c = q[2], d = q[3];
    R[0] = a * a - c * c - d * d + b * b; R[1] = 2 * (b * c - a * d);         R[2] = 2 * (b * d + a * c);
    R[3] = 2 * (b * c + a * d);         R[4] = a * a - d * d - b * b + c * c; R[5] = 2 * (c * d - a * b);
    R[6] = 2 * (b * d - a * c);         R[7] = 2 * (c * d + a * b);         R[8] = a * a - b * b - c * c + d * d;
}

}  // namespace

IbSolver::~IbSolver() {
    if (side_) (void)hipStreamSynchronize(side_);
    if (forced_) (void)hipEventDestroy(forced_);
    if (tail_done_) (void)hipEventDestroy(tail_done_);
    if (side_) (void)hipStreamDestroy(side_);
}

// The box that holds every marker's node block: the marker nodes span at most 2 rmax / dz + 3
// cells per axis (widths >= dz); a periodic axis may be covered whole when the body straddles its
// seam.  The smoothed profile of SPM_ellipse3D reaches sqrt(1.1025) semi-axes from the centre: the
// box is padded beyond the node blocks by what that exceeds of the marker radius (the library's
// other forms run over every cell).  Returns the cell count; ext and pad receive the extents.
int64_t IbSolver::box_extent(const ipl_ib_config& cfg, const int n[3], const int periodic[3], int dim, double dz,
                             double rmax, int ext[3], int* pad) {
    const bool pk4 = cfg.method == IPL_IB_DIRECT && cfg.kernel == IPL_IB_KERNEL_PESKIN4;
    const int span = pk4 ? 4 : 2;  // reach_hi - reach_lo
    int bpad = 0;
    if (cfg.spm == 1) {
        const double reach = std::sqrt(1.1025) * std::max(cfg.spm_a, cfg.spm_b);
        bpad = std::max(0, (int)std::ceil((reach - rmax) / dz) + 1);
    }
    int64_t vmax = 1;
    for (int a = 0; a < 3; ++a) {
        int e = (int)std::ceil(2.0 * rmax / dz) + 3 + span + 2 * bpad;
        if (dim == 2 && a == 1) e = 1;
        if (periodic[a] || e > n[a]) e = std::min(periodic[a] ? n[a] : e, n[a]);
        ext[a] = e;
        vmax *= e;
    }
    if (pad) *pad = bpad;
    return vmax;
}

// Multi-direct forcing runs its later passes in marker space where node weights are products of
// axis weights: every 3D block, and the 2D 5 x 5 block (the reference's 2D 3 x 3 block moves its
// (-1, -1) node to (+1, +1), which no product of axis weights gives); and where a node block meets
// another marker's through one periodic image only (periodic lattice axes of 9 cells or more).
bool IbSolver::marker_space_passes(const ipl_ib_config& cfg, int dim, const int n[3], const int periodic[3]) {
    if (cfg.method != IPL_IB_DIRECT || cfg.iterations <= 1) return false;
    if (dim == 2 && cfg.kernel != IPL_IB_KERNEL_PESKIN4) return false;
    for (int a = 0; a < 3; ++a)
        if (!(dim == 2 && a == 1) && periodic[a] && n[a] < 9) return false;
    return true;
}

// The device bytes attach() allocates, in the order it allocates them (each buffer at least 256 B,
// as the arena rounds empty requests).
int64_t IbSolver::estimate_bytes(const ipl_ib_config& cfg, const int n[3], const int periodic[3], int dim, double dz) {
    const int64_t M = cfg.n_markers;
    double rmax = 0;
    for (int64_t m = 0; m < M; ++m) {
        const double* op = cfg.marker_pos + 3 * m;
        rmax = std::max(rmax, std::sqrt(op[0] * op[0] + op[1] * op[1] + op[2] * op[2]));
    }
    int ext[3];
    const int64_t vmax = box_extent(cfg, n, periodic, dim, dz, rmax, ext, nullptr);
    const bool pk4 = cfg.method == IPL_IB_DIRECT && cfg.kernel == IPL_IB_KERNEL_PESKIN4;
    const int np = pk4 ? 5 : 3;
    const int64_t S = (int64_t)np * np * (dim == 3 ? np : 1);
    auto b = [](int64_t bytes) { return std::max<int64_t>(bytes, 256); };
    int64_t t = 0;
    for (int a = 0; a < 3; ++a) t += b((int64_t)(n[a] + 1) * 8);       // axis coordinates
    t += 6 * b(3 * M * 8) + b(3 * M * 4) + b(3 * M * S * 8);            // marker arrays, nodes, slot contributions
    if (cfg.method == IPL_IB_DIRECT) t += b(M * S * 8) + b(M * S * 4);  // first-pass slot weights and box indices
    if (marker_space_passes(cfg, dim, n, periodic))  // axis weights, 4 marker vectors, 2 x near lists
        t += b(M * ibk::kIbAxisWeights * 8) + 4 * b(3 * M * 8) + 2 * (b(M * ibk::kIbNearMax * 4) + b(M * 4));
    t += 2 * b(3 * vmax * 8) + 2 * b(vmax * 8) + 2 * b(vmax * 4);       // box force, velocity, 2 x (runs, marks)
    t += b(ibk::B_SIZE * 8) + 2 * (b(8 * 4) + b(M * 8));                // body, 2 x (box, sorted keys)
    return t;
}

Status IbSolver::attach(const ipl_ib_config& cfg, const IbGrid& g, const std::vector<double>* pos, hipStream_t st) {
    if (M_ > 0) return Status::err(IPL_ERR_ARGUMENT, "a body is already attached");
    mem_.bind(st);
    if (cfg.n_markers <= 0 || !cfg.marker_pos || !cfg.marker_normal)
        return Status::err(IPL_ERR_ARGUMENT, "markers missing");
    if (cfg.n_markers > kIbMaxMarkers) return Status::err(IPL_ERR_UNSUPPORTED, "more than 4096 markers");
    if (cfg.method != IPL_IB_STRESS && cfg.method != IPL_IB_DIRECT) return Status::err(IPL_ERR_ARGUMENT, "IB method");
    if (cfg.method == IPL_IB_STRESS && cfg.kernel != IPL_IB_KERNEL_REFERENCE)
        return Status::err(IPL_ERR_ARGUMENT, "the stress method uses the reference weights");
    if (cfg.kernel != IPL_IB_KERNEL_REFERENCE && cfg.kernel != IPL_IB_KERNEL_PESKIN4)
        return Status::err(IPL_ERR_ARGUMENT, "IB kernel");
    if (cfg.spm < 0 || cfg.spm > 3) return Status::err(IPL_ERR_ARGUMENT, "smoothed-profile form");
    if (cfg.spm == 1 && cfg.method != IPL_IB_STRESS)
        return Status::err(IPL_ERR_ARGUMENT, "SPM_ellipse3D runs with the stress method");
    if (cfg.spm && (!(cfg.spm_a > 0) || (cfg.spm != 3 && !(cfg.spm_b > 0))))
        return Status::err(IPL_ERR_ARGUMENT, "the smoothed profile needs positive semi-axes (radius)");
    if (!(cfg.mass > 0) || !(cfg.density > 0) || !(cfg.spacing > 0))
        return Status::err(IPL_ERR_ARGUMENT, "body mass, density and marker spacing must be positive");
    if (cfg.moving && !(cfg.inertia[0] > 0 && cfg.inertia[1] > 0 && cfg.inertia[2] > 0))
        return Status::err(IPL_ERR_ARGUMENT, "a moving body needs positive moments of inertia");
    g_ = g;
    cfg_ = cfg;
    st_ = st;
    M_ = cfg.n_markers;
    const bool pk4 = cfg.method == IPL_IB_DIRECT && cfg.kernel == IPL_IB_KERNEL_PESKIN4;
    reach_lo_ = pk4 ? -2 : -1;
    reach_hi_ = pk4 ? 2 : 1;
    const int np = reach_hi_ - reach_lo_ + 1;
    S_ = np * np * (g.dim == 3 ? np : 1);  // at most 125 <= the gather's lanes
    Status r;
    // axis coordinates; entry n holds the period (sum of widths) for the seam unwrap
    for (int a = 0; a < 3; ++a) {
        std::vector<double> t(pos[a].begin(), pos[a].end());
        t.push_back(g.period[a]);
        if (!(r = mem_.alloc_n(&pos_[a], t.size()))) return r;
        IPL_HIP_TRY(hipMemcpyAsync(pos_[a], t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, st));
        IPL_HIP_TRY(hipStreamSynchronize(st));
    }
    // body frame -> world: markers, unit normals, nearest cells (per-axis nearest centre, lower
    // index on ties: the first minimum of the drivers' exhaustive search, SphereFall.cu:250-257)
    double R[9];
    rotation_host(cfg.quat, R);
    std::vector<double> posw(3 * M_), nB(3 * M_), velw(3 * M_, 0.0);
    std::vector<int> node(3 * M_);
    double rmax = 0;
    for (int m = 0; m < M_; ++m) {
        const double* op = cfg.marker_pos + 3 * m;
        const double* on = cfg.marker_normal + 3 * m;
        const double nn = std::sqrt(on[0] * on[0] + on[1] * on[1] + on[2] * on[2]);
        rmax = std::max(rmax, std::sqrt(op[0] * op[0] + op[1] * op[1] + op[2] * op[2]));
        for (int i = 0; i < 3; ++i) {
            double p = 0, nv = 0;
            for (int j = 0; j < 3; ++j) { p += R[i * 3 + j] * op[j]; nv += R[i * 3 + j] * on[j]; }
            nB[3 * m + i] = nv / nn;
            posw[3 * m + i] = cfg.pos[i] + p;
        }
        for (int a = 0; a < 3; ++a) {
            int best = 0;
            double bd = 1e300;
            for (int x = 0; x < g.n[a]; ++x) {
                const double d = (pos[a][x] - posw[3 * m + a]) * (pos[a][x] - posw[3 * m + a]);
                if (d < bd) { bd = d; best = x; }
            }
            node[3 * m + a] = best;
        }
    }
    vmax_ = box_extent(cfg, g.n, g.periodic, g.dim, g.dz, rmax, box_max_, &bpad_);
    std::vector<double> body(ibk::B_SIZE, 0.0);
    for (int a = 0; a < 3; ++a) {
        body[ibk::B_POS + a] = cfg.pos[a];
        body[ibk::B_VEL + a] = cfg.vel[a];
        body[ibk::B_OMEGA + a] = cfg.omega[a];
    }
    for (int i = 0; i < 4; ++i) { body[ibk::B_QUAT + i] = cfg.quat[i]; body[ibk::B_QUATOLD + i] = cfg.quat[i]; }
    for (int i = 0; i < 9; ++i) body[ibk::B_R + i] = R[i];
    if (!(r = mem_.alloc_n(&oposw_, 3 * M_))) return r;
    if (!(r = mem_.alloc_n(&onB_, 3 * M_))) return r;
    if (!(r = mem_.alloc_n(&posw_, 3 * M_))) return r;
    if (!(r = mem_.alloc_n(&nB_, 3 * M_))) return r;
    if (!(r = mem_.alloc_n(&velw_, 3 * M_))) return r;
    if (!(r = mem_.alloc_n(&Gw_, 3 * M_))) return r;
    if (!(r = mem_.alloc_n(&node_, 3 * M_))) return r;
    if (!(r = mem_.alloc_n(&contrib_, (size_t)3 * M_ * S_))) return r;
    if (cfg.method == IPL_IB_DIRECT) {  // slot weights and box indices of the first forcing pass
        if (!(r = mem_.alloc_n(&wcache_, (size_t)M_ * S_))) return r;
        if (!(r = mem_.alloc_n(&bcache_, (size_t)M_ * S_))) return r;
    }
    if (marker_space_passes(cfg, g.dim, g.n, g.periodic)) {  // axis weights, velocity, pass forces, summed force
        if (!(r = mem_.alloc_n(&wax_, (size_t)M_ * ibk::kIbAxisWeights))) return r;
        if (!(r = mem_.alloc_n(&um_, (size_t)3 * M_))) return r;
        for (int q = 0; q < 2; ++q)
            if (!(r = mem_.alloc_n(&gm_[q], (size_t)3 * M_))) return r;
        if (!(r = mem_.alloc_n(&gtot_, (size_t)3 * M_))) return r;
        for (int q = 0; q < 2; ++q) {
            if (!(r = mem_.alloc_n(&near_[q], (size_t)M_ * ibk::kIbNearMax))) return r;
            if (!(r = mem_.alloc_n(&near_count_[q], (size_t)M_))) return r;
        }
    }
    if (!(r = mem_.alloc_n(&Fbox_, (size_t)3 * vmax_))) return r;
    if (!(r = mem_.alloc_n(&Ubox_, (size_t)3 * vmax_))) return r;
    for (int q = 0; q < 2; ++q)
        if (!(r = mem_.alloc_n(&range_[q], (size_t)vmax_))) return r;
    for (int q = 0; q < 2; ++q)
        if (!(r = mem_.alloc_n(&mark_[q], (size_t)vmax_))) return r;
    if (!(r = mem_.alloc_n(&body_, (size_t)ibk::B_SIZE))) return r;
    for (int q = 0; q < 2; ++q) {
        if (!(r = mem_.alloc_n(&box_[q], 8))) return r;
        if (!(r = mem_.alloc_n(&sorted_[q], (size_t)M_))) return r;
    }
    IPL_HIP_TRY(hipMemcpyAsync(oposw_, cfg.marker_pos, 3 * M_ * sizeof(double), hipMemcpyHostToDevice, st));
    IPL_HIP_TRY(hipMemcpyAsync(onB_, cfg.marker_normal, 3 * M_ * sizeof(double), hipMemcpyHostToDevice, st));
    IPL_HIP_TRY(hipMemcpyAsync(posw_, posw.data(), 3 * M_ * sizeof(double), hipMemcpyHostToDevice, st));
    IPL_HIP_TRY(hipMemcpyAsync(nB_, nB.data(), 3 * M_ * sizeof(double), hipMemcpyHostToDevice, st));
    IPL_HIP_TRY(hipMemcpyAsync(velw_, velw.data(), 3 * M_ * sizeof(double), hipMemcpyHostToDevice, st));
    IPL_HIP_TRY(hipMemsetAsync(Gw_, 0, 3 * M_ * sizeof(double), st));
    IPL_HIP_TRY(hipMemcpyAsync(node_, node.data(), 3 * M_ * sizeof(int), hipMemcpyHostToDevice, st));
    IPL_HIP_TRY(hipMemcpyAsync(body_, body.data(), ibk::B_SIZE * sizeof(double), hipMemcpyHostToDevice, st));
    for (int q = 0; q < 2; ++q) IPL_HIP_TRY(hipMemsetAsync(box_[q], 0, 8 * sizeof(int), st));
    for (int q = 0; q < 2; ++q)  // later cleared by k_ib_apply (moving body)
        IPL_HIP_TRY(hipMemsetAsync(mark_[q], 0, (size_t)vmax_ * sizeof(int), st));
    IPL_HIP_TRY(hipStreamSynchronize(st));
    ranked_[0] = ranked_[1] = false;
    weights_set_ = -1;
    // the stream of the step's tail (default priority, like every stream of the library)
    IPL_HIP_TRY(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    IPL_HIP_TRY(hipEventCreateWithFlags(&forced_, hipEventDisableTiming));
    IPL_HIP_TRY(hipEventCreateWithFlags(&tail_done_, hipEventDisableTiming));
    return Status::ok();
}

namespace {

ibk::IbArgs make_args(const IbGrid& g, const ipl_ib_config& cfg, int M, int S, int lo, int hi, int64_t vmax,
                      double* const pos[3], int bpad) {
    ibk::IbArgs A{};
    for (int a = 0; a < 3; ++a) { A.n[a] = g.n[a]; A.periodic[a] = g.periodic[a]; A.pos[a] = pos[a]; }
    A.s = g.s; A.o1 = g.o1; A.o2 = g.o2; A.ns = g.ns;
    A.dim = g.dim; A.Qs = g.Qs; A.scalar = g.scalar;
    A.ncell = (uint32_t)g.ncell;
    A.dz = g.dz; A.dt = g.dt; A.c = g.c; A.tau = g.tau;
    A.rho_L = g.rho_L; A.drho = g.rho_H - g.rho_L; A.beta_s = g.beta_s;
    A.M = M; A.S = S; A.lo = lo; A.hi = hi;
    A.block3 = cfg.kernel == IPL_IB_KERNEL_REFERENCE;
    A.method = cfg.method; A.kernel = cfg.kernel; A.moving = cfg.moving;
    const double c2 = g.c * g.c;
    const int dim = g.dim;
    if (cfg.method == IPL_IB_STRESS) {
        A.spread = std::pow(cfg.spacing / g.dz, dim - 1) / g.dz;                 // ibm.cu:239-241
        A.gwscale = cfg.rho_f * c2 / 3.0 * std::pow(cfg.spacing, dim - 1);      // ibm.cu:270-272
    } else {
        A.spread = std::pow(cfg.spacing / g.dz, dim - 1) * 3.0 / c2;             // ibm.cu:336-338
        A.wscale = cfg.rho_f * std::pow(cfg.spacing, dim - 1) * g.dz;           // ibm.cu:342
    }
    A.sina = std::sin(22.5 / 180.0 * 3.14159);  // the reference's literal angle (ibm.cu:201)
    A.inv_c = 1.0 / g.c;
    A.inv_c2 = 1.0 / c2;
    A.half_itau = 0.5 / g.tau;
    A.rho_f = cfg.rho_f; A.density = cfg.density; A.mass = cfg.mass;
    for (int a = 0; a < 3; ++a) { A.inertia[a] = cfg.inertia[a]; A.gravity[a] = cfg.gravity[a]; }
    A.vmax = vmax;
    A.spm = cfg.spm; A.bpad = bpad; A.spm_a = cfg.spm_a; A.spm_b = cfg.spm_b;
    A.mds = IbSolver::marker_space_passes(cfg, g.dim, g.n, g.periodic);
    A.kappa = c2 * g.dt * A.spread / 3.0;
    return A;
}

template <typename ST, typename MT>
void force_q(int Q, const ibk::IbArgs& A, int it, const ST* f, const int* node, const double* posw, const double* nB,
             const double* velw, double* Gw, double* contrib, const int* box, const double* Fbox, const double* Ubox,
             const double* wcache, const int* bcache, int* mark, double* um, double* gnext, double* gtot, hipStream_t st) {
#define IPL_IBF(QQ) IPL_LAUNCH_SYNC((ibk::k_ib_force<ST, MT, QQ>), A.M, ibk::kIbLanes, 0, st, A, it, f, node, posw, nB, velw, Gw, contrib, box, Fbox, Ubox, wcache, bcache, mark, um, gnext, gtot)
    if (Q == 9) IPL_IBF(9); else if (Q == 19) IPL_IBF(19); else IPL_IBF(27);
#undef IPL_IBF
}

template <typename ST, typename MT>
void ubase_q(int Q, const ibk::IbArgs& A, const ST* f, const ST* g, const MT* F, const int* box, double* Ubox, hipStream_t st) {
    const unsigned blocks = (unsigned)((A.vmax + 255) / 256);
    if (Q == 9) IPL_LAUNCH((ibk::k_ib_ubase<ST, MT, 9>), blocks, 256, 0, st, A, f, g, F, box, Ubox);
    else if (Q == 19) IPL_LAUNCH((ibk::k_ib_ubase<ST, MT, 19>), blocks, 256, 0, st, A, f, g, F, box, Ubox);
    else IPL_LAUNCH((ibk::k_ib_ubase<ST, MT, 27>), blocks, 256, 0, st, A, f, g, F, box, Ubox);
}

template <typename ST>
void apply_q(int Q, const ibk::IbArgs& A, ST* f, const int* box, const double* Fbox, int* mark, hipStream_t st) {
    const unsigned blocks = (unsigned)((A.vmax + 255) / 256);
    if (Q == 9) IPL_LAUNCH((ibk::k_ib_apply<ST, 9>), blocks, 256, 0, st, A, f, box, Fbox, mark);
    else if (Q == 19) IPL_LAUNCH((ibk::k_ib_apply<ST, 19>), blocks, 256, 0, st, A, f, box, Fbox, mark);
    else IPL_LAUNCH((ibk::k_ib_apply<ST, 27>), blocks, 256, 0, st, A, f, box, Fbox, mark);
}

}  // namespace

Status IbSolver::substep(void* f, const void* g, const void* F, int storage, int parity) {
    if (M_ == 0) return Status::ok();
    const int set = set_of(parity);
    Status r = join();  // the previous step's tail: body, points, this step's ranking
    if (r) r = ensure_ranked(parity);
    if (r && cfg_.spm) r = spm(f, g, F, storage, set);  // before the marker forcing (SphereFall.cu:376-377)
    // the forcing passes, with the tail forked after the last force pass (it ranks the next step's
    // set when the body moves), then update_velIBM
    if (r) r = forcing(f, g, F, storage, set, cfg_.moving ? set_of(parity + 1) : set);
    if (r) r = update_vel(f, storage, set);
    if (r && cfg_.moving) ranked_[set] = false;  // the nodes of set's ranking have moved
    return r;
}

Status IbSolver::join() {
    if (!tail_pending_) return Status::ok();
    IPL_HIP_TRY(hipStreamWaitEvent(st_, tail_done_, 0));
    tail_pending_ = false;
    return Status::ok();
}

Status IbSolver::ensure_ranked(int parity) {
    if (M_ == 0) return Status::ok();
    const int set = set_of(parity);
    const bool weighted = cfg_.method != IPL_IB_DIRECT || weights_set_ == set;  // the caches hold set's weights
    return ranked_[set] && weighted ? Status::ok() : rank(set, st_);
}

// One stage of the sub-step (ipl_ib_stage) on the context's stream: the stages that read the
// marker runs and the box rank the markers first, as the sub-step has them ranked.
Status IbSolver::stage(int which, void* f, const void* g, const void* F, int storage, int parity) {
    if (M_ == 0) return Status::err(IPL_ERR_ARGUMENT, "no body attached");
    const int set = set_of(parity);
    Status r = join();
    if (!r) return r;
    switch (which) {
    case IPL_IB_STAGE_SPM:
        if (!cfg_.spm) return Status::err(IPL_ERR_ARGUMENT, "the body has no smoothed-profile forcing");
        r = rank(set, st_);
        return r ? spm(f, g, F, storage, set) : r;
    case IPL_IB_STAGE_FORCE:
        r = rank(set, st_);
        return r ? forcing(f, g, F, storage, set, -1) : r;
    case IPL_IB_STAGE_UPDATE_VEL: return update_vel(f, storage, set);
    case IPL_IB_STAGE_BODY: return body(st_);
    case IPL_IB_STAGE_POINTS:
        r = points(st_);
        if (r && cfg_.moving) ranked_[0] = ranked_[1] = false;
        return r;
    default: return Status::err(IPL_ERR_ARGUMENT, "IB stage id");
    }
}

// markers sorted by nearest cell, the box of their node blocks and each box cell's marker run,
// into ranking set `set`
Status IbSolver::rank(int set, hipStream_t st) {
    const ibk::IbArgs A = make_args(g_, cfg_, M_, S_, reach_lo_, reach_hi_, vmax_, pos_, bpad_);
    IPL_LAUNCH(ibk::k_ib_clear_runs, (unsigned)((vmax_ + 255) / 256), 256, 0, st, A, range_[set]);
    constexpr int kPad = 2 * ibk::kIbRankParts;
    const int mpad = (M_ + kPad - 1) / kPad * kPad;
    IPL_LAUNCH_SYNC(ibk::k_ib_rank, (unsigned)((M_ + ibk::kIbRankMarkers - 1) / ibk::kIbRankMarkers), ibk::kIbRankThreads,
                    (size_t)mpad * sizeof(unsigned long long), st, A, node_, sorted_[set], box_[set], range_[set], mpad);
    if (cfg_.method == IPL_IB_DIRECT)  // the node weights of the forcing passes
        IPL_LAUNCH_SYNC(ibk::k_ib_weights, (unsigned)M_, 64, 0, st, A, node_, posw_, box_[set], wcache_, bcache_, wax_, mark_[set]);
    if (A.mds)  // the near-marker lists of the marker-space passes
        IPL_LAUNCH_SYNC(ibk::k_ib_near, (unsigned)M_, ibk::kIbLanes, 0, st, A, node_, near_[set], near_count_[set]);
    weights_set_ = set;
    IPL_HIP_TRY(hipGetLastError());
    ranked_[set] = true;
    return Status::ok();
}

// SPM_ellipse3D (SphereFall3D/SphereFall.cu:13-30) on the box
Status IbSolver::spm(void* f, const void* g, const void* F, int storage, int set) {
    const ibk::IbArgs A = make_args(g_, cfg_, M_, S_, reach_lo_, reach_hi_, vmax_, pos_, bpad_);
    const int64_t items = cfg_.spm == 1 ? vmax_ : (int64_t)g_.ncell;  // the box, or every cell
    const unsigned blocks = (unsigned)((items + 255) / 256);
    int* box = box_[set];
#define IPL_IBS(ST, MT, QQ) IPL_LAUNCH((ibk::k_ib_spm<ST, MT, QQ>), blocks, 256, 0, st_, A, (ST*)f, (const ST*)g, (const MT*)F, box, body_, velw_)
    if (storage == IPL_F64) { if (g_.Q == 9) IPL_IBS(double, double, 9); else if (g_.Q == 19) IPL_IBS(double, double, 19); else IPL_IBS(double, double, 27); }
    else { if (g_.Q == 9) IPL_IBS(float, float, 9); else if (g_.Q == 19) IPL_IBS(float, float, 19); else IPL_IBS(float, float, 27); }
#undef IPL_IBS
    IPL_HIP_TRY(hipGetLastError());
    return Status::ok();
}

// marker forces (get_IBMGw2 / IB_directForcing, repeated for multi-direct forcing) and their
// spread onto the box (needs ranking set `set` of the current nodes).  tail_set >= 0: once the last
// force pass is queued, the step's tail goes to the second stream (update_IBbody reads the final
// Gw; a moving body's update_IBpoint, search_IBlattice and ranking into tail_set follow), while
// the last gather pass runs here.
Status IbSolver::forcing(void* f, const void* g, const void* F, int storage, int set, int tail_set) {
    const ibk::IbArgs A = make_args(g_, cfg_, M_, S_, reach_lo_, reach_hi_, vmax_, pos_, bpad_);
    const int iters = cfg_.method == IPL_IB_DIRECT ? std::max(1, cfg_.iterations) : 1;
    int* box = box_[set];
    if (cfg_.method == IPL_IB_DIRECT) {
        if (storage == IPL_F64) ubase_q<double, double>(g_.Q, A, (const double*)f, (const double*)g, (const double*)F, box, Ubox_, st_);
        else ubase_q<float, float>(g_.Q, A, (const float*)f, (const float*)g, (const float*)F, box, Ubox_, st_);
    }
    // marker-space passes: the first force pass, the later passes in marker space, one gather of
    // the summed force
    const int spreads = A.mds ? 1 : iters;
    for (int it = 0; it < spreads; ++it) {
        if (storage == IPL_F64)
            force_q<double, double>(g_.Q, A, it, (const double*)f, node_, posw_, nB_, velw_, Gw_, contrib_, box, Fbox_, Ubox_,
                                    wcache_, bcache_, mark_[set], um_, gm_[0], gtot_, st_);
        else
            force_q<float, float>(g_.Q, A, it, (const float*)f, node_, posw_, nB_, velw_, Gw_, contrib_, box, Fbox_, Ubox_,
                                  wcache_, bcache_, mark_[set], um_, gm_[0], gtot_, st_);
        for (int p = 1; A.mds && p < iters; ++p)
            IPL_LAUNCH_SYNC(ibk::k_ib_correct, (unsigned)M_, ibk::kIbLanes, 0, st_, A, p == iters - 1 ? 1 : 0, node_,
                            near_[set], near_count_[set], velw_,
                            wax_, um_, gm_[(p - 1) & 1], gm_[p & 1], gtot_, Gw_, wcache_, bcache_, contrib_);
        const bool fork = it == spreads - 1 && tail_set >= 0 && !inline_tail_;
        if (fork) IPL_HIP_TRY(hipEventRecord(forced_, st_));
        IPL_LAUNCH_SYNC(ibk::k_ib_gather, (unsigned)vmax_, ibk::kIbGatherLanes, 0, st_, A, it, sorted_[set], range_[set], contrib_,
                        box, Fbox_, mark_[set]);
        if (fork) {  // queued after the last gather pass, which it does not wait for
            IPL_HIP_TRY(hipStreamWaitEvent(side_, forced_, 0));
            Status r = body(side_);
            if (r && cfg_.moving) r = points(side_);
            if (r && cfg_.moving) r = rank(tail_set, side_);
            if (!r) return r;
            IPL_HIP_TRY(hipEventRecord(tail_done_, side_));
            tail_pending_ = true;
        }
    }
    if (tail_set >= 0 && inline_tail_) {  // the tail after the last gather pass, on this stream
        Status r = body(st_);
        if (r && cfg_.moving) r = points(st_);
        if (r && cfg_.moving) r = rank(tail_set, st_);
        if (!r) return r;
    }
    IPL_HIP_TRY(hipGetLastError());
    return Status::ok();
}

// update_velIBM (ibm.cu:346-365) with the spread force of forcing(); clears the marks
Status IbSolver::update_vel(void* f, int storage, int set) {
    const ibk::IbArgs A = make_args(g_, cfg_, M_, S_, reach_lo_, reach_hi_, vmax_, pos_, bpad_);
    if (storage == IPL_F64) apply_q<double>(g_.Q, A, (double*)f, box_[set], Fbox_, mark_[set], st_);
    else apply_q<float>(g_.Q, A, (float*)f, box_[set], Fbox_, mark_[set], st_);
    IPL_HIP_TRY(hipGetLastError());
    return Status::ok();
}

// update_IBbody (ibm.cu:367-424)
Status IbSolver::body(hipStream_t st) {
    const ibk::IbArgs A = make_args(g_, cfg_, M_, S_, reach_lo_, reach_hi_, vmax_, pos_, bpad_);
    IPL_LAUNCH_SYNC(ibk::k_ib_body, 1, ibk::kIbBodyThreads, 0, st, A, Gw_, posw_, body_);
    IPL_HIP_TRY(hipGetLastError());
    return Status::ok();
}

// update_IBpoint + search_IBlattice (ibm.cu:426-463) of a moving body
Status IbSolver::points(hipStream_t st) {
    const ibk::IbArgs A = make_args(g_, cfg_, M_, S_, reach_lo_, reach_hi_, vmax_, pos_, bpad_);
    if (cfg_.moving) {
        const unsigned blocks = (unsigned)((M_ + 63) / 64);
        if (g_.Q == 9) IPL_LAUNCH((ibk::k_ib_points<9>), blocks, 64, 0, st, A, body_, oposw_, onB_, posw_, nB_, velw_, node_);
        else if (g_.Q == 19) IPL_LAUNCH((ibk::k_ib_points<19>), blocks, 64, 0, st, A, body_, oposw_, onB_, posw_, nB_, velw_, node_);
        else IPL_LAUNCH((ibk::k_ib_points<27>), blocks, 64, 0, st, A, body_, oposw_, onB_, posw_, nB_, velw_, node_);
    }
    IPL_HIP_TRY(hipGetLastError());
    return Status::ok();
}

Status IbSolver::check() {
    if (M_ == 0) return Status::ok();
    Status r = join();
    if (!r) return r;
    int flag[2] = {0, 0};
    for (int q = 0; q < 2; ++q) IPL_HIP_TRY(hipMemcpyAsync(&flag[q], box_[q] + 6, sizeof(int), hipMemcpyDeviceToHost, st_));
    IPL_HIP_TRY(hipStreamSynchronize(st_));
    if (flag[0] || flag[1]) return Status::err(IPL_ERR_UNSUPPORTED, "the marker node blocks outgrew the box sized at attach");
    return Status::ok();
}

Status IbSolver::body_state(double out[19]) {
    if (M_ == 0) return Status::err(IPL_ERR_ARGUMENT, "no body attached");
    double b[ibk::B_SIZE];
    IPL_HIP_TRY(hipMemcpyAsync(b, body_, sizeof(b), hipMemcpyDeviceToHost, st_));
    IPL_HIP_TRY(hipStreamSynchronize(st_));
    for (int a = 0; a < 3; ++a) { out[a] = b[ibk::B_POS + a]; out[3 + a] = b[ibk::B_VEL + a]; out[6 + a] = b[ibk::B_OMEGA + a]; }
    for (int i = 0; i < 4; ++i) out[9 + i] = b[ibk::B_QUAT + i];
    for (int a = 0; a < 3; ++a) { out[13 + a] = b[ibk::B_FB + a]; out[16 + a] = b[ibk::B_T + a]; }
    return Status::ok();
}

Status IbSolver::read_markers(int which, double* out, size_t count) {
    if (M_ == 0) return Status::err(IPL_ERR_ARGUMENT, "no body attached");
    if (which == IPL_IB_NODE) {
        if (count != (size_t)M_) return Status::err(IPL_ERR_ARGUMENT, "marker array size mismatch");
        std::vector<int> node(3 * M_);
        IPL_HIP_TRY(hipMemcpyAsync(node.data(), node_, node.size() * sizeof(int), hipMemcpyDeviceToHost, st_));
        IPL_HIP_TRY(hipStreamSynchronize(st_));
        for (int m = 0; m < M_; ++m)
            out[m] = (double)(((int64_t)node[3 * m] * g_.n[1] + node[3 * m + 1]) * g_.n[2] + node[3 * m + 2]);
        return Status::ok();
    }
    const double* src = which == IPL_IB_POS ? posw_ : which == IPL_IB_VEL ? velw_ : which == IPL_IB_FORCE ? Gw_ :
                        which == IPL_IB_NORMAL ? nB_ : nullptr;
    if (!src) return Status::err(IPL_ERR_ARGUMENT, "marker array id");
    if (count != (size_t)3 * M_) return Status::err(IPL_ERR_ARGUMENT, "marker array size mismatch");
    IPL_HIP_TRY(hipMemcpyAsync(out, src, count * sizeof(double), hipMemcpyDeviceToHost, st_));
    IPL_HIP_TRY(hipStreamSynchronize(st_));
    return Status::ok();
}

}  // namespace ipl
