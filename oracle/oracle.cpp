// oracle.cpp - FP64 CPU restatement of the reference's lattice Boltzmann step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP library in
// ip_cu_lbm_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it, and only as the checker (or as the timed CPU baseline).  The product never calls it.
//
// It restates, in the reference's own data model (array-of-structures populations f[i*Q+k],
// an int32 neighbour table neib[i*Q+k], one ghost cell appended per boundary cell), the
// kernels of function/{collision,adv,propagation,set_bound,basic}.cu and the step order of
// the case drivers.  Every device kernel becomes one OpenMP loop with the same per-thread
// body; citations are path:line into the reference.
//
// Deviations from the reference, all deliberate and shared with the GPU library:
//  - true FP64 everywhere: the reference's <double> instantiation rounds through float in
//    fminf/fmaxf/powf/sqrtf and a 'float tmp' (adv.cu:17-38, collision.cu:175,213-215);
//  - builder extensions the reference lacks: periodic axes (neighbour wrap + position unwrap
//    in the stencil), a single flux axis along y or z (the reference streams z exactly and
//    advects only x/y), D3Q19 and D3Q7 lattices, a 3-point stencil, gravity / beta_s as
//    parameters instead of the literals 9.81 / 0.824493.
//
// Parity status per function is listed at the end of this file (PARITY block).
//
// ORC_MUTANT (a compile-time number, 0 in every normal build) plants one known mistake, so that
// tests/test_oracle_pins.py can show that each closed-form pin fails on a plausible error in the
// function it pins.  The mutants are listed where they act; none is in the normal build.

#include "../include/ipl.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
constexpr int kMutant = ORC_MUTANT;

// ------------------------------------------------------------------------------------------
// Lattice constants: Items::setc (basic.cu:70-98).  D2Q9 lives in the x-z plane.  D3Q19 and
// D3Q7 are prefixes of the reference's D3Q27 ordering (rest, +-x/+-y/+-z, xy, xz, yz edges).
// ------------------------------------------------------------------------------------------
struct Lattice {
    int Q = 0;
    int dim = 0;
    int cx[27] = {0}, cy[27] = {0}, cz[27] = {0};
    double w[27] = {0};
    int kinv[27] = {0};
};

Lattice make_lattice(int Q) {
    Lattice L;
    L.Q = Q;
    if (Q == 9) {
        L.dim = 2;
        const int x[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
        const int z[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
        for (int k = 0; k < 9; ++k) { L.cx[k] = x[k]; L.cz[k] = z[k]; }
        L.w[0] = 4.0 / 9.0;
        for (int k = 1; k < 5; ++k) { L.w[k] = 1.0 / 9.0; L.w[k + 4] = 1.0 / 36.0; }
    } else {
        L.dim = 3;
        // basic.cu:89-94
        const int x[27] = {0, 1, 0, -1, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1};
        const int y[27] = {0, 0, 1, 0, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1};
        const int z[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1};
        for (int k = 0; k < Q; ++k) { L.cx[k] = x[k]; L.cy[k] = y[k]; L.cz[k] = z[k]; }
        if (Q == 27) {
            L.w[0] = 8.0 / 27.0;
            for (int k = 1; k < 7; ++k) L.w[k] = 2.0 / 27.0;
            for (int k = 7; k < 19; ++k) L.w[k] = 1.0 / 54.0;
            for (int k = 19; k < 27; ++k) L.w[k] = 1.0 / 216.0;
        } else if (Q == 19) {
            L.w[0] = 1.0 / 3.0;
            for (int k = 1; k < 7; ++k) L.w[k] = 1.0 / 18.0;
            for (int k = 7; k < 19; ++k) L.w[k] = 1.0 / 36.0;
        } else if (Q == 7) {
            L.w[0] = 1.0 / 4.0;
            for (int k = 1; k < 7; ++k) L.w[k] = 1.0 / 8.0;
        }
    }
    // set_k_inv (basic.cu:112-121): opposite direction
    for (int k = 0; k < Q; ++k)
        for (int j = 0; j < Q; ++j)
            if (L.cx[j] == -L.cx[k] && L.cy[j] == -L.cy[k] && L.cz[j] == -L.cz[k]) L.kinv[k] = j;
    return L;
}

// ------------------------------------------------------------------------------------------
// MRT moment matrices.  Rows are the moment polynomials of the reference's MRT_D3Q9.csv and
// MRT_D3Q27.csv (checked against those files by tests/test_oracle.py); D3Q19 uses the
// d'Humieres et al. (2002) basis in the same direction ordering.  M^-1 = M^T diag(1/|row|^2)
// because the rows are orthogonal.  set_M (basic.cu:171-201) then forms
// MM = M^-1 diag(S) M with the stress rates zeroed; the stress rows use 1/tau per cell.
// ------------------------------------------------------------------------------------------
void mrt_matrix(const Lattice& L, std::vector<double>& M, std::vector<int>& stress) {
    const int Q = L.Q;
    M.assign(Q * Q, 0.0);
    stress.clear();
    for (int k = 0; k < Q; ++k) {
        const double x = L.cx[k], y = L.cy[k], z = L.cz[k];
        const double c2 = x * x + y * y + z * z;
        double r[27];
        if (Q == 9) {
            r[0] = 1; r[1] = -4 + 3 * c2; r[2] = 4 - 10.5 * c2 + 4.5 * c2 * c2;
            r[3] = x; r[4] = (-5 + 3 * c2) * x; r[5] = z; r[6] = (-5 + 3 * c2) * z;
            r[7] = x * x - z * z; r[8] = x * z;
        } else if (Q == 27) {
            r[0] = 1; r[1] = x; r[2] = y; r[3] = z; r[4] = c2 - 2;
            r[5] = 3 * x * x - c2; r[6] = y * y - z * z; r[7] = x * y; r[8] = y * z; r[9] = x * z;
            r[10] = x * (3 * c2 - 7); r[11] = y * (3 * c2 - 7); r[12] = z * (3 * c2 - 7);
            const double p4 = (9 * c2 * c2 - 39 * c2 + 38) / 2;
            r[13] = x * p4; r[14] = y * p4; r[15] = z * p4;
            r[16] = (3 * c2 * c2 - 11 * c2 + 8) / 2;
            r[17] = (9 * c2 * c2 * c2 - 45 * c2 * c2 + 60 * c2 - 16) / 2;
            r[18] = (3 * x * x - c2) * (3 * c2 - 5); r[19] = (y * y - z * z) * (3 * c2 - 5);
            r[20] = x * y * (3 * c2 - 8); r[21] = y * z * (3 * c2 - 8); r[22] = x * z * (3 * c2 - 8);
            r[23] = x * (y * y - z * z); r[24] = y * (z * z - x * x); r[25] = z * (x * x - y * y);
            r[26] = x * y * z;
        } else {  // 19
            r[0] = 1; r[1] = 19 * c2 - 30; r[2] = (21 * c2 * c2 - 53 * c2 + 24) / 2;
            r[3] = x; r[4] = (5 * c2 - 9) * x; r[5] = y; r[6] = (5 * c2 - 9) * y;
            r[7] = z; r[8] = (5 * c2 - 9) * z;
            r[9] = 3 * x * x - c2; r[10] = (3 * c2 - 5) * (3 * x * x - c2);
            r[11] = y * y - z * z; r[12] = (3 * c2 - 5) * (y * y - z * z);
            r[13] = x * y; r[14] = y * z; r[15] = x * z;
            r[16] = x * (y * y - z * z); r[17] = y * (z * z - x * x); r[18] = z * (x * x - y * y);
        }
        for (int i = 0; i < Q; ++i) M[i * Q + k] = r[i];
    }
    if (Q == 9) stress = {7, 8};
    else if (Q == 27) stress = {5, 6, 7, 8, 9};
    else stress = {9, 11, 13, 14, 15};
}

void mrt_inverse(int Q, const std::vector<double>& M, std::vector<double>& Minv) {
    Minv.assign(Q * Q, 0.0);
    for (int i = 0; i < Q; ++i) {
        double n2 = 0;
        for (int k = 0; k < Q; ++k) n2 += M[i * Q + k] * M[i * Q + k];
        for (int k = 0; k < Q; ++k) Minv[k * Q + i] = M[i * Q + k] / n2;
    }
}

// Lagrange_interpolation (adv.cu:40-48): same product order.
void lagrange(double* coeff, const double* x, double xa, int n) {
    for (int i = 0; i < n; ++i) {
        coeff[i] = 1.0;
        for (int j = 0; j < i; ++j) coeff[i] *= (xa - x[j]) / (x[i] - x[j]);
        for (int j = i + 1; j < n; ++j) coeff[i] *= (xa - x[j]) / (x[i] - x[j]);
    }
}

// ------------------------------------------------------------------------------------------
// Reference direction tables for the 3D two-axis (x,y) corner scheme: set_cpm, set_cxypm
// (basic.cu:122-153).  Entries the reference leaves unset are 0 (= the cell itself).
// ------------------------------------------------------------------------------------------
struct RefTables {
    int cp[27] = {0}, cm[27] = {0}, cip[27] = {0}, cim[27] = {0};
    int cxp[27] = {0}, cxm[27] = {0}, cyp[27] = {0}, cym[27] = {0};
};

RefTables ref_tables_3d() {
    RefTables t;
    const int cp[27] = {0, 1, 2, 3, 4, 0, 0, 7, 8, 9, 10, 1, 1, 3, 3, 2, 2, 4, 4, 7, 7, 8, 8, 9, 9, 10, 10};
    const int cm[27] = {0, 3, 4, 1, 2, 0, 0, 10, 9, 8, 7, 3, 3, 1, 1, 4, 4, 2, 2, 10, 10, 9, 9, 8, 8, 7, 7};
    for (int k = 0; k < 27; ++k) { t.cp[k] = cp[k]; t.cm[k] = cm[k]; }
    const int cip[8] = {8, 8, 7, 7, 10, 10, 9, 9}, cim[8] = {9, 9, 10, 10, 7, 7, 8, 8};
    for (int k = 0; k < 8; ++k) { t.cip[19 + k] = cip[k]; t.cim[19 + k] = cim[k]; }
    const int cxp[27] = {0, 1, 0, 3, 0, 0, 0, 1, 1, 3, 3, 1, 1, 3, 3, 0, 0, 0, 0, 1, 1, 1, 1, 3, 3, 3, 3};
    const int cxm[27] = {0, 3, 0, 1, 0, 0, 0, 3, 3, 1, 1, 3, 3, 1, 1, 0, 0, 0, 0, 3, 3, 3, 3, 1, 1, 1, 1};
    const int cyp[27] = {0, 0, 2, 0, 4, 0, 0, 2, 4, 2, 4, 0, 0, 0, 0, 2, 2, 4, 4, 2, 2, 4, 4, 2, 2, 4, 4};
    const int cym[27] = {0, 0, 4, 0, 2, 0, 0, 4, 2, 4, 2, 0, 0, 0, 0, 4, 4, 2, 2, 4, 4, 2, 2, 4, 4, 2, 2};
    for (int k = 0; k < 27; ++k) { t.cxp[k] = cxp[k]; t.cxm[k] = cxm[k]; t.cyp[k] = cyp[k]; t.cym[k] = cym[k]; }
    return t;
}

// Population set of one transported field with the reference's buffer aliasing inside
// IP_process (adv.cu:246-256): fin lives in 'feq', finy in 'f', post-collision in 'ftmp'.
struct PopSet {
    int Q = 0;
    Lattice L;
    std::vector<double> f, ftmp, fout, feq, Fk;
    std::vector<int> nextB, nextK;  // G*Q, set_bound2D/3D
};

enum FluxMode { FLUX_NONE = 0, FLUX_ONE_AXIS = 1, FLUX_CORNER = 2 };

struct Oracle {
    ipl_config cfg{};
    Lattice L;        // flow lattice
    int Q = 0, dim = 0;
    int N = 0, G = 0, Ntot = 0;
    double dz = 1, dt = 1, c = 1;
    FluxMode flux_mode = FLUX_NONE;
    int flux_axis = 0;      // 0,1,2 for FLUX_ONE_AXIS
    int npts = 5;
    double period[3] = {0, 0, 0};
    RefTables rt;

    std::vector<int> neib;                       // Ntot*Q
    std::vector<double> pos[3], del[3];          // Ntot
    std::vector<double> p, u, v, w, rho, phi, S, tau, taus, Fx, Fy, Fz;
    std::vector<double> phiold, uold, vold, wold;
    std::vector<double> M, Minv, MM, Sr;
    std::vector<int> stress;

    PopSet fl, sc;   // flow and scalar populations
    bool has_scalar = false;
    std::string err;

    // grid index of calc cells (for the immersed boundary's node blocks)
    int n[3] = {0, 0, 0};
    std::vector<int> lnum;        // (x*ny + y)*nz + z -> calc id or -1
    std::vector<int> cidx[3];     // calc id -> coordinate index per axis

    // immersed boundary (ibm.cu)
    struct IB {
        bool active = false;
        ipl_ib_config cfg{};
        int M = 0;
        std::vector<double> oposw, onB, posw, nB, velw, Gw;  // 3M
        std::vector<int> lattice_id;                          // M
        double posB[3], velB[3], angleVB[3], quat[4], quatold[4], quaS[9], FB[3], Torque[3];
    } ib;

    // open boundary: cells reset to feq(p, u, v, w) at the end of every step
    std::vector<int> ob_cells;
    std::vector<double> ob_val;   // (p, u, v, w) per cell
};

inline int axis_dir(const Lattice& L, int axis, int sgn) {
    for (int k = 1; k < L.Q; ++k) {
        const int c[3] = {L.cx[k], L.cy[k], L.cz[k]};
        if (c[axis] == sgn && c[(axis + 1) % 3] == 0 && c[(axis + 2) % 3] == 0) return k;
    }
    return 0;
}
inline int comp(const Lattice& L, int k, int axis) {
    return axis == 0 ? L.cx[k] : (axis == 1 ? L.cy[k] : L.cz[k]);
}

// ------------------------------------------------------------------------------------------
// set_bound2D (set_bound.cu:25-58) and set_bound3D (set_bound.cu:60-147): slip remap tables
// per ghost and direction.  Restated with the same tables and the same tests.
// ------------------------------------------------------------------------------------------
void set_bound2D(Oracle& o, PopSet& ps) {
    const int Q = ps.Q, nb = o.N, QN = o.Q;
    int nei[2][2][9] = {}, kinv[2][2][9] = {};
    const Lattice& L = ps.L;
    for (int k = 0; k < 9; ++k) {
        nei[std::abs(L.cx[k])][std::abs(L.cz[k])][k] = 0;
        kinv[std::abs(L.cx[k])][std::abs(L.cz[k])][k] = L.kinv[k];
        kinv[0][0][k] = L.kinv[k];
    }
    nei[1][0][5] = 3; nei[0][1][5] = 4; nei[1][0][6] = 1; nei[0][1][6] = 4;
    nei[1][0][7] = 1; nei[0][1][7] = 2; nei[1][0][8] = 3; nei[0][1][8] = 2;
    kinv[1][0][5] = 8; kinv[0][1][5] = 6; kinv[1][0][6] = 7; kinv[0][1][6] = 5;
    kinv[1][0][7] = 6; kinv[0][1][7] = 8; kinv[1][0][8] = 5; kinv[0][1][8] = 7;
    ps.nextB.assign((size_t)o.G * Q, 0);
    ps.nextK.assign((size_t)o.G * Q, 0);
    for (int i = 0; i < o.G; ++i) {
        const size_t cid = (size_t)o.neib[(size_t)(nb + i) * QN] * QN;
        for (int k = 1; k < 9; ++k) {
            int cxx = L.cx[k], czz = L.cz[k];
            if (o.neib[cid + L.kinv[k]] >= nb) {
                if ((o.neib[cid + 1] >= nb && cxx == -1) || (o.neib[cid + 3] >= nb && cxx == 1)) cxx = 0;
                if ((o.neib[cid + 2] >= nb && czz == -1) || (o.neib[cid + 4] >= nb && czz == 1)) czz = 0;
            }
            ps.nextB[(size_t)i * Q + k] = nei[std::abs(cxx)][std::abs(czz)][k];
            ps.nextK[(size_t)i * Q + k] = kinv[std::abs(cxx)][std::abs(czz)][k];
        }
    }
}

void set_bound3D(Oracle& o, PopSet& ps) {
    const int Q = ps.Q, nb = o.N, QN = o.Q;
    const Lattice& L = ps.L;
    static int nei[2][2][2][27], kinv[2][2][2][27];
    std::memset(nei, 0, sizeof(nei));
    std::memset(kinv, 0, sizeof(kinv));
    const Lattice L27 = make_lattice(27);
    for (int k = 0; k < 27; ++k) {
        nei[std::abs(L27.cx[k])][std::abs(L27.cy[k])][std::abs(L27.cz[k])][k] = 0;
        kinv[std::abs(L27.cx[k])][std::abs(L27.cy[k])][std::abs(L27.cz[k])][k] = L27.kinv[k];
        kinv[0][0][0][k] = L27.kinv[k];
    }
    // set_bound.cu:130-143
    nei[1][0][0][7] = 3; nei[0][1][0][7] = 4; nei[1][0][0][8] = 3; nei[0][1][0][8] = 2;
    nei[1][0][0][9] = 1; nei[0][1][0][9] = 4; nei[1][0][0][10] = 1; nei[0][1][0][10] = 2;
    nei[1][0][0][11] = 3; nei[0][0][1][11] = 6; nei[1][0][0][12] = 3; nei[0][0][1][12] = 5;
    nei[1][0][0][13] = 1; nei[0][0][1][13] = 6; nei[1][0][0][14] = 1; nei[0][0][1][14] = 5;
    nei[0][1][0][15] = 4; nei[0][0][1][15] = 6; nei[0][1][0][16] = 4; nei[0][0][1][16] = 5;
    nei[0][1][0][17] = 2; nei[0][0][1][17] = 6; nei[0][1][0][18] = 2; nei[0][0][1][18] = 5;
    {
        const int t[8][6] = {{3, 4, 6, 10, 14, 18}, {3, 4, 5, 10, 13, 17}, {3, 2, 6, 9, 14, 16}, {3, 2, 5, 9, 13, 15},
                             {1, 4, 6, 8, 12, 18},  {1, 4, 5, 8, 11, 17},  {1, 2, 6, 7, 12, 16}, {1, 2, 5, 7, 11, 15}};
        for (int q = 0; q < 8; ++q) {
            const int k = 19 + q;
            nei[1][0][0][k] = t[q][0]; nei[0][1][0][k] = t[q][1]; nei[0][0][1][k] = t[q][2];
            nei[1][1][0][k] = t[q][3]; nei[1][0][1][k] = t[q][4]; nei[0][1][1][k] = t[q][5];
        }
    }
    // set_bound.cu:145-158
    kinv[1][0][0][7] = 8; kinv[0][1][0][7] = 9; kinv[1][0][0][8] = 7; kinv[0][1][0][8] = 10;
    kinv[1][0][0][9] = 10; kinv[0][1][0][9] = 7; kinv[1][0][0][10] = 9; kinv[0][1][0][10] = 8;
    kinv[1][0][0][11] = 12; kinv[0][0][1][11] = 13; kinv[1][0][0][12] = 11; kinv[0][0][1][12] = 14;
    kinv[1][0][0][13] = 14; kinv[0][0][1][13] = 11; kinv[1][0][0][14] = 13; kinv[0][0][1][14] = 12;
    kinv[0][1][0][15] = 16; kinv[0][0][1][15] = 17; kinv[0][1][0][16] = 15; kinv[0][0][1][16] = 18;
    kinv[0][1][0][17] = 18; kinv[0][0][1][17] = 15; kinv[0][1][0][18] = 17; kinv[0][0][1][18] = 16;
    {
        const int t[8][6] = {{22, 24, 25, 20, 21, 23}, {21, 23, 26, 19, 22, 24}, {20, 26, 23, 22, 19, 25},
                             {19, 25, 24, 21, 20, 26}, {26, 20, 21, 24, 25, 19}, {25, 19, 22, 23, 26, 20},
                             {24, 22, 19, 26, 23, 21}, {23, 21, 20, 25, 24, 22}};
        for (int q = 0; q < 8; ++q) {
            const int k = 19 + q;
            kinv[1][0][0][k] = t[q][0]; kinv[0][1][0][k] = t[q][1]; kinv[0][0][1][k] = t[q][2];
            kinv[1][1][0][k] = t[q][3]; kinv[1][0][1][k] = t[q][4]; kinv[0][1][1][k] = t[q][5];
        }
    }
    ps.nextB.assign((size_t)o.G * Q, 0);
    ps.nextK.assign((size_t)o.G * Q, 0);
    for (int i = 0; i < o.G; ++i) {
        const size_t cid = (size_t)o.neib[(size_t)(nb + i) * QN] * QN;
        auto nbv = [&](int d) { return d < QN ? o.neib[cid + d] : -1; };
        for (int k = 1; k < Q; ++k) {
            const int cxk = L.cx[k], cyk = L.cy[k], czk = L.cz[k];
            int cxx = cxk, cyy = cyk, czz = czk;
            if (o.neib[cid + L.kinv[k]] >= nb) {
                if ((nbv(1) >= nb && cxx == -1) || (nbv(3) >= nb && cxx == 1)) cxx = 0;
                if ((nbv(2) >= nb && cyy == -1) || (nbv(4) >= nb && cyy == 1)) cyy = 0;
                if ((nbv(5) >= nb && czz == -1) || (nbv(6) >= nb && czz == 1)) czz = 0;
                if (QN > 18) {  // edge tests read the xz / yz neighbours (set_bound.cu:477-501)
                    if ((nbv(11) < nb && cxk == -1 && czk == -1) || (nbv(14) < nb && cxk == 1 && czk == 1)) { cxx = cxk; czz = czk; }
                    if ((nbv(12) < nb && cxk == -1 && czk == 1) || (nbv(13) < nb && cxk == 1 && czk == -1)) { cxx = cxk; czz = czk; }
                    if ((nbv(15) < nb && cyk == -1 && czk == -1) || (nbv(18) < nb && cyk == 1 && czk == 1)) { cyy = cyk; czz = czk; }
                    if ((nbv(16) < nb && cyk == -1 && czk == 1) || (nbv(17) < nb && cyk == 1 && czk == -1)) { cyy = cyk; czz = czk; }
                    if ((nbv(11) >= nb && cxx == -1 && czz == -1) || (nbv(14) >= nb && cxx == 1 && czz == 1)) { cxx = 0; czz = 0; }
                    if ((nbv(12) >= nb && cxx == -1 && czz == 1) || (nbv(13) >= nb && cxx == 1 && czz == -1)) { cxx = 0; czz = 0; }
                    if ((nbv(15) >= nb && cyy == -1 && czz == -1) || (nbv(18) >= nb && cyy == 1 && czz == 1)) { cyy = 0; czz = 0; }
                    if ((nbv(16) >= nb && cyy == -1 && czz == 1) || (nbv(17) >= nb && cyy == 1 && czz == -1)) { cyy = 0; czz = 0; }
                }
            }
            ps.nextB[(size_t)i * Q + k] = nei[std::abs(cxx)][std::abs(cyy)][std::abs(czz)][k];
            ps.nextK[(size_t)i * Q + k] = kinv[std::abs(cxx)][std::abs(cyy)][std::abs(czz)][k];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Grid and neighbour build (LockExchange.cu:57-127) with set_neibghor_wall (basic.cu:203-237).
// ------------------------------------------------------------------------------------------
bool build(Oracle& o, const ipl_config& cfg, const ipl_init* init) {
    o.cfg = cfg;
    o.Q = cfg.lattice;
    if (o.Q != 9 && o.Q != 19 && o.Q != 27) { o.err = "lattice must be 9, 19 or 27"; return false; }
    o.L = make_lattice(o.Q);
    o.dim = o.L.dim;
    const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
    if (nx < 1 || ny < 1 || nz < 1 || (o.dim == 2 && ny != 1)) { o.err = "bad grid size"; return false; }
    o.dz = cfg.dz; o.dt = cfg.dt; o.c = cfg.dz / cfg.dt;
    const int n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        if (!cfg.pos[a] || !cfg.width[a]) { o.err = "missing axis geometry"; return false; }
        if (cfg.periodic[a]) {
            double L = 0;
            for (int i = 0; i < n[a]; ++i) L += cfg.width[a][i];
            o.period[a] = L;
        }
    }
    // flux mode
    int fa = cfg.flux_axes;
    if (fa == IPL_FLUX_REFERENCE) fa = (o.dim == 2) ? IPL_FLUX_X : (IPL_FLUX_X | IPL_FLUX_Y);
    if (fa == 0) o.flux_mode = FLUX_NONE;
    else if (fa == IPL_FLUX_X || fa == IPL_FLUX_Y || fa == IPL_FLUX_Z) {
        o.flux_mode = FLUX_ONE_AXIS;
        o.flux_axis = fa == IPL_FLUX_X ? 0 : (fa == IPL_FLUX_Y ? 1 : 2);
        if (o.dim == 2 && o.flux_axis == 1) { o.err = "no y flux axis in 2D"; return false; }
    } else if (fa == (IPL_FLUX_X | IPL_FLUX_Y) && o.dim == 3) {
        o.flux_mode = FLUX_CORNER;
        o.rt = ref_tables_3d();
    } else { o.err = "unsupported flux axes"; return false; }
    o.npts = cfg.stencil == 3 ? 3 : 5;
    // the reference's corner scheme is written for the 5 x 5 stencil only (adv.cu:100-145)
    if (o.flux_mode == FLUX_CORNER && o.npts != 5) { o.err = "the two-axis corner scheme needs the 5-point stencil"; return false; }
    // width >= dz on flux axes (the slab point must lie inside the cell, adv.cu:77,85-86)
    for (int a = 0; a < 3; ++a) {
        const bool isflux = (o.flux_mode == FLUX_ONE_AXIS && o.flux_axis == a) ||
                            (o.flux_mode == FLUX_CORNER && a < 2);
        if (!isflux) continue;
        for (int i = 0; i < n[a]; ++i)
            if (cfg.width[a][i] < o.dz * (1 - 1e-12)) { o.err = "cell width below lattice spacing on a flux axis"; return false; }
    }

    // lnum
    const size_t ncell = (size_t)nx * ny * nz;
    std::vector<int> lnum(ncell, -1);
    int N = 0;
    for (size_t q = 0; q < ncell; ++q)
        if (!(cfg.solid && cfg.solid[q])) lnum[q] = N++;
    o.N = N;
    const int Q = o.Q;
    std::vector<int> ci(N), cj(N), cl(N);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int l = 0; l < nz; ++l) {
                const int id = lnum[((size_t)i * ny + j) * nz + l];
                if (id >= 0) { ci[id] = i; cj[id] = j; cl[id] = l; }
            }
    o.n[0] = nx; o.n[1] = ny; o.n[2] = nz;
    o.lnum = lnum;
    o.cidx[0] = ci; o.cidx[1] = cj; o.cidx[2] = cl;
    // neib (LockExchange.cu:117-127), with periodic wrap as the builder extension
    o.neib.assign((size_t)N * Q, -1);
    for (int id = 0; id < N; ++id)
        for (int k = 0; k < Q; ++k) {
            int t[3] = {ci[id] + o.L.cx[k], cj[id] + o.L.cy[k], cl[id] + o.L.cz[k]};
            bool ok = true;
            for (int a = 0; a < 3; ++a) {
                if (t[a] < 0 || t[a] >= n[a]) {
                    if (cfg.periodic[a]) t[a] = (t[a] + n[a]) % n[a];
                    else ok = false;
                }
            }
            o.neib[(size_t)id * Q + k] = ok ? lnum[((size_t)t[0] * ny + t[1]) * nz + t[2]] : -1;
        }
    // geometry and fields of calc cells
    for (int a = 0; a < 3; ++a) { o.pos[a].resize(N); o.del[a].resize(N); }
    for (int id = 0; id < N; ++id) {
        const int ix[3] = {ci[id], cj[id], cl[id]};
        for (int a = 0; a < 3; ++a) { o.pos[a][id] = cfg.pos[a][ix[a]]; o.del[a][id] = cfg.width[a][ix[a]]; }
    }
    auto take = [&](const double* src, double dflt) {
        std::vector<double> v(N);
        for (int id = 0; id < N; ++id) v[id] = src ? src[id] : dflt;
        return v;
    };
    o.p = take(init ? init->p : nullptr, 0.0);
    o.u = take(init ? init->u : nullptr, 0.0);
    o.v = take(init ? init->v : nullptr, 0.0);
    o.w = take(init ? init->w : nullptr, 0.0);
    o.S = take(init ? init->S : nullptr, 0.0);
    o.phi = take(init ? init->phi : nullptr, 0.0);
    o.tau = take(init ? init->tau : nullptr, cfg.tau);
    o.taus = take(init ? init->taus : nullptr, cfg.taus);
    o.rho.resize(N);
    for (int id = 0; id < N; ++id)  // LockExchange.cu:93
        o.rho[id] = cfg.rho_L + o.phi[id] * (cfg.rho_H - cfg.rho_L) + o.S[id] * cfg.beta_s;
    o.Fx.assign(N, 0); o.Fy.assign(N, 0); o.Fz.assign(N, 0);
    o.phiold = o.phi; o.uold = o.u; o.vold = o.v; o.wold = o.w;

    // set_neibghor_wall (basic.cu:203-237): one ghost per calc cell with a missing link
    int G = 0;
    for (int id = 0; id < N; ++id) {
        bool missing = false;
        for (int k = 1; k < Q; ++k) if (o.neib[(size_t)id * Q + k] == -1) { missing = true; break; }
        if (!missing) continue;
        const int g = N + G;
        o.neib.push_back(id);
        for (int k = 1; k < Q; ++k) o.neib.push_back(g);
        o.p.push_back(o.p[id]); o.rho.push_back(o.rho[id]); o.phi.push_back(o.phi[id]);
        for (int a = 0; a < 3; ++a) { o.pos[a].push_back(o.pos[a][id]); o.del[a].push_back(o.del[a][id]); }
        o.u.push_back(0); o.v.push_back(0); o.w.push_back(0);
        for (int k = 1; k < Q; ++k)
            if (o.neib[(size_t)id * Q + k] == -1) {
                o.neib[(size_t)id * Q + k] = g;
                o.neib[(size_t)g * Q + o.L.kinv[k]] = id;
            }
        ++G;
    }
    o.G = G;
    o.Ntot = N + G;

    // MRT matrices, set_M (basic.cu:171-201)
    mrt_matrix(o.L, o.M, o.stress);
    mrt_inverse(Q, o.M, o.Minv);
    o.Sr.assign(Q, 0);
    for (int i = 0; i < Q; ++i) o.Sr[i] = cfg.s_rates[i];
    for (int s : o.stress) o.Sr[s] = 0;
    o.MM.assign(Q * Q, 0);
    for (int i = 0; i < Q; ++i)
        for (int j = 0; j < Q; ++j) {
            double acc = 0;
            for (int k = 0; k < Q; ++k) acc += o.Minv[i * Q + k] * (o.Sr[k] * o.M[k * Q + j]);
            o.MM[i * Q + j] = acc;
        }

    // populations: f = feq(p,u) (LockExchange.cu:133-139)
    auto setup_pops = [&](PopSet& ps, int Qp) {
        ps.Q = Qp;
        ps.L = make_lattice(Qp);
        const size_t sz = (size_t)o.Ntot * Qp;
        ps.f.assign(sz, 0); ps.ftmp.assign(sz, 0); ps.fout.assign(sz, 0); ps.feq.assign(sz, 0); ps.Fk.assign(sz, 0);
        if (o.dim == 2) set_bound2D(o, ps); else set_bound3D(o, ps);
    };
    setup_pops(o.fl, Q);
    const double c2 = o.c * o.c;
    for (int id = 0; id < N; ++id)
        for (int k = 0; k < Q; ++k) {
            const double t = (o.L.cx[k] * o.c * o.u[id] + o.L.cy[k] * o.c * o.v[id] + o.L.cz[k] * o.c * o.w[id]) / c2;
            o.fl.f[(size_t)id * Q + k] = o.L.w[k] * (o.p[id] + 3.0 * t + 4.5 * t * t -
                1.5 * (o.u[id] * o.u[id] + o.v[id] * o.v[id] + o.w[id] * o.w[id]) / c2);
        }
    o.has_scalar = cfg.scalar != IPL_SCALAR_NONE;
    if (o.has_scalar) {
        int Qs = cfg.scalar_lattice ? cfg.scalar_lattice : Q;
        if (!(Qs == Q || (Qs == 7 && o.dim == 3))) { o.err = "scalar lattice must equal the flow lattice or be 7 in 3D"; return false; }
        if (cfg.scalar == IPL_SCALAR_PHASE_FIELD && Qs != Q) { o.err = "phase field uses the flow lattice"; return false; }
        if (o.flux_mode == FLUX_CORNER && Qs == 7) { o.err = "D3Q7 scalar needs a single flux axis"; return false; }
        setup_pops(o.sc, Qs);
        const Lattice& Ls = o.sc.L;
        const bool linear = Qs == 7;
        const double cs2inv = linear ? 4.0 : 3.0;
        for (int id = 0; id < N; ++id) {
            const double s = cfg.scalar == IPL_SCALAR_SALINITY ? o.S[id] : o.phi[id];
            for (int k = 0; k < Qs; ++k) {
                const double t = (Ls.cx[k] * o.c * o.u[id] + Ls.cy[k] * o.c * o.v[id] + Ls.cz[k] * o.c * o.w[id]) / c2;
                double e = 1.0 + cs2inv * t;
                if (!linear) e += 4.5 * t * t - 1.5 * (o.u[id] * o.u[id] + o.v[id] * o.v[id] + o.w[id] * o.w[id]) / c2;
                o.sc.f[(size_t)id * Qs + k] = Ls.w[k] * s * e;
            }
        }
    }
    return true;
}

// ------------------------------------------------------------------------------------------
// Kernels of collision.cu, restated per cell.
// ------------------------------------------------------------------------------------------

// equ_f (collision.cu:3-14): feq of the flow
void equ_f(Oracle& o) {
    const int Q = o.Q; const double c2 = o.c * o.c;
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < o.N; ++id)
        for (int k = 0; k < Q; ++k) {
            const double t = (o.L.cx[k] * o.c * o.u[id] + o.L.cy[k] * o.c * o.v[id] + o.L.cz[k] * o.c * o.w[id]) / c2;
            o.fl.feq[(size_t)id * Q + k] = o.L.w[k] * (o.p[id] + 3.0 * t + 4.5 * t * t -
                1.5 * (o.u[id] * o.u[id] + o.v[id] * o.v[id] + o.w[id] * o.w[id]) / c2);
        }
}

// position difference a - b along axis; across a periodic seam the period is added
// (builder extension: the reference has no periodic axes)
inline double dpos(const Oracle& o, int axis, int a, int b) {
    double d = o.pos[axis][a] - o.pos[axis][b];
    if (o.cfg.periodic[axis] && d <= 0) d += o.period[axis];
    return d;
}

// Force (collision.cu:57-112): surface tension + buoyancy + pressure-gradient term into
// Fx,Fy,Fz (accumulating), then Fk = w dt (c.F) / (rho c^2 / 3).
void force(Oracle& o) {
    const int Q = o.Q; const double c2 = o.c * o.c;
    const int up = o.dim == 3 ? 5 : 2, down = o.dim == 3 ? 6 : 4;
    const double eps = o.dz / 1000;
    const double beta = 12.0 * o.cfg.sigma / o.cfg.pf_thickness, kap = 3.0 * o.cfg.sigma * o.cfg.pf_thickness / 2.0;
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < o.N; ++id) {
        const int* nb = &o.neib[(size_t)id * Q];
        const double* ph = o.phi.data(); const double* rh = o.rho.data();
        double nx_ = (ph[nb[1]] - ph[nb[3]]) / (dpos(o, 0, nb[1], nb[3]) + eps);
        double ny_ = 0;
        if (o.dim == 3) ny_ = (ph[nb[2]] - ph[nb[4]]) / (dpos(o, 1, nb[2], nb[4]) + eps);
        double nz_ = (ph[nb[up]] - ph[nb[down]]) / (dpos(o, 2, nb[up], nb[down]) + eps);
        // x term over delX^2, the 2/4 term over dz^2 as written in the reference
        // (collision.cu:70-71), also where 2/4 are the y directions of a 3D lattice
        const double lap = (ph[nb[1]] - 2 * ph[id] + ph[nb[3]]) / (o.del[0][id] * o.del[0][id]) +
                           (ph[nb[2]] - 2 * ph[id] + ph[nb[4]]) / (o.dz * o.dz);
        double st = 4 * beta * ph[id] * (ph[id] - 1.0) * (ph[id] - 0.5) - kap * lap;
        if (kMutant == 2) st = 4 * beta * ph[id] * (ph[id] - 1.0) * (ph[id] - 0.5) + kap * lap;  // mutant 2: sign of kappa
        if (kMutant == 12)  // mutant 12: the x term of the Laplacian over dz^2
            st = 4 * beta * ph[id] * (ph[id] - 1.0) * (ph[id] - 0.5) -
                 kap * ((ph[nb[1]] - 2 * ph[id] + ph[nb[3]]) / (o.dz * o.dz) + (ph[nb[2]] - 2 * ph[id] + ph[nb[4]]) / (o.dz * o.dz));
        o.Fx[id] += st * nx_; o.Fy[id] += st * ny_; o.Fz[id] += st * nz_;
        if (o.cfg.boussinesq) o.Fz[id] += -o.S[id] * o.cfg.beta_s * o.cfg.gravity;
        else o.Fz[id] += -o.rho[id] * o.cfg.gravity;
        nx_ = (rh[nb[1]] - rh[nb[3]]) / (dpos(o, 0, nb[1], nb[3]) + eps);
        ny_ = 0;
        if (o.dim == 3) ny_ = (rh[nb[2]] - rh[nb[4]]) / (dpos(o, 1, nb[2], nb[4]) + eps);
        nz_ = (rh[nb[up]] - rh[nb[down]]) / (dpos(o, 2, nb[up], nb[down]) + eps);
        const double psg = kMutant == 1 ? 1.0 : -1.0;  // mutant 1: sign of the pressure-gradient term
        o.Fx[id] += psg * o.p[id] * c2 / 3.0 * nx_;
        o.Fy[id] += psg * o.p[id] * c2 / 3.0 * ny_;
        o.Fz[id] += psg * o.p[id] * c2 / 3.0 * nz_;
        const double den = kMutant == 3 ? o.rho[id] * c2 : o.rho[id] * c2 / 3.0;  // mutant 3: 1/3 dropped
        for (int k = 0; k < Q; ++k)
            o.fl.Fk[(size_t)id * Q + k] = o.L.w[k] * o.dt *
                (o.L.cx[k] * o.c * o.Fx[id] + o.L.cy[k] * o.c * o.Fy[id] + o.L.cz[k] * o.c * o.Fz[id]) / den;
    }
}

// col_f_MRT (collision.cu:115-150) / col_f_SRT (collision.cu:151-165)
void collide(Oracle& o) {
    const int Q = o.Q;
    PopSet& P = o.fl;
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < o.N; ++id) {
        double* f = &P.f[(size_t)id * Q]; double* feq = &P.feq[(size_t)id * Q];
        double* ft = &P.ftmp[(size_t)id * Q]; const double* Fk = &P.Fk[(size_t)id * Q];
        if (o.cfg.collision == IPL_COLL_SRT) {
            for (int k = 0; k < Q; ++k) {
                f[k] = f[k] - (f[k] - feq[k] + Fk[k] / 2.0) / o.tau[id] + Fk[k];
                ft[k] = f[k];
            }
        } else {
            for (int k = 0; k < Q; ++k) feq[k] = (f[k] - feq[k] + Fk[k] / 2.0);
            double mles[27];
            for (int i = 0; i < Q; ++i) {
                double tmp = 0;
                for (int k = 0; k < Q; ++k) {
                    mles[k] = 0;
                    for (int s : o.stress) mles[k] += o.Minv[i * Q + s] * o.M[s * Q + k] / o.tau[id];
                }
                for (int k = 0; k < Q; ++k) tmp += (o.MM[i * Q + k] + mles[k]) * feq[k];
                f[i] = f[i] - tmp + Fk[i];
                ft[i] = f[i];
            }
        }
        for (int k = 0; k < Q; ++k) { feq[k] = 0; f[k] = 0; }
    }
}

// equ_g (collision.cu:16-27): w_k S (1 + 3t + 4.5t^2 - 1.5u^2/c^2), t = (c_k . u)/c^2, into the
// scalar populations of every calc cell (no driver calls it; the reference writes its feq
// argument).  S is the salinity, or phi with a phase field; linear with 1/cs^2 = 4 on D3Q7.
void equ_g(Oracle& o) {
    PopSet& P = o.sc; const int Q = P.Q; const Lattice& L = P.L;
    const double c2 = o.c * o.c;
    const bool linear = Q == 7;
    const double cs2inv = linear ? 4.0 : 3.0;
    const std::vector<double>& S = o.cfg.scalar == IPL_SCALAR_SALINITY ? o.S : o.phi;
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < o.N; ++id) {
        const double u2 = o.u[id] * o.u[id] + o.v[id] * o.v[id] + o.w[id] * o.w[id];
        for (int k = 0; k < Q; ++k) {
            const double t = (L.cx[k] * o.c * o.u[id] + L.cy[k] * o.c * o.v[id] + L.cz[k] * o.c * o.w[id]) / c2;
            double e = 1.0 + cs2inv * t;
            if (!linear) e += 4.5 * t * t - 1.5 * u2 / c2;
            P.f[(size_t)id * Q + k] = L.w[k] * S[id] * e;
        }
    }
}

// col_g_reg (collision.cu:167-188): regularized scalar collision.  On D3Q7 the equilibrium
// is linear and 1/cs^2 = 4 (builder extension); otherwise the reference's form.
void col_g_reg(Oracle& o) {
    PopSet& P = o.sc; const int Q = P.Q; const Lattice& L = P.L;
    const double c2 = o.c * o.c;
    const bool linear = Q == 7;
    const double cs2inv = linear && kMutant != 7 ? 4.0 : 3.0;  // mutant 7: 1/cs^2 = 3 on D3Q7
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < o.N; ++id) {
        double* g = &P.f[(size_t)id * Q]; double* geq = &P.feq[(size_t)id * Q]; double* gt = &P.ftmp[(size_t)id * Q];
        double gkx = 0, gky = 0, gkz = 0;
        const double u2 = o.u[id] * o.u[id] + o.v[id] * o.v[id] + o.w[id] * o.w[id];
        for (int k = 0; k < Q; ++k) {
            double t = (L.cx[k] * o.c * o.u[id] + L.cy[k] * o.c * o.v[id] + L.cz[k] * o.c * o.w[id]) / c2;
            if (kMutant == 6) t = 0;  // mutant 6: the equilibrium without the velocity
            double e = 1.0 + cs2inv * t;
            if (!linear) e += 4.5 * t * t - 1.5 * u2 / c2;
            geq[k] = L.w[k] * o.S[id] * e;
            gt[k] = g[k] - geq[k];
            gkx += gt[k] * L.cx[k] * o.c; gky += gt[k] * L.cy[k] * o.c; gkz += gt[k] * L.cz[k] * o.c;
        }
        for (int k = 0; k < Q; ++k) {
            gt[k] = L.w[k] * cs2inv * (gkx * L.cx[k] * o.c + gky * L.cy[k] * o.c + gkz * L.cz[k] * o.c) / c2;
            g[k] = geq[k] + (1.0 - 1.0 / o.taus[id]) * gt[k];
            gt[k] = g[k]; geq[k] = 0; g[k] = 0;
        }
    }
}

// col_PF (collision.cu:190-223): conservative Allen-Cahn collision; the reference hard-codes
// tau_phi = 0.58 and the source weight tau_phi - 0.5 = 0.08, here cfg.pf_tau.
void col_PF(Oracle& o) {
    PopSet& P = o.sc; const int Q = P.Q; const Lattice& L = P.L;
    const double c2 = o.c * o.c;
    const int up = o.dim == 3 ? 5 : 2, down = o.dim == 3 ? 6 : 4;
    const double eps = o.dz / 1000;
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < o.N; ++id) {
        const int* nb = &o.neib[(size_t)id * o.Q];
        const double* ph = o.phi.data();
        const double nx_ = (ph[nb[1]] - ph[nb[3]]) / (dpos(o, 0, nb[1], nb[3]) + eps);
        double ny_ = 0;
        if (o.dim == 3) ny_ = (ph[nb[2]] - ph[nb[4]]) / (dpos(o, 1, nb[2], nb[4]) + eps);
        const double nz_ = (ph[nb[up]] - ph[nb[down]]) / (dpos(o, 2, nb[up], nb[down]) + eps);
        const double nabu = (ph[id] * o.u[id] - o.phiold[id] * o.uold[id]) / o.dt;
        const double nabv = (ph[id] * o.v[id] - o.phiold[id] * o.vold[id]) / o.dt;
        const double nabw = (ph[id] * o.w[id] - o.phiold[id] * o.wold[id]) / o.dt;
        o.phiold[id] = ph[id]; o.uold[id] = o.u[id]; o.vold[id] = o.v[id]; o.wold[id] = o.w[id];
        double* g = &P.f[(size_t)id * Q]; double* geq = &P.feq[(size_t)id * Q]; double* gt = &P.ftmp[(size_t)id * Q];
        const double u2 = o.u[id] * o.u[id] + o.v[id] * o.v[id] + o.w[id] * o.w[id];
        const double nrm = std::sqrt(nx_ * nx_ + ny_ * ny_ + nz_ * nz_ + 1e-8);
        for (int k = 0; k < Q; ++k) {
            const double cxk = L.cx[k] * o.c, cyk = L.cy[k] * o.c, czk = L.cz[k] * o.c;
            const double t = (cxk * o.u[id] + cyk * o.v[id] + czk * o.w[id]) / c2;
            geq[k] = L.w[k] * ph[id] * (1.0 + 3.0 * t + 4.5 * t * t - 1.5 * u2 / c2);
            const double ssg = kMutant == 5 ? -1.0 : 1.0;  // mutant 5: sign of the sharpening term
            const double tsg = kMutant == 4 ? 0.0 : 1.0;   // mutant 4: d(phi u)/dt dropped
            const double Fkp = o.dt * L.w[k] * (
                ssg * (1.0 - 4.0 * (ph[id] - 0.5) * (ph[id] - 0.5)) / o.cfg.pf_thickness * (cxk * nx_ + cyk * ny_ + czk * nz_) / nrm
                + tsg * 3 * (cxk * nabu + cyk * nabv + czk * nabw) / c2);
            const double tp = o.cfg.pf_tau;  // 0.58 in the reference
            g[k] = g[k] - (g[k] - geq[k] - Fkp / 1.0 * (tp - 0.5)) / tp + Fkp * 0;
            gt[k] = g[k]; geq[k] = 0; g[k] = 0;
        }
    }
}

// update_scalar (collision.cu:225-236)
void update_scalar(Oracle& o, std::vector<double>& scalar) {
    PopSet& P = o.sc; const int Q = P.Q;
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < o.N; ++id) {
        double s = 0;
        for (int k = 0; k < Q; ++k) s += P.f[(size_t)id * Q + k];
        scalar[id] = s;
    }
}

// update_rho (collision.cu:238-257): macroscopic readout
void update_rho(Oracle& o) {
    const int Q = o.Q;
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < o.N; ++id) {
        double p = 0, ux = 0, uy = 0, uz = 0;
        for (int k = 0; k < Q; ++k) {
            const double f = o.fl.f[(size_t)id * Q + k];
            p += f; ux += o.L.cx[k] * o.c * f; uy += o.L.cy[k] * o.c * f; uz += o.L.cz[k] * o.c * f;
        }
        o.p[id] = p;
        o.rho[id] = o.cfg.rho_L + (o.cfg.rho_H - o.cfg.rho_L) * o.phi[id] + o.S[id] * o.cfg.beta_s;
        o.u[id] = ux + o.Fx[id] * o.dt / (2.0 * o.rho[id]);
        o.v[id] = uy + o.Fy[id] * o.dt / (2.0 * o.rho[id]);
        o.w[id] = uz + o.Fz[id] * o.dt / (2.0 * o.rho[id]);
    }
}

// set_wall_rho (basic.cu:254-261)
void set_wall_rho(Oracle& o, std::vector<double>& a) {
    for (int g = o.N; g < o.Ntot; ++g) a[g] = a[o.neib[(size_t)g * o.Q]];
}

// ------------------------------------------------------------------------------------------
// IP_process (adv.cu:245-257): the interpolated propagation, ten launches in the same order.
// ------------------------------------------------------------------------------------------

// set_posl_x / set_posl_y (adv.cu:15-28) for a stencil ordered along sign s, plus the
// periodic unwrap of the builder extension.
void fix_positions(const Oracle& o, int axis, double* x, int n, int s) {
    if (o.cfg.periodic[axis]) {
        const double Lp = o.period[axis];
        const int c = n / 2;
        for (int j = c + 1; j < n; ++j) while (s * (x[j] - x[j - 1]) <= 0) x[j] += s * Lp;
        for (int j = c - 1; j >= 0; --j) while (s * (x[j + 1] - x[j]) <= 0) x[j] -= s * Lp;
    }
    const double d = o.dz;
    if (n == 5) {
        if (s == 1) {
            x[1] = std::min(x[1], x[2] - d); x[0] = std::min(x[0], x[1] - d);
            x[3] = std::max(x[3], x[2] + d); x[4] = std::max(x[4], x[3] + d);
        } else {
            x[1] = std::max(x[1], x[2] + d); x[0] = std::max(x[0], x[1] + d);
            x[3] = std::min(x[3], x[2] - d); x[4] = std::min(x[4], x[3] - d);
        }
    } else {
        if (s == 1) { x[0] = std::min(x[0], x[1] - d); x[2] = std::max(x[2], x[1] + d); }
        else { x[0] = std::max(x[0], x[1] + d); x[2] = std::min(x[2], x[1] - d); }
    }
}

// 1D slab flux of cell id along axis for direction k: the x/y loops of set_out (adv.cu:76-99)
void flux_1d(Oracle& o, PopSet& P, int id, int k, int axis) {
    const int Q = P.Q, QN = o.Q;
    const int s = comp(P.L, k, axis);
    const int dp = axis_dir(o.L, axis, s), dm = axis_dir(o.L, axis, -s);
    const int cp = o.neib[(size_t)id * QN + dp], cm = o.neib[(size_t)id * QN + dm];
    const int cpp = o.neib[(size_t)cp * QN + dp], cmm = o.neib[(size_t)cm * QN + dm];
    const std::vector<double>& pa = o.pos[axis];
    const std::vector<double>& da = o.del[axis];
    const double xa = pa[id] + s * (da[id] - o.dz) / 2.0;
    double coeff[5], x[5];
    double F;
    if (o.npts == 5) {
        x[0] = pa[cmm]; x[1] = pa[cm]; x[2] = pa[id]; x[3] = pa[cp]; x[4] = pa[cpp];
        fix_positions(o, axis, x, 5, s);
        lagrange(coeff, x, xa, 5);
        F = coeff[0] * P.ftmp[(size_t)cmm * Q + k] + coeff[1] * P.ftmp[(size_t)cm * Q + k] +
            coeff[2] * P.ftmp[(size_t)id * Q + k] + coeff[3] * P.ftmp[(size_t)cp * Q + k] +
            coeff[4] * P.ftmp[(size_t)cpp * Q + k];
    } else {
        x[0] = pa[cm]; x[1] = pa[id]; x[2] = pa[cp];
        fix_positions(o, axis, x, 3, s);
        lagrange(coeff, x, xa, 3);
        F = coeff[0] * P.ftmp[(size_t)cm * Q + k] + coeff[1] * P.ftmp[(size_t)id * Q + k] +
            coeff[2] * P.ftmp[(size_t)cp * Q + k];
    }
    P.feq[(size_t)id * Q + k] = F * (o.dz / da[cp]);
    P.fout[(size_t)id * Q + k] = F * (o.dz / da[id]);
}

// Corner directions of the 3D x-y scheme: the num3 loop of set_out (adv.cu:100-145)
void flux_corner(Oracle& o, PopSet& P, int id, int kk) {
    const int Q = P.Q, QN = o.Q;
    const RefTables& t = o.rt;
    auto nb = [&](int cell, int d) { return o.neib[(size_t)cell * QN + d]; };
    int idf[25];
    idf[12] = id;
    idf[13] = nb(id, t.cxp[kk]); idf[14] = nb(idf[13], t.cxp[kk]);
    idf[11] = nb(id, t.cxm[kk]); idf[10] = nb(idf[11], t.cxm[kk]);
    idf[17] = nb(id, t.cyp[kk]); idf[22] = nb(idf[17], t.cyp[kk]);
    idf[7] = nb(id, t.cym[kk]); idf[2] = nb(idf[7], t.cym[kk]);
    idf[18] = nb(id, t.cp[kk]); idf[24] = nb(idf[18], t.cp[kk]);
    idf[6] = nb(id, t.cm[kk]); idf[0] = nb(idf[6], t.cm[kk]);
    idf[8] = nb(id, t.cip[kk]); idf[4] = nb(idf[8], t.cip[kk]);
    idf[16] = nb(id, t.cim[kk]); idf[20] = nb(idf[16], t.cim[kk]);
    idf[19] = nb(idf[14], t.cyp[kk]); idf[9] = nb(idf[14], t.cym[kk]);
    idf[15] = nb(idf[10], t.cyp[kk]); idf[5] = nb(idf[10], t.cym[kk]);
    idf[23] = nb(idf[22], t.cxp[kk]); idf[21] = nb(idf[22], t.cxm[kk]);
    idf[3] = nb(idf[2], t.cxp[kk]); idf[1] = nb(idf[2], t.cxm[kk]);
    const int sx = P.L.cx[kk], sy = P.L.cy[kk];
    double cxw[5], cyw[5], x[5];
    double xa = o.pos[0][id] + sx * (o.del[0][id] - o.dz) / 2.0;
    x[0] = o.pos[0][idf[10]]; x[1] = o.pos[0][idf[11]]; x[2] = o.pos[0][id]; x[3] = o.pos[0][idf[13]]; x[4] = o.pos[0][idf[14]];
    fix_positions(o, 0, x, 5, sx);
    lagrange(cxw, x, xa, 5);
    double fx = 0;
    for (int i = 0; i < 5; ++i) fx += cxw[i] * P.ftmp[(size_t)idf[i + 10] * Q + kk];
    xa = o.pos[1][id] + sy * (o.del[1][id] - o.dz) / 2.0;
    x[0] = o.pos[1][idf[2]]; x[1] = o.pos[1][idf[7]]; x[2] = o.pos[1][id]; x[3] = o.pos[1][idf[17]]; x[4] = o.pos[1][idf[22]];
    fix_positions(o, 1, x, 5, sy);
    lagrange(cyw, x, xa, 5);
    double fy = 0;
    for (int i = 0; i < 5; ++i) fy += cyw[i] * P.ftmp[(size_t)idf[2 + i * 5] * Q + kk];
    double fxy = 0;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) fxy += cxw[i] * cyw[j] * P.ftmp[(size_t)idf[i + j * 5] * Q + kk];
    const double dz = o.dz;
    const std::vector<double>& dX = o.del[0]; const std::vector<double>& dY = o.del[1];
    P.fout[(size_t)id * Q + kk] = fx * dz / dX[id] + fy * dz / dY[id] - (kMutant == 8 ? 0.0 : 1.0) * fxy * (dz * dz) / (dX[id] * dY[id]);
    const double a0 = fxy * (dz * dz) / (dX[idf[13]] * dY[idf[17]]);
    // mutant 9: the x width of the -y neighbour's flux taken at the +x neighbour
    const double a1 = (fy - fxy * dz / dX[kMutant == 9 ? idf[13] : idf[17]]) * dz / dY[idf[17]];
    const double a2 = (fx - fxy * dz / dY[idf[13]]) * dz / dX[idf[13]];
    double* fin = P.feq.data(); double* finy = P.f.data();
    #pragma omp atomic
    fin[(size_t)id * Q + kk] += a0;
    #pragma omp atomic
    fin[(size_t)nb(id, t.cxm[kk]) * Q + kk] += a1;
    #pragma omp atomic
    finy[(size_t)nb(id, t.cym[kk]) * Q + kk] += a2;
}

void ip_process(Oracle& o, PopSet& P) {
    const int Q = P.Q, QN = o.Q, N = o.N, Ntot = o.Ntot;
    const Lattice& L = P.L;
    // set_wall_f (adv.cu:3-13): ghost <- owner
    for (int g = N; g < Ntot; ++g) {
        const int ow = o.neib[(size_t)g * QN];
        for (int k = 0; k < Q; ++k) P.ftmp[(size_t)g * Q + k] = P.ftmp[(size_t)ow * Q + k];
    }
    // set_out (adv.cu:50-147)
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < N; ++id) {
        for (int k = 0; k < Q; ++k) {
            int nflux = 0, ax = -1;
            if (o.flux_mode == FLUX_ONE_AXIS) { if (comp(L, k, o.flux_axis) != 0) { nflux = 1; ax = o.flux_axis; } }
            else if (o.flux_mode == FLUX_CORNER) {
                const bool hx = L.cx[k] != 0, hy = L.cy[k] != 0;
                nflux = (int)hx + (int)hy; ax = hx ? 0 : 1;
            }
            if (nflux == 0) {
                P.fout[(size_t)id * Q + k] = P.ftmp[(size_t)id * Q + k];
                P.feq[(size_t)id * Q + k] = P.ftmp[(size_t)id * Q + k];
            } else if (nflux == 1) {
                flux_1d(o, P, id, k, ax);
            }
        }
        if (o.flux_mode == FLUX_CORNER)
            for (int k = 7; k < Q; ++k)
                if (L.cx[k] != 0 && L.cy[k] != 0) flux_corner(o, P, id, k);
    }
    // set_fin (adv.cu:148-175): 3D corner scheme only; ghosts pass their corner fluxes on
    if (o.flux_mode == FLUX_CORNER) {
        const int nk[8] = {9, 8, 10, 7, 7, 10, 8, 9};
        const int nkz[8] = {19, 20, 21, 22, 23, 24, 25, 26};
        const int nneib[8] = {1, 2, 1, 4, 3, 2, 3, 4};
        const int k3[12] = {7, 8, 9, 10, 19, 21, 23, 25, 20, 22, 24, 26};
        double* fin = P.feq.data(); double* finy = P.f.data();
        for (int g = N; g < Ntot; ++g)
            for (int k = 0; k < 4; ++k) {
                const int t0 = o.neib[(size_t)g * QN + nneib[k * 2 + 0]];
                const int t1 = o.neib[(size_t)g * QN + nneib[k * 2 + 1]];
                const double f0 = (double)(1 - t0 / g), f1 = (double)(1 - t1 / g);
                const int a = nk[2 * (k3[k] - 7) + 0], b = nk[2 * (k3[k] - 7) + 1];
                fin[(size_t)t0 * Q + a] += fin[(size_t)g * Q + k3[k]] * f0;
                finy[(size_t)t1 * Q + b] += finy[(size_t)g * Q + k3[k]] * f1;
                if (Q == 27) {
                    fin[(size_t)t0 * Q + nkz[2 * (a - 7) + 0]] += fin[(size_t)g * Q + k3[k + 4]] * f0;
                    fin[(size_t)t0 * Q + nkz[2 * (a - 7) + 1]] += fin[(size_t)g * Q + k3[k + 8]] * f0;
                    finy[(size_t)t1 * Q + nkz[2 * (b - 7) + 0]] += finy[(size_t)g * Q + k3[k + 4]] * f1;
                    finy[(size_t)t1 * Q + nkz[2 * (b - 7) + 1]] += finy[(size_t)g * Q + k3[k + 8]] * f1;
                }
            }
    }
    // reset_f (adv.cu:176-187)
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < N; ++id)
        for (int k = 0; k < Q; ++k) {
            P.feq[(size_t)id * Q + k] += P.f[(size_t)id * Q + k];
            P.f[(size_t)id * Q + k] = P.ftmp[(size_t)id * Q + k];
        }
    // set_wall_f_inv on f and on fout (adv.cu:188-232): ghost <- mirrored owner
    auto wall_inv = [&](std::vector<double>& a) {
        for (int g = N; g < Ntot; ++g) {
            const int ow = o.neib[(size_t)g * QN];
            if (o.flux_mode == FLUX_CORNER) {
                // xy mirror of axis / xz / yz directions, then corners through nextB/nextK
                int kxy[19] = {0, 3, 4, 1, 2, 0, 0, 0, 0, 0, 0, 13, 14, 11, 12, 17, 18, 15, 16};
                const int kn1[15] = {0, 1, 2, 3, 4, 5, 6, 11, 12, 13, 14, 15, 16, 17, 18};
                for (int q = 0; q < 15; ++q) a[(size_t)g * Q + kn1[q]] = a[(size_t)ow * Q + kxy[kn1[q]]];
                const int k3[12] = {7, 8, 9, 10, 19, 21, 23, 25, 20, 22, 24, 26};
                const int nkz[8] = {20, 19, 22, 21, 24, 23, 26, 25};
                const size_t wid = (size_t)(g - N) * Q;
                for (int q = 0; q < 4; ++q) {
                    const int nk = L.kinv[P.nextK[wid + L.kinv[k3[q]]]], neB = P.nextB[wid + L.kinv[k3[q]]];
                    const int src = o.neib[(size_t)ow * QN + neB];
                    a[(size_t)g * Q + k3[q]] = a[(size_t)src * Q + nk];
                    if (Q == 27) {
                        a[(size_t)g * Q + k3[q + 4]] = a[(size_t)src * Q + nkz[2 * (nk - 7) + 1]];
                        a[(size_t)g * Q + k3[q + 8]] = a[(size_t)src * Q + nkz[2 * (nk - 7) + 0]];
                    }
                }
            } else {
                // mirror along the flux axis: the 2D x table of adv.cu:196-199 for x
                for (int k = 1; k < Q; ++k) {
                    int mk = 0;
                    if (o.flux_mode == FLUX_ONE_AXIS) {
                        int c[3] = {L.cx[k], L.cy[k], L.cz[k]};
                        if (c[o.flux_axis] != 0) {
                            c[o.flux_axis] = -c[o.flux_axis];
                            for (int j = 0; j < Q; ++j)
                                if (L.cx[j] == c[0] && L.cy[j] == c[1] && L.cz[j] == c[2]) mk = j;
                        }
                    }
                    a[(size_t)g * Q + k] = a[(size_t)ow * Q + mk];
                }
            }
        }
    };
    wall_inv(P.f);
    wall_inv(P.fout);
    // set_tmp (adv.cu:234-244): ftmp[i,k] = f[n,k] + (fin[i,k] - fout[n,k]), n = neib[i, cp[k]]
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < N; ++id)
        for (int k = 0; k < Q; ++k) {
            int d = 0;
            if (o.flux_mode == FLUX_ONE_AXIS) {
                const int s = comp(L, k, o.flux_axis);
                if (s) d = axis_dir(o.L, o.flux_axis, s);
            } else if (o.flux_mode == FLUX_CORNER) d = o.rt.cp[k];
            const int n = o.neib[(size_t)id * QN + d];
            P.ftmp[(size_t)id * Q + k] = P.f[(size_t)n * Q + k] + (P.feq[(size_t)id * Q + k] - P.fout[(size_t)n * Q + k]);
        }
    // set_wall_propagation (set_bound.cu:149-168)
    for (int g = N; g < Ntot; ++g) {
        const int ow = o.neib[(size_t)g * QN];
        const size_t wid = (size_t)(g - N) * Q;
        for (int k = 1; k < Q; ++k) {
            if (o.cfg.wall_bc == IPL_BC_BOUNCE_BACK)
                P.ftmp[(size_t)g * Q + k] = P.ftmp[(size_t)ow * Q + L.kinv[k]];
            else {
                const int src = o.neib[(size_t)ow * QN + P.nextB[wid + k]];
                P.ftmp[(size_t)g * Q + k] = P.ftmp[(size_t)src * Q + P.nextK[wid + k]];
            }
        }
    }
    // propagation (propagation.cu:2-12)
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < N; ++id)
        for (int k = 0; k < Q; ++k)
            P.f[(size_t)id * Q + k] = P.ftmp[(size_t)o.neib[(size_t)id * QN + L.kinv[k]] * Q + k];
    // reset_wall_fin (basic.cu:274-284)
    for (int g = N; g < Ntot; ++g)
        for (int k = 0; k < Q; ++k) { P.feq[(size_t)g * Q + k] = 0; P.f[(size_t)g * Q + k] = 0; }
}

void reset_force(Oracle& o) {
    std::fill(o.Fx.begin(), o.Fx.end(), 0.0);
    std::fill(o.Fy.begin(), o.Fy.end(), 0.0);
    std::fill(o.Fz.begin(), o.Fz.end(), 0.0);
}

// ------------------------------------------------------------------------------------------
// Immersed boundary (function/ibm.cu) and the drivers' marker set-up (SphereFall.cu:197-261).
// One rigid body; markers are processed in index order, so every sum below has a fixed order
// (the reference spreads with atomicAdd, whose order is unspecified).
// ------------------------------------------------------------------------------------------

// rotation matrix (row-major) of the orientation quaternion, set_quaternionS (ibm.cu:16-28)
void ib_rotation(const double q[4], double R[9]) {
    const double a = q[0], b = q[1], c = q[2], d = q[3];
    R[0] = a * a - c * c - d * d + b * b; R[1] = 2 * (b * c - a * d);         R[2] = 2 * (b * d + a * c);
    R[3] = 2 * (b * c + a * d);         R[4] = a * a - d * d - b * b + c * c; R[5] = 2 * (c * d - a * b);
    R[6] = 2 * (b * d - a * c);         R[7] = 2 * (c * d + a * b);         R[8] = a * a - b * b - c * c + d * d;
}

// 4-point Peskin function of r = |x - x_node| / dz, weightFunction (ibm.cu:61-66)
double ib_peskin4(double r) {
    if (r < 1.0) return (3.0 - 2.0 * r + std::sqrt(1.0 + 4.0 * r - 4.0 * r * r)) / 8.0;
    if (r <= 2.0) return (5.0 - 2.0 * r - std::sqrt(-7.0 + 12.0 * r - 4.0 * r * r)) / 8.0;
    return 0.0;
}

// Calc id of the cell displaced by d from calc cell id, or -1 where the reference's neighbour is
// a ghost (outside a wall, or a masked cell).  pos receives the node coordinates: a ghost sits on
// its owner (basic.cu:221-225); across a periodic seam the coordinate is unwrapped.
int ib_node(const Oracle& o, int id, const int d[3], double pos[3]) {
    int c[3];
    double shift[3] = {0, 0, 0};
    bool ghost = false;
    for (int a = 0; a < 3; ++a) {
        c[a] = o.cidx[a][id] + d[a];
        if (c[a] < 0 || c[a] >= o.n[a]) {
            if (o.cfg.periodic[a]) {
                shift[a] = c[a] < 0 ? -o.period[a] : o.period[a];
                c[a] = c[a] < 0 ? c[a] + o.n[a] : c[a] - o.n[a];
            } else {
                ghost = true;
            }
        }
    }
    const int nid = ghost ? -1 : o.lnum[((size_t)c[0] * o.n[1] + c[1]) * o.n[2] + c[2]];
    for (int a = 0; a < 3; ++a) pos[a] = nid < 0 ? o.pos[a][id] : o.pos[a][nid] + shift[a];
    return nid;
}

// Displacement of node (i, j, l) of the 3x3(x3) block, the neighbour directions of idfIndex
// (ibm.cu:132-141).  In 2D the entry of (x-1, z-1) is direction 5 of D2Q9, i.e. (x+1, z+1);
// that node is kept as the reference has it.
void ib_block_offset(int dim, int i, int j, int l, int d[3]) {
    if (dim == 2) {
        d[0] = i - 1; d[1] = 0; d[2] = l - 1;
        if (i == 0 && l == 0) { d[0] = 1; d[2] = 1; }
    } else {
        d[0] = i - 1; d[1] = j - 1; d[2] = l - 1;
    }
}

struct IBWeights {
    int lo, hi;                 // node offsets lo..hi on each axis (y: 0..0 in 2D)
    double w[3][5];             // weights per axis, indexed by offset - lo
};

// Node weights of marker m on the block around its nearest cell.
//  stress  (get_IBMGw2, ibm.cu:146-189): x (and z in 2D) by the Peskin function on the three nodes,
//          masked at ghosts and, in 2D, renormalised; y, z (3D) by 3-point Lagrange.
//  direct  (IB_directForcing, ibm.cu:292-313): 3-point Lagrange on every axis.
//  peskin4 (direct, builder extension): the Peskin function on offsets -2..2 (its full support).
IBWeights ib_weights(const Oracle& o, int m, bool stress) {
    const auto& B = o.ib;
    const int id = B.lattice_id[m];
    const double* xw = &B.posw[3 * m];
    IBWeights W{};
    const bool pk4 = !stress && B.cfg.kernel == IPL_IB_KERNEL_PESKIN4;
    W.lo = pk4 ? -2 : -1; W.hi = pk4 ? 2 : 1;
    const int npt = W.hi - W.lo + 1;
    for (int a = 0; a < 3; ++a) {
        if (o.dim == 2 && a == 1) { for (int t = 0; t < npt; ++t) W.w[a][t] = (t == -W.lo) ? 1.0 : 0.0; continue; }
        double xs[5];
        int ids[5];
        for (int t = 0; t < npt; ++t) {
            int d[3] = {0, 0, 0};
            d[a] = W.lo + t;
            double pp[3];
            ids[t] = ib_node(o, id, d, pp);
            xs[t] = pp[a];
        }
        const bool peskin = pk4 || (stress && (a == 0 || (o.dim == 2 && a == 2)));
        if (peskin) {
            for (int t = 0; t < npt; ++t) W.w[a][t] = ib_peskin4(std::fabs(xw[a] - xs[t]) / o.dz) * (ids[t] >= 0 ? 1.0 : 0.0);
            if (stress && o.dim == 2) {
                double s = 0;
                for (int t = 0; t < npt; ++t) s += W.w[a][t];
                for (int t = 0; t < npt; ++t) W.w[a][t] /= s;
            }
        } else {
            lagrange(W.w[a], xs, xw[a], 3);
        }
    }
    return W;
}

// Calls fn(node id, weight) for every node of marker m's block in the reference's loop order
// (x outer, y, z inner); ghost nodes are skipped (the factor (id < num_calc) of ibm.cu:194).
template <typename Fn>
void ib_for_nodes(const Oracle& o, int m, const IBWeights& W, bool block3, Fn&& fn) {
    const int id = o.ib.lattice_id[m];
    const int jlo = o.dim == 2 ? 0 : W.lo, jhi = o.dim == 2 ? 0 : W.hi;
    for (int i = W.lo; i <= W.hi; ++i)
        for (int j = jlo; j <= jhi; ++j)
            for (int l = W.lo; l <= W.hi; ++l) {
                int d[3];
                if (block3) ib_block_offset(o.dim, i + 1, j + 1, l + 1, d);
                else { d[0] = i; d[1] = j; d[2] = l; }
                double pp[3];
                const int nid = ib_node(o, id, d, pp);
                const double wgt = W.w[0][i - W.lo] * W.w[1][j - W.lo] * W.w[2][l - W.lo];
                if (nid >= 0) fn(nid, wgt);
            }
}

// get_IBMGw2 (ibm.cu:124-274): f interpolated at the marker, split into the outer (fw1) and inner
// (fw2) populations by the marker normal, the wall stress difference spread to F, and the marker
// force Gw from the outer populations.
void ib_stress(Oracle& o) {
    auto& B = o.ib;
    const int Q = o.Q;
    const double c = o.c, c2 = c * c;
    const double alpha = 22.5 / 180.0 * 3.14159;  // the reference's literal
    const int dimension = o.dim;
    const double spread = std::pow(B.cfg.spacing / o.dz, dimension - 1) / o.dz;
    for (int m = 0; m < B.M; ++m) {
        const IBWeights W = ib_weights(o, m, true);
        std::vector<double> fw(Q, 0.0), fw1(Q), fw2(Q);
        ib_for_nodes(o, m, W, true, [&](int nid, double wgt) {
            for (int k = 0; k < Q; ++k) fw[k] += wgt * o.fl.f[(size_t)nid * Q + k];
        });
        const double U[3] = {B.velw[3 * m], B.velw[3 * m + 1], B.velw[3 * m + 2]};
        const double nv[3] = {B.nB[3 * m], B.nB[3 * m + 1], B.nB[3 * m + 2]};
        // the reference's lattice velocities are c e (items[IDX_cx], Items::setc, basic.cu:79-96)
        for (int k = 0; k < Q; ++k) {
            const double e[3] = {c * o.L.cx[k], c * o.L.cy[k], c * o.L.cz[k]};
            const double cn = (nv[0] * e[0] + nv[1] * e[1] + nv[2] * e[2]) / c;
            const double lim = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) / c * std::sin(alpha);
            const double bounce = fw[o.L.kinv[k]] + 6.0 * o.L.w[k] * (e[0] * U[0] + e[1] * U[1] + e[2] * U[2]) / c2;
            const double beta = std::fabs(cn) / (lim + 1e-10);
            const double mixed = (1.0 - beta) * fw[k] + beta * bounce;
            fw1[k] = cn <= 0 ? fw[k] : (cn <= lim ? mixed : bounce);
            fw2[k] = cn >= 0 ? fw[k] : (cn >= -lim ? mixed : bounce);
        }
        const double tau = o.tau[B.lattice_id[m]];
        // traction sigma . n of the populations g (ibm.cu:207-227): sigma_ab = -sum_k [ (1/(2 tau)) g_k delta_ab
        //   + (1 - 1/(2 tau)) 3 (g_k (c e_a - U_a)(c e_b - U_b) - (g_k - w_k) U_a U_b) / c^2 ]
        auto traction = [&](const std::vector<double>& g1, const std::vector<double>* g2, double out[3]) {
            for (int a = 0; a < 3; ++a) {
                double sg[3] = {0, 0, 0};
                for (int k = 0; k < Q; ++k) {
                    const double gk = g2 ? g1[k] - (*g2)[k] : g1[k];
                    const double e[3] = {c * o.L.cx[k], c * o.L.cy[k], c * o.L.cz[k]};
                    for (int b = 0; b < 3; ++b)
                        sg[b] -= 0.5 / tau * gk * (a == b ? 1.0 : 0.0) +
                                 (1.0 - 0.5 / tau) * 3 * (gk * (e[a] - U[a]) * (e[b] - U[b]) - (gk - o.L.w[k]) * U[a] * U[b]) / c2;
                }
                out[a] = sg[0] * nv[0] + sg[1] * nv[1] + sg[2] * nv[2];
            }
        };
        double g[3];
        traction(fw1, &fw2, g);
        for (int a = 0; a < 3; ++a) g[a] = -g[a];
        ib_for_nodes(o, m, W, true, [&](int nid, double wgt) {
            o.Fx[nid] += g[0] * spread * wgt;
            o.Fy[nid] += g[1] * spread * wgt;
            o.Fz[nid] += g[2] * spread * wgt;
        });
        double gw[3];
        traction(fw1, nullptr, gw);
        const double scale = B.cfg.rho_f * c2 / 3.0 * std::pow(B.cfg.spacing, dimension - 1);
        for (int a = 0; a < 3; ++a) B.Gw[3 * m + a] = gw[a] * scale;
    }
}

// IB_directForcing (ibm.cu:276-344): node velocity interpolated at the marker, force
// (U_w - u)/dt spread to F.  With iterations > 1 (multi-direct forcing, builder extension) the
// node velocities are corrected by the spread force, du = c^2 dt dF / 3 (what update_velIBM does
// to them, sum_k w_k (c e_k)(c e_k) = c^2 / 3), and the pass is repeated.
void ib_direct(Oracle& o) {
    auto& B = o.ib;
    const double c2 = o.c * o.c;
    const int dimension = o.dim;
    const double spread = std::pow(B.cfg.spacing / o.dz, dimension - 1) * 3.0 / c2;
    const double wscale = B.cfg.rho_f * std::pow(B.cfg.spacing, dimension - 1) * o.dz;
    const bool block3 = B.cfg.kernel != IPL_IB_KERNEL_PESKIN4;
    std::vector<double> vel[3] = {o.u, o.v, o.w};
    std::vector<IBWeights> W(B.M);
    for (int m = 0; m < B.M; ++m) W[m] = ib_weights(o, m, false);
    const int iters = std::max(1, B.cfg.iterations);
    std::vector<double> dF[3];
    for (int it = 0; it < iters; ++it) {
        for (int a = 0; a < 3; ++a) dF[a].assign(o.N, 0.0);
        std::vector<double> gw(3 * (size_t)B.M);
        for (int m = 0; m < B.M; ++m) {
            double um[3] = {0, 0, 0};
            ib_for_nodes(o, m, W[m], block3, [&](int nid, double wgt) {
                for (int a = 0; a < 3; ++a) um[a] += wgt * vel[a][nid];
            });
            for (int a = 0; a < 3; ++a) gw[3 * m + a] = (B.velw[3 * m + a] - um[a]) / o.dt;
        }
        for (int m = 0; m < B.M; ++m) {
            ib_for_nodes(o, m, W[m], block3, [&](int nid, double wgt) {
                for (int a = 0; a < 3; ++a) dF[a][nid] += gw[3 * m + a] * spread * wgt;
            });
            for (int a = 0; a < 3; ++a) B.Gw[3 * m + a] += -gw[3 * m + a] * wscale;
        }
        for (int id = 0; id < o.N; ++id) {
            o.Fx[id] += dF[0][id]; o.Fy[id] += dF[1][id]; o.Fz[id] += dF[2][id];
            for (int a = 0; a < 3; ++a) vel[a][id] += c2 * o.dt * dF[a][id] / 3.0;
        }
    }
}

// update_velIBM (ibm.cu:346-365): f += w dt (c e . F); p and u = sum c e f recomputed from f (no
// half-force term)
void ib_update_vel(Oracle& o) {
    const int Q = o.Q;
    const double c = o.c;
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < o.N; ++id) {
        double p = 0, ux = 0, uy = 0, uz = 0;
        for (int k = 0; k < Q; ++k) {
            double& f = o.fl.f[(size_t)id * Q + k];
            f = f + o.L.w[k] * o.dt * (c * o.L.cx[k] * o.Fx[id] + c * o.L.cy[k] * o.Fy[id] + c * o.L.cz[k] * o.Fz[id]);
            p += f; ux += c * o.L.cx[k] * f; uy += c * o.L.cy[k] * f; uz += c * o.L.cz[k] * f;
        }
        o.p[id] = p; o.u[id] = ux; o.v[id] = uy; o.w[id] = uz;
    }
}

// update_IBbody (ibm.cu:367-424): force and torque summed over the markers, buoyancy, and (for a
// moving body) second-order Adams-Bashforth steps of velocity, position, body-frame angular
// velocity and orientation quaternion.
void ib_body(Oracle& o) {
    auto& B = o.ib;
    const double dt = o.dt;
    double Rold[9], FBold[3], Told[3], velBold[3], angold[3];
    ib_rotation(B.quatold, Rold);
    for (int a = 0; a < 3; ++a) {
        FBold[a] = B.FB[a]; B.FB[a] = 0; Told[a] = B.Torque[a]; B.Torque[a] = 0;
        velBold[a] = B.velB[a]; angold[a] = B.angleVB[a];
    }
    for (int m = 0; m < B.M; ++m) {
        const double* G = &B.Gw[3 * m];
        const double r[3] = {B.posw[3 * m] - B.posB[0], B.posw[3 * m + 1] - B.posB[1], B.posw[3 * m + 2] - B.posB[2]};
        for (int a = 0; a < 3; ++a) B.FB[a] += G[a];
        B.Torque[0] += r[1] * G[2] - r[2] * G[1];
        B.Torque[1] += r[2] * G[0] - r[0] * G[2];
        B.Torque[2] += r[0] * G[1] - r[1] * G[0];
    }
    for (int a = 0; a < 3; ++a) B.FB[a] += (1.0 - B.cfg.rho_f / B.cfg.density) * B.cfg.mass * B.cfg.gravity[a];
    if (!B.cfg.moving) return;
    for (int a = 0; a < 3; ++a) {
        B.velB[a] += dt * (3.0 * B.FB[a] - FBold[a]) / B.cfg.mass / 2.0;
        B.posB[a] += dt * (3.0 * B.velB[a] - velBold[a]) / 2.0;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            B.angleVB[i] += dt * (3.0 * B.quaS[j * 3 + i] * B.Torque[j] - Rold[j * 3 + i] * Told[j]) / B.cfg.inertia[i] / 2.0;
    double q[4], w[3];
    for (int i = 0; i < 4; ++i) q[i] = B.quat[i];
    for (int i = 0; i < 3; ++i) w[i] = B.angleVB[i];
    // dq/dt = (1/2) Omega(w) q, Adams-Bashforth: q += dt (3 Omega(w) q - Omega(w_old) q_old) / 4
    B.quat[0] += dt * 3.0 * (-w[0] * q[1] - w[1] * q[2] - w[2] * q[3]) / 4.0;
    B.quat[1] += dt * 3.0 * (w[0] * q[0] + w[2] * q[2] - w[1] * q[3]) / 4.0;
    B.quat[2] += dt * 3.0 * (w[1] * q[0] - w[2] * q[1] + w[0] * q[3]) / 4.0;
    B.quat[3] += dt * 3.0 * (w[2] * q[0] + w[1] * q[1] - w[0] * q[2]) / 4.0;
    const double* qo = B.quatold;
    B.quat[0] -= dt * (-angold[0] * qo[1] - angold[1] * qo[2] - angold[2] * qo[3]) / 4.0;
    B.quat[1] -= dt * (angold[0] * qo[0] + angold[2] * qo[2] - angold[1] * qo[3]) / 4.0;
    B.quat[2] -= dt * (angold[1] * qo[0] - angold[2] * qo[1] + angold[0] * qo[3]) / 4.0;
    B.quat[3] -= dt * (angold[2] * qo[0] + angold[1] * qo[1] - angold[0] * qo[2]) / 4.0;
    for (int i = 0; i < 4; ++i) B.quatold[i] = q[i];
    ib_rotation(B.quat, B.quaS);
}

// update_IBpoint (ibm.cu:426-447) and search_IBlattice (ibm.cu:449-463)
void ib_points(Oracle& o) {
    auto& B = o.ib;
    const double* R = B.quaS;
    const double* w = B.angleVB;
    for (int m = 0; m < B.M; ++m) {
        const double* op = &B.oposw[3 * m];
        const double* on = &B.onB[3 * m];
        const double wxo[3] = {w[1] * op[2] - w[2] * op[1], w[2] * op[0] - w[0] * op[2], w[0] * op[1] - w[1] * op[0]};
        for (int i = 0; i < 3; ++i) {
            B.posw[3 * m + i] = B.posB[i];
            B.velw[3 * m + i] = B.velB[i];
            B.nB[3 * m + i] = 0;
            for (int j = 0; j < 3; ++j) {
                B.posw[3 * m + i] += R[i * 3 + j] * op[j];
                B.nB[3 * m + i] += R[i * 3 + j] * on[j];
            }
            B.velw[3 * m + i] += R[i * 3 + 0] * wxo[0] + R[i * 3 + 1] * wxo[1] + R[i * 3 + 2] * wxo[2];
        }
        // nearest node among the current one and its lattice neighbours (ghosts sit on their owner)
        const int id = B.lattice_id[m];
        auto d2 = [&](int c) {
            double s = 0;
            for (int a = 0; a < 3; ++a) s += (o.pos[a][c] - B.posw[3 * m + a]) * (o.pos[a][c] - B.posw[3 * m + a]);
            return s;
        };
        int near = id;
        double best = d2(id);
        for (int k = 1; k < o.Q; ++k) {
            const int nid = o.neib[(size_t)id * o.Q + k];
            const double dd = d2(nid);
            if (dd < best) { best = dd; near = nid; }
        }
        B.lattice_id[m] = near;
    }
}

// The IB sub-step after update_rho and resetF (SphereFall.cu:372-383), then resetF
// (SphereFall.cu:433).
// profile_s22 (SphereFall3D/SphereFall.cu:3-11): 1 inside, 0 outside, the driver's sine ramp
// between; the transition width is fixed at 0.1 whatever the argument
double spm_profile(double r, double Rada) {
    const double lim = 0.1;
    const double in_line = 1.0 - lim + std::pow(lim / 2, 2), out_line = 1.0 + lim + std::pow(lim / 2, 2);
    return 0 * (out_line < r) + (0.5 * (std::sin(3.14159 * r / lim / Rada) + 1.0)) * (in_line <= r && r <= out_line) +
           1.0 * (r < in_line);
}

// profile_s2 (ibm.cu:50-58), the profile of SPM_ellipse
double spm_profile_s2(double lim, double r) {
    const double in_line = 1.0 - lim + std::pow(lim / 2, 2), out_line = 1.0 + lim + std::pow(lim / 2, 2);
    return 0 * (out_line < r) + (0.5 * (std::sin(3.14159 * r / lim) + 1.0)) * (in_line <= r && r <= out_line) +
           1.0 * (r < in_line);
}

// profile_s (ibm.cu:44-48), the profile of SPM; dist = distance from the circle's centre
double spm_profile_s(double dz, double R, double dist) {
    const double r = R - 0.5 * dz - dist;
    return 0 * (r < -0.5 * dz) + (0.5 * (std::sin(3.14159 * r / dz) + 1.0)) * (std::fabs(r) <= 0.5 * dz) + 1.0 * (r > 0.5 * dz);
}

// The library's SPM_ellipse (ibm.cu:89-122, spm = 2) and SPM (ibm.cu:68-87, spm = 3): the second
// assignment of X1, Y1, Z1 (R d, untransposed) is the one SPM_ellipse uses, and on a 3D lattice
// its ellipse lies in the xy face; SPM's circle is in the x-y plane of its posx / posy arguments,
// here the lattice's first and last axes (x-z in 2D, where x-y would be a stripe), with the
// velocity of marker 0 and the force over dz.  f += w dt 3 (e c . F) / c^2 (SPM writes
// f = ftmp + ..., ftmp = f: the same once ftmp holds f).
void ib_spm_library(Oracle& o) {
    const auto& B = o.ib;
    const double a = B.cfg.spm_a, b = B.cfg.spm_b, c2 = o.c * o.c;
    const int Q = o.Q;
    for (int id = 0; id < o.N; ++id) {
        const double d[3] = {o.pos[0][id] - B.posB[0], o.pos[1][id] - B.posB[1], o.pos[2][id] - B.posB[2]};
        double pr, v[3];
        if (B.cfg.spm == 2) {
            double X1 = B.quaS[0] * d[0] + B.quaS[1] * d[1] + B.quaS[2] * d[2];
            double Y1 = B.quaS[3] * d[0] + B.quaS[4] * d[1] + B.quaS[5] * d[2];
            double Z1 = B.quaS[6] * d[0] + B.quaS[7] * d[1] + B.quaS[8] * d[2];
            if (kMutant == 13) {  // mutant 13: the transposed rotation (SPM_ellipse3D's) here
                X1 = B.quaS[0] * d[0] + B.quaS[3] * d[1] + B.quaS[6] * d[2];
                Y1 = B.quaS[1] * d[0] + B.quaS[4] * d[1] + B.quaS[7] * d[2];
                Z1 = B.quaS[2] * d[0] + B.quaS[5] * d[1] + B.quaS[8] * d[2];
            }
            if (o.dim == 3) Z1 = Y1;  // 3D (xy face)
            pr = spm_profile_s2(o.dz / a, std::pow(X1 / a, 2) + std::pow(Z1 / b, 2));
            for (int q = 0; q < 3; ++q) v[q] = B.velB[q];
        } else {
            const int t = o.dim == 3 ? 1 : 2;
            pr = spm_profile_s(o.dz, a, std::sqrt(std::pow(d[0], 2) + std::pow(d[t], 2)));
            for (int q = 0; q < 3; ++q) v[q] = kMutant == 14 ? B.velB[q] : B.velw[q];  // mutant 14: body velocity
        }
        if (pr == 0.0) continue;
        const double per = B.cfg.spm == 3 && kMutant != 15 ? o.dz : o.dt;  // mutant 15: SPM's force over dt
        const double fx = pr * (v[0] - o.u[id]) / per, fy = pr * (v[1] - o.v[id]) / per, fz = pr * (v[2] - o.w[id]) / per;
        for (int k = 0; k < Q; ++k)
            o.fl.f[(size_t)id * Q + k] += o.L.w[k] * o.dt * 3.0 * (o.L.cx[k] * o.c * fx + o.L.cy[k] * o.c * fy + o.L.cz[k] * o.c * fz) / c2;
    }
}

// SPM_ellipse3D (SphereFall3D/SphereFall.cu:13-30): forcing (v_B - u) / dt on the body interior,
// weighted by the profile of the body-frame ellipsoid distance, added to f
void ib_spm(Oracle& o) {
    const auto& B = o.ib;
    if (B.cfg.spm != 1) { ib_spm_library(o); return; }
    const double a = B.cfg.spm_a, b = B.cfg.spm_b, c2 = o.c * o.c;
    const int Q = o.Q;
    for (int id = 0; id < o.N; ++id) {
        const double d[3] = {o.pos[0][id] - B.posB[0], o.pos[1][id] - B.posB[1], o.pos[2][id] - B.posB[2]};
        double X1 = B.quaS[0] * d[0] + B.quaS[3] * d[1] + B.quaS[6] * d[2];
        double Y1 = B.quaS[1] * d[0] + B.quaS[4] * d[1] + B.quaS[7] * d[2];
        double Z1 = B.quaS[2] * d[0] + B.quaS[5] * d[1] + B.quaS[8] * d[2];
        if (kMutant == 10) {  // mutant 10: the rotation applied untransposed
            X1 = B.quaS[0] * d[0] + B.quaS[1] * d[1] + B.quaS[2] * d[2];
            Y1 = B.quaS[3] * d[0] + B.quaS[4] * d[1] + B.quaS[5] * d[2];
            Z1 = B.quaS[6] * d[0] + B.quaS[7] * d[1] + B.quaS[8] * d[2];
        }
        const double distance = std::pow(X1 / a, 2) + std::pow(Y1 / a, 2) + std::pow(Z1 / b, 2);
        const double pr = spm_profile(distance, a);
        if (pr == 0.0) continue;
        const double fx = pr * (B.velB[0] - o.u[id]) / o.dt, fy = pr * (B.velB[1] - o.v[id]) / o.dt,
                     fz = pr * (B.velB[2] - o.w[id]) / o.dt;
        for (int k = 0; k < Q; ++k)
            o.fl.f[(size_t)id * Q + k] += o.L.w[k] * o.dt * 3.0 * (o.L.cx[k] * o.c * fx + o.L.cy[k] * o.c * fy + o.L.cz[k] * o.c * fz) / c2;
    }
}

void ib_substep(Oracle& o) {
    auto& B = o.ib;
    reset_force(o);
    std::fill(B.Gw.begin(), B.Gw.end(), 0.0);
    if (B.cfg.spm) ib_spm(o);
    if (B.cfg.method == IPL_IB_STRESS) ib_stress(o);
    else ib_direct(o);
    ib_update_vel(o);
    ib_body(o);
    if (B.cfg.moving) ib_points(o);
    reset_force(o);
}

// Marker set-up (SphereFall.cu:203-258): marker positions and unit normals from the body state,
// nearest calc cell by exhaustive search (first minimum in calc order).
bool ib_attach(Oracle& o, const ipl_ib_config& cfg) {
    auto& B = o.ib;
    if (cfg.n_markers <= 0 || !cfg.marker_pos || !cfg.marker_normal) { o.err = "markers missing"; return false; }
    if (cfg.method == IPL_IB_STRESS && cfg.kernel != IPL_IB_KERNEL_REFERENCE) { o.err = "the stress method uses the reference weights"; return false; }
    B.cfg = cfg;
    B.M = cfg.n_markers;
    B.oposw.assign(cfg.marker_pos, cfg.marker_pos + 3 * B.M);
    B.onB.assign(cfg.marker_normal, cfg.marker_normal + 3 * B.M);
    B.posw.assign(3 * B.M, 0); B.nB.assign(3 * B.M, 0); B.velw.assign(3 * B.M, 0); B.Gw.assign(3 * B.M, 0);
    B.lattice_id.assign(B.M, 0);
    for (int a = 0; a < 3; ++a) {
        B.posB[a] = cfg.pos[a]; B.velB[a] = cfg.vel[a]; B.angleVB[a] = cfg.omega[a]; B.FB[a] = 0; B.Torque[a] = 0;
    }
    for (int i = 0; i < 4; ++i) { B.quat[i] = cfg.quat[i]; B.quatold[i] = cfg.quat[i]; }
    ib_rotation(B.quat, B.quaS);
    for (int m = 0; m < B.M; ++m) {
        const double* op = &B.oposw[3 * m];
        const double* on = &B.onB[3 * m];
        const double nn = std::sqrt(on[0] * on[0] + on[1] * on[1] + on[2] * on[2]);
        for (int i = 0; i < 3; ++i) {
            double p = 0, nv = 0;
            for (int j = 0; j < 3; ++j) { p += B.quaS[i * 3 + j] * op[j]; nv += B.quaS[i * 3 + j] * on[j]; }
            B.nB[3 * m + i] = nv / nn;
            B.posw[3 * m + i] = B.posB[i] + p;
        }
        // the drivers leave velw at zero until the first update_IBpoint
        double best = 1e300;
        int near = 0;
        for (int id = 0; id < o.N; ++id) {
            double s = 0;
            for (int a = 0; a < 3; ++a) s += (o.pos[a][id] - B.posw[3 * m + a]) * (o.pos[a][id] - B.posw[3 * m + a]);
            if (s < best) { best = s; near = id; }
        }
        B.lattice_id[m] = near;
    }
    B.active = true;
    return true;
}

void stage(Oracle& o, int s) {
    switch (s) {
    case IPL_STAGE_EQU_F: equ_f(o); break;
    case IPL_STAGE_FORCE: force(o); break;
    case IPL_STAGE_COLLIDE: collide(o); break;
    case IPL_STAGE_IP_F: ip_process(o, o.fl); break;
    case IPL_STAGE_SCALAR_COLLIDE:
        if (o.cfg.scalar == IPL_SCALAR_SALINITY) col_g_reg(o);
        else if (o.cfg.scalar == IPL_SCALAR_PHASE_FIELD) col_PF(o);
        break;
    case IPL_STAGE_IP_G: if (o.has_scalar) ip_process(o, o.sc); break;
    case IPL_STAGE_UPDATE_SCALAR:
        if (o.cfg.scalar == IPL_SCALAR_SALINITY) update_scalar(o, o.S);
        else if (o.cfg.scalar == IPL_SCALAR_PHASE_FIELD) update_scalar(o, o.phi);
        break;
    case IPL_STAGE_UPDATE_RHO: update_rho(o); break;
    case IPL_STAGE_SET_WALL_RHO:
        set_wall_rho(o, o.rho);
        if (o.cfg.scalar == IPL_SCALAR_PHASE_FIELD) set_wall_rho(o, o.phi);
        break;
    case IPL_STAGE_RESET_F: reset_force(o); break;
    case IPL_STAGE_EQU_G: if (o.has_scalar) equ_g(o); break;
    }
}

// One step in driver order (LockExchange.cu:220-278; Zhou_soliton4B1.cu:220-277)
// set_velocity1 + set_wall_boundary1 (Karman_vortexxz.cu:38-57): the macroscopic fields of the
// listed cells take the prescribed values and their populations the equilibrium of them (the
// reference's f form, collision.cu:3-14)
void open_boundary(Oracle& o) {
    const int Q = o.Q;
    const double c2 = o.c * o.c;
    for (size_t n = 0; n < o.ob_cells.size(); ++n) {
        const int id = o.ob_cells[n];
        const double* q = &o.ob_val[4 * n];
        o.p[id] = q[0]; o.u[id] = q[1]; o.v[id] = q[2]; o.w[id] = q[3];
        for (int k = 0; k < Q; ++k) {
            const double t = (o.L.cx[k] * o.c * o.u[id] + o.L.cy[k] * o.c * o.v[id] + o.L.cz[k] * o.c * o.w[id]) / c2;
            o.fl.f[(size_t)id * Q + k] = o.L.w[k] * (o.p[id] + 3.0 * t + 4.5 * t * t -
                1.5 * (o.u[id] * o.u[id] + o.v[id] * o.v[id] + o.w[id] * o.w[id]) / c2);
        }
    }
}

// u_tau of Spalding's law y+ = u+ + e^{-kB}(e^{k u+} - 1 - k u+ - (k u+)^2/2 - (k u+)^3/6),
// u+ = |u| / u_tau, y+ = y u_tau / nu: 15 Newton steps from 0.11 |u| (collision.cu:37-50)
double spalding_utau(double abs_u, double eta, double y_dist, double nu) {
    const double kap = 0.41, Beta = 5.5;
    double utau = abs_u * 0.11;
    for (int i = 0; i < 15; ++i) {
        const double tmp = kap * abs_u / (utau + eta);
        const double eps = -((abs_u / (utau + eta) - y_dist * utau / nu +
                              std::exp(-kap * Beta) * (std::exp(tmp) - 1.0 - tmp - std::pow(tmp, 2) / 2.0 - std::pow(tmp, 3) / 6.0)) /
                             (-y_dist / nu - abs_u / (std::pow(utau, 2) + eta) +
                              std::exp(-kap * Beta) * tmp / (utau + eta) * (-std::exp(tmp) + 1.0 + tmp + std::pow(tmp, 2) / 2.0) + eta));
        utau += eps;
    }
    return utau;
}

// wall_function (collision.cu:28-55) on one face's wall cells: Spalding's law solved for u_tau by
// 15 Newton steps, wall shear added to the two tangential force components.  (x, y, z) picks the
// wall normal; vx, vy are the tangential velocities, vz the normal one; Fa, Fb receive the shear.
void wall_function(Oracle& o, int x, int y, int z, const std::vector<int>& wall, const std::vector<double>& vx,
                   const std::vector<double>& vy, const std::vector<double>& vz, std::vector<double>& Fa,
                   std::vector<double>& Fb) {
    for (size_t n = 0; n < wall.size(); ++n) {
        const int id = wall[n];
        const double volume = o.del[0][id] * o.del[1][id] * o.dz;
        const double abs_u = std::sqrt(vx[id] * vx[id] + vy[id] * vy[id] + vz[id] * vz[id]);
        const double eta = abs_u * std::pow(10.0, -7) + std::pow(10.0, -20);
        const double kap = 0.41, Beta = 5.5;
        const double y_dist = x * o.del[0][id] / 2.0 + y * o.del[1][id] / 2.0 + z * o.dz / 2.0;
        const double area = volume / (y_dist * 2.0);
        (void)kap; (void)Beta;
        const double utau = spalding_utau(abs_u, eta, y_dist, o.cfg.wall_nu);
        Fa[id] += -vx[id] / (abs_u + eta) * std::pow(utau, 2.0) * o.rho[id] * (area / volume);
        Fb[id] += -vy[id] / (abs_u + eta) * std::pow(utau, 2.0) * o.rho[id] * (area / volume);
    }
}

// LES (basic.cu:285-339): coherent-structure model.  Velocity gradients by central differences
// over the neighbour table (a ghost has zero velocity and its owner's position; across a periodic
// seam the period is added, as in Force), the model coefficient from Q / E clipped to [-1, 1],
// and tau, tau_s of every calc cell from the molecular viscosity plus the eddy viscosity.  In 2D
// the z differences use the same neighbour twice (up = down = 2) and vanish, as in the reference.
void les(Oracle& o) {
    const int Q = o.Q;
    const int up = Q == 9 ? 2 : 5, down = Q == 9 ? 2 : 6;
    const double dz = o.dz;
    std::vector<double> tau(o.N), taus(o.N);
    #pragma omp parallel for schedule(static)
    for (int id = 0; id < o.N; ++id) {
        const int* nb = &o.neib[(size_t)id * Q];
        auto vel = [&](const std::vector<double>& a, int j) { return j < o.N ? a[j] : 0.0; };
        auto grad = [&](const std::vector<double>& a, int jp, int jm, int axis) {
            return (vel(a, jp) - vel(a, jm)) / (dpos(o, axis, jp, jm) + dz / 1000);
        };
        const double nxx = grad(o.u, nb[1], nb[3], 0), nxy = grad(o.v, nb[1], nb[3], 0), nxz = grad(o.w, nb[1], nb[3], 0);
        double nyx = 0, nyy = 0, nyz = 0;
        if (Q != 9) { nyx = grad(o.u, nb[2], nb[4], 1); nyy = grad(o.v, nb[2], nb[4], 1); nyz = grad(o.w, nb[2], nb[4], 1); }
        const double nzx = grad(o.u, nb[up], nb[down], 2), nzy = grad(o.v, nb[up], nb[down], 2), nzz = grad(o.w, nb[up], nb[down], 2);
        const double arr[9] = {(nxx + nxx) / 2.0, (nxy + nyx) / 2.0, (nxz + nzx) / 2.0, (nyx + nxy) / 2.0, (nyy + nyy) / 2.0,
                               (nyz + nzy) / 2.0, (nzx + nxz) / 2.0, (nzy + nyz) / 2.0, (nzz + nzz) / 2.0};
        const double Qi = -1.0 / 2.0 * (nxx * nxx + nxy * nyx + nxz * nzx + nyx * nxy + nyy * nyy + nyz * nzy +
                                        nzx * nxz + nzy * nyz + nzz * nzz);
        const double E = 1.0 / (2.0 * std::pow(nxx + nxy + nxz + nyx + nyy + nyz + nzx + nzy + nzz, 2) + std::pow(10.0, -40));
        const double F_cs = std::max(-1.0, std::min(Qi / E, 1.0));
        const double C = 0.05 * std::pow(std::fabs(F_cs), 3.0 / 2.0);
        double S = 0;
        for (int k = 0; k < 9; ++k) S += std::sqrt(2.0 * std::pow(arr[k], 2));
        const double c2dt = o.c * o.c * o.dt;
        tau[id] = 3.0 * ((o.cfg.mu_L + o.phi[id] * (o.cfg.mu_H - o.cfg.mu_L)) / o.rho[id] + C * S * std::pow(dz, 2)) / c2dt + 0.5;
        taus[id] = 3.0 * (std::pow(10.0, -9) + C * S * std::pow(dz, 2) / 1000) / c2dt + 0.5;
    }
    for (int id = 0; id < o.N; ++id) { o.tau[id] = tau[id]; o.taus[id] = taus[id]; }
}

// the driver's six calls (LockExchange.cu:214-219): faces +x, -x, +z, -z, +y, -y
void wall_functions(Oracle& o) {
    int dirs[7];
    dirs[1] = 1; dirs[3] = 3;
    if (o.Q == 9) { dirs[2] = 0; dirs[4] = 0; dirs[5] = 2; dirs[6] = 4; }
    else { dirs[2] = 2; dirs[4] = 4; dirs[5] = 5; dirs[6] = 6; }
    std::vector<int> wall[7];
    for (int f = 1; f <= 6; ++f)
        for (int i = 0; i < o.N; ++i)
            if (o.neib[(size_t)i * o.Q + dirs[f]] >= o.N) wall[f].push_back(i);
    wall_function(o, 1, 0, 0, wall[1], o.v, o.w, o.u, o.Fy, o.Fz);
    wall_function(o, 1, 0, 0, wall[3], o.v, o.w, o.u, o.Fy, o.Fz);
    wall_function(o, 0, 0, 1, wall[5], o.u, o.v, o.w, o.Fx, o.Fy);
    wall_function(o, 0, 0, 1, wall[6], o.u, o.v, o.w, o.Fx, o.Fy);
    wall_function(o, 0, 1, 0, wall[2], o.u, o.w, o.v, o.Fx, o.Fz);
    wall_function(o, 0, 1, 0, wall[4], o.u, o.w, o.v, o.Fx, o.Fz);
}

void step(Oracle& o, bool with_ib = true) {
    if (o.cfg.wall_function) wall_functions(o);
    for (int s = IPL_STAGE_EQU_F; s <= IPL_STAGE_RESET_F; ++s) {
        stage(o, s);
        if (s == IPL_STAGE_UPDATE_RHO && o.cfg.les) les(o);
    }
    if (o.ib.active && with_ib) ib_substep(o);
    if (!o.ob_cells.empty()) open_boundary(o);
}

// One stage of the IB sub-step (ipl_ib_stage): FLOW = a step without it, then the kernels of
// ib_substep in its order; FORCE clears F and Gw first and POINTS clears F after, as ib_substep.
int ib_stage(Oracle& o, int which) {
    auto& B = o.ib;
    if (!B.active) return IPL_ERR_ARGUMENT;
    switch (which) {
    case IPL_IB_STAGE_FLOW: step(o, false); break;
    case IPL_IB_STAGE_SPM: if (!B.cfg.spm) return IPL_ERR_ARGUMENT; ib_spm(o); break;
    case IPL_IB_STAGE_FORCE:
        reset_force(o);
        std::fill(B.Gw.begin(), B.Gw.end(), 0.0);
        if (B.cfg.method == IPL_IB_STRESS) ib_stress(o);
        else ib_direct(o);
        break;
    case IPL_IB_STAGE_UPDATE_VEL: ib_update_vel(o); break;
    case IPL_IB_STAGE_BODY: ib_body(o); break;
    case IPL_IB_STAGE_POINTS: if (B.cfg.moving) ib_points(o); reset_force(o); break;
    default: return IPL_ERR_ARGUMENT;
    }
    return 0;
}

}  // namespace

extern "C" {

void* orc_create(const ipl_config* cfg, const ipl_init* init, char* err, int errlen) {
    Oracle* o = new Oracle();
    if (!build(*o, *cfg, init)) {
        if (err && errlen > 0) std::snprintf(err, errlen, "%s", o->err.c_str());
        delete o;
        return nullptr;
    }
    return o;
}

void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }

int orc_step(void* h, int n) {
    Oracle& o = *static_cast<Oracle*>(h);
    for (int i = 0; i < n; ++i) step(o);
    return 0;
}

double orc_time_steps(void* h, int n) {
    Oracle& o = *static_cast<Oracle*>(h);
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n; ++i) step(o);
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

int orc_ib_stage(void* h, int which) { return ib_stage(*static_cast<Oracle*>(h), which); }

int orc_stage(void* h, int s) {
    stage(*static_cast<Oracle*>(h), s);
    return 0;
}

int orc_info(void* h, int64_t* n_calc, int64_t* n_ghost) {
    Oracle& o = *static_cast<Oracle*>(h);
    *n_calc = o.N; *n_ghost = o.G;
    return 0;
}

static std::vector<double>* field_ptr(Oracle& o, int field) {
    switch (field) {
    case IPL_FIELD_P: return &o.p; case IPL_FIELD_U: return &o.u; case IPL_FIELD_V: return &o.v;
    case IPL_FIELD_W: return &o.w; case IPL_FIELD_RHO: return &o.rho; case IPL_FIELD_S: return &o.S;
    case IPL_FIELD_PHI: return &o.phi; case IPL_FIELD_FX: return &o.Fx; case IPL_FIELD_FY: return &o.Fy;
    case IPL_FIELD_FZ: return &o.Fz; case IPL_FIELD_TAU: return &o.tau; case IPL_FIELD_TAUS: return &o.taus;
    default: return nullptr;
    }
}

int orc_read(void* h, int field, double* out, size_t count) {
    Oracle& o = *static_cast<Oracle*>(h);
    if (field == IPL_FIELD_F || field == IPL_FIELD_G) {
        PopSet& P = field == IPL_FIELD_F ? o.fl : o.sc;
        if (field == IPL_FIELD_G && !o.has_scalar) return IPL_ERR_ARGUMENT;
        if (count != (size_t)P.Q * o.N) return IPL_ERR_ARGUMENT;
        for (int k = 0; k < P.Q; ++k)
            for (int id = 0; id < o.N; ++id) out[(size_t)k * o.N + id] = P.f[(size_t)id * P.Q + k];
        return 0;
    }
    std::vector<double>* v = field_ptr(o, field);
    if (!v || count != (size_t)o.N) return IPL_ERR_ARGUMENT;
    std::copy(v->begin(), v->begin() + o.N, out);
    return 0;
}

int orc_write(void* h, int field, const double* in, size_t count) {
    Oracle& o = *static_cast<Oracle*>(h);
    if (field == IPL_FIELD_F || field == IPL_FIELD_G) {
        PopSet& P = field == IPL_FIELD_F ? o.fl : o.sc;
        if (count != (size_t)P.Q * o.N) return IPL_ERR_ARGUMENT;
        for (int k = 0; k < P.Q; ++k)
            for (int id = 0; id < o.N; ++id) P.f[(size_t)id * P.Q + k] = in[(size_t)k * o.N + id];
        return 0;
    }
    std::vector<double>* v = field_ptr(o, field);
    if (!v || count != (size_t)o.N) return IPL_ERR_ARGUMENT;
    std::copy(in, in + count, v->begin());
    return 0;
}

// output() global sums (basic.cu:9-13): volume = delX delY / dz^2
int orc_global_sums(void* h, double out[3]) {
    Oracle& o = *static_cast<Oracle*>(h);
    double tp = 0, ts = 0, tf = 0;
    for (int i = 0; i < o.N; ++i) {
        const double vol = o.del[0][i] * o.del[1][i] * o.del[2][i] / (o.dz * o.dz * o.dz);
        tp += o.p[i] * vol; ts += o.S[i] * vol; tf += o.phi[i] * vol;
    }
    out[0] = tp; out[1] = ts; out[2] = tf;
    return 0;
}

// Builder tables for cross-checks: 0 = neib (Ntot*Q), 1 = nextB, 2 = nextK (G*Q, flow lattice)
int64_t orc_table(void* h, int which, int32_t* out, int64_t count) {
    Oracle& o = *static_cast<Oracle*>(h);
    const std::vector<int>& t = which == 0 ? o.neib : (which == 1 ? o.fl.nextB : o.fl.nextK);
    if (!out) return (int64_t)t.size();
    if (count != (int64_t)t.size()) return -1;
    std::copy(t.begin(), t.end(), out);
    return count;
}

void orc_lagrange(const double* x, double xa, int n, double* coeff) { lagrange(coeff, x, xa, n); }

// reset_items_dt + reset_pressure (basic.cu:100-110, 262-273): dt / alpha, c * alpha (the
// lattice vectors are c times the integer directions), p / alpha^2, f = feq(p, u) with the new c
int orc_change_dt(void* h, double alpha) {
    Oracle& o = *static_cast<Oracle*>(h);
    o.dt /= alpha;
    o.c *= alpha;
    o.cfg.dt = o.dt;
    const int Q = o.Q;
    for (int id = 0; id < o.N; ++id) {
        o.p[id] /= kMutant == 11 ? alpha : alpha * alpha;  // mutant 11: p / alpha
        for (int k = 0; k < Q; ++k) {
            const double tmp = (o.L.cx[k] * o.c * o.u[id] + o.L.cy[k] * o.c * o.v[id] + o.L.cz[k] * o.c * o.w[id]) / (o.c * o.c);
            o.fl.f[(size_t)id * Q + k] = o.L.w[k] * (o.p[id] + 3.0 * tmp + 4.5 * tmp * tmp -
                1.5 * (o.u[id] * o.u[id] + o.v[id] * o.v[id] + o.w[id] * o.w[id]) / (o.c * o.c));
        }
    }
    return 0;
}

// the LES pass alone on the current macroscopic fields (for its closed-form check)
int orc_les(void* h) {
    les(*static_cast<Oracle*>(h));
    return 0;
}

double orc_spalding_utau(double abs_u, double y_dist, double nu) {
    return spalding_utau(abs_u, abs_u * std::pow(10.0, -7) + std::pow(10.0, -20), y_dist, nu);
}

int orc_set_open_boundary(void* h, const int* cells, int n, const double* p, const double* u, const double* v,
                          const double* w) {
    Oracle& o = *static_cast<Oracle*>(h);
    o.ob_cells.assign(cells, cells + n);
    o.ob_val.resize((size_t)4 * n);
    for (int i = 0; i < n; ++i) {
        if (cells[i] < 0 || cells[i] >= o.N) return IPL_ERR_ARGUMENT;
        o.ob_val[4 * i] = p[i]; o.ob_val[4 * i + 1] = u[i]; o.ob_val[4 * i + 2] = v[i]; o.ob_val[4 * i + 3] = w[i];
    }
    return 0;
}

// hydrostatic_pressure (basic.cu:239-252): p of every calc cell from its z position
int orc_hydrostatic_pressure(void* h, int boussinesq, double center_z, double* p_out) {
    const Oracle& o = *static_cast<Oracle*>(h);
    for (int i = o.N - 1; i >= 0; --i) {
        if (boussinesq) p_out[i] = 3 * 9.81 * (o.cfg.nz / 2.0 * o.dz - o.pos[2][i]) / std::pow(o.c, 2) / 1000;
        else p_out[i] = 3 * 9.81 * (center_z - o.pos[2][i]) / std::pow(o.c, 2);
    }
    return 0;
}

// set_walln (set_bound.cu:8-23) on the oracle's neighbour table; returns the number of cells
int orc_wall_cells(void* h, int face, int* out) {
    const Oracle& o = *static_cast<Oracle*>(h);
    int dir1 = 1, dir2, dir3 = 3, dir4, dir5, dir6;
    if (o.Q == 9) { dir2 = 0; dir4 = 0; dir5 = 2; dir6 = 4; }
    else { dir2 = 2; dir4 = 4; dir5 = 5; dir6 = 6; }
    const int dirs[7] = {0, dir1, dir2, dir3, dir4, dir5, dir6};
    int m = 0;
    for (int i = 0; i < o.N; ++i)
        if (o.neib[(size_t)i * o.Q + dirs[face]] >= o.N) out[m++] = i;
    return m;
}

int orc_ib_attach(void* h, const ipl_ib_config* cfg, char* err, int errlen) {
    Oracle& o = *static_cast<Oracle*>(h);
    if (!ib_attach(o, *cfg)) {
        if (err && errlen > 0) std::snprintf(err, errlen, "%s", o.err.c_str());
        return IPL_ERR_ARGUMENT;
    }
    return 0;
}

int orc_ib_body_state(void* h, double* out) {
    const auto& B = static_cast<Oracle*>(h)->ib;
    for (int a = 0; a < 3; ++a) { out[a] = B.posB[a]; out[3 + a] = B.velB[a]; out[6 + a] = B.angleVB[a]; }
    for (int i = 0; i < 4; ++i) out[9 + i] = B.quat[i];
    for (int a = 0; a < 3; ++a) { out[13 + a] = B.FB[a]; out[16 + a] = B.Torque[a]; }
    return 0;
}

int orc_ib_read_markers(void* h, int which, double* out, size_t count) {
    const auto& B = static_cast<Oracle*>(h)->ib;
    if (which == IPL_IB_NODE) {
        if (count != (size_t)B.M) return IPL_ERR_ARGUMENT;
        const Oracle& o = *static_cast<Oracle*>(h);
        for (int m = 0; m < B.M; ++m) {
            const int id = B.lattice_id[m];
            out[m] = (double)(((size_t)o.cidx[0][id] * o.n[1] + o.cidx[1][id]) * o.n[2] + o.cidx[2][id]);
        }
        return 0;
    }
    const std::vector<double>* v = which == IPL_IB_POS ? &B.posw : which == IPL_IB_VEL ? &B.velw :
                                   which == IPL_IB_FORCE ? &B.Gw : which == IPL_IB_NORMAL ? &B.nB : nullptr;
    if (!v || count != v->size()) return IPL_ERR_ARGUMENT;
    std::copy(v->begin(), v->end(), out);
    return 0;
}

double orc_ib_peskin4(double r) { return ib_peskin4(r); }
void orc_ib_rotation(const double* q, double* R) { ib_rotation(q, R); }

int orc_mrt(int Q, double* M, double* Minv, int32_t* stress, int32_t* nstress) {
    if (Q != 9 && Q != 19 && Q != 27) return -1;
    Lattice L = make_lattice(Q);
    std::vector<double> m, mi; std::vector<int> st;
    mrt_matrix(L, m, st);
    mrt_inverse(Q, m, mi);
    std::copy(m.begin(), m.end(), M);
    std::copy(mi.begin(), mi.end(), Minv);
    for (size_t i = 0; i < st.size(); ++i) stress[i] = st[i];
    *nstress = (int32_t)st.size();
    return 0;
}

int orc_lattice(int Q, int32_t* c, double* w) {
    if (Q != 7 && Q != 9 && Q != 19 && Q != 27) return -1;
    Lattice L = make_lattice(Q);
    for (int k = 0; k < Q; ++k) { c[3 * k] = L.cx[k]; c[3 * k + 1] = L.cy[k]; c[3 * k + 2] = L.cz[k]; w[k] = L.w[k]; }
    return 0;
}

}  // extern "C"

/* PARITY
 * Pinned (tests/test_oracle.py): MRT rows vs the reference's MRT_D3Q9/MRT_D3Q27.csv and M^-1 vs
 * Minv_*.csv; lattice moments (sum w = 1, sum w c c = I/3); Lagrange weights (partition of unity,
 * exact on polynomials of degree < n, unit vector at a node); equilibrium moments; a rest state
 * is a fixed point; MRT with all rates 1/tau equals SRT; uniform grid = exact shift; periodic
 * translation invariance; volume-weighted mass conservation with walls; Poiseuille profile with
 * bounce-back; slip keeps a plug flow; set_bound2D tables against the mirror rule.
 * Immersed boundary (tests/test_oracle_ibm.py): weightFunction's Peskin moments (sum 1, first
 * moment 0, sum of squares 3/8, support); set_quaternionS against scipy's rotation; the marker
 * generator against the reference's sphere_points_n7.csv; the stress method under a uniform p
 * gives the closed-form pressure traction and leaves the fluid at rest; direct forcing in a
 * uniform stream gives (u0/dt) rho_f dIBM^2 dz per marker and the matching momentum loss.
 * update_IBbody / update_IBpoint / search_IBlattice: the first step of a body in fluid at rest
 * (buoyancy only) against the AB2 closed form, rigid marker motion and brute-force nearest cells;
 * later steps (torque, quaternion AB2) only through the GPU comparison.
 * Open boundary (set_velocity1 + set_wall_boundary1): a uniform stream held by it stays uniform
 * and the open cells hold exactly the reference's equilibrium (tests/test_boundary_setters.py).
 * hydrostatic_pressure: closed form on a stretched grid.  set_walln: brute force over
 * coordinates, with and without solid cells.
 * wall_function: its Newton solve against scipy's root of Spalding's law (within the reference's
 * regularisation); the face mapping and the force application are pinned only through the GPU
 * comparison: parity unpinned against the reference itself.
 * LES: a linear velocity field (exact central differences) gives the closed-form tau and tau_s
 * at every interior cell (tests/test_boundary_setters.py::test_les_closed_form).
 * tests/test_oracle_pins.py, each pin also shown to fail on a planted mistake (ORC_MUTANT):
 *   Force: Guo forcing, u_z = -(n + 1/2) g dt in a periodic box; the pressure-gradient term on a
 *     linear rho over a stretched axis; the chemical potential of the tanh profile (zero at
 *     second order for the equilibrium width, closed form for twice the width, x over delX^2);
 *   col_PF: a flat interface stays put with sum phi vol conserved; a slab translates with a
 *     uniform stream; in an accelerating stream it follows the fluid's lattice momentum (the
 *     d(phi u)/dt term);
 *   col_g_reg: an advected sine mode decays as exp(-D k^2 t) and moves at u (D3Q7 and same-Q);
 *   corner scheme (set_out num3, set_fin): equals the one-axis scheme along x (y uniform) and
 *     along y (x uniform), exact streaming on a uniform grid, sum p vol conserved with slip walls
 *     on a grid refined in x and y;
 *   SPM_ellipse3D: v_B exactly inside a rotated ellipsoid (scipy's body frame), rest outside;
 *   reset_items_dt + reset_pressure: p c^2, momentum and momentum flux kept, shear-wave decay
 *     at the kept lattice viscosity.
 * Parity unpinned: the reference's own FP32 trajectories on NVIDIA hardware (CUDA device math
 * and atomicAdd ordering cannot run here).
 */
