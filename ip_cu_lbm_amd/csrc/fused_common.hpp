// fused_common.hpp - parameter blocks and launcher declarations of the fused path.
#pragma once

#include "hip_compat.hpp"

#include <cstddef>
#include <cstdint>

#include "../../include/ipl.h"

namespace ipl {
namespace fusedk {

struct FGeom {
    int n[3];
    int periodic[3];
    int s, o1, o2;   // stream axis and the two line axes
    int ns;          // cells per line
    int64_t lines, ncell;
    int tpl;         // work-items per line slot, ns rounded up to whole wavefronts (64):
                     // work-item i of a slot owns cell i of its line
    int lpw;         // lines per workgroup: consecutive lines of the XCD order share a workgroup
    int groups;      // workgroups of one step, ceil(lines / lpw)
    int lds_line;    // LDS elements (of the arithmetic type) of one line's rows
    int wall_bc;
    int flux;        // 1 = interpolated flux along s
    int need_grad;   // the scalar-field pass runs (phi, and rho beside it, for the gradients)
    int grad_rho;    // rho varies: the pressure-gradient term of Force is nonzero
    int inline_rho;  // salinity: the step forms the neighbours' rho from their g rows itself
                     // (along the line through LDS), and no scalar-field pass runs
    int inline_phi;  // 2D phase field: the step forms the neighbours' phi (and rho) from their g
                     // rows itself, like inline_rho, and no scalar-field pass runs
    int grad_phi;    // phase field: grad phi and its Laplacian (Force, col_PF)
    int wall_fn;     // wall-function shear on the cells next to walls (collision.cu:28-55)
    int les;         // per-cell tau, tau_s rows set by the LES pass (basic.cu:285-339)
    int zero_force;  // no gravity, gradient or wall term: F stays zero, its rows are neither
                     // read nor written by the step (they hold the zeros of the set-up)
    int half_force;  // u = sum e c f + F_old dt / (2 rho) (update_rho); 0 after an immersed-boundary
                     // sub-step, whose update_velIBM leaves u = sum e f (ibm.cu:356-360)
    int scalar;
    float inv_scale; // FP16 storage: stored value = f / w_k * scale
    float scale;
};

template <typename MT>
struct FTables {
    const MT* W;       // [sign(0:+,1:-)][j][ns]
    const MT* rdel;    // [ns] dz / width along s
    const MT* invd[3]; // per axis index: 1 / (pos[+1] - pos[-1] + dz/1000), seam-corrected
    const MT* inv_wx2; // per x index: 1 / width_x^2
    const MT* wid[3];  // per axis index: cell width (wall function)
};

template <typename ST, typename MT>
struct FBufs {
    const ST* f_src; ST* f_dst;
    const ST* g_src; ST* g_dst;
    const MT* F_old; MT* F_new;   // [3][ncell]
    const MT* rho_n;              // [ncell] rho of step n (gradient pass), or null
    const MT* phi_n;              // [ncell] phi of step n, or null
    MT *phiold, *uold, *vold, *wold;
    const MT *tau, *taus;         // [ncell] per-cell tau, tau_s (LES), or null
};

struct FParams {
    double dz, dt, c;
    double rates[27];
    double rho_L, rho_H, beta_s, gravity, sigma, pf_thickness, pf_tau;
    double tau, taus;
    int boussinesq, collision;
    // FP64 host-precomputed factors, so that the step divides once per cell (1 / rho)
    double inv_c, inv_c2, half_dt, fk_coef;  // 1/c, 1/c^2, dt/2, 3 dt / c
    double itau, om_s;                       // 1/tau, 1 - 1/tau_s
    double rnorm[27];                        // MRT: rate of row r (1/tau on stress rows) / |row r|^2
    double press;                            // c^2 / 3
    double st_beta, kap, inv_dz2;            // 4 * 12 sigma / W, 3 sigma W / 2, 1/dz^2
    double inv_dt, inv_pfW, pf_itau, pf_fcoef; // 1/dt, 1/W, 1/tau_phi, tau_phi - 1/2
    double wall_nu;                          // wall function: kinematic viscosity
    double inv_n2[27];                       // 1 / |moment row r|^2 (LES: per-cell stress rates)
    double mu_L, mu_H;                       // LES: dynamic viscosities
};

// The coefficients of the step in the arithmetic type, cast once on the host (the kernels then
// hold no FP64 constants and convert nothing).
template <typename MT>
struct FCoef {
    MT c, inv_c, inv_c2, half_dt, fk_coef, dt;
    MT uc;                             // velocity scale of sum e f: c
    MT itau, om_s, rnorm[27];
    MT wall_nu, dz;                    // kinematic viscosity of the wall function's log law; dz
    MT inv_n2[27], mu_L, mu_H;         // LES: 1 / |row r|^2, dynamic viscosities
    MT rho_L, drho, beta_s, gravity;   // drho = rho_H - rho_L (in MT)
    MT press, st_beta, kap, inv_dz2;
    MT inv_dt, inv_pfW, pf_itau, pf_fcoef;
    int boussinesq, collision;
};

template <typename MT>
inline FCoef<MT> make_coef(const FParams& P) {
    FCoef<MT> k{};
    k.c = (MT)P.c; k.inv_c = (MT)P.inv_c; k.inv_c2 = (MT)P.inv_c2; k.half_dt = (MT)P.half_dt;
    k.fk_coef = (MT)P.fk_coef; k.dt = (MT)P.dt;
    k.uc = (MT)P.c;
    k.itau = (MT)P.itau; k.om_s = (MT)P.om_s;
    for (int r = 0; r < 27; ++r) k.rnorm[r] = (MT)P.rnorm[r];
    k.rho_L = (MT)P.rho_L; k.drho = (MT)P.rho_H - (MT)P.rho_L; k.beta_s = (MT)P.beta_s; k.gravity = (MT)P.gravity;
    k.press = (MT)P.press; k.st_beta = (MT)P.st_beta; k.kap = (MT)P.kap; k.inv_dz2 = (MT)P.inv_dz2;
    k.inv_dt = (MT)P.inv_dt; k.inv_pfW = (MT)P.inv_pfW; k.pf_itau = (MT)P.pf_itau; k.pf_fcoef = (MT)P.pf_fcoef;
    k.boussinesq = P.boussinesq; k.collision = P.collision;
    k.wall_nu = (MT)P.wall_nu; k.dz = (MT)P.dz;
    for (int r = 0; r < 27; ++r) k.inv_n2[r] = (MT)P.inv_n2[r];
    k.mu_L = (MT)P.mu_L; k.mu_H = (MT)P.mu_H;
    return k;
}

// One job of a fused step launch (device memory): its geometry, tables, buffer roles at even and
// odd launches, coefficients, and the first workgroup of its lines.  A context stepped alone keeps
// a table of one job; batched launches build tables of many.
template <typename ST, typename MT>
struct BatchJob {
    FGeom G;
    FTables<MT> tb;
    FBufs<ST, MT> b[2];
    FCoef<MT> P;
    int first_block;
    int pad_;
};
// The initial fields of a context on the device (k_init_state): rows [p, u, v, w, scalar (S or
// phi), tau, tau_s][ncell] in canonical cell order; a row whose bit is clear in `present` holds
// dflt[r] everywhere and is not read.
struct InitArgs {
    const double* rows;
    unsigned present;
    double dflt[7];
    const double* wid[3];  // FP64 cell widths per axis index
    double dz, c;
};

template <typename ST, typename MT, int Q, int QS>
void fused_launch_init(const FGeom& G, const InitArgs& A, ST* f, ST* g, double* vol, MT* phiold, MT* uold, MT* vold,
                       MT* wold, MT* tau, MT* taus, hipStream_t st);
template <typename ST, typename MT, int Q, int QS>
void fused_launch_step(const BatchJob<ST, MT>* jobs, int njobs, int64_t blocks, int step, int npts, int s, bool flux,
                             bool ext, bool nt, unsigned block, size_t lds, hipStream_t st);
template <typename ST, typename MT, int Q, int QS>
size_t fused_lds_bytes(int s, int ns);
template <typename ST, typename MT, int QS>
void fused_launch_scalar_fields(const FGeom& G, const ST* g_src, MT* rho_n, MT* phi_n, const FCoef<MT>& P, hipStream_t st);
template <typename ST, typename MT, int Q, int QS>
void fused_launch_les(const FGeom& G, const FTables<MT>& tb, const ST* f, const ST* g, const MT* F, MT* vel, MT* nu,
                      MT* tau, MT* taus, const FCoef<MT>& P, unsigned block, hipStream_t st);
template <typename ST, typename MT, int Q, int QS>
void fused_launch_readout(const FGeom& G, const ST* f_src, const ST* g_src, const MT* F_old, double* out,
                          const FParams& P, hipStream_t st);
// to_double: population row q of src -> dst (FP64); otherwise dst (FP64) -> row q of wdst
template <typename ST, typename MT, int Q>
void fused_launch_convert(const FGeom& G, const ST* src, double* dst, int q, bool to_double, ST* wdst, hipStream_t st);

}  // namespace fusedk
}  // namespace ipl
