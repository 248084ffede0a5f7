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
    const MT* rdel;    // [ns] dz / width along s
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

struct FParams {
    double dz, dt, c;
    double rates[27];
    double rho_L, rho_H, beta_s, gravity, sigma, pf_thickness, pf_tau;
    double tau, taus;
    int boussinesq, collision;
};

template <typename ST, typename MT, int Q, int QS>
void fused_launch_step(const FGeom& G, const FTables<MT>& tb, const FBufs<ST, MT>& b, const FParams& P, int npts,
                       unsigned block, size_t lds, hipStream_t st);
template <typename ST, typename MT, int Q, int QS>
size_t fused_lds_bytes(int s, int ns);
template <typename ST, typename MT, int QS>
void fused_launch_scalar_fields(const FGeom& G, const ST* g_src, MT* rho_n, MT* phi_n, const FParams& P, hipStream_t st);
template <typename ST, typename MT, int Q, int QS>
void fused_launch_readout(const FGeom& G, const ST* f_src, const ST* g_src, const MT* F_old, double* out,
                          const FParams& P, hipStream_t st);
// to_double: population row q of src -> dst (FP64); otherwise dst (FP64) -> row q of wdst
template <typename ST, typename MT, int Q>
void fused_launch_convert(const FGeom& G, const ST* src, double* dst, int q, bool to_double, ST* wdst, hipStream_t st);

}  // namespace fusedk
}  // namespace ipl
