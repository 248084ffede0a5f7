// plain.hpp - device views and run parameters shared by the HIP kernels and the host API.
#pragma once

#include "hip_compat.hpp"

#include <cstdint>

#include "../../include/ipl.h"
#include "graph.hpp"

namespace ipl {

// Scalar run parameters, passed by value to every kernel.
struct Params {
    double dz, dt, c;
    double rates[27];
    double rho_L, rho_H, beta_s, gravity, sigma, pf_thickness, pf_tau;
    int boussinesq, collision, scalar, wall_bc;
    int flux_mode, flux_axis, npts;
    int periodic[3];
    double period[3];
};

// Device pointers of the plain path.  Populations are [k][Ntot]; cell fields are [Ntot].
template <typename T>
struct PlainView {
    int64_t N, Ntot;
    T *f, *ftmp, *fout, *feq, *Fk;
    T *g, *gtmp, *gout, *geq;
    T *p, *u, *v, *w, *rho, *phi, *S, *tau, *taus, *Fx, *Fy, *Fz;
    T *phiold, *uold, *vold, *wold;
    T *pos[3], *del[3], *rdel[3];
    T *W;  // stencil weights [(axis*2 + sign)*npts + j][N]
    const int32_t *neib, *nextB, *nextK, *snextB, *snextK;
};

template <typename T>
int plain_stage(const PlainView<T>& s, const Params& P, int Q, int Qs, int stage, hipStream_t st);

}  // namespace ipl
