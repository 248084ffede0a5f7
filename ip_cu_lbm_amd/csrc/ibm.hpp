// ibm.hpp - immersed boundary on the fused path's line layout (function/ibm.cu).
//
// One rigid body carried by Lagrangian markers.  Every step, after the fused step has written
// f^{n+1} (and the Force field F^n), the sub-step of the drivers (SphereFall.cu:372-383) runs:
//   k_ib_rank    markers sorted by (nearest cell, index) by counting ranks, every workgroup
//                holding all markers' cells in LDS; the bounding box of their node blocks and the
//                run of sorted markers nearest to each box cell
//   k_ib_weights (direct forcing, with the ranking) every marker's node weights: slot weights and
//                box indices, axis weights, and the marks of the box cells its block covers
//   k_ib_near    (multi-direct forcing, with the ranking) each marker's list of the markers whose
//                node blocks can meet its own
//   k_ib_ubase   (direct forcing) update_rho's velocity at every cell of the box
//   k_ib_force   one 64-lane workgroup per marker, lanes over the block slots: interpolation of f
//                (stress method, get_IBMGw2, with its own node weights) or of the node velocity
//                (direct forcing, IB_directForcing) reduced in a fixed tree, marker force (stress
//                method: one lane per direction, moment sums in a fixed tree), and the
//                contribution of every block slot to the spread force
//   k_ib_correct (multi-direct forcing) a later pass in marker space: the velocity change at the
//                marker is the overlap of its node block with every near marker's, weighted by
//                their last forces (the spread is linear); the last pass writes the slots' shares
//                of the force summed over the passes.  Where node weights are not products of
//                axis weights, force + gather are repeated per pass instead.
//   k_ib_gather  one wavefront per cell of the box: the spread force of the cell, summed over
//                markers in index order per slot and over slots in a fixed tree (no atomics: the
//                sum has one order whatever the scheduling)
//   k_ib_apply   update_velIBM on the cells of the box: f += w dt (e . F); clears a moving body's
//                marks
//   k_ib_body    one workgroup: FP64 force / torque sums in a fixed tree, buoyancy, Adams-Bashforth
//                update of the body (update_IBbody)
//   k_ib_points  one work-item per marker: update_IBpoint and search_IBlattice
// All marker and body arithmetic is FP64 whatever the population storage.
#pragma once

#include "hip_compat.hpp"

#include <cstdint>
#include <vector>

#include "../../include/ipl.h"
#include "engine.hpp"

namespace ipl {

// What the IB kernels need to know of the grid and the fused layout.
struct IbGrid {
    int n[3];
    int periodic[3];
    int s, o1, o2, ns;   // cell = ((c[o1] * n[o2]) + c[o2]) * ns + c[s]
    int64_t ncell;
    int dim, Q, Qs, scalar;
    double dz, dt, c, tau;
    double rho_L, rho_H, beta_s;
    double period[3];    // sum of widths of periodic axes
};

class IbSolver {
public:
    ~IbSolver();
    // Uploads markers and body, finds the nearest cells; pos[a] = cell-centre coordinates per axis.
    Status attach(const ipl_ib_config& cfg, const IbGrid& g, const std::vector<double>* pos, hipStream_t st);
    // The IB sub-step on the populations f of the fused path (storage F64 or F32), with F the Force
    // field of the step (same type as the arithmetic) and g the scalar populations (or null).
    // parity: the step's buffer parity.  The step's last kernels (update_IBbody and, for a moving
    // body, update_IBpoint, search_IBlattice and the next step's ranking) run on a second stream
    // beside the last gather pass, update_velIBM and the next fused step; a moving body's ranking
    // is therefore held twice, the step of parity p reading set p while set 1 - p is written.
    Status substep(void* f, const void* g, const void* F, int storage, int parity);
    // The second stream's work joined into the context's stream (the end of every launch of steps).
    Status join();
    // the tail on the context's stream after the last gather pass (no second stream)
    void set_inline_tail(bool on) { inline_tail_ = on; }
    bool inline_tail() const { return inline_tail_; }
    // The ranking the step of this parity reads, before a graph of steps is captured.
    Status ensure_ranked(int parity);
    // One stage of the sub-step (IPL_IB_STAGE_*) on the current state (on the context's stream).
    Status stage(int which, void* f, const void* g, const void* F, int storage, int parity);
    Status body_state(double out[19]);
    // IPL_ERR_UNSUPPORTED when a marker block left the bound the box was sized for
    Status check();
    Status read_markers(int which, double* out, size_t count);
    bool active() const { return M_ > 0; }
    int64_t device_bytes() const { return (int64_t)mem_.total(); }
    // device bytes attach() allocates for this body on a grid of n cells per axis
    static int64_t estimate_bytes(const ipl_ib_config& cfg, const int n[3], const int periodic[3], int dim, double dz);
    // multi-direct forcing's passes after the first run in marker space (k_ib_correct)
    static bool marker_space_passes(const ipl_ib_config& cfg, int dim, const int n[3], const int periodic[3]);
    static int64_t box_extent(const ipl_ib_config& cfg, const int n[3], const int periodic[3], int dim, double dz,
                              double rmax, int ext[3], int* pad);

    // slots of a marker's node block
    int slots() const { return S_; }
    int markers() const { return M_; }

private:
    int set_of(int parity) const { return cfg_.moving ? parity & 1 : 0; }
    Status rank(int set, hipStream_t st);
    Status spm(void* f, const void* g, const void* F, int storage, int set);
    // the forcing passes; tail_set >= 0: after the last force pass the step's tail runs on the
    // second stream, ranking into tail_set
    Status forcing(void* f, const void* g, const void* F, int storage, int set, int tail_set);
    Status update_vel(void* f, int storage, int set);
    Status body(hipStream_t st);
    Status points(hipStream_t st);

    IbGrid g_{};
    ipl_ib_config cfg_{};
    hipStream_t st_ = nullptr;
    int M_ = 0, S_ = 0, reach_lo_ = -1, reach_hi_ = 1;
    int bpad_ = 0;  // extra box cells on every side for the smoothed-profile forcing
    int64_t vmax_ = 0;
    int box_max_[3] = {0, 0, 0};
    DeviceArena mem_;
    double *pos_[3] = {nullptr, nullptr, nullptr};
    double *oposw_ = nullptr, *onB_ = nullptr, *posw_ = nullptr, *nB_ = nullptr, *velw_ = nullptr, *Gw_ = nullptr;
    double *contrib_ = nullptr, *Fbox_ = nullptr, *Ubox_ = nullptr, *body_ = nullptr;
    double* wcache_ = nullptr;  // direct forcing: slot weights [M][S] (computed with the ranking)
    int* bcache_ = nullptr;     // and their box indices (-2: ghost node)
    // marker-space passes: axis weights [M][3][5], interpolated velocity, the forces of the last two
    // passes and their sum [M][3]
    double *wax_ = nullptr, *um_ = nullptr, *gm_[2] = {nullptr, nullptr}, *gtot_ = nullptr;
    // and, per ranking set, each marker's list of markers whose node blocks can meet its own
    int *near_[2] = {nullptr, nullptr}, *near_count_[2] = {nullptr, nullptr};
    int* node_ = nullptr;
    // the ranking (a set per step parity for a moving body, one for a fixed one): the box of the
    // node blocks (+ an overflow flag), each box cell's run of sorted markers, the sorted keys
    int* box_[2] = {nullptr, nullptr};
    int2* range_[2] = {nullptr, nullptr};
    unsigned long long* sorted_[2] = {nullptr, nullptr};
    bool ranked_[2] = {false, false};  // the set holds the ranking of the current nodes (in stream order)
    // per ranking set: box cells some marker block covers (set by the ranking's weights for direct
    // forcing, by the force pass for the stress method; cleared by k_ib_apply for a moving body)
    int* mark_[2] = {nullptr, nullptr};
    int weights_set_ = -1;      // the ranking set whose node weights wcache_, bcache_, wax_ hold
    hipStream_t side_ = nullptr;       // the step's tail
    hipEvent_t forced_ = nullptr;      // the last force pass has been queued (Gw final)
    hipEvent_t tail_done_ = nullptr;   // the tail has been queued
    bool tail_pending_ = false;        // the context's stream has not yet waited for it
    bool inline_tail_ = false;         // the tail runs on the context's stream instead
};

constexpr int kIbMaxMarkers = 4096;  // the rank kernel holds every marker's key in 32 KiB of LDS

}  // namespace ipl
