// graph.hpp - host-side build of the cell graph the kernels run on.
//
// The reference's data model is kept because the interpolated propagation is defined on it:
// calc cells in lnum order, one ghost cell appended per calc cell that has a missing link
// (set_neibghor_wall, basic.cu:203-237), a neighbour table, and per-ghost slip remap tables
// (set_bound2D/3D, set_bound.cu:25-147).  Everything is laid out structure-of-arrays
// ([direction][cell]) so that one thread per cell reads coalesced rows, and the Lagrange
// weights of the slab reconstruction (adv.cu:40-48,76-99) are precomputed here in FP64 once,
// instead of 20 divisions per direction per cell per step.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ipl.h"

namespace ipl {

enum FluxMode { FLUX_MODE_NONE = 0, FLUX_MODE_ONE_AXIS = 1, FLUX_MODE_CORNER = 2 };

struct DirTable {
    int Q = 0;
    int cx[27] = {0}, cy[27] = {0}, cz[27] = {0};
    double w[27] = {0};
    int opp[27] = {0};
    int comp(int k, int a) const { return a == 0 ? cx[k] : (a == 1 ? cy[k] : cz[k]); }
    int find(int x, int y, int z) const;
};
DirTable make_dirs(int Q);

struct Graph {
    int Q = 0, Qs = 0, dim = 0;
    int64_t N = 0, G = 0, Ntot = 0;
    int n[3] = {0, 0, 0};
    FluxMode flux_mode = FLUX_MODE_NONE;
    int flux_axis = -1;        // one-axis mode
    int npts = 5;
    double dz = 1, dt = 1, c = 1;
    double period[3] = {0, 0, 0};
    DirTable dirs, sdirs;

    std::vector<int32_t> cell_ijk;          // 3*N canonical coordinates of calc cells
    std::vector<int32_t> neib;              // [k][Ntot], flow lattice
    std::vector<int32_t> nextB, nextK;      // [k][G], flow lattice
    std::vector<int32_t> snextB, snextK;    // [k][G], scalar lattice
    std::vector<double> pos[3], del[3];     // [Ntot]
    // Lagrange weights of the slab reconstruction: [(ax*2 + (s<0))*npts + j][N] for the flux
    // axes (ax = 0..2), zero rows for axes that carry no flux.
    std::vector<double> weights;
    bool axis_has_weights[3] = {false, false, false};
};

// Builds the graph; returns false and fills err on an invalid configuration.
bool build_graph(const ipl_config& cfg, Graph& g, std::string& err);
// The checks and the scalar fields of build_graph (lattices, sizes, flux mode, stencil, calc-cell
// count) without its tables: enough to decide the path and size a context.
bool graph_header(const ipl_config& cfg, Graph& g, std::string& err);
// A box domain (no cell mask) as the fused path needs it: the header, the calc cells' coordinates
// (canonical order) and the ghost count, with no neighbour, slip or weight tables.
bool build_box_graph(const ipl_config& cfg, Graph& g, std::string& err);
// Lagrange weights of the slab reconstruction along `axis` for the cells of one line along it,
// [sign (0: +, 1: -)][j][n_axis]: those of build_graph, which on a box depend only on the index
// along the axis (computed on a one-line copy of the configuration).
bool axis_line_weights(const ipl_config& cfg, int axis, std::vector<double>& w, std::string& err);

// Lagrange_interpolation (adv.cu:40-48) with the same product order.
void lagrange_weights(double* coeff, const double* x, double xa, int n);

}  // namespace ipl
