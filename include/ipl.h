/*
 * ipl.h - C ABI of the MI355X-native interpolation-based lattice Boltzmann step.
 *
 * The reference (Kasahara1341/IP_CU_LBM) has no plugin or FFI layer: its boundary is the
 * CUDA C++ template kernel library declared in all.hpp:39-156, linked into one driver main()
 * per case.  This header replaces that library for the driver run loop
 * (LockExchange/LockExchange.cu:199-279 and its siblings): init, N x step, readout.
 *
 * Conventions
 *   - every entry point returns int status (IPL_OK == 0); ipl_error_string() names the code,
 *     ipl_last_error_message() gives the detail of the last failure on a context;
 *   - the library owns device memory; host pointers are borrowed for the duration of a call;
 *   - one context is not thread-safe; distinct contexts are independent;
 *   - host arrays of cell fields are in canonical cell order: x slowest, then y, then z
 *     fastest, solid cells skipped (the reference's lnum order, LockExchange.cu:57-74).
 */
#ifndef IPL_H
#define IPL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------- */
enum {
    IPL_OK = 0,
    IPL_ERR_INVALID_CONFIG = 1,   /* rejected at init (e.g. cell width below the lattice spacing) */
    IPL_ERR_HIP = 2,              /* a HIP runtime call failed */
    IPL_ERR_OUT_OF_MEMORY = 3,
    IPL_ERR_NONFINITE = 4,        /* NaN/Inf detected in the macroscopic fields */
    IPL_ERR_UNSUPPORTED = 5,      /* the requested path cannot run this configuration */
    IPL_ERR_ARGUMENT = 6,         /* bad argument to a call (size, field id, ...) */
    IPL_ERR_BUDGET = 7            /* batch admission: the job does not fit the memory budget */
};

/* ---- configuration enums ------------------------------------------------------------------- */
enum { IPL_BC_SLIP = 0, IPL_BC_BOUNCE_BACK = 1 };            /* IP_process 'slip' argument, adv.cu:246 */
enum { IPL_COLL_MRT = 0, IPL_COLL_SRT = 1 };                 /* col_f_MRT / col_f_SRT, collision.cu:115,151 */
enum { IPL_SCALAR_NONE = 0, IPL_SCALAR_SALINITY = 1, IPL_SCALAR_PHASE_FIELD = 2 };
enum { IPL_F64 = 0, IPL_F32 = 1, IPL_F16 = 2 };              /* population storage type */
enum { IPL_PATH_AUTO = 0, IPL_PATH_PLAIN = 1, IPL_PATH_FUSED = 2 };
/* flux_axes: bit 0 = x, bit 1 = y, bit 2 = z; IPL_FLUX_REFERENCE selects the reference's
 * choice (x in 2D, x and y with the two-axis corner scheme in 3D, adv.cu:50-147). */
enum { IPL_FLUX_NONE = 0, IPL_FLUX_X = 1, IPL_FLUX_Y = 2, IPL_FLUX_Z = 4, IPL_FLUX_REFERENCE = -1 };

/* fields readable / writable with ipl_read / ipl_write */
enum {
    IPL_FIELD_P = 0,      /* pressure-like zeroth moment p = sum f            (update_rho, collision.cu:238) */
    IPL_FIELD_U = 1, IPL_FIELD_V = 2, IPL_FIELD_W = 3,
    IPL_FIELD_RHO = 4,    /* mass density rho_L + (rho_H - rho_L) phi + beta_s S */
    IPL_FIELD_S = 5,      /* salinity */
    IPL_FIELD_PHI = 6,    /* phase field */
    IPL_FIELD_FX = 7, IPL_FIELD_FY = 8, IPL_FIELD_FZ = 9,
    IPL_FIELD_TAU = 10, IPL_FIELD_TAUS = 11,
    IPL_FIELD_F = 12,     /* flow populations, [k][cell] (k-major), cells canonical */
    IPL_FIELD_G = 13,     /* scalar populations, [k][cell] */
    IPL_FIELD_COUNT = 14
};

/* Run configuration.  Mirrors the reference's Items class plus items[] parameter block
 * (all.hpp:31-37, itemIndex.hpp:4-33) and the constants each driver hard-codes
 * (e.g. LockExchange.cu:13-15,77,223). */
typedef struct ipl_config {
    int32_t lattice;          /* 9 (D2Q9, x-z plane), 19 (D3Q19) or 27 (D3Q27) */
    int32_t nx, ny, nz;       /* cells per axis; ny == 1 for D2Q9 */
    double dz;                /* lattice spacing = width of a uniform cell (items[IDX_dz]) */
    double dt;                /* time step; c = dz / dt (basic.cu:71) */
    const double* pos[3];     /* cell-centre coordinate per axis index (length nx, ny, nz) */
    const double* width[3];   /* cell width per axis index; must be >= dz on flux axes */
    const uint8_t* solid;     /* optional nx*ny*nz mask in (x,y,z) order, 1 = excluded; NULL = none */
    int32_t periodic[3];      /* 1 = periodic axis (no walls), 0 = walls at both ends */
    int32_t wall_bc;          /* IPL_BC_SLIP or IPL_BC_BOUNCE_BACK */
    int32_t flux_axes;        /* see IPL_FLUX_* */
    int32_t stencil;          /* 5 (reference, centred) or 3 (cm, c, cp) Lagrange points */
    int32_t collision;        /* IPL_COLL_MRT or IPL_COLL_SRT */
    double s_rates[27];       /* MRT rates in moment order; stress rows use 1/tau per cell */
    double tau;               /* uniform tau when init->tau is NULL */
    double taus;              /* uniform scalar tau when init->taus is NULL */
    double rho_L, rho_H;      /* mass density of phi = 0 / phi = 1 (update_rho, collision.cu:251) */
    double beta_s;            /* density per unit salinity (0.824493 in the reference) */
    double gravity;           /* acceleration along -z (9.81 in the reference, collision.cu:78-79) */
    int32_t boussinesq;       /* 1: Fz -= S beta_s g ; 0: Fz -= rho g */
    double sigma;             /* surface tension (items[IDX_sigma]) */
    double pf_thickness;      /* phase-field interface width W (items[IDX_PFthick]) */
    int32_t scalar;           /* IPL_SCALAR_* */
    int32_t scalar_lattice;   /* 0 = same lattice as the flow; 7 = D3Q7 (3D only) */
    int32_t storage;          /* IPL_F64 / IPL_F32 / IPL_F16 (FP16 only on the fused path) */
    int32_t path;             /* IPL_PATH_* */
    double f16_scale;         /* FP16 storage: populations stored as (f - w p0) * f16_scale; 0 = auto */
    int32_t check_finite;     /* 1: ipl_step checks p and u for NaN/Inf after the last step */
    double pf_tau;            /* phase-field relaxation time; the reference hard-codes 0.58
                                 (collision.cu:218), i.e. mobility (pf_tau - 0.5) c^2 dt / 3 */
    int32_t wall_function;    /* 1: Spalding log-law wall shear on the cells next to every wall,
                                 added to F before Force (wall_function, collision.cu:28-55, called
                                 for the six faces at LockExchange.cu:214-219); fused path */
    int32_t workgroup;        /* fused path: work-items per workgroup of the step, a multiple of 64
                                 (at most 512) holding as many line slots (a line's cells rounded up
                                 to whole 64-wide wavefronts) as fit; 0 = one line per workgroup.
                                 Every size gives the same results.  No counterpart in the
                                 reference (blocks of 64 threads there, LockExchange.cu:195-196). */
    double wall_nu;           /* kinematic viscosity of the log law (items[IDX_nu]) */
    int32_t les;              /* 1: coherent-structure LES sets tau and tau_s of every cell after
                                 each step (LES, basic.cu:285-339, called after update_rho at
                                 Boegman2005run18IB.cu:339); fused path */
    int32_t line_axis;        /* fused path, grids without a stretched axis: the axis its lines (one
                                 workgroup each) run along, 1/2/3 = x/y/z; 0 = chosen by the
                                 library (the axis whose lines fill their wavefronts best).  An
                                 axis that cannot hold a line makes the fused path unavailable.
                                 No counterpart in the reference (one thread per cell there). */
    double mu_L, mu_H;        /* LES: dynamic viscosity of phi = 0 / phi = 1 (muL, muH) */
} ipl_config;

/* Initial state, canonical cell order, length = number of calc cells.  NULL entries default to
 * zero (p, u, v, w, S), to 0 (phi) or to the config scalars (tau, taus).  Populations start at
 * equilibrium: f = feq(p, u) (LockExchange.cu:133-139); g = geq(S or phi, u). */
typedef struct ipl_init {
    const double* p;
    const double* u;
    const double* v;
    const double* w;
    const double* S;
    const double* phi;
    const double* tau;
    const double* taus;
} ipl_init;

typedef struct ipl_ctx ipl_ctx;

/* Geometry, neighbour / ghost tables, boundary remap tables, MRT matrices, initial populations;
 * replaces the driver prologue: input_items + Items::setc + set_M + grid/neib build +
 * set_neibghor_wall + set_bound2D/3D + cuMallocCopy (basic.cu:57-98,171-237;
 * set_bound.cu:25-147; cuda_mem.cu:3-7; LockExchange.cu:12-189). */
int ipl_create(const ipl_config* cfg, const ipl_init* init, ipl_ctx** out);

/* Advances nsteps steps: equ_f, Force, collision, IP_process(f), scalar collision,
 * IP_process(g), update_scalar, update_rho, set_wall_rho, resetF (LockExchange.cu:220-278,
 * Zhou_soliton4B1.cu:220-277 with its stale call sites repaired). */
int ipl_step(ipl_ctx* ctx, int32_t nsteps);

/* Same as ipl_step but returns the device time of the steps (HIP events on the context's
 * stream), in milliseconds. */
int ipl_step_timed(ipl_ctx* ctx, int32_t nsteps, double* ms);

/* Runs nsteps steps with every kernel launch bracketed by HIP events on the context's stream.
 * out[0] = mean duration (ms) of one launch of the dominant kernel (the fused step kernel; on the
 * plain path the whole step), out[1] = mean per-step time (ms) of the remaining kernels,
 * out[2] = total (ms).  The reference times its whole loop with a host clock only
 * (LockExchange.cu:197,284-291). */
int ipl_profile_steps(ipl_ctx* ctx, int32_t nsteps, double out[3]);

/* Device -> host copy of a field in canonical cell order (the reference's save-interval
 * cudaMemcpy, LockExchange.cu:245-256).  bytes must equal the field size in doubles. */
int ipl_read(ipl_ctx* ctx, int32_t field, double* host, size_t count);

/* Host -> device overwrite of a field (canonical order, doubles). */
int ipl_write(ipl_ctx* ctx, int32_t field, const double* host, size_t count);

/* Volume-weighted global sums in FP64 of p, S and phi over calc cells: the conservation
 * monitor printed by output() (basic.cu:9-15). */
int ipl_global_sums(ipl_ctx* ctx, double out[3]);

/* The sums of ipl_global_sums without a device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The partial sums go to a file-scope device array of
 * 3 x 1024 + 3 doubles (24,600 bytes) that every GPU the library runs on holds until the process
 * ends; the reduction grid is fixed, so the sums of any context fit it, and a reduction that would
 * need more is refused with IPL_ERR_ARGUMENT.  The context's device must be current.  Every call
 * on one GPU shares that GPU's array (its contents are the last call's): calls that overlap on
 * one GPU, from different threads or contexts, race.  Same kernels and order as ipl_global_sums,
 * so the two return the same bits. */
int ipl_global_sums_noalloc(ipl_ctx* ctx, double out[3]);

/* Why the context runs on the plain path: the fused path's refusal under IPL_PATH_AUTO (a cell mask,
 * the two-axis corner scheme, lines over 512 cells, a non-uniform tau, ...), or that the plain path
 * was requested; "" on the fused path.  The string lives as long as the context. */
const char* ipl_path_reason(ipl_ctx* ctx);

/* Number of calc cells, ghost cells and population storage bytes of a context. */
int ipl_info(ipl_ctx* ctx, int64_t* n_calc, int64_t* n_ghost, int64_t* device_bytes, int32_t* path);

int ipl_destroy(ipl_ctx* ctx);

const char* ipl_error_string(int code);
const char* ipl_last_error_message(ipl_ctx* ctx);

/* ---- stage-wise entry points (parity harness; plain path only) ----------------------------- */
enum {
    IPL_STAGE_EQU_F = 0,        /* collision.cu:3-14 */
    IPL_STAGE_FORCE = 1,        /* collision.cu:57-112 */
    IPL_STAGE_COLLIDE = 2,      /* col_f_MRT / col_f_SRT, collision.cu:115-165 */
    IPL_STAGE_IP_F = 3,         /* IP_process on f, adv.cu:245-257 */
    IPL_STAGE_SCALAR_COLLIDE = 4, /* col_g_reg / col_PF, collision.cu:167-223 */
    IPL_STAGE_IP_G = 5,         /* IP_process on g */
    IPL_STAGE_UPDATE_SCALAR = 6,/* collision.cu:225-236 */
    IPL_STAGE_UPDATE_RHO = 7,   /* collision.cu:238-257 */
    IPL_STAGE_SET_WALL_RHO = 8, /* basic.cu:254-261 (rho, and phi with a phase field) */
    IPL_STAGE_RESET_F = 9,      /* resetF, collision.cu:258-264 */
    IPL_STAGE_EQU_G = 10        /* equ_g, collision.cu:16-27: the scalar populations of every calc
                                   cell <- the equilibrium of its scalar (S, or phi with a phase
                                   field) and u; linear on D3Q7.  Not part of the step (no driver
                                   calls it); the reference writes its feq argument, here g */
};
int ipl_stage(ipl_ctx* ctx, int32_t stage);

/* ---- immersed boundary: one rigid body carried by Lagrangian markers (function/ibm.cu) -------
 * Replaces the drivers' marker set-up (SphereFall3D/SphereFall.cu:197-261) and adds the IB
 * sub-step to every ipl_step, in the drivers' order (SphereFall.cu:372-383, after update_rho and
 * resetF): marker forces and spreading, update_velIBM, update_IBbody, update_IBpoint,
 * search_IBlattice. */
enum ipl_ib_method {
    IPL_IB_STRESS = 0,  /* stress-based wall force of Suzuki & Yoshino, get_IBMGw2 (ibm.cu:124-274) */
    IPL_IB_DIRECT = 1   /* direct forcing, IB_directForcing (ibm.cu:276-344); iterations > 1 repeats
                           interpolate-force-spread with corrected node velocities (multi-direct forcing) */
};
enum ipl_ib_kernel {
    IPL_IB_KERNEL_REFERENCE = 0, /* the reference's three nodes per axis: STRESS weights x (and z in 2D)
                                    by the truncated 4-point Peskin function weightFunction (ibm.cu:61-66),
                                    y and z (3D) by 3-point Lagrange; DIRECT uses 3-point Lagrange */
    IPL_IB_KERNEL_PESKIN4 = 1    /* DIRECT only: the full 4-point Peskin support on every axis */
};
typedef struct ipl_ib_config {
    int32_t method;               /* IPL_IB_STRESS or IPL_IB_DIRECT */
    int32_t kernel;               /* IPL_IB_KERNEL_* */
    int32_t iterations;           /* DIRECT: forcing passes per step (1 = IB_directForcing) */
    int32_t moving;               /* 1: update_IBbody integrates the body; 0: fixed body (forces still summed) */
    int32_t n_markers;
    const double* marker_pos;     /* 3*n offsets from the body centre in the body frame (oposw) */
    const double* marker_normal;  /* 3*n outward unit normals in the body frame (onB_vec) */
    double spacing;               /* marker spacing dIBM (items[IDX_dIBM], SphereFall.cu:277) */
    double rho_f;                 /* fluid density: rhof of get_IBMGw2 / update_IBbody; the weight of the
                                     body force of direct forcing (the literal 1000 of ibm.cu:342) */
    double density;               /* body density densB */
    double mass;                  /* body mass massB */
    double inertia[3];            /* principal moments of inertia (IB) */
    double gravity[3];            /* acceleration in the buoyancy term of update_IBbody: the reference
                                     adds (1 - rhof/densB) m 9.81 along +z (ibm.cu:391) */
    double pos[3], vel[3];        /* initial body centre and velocity */
    double omega[3];              /* initial angular velocity, body frame (angleVB) */
    double quat[4];               /* initial orientation quaternion (q0, q1, q2, q3) */
    int32_t spm;                  /* smoothed-profile forcing of the body interior before the marker
                                     forcing, f += w dt 3 (e c . F) / c^2 with F = profile (v - u) / dt:
                                     1 = SPM_ellipse3D (SphereFall3D/SphereFall.cu:13-30, called at
                                     :376; stress method): ellipsoid a, a, b in the body frame;
                                     2 = SPM_ellipse (ibm.cu:89-122): ellipse a, b over the body-frame
                                     x and z (x and y on a 3D lattice), the reference's rotation as
                                     written (untransposed), sine ramp of width dz/a;
                                     3 = SPM (ibm.cu:68-87): circle of radius a about the body centre in
                                     the x-z plane (x-y on a 3D lattice), ramp over +-dz/2 about
                                     a - dz/2, v = the velocity of marker 0, F over dz.
                                     The library's two forms are commented out in every driver. */
    int32_t reserved0;
    double spm_a, spm_b;          /* SPM semi-axes (form 1: body frame x, y: a; z: b; form 3: radius a) */
    int32_t reserved[2];
} ipl_ib_config;

/* The immersed-boundary kernels of the drivers' loop one at a time, for stage-wise parity
 * (SphereFall.cu:372-383).  IPL_IB_STAGE_FLOW runs one step without its IB sub-step (every kernel
 * before get_IBMGw2; with an open boundary its reset comes at the end of this stage); the other
 * stages then act on the current state in the drivers' order: SPM (the body's smoothed-profile
 * forcing, SPM_ellipse3D, SphereFall.cu:13-30), FORCE (get_IBMGw2, ibm.cu:124-274, or
 * IB_directForcing, ibm.cu:276-344, repeated for multi-direct forcing, spread onto the grid),
 * UPDATE_VEL (update_velIBM, ibm.cu:346-365, with the force FORCE spread), BODY (update_IBbody,
 * ibm.cu:367-424) and POINTS (update_IBpoint + search_IBlattice, ibm.cu:426-463, of a moving
 * body).  FLOW then SPM (if configured), FORCE, UPDATE_VEL, BODY, POINTS is one ipl_step.  Fused
 * path, body attached. */
enum {
    IPL_IB_STAGE_FLOW = 0, IPL_IB_STAGE_SPM = 1, IPL_IB_STAGE_FORCE = 2, IPL_IB_STAGE_UPDATE_VEL = 3,
    IPL_IB_STAGE_BODY = 4, IPL_IB_STAGE_POINTS = 5
};
int ipl_ib_stage(ipl_ctx* ctx, int32_t stage);

/* Attaches the body to a context (once, before the first step); nodes of a marker must lie inside
 * the box: the reference's ghost positions make the 3-point Lagrange weights singular there. */
int ipl_ib_attach(ipl_ctx* ctx, const ipl_ib_config* ib);
/* FP64 body state after the last step: pos[3] vel[3] omega[3] quat[4] force[3] torque[3] (19 values;
 * force = sum of marker forces + buoyancy, torque about the centre, as in update_IBbody). */
int ipl_ib_body_state(ipl_ctx* ctx, double out[19]);
/* Marker arrays (IB_csv, ibm.cu:3-14): positions, velocities, forces Gw, normals (3 per marker) or
 * the nearest calc cell of each marker (1 per marker, canonical index). */
enum ipl_ib_array { IPL_IB_POS = 0, IPL_IB_VEL = 1, IPL_IB_FORCE = 2, IPL_IB_NORMAL = 3, IPL_IB_NODE = 4 };
int ipl_ib_read_markers(ipl_ctx* ctx, int32_t which, double* out, size_t count);

/* ---- batches: many independent simulations advanced together -----------------------------
 * A batch holds contexts (jobs) of any sizes and configurations.  With batched launches (the
 * default) the members whose step is the fused kernel alone are grouped by kernel variant
 * (storage, lattices, stencil, line axis, line length, LDS) and every group advances by one
 * launch per step whose workgroups cover the lines of all its jobs (a job table on the device;
 * a run of at least 20 steps of one group replays a captured HIP graph of 20 launches, kept for
 * later runs of the same shape).  Members with more in their step (an immersed body, LES, the wall function, an open
 * boundary, a 3D phase field) step alone on their own streams (graph replay), overlapping with
 * the batched launches.  Members keep their own buffers: a job's results are the same alone and
 * in any mix.  The reference runs one case per process and has no counterpart; batches are what a
 * job server steps. */
enum { IPL_BATCH_LAUNCH_PER_MEMBER = 0, IPL_BATCH_LAUNCH_BATCHED = 1 };
typedef struct ipl_batch ipl_batch;
int ipl_batch_create(ipl_batch** out);
/* A member shares the device with the others (ipl_set_shared(ctx, 1) until it is removed). */
int ipl_batch_add(ipl_batch* batch, ipl_ctx* ctx);
int ipl_batch_remove(ipl_batch* batch, ipl_ctx* ctx);
/* Advances member i by nsteps[i] steps (members in the order they were added), or every member by
 * nsteps[0] when all == 1. */
int ipl_batch_step(ipl_batch* batch, const int32_t* nsteps, int32_t all);
int ipl_batch_destroy(ipl_batch* batch);       /* the members stay valid */
/* IPL_BATCH_LAUNCH_BATCHED (default) or IPL_BATCH_LAUNCH_PER_MEMBER (every member on its own stream). */
int ipl_batch_set_launch(ipl_batch* batch, int32_t launch);
/* Members whose step has more than max_lines lines (workgroups) step on their own streams even
 * with batched launches: their launches fill the device by themselves.  0: no limit. */
int ipl_batch_set_max_lines(ipl_batch* batch, int64_t max_lines);
/* Totals since ipl_batch_create (diagnostics): out[0] batched kernel launches (those inside
 * graph replays included), out[1] member-steps they advanced, out[2] member-steps launched on the
 * members' own streams, out[3] graphs the batched launches captured. */
int ipl_batch_counters(ipl_batch* batch, int64_t out[4]);
const char* ipl_batch_last_error_message(ipl_batch* batch);

/* Device bytes ipl_create would allocate for cfg (population, force and field buffers, tables and
 * staging) on the path it would take (the fused path's own checks decide, as in ipl_create), for
 * admission against a memory budget; no device work. */
int ipl_estimate_device_bytes(const ipl_config* cfg, int64_t* bytes);
/* Device bytes ipl_ib_attach would add for the body ib on the grid of cfg (markers, node-block
 * box and body state); no device work. */
int ipl_estimate_ib_bytes(const ipl_config* cfg, const ipl_ib_config* ib, int64_t* bytes);
/* 0: launch every step directly; 1 (default): replay captured graphs where the path supports it. */
int ipl_set_graph(ipl_ctx* ctx, int32_t on);
/* Captures the step graphs of the fused path now (one per ping-pong buffer parity; recorded, not
 * run) instead of at the first ipl_step that uses them, so that later steps, e.g. a timed region
 * or a serving chunk, pay no capture.  A context captures each graph once unless its set-up
 * changes (immersed body attached, open boundary, dt change).  No-op on the plain path.  The
 * reference launches every kernel directly and has no counterpart (LockExchange.cu:220-278). */
int ipl_prepare(ipl_ctx* ctx);
/* Number of graph captures the context has made so far (diagnostics). */
int ipl_graph_captures(ipl_ctx* ctx, int32_t* count);
/* on = 1: the context shares the device with other contexts stepping at the same time (a job
 * server's jobs), so an immersed body's step tail (body update, marker points, next ranking) runs
 * on the context's own stream after the forcing instead of on a second stream beside the next
 * step, whose fork and join cost more than they hide beside other contexts' work; on = 0 (the
 * default): the second stream.  ipl_batch_add sets it, ipl_batch_remove clears it.  Set it before
 * ipl_prepare: a change drops the captured step graphs.  The reference steps one case per process
 * and has no counterpart. */
int ipl_set_shared(ipl_ctx* ctx, int32_t on);

/* Open (inflow / outflow) boundary, replacing set_velocity1 + set_wall_boundary1
 * (Karman_vortex_2D/Karman_vortexxz.cu:38-57, called at :391-392 and :425-430): at the end of
 * every step the listed cells (canonical indices) are reset to the equilibrium of the given
 * p, u, v, w (per cell), and the velocity of later steps starts from those values.  The driver
 * fixes the inlet velocity with set_velocity1 and re-imposes each open cell's own p, u, which
 * stay at their initial values while no force acts there; this call takes those values directly.
 * n = 0 removes the boundary.  Replaces an earlier list. */
int ipl_set_open_boundary(ipl_ctx* ctx, const int32_t* cells, int64_t n, const double* p, const double* u,
                          const double* v, const double* w);

/* Time-step change (reset_items_dt + reset_pressure, basic.cu:100-110 and 262-273; the driver's
 * high-resolution start at LockExchange.cu:201-211): dt /= alpha, c *= alpha, p /= alpha^2 and
 * every cell's f reset to feq(p, u) with the new c; tau is kept.  Fused path, no immersed body. */
int ipl_change_dt(ipl_ctx* ctx, double alpha);

/* ---- file writers (host side, after a readout of the fields) -------------------------------- */
/* output() (basic.cu:3-30): "num_calc,N", the 15-column header and one "%f" row per calc cell
 * (x, y, z, delx, dely, pressure, salinity, phi, rho, u, v, w, Fx, Fy, Fz), the format the
 * reference's post-processing reads (data/Fakhari_LBM%04d.csv). */
int ipl_write_fields_csv(ipl_ctx* ctx, const char* path);
/* IB_csv (ibm.cu:3-14): "x,y,z,velx,vely,velz,ax,ay,az" and one row per marker (IB_point/ibm_%04d.csv). */
int ipl_write_markers_csv(ipl_ctx* ctx, const char* path);
/* out_C_D (SphereFall3D/SphereFall.cu:32-42, Karman_vortex_2D/Karman_vortex2D.cu:81-91): the
 * "C_D,C_time" header and one row per record, the two columns in the C++ stream's default format
 * (6 significant digits).  The drivers record a drag coefficient (Karman_vortex2D.cu:378-379) or
 * the body's settling velocity (SphereFall.cu:424) at every save interval; n = 0 writes the header. */
int ipl_write_cd_csv(const char* path, const double* c_d, const double* time, int64_t n);
/* The VTU of LockExchange/make_VTK.py: one VTK_HEXAHEDRON per calc cell (corners at the cell
 * faces, shared between neighbours), cell data salinity, phi, rho, pressure, velocity, force;
 * ASCII, written directly instead of through a CSV and pyvista. */
int ipl_write_vtu(ipl_ctx* ctx, const char* path);

/* ---- host-only set-up helpers (no device work) --------------------------------------------- */
/* hydrostatic_pressure (basic.cu:239-252): the initial p of every calc cell (canonical order),
 * p = 3 g (nz dz / 2 - z) / c^2 / 1000 with boussinesq != 0, else 3 g (center_z - z) / c^2,
 * g = 9.81.  The reference's own f from that function is overwritten by its only caller with
 * feq(p, u) (Zhou_soliton4B2.cu:137-145); pass p as the initial field to get the same start. */
int ipl_hydrostatic_pressure(const ipl_config* cfg, int32_t boussinesq, double center_z, double* p_out, int64_t count);
/* set_walln (set_bound.cu:8-23): the calc cells (canonical indices, ascending) whose neighbour in
 * the face's direction is a ghost: face 1 = +x, 2 = +y, 3 = -x, 4 = -y, 5 = +z, 6 = -z (D2Q9:
 * faces 2 and 4 are empty).  *count holds the capacity of out on entry and the number of cells
 * on return; with out == NULL only the number is returned. */
int ipl_wall_cells(const ipl_config* cfg, int32_t face, int32_t* out, int64_t* count);

/* ---- host-only diagnostics ------------------------------------------------------------------
 * Builds the cell graph of cfg on the host (no device work) and copies one table out:
 * which = 0 neighbour table [k][n_calc + n_ghost] (ghosts after calc cells, basic.cu:203-237),
 * 1 / 2 = slip remap nextB / nextK [k][n_ghost] (set_bound.cu:25-147), 3 = stencil weights
 * [(axis*2 + sign)*npts + j][n_calc].  With out == NULL returns the table length in *count. */
int ipl_build_tables(const ipl_config* cfg, int32_t which, void* out, int64_t* count);

#ifdef __cplusplus
}
#endif
#endif /* IPL_H */
