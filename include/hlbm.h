/*
 * hlbm.h -- C ABI of the MI355X (gfx950) D3Q19 lattice Boltzmann engine.
 *
 * This is the drop-in boundary for hippoLBM's hot path.  In the reference each
 * hot-path step is an Onika operator (C++ class + ADD_SLOT + execute()) that
 * mutates LBMFields<19> in place; this library replaces the bodies of those
 * operators.  A host-side operator layer (hippolbm_amd/host/operators.cpp, or
 * an Onika plugin) keeps the reference's operator names and slots and calls
 * the entry points below.
 *
 * Conventions
 *   - every function returns HLBM_OK (0) or a negative status; no exception
 *     crosses the ABI; hlbm_last_error() returns a message for the calling
 *     thread's last failure;
 *   - one host thread per domain handle at a time (thread-compatible);
 *   - `stream` arguments are hipStream_t passed as void* (NULL = the legacy
 *     default stream);
 *   - host pointers are plain C arrays; the populations stay resident in HBM
 *     between calls;
 *   - "node layout": Real nodes only, x fastest, f[i*N + (z*ny + y)*nx + x],
 *     N = nx*ny*nz;  "reference layout": hippoLBM's LBMFields<19> layout over
 *     the All box with two ghost layers per side,
 *     f[i*N_all + (z*Ly + y)*Lx + x], L = n + 4
 *     (src/grid/include/hippoLBM/grid/field_view.hpp:56-72,
 *      include/hippoLBM/core/box3d.hpp:127-132).
 */
#ifndef HLBM_H
#define HLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define HLBM_OK 0
#define HLBM_ERR_ARG (-1)         /* invalid argument                           */
#define HLBM_ERR_HIP (-2)         /* a HIP runtime call failed                  */
#define HLBM_ERR_UNSUPPORTED (-3) /* operator combination the engine rejects    */
#define HLBM_ERR_NOMEM (-4)       /* device allocation failed                   */
#define HLBM_ERR_NODEVICE (-5)    /* no HIP device visible                      */

/* ---- enums ------------------------------------------------------------ */
enum hlbm_storage { HLBM_F64 = 0, HLBM_F32 = 1, HLBM_F16 = 2, HLBM_BF16 = 3 };
enum hlbm_pattern { HLBM_PATTERN_AB = 0, HLBM_PATTERN_AA = 1 };
enum hlbm_collision { HLBM_BGK = 0, HLBM_MRT = 1 };
enum hlbm_layout { HLBM_LAYOUT_NODES = 0, HLBM_LAYOUT_REF = 1 };

/* boundary operators, in the reference's operator-name vocabulary */
enum hlbm_bc_kind {
  HLBM_BC_PRE_BOUNCE_BACK = 1,   /* "pre_bounce_back"   src/bcs/pre_bounce_back.cu:89-118   */
  HLBM_BC_POST_BOUNCE_BACK = 2,  /* "post_bounce_back"  src/bcs/post_bounce_back.cu:101-123 */
  HLBM_BC_WALL_BOUNCE_BACK = 3,  /* "wall_bounce_back"  src/bcs/wall_bounce_back.cu:58-73   */
  HLBM_BC_CAVITY_Z_0 = 4,        /* "cavity_z_0"        src/bcs/cavity.cu:66-99             */
  HLBM_BC_CAVITY_Z_L = 5,        /* "cavity_z_l"        src/bcs/cavity.cu:66-99             */
  HLBM_BC_LID_DRIVEN_CAVITY = 6, /* "lid_driven_cavity" src/bcs/lid_driven_cavity.cu:69-93  */
  HLBM_BC_NEUMANN = 7,           /* "neumann"           src/bcs/neuman.cu:69-122            */
  /* Open-boundary operators with no counterpart in the reference (the porous
   * channel's inlet and outlet); post_stream_bcs or boundary_conditions only,
   * on the planes of a non-periodic axis.
   *   pressure_inlet: Zou/He density boundary, U[0] = prescribed density,
   *     zero tangential velocity (the neumann formulas with u_n from rho).
   *   outflow: zero-gradient outlet, the populations entering from outside
   *     copy their post-stream values at the first interior node. */
  HLBM_BC_PRESSURE_INLET = 8,    /* "pressure_inlet"                                        */
  HLBM_BC_OUTFLOW = 9,           /* "outflow"                                               */
  /* set_distribution in start_iteration (src/grid/set_distribution.cu:69-109,
   * example/ball.msp:23-26): at the start of every step, f_i = U[0] * w_i on the
   * nodes of the box `bounds` (plane = 1) or on every node (plane = 0); stage
   * HLBM_STAGE_START only, at most 4 entries, applied in list order. */
  HLBM_BC_SET_DISTRIBUTION = 10  /* "set_distribution"                                      */
};
/* operator lists of the compute loop (data/config/config_compute_loop.msp:23-38):
 * start_iteration runs before macro_variables and the collision */
enum hlbm_bc_stage {
  HLBM_STAGE_START = -1, HLBM_STAGE_PRE_STREAM = 0, HLBM_STAGE_POST_STREAM = 1, HLBM_STAGE_BOUNDARY = 2
};
/* plane traversals (src/grid/include/hippoLBM/grid/grid_region.hpp:75-77) */
enum hlbm_plane {
  HLBM_PLAN_XY_0 = 0, HLBM_PLAN_XY_L = 1, HLBM_PLAN_XZ_0 = 2,
  HLBM_PLAN_XZ_L = 3, HLBM_PLAN_YZ_0 = 4, HLBM_PLAN_YZ_L = 5
};

/* hlbm_step flags */
#define HLBM_STEP_WRITE_MACRO 1u /* the last kernel of the call also writes rho/u (macro_variables) */

/* ---- descriptors ------------------------------------------------------ */
typedef struct hlbm_domain hlbm_domain;

typedef struct {
  int32_t nodes[3];    /* grid nodes per axis = cell_dims + 1 (src/grid/domain_lbm.cpp:68)     */
  int32_t periodic[3]; /* periodic axes (domain slot `periodic`)                               */
  int32_t storage;     /* enum hlbm_storage; HLBM_F64 is the bit-faithful drop-in              */
  int32_t pattern;     /* enum hlbm_pattern                                                    */
  int32_t collision;   /* enum hlbm_collision                                                  */
  int32_t mrt_custom;  /* 0: hippoLBM's hard-coded MRT rates (mrt.hpp:43-47); 1: mrt_rates    */
  double tau;          /* LBMParameters::tau_; must be > 0.5 (nu = (tau - 0.5)/3 > 0)          */
  double fext[3];      /* LBMParameters::Fext_ (LBM units)                                     */
  double dx;           /* grid spacing, used by obstacle rasterization                         */
  double mrt_rates[10];/* s1, s2, s4, s9, s10, s13, s16, weps, wepsj, wxx                     */
  /* FP32/FP16/BF16 storage holds f_i minus the collision's rest state: w_i for
   * BGK, w_i + (weps - 3) * {1/21, -1/63, 1/252} (rest/axis/diagonal) for MRT,
   * whose rates relax the epsilon moment at rest to weps (mrt.hpp:96-100). */
  int32_t round_nearest;/* FP16/BF16 storage: 0 = stochastic rounding of every stored
                           population (the default: unbiased, so rounding errors do not
                           accumulate in steady flow), 1 = round to nearest even          */
  int32_t pad;
} hlbm_domain_desc;

typedef struct {
  int32_t kind;  /* enum hlbm_bc_kind                                                           */
  int32_t stage; /* enum hlbm_bc_stage: which list of the compute loop the operator sits in    */
  int32_t plane; /* enum hlbm_plane, for lid_driven_cavity / neumann (one entry per region)    */
  int32_t pad;
  double U[3];   /* velocity in LBM units for lid/neumann (already divided by celerity);
                    raw operator value for cavity_z_* (hippoLBM applies no conversion there) */
  double bounds[6]; /* set_distribution: {xmin,ymin,zmin,xmax,ymax,zmax} (physical units)    */
} hlbm_bc;

typedef struct {
  int32_t kind; /* 0 = ball (register_solid_ball), 1 = wall (register_solid_wall),
                   2 = quadric (register_quadrics)                                             */
  int32_t pad;
  double p0[3]; /* ball: center; wall: bmin                                                    */
  double p1[3]; /* wall: bmax                                                                  */
  double radius;
  double q[16]; /* quadric: 4x4 matrix Q, row-major, transform already applied
                   (register_quadrics.cpp:72-84); solid where [p,1] Q [p,1]^T <= 0            */
} hlbm_obstacle;

/* ---- library ---------------------------------------------------------- */
const char* hlbm_last_error(void);
const char* hlbm_version(void);
/* Number of visible HIP devices (0 on a host without GPU); never allocates. */
int hlbm_device_count(int32_t* count);

/* ---- domain lifetime: replaces `domain` + `build_traversal` + `define_grid_3dq19`
 *      (src/grid/domain_lbm.cpp:66-110, src/grid/build_traversal.cpp:58-62,
 *       src/grid/define_lbm_fields.cpp:58-86) for one single-rank subdomain.
 *      Populations start at zero, flags at fluid, rho/u at zero. */
int hlbm_domain_create(const hlbm_domain_desc* desc, hlbm_domain** out);
int hlbm_domain_destroy(hlbm_domain* h);
/* Device bytes held by the domain (populations + flags + macro + scratch). */
int hlbm_domain_bytes(const hlbm_domain* h, uint64_t* bytes);

/* ---- boundary-condition plan: the pre_stream_bcs / post_stream_bcs /
 *      boundary_conditions operator lists, in list order.  Replaces the
 *      execute() bodies of pre_bounce_back, post_bounce_back, wall_bounce_back,
 *      cavity_z_0/l, lid_driven_cavity and neumann.  Call again to change it. */
int hlbm_set_bcs(hlbm_domain* h, const hlbm_bc* list, int32_t n);
/* Change tau / Fext (lbm_parameters) without recreating the domain. */
int hlbm_set_params(hlbm_domain* h, double tau, const double fext[3]);
/* Select the collision operator the `collision` slot names: HLBM_BGK
 * (src/collision/bgk.cu:77-94) or HLBM_MRT (src/collision/mrt.cu:71-89);
 * mrt_rates = NULL keeps the reference's hard-coded rates (mrt.hpp:44-47),
 * otherwise [s1 s2 s4 s9 s10 s13 s16 weps wepsj wxx].  A pending streaming is
 * applied first; shifted storage moves to the new rest state. */
int hlbm_set_collision(hlbm_domain* h, int32_t collision, const double* mrt_rates);

/* ---- initial state ----------------------------------------------------- */
/* init_obst: every node fluid (src/grid/init_obst.cu:54-60). */
int hlbm_init_obst(hlbm_domain* h);
/* update_obstacles for a list of balls/walls/quadrics in obstacle-id order,
 * rasterized on the device with the reference's rules
 * (src/obstacle/update_obstacles.cu:32-71, obstacle.hpp:13-17,61-64,97,136-138),
 * including the first ghost layer of the periodic axes at its own positions:
 * with wall_bounce_back, a wall across a periodic face bounces off that ghost
 * as in the reference (src/bcs/include/hippoLBM/bcs/bounce_back.hpp:163-185).
 * A wall ghost over a fluid periodic image next to a wall node makes
 * hlbm_step fail with HLBM_ERR_UNSUPPORTED. */
int hlbm_add_obstacles(hlbm_domain* h, const hlbm_obstacle* list, int32_t n);
/* Obstacle field as int32 -1 (fluid) / -2 (wall), node or reference layout.
 * set, reference layout: the first ghost layer of the periodic axes is taken
 * as given; node layout: those ghosts take their periodic images' flags.
 * get writes the Real nodes only. */
int hlbm_set_flags(hlbm_domain* h, const int32_t* obst, int32_t layout);
int hlbm_get_flags(hlbm_domain* h, int32_t* obst, int32_t layout);
/* set_distribution: f_i = value*w_i everywhere (bounds == NULL) or on the
 * nodes of the box bounds = {xmin,ymin,zmin,xmax,ymax,zmax}
 * (src/grid/set_distribution.cu:69-109). */
int hlbm_set_distribution(hlbm_domain* h, double value, const double* bounds);
/* init_double_couette (src/prepro/init_double_couette.cu:80-120,
 * src/prepro/include/hippoLBM/prepro/double_couette.hpp:26-45): every node
 * gets the equilibrium at rho = 1 of u = U - dU * k, k the node coordinate
 * along dim (0 x, 1 y, 2 z) and dU = U / (0.5 * (n_dim - 1)); U in LBM units
 * (the operator's U divided by celerity). */
int hlbm_init_double_couette(hlbm_domain* h, int32_t dim, const double U[3]);
/* Populations in FP64 (converted from the storage type; the shifted storage
 * adds w_i back).  get writes the current state f^T as the reference holds it
 * after T time steps (post-stream, post-BC); in reference layout the periodic
 * ghost layers are filled as update_ghost would, other ghost entries are left
 * untouched. */
int hlbm_set_populations_f64(hlbm_domain* h, const double* f, int32_t layout);
int hlbm_get_populations_f64(hlbm_domain* h, double* f, int32_t layout);

/* ---- time stepping ------------------------------------------------------
 * Advance nsteps reference time steps: macro_variables + collision +
 * pre_stream_bcs + streaming (+ update_ghost) + post_stream_bcs +
 * boundary_conditions (data/config/config_compute_loop.msp:28-36), fused into
 * one pull kernel per step.  Asynchronous on `stream`. */
int hlbm_step(hlbm_domain* h, int32_t nsteps, uint32_t flags, void* stream);
/* Steps taken since creation / the last set of populations. */
int hlbm_timestep(const hlbm_domain* h, int64_t* t);
int hlbm_synchronize(hlbm_domain* h);

/* ---- macroscopic fields and diagnostics -------------------------------- */
/* macro_variables of the current state (src/collision/macro_variables.cu:71-90):
 * fills the domain's rho/u fields; copies them out when rho/u are non-NULL
 * (rho: N values, u: 3*N SoA, node or reference layout). */
int hlbm_macro(hlbm_domain* h, double* rho, double* u, int32_t layout);
/* Copy the domain's rho/u fields as they are (the m0_/m1_ slots). */
int hlbm_get_macro(hlbm_domain* h, double* rho, double* u, int32_t layout);
/* simulation_state over the rho/u fields (src/io/simulation_state.cu:73-98),
 * FP64 on the device: out[0] = sum rho, out[1] = min |u|, out[2] = max |u|,
 * out[3..5] = sum rho*u (momentum), out[6] = number of fluid nodes. */
int hlbm_reduce_stats(hlbm_domain* h, double out[7]);

/* plane_velocity_profile (src/check/plane_velocity_profile.cu:81-101,175-205):
 * for each plane perpendicular to dim (0 x, 1 y, 2 z), over the fluid nodes
 * of the domain's u field as it is held (the m1_ slot), v = |scale*u|:
 * sum[p], count[p], vmin[p], vmax[p], one entry per node plane (nodes[dim]
 * entries).  A plane without fluid gives count 0, vmin DBL_MAX, vmax -DBL_MAX,
 * as the reference's reset values.  Sums are folded on the device in a fixed
 * order (deterministic); the reference adds with atomics. */
int hlbm_plane_profile(hlbm_domain* h, int32_t dim, double scale, double* sum, int32_t* count, double* vmin,
                       double* vmax);
/* The same statistics without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Asynchronous on `stream`; writes nodes[dim] records into the caller's DEVICE
 * array `out`, bit for bit what hlbm_plane_profile returns.  The per-column
 * partials live in a library-owned static device array of
 * HLBM_PROFILE_SCRATCH_COLUMNS records (32 bytes each, 16 MiB), one per GPU for
 * the life of the process: a domain with more than that many columns,
 * nodes[0] * (dim == 1 ? nodes[1] : nodes[2]), is refused with HLBM_ERR_ARG.
 * Every call on a GPU shares that array, so calls that overlap in time on one
 * GPU (several streams or host threads) race; calls on one stream are ordered. */
#define HLBM_PROFILE_SCRATCH_COLUMNS (1 << 19)
typedef struct {
  double sum, vmin, vmax;
  int32_t count, pad;
} hlbm_plane_stat;
int hlbm_plane_profile_noalloc(hlbm_domain* h, int32_t dim, double scale, hlbm_plane_stat* out, void* stream);

/* plot_line_velocity (src/check/plot_over_line.cu:77-90,163-172): scale*u of
 * the domain's u field at the n nodes start + k*e_axis (node coordinates,
 * 0-based), written u[3k + c]; nodes outside the domain give 0. */
int hlbm_line_velocity(hlbm_domain* h, const int32_t start[3], int32_t axis, int32_t n, double scale, double* u);

/* update_ghost (src/grid/update_ghost.cu:66-73): on one GPU there is no halo;
 * this fills the periodic ghost layers of a reference-layout FP64 population
 * array from its Real nodes, on the host array given.  It exists so an Onika
 * deck that calls update_ghost explicitly keeps working. */
int hlbm_ghost_update_ref(const hlbm_domain* h, double* f_ref);

/* ---- batched domains and the solve server -------------------------------
 * Many independent domains of uneven sizes advance together: one kernel launch
 * per collision model and time step covers a tile table of every running
 * domain (one workgroup per 1024-node tile), with per-domain tau / force /
 * boundary tables.  Chunks of `chunk_steps` steps are captured as a HIP graph
 * and replayed (captured again when the tile count outgrows the launch).
 * Requests ("solves") are submitted at any time; each hlbm_batch_run() call
 * admits queued solves whose populations fit the HBM budget, runs one chunk,
 * and retires finished solves (their rho/u are copied to the host and their
 * device memory is released).  This is the serving counterpart of running
 * hippoLBM decks one after another; it has no operator of its own in the
 * reference.  Boundary conditions per solve: periodic or closed faces
 * (pre_bounce_back + post_bounce_back on every closed face), optional moving
 * walls (cavity_z_0 / cavity_z_l), obstacles with wall_bounce_back, body force,
 * BGK or MRT. */
typedef struct hlbm_batch hlbm_batch;

typedef struct {
  int32_t storage;      /* enum hlbm_storage, one per batch                         */
  int32_t chunk_steps;  /* steps per hlbm_batch_run call (graph length)             */
  int32_t use_graph;    /* 1: replay a captured HIP graph per chunk                 */
  int32_t round_nearest;/* as hlbm_domain_desc.round_nearest                        */
  uint64_t budget_bytes;/* HBM for populations + flags of all resident solves       */
} hlbm_batch_desc;

typedef struct {
  int32_t nodes[3];
  int32_t periodic[3];
  int32_t collision;    /* enum hlbm_collision                                       */
  int32_t steps;        /* time steps to run                                         */
  int32_t wall_bb;      /* wall_bounce_back on the obstacles                         */
  int32_t n_obstacles;
  int32_t cavity_on[2]; /* cavity_z_0 / cavity_z_l moving walls (z must be closed)   */
  int32_t keep_populations; /* also return the final populations (FP64, node layout) */
  int32_t mrt_custom;
  double tau;           /* > 0.5                                                     */
  double fext[3];
  double dx;
  double init_rho;      /* initial state f_i = init_rho * w_i                        */
  double cavity_u[2][3];
  double mrt_rates[10];
  const hlbm_obstacle* obstacles; /* copied at submit                                */
  const double* init_f; /* optional initial populations, FP64 node layout SoA        *
                         * (f[i*N + (z*ny + y)*nx + x]); NULL = init_rho * w_i;      *
                         * copied at submit                                          */
} hlbm_solve_desc;

/* A solve ends DONE, or FAILED when its final state holds a NaN or an infinity:
 * at retirement the simulation_state reduction (src/io/simulation_state.cu:73-98,
 * FP64: sum rho, min/max |u|, momentum) runs over the solve alone, and a
 * non-finite sum marks that solve only; the other solves of the batch are not
 * affected.  A FAILED solve's fields can still be read with hlbm_batch_result. */
enum hlbm_solve_state { HLBM_SOLVE_QUEUED = 0, HLBM_SOLVE_RUNNING = 1, HLBM_SOLVE_DONE = 2, HLBM_SOLVE_FAILED = 3 };

typedef struct {
  int64_t queued, running, done, chunks;
  uint64_t bytes_in_use, budget_bytes;
  double node_updates; /* lattice node updates performed so far                     */
  double gpu_ms;       /* device time of all chunks (HIP events)                    */
  uint64_t peak_bytes_in_use; /* the largest bytes_in_use after any admission        */
  int64_t relocated;   /* arena pairs replaced by the placement check at create     */
} hlbm_batch_info;

/* Arenas of 8 GiB and more are placement-checked at create: the streaming
 * pattern of a step is timed on them and on up to three fresh arena pairs (held
 * together while the check runs, each when its size plus 4 GiB is free), and the
 * fastest pair is kept (HLBM_PLACEMENT=0 turns it off).  The batch also holds a
 * pinned host pool of min(2 GiB, 32 bytes per budgeted node) into which a
 * retiring solve's rho, u and statistics are copied on the compute stream. */
int hlbm_batch_create(const hlbm_batch_desc* desc, hlbm_batch** out);
int hlbm_batch_destroy(hlbm_batch* b);
/* Queue a solve; fails with HLBM_ERR_NOMEM if it can never fit the budget. */
int hlbm_batch_submit(hlbm_batch* b, const hlbm_solve_desc* s, int64_t* id);
/* Admit what fits, run one chunk, retire finished solves. */
int hlbm_batch_run(hlbm_batch* b, int32_t* n_finished);
int hlbm_batch_status(hlbm_batch* b, int64_t id, int32_t* state, int64_t* steps_done);
/* Seconds on the library's steady clock: submission, admission, completion. */
int hlbm_batch_times(hlbm_batch* b, int64_t id, double* t_submit, double* t_admit, double* t_done);
/* A finished (DONE or FAILED) solve's simulation_state: out[0] sum rho over the
 * nodes, out[1] min |u|, out[2] max |u|, out[3..5] sum rho u, out[6] fluid nodes
 * (src/io/simulation_state.cu:73-98; walls hold rho = 0, u = 0). */
int hlbm_batch_stats(hlbm_batch* b, int64_t id, double out[7]);
/* Copy a finished solve's rho (N), u (3N) and, if kept, f (19N), then forget it
 * (any of the three pointers may be NULL: a call with all three NULL only
 * releases the solve). */
int hlbm_batch_result(hlbm_batch* b, int64_t id, double* rho, double* u, double* f);
int hlbm_batch_info_get(hlbm_batch* b, hlbm_batch_info* out);
/* The library's steady clock, seconds (same clock as hlbm_batch_times). */
double hlbm_clock(void);

/* ---- launch-shape autotuner (no counterpart in the reference) -----------
 * A domain with boundary-operator planes first tries the node layouts that
 * put each such plane's axis last (hlbm_set_layout), timed with the default
 * shape, and keeps the fastest.  Then, in that layout:
 * the step kernels take a tile shape (block_x nodes along x by block_y rows),
 * a count of nodes per thread along y (npt, 1 or 2), an XCD-aware tile order
 * and a bulk-kernel variant: 0 = plane bases in scalar registers with 32-bit
 * lane offsets and a per-node bounce-back mask (fewest instructions), 1 =
 * 64-bit per-lane addresses and neighbour flags (its loads schedule more
 * freely; npt is 1).  hlbm_autotune times every candidate on scratch memory (the domain's
 * populations, phase and step count are left as they were: A/B domains write
 * only the buffer the next step overwrites; AA domains work on a temporary
 * copy), keeps the fastest for this domain and caches it per (dims, storage,
 * collision, pattern, boundary program, device) so later domains of the same
 * shape skip the search (out->candidates == 0).  Each candidate is timed over
 * `reps` steps, or over as many as take about 5 ms when `reps` steps of the
 * default shape take less (at most 500).  First, a domain with 256 MiB
 * of populations or more checks its placement: the step speed
 * depends on where the population buffers land in HBM (up to ~8 %, fixed per
 * allocation), so up to three fresh buffers (pairs for A/B) are timed and the state moves
 * to a faster one (out->relocated; rejected buffers stay allocated until the
 * search ends so that each try lands elsewhere, so it uses up to four times the
 * population memory while free memory allows, keeping 4 GiB free;
 * HLBM_PLACEMENT=0 turns it off).  Results never depend on the
 * shape or the placement: every node runs the same arithmetic. */
typedef struct {
  int32_t block_x, block_y, npt, xcd_remap; /* chosen shape */
  int32_t candidates;                       /* shapes timed; 0 = taken from the cache */
  int32_t variant;                          /* bulk-kernel variant, 0 or 1 */
  int32_t relocated;                        /* population buffers moved by the placement check */
  int32_t order;                            /* chosen node layout, 100*order[0] + 10*order[1] + order[2] */
  double ms_best;                           /* per step, chosen shape */
  double ms_default;                        /* per step, the default 128x2 / npt 1 / remap */
} hlbm_tune_info;

int hlbm_autotune(hlbm_domain* h, int32_t reps, void* stream, hlbm_tune_info* out);
/* Set the shape by hand: block_x*block_y a multiple of 64 and <= 256, npt 1|2,
 * variant 0|1 (domains above 2^32 bytes per population plane, or with
 * start_iteration boxes, always take variant 1). */
int hlbm_set_launch(hlbm_domain* h, int32_t block_x, int32_t block_y, int32_t npt, int32_t xcd_remap,
                    int32_t variant);

/* Node layout in HBM (no counterpart in the reference, whose fields are x
 * fastest, box3d.hpp:127-132): order = the axes from fastest to slowest in
 * memory, a permutation of {0 (x), 1 (y), 2 (z)}; {0, 1, 2} at creation.  The
 * bulk kernel's lanes run along order[0].  Putting the axis of x-normal (or
 * y-normal) boundary planes last makes those planes contiguous, so the
 * boundary-operator kernel reads and writes whole cache lines.  The state,
 * flags and rho/u fields move to the new layout (a pending streaming is
 * applied first); every host-side array of this ABI stays in node layout (x
 * fastest), and results do not depend on the layout.  hlbm_autotune tries the
 * layouts that put a boundary-plane axis last. */
int hlbm_set_layout(hlbm_domain* h, const int32_t order[3]);
int hlbm_get_layout(const hlbm_domain* h, int32_t order[3]);

/* ---- measurement helper (no counterpart in the reference) --------------
 * Times `reps` streaming copies of `bytes` bytes between two scratch buffers
 * it allocates and frees itself (kernel "hbm_copy_kernel"): the achievable
 * HBM bandwidth and the calibration point for the HBM byte counters. */
int hlbm_copy_bench(uint64_t bytes, int32_t reps, void* stream, double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* HLBM_H */
