/*
 * lbm_amd.h -- C ABI of the MI355X-native D2Q9 lattice-Boltzmann engine.
 *
 * This is the drop-in boundary under the LLALBM parallelization-policy API.
 * A policy class (HIPPolicy<2, Configuration>, see the C++ headers shipped in
 * the package) forwards the static methods that Lattice::perform_lbm calls to
 * these entry points.  Populations stay on the GPU between calls.  Host
 * arrays are owned by the caller and copied.  Every function returns an
 * lbm_status; lbm_last_error() describes the last failure of the calling
 * thread.  One host thread drives a context; calls on one context are
 * stream-ordered and not re-entrant.
 *
 * Layout conventions follow the reference's Eigen ColMajor tensors:
 *   populations (NR, NC, 9): element (i, j, d) at  i + NR*(j + NC*d)
 *   density     (NR, NC)   : element (i, j)    at  i + NR*j
 *   velocity    (NR, NC, 2): element (i, j, k) at  i + NR*(j + NC*k)
 *   node class  (NR, NC)   : uint8 per node, values of lbm_node_class
 * with i = row (0 = top) and j = column, as in the reference
 * (llalbm/src/core/lattices/Lattice.hpp:58-70, NodeIdentification.hpp:49-51).
 */
#ifndef LBM_AMD_H
#define LBM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_ctx lbm_ctx;

typedef enum lbm_status {
    LBM_OK = 0,
    LBM_ERR_INVALID = 1,     /* bad argument or geometry */
    LBM_ERR_HIP = 2,         /* HIP runtime failure */
    LBM_ERR_NAN = 3,         /* a NaN density was produced (reference: LLALBM_DEBUG check, Lattice.hpp:412-422) */
    LBM_ERR_STATE = 4,       /* call out of order (e.g. step before geometry) */
    LBM_ERR_UNSUPPORTED = 5, /* valid in the reference, not handled here (documented cases) */
    LBM_ERR_NOMEM = 6,
    LBM_ERR_RANGE = 7        /* FP16 storage: a scaled population reached the FP16 limit (65504); the storage
                                scale is too large for the flow (no reference counterpart) */
} lbm_status;

/* Node classes = the reference's InputNodeType (llalbm/src/utils/aliases.hpp:41-48). */
typedef enum lbm_node_class {
    LBM_FLUID = 0,
    LBM_SOLID = 1,
    LBM_BOUNDARY = 2,
    LBM_INLET = 3,
    LBM_OUTLET = 4,
    LBM_OBSTACLE = 5
} lbm_node_class;

typedef enum lbm_collision {
    LBM_BGK = 0, /* BGKCollision.hpp:119-155 */
    LBM_TRT = 1  /* TRTCollision.hpp:146-191 */
} lbm_collision;

typedef enum lbm_open_policy {
    LBM_ZOUHE_VELOCITY = 0, /* ZouHePolicy.hpp:58-123 */
    LBM_ZOUHE_PRESSURE = 1, /* no reference counterpart: prescribed density */
    LBM_OPEN_NOOP = 2       /* OpenBoundaryPolicy.hpp:58-61 */
} lbm_open_policy;

typedef enum lbm_storage {
    LBM_STORE_F64 = 0,         /* plain FP64: the drop-in, parity path */
    LBM_STORE_F32_SHIFTED = 1, /* f - w in FP32, FP32 collision, FP64 reductions */
    LBM_STORE_F16_SHIFTED = 2  /* (f - w) * scale in FP16, FP32 collision, FP64 reductions */
} lbm_storage;

typedef enum lbm_step_flags {
    LBM_STEP_MACRO = 1u,    /* the last step of the call also stores rho/u (reference global_rho/global_u) */
    LBM_STEP_CHECK = 2u,    /* synchronise after the call and report LBM_ERR_NAN (or LBM_ERR_RANGE) */
    LBM_STEP_NO_GRAPH = 4u  /* launch every step eagerly instead of replaying a captured HIP graph */
} lbm_step_flags;

typedef struct lbm_desc {
    int64_t n_rows;        /* NR (height), >= 3 */
    int64_t n_cols;        /* NC (width),  >= 3 */
    int32_t collision;     /* lbm_collision */
    int32_t storage;       /* lbm_storage */
    double tau;            /* BGK relaxation time (BGKCollisionPolicy<2>::tau, BGKCollision.hpp:67) */
    double tau_even;       /* TRT tau+ (TRTCollision.hpp:69) */
    double tau_odd;        /* TRT tau- (TRTCollision.hpp:70) */
    int32_t inlet_policy;  /* lbm_open_policy of the inlet list  (LatticeConfiguration inlet_policy_t) */
    int32_t outlet_policy; /* lbm_open_policy of the outlet list (LatticeConfiguration outlet_policy_t) */
    double f16_scale;      /* FP16 storage scale (a power of two in [1, 2^24]); 0 = chosen from the prescribed
                              values of the first step (see lbm_init_equilibrium) */
    int32_t device;        /* HIP device ordinal */
    int32_t options;       /* lbm_options bits (0 = none) */
} lbm_desc;

typedef enum lbm_options {
    /* FP16 storage rounds stochastically by default: slowly varying fields
     * otherwise round the same way step after step and the errors add up
     * coherently (4096^2 cavity, 200k steps: velocity L2 error 0.44 rounded
     * to nearest, 2.8e-4 stochastic).  This bit selects round-to-nearest.
     * No reference counterpart. */
    LBM_OPT_F16_ROUND_TO_NEAREST = 1,
    /* Populations are stored blocked (the 9 directions of 64 consecutive
     * wet nodes together) whenever a buffer of 9 * plane elements stays
     * below 4 GiB, else in 9 planes.  This bit selects the planes for any
     * size.  Results are the same in both layouts, bit for bit; the
     * blocked one is faster.  No reference counterpart. */
    LBM_OPT_PLANAR_LAYOUT = 2
} lbm_options;

/* Library version string. */
const char *lbm_version(void);

/* Message of the last failure on the calling thread ("" if none). */
const char *lbm_last_error(void);

/* Creates a context.  Replaces constructing Lattice<Policy> with its three
 * population tensors (Lattice.hpp:173-205) and the static parameter setters
 * BGKCollisionPolicy<2>::initialize (BGKCollision.hpp:75-88) and
 * TRTCollisionPolicy<2>::initialize (TRTCollision.hpp:83-93). */
int lbm_create(const lbm_desc *desc, lbm_ctx **out);

/* Frees all device memory of the context. */
int lbm_destroy(lbm_ctx *ctx);

/* Changes the relaxation parameters (BGKCollisionPolicy<2>::initialize,
 * TRTCollisionPolicy<2>::initialize / enforce_magic_parameter,
 * TRTCollision.hpp:128-134). */
int lbm_set_relaxation(lbm_ctx *ctx, double tau, double tau_even, double tau_odd);

/* Makes the wet obstacles partially saturated bounce-back nodes
 * (PSBounceBackPolicy<2> as the lattice's obstacle policy,
 * PSBounceBackPolicy.hpp:177-297): each obstacle o with weight b sends
 * into its wet neighbour n = o + c_d
 *   p_q - (1 - b) (p_q - e_q) / tau + b ((P_d(n) - e_d(n)) - (p_q - 2 w_d))
 * with q = opp d, p_q = F*_q(n), e the equilibrium of n, instead of the
 * plain bounce-back p_q.  tau is the policy's own relaxation time
 * (PSBounceBackPolicy<2>::tau, set by its initialize / allowed_tau);
 * b[k] is the weight of the k-th wet obstacle in list order
 * (ObstaclePoint<2>::b, Lattice.hpp:245-278), NULL = 1 for every obstacle
 * (the value the reference's weights keep).  tau <= 0 switches it off.
 * FP64 storage only (LBM_ERR_UNSUPPORTED otherwise); every node then
 * takes the general path of the step.  Set after lbm_set_geometry and
 * before the first step of a run (LBM_ERR_STATE otherwise); not for
 * ensemble samples. */
int lbm_set_partial_bounce_back(lbm_ctx *ctx, double tau, const double *b);

/* Installs the geometry from a class plane (ColMajor NR*NC).  Replaces
 * Lattice::reinit (Lattice.hpp:453-481) with the node lists that
 * build_lattice / read_lattice_file produce (ConstructionTools.hpp:54-201,
 * LatticeReader.hpp:367-522), including identify_node_type and
 * identify_obstacle_propagation (NodeIdentification.hpp:45-242).
 * Inlet/outlet nodes must lie on the domain perimeter (the reference
 * asserts it, LatticeReader.hpp:288-300); SOLID nodes next to a wet node
 * are LBM_ERR_UNSUPPORTED.  Size limits (LBM_ERR_UNSUPPORTED, no
 * counterpart in the reference): the kernels address populations with
 * 32-bit byte offsets spanning three planes, so 3 * plane * element size
 * must stay below 4 GiB (plane = wet nodes + 1 rounded up to 64; about
 * 179M wet nodes for FP64, twice that for FP32, four times for FP16), and
 * wave indices are 32-bit, so at most 2^30 64-row segments
 * (ceil(NR / 64) * NC).  Lattices up to a third of that (a whole buffer,
 * 9 * plane * element size, below 4 GiB) are stored blocked unless the
 * context was created with LBM_OPT_PLANAR_LAYOUT; larger ones in planes. */
int lbm_set_geometry(lbm_ctx *ctx, const uint8_t *node_class);

/* Population layout the geometry got (LBM_ERR_STATE before any):
 * LBM_LAYOUT_BLOCKED or LBM_LAYOUT_PLANES (see LBM_OPT_PLANAR_LAYOUT).  No
 * counterpart in the reference. */
enum { LBM_LAYOUT_BLOCKED = 0, LBM_LAYOUT_PLANES = 1 };
int lbm_get_layout(lbm_ctx *ctx, int32_t *layout);

/* Number of nodes per class after set_geometry: out[0..5] indexed by
 * lbm_node_class, out[6] = wet obstacle nodes (directions mask != 0). */
int lbm_node_counts(lbm_ctx *ctx, int64_t out[7]);

/* Inlet (which = 0) or outlet (which = 1) nodes in list order -- the
 * row-major order of the reference's lists -- as (i, j) pairs; *count
 * receives the number of nodes (ij may be NULL to query the count). */
int lbm_open_nodes(lbm_ctx *ctx, int which, int64_t *ij, int64_t *count);

/* Prescribed values of the inlet/outlet nodes for the coming step, in
 * list order.  ux/uy are the velocity the reference's update functions
 * return (VelocityInitializer<2>::update_nodes,
 * InletOutletInitializer.hpp:90-129); rho is the density of a
 * LBM_ZOUHE_PRESSURE list.  NULL leaves a component unchanged.  Values
 * persist until set again. */
int lbm_set_open_values(lbm_ctx *ctx, int which, const double *ux, const double *uy, const double *rho);

/* Start of Lattice::perform_lbm (Lattice.hpp:294-328): rho = 1, u = 0 and
 * populations = feq(1, 0) on the wet nodes; the step counter restarts at 0.
 * FP16 storage with f16_scale = 0 picks its scale from the prescribed
 * values here, and again from any values set after this call and before
 * the first step (the state is then exactly zero in every scale), so the
 * values of step 0 may be set before or after this call. */
int lbm_init_equilibrium(lbm_ctx *ctx);

/* Replaces the population state (NR*NC*9, reference layout).  Only wet
 * nodes are read.  Does not change the step counter. */
int lbm_upload_populations(lbm_ctx *ctx, const double *populations);

/* Current populations in the reference layout -- the tensor
 * Lattice::get_populations returns after the completed steps
 * (Lattice.hpp:492).  Non-wet nodes read 0 (the reference holds
 * order-dependent scratch there). */
int lbm_download_populations(lbm_ctx *ctx, double *populations);

/* Post-collision populations of the last completed step (the
 * after_collision_populations tensor, Lattice.hpp:136-145); non-wet nodes
 * read 0.  LBM_ERR_STATE before the first step. */
int lbm_download_after_collision(lbm_ctx *ctx, double *populations);

/* Advances n time steps.  Replaces the loop body of Lattice::perform_lbm
 * (Lattice.hpp:338-423): update_nodes, update_macro, calc_equilibrium,
 * collide, stream, wall/obstacle bounce-back, inlet/outlet boundaries --
 * fused into one kernel per step.  flags: lbm_step_flags. */
int lbm_step(lbm_ctx *ctx, int64_t n, uint32_t flags);

/* Number of completed steps. */
int lbm_get_time(lbm_ctx *ctx, int64_t *t);

/* global_rho / global_u as the reference holds them after the last step
 * called with LBM_STEP_MACRO (Lattice::get_global_rho/get_global_u,
 * Lattice.hpp:499,506): moments of the pre-collision populations on fluid
 * nodes, rho = 1 and the prescribed u on velocity-prescribed inlet/outlet
 * nodes, rho = 1, u = 0 elsewhere. */
int lbm_download_macro(lbm_ctx *ctx, double *rho, double *u);

/* Asynchronous lbm_download_macro for output on save steps
 * (Lattice.hpp:370-374 writes u_x/u_y every save_step-th step;
 * LatticeWriter.hpp:61-93).  Queues the copies of the fields of the last
 * LBM_STEP_MACRO step into rho and u (either may be NULL) on a copy stream
 * of the context and returns at once; *ticket names the copy for
 * lbm_wait.  The context keeps two device copies of the fields, so the
 * copy overlaps the steps that follow; a third LBM_STEP_MACRO step waits
 * for the copy of the first.  rho and u must stay valid until the ticket
 * completes; with pageable host memory HIP may stage the copy
 * synchronously, page-locked memory (lbm_host_alloc) keeps it
 * asynchronous. */
int lbm_download_macro_async(lbm_ctx *ctx, double *rho, double *u, int64_t *ticket);

/* Blocks until the work named by a ticket of this context has completed
 * (tickets older than the last 64 are complete already). */
int lbm_wait(lbm_ctx *ctx, int64_t ticket);

/* Page-locked host memory for asynchronous downloads (hipHostMalloc);
 * lbm_host_free releases it.  No reference counterpart. */
int lbm_host_alloc(int64_t bytes, void **ptr);
int lbm_host_free(void *ptr);

/* FlowAnalyzer<2>::compute_flow_properties (FlowAnalyzer.hpp:348-451) on
 * the post-collision populations of the last step, over n points given as
 * (i, j) pairs with their 9-bit direction masks (NULL, NULL = every wet
 * obstacle node, in list order).  lift_drag[0] = total lift,
 * lift_drag[1] = total drag, reduced in FP64 on the GPU.  per_point
 * (optional, 2*n) receives (lift, drag) of each point.  A direction bit
 * must name a wet neighbour (fluid, inlet or outlet) inside the lattice,
 * as the reference's direction masks do (NodeIdentification.hpp:137-211);
 * other masks are LBM_ERR_INVALID. */
int lbm_lift_drag(lbm_ctx *ctx, const int64_t *ij, const int32_t *masks, int64_t n, double lift_drag[2],
                  double *per_point);

/* lbm_lift_drag without waiting: the points are copied before the call
 * returns, the results arrive in lift_drag / per_point (which must stay
 * valid; page-locked memory keeps the copy asynchronous) when *ticket
 * completes (lbm_wait).  Used on analysis steps so the steps after them
 * queue behind the reduction instead of waiting for it. */
int lbm_lift_drag_async(lbm_ctx *ctx, const int64_t *ij, const int32_t *masks, int64_t n, double *lift_drag,
                        double *per_point, int64_t *ticket);

/* FP64 reductions over the current populations: out[0] = mass of the wet
 * nodes, out[1] = x momentum and out[2] = y momentum of the fluid nodes,
 * out[3] = mass of the fluid nodes, out[4] = number of fluid nodes.  No
 * reference counterpart (the reference has no reductions besides
 * FlowAnalyzer); used for mass drift and Darcy permeability. */
int lbm_reduce(lbm_ctx *ctx, double out[5]);

/* lbm_reduce without allocating and without waiting: the five sums go to
 * `out`, 5 doubles in DEVICE memory, in stream order on the context's
 * stream (lbm_get_stream), with the same kernels and the same bits as
 * lbm_reduce.  No counterpart in the reference: for callers that cannot
 * allocate -- code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc.  The block partials live in one fixed scratch
 * array per device (5 MiB, held until the process ends), so the lattice
 * may have at most 131072 reduction blocks, ceil(NR / 256) * NC (e.g.
 * 4096 x 8192); larger ones are LBM_ERR_UNSUPPORTED.  All calls on one
 * device share that array: calls whose kernels can overlap (contexts on
 * different streams not ordered by the caller, or threads) race, and the
 * caller must order them. */
int lbm_reduce_device(lbm_ctx *ctx, double *out);

/* Waits for all work queued on the context. */
int lbm_synchronize(lbm_ctx *ctx);

/* The HIP stream the context queues its work on (hipStream_t). */
int lbm_get_stream(lbm_ctx *ctx, void **stream);

/* Storage scale in use (1 unless LBM_STORE_F16_SHIFTED). */
int lbm_get_scale(lbm_ctx *ctx, double *scale);

/* Launch shape of the step kernel: `segments` 64-row segments per wave
 * (1, 2, 4 or 8; 0 = the storage format's default), or LBM_LAUNCH_WET_RUNS:
 * one wave per run of up to 64 consecutive wet rows of a column, so no lane
 * idles on a solid node, or LBM_LAUNCH_RUN_TABLE: the same runs with every
 * lane's neighbour addressing read from a table built on first use (4 bytes
 * per wet node and step more traffic, far fewer instructions per wave).
 * Results are the same bits for every shape.  No counterpart in the
 * reference (a tuning knob of this engine). */
enum { LBM_LAUNCH_WET_RUNS = 64, LBM_LAUNCH_RUN_TABLE = 65 };
int lbm_set_launch(lbm_ctx *ctx, int32_t segments);
int lbm_get_launch(lbm_ctx *ctx, int32_t *segments);

/* Occupancy of the step kernels: at most `groups_per_cu` work-groups (of
 * four waves) per compute unit, 2 to 7, or 0 for as many as their
 * registers allow (8 for most shapes).  Fewer waves in flight can move more
 * bytes per second.  Until set or tuned it is 7 (FP32, FP16 storage) or 5
 * (FP64), the best for the default shapes on MI355X; lbm_get_occupancy
 * reports the value in use.  Results are the same bits at every occupancy.
 * No counterpart in the reference. */
int lbm_set_occupancy(lbm_ctx *ctx, int32_t groups_per_cu);
int lbm_get_occupancy(lbm_ctx *ctx, int32_t *groups_per_cu);

/* Times each launch shape (1, 2, 4, 8 segments, wet runs, tabled wet runs)
 * at each occupancy (no limit, 7 down to 3 work-groups per compute unit)
 * over `steps` launches (every launch rewrites the spare buffer from the
 * current state, so the state does not change) and keeps the fastest
 * pair.  Reports the shape and its time per step (outputs may be NULL);
 * lbm_get_occupancy reports the occupancy.  LBM_ERR_STATE for a context
 * that belongs to an ensemble (the ensemble has its own,
 * lbm_batch_autotune) and for a lattice with partially saturated
 * bounce-back (every launch advances its carried obstacle values).  No
 * counterpart in the reference. */
int lbm_autotune(lbm_ctx *ctx, int32_t steps, int32_t *segments, double *ms_per_step);

/* ------------------------------------------------------------------ */
/* Ensembles of independent lattices                                   */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference, which advances one Lattice per
 * process (Lattice.hpp:287-425).  An ensemble groups lattices that were
 * set up as usual (lbm_create, lbm_set_geometry, lbm_set_open_values,
 * lbm_init_equilibrium or lbm_upload_populations) and advances all of them
 * with one step-kernel launch per time step over a table of tiles
 * (sample, column, 64*R rows) -- one wave per tile by default, or a
 * persistent grid that walks the table -- plus one small launch for the
 * samples' perimeter segments.  Each sample's results are those of
 * lbm_step on that sample alone, bit for bit.
 *
 * Rules: all samples share the device and the storage format and are in
 * the same state (all freshly initialised/uploaded, or all stepped); a
 * context belongs to at most one ensemble; while it does, lbm_step and
 * lbm_set_geometry on it fail with LBM_ERR_STATE and lbm_destroy refuses
 * it.  Everything else (prescribed values, downloads, reductions, lift and
 * drag) stays available per sample and is ordered after the ensemble's
 * steps. */
typedef struct lbm_batch lbm_batch;

int lbm_batch_create(lbm_ctx *const *samples, int64_t n_samples, lbm_batch **out);
int lbm_batch_destroy(lbm_batch *batch);

/* Tile shape: `segments` 64-row segments per wave (1, 2, 4 or 8), and
 * either a persistent grid of `blocks` work-groups of 256 threads that
 * walks the tile table (rounded up to a multiple of 8, one share per XCD)
 * or, with blocks = 0 (the default), one wave per tile.
 * segments = LBM_LAUNCH_WET_RUNS (blocks = 0) takes one wave per run of up
 * to 64 consecutive wet rows of a sample's column, over every sample's run
 * table (as lbm_set_launch for one lattice). */
int lbm_batch_set_launch(lbm_batch *batch, int32_t segments, int64_t blocks);
int lbm_batch_get_launch(lbm_batch *batch, int32_t *segments, int64_t *blocks);

/* Autotuner: times every candidate tile shape and grid (one wave per tile,
 * or 2, 4, 8, 16 persistent work-groups per compute unit; and wet-run
 * waves) over `steps`
 * launches each (a launch rewrites the spare buffer from the current
 * state, so the state does not change) and keeps the fastest.  Reports
 * the choice and its time per step (any output may be NULL). */
int lbm_batch_autotune(lbm_batch *batch, int32_t steps, int32_t *segments, int64_t *blocks,
                       double *ms_per_step);

/* Advances every sample n time steps.  Runs of steps with unchanged
 * prescribed values replay a captured HIP graph of the step loop unless
 * LBM_STEP_NO_GRAPH is given; LBM_STEP_CHECK reports LBM_ERR_NAN with the
 * first sample whose density became NaN.  LBM_STEP_MACRO is not
 * supported here (step the sample alone for its fields). */
int lbm_batch_step(lbm_batch *batch, int64_t n, uint32_t flags);

/* Per-sample FP64 reductions of the current pre-collision state, as
 * lbm_reduce: out[5*s + k] for sample s. */
int lbm_batch_reduce(lbm_batch *batch, double *out);

/* Total wet nodes over the samples; tiles in the current table. */
int lbm_batch_info(lbm_batch *batch, int64_t *wet_nodes, int64_t *tiles);
int lbm_batch_get_stream(lbm_batch *batch, void **stream);
int lbm_batch_synchronize(lbm_batch *batch);

/* ------------------------------------------------------------------ */
/* Measurement (no reference counterpart)                              */
/* ------------------------------------------------------------------ */

/* The practical HBM ceiling of a plain stream on `device`: the best rate
 * (read + write bytes per second / 1e9) of a 16-byte-per-lane copy kernel
 * over `bytes` bytes (rounded down to 16), with and without streaming
 * stores, each timed over `reps` launches.  Allocates and frees two
 * buffers of that size.  Benchmarks set the step kernels' achieved
 * bandwidth against it. */
int lbm_copy_bandwidth(int32_t device, int64_t bytes, int32_t reps, double *gbs);

/* The rate (read + write bytes per second / 1e9, 18 elements moved per
 * node) of a D2Q9 step's access pattern without its arithmetic, on a dense
 * n_rows x n_cols lattice with periodic wrap (n_rows * n_cols a multiple of
 * 64), `elem_bytes` 2, 4 or 8, timed over `reps` launches:
 *   LBM_PATTERN_AB       pull between two buffers of 9 planes (the step
 *                        kernels' layout);
 *   LBM_PATTERN_AA       the AA in-place pattern on one buffer, even and
 *                        odd steps alternating;
 *   LBM_PATTERN_AA_EVEN  its even step alone (own slots, swapped);
 *   LBM_PATTERN_AA_ODD   its odd step alone (neighbours' slots);
 *   LBM_PATTERN_BLOCKED  pull with the 9 populations of 64 consecutive
 *                        nodes stored together.
 * Allocates and frees two buffers of 9 * n_rows * n_cols * elem_bytes
 * bytes.  No counterpart in the reference (a measurement of this engine's
 * layout choices). */
enum { LBM_PATTERN_AB = 0, LBM_PATTERN_AA = 1, LBM_PATTERN_AA_EVEN = 2, LBM_PATTERN_AA_ODD = 3, LBM_PATTERN_BLOCKED = 4 };
int lbm_pattern_bandwidth(int32_t device, int32_t pattern, int32_t elem_bytes, int64_t n_rows, int64_t n_cols,
                          int32_t reps, double *gbs);

/* ------------------------------------------------------------------ */
/* Geometry helpers (host only; write a class plane, ColMajor NR*NC)   */
/* ------------------------------------------------------------------ */

/* Lid-driven cavity: BOUNDARY perimeter, INLET on row 0, columns
 * 1..NC-2.  Same node sets as ConstructionInfo::add_perimeter_nodes +
 * add_nodes_interval({0,1},{0,NC-2},INLET) (ConstructionInfo.hpp:147-281)
 * and as test/assets/lid-lattice.txt. */
int lbm_geometry_cavity(int64_t nr, int64_t nc, uint8_t *node_class);

/* Channel: BOUNDARY rows 0 and NR-1, INLET column 0 and OUTLET column
 * NC-1 on rows 1..NR-2 (examples/3_Vortex.cpp:36-43). */
int lbm_geometry_channel(int64_t nr, int64_t nc, uint8_t *node_class);

/* Marks a disc of OBSTACLE nodes: integer (i-ci)^2 + (j-cj)^2 <= r^2, as
 * ConstructionInfo::add_obstacle_hyper_sphere (ConstructionInfo.hpp:374-451);
 * fails with LBM_ERR_INVALID, like the reference, when the disc would
 * leave rows/columns 1..N-2.  *added (optional) = nodes newly marked. */
int lbm_geometry_add_disc(int64_t nr, int64_t nc, uint8_t *node_class, int64_t ci, int64_t cj, int64_t r,
                          int64_t *added);

/* Random porous medium (no reference counterpart; the reference's
 * add_obstacle_random_spheres draws a fixed count with libstdc++'s
 * distribution, ConstructionInfo.hpp:462-486): discs with integer radius
 * uniform in [rmin, rmax] and centres uniform over the domain, drawn from
 * mt19937_64(seed) with the mapping lo + floor(x * (hi-lo+1) / 2^64);
 * a disc leaving rows/columns 1..N-2 is redrawn; stops when the OBSTACLE
 * fraction of all nodes reaches target_solid.  *solid receives the
 * fraction reached, *ndiscs the number of discs placed. */
int lbm_geometry_porous(int64_t nr, int64_t nc, uint8_t *node_class, uint64_t seed, int64_t rmin, int64_t rmax,
                        double target_solid, double *solid, int64_t *ndiscs);

#ifdef __cplusplus
}
#endif

#endif /* LBM_AMD_H */
