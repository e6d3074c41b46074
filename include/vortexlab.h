/*
 * vortexlab-mi355x: C ABI of the MI355X-native D2Q9 lattice-Boltzmann time step.
 *
 * The reference (a Rust crate) has no FFI layer: its boundary is the public Rust API
 * Builder -> Sim -> Channel.  This header is that API restated as plain C, so a Rust
 * `extern "C"` binding (or ctypes, JNI, ...) can drive the GPU solver with the same
 * call sequence.  Each declaration cites the reference item it replaces
 * (path:line into the reference crate).
 *
 * Conventions
 *  - Every fallible call returns vl_status (0 = VL_OK).  The reference panics on misuse
 *    (src/sim.rs:85, 101, 122-124, 222, 261, 303); here those become VL_EINVAL with a
 *    message in vl_last_error() (thread-local).
 *  - Divergence (the reference's `!cd.is_finite()` check, src/cases/tunnel.rs:78-84,
 *    src/cases/cylinder.rs:179-186) is VL_ENONFINITE; the step index is in the message.
 *  - Handles are single-caller: one thread at a time per handle.  Each vl_sim owns one
 *    HIP stream; all work is stream-ordered and only readbacks synchronise.
 *  - Arrays handed in are copied; the library owns all device memory.
 *  - Populations use the reference layout in every host array: f[i * nx * ny + y * nx + x],
 *    i = 0..8 in the reference's direction numbering (src/lattice.rs:3-24).
 */
#ifndef VORTEXLAB_H
#define VORTEXLAB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int vl_status;
#define VL_OK 0
#define VL_EINVAL 1      /* bad argument or builder misuse (reference: panic/assert) */
#define VL_ENOMEM 2      /* host or device allocation failed */
#define VL_EDEVICE 3     /* a HIP call failed (no GPU, launch failure, ...) */
#define VL_ENONFINITE 4  /* the run diverged (non-finite force coefficient) */

#define VL_ABI_VERSION 1

/* Storage precision of the populations.  F64 stores f_i; F32/F16 store the deviation
 * f_i - w_i (F16 scaled by 2^10) and collide in FP32.  Forces, the outlet controller and
 * all statistics stay FP64 for every precision. */
#define VL_F64 0
#define VL_F32 1
#define VL_F16 2

/* Collision models: src/lattice.rs:28-45. */
#define VL_BGK 0
#define VL_TRT 1

/* Last error message of the calling thread ("" if none). */
const char* vl_last_error(void);
int vl_abi_version(void);
/* Number of HIP devices visible (0 without a GPU); never fails. */
int vl_device_count(void);

/* ------------------------------------------------------------------------------------
 * Shapes: src/geometry.rs:12-95, 187-279.  A shape only answers "is (x, y) inside?".
 * Composite constructors take ownership of `inner`.
 * ---------------------------------------------------------------------------------- */
typedef struct vl_shape vl_shape;
typedef int (*vl_inside_fn)(void* ctx, double x, double y);

vl_shape* vl_shape_circle(double cx, double cy, double r);                 /* geometry.rs:16-27 */
vl_shape* vl_shape_polygon(const double* xy, size_t npts);                 /* geometry.rs:30-49 */
vl_shape* vl_shape_rounded_square(double cx, double cy, double half, double r); /* :52-66 */
vl_shape* vl_shape_rotated(vl_shape* inner, double cx, double cy, double angle); /* :69-82 */
vl_shape* vl_shape_translated(vl_shape* inner, double dx, double dy);      /* geometry.rs:85-95 */
/* Square section of side d with a corner treatment (geometry.rs:187-261):
 * corner 0 sharp, 1 chamfered(b), 2 rounded(b), 3 recessed(b), 4 double-recessed(b). */
vl_shape* vl_shape_section(int corner, double b, double cx, double cy, double d, double angle);
/* "Everything outside the band y0 < y < y1" (src/cases/poiseuille.rs:12-21). */
vl_shape* vl_shape_outside_band(double y0, double y1);
/* Caller-supplied predicate (the Rust `Box<dyn Shape>`); called only at build time. */
vl_shape* vl_shape_callback(vl_inside_fn fn, void* ctx);
int vl_shape_inside(const vl_shape* s, double x, double y);
/* Fraction of the link (x,y)->(x+dx,y+dy) in the fluid: geometry.rs:268-279. */
double vl_cut_fraction(const vl_shape* s, double x, double y, double dx, double dy);
void vl_shape_free(vl_shape* s);

/* ------------------------------------------------------------------------------------
 * Wall velocity functions (the reference's VelocityFn closures, src/sim.rs:33-34),
 * evaluated only at build time at the wall point of each link.
 * ---------------------------------------------------------------------------------- */
typedef void (*vl_velocity_fn)(void* ctx, double x, double y, double* out2);
#define VL_VEL_REST 0      /* wall at rest (velocity: None) */
#define VL_VEL_CONST 1     /* [p0, p1] */
#define VL_VEL_PARABOLA 2  /* channel inlet profile, u_mean = p0, h = p1 (cases/channel.rs:95-103) */
#define VL_VEL_SPIN 3      /* rotation about (p0, p1): [-(y - p1), x - p0] (cases/channel.rs:122) */
#define VL_VEL_CALLBACK 4  /* fn(ctx, x, y, out) */
typedef struct {
  int kind;
  double p[4];
  vl_velocity_fn fn;
  void* ctx;
} vl_velocity;

/* ------------------------------------------------------------------------------------
 * Builder: src/sim.rs:70-137.
 * ---------------------------------------------------------------------------------- */
typedef struct vl_builder vl_builder;
vl_builder* vl_builder_new(size_t nx, size_t ny);                          /* sim.rs:84-97 */
/* pub fields nx, ny, periodic_x/y, tau, collision, accel (sim.rs:70-81). */
vl_status vl_builder_config(vl_builder* b, int periodic_x, int periodic_y, double tau,
                            int collision, double magic, double ax, double ay);
/* Kind::Wall(Wall{velocity, vgroup, fgroup, shape}) (sim.rs:36-54, 100-104).  Returns the
 * kind id (>0) or -1.  fgroup < 0 means "no force bookkeeping"; `shape` is consumed. */
int vl_builder_add_wall(vl_builder* b, const vl_velocity* vel, int vgroup, int fgroup,
                        vl_shape* shape);
int vl_builder_add_pressure(vl_builder* b, double rho);                    /* Kind::Pressure */
int vl_builder_add_radiating(vl_builder* b, double rho);                   /* Kind::Radiating */
vl_status vl_builder_fill_rect(vl_builder* b, size_t x0, size_t x1, size_t y0, size_t y1,
                               int kind);                                  /* sim.rs:111-118 */
vl_status vl_builder_fill_shape(vl_builder* b, int kind);                  /* sim.rs:120-132 */
void vl_builder_free(vl_builder* b);

/* ------------------------------------------------------------------------------------
 * Sim: src/sim.rs:180-673 on one GPU.
 * ---------------------------------------------------------------------------------- */
typedef struct vl_sim vl_sim;

typedef struct {
  int64_t nx, ny;
  int64_t fluid_nodes;   /* sim.rs:377-379 */
  int64_t link_count;    /* sim.rs:386-388 */
  int64_t bnode_count;
  int64_t pnode_count;
  int64_t force_groups;  /* len of Sim::force */
  int64_t t;             /* Sim::t */
  int32_t precision;
  int32_t collision;
  double omega_plus, omega_minus;
  double outlet_density_offset; /* Sim::outlet_density_offset (last step) */
  int64_t device_bytes;
  int32_t split;  /* 1: boundary and bulk run as two concurrent kernels; 0: one fused kernel */
  int32_t pass_steps;      /* K: groups of K steps run one K-step pass over the covered interior (1: off) */
  int64_t pass_nodes;      /* interior nodes the pass covers */
  int64_t pass_tasks;      /* its tasks (one wave each) */
  int64_t listed_nodes;    /* interior nodes stepped one step at a time beside it */
  int32_t pass_rows;       /* output rows per task, at most */
  int32_t pass_joined_tasks; /* tasks in joined strips (a workgroup's waves sweep one strip) */
  int64_t pass_far_nodes;  /* covered nodes of the far tasks (their pass overlaps the single steps) */
  int64_t pass_far_tasks;
} vl_sim_info;

/* Per-step host controls (what Channel::step sets before Sim::step, cases/channel.rs:136-147).
 * target is used only when has_target != 0 (Sim::outflow_target = Some(target)). */
typedef struct {
  double vscale[4];
  double target;
  int32_t has_target;
  int32_t pad_;
} vl_ctl;

/* Builder::build (sim.rs:134-136, from_builder :219-354).  The builder is consumed
 * (freed) whatever the outcome.  State is initialised to rest (rho 1, u 0). */
vl_status vl_sim_create(vl_builder* b, int precision, vl_sim** out);
void vl_sim_destroy(vl_sim* s);
vl_status vl_sim_info_get(vl_sim* s, vl_sim_info* out);
/* Sim::init (sim.rs:357-371): equilibrium of per-node (rho, ux, uy), arrays of nx*ny;
 * non-fluid cells get (1, 0, 0).  Resets t to 0. */
vl_status vl_sim_init_macro(vl_sim* s, const double* rho, const double* ux, const double* uy);
/* Sim::vscale and Sim::outflow_target (sim.rs:195-208). */
vl_status vl_sim_set_vscale(vl_sim* s, const double* vscale4);
vl_status vl_sim_set_outflow_target(vl_sim* s, int enabled, double target);
/* nsteps x Sim::step (sim.rs:391-420) with the current vscale / outflow_target. */
vl_status vl_sim_step(vl_sim* s, int64_t nsteps);
/* nsteps steps with per-step controls ctl[k]; if force_hist is non-NULL it receives
 * Sim::force after every step: force_hist[(k * force_groups + g) * 2 + {0,1}]. */
vl_status vl_sim_run(vl_sim* s, int64_t nsteps, const vl_ctl* ctl, double* force_hist);
/* vl_sim_run without any device allocation during the call (no reference counterpart; for callers
 * that cannot allocate: code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as vl_sim_run, bit for bit.  The per-step control table
 * and the force-history staging live in a fixed scratch area of the library (one per GPU, about
 * 450 KB, set up when the first simulation is created on that GPU) instead of the handle's own
 * buffers, and steps are launched directly rather than through step-window graphs (which would be
 * instantiated on first use).  At most 4096 steps per call and 4 force groups, else VL_EINVAL;
 * VL_EINVAL also while kernel timing is on.  Every handle on a GPU shares its scratch area: calls
 * that overlap on one GPU race. */
vl_status vl_sim_run_noalloc(vl_sim* s, int64_t nsteps, const vl_ctl* ctl, double* force_hist);
/* Instantiates, without running anything, the step graphs a vl_sim_run of nsteps steps from the
 * current state would replay (no reference counterpart: the reference has no launch graphs).
 * vl_sim_run replays whole step windows (30 or 32 steps) on its own; the whole groups of K steps
 * of a shorter run are replayed as one graph only when this call has instantiated it, and
 * launched one by one otherwise.  Results are the same bits either way.  For callers that time
 * or repeat short runs. */
vl_status vl_sim_prepare(vl_sim* s, int64_t nsteps);
/* The same for a run of nsteps steps that starts skip steps from the current state (no reference
 * counterpart), so that a caller can prepare a timed run before its warmup steps. */
vl_status vl_sim_prepare_after(vl_sim* s, int64_t skip, int64_t nsteps);
/* Sim::force after the last step (force_groups * 2 doubles). */
vl_status vl_sim_get_force(vl_sim* s, double* out);
/* Populations of A (the current state) in the reference layout, as f_i in FP64. */
vl_status vl_sim_read(vl_sim* s, double* f);
vl_status vl_sim_write(vl_sim* s, const double* f);
/* Sim::macros for every node (sim.rs:527-534): rho, ux, uy arrays of nx*ny. */
vl_status vl_sim_macros(vl_sim* s, double* rho, double* ux, double* uy);
vl_status vl_sim_state_hash(vl_sim* s, uint64_t* out);          /* sim.rs:654-668 */
vl_status vl_sim_total_mass(vl_sim* s, double* out);            /* sim.rs:538-548 */
vl_status vl_sim_total_momentum(vl_sim* s, double* out2);       /* sim.rs:551-563 */
vl_status vl_sim_max_speed(vl_sim* s, double* out);             /* sim.rs:566-579 */
/* Sim::vorticity (sim.rs:581-620): dv/dx - du/dy per node, nx*ny floats in node order
 * (one-sided differences next to non-fluid cells and the domain edge, 0 in solids). */
vl_status vl_sim_vorticity(vl_sim* s, float* out);
vl_status vl_sim_is_fluid(vl_sim* s, int64_t x, int64_t y, int* out); /* sim.rs:381-383 */
vl_status vl_sim_periodic(vl_sim* s, int* px, int* py);          /* sim.rs:670-672 */
/* Kernel timing with HIP events on the sim's own stream (no reference counterpart; the
 * reference times wall clock only, src/cases/cylinder.rs:224-232).  Switching it on resets
 * the counters; bulk_ms sums the duration of every bulk stream-collide launch, total_ms the
 * duration of whole steps (every kernel of the step). */
vl_status vl_sim_set_timing(vl_sim* s, int on);
/* Best-of-`iters` bandwidth of a plain device-to-device copy kernel of `bytes` (read + write
 * bytes counted): the achievable-HBM yardstick reported beside the kernel figures. */
vl_status vl_device_copy_bandwidth(size_t bytes, int iters, double* gbs);
vl_status vl_sim_get_timing(vl_sim* s, double* bulk_ms, int64_t* bulk_launches, double* total_ms,
                            int64_t* steps);
/* The K-step pass plan of one domain, host only, no device needed (no reference counterpart:
 * the planner behind the K-step passes that vl_sim_create and vl_batch_build make, exposed for
 * tests and diagnostics).  flags: nx * ny bytes, 0 interior, 1 boundary, 2 solid; k steps per
 * pass (2..5), v columns per lane (1, 2), tasks of at most `rows` output rows, strips from column
 * x0 (the halo plus a multiple of v).  join != 0 (with waves >= 2) plans joined strips, each
 * swept by a workgroup of `waves` waves together, where they fit.  waves > 0 also sets the
 * lockstep bits as vl_sim_create does for workgroups of that many waves, far and near tasks
 * launched apart; waves == 0 leaves them clear.  Writes up to max_tasks tasks, 4 int32 each
 * (first output column, first row, rows, output columns | ring-edge bits << 8: first row, last
 * row, first column, last column | lockstep << 12 | joined << 13 | the wave's first output column
 * less its first loaded column << 16, for joined tasks), in the order far joined groups, far
 * single-wave tasks, near joined groups, near single-wave tasks; and up to max_listed listed
 * nodes (y << 16 | x); counts (6 values) = {tasks, far tasks, listed nodes, covered nodes, tasks
 * in far joined groups, tasks in near joined groups}. */
vl_status vl_plan_passes(const uint8_t* flags, int nx, int ny, int k, int v, int rows, int x0, int waves,
                         int join, int32_t* tasks, int64_t max_tasks, uint32_t* listed, int64_t max_listed,
                         int64_t* counts);
/* The lockstep bits of a pass task list, host only (no reference counterpart; the grouping that
 * vl_sim_create and vl_batch_build apply before the pass kernels run).  tasks: ntasks x 4 int32
 * as vl_plan_passes writes them; owner: the case of each task (NULL: one case); ranges: nranges
 * pairs [begin, end) of the launches, each run as workgroups of `waves` consecutive tasks from
 * begin.  Unless check_only, sets bit 12 of tasks[4i+3] on exactly the full workgroups whose
 * tasks share first row, row count and owner, and clears it elsewhere.  Then checks that every
 * workgroup of every range holds no marked task or only such a full group: VL_EINVAL if not
 * (a workgroup whose waves met different barriers would hang).  Joined tasks (bit 13, from
 * vl_plan_passes) keep their bits and must fill whole workgroups whose waves load adjacent runs
 * of 64 or 128 columns. */
vl_status vl_pass_lockstep(int32_t* tasks, const int32_t* owner, int64_t ntasks, const int64_t* ranges, int nranges,
                           int waves, int check_only);

/* ------------------------------------------------------------------------------------
 * Channel: src/cases/channel.rs:40-155 (a body in a channel; the per-step driver).
 * ---------------------------------------------------------------------------------- */
#define VL_INFLOW_PARABOLIC 0
#define VL_INFLOW_UNIFORM 1
#define VL_INFLOW_UNIFORM_OPEN 2
typedef struct {
  double cx, cy, omega;
  uint64_t start, duration;
} vl_kick;
typedef struct {
  int64_t nx, ny;
  int32_t inflow;      /* VL_INFLOW_* */
  int32_t collision;   /* VL_BGK / VL_TRT */
  double u;            /* u_mean (parabolic) or u (uniform) */
  double nu;           /* lattice viscosity */
  double magic;        /* TRT Lambda */
  uint64_t ramp_steps;
  int32_t has_kick;
  int32_t pad_;
  vl_kick kick;
} vl_channel_spec;

typedef struct vl_channel vl_channel;
/* Channel::new (channel.rs:89-134); `body` is consumed. */
vl_status vl_channel_create(const vl_channel_spec* spec, vl_shape* body, int precision,
                            vl_channel** out);
void vl_channel_destroy(vl_channel* c);
vl_sim* vl_channel_sim(vl_channel* c);
/* Channel::absorbing_outlet (channel.rs:84-85). */
vl_status vl_channel_set_absorbing(vl_channel* c, int absorbing);
/* nsteps x Channel::step (channel.rs:136-148).  If coef is non-NULL it receives the drag
 * and lift coefficients 2F/(u_ref^2 d) after every step (coef[2k], coef[2k+1]), as
 * Channel::coefficients (channel.rs:151-154).  A non-finite drag coefficient returns
 * VL_ENONFINITE (the history up to that step is still written). */
vl_status vl_channel_run(vl_channel* c, int64_t nsteps, double d, double* coef);
vl_status vl_channel_coefficients(vl_channel* c, double d, double* cd, double* cl);

/* ------------------------------------------------------------------------------------
 * Batched cases (no reference counterpart: the reference runs one case at a time,
 * src/cases/tunnel.rs:152-192).  Each case is a Channel (spec + body outline) run for its own
 * number of steps; all cases step together on one GPU over a packed tile table: the covered
 * interior of every case in K-step passes, its boundary and remaining interior nodes in single
 * steps, in replayed step-window graphs.  A case's results are bit-identical to the same case
 * run alone with vl_channel_run in the same precision.
 * ---------------------------------------------------------------------------------- */
typedef struct vl_batch vl_batch;
typedef struct {
  /* lattice_updates: node updates performed so far, each case's fluid nodes times the steps it
   * completed (a retired case counts the steps it ran). */
  int64_t ncases, total_nodes, fluid_nodes, lattice_updates, max_steps, t;
  int32_t boundary_tiles, bulk_tiles, tile_w, tile_h, grid_bulk, grid_boundary;
  double last_run_seconds;
  int64_t device_bytes;
  double autotune_ms[3][4]; /* time of a trial step 0 (boundary kernel beside the bulk) per
                               (tile shape 64x4 / 128x2 / 256x1, bulk grid: one workgroup per
                               tile / 4 / 8 / 16 persistent per CU); only the first column is
                               timed unless VL_BATCH_AUTOTUNE=full, the others stay 0 */
  /* K-step passes (pass_steps > 1): a case that runs a whole group of K steps advances its
   * covered interior (pass_nodes over all cases, in pass_tasks strips of up to pass_rows rows)
   * K steps per pass and its other interior nodes (listed_nodes) one step at a time. */
  int32_t pass_steps, pass_rows, pass_tasks, listed_tiles;
  int64_t pass_nodes, listed_nodes;
  /* Set-up (vl_batch_build): 1 when the cell kinds, node flags and boundary tables were built on
   * the device (every body a polygon or a circle), 0 when by host threads; the time spent on the
   * cases and their tables, the K-step pass plans, packing and uploads, and the autotuner (ms). */
  int32_t setup_on_device, pad_;
  double setup_ms[4];
} vl_batch_info;
vl_status vl_batch_create(int precision, vl_batch** out);
/* Adds a case (the channel of vl_channel_create; `body` consumed) that runs `steps` steps;
 * coefficients use the reference length `d`.  Returns the case id, or -1. */
int vl_batch_add(vl_batch* b, const vl_channel_spec* spec, vl_shape* body, int64_t steps, double d);
/* Host set-up of every case (in parallel), device packing, and the tile-shape autotuner. */
vl_status vl_batch_build(vl_batch* b);
/* Advances every live case by up to max_steps steps (< 0: until all cases are done).  A case
 * whose drag turns non-finite is retired (status 1) while the others go on. */
vl_status vl_batch_run(vl_batch* b, int64_t max_steps);
vl_status vl_batch_info_get(vl_batch* b, vl_batch_info* out);
/* Kernel timing of the batch's own stepping (no reference counterpart: the reference's
 * `bench` times whole runs only, src/main.rs:448).  While on, vl_batch_run launches every step
 * directly, as a single step (no K-step passes), with HIP events around the boundary kernel
 * (side stream), the bulk kernel (main stream) and the whole step; vl_batch_get_timing returns
 * the sums in ms since the last vl_batch_set_timing call. */
vl_status vl_batch_set_timing(vl_batch* b, int on);
vl_status vl_batch_get_timing(vl_batch* b, double* boundary_ms, double* bulk_ms, double* step_ms, int64_t* steps);
/* (cd, cl) per step of case `id` (2 * steps doubles), the steps completed, and its status
 * (0 running or finished, 1 diverged). */
vl_status vl_batch_case_result(vl_batch* b, int id, double* coef, int64_t* steps_done, int* status);
vl_status vl_batch_read_case(vl_batch* b, int id, double* f);
/* Grid, fluid-node count and requested steps of case `id` (after vl_batch_build); any output
 * pointer may be NULL.  Like Sim::fluid_nodes (src/sim.rs:640) for one case of the batch. */
vl_status vl_batch_case_info(vl_batch* b, int id, int64_t* nx, int64_t* ny, int64_t* fluid_nodes, int64_t* steps);
/* FNV-1a digests of case `id`'s device tables after vl_batch_build, for comparing two builds (no
 * reference counterpart): its pitched node flags, then bx, by, bfirst, dir, kind, vgroup, lflags,
 * lnx, lny, fslot, q, uwx, uwy, cu6, rho, px, py, pnx, pny, gfirst, and the counts (nodes, links,
 * outlet nodes, outlet-controller nodes, force slots, force groups, fluid nodes): 22 values, the
 * first n written.  Equal digests for the device and the host set-up (VL_BATCH_SETUP=host) of the
 * same case mean byte-identical tables. */
vl_status vl_batch_table_digests(vl_batch* b, int id, uint64_t* out, int n);
void vl_batch_destroy(vl_batch* b);

/* ------------------------------------------------------------------------------------
 * Surrogate network (no reference counterpart): a 64x64 body mask plus the scaled log Re ->
 * (mean Cd, Cl', St), trained and evaluated by hand-written HIP kernels in FP32.  Layout:
 * 4 x [conv 3x3 + ReLU + 2x2 max-pool] (1-32-32-64-64 channels), FC 1025-128 + ReLU, FC 128-3;
 * parameters are one flat FP32 vector (per conv: weights [co][ci][3][3] then biases; then FC1
 * weights [128][1025], biases; FC2 weights [3][128], biases).  Reductions run in a fixed order,
 * so training is bitwise reproducible from the seed.
 * ---------------------------------------------------------------------------------- */
typedef struct vl_nn vl_nn;
vl_status vl_nn_create(int max_batch, uint64_t seed, vl_nn** out);
void vl_nn_destroy(vl_nn* q);
int64_t vl_nn_param_count(void);
vl_status vl_nn_get_params(vl_nn* q, float* out);
vl_status vl_nn_set_params(vl_nn* q, const float* in);
/* masks: n x 64 x 64 (row-major), feat: n scaled log Re, out: n x 3. */
vl_status vl_nn_forward(vl_nn* q, const float* masks, const float* feat, int n, float* out);
/* Mean-squared-error loss of one batch and the gradient of every parameter (no update). */
vl_status vl_nn_gradients(vl_nn* q, const float* masks, const float* feat, const float* targets, int n,
                          float* loss, float* grad);
/* Training data kept on the device; epochs gather their batches there. */
vl_status vl_nn_set_dataset(vl_nn* q, const float* masks, const float* feat, const float* targets,
                            int64_t count);
vl_status vl_nn_train_epoch(vl_nn* q, const int32_t* order, int64_t count, int batch, float lr,
                            float* mean_loss);
vl_status vl_nn_predict_dataset(vl_nn* q, const int32_t* order, int64_t count, float* out);

/* ------------------------------------------------------------------------------------
 * Force-history statistics: src/cases/channel.rs:157-207 and src/signal.rs (host, FP64).
 * ---------------------------------------------------------------------------------- */
typedef struct {
  double cd_mean, cd_rms, cl_mean, cl_rms, st, st_crossings;
  int32_t periodic;
  int32_t pad_;
} vl_stats;
/* Series::stats (channel.rs:182-206) of a history sampled every dt, from time `from`. */
vl_status vl_series_stats(const double* cd, const double* cl, size_t n, double dt,
                          double from, vl_stats* out);
/* signal::dominant_frequency (signal.rs:92-107): out2 = (frequency, amplitude). */
vl_status vl_dominant_frequency(const double* x, size_t n, double dt, double* out2);
/* signal::peaks (signal.rs:109-123): local maxima above the mean with parabolic refinement,
 * (fractional index, value); `count` = all peaks, the first `cap` are written. */
vl_status vl_peaks(const double* x, size_t n, double* pos, double* val, size_t cap, size_t* count);
/* signal::crossing_frequency (signal.rs:127-140); returns 0 when fewer than two crossings. */
vl_status vl_crossing_frequency(const double* x, size_t n, double dt, double* out, int* found);
/* In-place radix-2 FFT (signal.rs:6-43); n must be a power of two. */
vl_status vl_fft(double* re, double* im, size_t n);

#ifdef __cplusplus
}
#endif
#endif /* VORTEXLAB_H */
