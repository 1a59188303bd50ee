/*
 * lbm_hip.h — C ABI of the MI355X-native D2Q9 lattice Boltzmann step.
 *
 * The library replaces the simulation core of the lbm-wgpu crate (the Rust type
 * `LBM` in lbm-wgpu/src/lbm.rs and its eight WGSL compute passes) with one fused
 * HIP pull-stream kernel per step on a gfx950 GPU.  Each entry point below names
 * the reference method it replaces.  Plain C types only: no HIP or torch types
 * cross this boundary, and no C++ exception escapes it.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative LBM_E* code; the text of
 *     the last error on a grid is returned by lbm_last_error(grid) (or
 *     lbm_last_error(NULL) for a failed lbm_create).
 *   - A grid owns its device memory and one HIP stream.  Host pointers are
 *     borrowed for the duration of the call only.  Calls on one grid must be
 *     serialised by the caller; distinct grids may be used from distinct threads.
 *   - Cells are indexed x + y * nx (row-major, row y = 0 at the top), like the
 *     reference (lbm-wgpu/src/lbm.rs:1517-1519).  Host planes are dense nx * ny.
 *   - Populations are numbered in the reference's plane order
 *     k = 0 NW, 1 N, 2 NE, 3 W, 4 C (rest), 5 E, 6 SW, 7 S, 8 SE
 *     (lbm-wgpu/src/lbm.rs:632-640); N moves to row y-1.
 *   - jx/jy are momenta; jy is north-positive (N - S), as in the reference's
 *     density bind group (rewritten_shaders/pre_collision/{corner,cardinal}_pre_collision.wgsl).
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
#define LBM_OK 0
#define LBM_EINVAL (-1)     /* invalid argument (size, index, enum, null pointer) */
#define LBM_EHIP (-2)       /* a HIP runtime call failed */
#define LBM_ENOMEM (-3)     /* device or host allocation failed */
#define LBM_ENONFINITE (-4) /* a health check found NaN or Inf */
#define LBM_ENODEV (-5)     /* no usable gfx950 device */

/* ---- enums ----------------------------------------------------------------- */
/* storage of the distributions; arithmetic is always f32 */
#define LBM_PRECISION_F32 0  /* 9 x f32 per cell, 72 B per cell update */
#define LBM_PRECISION_F16S 1 /* f_k - w_k as binary16 (round to nearest even), 36 B */
/* f_k - feq_k(1, (u_in, 0)) as binary16, 36 B: shifted by the free-stream
 * equilibrium of the inflow speed u_in of the last uniform reset (lbm_create's
 * u_init, lbm_reset_equilibrium's ux), so a flow near the free stream stores
 * small values and loses less to binary16 rounding.  The same kernels as F16S;
 * a reset to another speed re-centres the storage (both sets are refilled). */
#define LBM_PRECISION_F16E 2

/* outlet treatment */
#define LBM_BOUNDARY_REFERENCE 0     /* the reference's flattened-index wrap at x = nx-1 */
#define LBM_BOUNDARY_ZERO_GRADIENT 1 /* zero-gradient outflow at x = nx-1 */

/* summary fields, in the order of the reference enum SummaryStat (lbm.rs:10-16) */
#define LBM_SUMMARY_CURL 0
#define LBM_SUMMARY_UX 1
#define LBM_SUMMARY_UY 2
#define LBM_SUMMARY_RHO 3
#define LBM_SUMMARY_SPEED 4

/* colour maps, in the order of the reference's ColorMap enum (lbm.rs:20-24) */
#define LBM_COLORMAP_INFERNO 0
#define LBM_COLORMAP_VIRIDIS 1
#define LBM_COLORMAP_JET 2

typedef struct lbm_grid lbm_grid;

typedef struct lbm_config {
    uint32_t nx, ny;      /* grid size; nx, ny >= 3 */
    float omega;          /* BGK relaxation rate, 1/tau */
    float u_init;         /* initial uniform x velocity (LBM::new uses 0.1, lbm.rs:740) */
    int32_t precision;    /* LBM_PRECISION_* */
    int32_t boundary;     /* LBM_BOUNDARY_* */
    int32_t device;       /* HIP device ordinal */
    uint32_t row_pad;     /* extra elements per row beyond the 64-aligned width (0 = default) */
} lbm_config;

/* ---- lifecycle --------------------------------------------------------------- */

/* Fill *cfg with defaults: omega 1/(0.5 + 0.3) (lib.rs:39), u_init 0.1, f32,
 * reference boundary, device 0. */
int lbm_config_default(lbm_config *cfg, uint32_t nx, uint32_t ny);

/* Replaces LBM::new (lbm.rs:726-1049): allocates the two population sets and the
 * mask, fills both sets with feq(rho 1, u (u_init, 0)) (lbm.rs:740-745), paints
 * rows 0 and ny-1 solid (init_barrier, lbm.rs:595-605), zeroes the fields and
 * the step counter. */
int lbm_create(const lbm_config *cfg, lbm_grid **out);
int lbm_destroy(lbm_grid *g);
const char *lbm_last_error(const lbm_grid *g);

/* ---- parameters and resets --------------------------------------------------- */

/* Replaces LBM::update_omega_buffer (lbm.rs:1358-1360). */
int lbm_set_omega(lbm_grid *g, float omega);
/* omega = 1 / (3 nu + 0.5), as the viscosity slider does (lib.rs:181-186). */
int lbm_set_viscosity(lbm_grid *g, float nu);
int lbm_get_omega(const lbm_grid *g, float *omega);

/* Replaces LBM::reset_to_equilibrium / LBM::custom_speed (lbm.rs:1076-1102):
 * both sets := feq(rho, (ux, uy)) computed like set_equil (lbm.rs:611-643), the
 * step counter := 0, and the fields := the pre-collision moments of that state
 * (which, as in the reference, leave the rest population out of rho). */
int lbm_reset_equilibrium(lbm_grid *g, float rho, float ux, float uy);

/* Replaces LBM::single_cell (lbm.rs:1482-1515): both sets := feq(1, 0) with the
 * value 4.0 in plane k at the reference's fixed cell for k; step counter := 0;
 * the fields are left as they are.  k outside 0..8 leaves a plain rest state. */
int lbm_single_cell(lbm_grid *g, int k);

/* Set population k of the current state at (x, y) to v (no reference
 * counterpart; a general form of single_cell's pulse). */
int lbm_set_pulse(lbm_grid *g, int k, uint32_t x, uint32_t y, float v);

/* ---- obstacles ---------------------------------------------------------------- */

/* Replaces LBM::reset_barrier (lbm.rs:1362-1365): only rows 0 and ny-1 solid. */
int lbm_reset_barrier(lbm_grid *g);

/* Replaces LBM::draw_shape / draw_barrier_updates (lbm.rs:1337-1356) with the
 * same wire format as get_points_vector (barrier_shapes/merge_shapes.rs:12-22):
 * n_u32 words [idx0, val0, idx1, val1, ...], idx = x + y * nx, and a cell is solid
 * iff its value is 1 (stream/e_w_stream.wgsl:31).  Pairs apply in list order, so the
 * last pair for a cell wins.  An odd length or an index >= nx*ny is LBM_EINVAL
 * and changes nothing; an empty list is a no-op (the reference underflows on it,
 * lbm.rs:1343).  Populations are not touched. */
int lbm_paint(lbm_grid *g, const uint32_t *pairs, size_t n_u32);

/* Whole-mask upload / download, one byte per cell, 1 = solid. */
int lbm_set_mask(lbm_grid *g, const uint8_t *solid);
int lbm_get_mask(lbm_grid *g, uint8_t *solid);

/* ---- stepping ------------------------------------------------------------------ */

/* Replaces LBM::iterate's loop over compute_step (lbm.rs:1065-1068, 1112-1134):
 * enqueues n steps on the grid's stream and returns without waiting.  The
 * fields of the last step (the reference's density bind group) are written by
 * the last launch. */
int lbm_step(lbm_grid *g, uint32_t n);
/* Blocks until the grid's stream is idle. */
int lbm_sync(lbm_grid *g);

/* Step-kernel variants (no reference counterpart): how the fused step moves
 * memory -- cells per thread, non-temporal access, block order, or (8, 9) two
 * steps per launch through LDS, which reads and writes the state once per two
 * steps (row blocks of 8 or 16).  All variants compute the same bits.  The two-step variants need two
 * more state sets: lbm_set_variant allocates them when it selects one (and
 * fails, changing nothing, when it cannot) and frees them when it selects
 * another.  lbm_autotune times each variant for `steps` steps on the grid's
 * own state, restores that state exactly, selects the fastest, and returns it
 * in *best and the per-step time of every variant in ms[LBM_NUM_VARIANTS]
 * (either pointer may be NULL; +inf for a two-step variant whose sets could
 * not be allocated). */
#define LBM_NUM_VARIANTS 10
int lbm_set_variant(lbm_grid *g, int variant);
int lbm_get_variant(const lbm_grid *g, int *variant);
int lbm_autotune(lbm_grid *g, uint32_t steps, int *best, float *ms);
/* lbm_autotune through a JSON file (no reference counterpart): the file maps
 * (device architecture and CU count, shape class = the grid's width and height
 * rounded up to powers of two, precision, boundary, number of variants) to the
 * variant that was fastest.  A key found in the file selects its variant
 * without timing (*hit = 1); otherwise the call tunes as lbm_autotune and
 * adds the key (*hit = 0), rewriting the file through a temporary file and a
 * rename.  A cached two-step variant that cannot get its spare sets counts as
 * a miss.  LBM_EINVAL when the file cannot be written (the tuned variant is
 * selected all the same).  best and hit may be NULL. */
int lbm_autotune_cached(lbm_grid *g, const char *cache_path, uint32_t steps, int *best, int *hit);
/* lbm_step(g, n) bracketed by two hipEvents recorded on the grid's stream;
 * waits for the second and returns the elapsed device time in *ms.  (No
 * reference counterpart: the reference has no timing, lbm.rs:1166-1172.) */
int lbm_step_timed(lbm_grid *g, uint32_t n, float *ms);
/* Replaces LBM::get_compute_num (lbm.rs:1170-1172). */
int lbm_step_count(const lbm_grid *g, uint64_t *n);
/* The grid's hipStream_t, as an opaque pointer. */
int lbm_get_stream(const lbm_grid *g, void **stream);

/* ---- fields ---------------------------------------------------------------------- */

/* Momentum and density planes.  lagged != 0: the renderer's view, i.e. the
 * pre-collision moments computed in the last executed step (one step behind the
 * state; lbm.rs:1065-1074).  lagged == 0: the moments of the current state. */
int lbm_read_moments(lbm_grid *g, float *jx, float *jy, float *rho, int lagged);

/* rho, velocity u = j / rho and the reference's curl
 * 10 (jy[x+1] - jy[x-1] - jx[y-1] + jx[y+1]) / rho (summary_stats/curl.wgsl:49),
 * which is 0 where the reference skips it (x = 0, y = ny-1).  Any pointer may be
 * NULL. */
int lbm_read_fields(lbm_grid *g, float *rho, float *ux, float *uy, float *curl, int lagged);

/* Replaces LBM::calculate_summary (lbm.rs:1051-1059): one renderer field into a
 * persistent output plane that, like the reference's, keeps values the pass
 * skips; kind is LBM_SUMMARY_*. */
int lbm_summary(lbm_grid *g, int kind, float *out);
/* Replaces calculate_summary + color_map (lbm.rs:1051-1059, lbm.rs:1299-1305;
 * color_map/{inferno,viridis,jet}.wgsl), fused in one pass: the summary plane
 * of `kind` (kept as lbm_summary keeps it) mapped to colours, rgb = nx*ny*3
 * floats (R, G, B per cell, rows of nx cells, row 0 first), solid cells black.
 * The reference renders the same colours one quad per cell. */
int lbm_render(lbm_grid *g, int kind, int colormap, float *rgb);

/* FP64 sums of rho, jx, jy over the fluid cells (mask != 1) of the current
 * state.  Returns LBM_ENONFINITE if any sum is NaN or Inf. */
int lbm_conservation(lbm_grid *g, double out[3]);

/* lbm_conservation without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate -- code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc -- where lbm_conservation
 * would allocate its moment planes and partial-sum buffer on first use.  The
 * partial sums go to a fixed device array of the library, 2^18 reduction
 * blocks (6 MB) per device, which holds grids whose moments pass has at most
 * that many blocks (f32: ceil(nx / 256) * ny; binary16: ceil(nx / 512) * ny
 * once stepped); a larger grid is LBM_EINVAL.  Same pass and order as
 * lbm_conservation: the same bits.  Every call on one device shares the array:
 * calls that overlap on one device (grids on several threads) race. */
int lbm_conservation_static(lbm_grid *g, double out[3]);

/* ---- state ------------------------------------------------------------------------- */

/* The current state S_n as 9 dense f32 planes (9 * nx * ny floats): the
 * reference's planes k != 4 of set n % 2 plus the rest plane (lbm.rs:775). */
int lbm_read_f(lbm_grid *g, float *f9);
/* Write S_n := f9 into both sets (as the reference's resets do); the step count
 * and the fields are left as they are. */
int lbm_write_f(lbm_grid *g, const float *f9);

/* Checkpoint: an opaque, self-describing byte image of the whole grid state
 * (both sets, mask, fields, omega, step count).  Loading it into a grid of the
 * same size, precision and boundary mode continues bit for bit. */
int lbm_state_size(const lbm_grid *g, size_t *bytes);
int lbm_save_state(lbm_grid *g, void *buf, size_t bytes);
int lbm_load_state(lbm_grid *g, const void *buf, size_t bytes);

/* ---- batched domains (no reference counterpart: the reference steps one grid) -------- */

/* Many independent grids of uneven size stepped together: each domain follows
 * the single-grid semantics above (walls on rows 0 and ny-1, equilibrium inlet,
 * the selected outlet) with its own omega and inflow speed, and equals a single
 * grid of the same size and mask bit for bit.  All domains share one precision,
 * one boundary mode and one step count.  A step launches one interior and one
 * boundary kernel for all domains; lbm_batch_step replays them from a hipGraph
 * two steps at a time. */
typedef struct lbm_batch lbm_batch;

typedef struct lbm_domain {
    uint32_t nx, ny;
    float omega;  /* 1/tau */
    float u_in;   /* uniform x velocity of the initial / inflow equilibrium */
} lbm_domain;

#define LBM_BATCH_NUM_VARIANTS 8

int lbm_batch_create(const lbm_domain *domains, uint32_t n, int32_t precision, int32_t boundary,
                     int32_t device, lbm_batch **out);
int lbm_batch_destroy(lbm_batch *b);
const char *lbm_batch_last_error(const lbm_batch *b);
/* every domain := feq(1, (u_in, 0)) in both sets; step count := 0 */
int lbm_batch_reset(lbm_batch *b);
/* whole-domain mask, one byte per cell, 1 = solid (like lbm_set_mask) */
int lbm_batch_set_mask(lbm_batch *b, uint32_t domain, const uint8_t *solid);
int lbm_batch_step(lbm_batch *b, uint32_t n);
int lbm_batch_step_timed(lbm_batch *b, uint32_t n, float *ms);
int lbm_batch_sync(lbm_batch *b);
int lbm_batch_step_count(const lbm_batch *b, uint64_t *n);
/* total cells, interior tiles and boundary segments of the current work lists */
int lbm_batch_info(lbm_batch *b, uint64_t *cells, uint32_t *interior_tiles, uint32_t *boundary_segments);
/* hipGraph replay on (default) or off */
int lbm_batch_set_graph(lbm_batch *b, int enable);
/* Batch variants: 0..5 the interior kernel's tile (cells per thread 1, 2, 4
 * -- 2, 4, 8 for binary16 -- each in list order or chunked); 6, 7 two steps
 * per launch for every domain (row chunks of 8 / 16), which needs two more
 * state sets: lbm_batch_set_variant allocates them (and fails, changing
 * nothing, when it cannot) and frees them when another variant is selected.
 * All variants give the same bits.  lbm_batch_autotune as lbm_autotune (+inf
 * for a two-step variant without memory or with steps < 3). */
int lbm_batch_set_variant(lbm_batch *b, int variant);
int lbm_batch_get_variant(const lbm_batch *b, int *variant);
int lbm_batch_autotune(lbm_batch *b, uint32_t steps, int *best, float *ms);
/* as lbm_autotune_cached; the shape class is the domain count and the total
 * cells in MiB-cells, each rounded up to a power of two */
int lbm_batch_autotune_cached(lbm_batch *b, const char *cache_path, uint32_t steps, int *best, int *hit);
/* one domain's current state (9 dense nx*ny f32 planes) and renderer (lagged) moments */
int lbm_batch_read_f(lbm_batch *b, uint32_t domain, float *f9);
int lbm_batch_read_moments(lbm_batch *b, uint32_t domain, float *jx, float *jy, float *rho);

/* ---- FP64 solver (no reference counterpart) ------------------------------------------ */

/* The textbook FP64 D2Q9 scheme (BGK collide of every cell, then a pull stream
 * with half-way bounce-back) with the reference step's boundary rules, on the
 * GPU: the yardstick for the error of the f32 and FP16-shifted grids at full
 * size.  Starts at feq(1, (u_in, 0)) with walls on rows 0 and ny - 1. */
typedef struct lbm_f64 lbm_f64;
const char *lbm_f64_last_error(void);
int lbm_f64_create(uint32_t nx, uint32_t ny, double omega, double u_in, int32_t boundary, int32_t device,
                   lbm_f64 **out);
int lbm_f64_destroy(lbm_f64 *h);
/* one byte per cell, 1 = solid */
int lbm_f64_set_mask(lbm_f64 *h, const uint8_t *solid);
int lbm_f64_step(lbm_f64 *h, uint32_t n);
int lbm_f64_sync(lbm_f64 *h);
int lbm_f64_step_count(const lbm_f64 *h, uint64_t *n);
/* rho, ux, uy (north-positive) of the current state, nx*ny doubles each (any may be NULL) */
int lbm_f64_macro(lbm_f64 *h, double *rho, double *ux, double *uy);
/* the current (pre-collision) state, 9 dense nx*ny planes */
int lbm_f64_read_f(lbm_f64 *h, double *f9);

/* ---- learned surrogate (no reference counterpart) --------------------------------- */

/* Text of the last error of an lbm_unet_* / lbm_conv3x3_test call on this thread. */
const char *lbm_unet_last_error(void);

/* A 4-level U-Net (3x3 convs, 32-64-128-256 channels, nearest-upsample + 1x1
 * convs and skip concatenation on the way up, 1x1 head) with BF16 weights and
 * activations and FP32 accumulation on the matrix cores, for `batch` images of
 * height x width (height % 8 == 0, width % 256 == 0).  Input per pixel (NHWC,
 * 8 f32 on the host): solid fraction, (rho - 1) / s_rho, ux / s_u, uy / s_u,
 * (tau - 1/2) / 0.2, u_in / s_u, 0, 0.  Output: the change of (rho, ux, uy) over
 * one jump in units d_rho, d_u.  Parameters are f32 masters (weights
 * [taps][cout][cin], then the bias, per layer). */
typedef struct lbm_unet lbm_unet;
int lbm_unet_create(int batch, int height, int width, uint64_t seed, int device, lbm_unet **out);
int lbm_unet_destroy(lbm_unet *net);
int lbm_unet_num_params(const lbm_unet *net, size_t *count);
int lbm_unet_get_params(lbm_unet *net, float *params);
int lbm_unet_set_params(lbm_unet *net, const float *params);
/* x: batch*height*width*8 floats; y: batch*height*width*3 floats */
int lbm_unet_forward(lbm_unet *net, const float *x, float *y);
/* masked mean-square loss sum_p w_p |y_p - target_p|^2 / sum_p w_p and its
 * gradient with respect to every parameter (grad may be NULL) */
int lbm_unet_loss_grad(lbm_unet *net, const float *x, const float *target, const float *weights, float *loss,
                       float *grad);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss; *loss = the loss before the step */
int lbm_unet_train_step(lbm_unet *net, const float *x, const float *target, const float *weights, float lr,
                        float *loss);
/* device time of one forward pass (ms), averaged over reps */
int lbm_unet_time_forward(lbm_unet *net, int reps, float *ms);
/* lbm_surrogate_jump replays the forward pass from a hipGraph captured on the
 * grid's stream (default on); off: plain launches.  Results are identical. */
int lbm_unet_set_graph(lbm_unet *net, int enable);

/* Surrogate jump on a grid (same device): the network sees the current state
 * with rho - 1 and u in units s_rho, s_u and predicts the change of rho and u
 * over `steps` steps in units d_rho, d_u; every streamed fluid cell (solid
 * cells, the inlet column and the last row keep their populations) gets the
 * equilibrium of its own rho, u plus that change, and the step count advances
 * by `steps`.  Net shape (1, ny, nx): one network
 * pixel per cell.  Net shape (1, ny / 2, nx / 2): each pixel sees a 2 x 2 block
 * (solid fraction, fluid means of rho and u) and the change is bilinearly
 * interpolated back to the cells. */
int lbm_surrogate_jump(lbm_grid *g, lbm_unet *net, uint32_t steps, float s_rho, float s_u, float d_rho,
                       float d_u);
/* Device-side copy of the whole evolving state (one slot) and its restoration:
 * the rollout's fallback. */
int lbm_snapshot(lbm_grid *g);
int lbm_restore(lbm_grid *g);
/* Velocity residual: mark the current velocity field, then later
 * rel = |u - u_mark| / |u_mark| over fluid cells (FP64 sums on the device);
 * rel is +inf when either field holds a NaN or Inf. */
int lbm_velocity_mark(lbm_grid *g);
int lbm_velocity_change(lbm_grid *g, double *rel);
/* The rollout's check.  lbm_velocity_mark_slot(g, s) marks the current
 * velocity field in slot s (0 or 1; lbm_velocity_mark is slot 0).  With slot 0
 * holding the field at a cycle's start and slot 1 the surrogate's forecast for
 * its end, and the grid now at the plain LBM's end of the same cycle:
 * out[0] = |u - m1| / |u - m0| (the forecast's error over that of persistence:
 * 0 perfect, 1 no better than predicting no change), out[1] = |u - m1| / |m1|
 * (its relative error), over fluid cells with FP64 sums on the device; +inf
 * when a field holds a NaN or Inf. */
int lbm_velocity_mark_slot(lbm_grid *g, int slot);
int lbm_velocity_skill(lbm_grid *g, double out[2]);
/* The surrogate's 3x3 convolution primitives on host arrays, for testing: NHWC
 * f32 tensors rounded to bf16 on the device, f32 accumulation on the matrix
 * cores.  w is [9][Cout][Cin] (tap = 3 dy + dx).  mode 0: out (N,H,W,Cout) =
 * conv(in) + bias (bias may be NULL); mode 1: out (N,H,W,Cin) = input gradient
 * for the output gradient `in` (N,H,W,Cout); mode 2: out [9][Cout][Cin] = weight
 * gradient for input `in` and output gradient `g`.  W, Cin, Cout must be
 * multiples of 32.  mode 3: out[0] = milliseconds per launch of the mode-0
 * convolution (bf16 output, ReLU), averaged over 20 launches.  Bits 4..7 of
 * mode force the block shape of modes 0, 1 and 3 (0 = chosen by size; 1..13 =
 * (rows per wave, 32-channel tiles per wave, waves along channels) of
 * (1,1,1) (2,1,1) (4,1,1) (1,2,1) (2,2,1) (1,1,2) (2,1,2) (1,2,2) (2,2,2)
 * (1,1,4) (2,1,4) (1,2,4) (2,2,4)), bits 8..10 the split of the input channels
 * over blocks (0 = chosen by size, or 1, 2, 4; at most the number of 32-channel
 * input chunks), bit 11 the kernel that stages the weights in LDS (block shapes
 * with at most 64 output channels; a forced shape otherwise runs the kernel
 * that reads weights per wave), bit 13 that kernel with 64-channel input
 * chunks (block shape (1,1,1) only, input channels a multiple of 64).  Bit 12
 * with mode 0: a bf16 output and its
 * 2 x 2 average pool, returned as out[0, N H W Cout) and out[N H W Cout,
 * + N H W Cout / 4) (H even). */
int lbm_conv3x3_test(int mode, int n, int h, int w, int cin, int cout, const float *in, const float *g,
                     const float *weights, const float *bias, float *out, int device);

/* ---- seeded obstacle generator (host only, no GPU) --------------------------------- */

/* SplitMix64 (Steele, Lea and Flood 2014).  Returns the next 64-bit output and
 * advances *state. */
uint64_t lbm_splitmix64(uint64_t *state);

/* Paint n_circles filled circles into solid[] (nx * ny bytes; existing bytes are
 * kept, so callers pass a zeroed or wall-painted array).  For each circle, in
 * this order: r = r_min + u(r_max - r_min + 1), cx = r + 1 + u(nx - r - 1),
 * cy = u(ny), with u(m) = (splitmix64() >> 32) * m >> 32.  Cells with
 * (x - cx)^2 + (y - cy)^2 <= r^2 become 1; circles are clipped at the grid edge
 * and keep column 0 free (cx - r >= 1). */
int lbm_gen_circles(uint8_t *solid, uint32_t nx, uint32_t ny, uint64_t seed,
                    uint32_t n_circles, uint32_t r_min, uint32_t r_max);

/* n batch domains from one seed: nx = 128 + 32 u(29), ny = 64 + 32 u(15),
 * tau = 0.52 + 0.28 d, u_in = 0.02 + 0.08 d, drawn in that order per domain, with
 * d = (splitmix64() >> 11) * 2^-53 and u as for lbm_gen_circles; omega = 1/tau. */
int lbm_gen_batch(uint64_t seed, uint32_t n, lbm_domain *out);

/* n_min + u(n_max - n_min + 1) shapes into solid[] (nx, ny >= 16), each a circle
 * (u(2) == 0: r = 3 + u(min(nx, ny) / 8), cx = r + 1 + u(nx - r - 1), cy = u(ny)) or an
 * axis-aligned rectangle (w = 2 + u(nx / 8), h = 2 + u(ny / 4), x0 = 1 + u(nx - w - 1),
 * y0 = u(ny - h)); column 0 stays free. */
int lbm_gen_shapes(uint8_t *solid, uint32_t nx, uint32_t ny, uint64_t seed, uint32_t n_min, uint32_t n_max);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
