/* libdc_hip.so -- entry points beyond the reference's backend ABI.
 *
 * The reference backend ABI (propagate.h, device_data.h) is what the
 * reference's MPI worker calls; everything here is this library's own API for
 * callers that want the MI355X-specific forms: a fused whole-step launch,
 * device-resident time loops with HIP-graph replay, on-device halo packing,
 * FP16 storage, batched shots, and the host-side model setup the reference
 * does in C (src/precomp.c, src/boundary.c, src/calculate_source.c).
 *
 * Errors in these calls, like the reference ABI, print to stderr and exit(1),
 * except where a function is documented to return a status.
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "device_data.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------- kernels ---- */

/* Stencil kernels selectable for dc_propagate and the fused step.
 *  DCX_KERNEL_FAST   2.5-D z-streaming kernel (LDS-staged xy tile with a
 *                    4-cell apron, 8-slot register plane rings, separable
 *                    cross derivatives, fused multiply-add).  Results agree
 *                    with the reference to ~1e-7 relative per step.
 *  DCX_KERNEL_EXACT  one thread per cell, the reference's operation order
 *                    with no contraction: bit-identical to the reference's
 *                    OpenMP path (include/sample_compute.h:105-185). */
enum { DCX_KERNEL_FAST = 0, DCX_KERNEL_EXACT = 1 };

/* Select the kernel for this struct; returns the previous choice. */
int dcx_set_kernel(dc_device_data *data, int kernel);

/* Tile shape of the fast kernel: (tile_x, tile_y) one of (32,16), (32,32),
 * (64,16), (64,8); z_chunk >= 1 planes per workgroup, or 0 (the default) to let each
 * launch pick the longest chunk, from 32 to 512 planes, that gives two tiles
 * per compute unit.  Returns 0, or -1 (and changes nothing) for an
 * unsupported shape. */
int dcx_set_tiling(dc_device_data *data, int tile_x, int tile_y, int z_chunk);

/* Fast kernel prefetch depth: 1 = the staged plane one plane ahead and the
 * coefficients at the top of their step; 2 = the staged plane two planes
 * ahead and the coefficients one step ahead; 3 = as 2 with the coefficients
 * and old fields two steps ahead (32x16 FP32 only; other forms use 2).
 * The default is 3.  Returns the previous depth. */
int dcx_set_prefetch_depth(dc_device_data *data, int planes);

/* Fast kernel: which XCD (and L2) runs which tile.  0: the hardware's round
 * robin over workgroup ids, so x-neighbour tiles run on different XCDs.
 * 1: each XCD takes a contiguous run of tiles.  2: XCD k takes the whole
 * tile rows r = k mod 8 (when the row count is a multiple of 8), so x
 * neighbours share an L2 while the rows in flight stay one band.  -1 (the
 * default): 0 for FP32 fields and coefficients, 2 for every compact form.
 * On the 512^3 block mode 2 is 7% slower than 0 for plain FP32 and 1-5.5%
 * faster for the compact forms; mode 1 is slower for FP32 and the axis form.
 * Results are the same bits in every mode.  Returns the previous mode. */
int dcx_set_xcd_remap(dc_device_data *data, int mode);

/* Wait for all queued work of this struct. */
void dcx_synchronize(dc_device_data *data);

/* Copy one field (0 pp, 1 pc, 2 qp, 3 qc) of the current pointer state to or
 * from host memory (sizes[0]*sizes[1]*sizes[2] floats, x fastest). */
void dcx_download(dc_device_data *data, int which, float *host);
void dcx_upload(dc_device_data *data, int which, const float *host);

/* ----------------------------------------------------- storage forms ---- */

/* Field storage:
 *  DCX_STORE_F32   leapfrog on the FP32 pp/pc/qp/qc arrays (the reference's
 *                  form; the only one dc_propagate accepts).
 *  DCX_STORE_F16   time-difference form: P = p_n and D = p_n - p_{n-1} per
 *                  field, both stored as IEEE half in units of a power-of-two
 *                  field scale (D with its own power-of-two factor); the
 *                  update D += dt^2 rhs, P += D runs in FP32 registers.
 *  DCX_STORE_BF16  the same with bfloat16 storage.
 * Coefficient storage:
 *  DCX_COEF_F32    the ten FP32 arrays.
 *  DCX_COEF_F16    the ten arrays as half, each scaled by a power of two.
 *  DCX_COEF_U16    16-bit codes per array: affine over [min,max] for the six
 *                  angle terms, affine in log2 for the four v^2 terms (code
 *                  0 = exactly zero); ~14x finer than half on v^2.
 *  DCX_COEF_A16    the six angle terms as DCX_COEF_U16 codes them, the four
 *                  v^2 terms in FP32 (FP32 field storage only).
 *  DCX_COEF_N16    the TTI form of the ten arrays (src/precomp.c:85-105):
 *                  the six angle terms are n_i n_j (x2 off the diagonal) of
 *                  the symmetry axis n = (sin t cos f, sin t sin f, cos t),
 *                  and v2px, v2pn, v2sz are v2pz times 1 + 2 eps, 1 + 2 delta
 *                  and (vsv/vpz)^2.  Stored as n's three components and the
 *                  three ratios in 16-bit affine codes beside the FP32 v2pz
 *                  array (16 B per cell instead of 40); the kernel rebuilds
 *                  the ten terms in registers.  FP32 field storage only, and
 *                  only for arrays of that form: dcx_set_storage returns -1
 *                  when a decoded term would differ from the array's value by
 *                  more than 1e-4 (angle terms, absolute) or 1e-4 relative
 *                  (v^2 terms). */
enum { DCX_STORE_F32 = 0, DCX_STORE_F16 = 1, DCX_STORE_BF16 = 2 };
enum { DCX_COEF_F32 = 0, DCX_COEF_F16 = 1, DCX_COEF_U16 = 2, DCX_COEF_A16 = 3,
       DCX_COEF_N16 = 4 };

/* Convert the struct to the given storage (encoding the current fields and
 * coefficients; field_scale_log2 sets the field scale 2^k for half storage).
 * dcx_step / dcx_run then run the matching kernel; dcx_download,
 * dc_device_data_get_results and going back to DCX_STORE_F32 decode.  The
 * fast kernel only.  Returns 0, or -1 for an unknown combination. */
int dcx_set_storage(dc_device_data *data, int field_store, int coef_store,
                    int field_scale_log2);

/* Half field storage is range-checked.  At conversion (dcx_set_storage) the
 * field scale and the time-difference factor are powers of two chosen from
 * the fields' own largest |p|, |q| and |p_n - p_{n-1}| so that they sit at
 * <= 2^12 field units, 16x below half's largest finite value (the caller's
 * field_scale_log2, and 2^3 for the difference, where the fields are still
 * zero).  Every 16-bit store saturates at +-65504 instead of becoming inf or
 * NaN, and sets a device flag when it did (bfloat16: when the value was not
 * finite).  dcx_storage_overflow returns 1 when any store has done so since
 * the last conversion (it waits for the struct's queued work), else 0;
 * dcx_run, dcx_synchronize, dcx_download and dc_device_data_get_results print
 * a one-line warning on stderr the first time it happens.  The half field
 * forms' rounding makes the leapfrog grow slowly (about 2x per 700 steps after
 * ~1500 steps on the octant model), so long runs do reach the range. */
int dcx_storage_overflow(dc_device_data *data);

/* Set both time levels of p and q to zero (a run from rest) in the struct's
 * current storage form, on its stream with a fill kernel (no memset node, so
 * it may be captured in a graph).  Half storage keeps its scales, and its
 * overflow flag is cleared.  Coefficients are untouched. */
void dcx_zero_fields(dc_device_data *data);

/* Bytes per cell-update the fast kernel moves (algorithmic, no apron) for
 * the struct's current storage. */
int dcx_bytes_per_cell(const dc_device_data *data);

/* ------------------------------------------------------ fused steps ---- */

/* One whole time step of one rank with no neighbours, in one launch:
 * update every cell of [4, sizes-4)^3, then swap (pp<->pc, qp<->qc).
 * When src_index >= 0 and add_next is nonzero, src_next is added to the new
 * pc and qc at src_index as the launch writes them -- the same float
 * additions dc_device_add_source would do before the next step, so a run of
 * dcx_step calls equals the reference's loop (src/worker.c:466-511). */
void dcx_step(dc_device_data *data, float dx, float dy, float dz, float dt,
              long long src_index, float src_next, int add_next);

/* n_steps of the reference's one-rank loop: before step it, add src_amp[it]
 * to pc and qc at src_index (skipped when src_index < 0 or src_amp is NULL),
 * update, swap.  With use_graph nonzero the steps are captured once into a
 * HIP graph (two step variants for the two pointer parities) and replayed.
 * Returns the elapsed device time of the steps in milliseconds, measured with
 * HIP events on the struct's stream. */
double dcx_run(dc_device_data *data, float dx, float dy, float dz, float dt,
               unsigned int n_steps, long long src_index,
               const float *src_amp, int use_graph);

/* dcx_run that also samples p at n_rec receiver cells (linear indices
 * x + y*sx + z*sx*sy of the caller's layout) after every step:
 * traces[step * n_rec + r] (n_steps * n_rec floats, host). */
double dcx_run_traces(dc_device_data *data, float dx, float dy, float dz,
                      float dt, unsigned int n_steps, long long src_index,
                      const float *src_amp, int n_rec,
                      const long long *receivers, float *traces,
                      int use_graph);

/* Average duration (ms) of the stencil launches of the last dcx_run, from
 * HIP events recorded around each launch when timing_detail was enabled. */
void dcx_set_timing_detail(dc_device_data *data, int enabled);
double dcx_last_kernel_ms(const dc_device_data *data);

/* Average duration (ms) of `reps` pointwise passes over the compute region
 * that move the time step's bytes without its stencil: each cell reads pc,
 * qc, pp, qp and the ten coefficients and writes pp and qp (64 B per cell,
 * FP32 storage only).  Overwrites pp and qp.  The rate of this pass is the
 * ceiling the memory system sets for a step of the same bytes. */
double dcx_stream_probe(dc_device_data *data, int reps);

/* ------------------------------------------------- halo boxes on device -- */

/* Pack nbox boxes of the device array `from` (a field of data, laid out with
 * the struct's sizes) into the device buffer `dev_buf`, box after box, each x
 * fastest; box k is [starts[3k..3k+2], ends[3k..3k+2]).  One launch. */
void dcx_pack_boxes(dc_device_data *data, float *dev_buf, int nbox,
                    const size_t *starts, const size_t *ends,
                    const float *from);

/* The inverse: scatter dev_buf into the boxes of `to`. */
void dcx_unpack_boxes(dc_device_data *data, const float *dev_buf, int nbox,
                      const size_t *starts, const size_t *ends, float *to);

/* Register a loopback halo exchange that every dcx_step / dcx_run step
 * performs after its stencil launch: box k of the new pp and qp,
 * [send_starts[3k..], send_ends[3k..]), is packed into one device buffer (one
 * launch for all boxes and both fields) and unpacked into
 * [recv_starts[3k..], recv_ends[3k..]) of the same arrays (a second launch).
 * Send and receive boxes must have equal shapes.  This is the on-device half
 * of a multi-GPU halo exchange with the transfer taken out; nbox = 0
 * removes it.  As in the reference's worker (src/worker.c:466-511), the
 * received halos never carry the next step's source: when the source cell
 * lies in a send box, dcx_step / dcx_run add the next amplitude after the
 * exchange (one small launch) instead of inside the stencil launch. */
void dcx_set_halo_loopback(dc_device_data *data, int nbox,
                           const size_t *send_starts, const size_t *send_ends,
                           const size_t *recv_starts, const size_t *recv_ends);

/* dc_device_extract_halo_face (dir = 0: the box of `array` into the host
 * `buffer`) and dc_device_insert_halo_face (dir = 1: `buffer` into the box)
 * without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate (tight exchange loops, harnesses that forbid
 * hipMalloc; it waits for its copy, so it is not for graph capture).  The
 * two reference
 * entry points stage the packed box in a per-struct device buffer they
 * allocate on first use and grow on demand; this one stages it in a fixed
 * 16 MiB array that every device the library runs on holds until the process
 * ends, so a box may hold at most 4,194,304 cells (larger boxes print an
 * error and exit 1).  Same arguments, checks and element order as the
 * reference entry points; returns when the copy is complete.  All calls on
 * one GPU share that array: calls that overlap on one GPU (from several host
 * threads) race. */
void dcx_halo_face_noalloc(dc_device_data *data, float *buffer,
                           const size_t start_coords[3],
                           const size_t end_coords[3], const size_t sizes[3],
                           float *array, int dir);

/* Copy the whole device allocation of field src over that of field dst
 * (0 pp, 1 pc, 2 qp, 3 qc; padding included) with the library's box-copy
 * kernel: a contiguous, aligned streaming copy of known size, used to
 * calibrate HBM traffic counters.  Returns the bytes read (= bytes written). */
size_t dcx_copy_field(dc_device_data *data, int dst, int src);

/* Device scratch buffer owned by the struct (grown on demand, freed with the
 * struct), for the packed halos. */
float *dcx_scratch(dc_device_data *data, size_t n_floats);

/* ------------------------------------------------------ batched shots -- */

/* Many independent one-rank volumes ("shots": a survey's sources, each with
 * its own aperture of the model) advanced together.  Each shot is a
 * dc_device_data of its own (its own sizes, coefficients and fields; one
 * storage form for all; no halo loopback); dcx_batch_create moves them all
 * onto one stream of the batch, dcx_batch_free gives them streams of their
 * own again.  No reference counterpart: the reference runs one shot per
 * program (src/main.c). */
typedef struct dcx_batch dcx_batch;
dcx_batch *dcx_batch_create(dc_device_data **shots, int n);
void dcx_batch_free(dcx_batch *batch);

/* Z chunk of the 32x16-cell tiles, prefetch depth (1, 2 or 3) and
 * persistent workgroups per CU of the batched launch (1..64; 0 = one
 * workgroup per tile).  Returns 0, or -1 for an unsupported value (nothing
 * changed). */
int dcx_batch_set_tiling(dcx_batch *batch, int z_chunk, int prefetch_planes,
                         int blocks_per_cu);

/* enabled != 0: the next dcx_batch_run picks the tiling itself, timing
 * candidates on its own first steps (z chunk 32..512, then prefetch depth
 * 1 or 3 against 2,
 * then workgroups per CU 1, 2, 4 or one per tile; two steps each).  Every
 * tiling computes the same values, so the run's results do not depend on
 * it.  The choice stays set for later runs (dcx_batch_get_tiling). */
int dcx_batch_set_autotune(dcx_batch *batch, int enabled);
/* out = {z chunk, prefetch planes, workgroups per CU (0 = one per tile)}. */
void dcx_batch_get_tiling(const dcx_batch *batch, int out[3]);
/* Device milliseconds per step of the last dcx_batch_run's steps after the
 * autotuning trials. */
double dcx_batch_last_step_ms(const dcx_batch *batch);

/* n_steps full time steps of every shot: per step ONE persistent launch that
 * walks the tiles of all shots.  d4[4k..4k+3] = dx, dy, dz, dt of shot k;
 * src_index[k] = shot k's source cell (linear index of its arrays, < 0 for
 * none; the array may be NULL), src_amp[k * n_steps + it] = its amplitude at
 * step it.  Each shot's fields end as n_steps dcx_step calls on it alone
 * would leave them, bit for bit.  use_graph != 0 captures the loop into one
 * HIP graph.  Returns the elapsed milliseconds on the device. */
double dcx_batch_run(dcx_batch *batch, unsigned int n_steps, const float *d4,
                     const long long *src_index, const float *src_amp,
                     int use_graph);

/* ------------------------------------------- learned absorbing layer -- */

/* A small network on every cell of a thin layer (`width` cells, 1..8) along
 * the six faces of the compute region that corrects the new p and q the
 * stencil writes there, in place of a thick absorbing shell.  No reference
 * counterpart: the reference absorbs with its random-velocity shell
 * (src/boundary.c:7-104), which stays the fallback (set no network).
 *
 * Input of a layer cell (64 values): the 3x3x3 patch of the time
 * differences p^n - p^{n-1} and q^n - q^{n-1} around it (54 values), taken
 * in the frame of its nearest face (axis order: outward normal, then the two
 * other axes in increasing order; p's 27 then q's 27, the last axis
 * fastest) and divided by their largest magnitude s; a one-hot of the
 * cell's depth in the layer (inputs 54..61); the cell's Courant numbers
 * sqrt(v2px) * c and sqrt(v2pz) * c with c = dt * sqrt(1/dx^2) (inputs 62,
 * 63; 1/dx^2 formed in float).  Network: 64 -> 64 (ReLU) ->
 * 64 (ReLU) -> 2, FP16 weights and activations on MFMA, FP32 accumulation.
 * Correction of the new p and q of the cell, by output form
 * (dcx_mlp_set_output):
 *   DCX_MLP_FREE     (dp, dq) = s * y (0 where s = 0);
 *   DCX_MLP_DAMPING  dp = -sigma_max * sigmoid(y0) * (p^n - p^{n-1}) at the
 *                    cell (dq likewise with y1): a learned damping that can
 *                    only take energy out (0 < sigma_max < 1).
 *
 * Parameters (dcx_mlp_nparam() floats): W1 [64][64] (out, in), b1 [64],
 * W2 [64][64], b2 [64], W3 [2][64], b3 [2]. */
typedef struct dcx_mlp dcx_mlp;
typedef struct dcx_samples dcx_samples;

enum { DCX_MLP_FREE = 0, DCX_MLP_DAMPING = 1 };

int dcx_mlp_nparam(void);
/* A network with zero parameters and the free output form. */
dcx_mlp *dcx_mlp_create(int device);
/* Output form and sigma_max (DCX_MLP_DAMPING: 0 < sigma_max < 1); the
 * training loss compares the prediction of that form, divided by s, with
 * the targets.  Returns 0, or -1 (nothing changed). */
int dcx_mlp_set_output(dcx_mlp *mlp, int mode, float sigma_max);
void dcx_mlp_free(dcx_mlp *mlp);
/* Set the parameters (and reset the optimiser state) / read them back. */
void dcx_mlp_set_params(dcx_mlp *mlp, const float *params);
void dcx_mlp_get_params(const dcx_mlp *mlp, float *params);

/* Training samples on the device: inputs (stored as FP16), targets (the
 * correction divided by s) and s. */
dcx_samples *dcx_samples_create(int device, long long capacity);
void dcx_samples_free(dcx_samples *samples);
long long dcx_samples_count(const dcx_samples *samples);
void dcx_samples_clear(dcx_samples *samples);
/* Append n samples given as FP32 inputs [n][64] and targets [n][2] (s = 1). */
void dcx_samples_upload(dcx_samples *samples, long long n, const float *x,
                        const float *t);
/* Copy samples [first, first + n) to the host: inputs [n][64], targets
 * [n][2], s, and the cell (x | y << 10 | z << 20 in compute coordinates of
 * the small struct, 0 for uploaded samples); any pointer may be NULL. */
void dcx_samples_download(const dcx_samples *samples, long long first,
                          long long n, float *x, float *t, float *scale,
                          unsigned int *codes);

/* Raw network outputs y of every stored sample (before the output form):
 * y [count][2]. */
void dcx_mlp_forward(dcx_mlp *mlp, const dcx_samples *samples, float *y);
/* Gradient of the mean over all stored samples of (u0 - t0)^2 + (u1 - t1)^2
 * with respect to the parameters, without a step, where u is the output
 * form's correction divided by s (free: u = y; damping: u = -sigma_max *
 * sigmoid(y) * x_c with x_c the cell's own normalised time difference,
 * input 13 for p, 40 for q); returns that mean loss. */
double dcx_mlp_gradient(dcx_mlp *mlp, const dcx_samples *samples, float *grad);
/* n_iter Adam steps (beta 0.9/0.999, eps 1e-8) on minibatches of `batch`
 * samples drawn uniformly with replacement (seeded); loss_per_iter (may be
 * NULL) receives each minibatch's mean loss.  Returns the mean over the
 * iterations. */
double dcx_mlp_train(dcx_mlp *mlp, const dcx_samples *samples, int n_iter,
                     int batch, float lr, long long seed, float *loss_per_iter);

/* Attach the network to a struct (FP32 storage): every dcx_step / dcx_run
 * step then corrects its layer cells.  NULL or width 0 detaches it.  Returns
 * the number of layer cells. */
int dcx_set_learned_layer(dc_device_data *data, dcx_mlp *mlp, int width);

/* The frame condition of a whole-domain run.  DCX_WALL_ZERO (the default)
 * leaves the 4-cell frame beyond the compute region as it is: zero, the
 * rigid wall of the reference's outer frame (src/boundary.c:7-104), unless
 * a wall has filled it (switching back keeps the filled values;
 * dcx_zero_fields restarts from rest).
 * DCX_WALL_ONE_WAY: after each dcx_step / dcx_run step (after the learned
 * layer's corrections) the frame cells of the new p and q are filled from
 * the outgoing wave by the first-order one-way condition, outward one frame
 * layer at a time:
 *   u_F^{n+1} = u_I^n + r (u_I^{n+1} - u_F^n),  r = (c dt - h)/(c dt + h),
 * I being F's neighbour towards the compute region, h the spacing along the
 * face normal and c^2 = v2px (1 - c_nn) + v2pz c_nn at the last compute cell
 * (c_nn = cxx, cyy or czz for x, y or z faces).  DCX_WALL_ONE_WAY2: the
 * second-order condition, the first-order operator applied twice, as the
 * first-order condition on the first-order residual; with
 * pred(a, b, c) = a + r (b - c) and v_G^m = u_G^m - pred(u_{G-1}^{m-1},
 * u_{G-1}^m, u_G^{m-1}) (G-1 one cell towards the compute region):
 *   u_F^{n+1} = pred(u_I^n, u_I^{n+1}, u_F^n) + pred(v_I^n, v_I^{n+1}, v_F^n).
 * It keeps level n-1 of the last two compute layers in a buffer of the
 * struct.  c is computed at every face cell when the wall is set (the
 * coefficients are read-only afterwards; dcx_set_storage is refused while a
 * one-way wall is set).
 * x faces first (compute range of y and z), then y faces (all x), then z
 * faces (all x and y).  No reference counterpart.  FP32 storage, no
 * registered halo loopback, not a batch shot, at least two compute cells
 * per axis.  Returns 0. */
enum { DCX_WALL_ZERO = 0, DCX_WALL_ONE_WAY = 1, DCX_WALL_ONE_WAY2 = 2 };
int dcx_set_wall(dc_device_data *data, int mode);

/* One training-data step.  `small` is a window of `padded`: its compute
 * region starts at padded cell origin[3] (>= 4 per axis).
 * flags: DCX_SAMPLE_ROLLOUT and/or DCX_SAMPLE_TAPER.
 * Without DCX_SAMPLE_ROLLOUT (teacher forcing):
 *  1. copies the padded fields over small's compute region (small's 4-cell
 *     frame stays zero: small has no cells beyond its edge),
 *  2. stores the inputs of n_cells random cells of small's layer,
 *  3. steps small once without a source and without swapping;
 * with DCX_SAMPLE_TAPER the copied layer cells are first multiplied by the
 * taper a(d) = sin^2(pi (d + 1) / (2 (width + 1))) of their depth d.
 * With DCX_SAMPLE_ROLLOUT (small runs on its own state, e.g. with the
 * network attached):
 *  2. stores the inputs of n_cells random layer cells of small's fields,
 *  3. runs small's whole step (dcx_step, with padded's source moved into its
 *     coordinates when inside), noting its stencil's new p, q before the
 *     learned corrections;
 * then in both:
 *  4. runs dcx_step on padded (with its source arguments),
 *  5. stores target = (a * padded's new p, q - small's stencil's new p, q) / s
 * at those cells (a = the taper with DCX_SAMPLE_TAPER, else 1): the
 * correction that makes small's step match padded's (tapered) one.
 * Returns the new sample count. */
enum { DCX_SAMPLE_ROLLOUT = 1, DCX_SAMPLE_TAPER = 2 };
long long dcx_learned_sample_step(dc_device_data *padded, dc_device_data *small,
                                  const long long origin[3], int width,
                                  int n_cells, unsigned int seed,
                                  dcx_samples *samples, float dx, float dy,
                                  float dz, float dt, long long src_index,
                                  float src_next, int add_next, int flags);

/* ---------------------------------------------- whole program, 1 GPU ---- */

/* The reference program (src/main.c) on one GPU: global grid
 * size + 2*absorption + 2*STENCIL per axis, the default TTI model with the
 * random-velocity boundary, a Ricker source at the grid centre, and
 * ceil(time_max/dt) steps (float division, src/main.c:132).  The grid is cut
 * into topology[0]*topology[1]*topology[2] subdomains exactly as the
 * reference's coordinator cuts it for that many ranks
 * (src/coordinator.c:19-114); all of them live on the one GPU and exchange
 * their 4-cell halos by on-device packing and unpacking each step, in the
 * reference's step order (src/worker.c:466-511).
 * out_pc/out_qc (may be NULL) receive the global fields in the layout of the
 * reference's output file (src/coordinator.c:116-219: only the compute
 * region written, the 4-cell frame zero).  Returns 0 on success.
 * stats[0] = elapsed seconds of the time loop, stats[1] = iterations,
 * stats[2] = Mcell-updates/s over the global compute region. */
typedef struct {
  size_t size_x, size_y, size_z;
  float dx, dy, dz, dt, time_max;
  size_t absorption;
  int topology[3];
  int kernel; /* DCX_KERNEL_* */
} dcx_program_args;

int dcx_program_run(const dcx_program_args *args, float *out_pc,
                    float *out_qc, double stats[3]);

/* The rank topology the reference program takes for n ranks: src/main.c:89-90
 * calls MPI_Dims_create(n, 3, dims) with dims zeroed.  Of the factorisations
 * n = dims[0]*dims[1]*dims[2] with dims non-increasing, the one with the least
 * dims[0] - dims[2], and of equal spreads the larger dims[2] (MPICH's
 * result).  Host only, no device work.  Returns 0, or -1 for n < 1. */
int dcx_dims_create(int n, int dims[3]);

/* Write the two global fields as the reference's output file: pc then qc,
 * float32, x fastest (src/coordinator.c:116-219).  Returns 0 on success. */
int dcx_write_dc_file(const char *path, size_t gx, size_t gy, size_t gz,
                      const float *pc, const float *qc);

/* ---------------------------------------------------- host model setup -- */

/* src/precomp.c:156-169: the reference's default homogeneous model. */
void dcx_default_model(size_t n, float *vpz, float *vsv, float *epsilon,
                       float *delta, float *phi, float *theta);

/* src/precomp.c:85-105: the ten coefficient arrays in dc_precomp_vars order
 * (trig in double, rounded to float before the products). */
void dcx_precompute(size_t n, const float *vpz, const float *vsv,
                    const float *epsilon, const float *delta, const float *phi,
                    const float *theta, float *const out[10]);

/* dc_device_data_init with the reference worker's model setup
 * (src/worker.c:49-156) computed on the device instead of passed in: the
 * default TTI model (src/precomp.c:156-169) in every local cell, the
 * random-velocity shell (src/boundary.c:7-104) over the rank's box of the
 * global grid, and the ten coefficient arrays (src/precomp.c:85-105), one
 * thread per local cell, each shell cell finding its own rand_r draw in
 * closed form instead of the reference's walk over the whole global grid.
 * The default model's trigonometry is one (theta, phi) pair, done in FP64 on
 * the host.  process->sizes gives the local grid (halo included) and
 * process->pp/pc/qp/qc the initial fields (NULL: zero); precomp_vars and
 * anisotropy_vars are not read.  global_sizes, start_coords (global
 * coordinates of local cell 0) and problem_sizes are the arguments the
 * reference worker passes to randomVelocityBoundaryPartition with
 * bord = STENCIL (src/worker.c:94-103); *seed is the generator state before
 * the walk (0 in the reference) and is left where the walk leaves it.  The
 * coefficient arrays equal, bit for bit, those dc_device_data_init would
 * upload from the reference's host setup. */
dc_device_data *dcx_device_data_init_model(dc_process_t *process,
                                           const int global_sizes[3],
                                           const int start_coords[3],
                                           const int problem_sizes[3],
                                           int absorption, unsigned int *seed);

/* Copy coefficient array k (dc_precomp_vars order, 0 = ch1dxx ... 9 = v2pn,
 * FP32) of the struct to host memory (sizes[0]*sizes[1]*sizes[2] floats, x
 * fastest). */
void dcx_download_coef(dc_device_data *data, int k, float *host);

/* src/boundary.c:7-104: random-velocity boundary over one local box. */
void dcx_random_boundary(int local_sx, int local_sy, int local_sz,
                         int global_sx, int global_sy, int global_sz,
                         int start_x, int start_y, int start_z, int nx, int ny,
                         int nz, int bord, int absorb, float *vpz, float *vsv,
                         unsigned int *seed);

/* src/calculate_source.c:10-17 with the cut-off frequency as a parameter
 * (the reference uses fcut = 40; a Ricker of peak frequency f has
 * fcut = 3f). */
float dcx_ricker(float dt, int it, double fcut);

#ifdef __cplusplus
}
#endif
