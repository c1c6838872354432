// Internal declarations shared by the HIP translation units of libdc_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "dc_hip.h"
#include "device_data.h"

#define DCX_CHECK(expr)                                                        \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "[dc_hip] HIP error %s (%s) at %s:%d: %s\n",        \
                   hipGetErrorName(e_), hipGetErrorString(e_), __FILE__,       \
                   __LINE__, #expr);                                           \
      std::exit(1);                                                            \
    }                                                                          \
  } while (0)

#define DCX_FAIL(...)                                                          \
  do {                                                                         \
    std::fprintf(stderr, "[dc_hip] error: ");                                  \
    std::fprintf(stderr, __VA_ARGS__);                                         \
    std::fprintf(stderr, "\n");                                                \
    std::exit(1);                                                              \
  } while (0)

namespace dcx {

// Coefficient arrays in dc_precomp_vars order.
enum { CXX, CYY, CZZ, CXY, CYZ, CXZ, V2PX, V2PZ, V2SZ, V2PN, NCOEF };

struct Coef {
  const float *c[NCOEF];
};

// Half-open box of cell coordinates.
struct Box {
  int64_t x0, y0, z0, x1, y1, z1;
};

// Arguments of one stencil launch over one array set.
struct StepArgs {
  float *pp;
  const float *pc;
  float *qp;
  const float *qc;
  Coef coef;
  int64_t ys, zs;  // y and z strides in elements (x stride is 1)
  Box box;         // cells to update
  float dxx, dyy, dzz, dxy, dxz, dyz;  // 1/(dx*dx) ... formed in float
  float dt;
  // Fused source: when src_z >= 0, add src_amp to the new p and q of the
  // cell at plane src_z, in-plane offset src_xy (x + y*ys).
  int64_t src_z, src_xy;
  float src_amp;
  // Batched launches read the source amplitude of step it+1 from this
  // per-shot device table (field units) instead of src_amp.
  const float *src_tab;
  // Storage-generic views for k_fast (set by make_step_args).  Field arrays
  // hold values in field units (true value x fscale).  Leapfrog storage:
  // sf = pc/qc (read), of = pp/qp (read, then overwritten with the new
  // level).  Time-difference storage: sf = P (current level, read), of = D =
  // dscale x (P_n - P_{n-1}) (read-modify-write), nf = the new P.
  const void *sf[2];
  void *of[2];
  void *nf[2];
  const void *cf[NCOEF];
  // Layout of the arrays cf points to: the pitched layout of the fields
  // (cw == 0: offset x + y*ys + z*zs), or a dense copy of a band of cw cells
  // in x that starts at x = cx0 (offset (x - cx0) + y*cw + z*czs, with x
  // clamped into the band).
  int cx0, cw;
  int64_t czs;
  float c_lo[NCOEF], c_st[NCOEF];  // coefficient decode parameters
  float dscale;                    // time-difference units per field unit
  // 16-bit field storage: device flag set when a stored value saturated or
  // was not finite (null for FP32 fields).
  int *ovf;
  int fstore, cstore;              // DCX_STORE_*, DCX_COEF_*
  int zchunk;     // planes per workgroup (fast kernel)
  int xcd_remap;  // tile order over the XCDs: 0 round robin, 1 runs, 2 rows
  int deep;       // fast kernel prefetch depth - 1 (0, 1, 2)
  int gx, gy;     // tiles along x and y (set by the launcher)
};

void launch_exact(const StepArgs &a, hipStream_t s);
void launch_fast(const StepArgs &a, int tile_x, int tile_y, hipStream_t s);
// Number of tiles one fast launch over a's box has (tile_x x tile_y x zchunk).
int64_t fast_tiles(const StepArgs &a, int tile_x, int tile_y);
// One launch stepping many volumes: shots[k] (device array) with their tile
// counts' prefix sums prefix[0..nshots] (device), `grid` persistent
// workgroups looping over all tiles; it/n_steps select each shot's source
// amplitude from its src_tab.  All shots share the storage form of shots[0].
void launch_fast_batch(const StepArgs *shots, const int64_t *prefix,
                       int nshots, int64_t total, int fstore, int cstore,
                       int deep, int tile_x, int tile_y, int grid, int it,
                       int n_steps, hipStream_t s);
bool fast_tiling_supported(int tile_x, int tile_y);
void launch_add_source(float *pc, float *qc, int64_t index, float amp,
                       hipStream_t s);
// Pointwise pass with the step's byte pattern and no stencil (FP32 arrays).
void launch_stream_probe(const StepArgs &a, hipStream_t s);

// One box copy between a strided array and a packed buffer (elements of
// esize bytes: 4 for FP32 fields, 2 for 16-bit storage).
struct BoxCopy {
  float *arr;        // strided array (reinterpreted for esize 2)
  int64_t ys, zs;    // its strides
  int64_t x0, y0, z0;
  int64_t nx, ny, nz;
  int64_t off;       // offset of the box in the packed buffer (elements)
  int esize;
};

// dir = 0: pack (array -> buffer), dir = 1: unpack (buffer -> array).
// boxes is a device array of nbox entries; total = sum of box volumes.
void launch_box_copy(const BoxCopy *boxes, int nbox, int64_t total, float *buf,
                     int dir, hipStream_t s);
// Single box of any volume, passed by value.
void launch_box_copy1(const BoxCopy &box, float *buf, int dir, hipStream_t s);
// Fails (message, exit 1) for a box launch_box_copy cannot take: its in-box
// coordinates are 32-bit, so each box must hold fewer than 2^32 cells.
void check_multi_box(const BoxCopy &b);

// dc_propagate boxes one struct queues before launching them together (the
// reference's worker makes seven calls per step, src/worker.c:321-402).
constexpr int kMaxPending = 8;

// Several boxes of one array set in one fast-kernel launch: tile prefix sums,
// and per box its bounds, tile grid, z chunk and coefficient arrays (the
// pitched arrays, or an x band's dense copy).  Passed by value.
struct MultiBox {
  int n;
  int64_t prefix[kMaxPending + 1];
  Box box[kMaxPending];
  int gx[kMaxPending], gy[kMaxPending], zchunk[kMaxPending];
  int cx0[kMaxPending], cw[kMaxPending];
  int64_t czs[kMaxPending];
  const void *cf[kMaxPending][NCOEF];
};
// One launch of the 32 x 16 fast kernel over the boxes of m, which share
// a's arrays, spacings and dt (a.box and the per-box fields are ignored).
void launch_fast_multi(const StepArgs &a, const MultiBox &m, hipStream_t s);

// Library state behind a dc_device_data pointer.  `abi` must stay first: the
// caller holds &state->abi.
struct State {
  dc_device_data abi;
  uint64_t magic;
  int device;
  hipStream_t stream;
  bool own_stream;  // false once a driver has put the struct on its stream
  int64_t sizes[3];
  int64_t n;
  // Device layout: rows pitched to a multiple of 32 floats (128 B) and the
  // x = 0 element of every row at offset kPad, so the first updated cell
  // (x = STENCIL) of every row starts a 128-byte cache line.
  int64_t ys, zs;
  int64_t alloc_floats;  // per array, including the front pad
  float *field_alloc[4];
  float *field_mem[4];   // field_alloc + kPad
  float *coef_alloc;     // ten arrays of alloc_floats each
  // Dense copies of the ten FP32 coefficient arrays over the x bands
  // [STENCIL, 2*STENCIL) and [sx - 2*STENCIL, sx - STENCIL), all y and z:
  // [side][coefficient][z][y][STENCIL].  The reference worker's x boundary
  // slabs read them instead of 16 bytes from every 128-byte line of the
  // pitched arrays.  Null when sx < 4*STENCIL.
  float *xband;
  float *scratch;
  size_t scratch_n;
  float *stage;  // device staging for the ABI halo copies
  size_t stage_n;
  int kernel;
  int tile_x, tile_y;
  int zchunk;  // planes per workgroup of the fast kernel; 0 = chosen per launch
  int n_cu;    // compute units of the struct's device
  int xcd_remap;
  int deep;
  int timing_detail;
  double last_kernel_ms;
  hipEvent_t ev[2];
  // Loopback halo exchange run after every fused step (dcx_set_halo_loopback):
  // box tables [parity][pack|unpack] in device memory and the packed buffer.
  int halo_nbox;       // table entries per parity and direction
  int64_t halo_total;  // packed elements
  BoxCopy *halo_tab;
  float *halo_buf;
  int halo_nuser;      // boxes registered by dcx_set_halo_loopback
  size_t *halo_user;   // 12 per box: send start/end, receive start/end
  // Compact storage (dcx_set_storage): 16-bit P (two levels) and D per
  // field, 16-bit coefficients, and their decode parameters.
  int fstore, cstore;
  float fscale, dscale;
  // Overflow flag of 16-bit field storage (device int, set by the kernels)
  // and whether the host has reported it since the conversion.
  int *ovf;
  bool ovf_reported;
  uint16_t *P16[2][2];  // [field][level]
  uint16_t *D16[2];
  int pcur;             // level of P16 holding the current P
  uint16_t *C16;        // ten arrays of alloc_floats elements
  float c_lo[NCOEF], c_st[NCOEF];
  int halo_esize;       // element size the loopback tables were built for
  // Learned absorbing layer (dcx_set_learned_layer): the network, the layer
  // width, its cell count, its tiles and their corrections.
  const void *layer_mlp;
  int layer_w, layer_n, layer_ntiles;
  void *layer_tiles;   // 16 x 16 face-plane tiles of the layer
  float *layer_delta;  // corrections, per tile cell
  // Per tile cell, fixed while the layer is set: the cell's index in its
  // tile's staged box (-1 where the cell is not a layer cell of that face)
  // and sqrt(v2px), sqrt(v2pz) there.
  int16_t *layer_lidx;
  float2 *layer_root;
  // Layer cells used when the struct is the small domain of
  // dcx_learned_sample_step.
  int sample_w, sample_n;
  uint32_t *sample_cells;
  // The batch the struct belongs to (dcx_batch_create), if any.
  const void *batch;
  // Frame condition after each fused step (dcx_set_wall): DCX_WALL_ZERO
  // leaves the frame as it is, DCX_WALL_ONE_WAY / DCX_WALL_ONE_WAY2 fill it
  // from the outgoing wave; the second-order condition keeps level n-1 of
  // the last two compute layers of every face in wall_aux.
  int wall;
  float *wall_aux;
  // dc_propagate calls not yet launched: their boxes and the launch
  // arguments they share (arrays, spacings, dt).  Every other entry point
  // launches them first (state_of), as one launch (flush_pending).
  int npend;
  Box pend[kMaxPending];
  StepArgs pend_args;
};
// Count of structs a network is attached to (it cannot be freed before).
void mlp_attach(const void *mlp, int delta);

constexpr uint64_t kMagic = 0x64635f6869703031ull;  // "dc_hip01"
constexpr int64_t kRowAlign = 32;  // floats per 128-byte cache line
constexpr int64_t kPad = (kRowAlign - STENCIL % kRowAlign) % kRowAlign;

// x at which the fast kernel's tile grid starts for a box that starts at x0:
// the largest x <= x0 whose element opens a 128-byte line of the pitched
// arrays (x = STENCIL + 32k), so that each 32-element tile row lies in one
// line whatever the box.  The kernel masks the cells before x0.  Every box
// starts at x0 >= STENCIL, so the tile's apron starts at x >= 0.
__host__ __device__ inline int64_t tile_origin_x(int64_t x0) {
  return x0 >= STENCIL ? x0 - ((x0 - STENCIL) & (kRowAlign - 1)) : x0;
}

// Device offset of the caller's linear index x + y*sx + z*sx*sy.
int64_t device_index(const State *st, int64_t logical);

// The struct behind a caller's handle (fails for a foreign pointer), with its
// queued dc_propagate boxes launched first; state_peek leaves them queued.
State *state_of(dc_device_data *d);
const State *state_of(const dc_device_data *d);
State *state_peek(dc_device_data *d);
// Launch the queued dc_propagate boxes of st (none: no-op).
void flush_pending(State *st);

// Host checks of the reference ABI's box calls (each fails with a message and
// exit 1): the caller's sizes equal the struct's; the token names one of the
// struct's four fields (returned as its device array); the box [start, end)
// lies in the local grid (returned as a copy record of `arr`, offset 0).
void check_sizes(const State *st, const size_t sizes[3]);
// The two halves of dc_device_data_init.  create_state: a struct for
// process's local grid on its device (stream, pitched layout, the four fields
// uploaded from process->pp ... or zero where NULL, through the packed
// device buffer tmp of n floats or NULL, and zeroed coefficient arrays);
// the caller then fills the coefficient arrays (coef_array(st, k), pitched,
// dc_precomp_vars order) on the struct's stream; finish_state makes the
// x-band copies, waits, and returns the caller's handle.
State *create_state(const dc_process_t *process, float *tmp);
float *coef_array(const State *st, int k);
dc_device_data *finish_state(State *st);
// Resolve, once per device (the current one), the device address of the
// staging array dcx_halo_face_noalloc uses.
void noalloc_init(int device);
float *field_of(const State *st, const float *token);
BoxCopy box_of(const State *st, const size_t start[3], const size_t end[3],
               float *arr);

// Build StepArgs for the current pointer state of `st` over `box`.
StepArgs make_step_args(const State *st, const Box &box, float dx, float dy,
                        float dz, float dt);

// Launch the selected stencil kernel.
void launch_stencil(const State *st, const StepArgs &a);

// Compact storage (storage.hip).
// DCX_COEF_N16's code arrays in the C16 allocation (12 B per cell): a dword
// of (a, b) codes, then a qword of (c, rx, rn, rs) codes, each laid out as
// the fields (element x = 0 of row 0 at index 0, kPad elements in).
inline uint32_t *n16_ab(const State *st) {
  return reinterpret_cast<uint32_t *>(st->C16 - kPad) + kPad;
}
inline uint64_t *n16_cr(const State *st) {
  return reinterpret_cast<uint64_t *>(
             reinterpret_cast<char *>(st->C16 - kPad) + 4 * st->alloc_floats) +
         kPad;
}
void storage_free(State *st);
// 16-bit field storage: whether a stored value saturated or was not finite
// since the last conversion (waits for the struct's stream); report_overflow
// prints a one-line warning the first time it is true.
bool storage_overflowed(State *st);
void report_overflow(State *st);
void storage_download(const State *st, int which, float *host);
void storage_add_source(State *st, int64_t dev_index, float amp);
// (Re)build the device tables of the registered loopback halo boxes for the
// struct's current storage (device_data.hip).
void build_halo_tables(State *st);

// Move a struct onto a caller-owned stream (the struct's own stream is
// destroyed); used by drivers that order several subdomains on one stream.
void adopt_stream(State *st, hipStream_t s);
// Learned layer hooks around a step's stencil launch (no-ops when unset):
// corrections from the current fields, then added to the new ones.
void layer_infer(State *st, const StepArgs &a);
// One whole step as dcx_step runs it; after_stencil(ctx), when given, runs
// between the stencil launch and the learned-layer corrections.
void fused_step(State *st, float dx, float dy, float dz, float dt,
                long long src_index, float src_next, int add_next,
                void (*after_stencil)(void *), void *ctx);
void layer_apply(State *st);
void layer_free(State *st);
// The one-way wall (wall.hip; no-ops when the struct's wall is
// DCX_WALL_ZERO): wall_save keeps what the second-order condition needs of
// level n-1 before the stencil overwrites it, wall_apply fills the frame of
// the new level after the stencil, wall_free releases the wall's buffer.
void wall_save(State *st);
void wall_apply(State *st, float dx, float dy, float dz, float dt);
void wall_free(State *st);
// Host array (x fastest) <-> pitched device array through a packed device
// buffer tmp of n floats (NULL: allocated and freed inside).
void host_to_pitched(const State *st, float *dev, const float *host, float *tmp);
void pitched_to_host(const State *st, float *host, const float *dev, float *tmp);

}  // namespace dcx
