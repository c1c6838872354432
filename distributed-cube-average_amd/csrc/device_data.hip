// The reference's backend ABI (include/propagate.h, include/device_data.h) on
// one MI355X, plus the library's own dcx_* entry points for device-resident
// time loops.  All work of one dc_device_data is ordered on one HIP stream.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dc_internal.hpp"
#include "propagate.h"

namespace dcx {

State *state_peek(dc_device_data *d) {
  if (d == nullptr) DCX_FAIL("null dc_device_data");
  State *st = reinterpret_cast<State *>(d);
  if (st->magic != kMagic)
    DCX_FAIL("dc_device_data was not created by dc_device_data_init");
  return st;
}

State *state_of(dc_device_data *d) {
  State *st = state_peek(d);
  flush_pending(st);
  return st;
}

const State *state_of(const dc_device_data *d) {
  return state_of(const_cast<dc_device_data *>(d));
}

void check_sizes(const State *st, const size_t sizes[3]) {
  for (int k = 0; k < 3; ++k)
    if ((int64_t)sizes[k] != st->sizes[k])
      DCX_FAIL("sizes (%zu,%zu,%zu) differ from the initialised (%lld,%lld,%lld)",
               sizes[0], sizes[1], sizes[2], (long long)st->sizes[0],
               (long long)st->sizes[1], (long long)st->sizes[2]);
}

// Which device array a caller's field token names.
float *field_of(const State *st, const float *token) {
  const dc_device_data &d = st->abi;
  if (token == d.pp || token == d.pc || token == d.qp || token == d.qc)
    return const_cast<float *>(token);
  DCX_FAIL("array is not one of this dc_device_data's fields");
}

// z chunk of a fast launch when the struct leaves it to the library: the
// longest chunk (at most 512 planes) that still gives the launch two tiles
// per compute unit, and never below 32 planes (each chunk re-reads 8 planes
// to fill its z rings).  A 512^3 box keeps whole 512-plane columns; a 200^3
// shot, 91 columns of 32 x 16 cells, gets 34-plane chunks.
static int auto_zchunk(const State *st, const Box &box) {
  const int64_t nx = box.x1 - box.x0, ny = box.y1 - box.y0, nz = box.z1 - box.z0;
  if (nx <= 0 || ny <= 0 || nz <= 0) return 512;
  const int64_t cols = ((box.x1 - tile_origin_x(box.x0) + st->tile_x - 1) / st->tile_x) *
                       ((ny + st->tile_y - 1) / st->tile_y);
  const int64_t target = 2 * (int64_t)(st->n_cu > 0 ? st->n_cu : 256);
  const int64_t chunks = (target + cols - 1) / cols;
  const int64_t z = (nz + chunks - 1) / chunks;
  return (int)(z < 32 ? 32 : z > 512 ? 512 : z);
}

StepArgs make_step_args(const State *st, const Box &box, float dx, float dy,
                        float dz, float dt) {
  StepArgs a{};
  a.pp = st->abi.pp;
  a.pc = st->abi.pc;
  a.qp = st->abi.qp;
  a.qc = st->abi.qc;
  const dc_precomp_vars &pv = st->abi.precomp_vars;
  const float *cs[NCOEF] = {pv.ch1dxx, pv.ch1dyy, pv.ch1dzz, pv.ch1dxy,
                            pv.ch1dyz, pv.ch1dxz, pv.v2px,   pv.v2pz,
                            pv.v2sz,   pv.v2pn};
  for (int k = 0; k < NCOEF; ++k) a.coef.c[k] = cs[k];
  a.ys = st->ys;
  a.zs = st->zs;
  a.box = box;
  // Formed in float as include/sample_compute.h:128-133 forms them.
  a.dxx = 1.0f / (dx * dx);
  a.dyy = 1.0f / (dy * dy);
  a.dzz = 1.0f / (dz * dz);
  a.dxy = 1.0f / (dx * dy);
  a.dxz = 1.0f / (dx * dz);
  a.dyz = 1.0f / (dy * dz);
  a.dt = dt;
  a.src_z = -1;
  a.src_xy = -1;
  a.src_amp = 0.0f;
  a.zchunk = st->zchunk > 0 ? st->zchunk : auto_zchunk(st, box);
  // The default tile order (-1) depends on the storage: plain FP32 runs
  // fastest round robin, every compact form with whole tile rows per XCD.
  a.xcd_remap = st->xcd_remap >= 0 ? st->xcd_remap
                : st->fstore == DCX_STORE_F32 && st->cstore == DCX_COEF_F32 ? 0
                                                                             : 2;
  a.deep = st->deep;
  a.fstore = st->fstore;
  a.cstore = st->cstore;
  a.dscale = 1.0f;
  if (st->fstore == DCX_STORE_F32) {
    a.sf[0] = a.pc;
    a.sf[1] = a.qc;
    a.of[0] = a.pp;
    a.of[1] = a.qp;
    a.nf[0] = a.nf[1] = nullptr;
  } else {
    a.sf[0] = st->P16[0][st->pcur];
    a.sf[1] = st->P16[1][st->pcur];
    a.nf[0] = st->P16[0][1 - st->pcur];
    a.nf[1] = st->P16[1][1 - st->pcur];
    a.of[0] = st->D16[0];
    a.of[1] = st->D16[1];
    a.dscale = st->dscale;
    a.ovf = st->ovf;
  }
  for (int k = 0; k < NCOEF; ++k) {
    const bool f32 = st->cstore == DCX_COEF_F32 ||
                     (st->cstore == DCX_COEF_A16 && k >= V2PX);
    a.cf[k] = f32 ? (const void *)a.coef.c[k]
                  : (const void *)(st->C16 + (size_t)k * st->alloc_floats);
    a.c_lo[k] = st->c_lo[k];
    a.c_st[k] = st->c_st[k];
  }
  if (st->cstore == DCX_COEF_N16) {
    // The axis form's dword (a, b) and qword (c, rx, rn, rs) code arrays
    // share the C16 allocation; v2pz stays the FP32 array.  The kernel reads
    // (a, b) once per cell, through slot CXX.
    const void *ab = n16_ab(st), *cr = n16_cr(st);
    for (int k = 0; k < NCOEF; ++k) a.cf[k] = nullptr;
    a.cf[CXX] = ab;
    a.cf[CXZ] = a.cf[CYZ] = cr;
    a.cf[V2PZ] = a.coef.c[V2PZ];
    // The kernel decodes the axis as sqrt(2) (n_x/dx, n_y/dy, n_z/dz): a
    // product of two components is then the angle term over d_a d_b (2 n_a
    // n_b off the diagonal; twice n_a^2 on it, which the kernel's
    // half-weight second differences undo).
    const double inv[3] = {std::sqrt(2.0) / dx, std::sqrt(2.0) / dy,
                           std::sqrt(2.0) / dz};
    for (int k = 0; k < 3; ++k) {
      a.c_lo[k] = (float)(st->c_lo[k] * inv[k]);
      a.c_st[k] = (float)(st->c_st[k] * inv[k]);
    }
  }
  return a;
}

// The fast kernel's arguments for a's box: a box inside one of the two
// 4-cell x bands (the reference worker's x boundary slabs) reads its FP32
// coefficients from the band's dense copy.
static StepArgs with_band(const State *st, const StepArgs &a) {
  if (st->xband && a.cstore == DCX_COEF_F32) {
    const int64_t band = STENCIL * st->sizes[1] * st->sizes[2];
    for (int side = 0; side < 2; ++side) {
      const int64_t bx0 = side ? st->sizes[0] - 2 * STENCIL : STENCIL;
      if (a.box.x0 >= bx0 && a.box.x1 <= bx0 + STENCIL) {
        StepArgs b = a;
        for (int k = 0; k < NCOEF; ++k)
          b.cf[k] = st->xband + ((size_t)side * NCOEF + k) * (size_t)band;
        b.cx0 = (int)bx0;
        b.cw = STENCIL;
        b.czs = STENCIL * st->sizes[1];
        return b;
      }
    }
  }
  return a;
}

void launch_stencil(const State *st, const StepArgs &a) {
  if (st->kernel == DCX_KERNEL_EXACT) {
    if (st->fstore != DCX_STORE_F32 || st->cstore != DCX_COEF_F32)
      DCX_FAIL("the exact kernel runs on FP32 storage only");
    launch_exact(a, st->stream);
    return;
  }
  launch_fast(with_band(st, a), st->tile_x, st->tile_y, st->stream);
}

static int64_t box_volume(const Box &b) {
  return (b.x1 - b.x0) * (b.y1 - b.y0) * (b.z1 - b.z0);
}

static bool boxes_overlap(const Box &a, const Box &b) {
  return a.x0 < b.x1 && b.x0 < a.x1 && a.y0 < b.y1 && b.y0 < a.y1 &&
         a.z0 < b.z1 && b.z0 < a.z1;
}

// Launch the queued dc_propagate boxes.  They are pairwise disjoint (a box
// that overlaps a queued one flushes the queue first) and every cell's update
// reads only the current fields, so their order does not matter.  When the
// boxes tile their bounding box exactly -- the reference worker's six slabs
// and interior cover the compute region -- they run as one launch over it;
// otherwise the fast kernel takes them all in one multi-box launch (the
// exact kernel, or forms without one, run them one by one).  The results are
// those of the separate calls, bit for bit.
void flush_pending(State *st) {
  const int n = st->npend;
  if (n == 0) return;
  st->npend = 0;
  StepArgs a = st->pend_args;
  Box hull = st->pend[0];
  int64_t vol = 0;
  for (int k = 0; k < n; ++k) {
    const Box &b = st->pend[k];
    hull.x0 = std::min(hull.x0, b.x0);
    hull.y0 = std::min(hull.y0, b.y0);
    hull.z0 = std::min(hull.z0, b.z0);
    hull.x1 = std::max(hull.x1, b.x1);
    hull.y1 = std::max(hull.y1, b.y1);
    hull.z1 = std::max(hull.z1, b.z1);
    vol += box_volume(b);
  }
  if (n == 1 || vol == box_volume(hull)) {
    a.box = n == 1 ? st->pend[0] : hull;
    if (n > 1 && st->zchunk == 0) a.zchunk = auto_zchunk(st, a.box);
    launch_stencil(st, a);
    return;
  }
  const bool multi = st->kernel == DCX_KERNEL_FAST && st->tile_x == 32 &&
                     st->tile_y == 16 && a.fstore == DCX_STORE_F32 &&
                     a.cstore == DCX_COEF_F32;
  if (!multi) {
    for (int k = 0; k < n; ++k) {
      a.box = st->pend[k];
      if (st->zchunk == 0) a.zchunk = auto_zchunk(st, a.box);
      launch_stencil(st, a);
    }
    return;
  }
  MultiBox m{};
  m.n = n;
  for (int k = 0; k < n; ++k) {
    StepArgs b = a;
    b.box = st->pend[k];
    b.zchunk = st->zchunk > 0 ? st->zchunk : auto_zchunk(st, b.box);
    b = with_band(st, b);
    m.box[k] = b.box;
    m.zchunk[k] = b.zchunk;
    m.gx[k] = (int)((b.box.x1 - tile_origin_x(b.box.x0) + 31) / 32);
    m.gy[k] = (int)((b.box.y1 - b.box.y0 + 15) / 16);
    m.cx0[k] = b.cx0;
    m.cw[k] = b.cw;
    m.czs[k] = b.czs;
    for (int c = 0; c < NCOEF; ++c) m.cf[k][c] = b.cf[c];
    m.prefix[k + 1] = m.prefix[k] + fast_tiles(b, 32, 16);
  }
  launch_fast_multi(a, m, st->stream);
}

void adopt_stream(State *st, hipStream_t s) {
  DCX_CHECK(hipStreamSynchronize(st->stream));
  if (st->own_stream) DCX_CHECK(hipStreamDestroy(st->stream));
  st->stream = s;
  st->own_stream = false;
}

int64_t device_index(const State *st, int64_t logical) {
  const int64_t plane = st->sizes[0] * st->sizes[1];
  const int64_t z = logical / plane, r = logical % plane;
  const int64_t y = r / st->sizes[0], x = r % st->sizes[0];
  return x + y * st->ys + z * st->zs;
}

static void set_source(const State *st, StepArgs &a, long long index,
                       float amp) {
  if (index < 0) return;
  if (index >= st->n) DCX_FAIL("source index %lld out of range", index);
  const int64_t i = device_index(st, index);
  a.src_z = i / st->zs;
  a.src_xy = i % st->zs;
  a.src_amp = st->fstore == DCX_STORE_F32 ? amp : amp * st->fscale;
}

// Whether the cell of logical index `index` lies in a send box of the
// registered loopback.  The reference's worker adds step it+1's source to pc
// after the halos of step it have been sent and received
// (src/worker.c:466-511), so a neighbour never receives it.  The fused step
// adds it as the stencil writes the new field, before the loopback packs
// that field; for a source cell inside a send box the step therefore adds it
// after the loopback instead (add_deferred_source).
static bool source_in_send_box(const State *st, long long index) {
  if (index < 0 || st->halo_nuser == 0) return false;
  const int64_t plane = st->sizes[0] * st->sizes[1];
  const int64_t c[3] = {index % st->sizes[0], (index % plane) / st->sizes[0],
                        index / plane};
  for (int k = 0; k < st->halo_nuser; ++k) {
    const size_t *u = st->halo_user + 12 * k;
    bool in = true;
    for (int a = 0; a < 3; ++a)
      in = in && c[a] >= (int64_t)u[a] && c[a] < (int64_t)u[3 + a];
    if (in) return true;
  }
  return false;
}

// Add a source amplitude to the current fields (pc, qc) after a step's
// loopback: dc_device_add_source's additions.
static void add_deferred_source(State *st, long long index, float amp) {
  if (st->fstore == DCX_STORE_F32)
    launch_add_source(st->abi.pc, st->abi.qc, device_index(st, index), amp,
                      st->stream);
  else
    storage_add_source(st, device_index(st, index), amp);
}

// Advance the struct one step: swap the leapfrog pointers, or flip the
// current level of compact storage.
static void advance(State *st) {
  if (st->fstore == DCX_STORE_F32)
    dc_device_swap_arrays(&st->abi);
  else
    st->pcur ^= 1;
}

// Host array (x fastest, rows of sx) <-> pitched device array.  The host side
// moves as one contiguous copy through a packed device buffer (`tmp`, at
// least sizes[0]*sizes[1]*sizes[2] floats, or NULL to allocate one) and the
// pitch is applied by a device-to-device 2-D copy: a 2-D copy straight from
// pageable host memory goes row by row and runs at a small fraction of the
// link's rate.
void host_to_pitched(const State *st, float *dev, const float *host, float *tmp) {
  const size_t n = (size_t)st->n;
  float *buf = tmp;
  if (buf == nullptr) DCX_CHECK(hipMalloc(&buf, n * sizeof(float)));
  DCX_CHECK(hipMemcpyAsync(buf, host, n * sizeof(float), hipMemcpyHostToDevice,
                           st->stream));
  DCX_CHECK(hipMemcpy2DAsync(dev, st->ys * sizeof(float), buf,
                             st->sizes[0] * sizeof(float),
                             st->sizes[0] * sizeof(float),
                             st->sizes[1] * st->sizes[2],
                             hipMemcpyDeviceToDevice, st->stream));
  if (tmp == nullptr) {
    DCX_CHECK(hipStreamSynchronize(st->stream));
    DCX_CHECK(hipFree(buf));
  }
}

void pitched_to_host(const State *st, float *host, const float *dev, float *tmp) {
  const size_t n = (size_t)st->n;
  float *buf = tmp;
  if (buf == nullptr) DCX_CHECK(hipMalloc(&buf, n * sizeof(float)));
  DCX_CHECK(hipMemcpy2DAsync(buf, st->sizes[0] * sizeof(float), dev,
                             st->ys * sizeof(float),
                             st->sizes[0] * sizeof(float),
                             st->sizes[1] * st->sizes[2],
                             hipMemcpyDeviceToDevice, st->stream));
  DCX_CHECK(hipMemcpyAsync(host, buf, n * sizeof(float), hipMemcpyDeviceToHost,
                           st->stream));
  DCX_CHECK(hipStreamSynchronize(st->stream));
  if (tmp == nullptr) DCX_CHECK(hipFree(buf));
}

static void copy_in(const State *st, float *dev, const float *host,
                    float *tmp = nullptr) {
  host_to_pitched(st, dev, host, tmp);
}

static void copy_out(const State *st, float *host, const float *dev) {
  pitched_to_host(st, host, dev, nullptr);
}

static float *ensure(float *&buf, size_t &cap, size_t n) {
  if (n > cap) {
    if (buf) DCX_CHECK(hipFree(buf));
    DCX_CHECK(hipMalloc(&buf, n * sizeof(float)));
    cap = n;
  }
  return buf;
}

// Device ordinal of a rank: the number of earlier ranks on the same host,
// modulo the device count (the rule of src/device_data.cu:18-33).
static int select_device(const dc_process_t *p) {
  int count = 0;
  DCX_CHECK(hipGetDeviceCount(&count));
  if (count <= 0) DCX_FAIL("no HIP device");
  if (p->hostnames == nullptr || p->num_workers == 0) return 0;
  const size_t len = 256;
  const char *mine = p->hostnames + len * (size_t)p->rank;
  int same = 0;
  for (int i = 0; i < (int)p->num_workers && i != p->rank; i++)
    if (std::strncmp(mine, p->hostnames + len * (size_t)i, len) == 0) same++;
  return same % count;
}

BoxCopy box_of(const State *st, const size_t start[3],
                      const size_t end[3], float *arr) {
  for (int k = 0; k < 3; ++k)
    if (start[k] > end[k] || (int64_t)end[k] > st->sizes[k])
      DCX_FAIL("box [%zu,%zu)x[%zu,%zu)x[%zu,%zu) outside the local grid",
               start[0], end[0], start[1], end[1], start[2], end[2]);
  BoxCopy b{};
  b.arr = arr;
  b.ys = st->ys;
  b.zs = st->zs;
  b.x0 = (int64_t)start[0];
  b.y0 = (int64_t)start[1];
  b.z0 = (int64_t)start[2];
  b.nx = (int64_t)(end[0] - start[0]);
  b.ny = (int64_t)(end[1] - start[1]);
  b.nz = (int64_t)(end[2] - start[2]);
  b.off = 0;
  return b;
}

// Entries: for each registered box, p then q; tables for both parities.  The
// arrays are those a step writes its new level to: pp/qp at the pointer
// parity (FP32 leapfrog), or the P level after the current one (compact).
void build_halo_tables(State *st) {
  DCX_CHECK(hipStreamSynchronize(st->stream));
  if (st->halo_tab) DCX_CHECK(hipFree(st->halo_tab));
  if (st->halo_buf) DCX_CHECK(hipFree(st->halo_buf));
  st->halo_tab = nullptr;
  st->halo_buf = nullptr;
  st->halo_nbox = 0;
  st->halo_total = 0;
  const int nbox = st->halo_nuser;
  if (nbox <= 0) return;
  const bool f32 = st->fstore == DCX_STORE_F32;
  std::vector<BoxCopy> tab((size_t)4 * 2 * nbox);
  int64_t total = 0;
  const int nb = 2 * nbox;
  for (int k = 0; k < nbox; ++k) {
    const size_t *u = st->halo_user + 12 * k;
    for (int f = 0; f < 2; ++f) {
      for (int par = 0; par < 2; ++par) {
        float *arr = f32 ? st->field_mem[2 * f + par]
                         : reinterpret_cast<float *>(st->P16[f][1 - par]);
        BoxCopy pb = box_of(st, u, u + 3, arr);
        BoxCopy ub = box_of(st, u + 6, u + 9, arr);
        if (pb.nx != ub.nx || pb.ny != ub.ny || pb.nz != ub.nz)
          DCX_FAIL("send and receive box %d differ in shape", k);
        pb.esize = ub.esize = f32 ? 4 : 2;
        pb.off = ub.off = total;
        check_multi_box(pb);
        tab[(size_t)(2 * par) * nb + 2 * k + f] = pb;
        tab[(size_t)(2 * par + 1) * nb + 2 * k + f] = ub;
      }
      const BoxCopy &b = tab[2 * k + f];
      total += b.nx * b.ny * b.nz;
    }
  }
  DCX_CHECK(hipMalloc(&st->halo_tab, tab.size() * sizeof(BoxCopy)));
  DCX_CHECK(hipMemcpy(st->halo_tab, tab.data(), tab.size() * sizeof(BoxCopy),
                      hipMemcpyHostToDevice));
  DCX_CHECK(hipMalloc(&st->halo_buf, (size_t)total * sizeof(float)));
  st->halo_nbox = nb;
  st->halo_total = total;
}

}  // namespace dcx

using namespace dcx;

extern "C" {

dc_device_data *dc_device_data_init(dc_process_t *process) {
  if (process == nullptr) DCX_FAIL("null process");
  const dc_precomp_vars &pv = process->precomp_vars;
  const float *hc[NCOEF] = {pv.ch1dxx, pv.ch1dyy, pv.ch1dzz, pv.ch1dxy,
                            pv.ch1dyz, pv.ch1dxz, pv.v2px,   pv.v2pz,
                            pv.v2sz,   pv.v2pn};
  for (int k = 0; k < NCOEF; ++k)
    if (hc[k] == nullptr) DCX_FAIL("precomp_vars array %d is NULL", k);
  const size_t n = process->sizes[0] * process->sizes[1] * process->sizes[2];
  float *tmp = nullptr;  // packed staging for the uploads
  if (n > 0) DCX_CHECK(hipMalloc(&tmp, n * sizeof(float)));
  State *st = create_state(process, tmp);
  for (int k = 0; k < NCOEF; ++k) copy_in(st, coef_array(st, k), hc[k], tmp);
  dc_device_data *d = finish_state(st);
  DCX_CHECK(hipFree(tmp));
  return d;
}

}  // extern "C"

namespace dcx {

float *coef_array(const State *st, int k) {
  return st->coef_alloc + (size_t)k * st->alloc_floats + kPad;
}

State *create_state(const dc_process_t *process, float *tmp) {
  State *st = static_cast<State *>(std::calloc(1, sizeof(State)));
  if (st == nullptr)
    DCX_FAIL("OOM: could not allocate memory for device_data in "
             "dc_device_data_init");
  st->magic = kMagic;
  st->device = select_device(process);
  DCX_CHECK(hipSetDevice(st->device));
  noalloc_init(st->device);
  DCX_CHECK(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
  st->own_stream = true;
  DCX_CHECK(hipEventCreate(&st->ev[0]));
  DCX_CHECK(hipEventCreate(&st->ev[1]));
  for (int k = 0; k < 3; ++k) {
    if (process->sizes[k] < 1) DCX_FAIL("empty local grid");
    st->sizes[k] = (int64_t)process->sizes[k];
  }
  st->n = st->sizes[0] * st->sizes[1] * st->sizes[2];
  if (st->sizes[0] * st->sizes[1] >= (int64_t(1) << 31))
    DCX_FAIL("one xy plane must hold fewer than 2^31 cells");
  st->ys = (st->sizes[0] + kRowAlign - 1) / kRowAlign * kRowAlign;
  st->zs = st->ys * st->sizes[1];
  st->alloc_floats = (kPad + st->zs * st->sizes[2] + 63) / 64 * 64;
  st->kernel = DCX_KERNEL_FAST;
  st->tile_x = 32;
  st->tile_y = 16;
  st->zchunk = 0;
  DCX_CHECK(hipDeviceGetAttribute(&st->n_cu, hipDeviceAttributeMultiprocessorCount,
                                  st->device));
  st->xcd_remap = -1;
  // Three-level prefetch: taken by the 32x16 FP32 kernels (1.635 against
  // 1.673 ms per 512^3 launch with two levels); every other kernel runs two.
  st->deep = 2;
  const size_t abytes = (size_t)st->alloc_floats * sizeof(float);

  float *host_fields[4] = {process->pp, process->pc, process->qp, process->qc};
  for (int k = 0; k < 4; ++k) {
    DCX_CHECK(hipMalloc(&st->field_alloc[k], abytes));
    DCX_CHECK(hipMemsetAsync(st->field_alloc[k], 0, abytes, st->stream));
    st->field_mem[k] = st->field_alloc[k] + kPad;
    if (host_fields[k]) copy_in(st, st->field_mem[k], host_fields[k], tmp);
  }
  DCX_CHECK(hipMalloc(&st->coef_alloc, abytes * NCOEF));
  DCX_CHECK(hipMemsetAsync(st->coef_alloc, 0, abytes * NCOEF, st->stream));
  return st;
}

dc_device_data *finish_state(State *st) {
  float *dc[NCOEF];
  for (int k = 0; k < NCOEF; ++k) dc[k] = coef_array(st, k);
  // Dense copies of the coefficients over the two x bands (one 2-D copy of
  // 16-byte rows per array and band).
  if (st->sizes[0] >= 4 * STENCIL) {
    const size_t band = (size_t)STENCIL * st->sizes[1] * st->sizes[2];
    DCX_CHECK(hipMalloc(&st->xband, 2 * NCOEF * band * sizeof(float)));
    for (int side = 0; side < 2; ++side) {
      const int64_t bx0 = side ? st->sizes[0] - 2 * STENCIL : STENCIL;
      for (int k = 0; k < NCOEF; ++k)
        DCX_CHECK(hipMemcpy2DAsync(
            st->xband + ((size_t)side * NCOEF + k) * band,
            STENCIL * sizeof(float), dc[k] + bx0, st->ys * sizeof(float),
            STENCIL * sizeof(float), st->sizes[1] * st->sizes[2],
            hipMemcpyDeviceToDevice, st->stream));
    }
  }
  DCX_CHECK(hipStreamSynchronize(st->stream));

  dc_device_data &d = st->abi;
  d.pp = st->field_mem[0];
  d.pc = st->field_mem[1];
  d.qp = st->field_mem[2];
  d.qc = st->field_mem[3];
  d.vpz = nullptr;
  d.vsv = nullptr;
  d.precomp_vars = dc_precomp_vars{dc[0], dc[1], dc[2], dc[3], dc[4],
                                   dc[5], dc[6], dc[7], dc[8], dc[9]};
  return &d;
}

}  // namespace dcx

extern "C" {

void dc_device_data_free(dc_device_data *data) {
  if (data == nullptr) return;
  State *st = state_of(data);
  if (st->batch) DCX_FAIL("free the batch (dcx_batch_free) before its shots");
  DCX_CHECK(hipSetDevice(st->device));
  DCX_CHECK(hipStreamSynchronize(st->stream));
  for (int k = 0; k < 4; ++k) DCX_CHECK(hipFree(st->field_alloc[k]));
  DCX_CHECK(hipFree(st->coef_alloc));
  if (st->xband) DCX_CHECK(hipFree(st->xband));
  if (st->scratch) DCX_CHECK(hipFree(st->scratch));
  if (st->stage) DCX_CHECK(hipFree(st->stage));
  if (st->halo_tab) DCX_CHECK(hipFree(st->halo_tab));
  if (st->halo_buf) DCX_CHECK(hipFree(st->halo_buf));
  std::free(st->halo_user);
  layer_free(st);
  wall_free(st);
  storage_free(st);
  DCX_CHECK(hipEventDestroy(st->ev[0]));
  DCX_CHECK(hipEventDestroy(st->ev[1]));
  if (st->own_stream) DCX_CHECK(hipStreamDestroy(st->stream));
  st->magic = 0;
  std::free(st);
}

void dc_device_data_get_results(dc_process_t *process, dc_device_data *data) {
  State *st = state_of(data);
  if (process->pc == nullptr || process->qc == nullptr)
    DCX_FAIL("process->pc/qc must be host buffers of the local size");
  if (st->fstore != DCX_STORE_F32) {
    report_overflow(st);
    storage_download(st, 1, process->pc);
    storage_download(st, 3, process->qc);
    return;
  }
  copy_out(st, process->pc, data->pc);
  copy_out(st, process->qc, data->qc);
  DCX_CHECK(hipStreamSynchronize(st->stream));
}

void dc_device_swap_arrays(dc_device_data *data) {
  // Host pointers only: queued boxes keep the arrays they were queued with.
  state_peek(data);
  float *t = data->pp;
  data->pp = data->pc;
  data->pc = t;
  t = data->qp;
  data->qp = data->qc;
  data->qc = t;
}

void dc_device_add_source(dc_device_data *data, size_t index, float source) {
  State *st = state_of(data);
  if ((int64_t)index >= st->n) DCX_FAIL("source index %zu out of range", index);
  if (st->fstore != DCX_STORE_F32) {
    storage_add_source(st, device_index(st, (int64_t)index), source);
    return;
  }
  launch_add_source(data->pc, data->qc, device_index(st, (int64_t)index),
                    source, st->stream);
}


void dc_device_extract_halo_face(dc_device_data *data, float *buffer,
                                 const size_t start_coords[DIMENSIONS],
                                 const size_t end_coords[DIMENSIONS],
                                 const size_t sizes[DIMENSIONS],
                                 const float *from_array) {
  State *st = state_of(data);
  check_sizes(st, sizes);
  const BoxCopy b = box_of(st, start_coords, end_coords, field_of(st, from_array));
  const size_t n = (size_t)(b.nx * b.ny * b.nz);
  if (n == 0) return;
  float *stage = ensure(st->stage, st->stage_n, n);
  launch_box_copy1(b, stage, 0, st->stream);
  DCX_CHECK(hipMemcpyAsync(buffer, stage, n * sizeof(float),
                           hipMemcpyDeviceToHost, st->stream));
  DCX_CHECK(hipStreamSynchronize(st->stream));
}

void dc_device_insert_halo_face(dc_device_data *data, const float *buffer,
                                const size_t start_coords[DIMENSIONS],
                                const size_t end_coords[DIMENSIONS],
                                const size_t sizes[DIMENSIONS],
                                float *to_array) {
  State *st = state_of(data);
  check_sizes(st, sizes);
  const BoxCopy b = box_of(st, start_coords, end_coords, field_of(st, to_array));
  const size_t n = (size_t)(b.nx * b.ny * b.nz);
  if (n == 0) return;
  // The staging buffer may still feed an earlier queued unpack: wait first.
  DCX_CHECK(hipStreamSynchronize(st->stream));
  float *stage = ensure(st->stage, st->stage_n, n);
  DCX_CHECK(hipMemcpyAsync(stage, buffer, n * sizeof(float),
                           hipMemcpyHostToDevice, st->stream));
  // The caller may free `buffer` as soon as this returns.
  DCX_CHECK(hipStreamSynchronize(st->stream));
  launch_box_copy1(b, stage, 1, st->stream);
}

void dc_propagate(const size_t start_coords[DIMENSIONS],
                  const size_t end_coords[DIMENSIONS],
                  const size_t sizes[DIMENSIONS],
                  const int process_coordinates[DIMENSIONS],
                  const int topology[DIMENSIONS], dc_device_data *data,
                  const float dx, const float dy, const float dz,
                  const float dt) {
  (void)process_coordinates;
  (void)topology;
  State *st = state_peek(data);  // queued boxes stay queued
  check_sizes(st, sizes);
  if (st->fstore != DCX_STORE_F32)
    DCX_FAIL("dc_propagate runs on FP32 field storage (dcx_set_storage)");
  for (int k = 0; k < 3; ++k)
    if (end_coords[k] <= start_coords[k]) return;
  for (int k = 0; k < 3; ++k)
    if (start_coords[k] < STENCIL || (int64_t)end_coords[k] > st->sizes[k] - STENCIL)
      DCX_FAIL("dc_propagate box [%zu,%zu)x[%zu,%zu)x[%zu,%zu) must lie in "
               "[%d, size-%d) on every axis",
               start_coords[0], end_coords[0], start_coords[1], end_coords[1],
               start_coords[2], end_coords[2], STENCIL, STENCIL);
  const Box box{(int64_t)start_coords[0], (int64_t)start_coords[1],
                (int64_t)start_coords[2], (int64_t)end_coords[0],
                (int64_t)end_coords[1], (int64_t)end_coords[2]};
  // Queue the box; the next entry point that needs the fields launches the
  // queue (flush_pending).  A box that cannot join the queue -- other
  // arrays (after a swap), other spacings or dt, an overlap with a queued
  // box, a full queue -- launches the queue first.
  const StepArgs a = make_step_args(st, box, dx, dy, dz, dt);
  if (st->npend > 0) {
    const StepArgs &q = st->pend_args;
    bool join = st->npend < kMaxPending && a.pp == q.pp && a.pc == q.pc &&
                a.qp == q.qp && a.qc == q.qc && a.dxx == q.dxx &&
                a.dyy == q.dyy && a.dzz == q.dzz && a.dxy == q.dxy &&
                a.dxz == q.dxz && a.dyz == q.dyz && a.dt == q.dt;
    for (int k = 0; join && k < st->npend; ++k)
      join = !boxes_overlap(st->pend[k], box);
    if (!join) flush_pending(st);
  }
  if (st->npend == 0) st->pend_args = a;
  st->pend[st->npend++] = box;
}

int dcx_set_kernel(dc_device_data *data, int kernel) {
  State *st = state_of(data);
  if (kernel != DCX_KERNEL_FAST && kernel != DCX_KERNEL_EXACT)
    DCX_FAIL("unknown kernel %d", kernel);
  if (st->batch && kernel != DCX_KERNEL_FAST)
    DCX_FAIL("dcx_set_kernel on a batch shot: batched launches run the fast "
             "kernel only (dcx_batch_free first)");
  const int prev = st->kernel;
  st->kernel = kernel;
  return prev;
}

int dcx_set_tiling(dc_device_data *data, int tile_x, int tile_y, int z_chunk) {
  State *st = state_of(data);
  if (!fast_tiling_supported(tile_x, tile_y) || z_chunk < 0) return -1;
  st->tile_x = tile_x;
  st->tile_y = tile_y;
  st->zchunk = z_chunk;
  return 0;
}

int dcx_set_prefetch_depth(dc_device_data *data, int planes) {
  State *st = state_of(data);
  const int prev = st->deep + 1;
  if (planes < 1 || planes > 3) DCX_FAIL("prefetch depth must be 1, 2 or 3");
  st->deep = planes - 1;
  return prev;
}

int dcx_set_xcd_remap(dc_device_data *data, int mode) {
  State *st = state_of(data);
  const int prev = st->xcd_remap;
  st->xcd_remap = mode < 0 ? -1 : mode == 2 ? 2 : mode != 0;
  return prev;
}

void dcx_synchronize(dc_device_data *data) {
  State *st = state_of(data);
  DCX_CHECK(hipStreamSynchronize(st->stream));
  report_overflow(st);
}

static float *field_index(State *st, int which) {
  switch (which) {
    case 0: return st->abi.pp;
    case 1: return st->abi.pc;
    case 2: return st->abi.qp;
    case 3: return st->abi.qc;
  }
  DCX_FAIL("field index %d not in 0..3", which);
}

void dcx_download(dc_device_data *data, int which, float *host) {
  State *st = state_of(data);
  if (st->fstore != DCX_STORE_F32) {
    if (which < 0 || which > 3) DCX_FAIL("field index %d not in 0..3", which);
    DCX_CHECK(hipStreamSynchronize(st->stream));
    report_overflow(st);
    storage_download(st, which, host);
    return;
  }
  copy_out(st, host, field_index(st, which));
  DCX_CHECK(hipStreamSynchronize(st->stream));
}

void dcx_download_coef(dc_device_data *data, int k, float *host) {
  State *st = state_of(data);
  if (k < 0 || k >= NCOEF) DCX_FAIL("coefficient index %d not in 0..9", k);
  copy_out(st, host, coef_array(st, k));
  DCX_CHECK(hipStreamSynchronize(st->stream));
}

void dcx_upload(dc_device_data *data, int which, const float *host) {
  State *st = state_of(data);
  if (st->fstore != DCX_STORE_F32)
    DCX_FAIL("dcx_upload needs FP32 field storage (dcx_set_storage)");
  copy_in(st, field_index(st, which), host);
  DCX_CHECK(hipStreamSynchronize(st->stream));
}

static Box full_box(const State *st) {
  return Box{STENCIL,
             STENCIL,
             STENCIL,
             st->sizes[0] - STENCIL,
             st->sizes[1] - STENCIL,
             st->sizes[2] - STENCIL};
}

// Pack the registered send boxes of the new pp/qp and unpack them into the
// receive boxes (no-op when no loopback is registered).
static void halo_loopback(State *st) {
  if (st->halo_nbox == 0) return;
  const int par = st->fstore == DCX_STORE_F32
                      ? (st->abi.pp == st->field_mem[0] ? 0 : 1)
                      : st->pcur;
  const int nb = st->halo_nbox;
  launch_box_copy(st->halo_tab + (2 * par) * nb, nb, st->halo_total,
                  st->halo_buf, 0, st->stream);
  launch_box_copy(st->halo_tab + (2 * par + 1) * nb, nb, st->halo_total,
                  st->halo_buf, 1, st->stream);
}

}  // extern "C"

namespace dcx {
void fused_step(State *st, float dx, float dy, float dz, float dt,
                long long src_index, float src_next, int add_next,
                void (*after_stencil)(void *), void *ctx) {
  StepArgs a = make_step_args(st, full_box(st), dx, dy, dz, dt);
  const bool defer = add_next && source_in_send_box(st, src_index);
  if (add_next && !defer) set_source(st, a, src_index, src_next);
  layer_infer(st, a);
  wall_save(st);
  launch_stencil(st, a);
  if (after_stencil) after_stencil(ctx);
  layer_apply(st);
  wall_apply(st, dx, dy, dz, dt);
  halo_loopback(st);
  advance(st);
  if (defer) add_deferred_source(st, src_index, src_next);
}
}  // namespace dcx

extern "C" {

void dcx_step(dc_device_data *data, float dx, float dy, float dz, float dt,
              long long src_index, float src_next, int add_next) {
  fused_step(state_of(data), dx, dy, dz, dt, src_index, src_next, add_next,
             nullptr, nullptr);
}

void dcx_set_timing_detail(dc_device_data *data, int enabled) {
  state_of(data)->timing_detail = enabled != 0;
}

double dcx_last_kernel_ms(const dc_device_data *data) {
  return state_of(data)->last_kernel_ms;
}

double dcx_stream_probe(dc_device_data *data, int reps) {
  State *st = state_of(data);
  if (st->fstore != DCX_STORE_F32 || st->cstore != DCX_COEF_F32)
    DCX_FAIL("the stream probe reads FP32 fields and coefficients");
  if (reps < 1) DCX_FAIL("reps must be positive");
  const StepArgs a = make_step_args(st, full_box(st), 1.0f, 1.0f, 1.0f, 0.0f);
  launch_stream_probe(a, st->stream);  // untimed first pass
  DCX_CHECK(hipEventRecord(st->ev[0], st->stream));
  for (int r = 0; r < reps; ++r) launch_stream_probe(a, st->stream);
  DCX_CHECK(hipEventRecord(st->ev[1], st->stream));
  DCX_CHECK(hipEventSynchronize(st->ev[1]));
  float ms = 0.0f;
  DCX_CHECK(hipEventElapsedTime(&ms, st->ev[0], st->ev[1]));
  return (double)ms / reps;
}

}  // extern "C"

// Receiver sampling after each step: out[step * n + r] = p at dev_idx[r]
// (field units converted back to true values).
template <int FS>
__global__ void k_gather(const void *field, const int64_t *__restrict__ idx,
                         int n, float *__restrict__ out, float inv_scale) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  float v;
  if (FS == DCX_STORE_F32) {
    v = static_cast<const float *>(field)[idx[r]];
  } else if (FS == DCX_STORE_F16) {
    v = (float)__builtin_bit_cast(_Float16, static_cast<const uint16_t *>(field)[idx[r]]);
  } else {
    v = __uint_as_float((uint32_t)static_cast<const uint16_t *>(field)[idx[r]] << 16);
  }
  out[r] = v * inv_scale;
}

struct TraceCtx {
  int n;
  const int64_t *idx;  // device offsets
  float *buf;          // device, n_steps * n
};

static void gather(State *st, const TraceCtx &tc, unsigned it) {
  const void *field = st->fstore == DCX_STORE_F32
                          ? (const void *)st->abi.pc
                          : (const void *)st->P16[0][st->pcur];
  const float inv = st->fstore == DCX_STORE_F32 ? 1.0f : 1.0f / st->fscale;
  float *out = tc.buf + (size_t)it * tc.n;
  const dim3 g((tc.n + 255) / 256), b(256);
  if (st->fstore == DCX_STORE_F32)
    hipLaunchKernelGGL(k_gather<DCX_STORE_F32>, g, b, 0, st->stream, field, tc.idx, tc.n, out, inv);
  else if (st->fstore == DCX_STORE_F16)
    hipLaunchKernelGGL(k_gather<DCX_STORE_F16>, g, b, 0, st->stream, field, tc.idx, tc.n, out, inv);
  else
    hipLaunchKernelGGL(k_gather<DCX_STORE_BF16>, g, b, 0, st->stream, field, tc.idx, tc.n, out, inv);
  DCX_CHECK(hipGetLastError());
}

static double run_impl(State *st, float dx, float dy, float dz, float dt,
                       unsigned int n_steps, long long src_index,
                       const float *src_amp, int use_graph, const TraceCtx *tc) {
  dc_device_data *data = &st->abi;
  const bool src = src_index >= 0 && src_amp != nullptr;
  if (src && src_index >= st->n) DCX_FAIL("source index out of range");
  const bool defer = src && source_in_send_box(st, src_index);
  // Step it writes the fields that step it+1 reads as pc/qc; the source of
  // step it+1 is added as they are written (see dcx_step).
  auto enqueue = [&](std::vector<hipEvent_t> *evs) {
    if (src && n_steps > 0) {
      if (st->fstore == DCX_STORE_F32)
        launch_add_source(st->abi.pc, st->abi.qc, device_index(st, src_index),
                          src_amp[0], st->stream);
      else
        storage_add_source(st, device_index(st, src_index), src_amp[0]);
    }
    for (unsigned it = 0; it < n_steps; ++it) {
      StepArgs a = make_step_args(st, full_box(st), dx, dy, dz, dt);
      if (src && !defer && it + 1 < n_steps)
        set_source(st, a, src_index, src_amp[it + 1]);
      layer_infer(st, a);
      wall_save(st);
      if (evs) DCX_CHECK(hipEventRecord((*evs)[2 * it], st->stream));
      launch_stencil(st, a);
      if (evs) DCX_CHECK(hipEventRecord((*evs)[2 * it + 1], st->stream));
      layer_apply(st);
      wall_apply(st, dx, dy, dz, dt);
      halo_loopback(st);
      advance(st);
      if (tc) gather(st, *tc, it);
      if (defer && it + 1 < n_steps)
        add_deferred_source(st, src_index, src_amp[it + 1]);
    }
  };
  float ms = 0.f;
  if (use_graph) {
    // Capture the whole loop, then replay it once.  Pointer swaps happen on
    // the host during capture; the captured graph already encodes them, so
    // the replay starts from the pre-capture state and ends in the captured
    // final state.
    dc_device_data before = st->abi;
    const int pcur_before = st->pcur;
    hipGraph_t g;
    hipGraphExec_t ge;
    DCX_CHECK(hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal));
    enqueue(nullptr);
    DCX_CHECK(hipStreamEndCapture(st->stream, &g));
    DCX_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    dc_device_data after = st->abi;
    const int pcur_after = st->pcur;
    st->abi = before;
    st->pcur = pcur_before;
    DCX_CHECK(hipEventRecord(st->ev[0], st->stream));
    DCX_CHECK(hipGraphLaunch(ge, st->stream));
    DCX_CHECK(hipEventRecord(st->ev[1], st->stream));
    DCX_CHECK(hipEventSynchronize(st->ev[1]));
    st->abi = after;
    st->pcur = pcur_after;
    DCX_CHECK(hipEventElapsedTime(&ms, st->ev[0], st->ev[1]));
    DCX_CHECK(hipGraphExecDestroy(ge));
    DCX_CHECK(hipGraphDestroy(g));
  } else if (st->timing_detail && n_steps > 0) {
    std::vector<hipEvent_t> evs(2 * (size_t)n_steps);
    for (auto &e : evs) DCX_CHECK(hipEventCreate(&e));
    DCX_CHECK(hipEventRecord(st->ev[0], st->stream));
    enqueue(&evs);
    DCX_CHECK(hipEventRecord(st->ev[1], st->stream));
    DCX_CHECK(hipEventSynchronize(st->ev[1]));
    DCX_CHECK(hipEventElapsedTime(&ms, st->ev[0], st->ev[1]));
    double sum = 0.0;
    for (unsigned it = 0; it < n_steps; ++it) {
      float k = 0.f;
      DCX_CHECK(hipEventElapsedTime(&k, evs[2 * it], evs[2 * it + 1]));
      sum += k;
    }
    st->last_kernel_ms = sum / n_steps;
    for (auto &e : evs) DCX_CHECK(hipEventDestroy(e));
  } else {
    DCX_CHECK(hipEventRecord(st->ev[0], st->stream));
    enqueue(nullptr);
    DCX_CHECK(hipEventRecord(st->ev[1], st->stream));
    DCX_CHECK(hipEventSynchronize(st->ev[1]));
    DCX_CHECK(hipEventElapsedTime(&ms, st->ev[0], st->ev[1]));
  }
  (void)data;
  report_overflow(st);
  return (double)ms;
}

extern "C" {

double dcx_run(dc_device_data *data, float dx, float dy, float dz, float dt,
               unsigned int n_steps, long long src_index, const float *src_amp,
               int use_graph) {
  return run_impl(state_of(data), dx, dy, dz, dt, n_steps, src_index, src_amp,
                  use_graph, nullptr);
}

double dcx_run_traces(dc_device_data *data, float dx, float dy, float dz,
                      float dt, unsigned int n_steps, long long src_index,
                      const float *src_amp, int n_rec,
                      const long long *receivers, float *traces,
                      int use_graph) {
  State *st = state_of(data);
  if (n_rec <= 0 || n_steps == 0)
    return run_impl(st, dx, dy, dz, dt, n_steps, src_index, src_amp, use_graph,
                    nullptr);
  std::vector<int64_t> idx(n_rec);
  for (int r = 0; r < n_rec; ++r) {
    if (receivers[r] < 0 || receivers[r] >= st->n)
      DCX_FAIL("receiver %d index %lld out of range", r, receivers[r]);
    idx[r] = device_index(st, receivers[r]);
  }
  TraceCtx tc{n_rec, nullptr, nullptr};
  int64_t *d_idx = nullptr;
  float *d_buf = nullptr;
  DCX_CHECK(hipMalloc(&d_idx, n_rec * sizeof(int64_t)));
  DCX_CHECK(hipMalloc(&d_buf, (size_t)n_rec * n_steps * sizeof(float)));
  DCX_CHECK(hipMemcpy(d_idx, idx.data(), n_rec * sizeof(int64_t),
                      hipMemcpyHostToDevice));
  tc.idx = d_idx;
  tc.buf = d_buf;
  const double ms = run_impl(st, dx, dy, dz, dt, n_steps, src_index, src_amp,
                             use_graph, &tc);
  DCX_CHECK(hipMemcpy(traces, d_buf, (size_t)n_rec * n_steps * sizeof(float),
                      hipMemcpyDeviceToHost));
  DCX_CHECK(hipFree(d_idx));
  DCX_CHECK(hipFree(d_buf));
  return ms;
}

float *dcx_scratch(dc_device_data *data, size_t n_floats) {
  State *st = state_of(data);
  DCX_CHECK(hipStreamSynchronize(st->stream));
  return ensure(st->scratch, st->scratch_n, n_floats);
}

static void boxes_of(State *st, int nbox, const size_t *starts,
                     const size_t *ends, float *arr, std::vector<BoxCopy> &out,
                     int64_t &total) {
  out.resize(nbox);
  total = 0;
  for (int k = 0; k < nbox; ++k) {
    out[k] = box_of(st, starts + 3 * k, ends + 3 * k, arr);
    check_multi_box(out[k]);
    out[k].off = total;
    total += out[k].nx * out[k].ny * out[k].nz;
  }
}

static void copy_boxes(State *st, float *dev_buf, int nbox, const size_t *starts,
                       const size_t *ends, float *arr, int dir) {
  std::vector<BoxCopy> boxes;
  int64_t total = 0;
  boxes_of(st, nbox, starts, ends, arr, boxes, total);
  if (total == 0) return;
  BoxCopy *d_boxes = nullptr;
  DCX_CHECK(hipMallocAsync((void **)&d_boxes, nbox * sizeof(BoxCopy), st->stream));
  DCX_CHECK(hipMemcpyAsync(d_boxes, boxes.data(), nbox * sizeof(BoxCopy),
                           hipMemcpyHostToDevice, st->stream));
  launch_box_copy(d_boxes, nbox, total, dev_buf, dir, st->stream);
  DCX_CHECK(hipFreeAsync(d_boxes, st->stream));
  DCX_CHECK(hipStreamSynchronize(st->stream));
}

void dcx_set_halo_loopback(dc_device_data *data, int nbox,
                           const size_t *send_starts, const size_t *send_ends,
                           const size_t *recv_starts, const size_t *recv_ends) {
  State *st = state_of(data);
  if (st->batch)
    DCX_FAIL("dcx_set_halo_loopback on a batch shot: batched launches carry "
             "no halo exchange (dcx_batch_free first)");
  if (nbox > 0 && st->wall != DCX_WALL_ZERO)
    DCX_FAIL("dcx_set_halo_loopback with a one-way wall set: the receive "
             "boxes lie in the frame the wall fills");
  std::free(st->halo_user);
  st->halo_user = nullptr;
  st->halo_nuser = nbox > 0 ? nbox : 0;
  if (nbox > 0) {
    st->halo_user = static_cast<size_t *>(std::malloc(sizeof(size_t) * 12 * nbox));
    if (!st->halo_user) DCX_FAIL("OOM");
    for (int k = 0; k < nbox; ++k)
      for (int a = 0; a < 3; ++a) {
        st->halo_user[12 * k + a] = send_starts[3 * k + a];
        st->halo_user[12 * k + 3 + a] = send_ends[3 * k + a];
        st->halo_user[12 * k + 6 + a] = recv_starts[3 * k + a];
        st->halo_user[12 * k + 9 + a] = recv_ends[3 * k + a];
      }
  }
  build_halo_tables(st);
}

size_t dcx_copy_field(dc_device_data *data, int dst, int src) {
  State *st = state_of(data);
  const int d = dst == 0 ? 0 : dst == 1 ? 1 : dst == 2 ? 2 : 3;
  const int s = src == 0 ? 0 : src == 1 ? 1 : src == 2 ? 2 : 3;
  if (d != dst || s != src || d == s) DCX_FAIL("bad field pair %d <- %d", dst, src);
  // Whole allocations, contiguous and 256-byte aligned: every byte of both
  // is streamed exactly once.
  BoxCopy b{};
  b.arr = st->field_alloc[s];
  b.nx = st->alloc_floats;
  b.ny = 1;
  b.nz = 1;
  launch_box_copy1(b, st->field_alloc[d], 0, st->stream);
  return (size_t)st->alloc_floats * sizeof(float);
}

void dcx_pack_boxes(dc_device_data *data, float *dev_buf, int nbox,
                    const size_t *starts, const size_t *ends,
                    const float *from) {
  State *st = state_of(data);
  copy_boxes(st, dev_buf, nbox, starts, ends, field_of(st, from), 0);
}

void dcx_unpack_boxes(dc_device_data *data, const float *dev_buf, int nbox,
                      const size_t *starts, const size_t *ends, float *to) {
  State *st = state_of(data);
  copy_boxes(st, const_cast<float *>(dev_buf), nbox, starts, ends,
             field_of(st, to), 1);
}

}  // extern "C"
