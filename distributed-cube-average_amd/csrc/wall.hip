// One-way wall for a whole-domain run (dcx_set_wall).
//
// The reference keeps the 4-cell frame beyond the compute region at zero
// (its outer frame of the random-velocity boundary, src/boundary.c:7-104):
// a rigid wall that returns every wave that reaches it.  With the wall on,
// each step fills the frame cells of the new level from the outgoing wave
// instead, by Higdon's discrete one-way conditions, outward one frame layer
// at a time.  With K the shift to F's neighbour towards the compute region
// and Z the shift back one time level, the first-order condition at frame
// cell F of the new level n+1 is
//   B u = 0,  B = (1 - r K) + Z (r - K),  r = (c dt - h) / (c dt + h):
//   u_F^{n+1} = u_I^n + r (u_I^{n+1} - u_F^n),
// and the second-order one is B B u = 0, evaluated as the first-order
// condition on the first-order residual v = B u, with
// pred(a, b, c) = a + r (b - c):
//   u_F^{n+1} = pred(u_I^n, u_I^{n+1}, u_F^n) + pred(v_I^n, v_I^{n+1}, v_F^n),
//   v_G^m = u_G^m - pred(u_{G-1}^{m-1}, u_{G-1}^m, u_G^{m-1})
// (G-1 one cell towards the compute region): every term a difference of
// neighbouring values, so a field at rest stays exactly at rest.  h is the
// spacing along the face normal and c the qP speed along that normal at the
// wall, c^2 = v2px (1 - c_nn) + v2pz c_nn, c_nn the angle term of the
// normal's axis (cxx, cyy or czz: the squared cosine between the normal and
// the symmetry axis), read at the last compute cell.  B is exact for
// outgoing waves u = f(x - c t) with f linear or quadratic (B B for any
// wave u = f linear, whatever r); a plane wave at
// angle a to the normal comes back with amplitude ((1-cos a)/(1+cos a))^k
// for order k in the continuum.
//
// The faces are filled axis by axis: x faces over the compute range of y and
// z, then y faces over x's whole range (its frames included), then z faces
// over the whole x and y ranges, so that edge and corner frame cells take
// the condition along the last axis whose frame they are in.  Every frame
// cell belongs to one of these passes only.
//
// Each face cell's c is computed once, when the wall is set (the
// coefficients do not change afterwards), into the struct's wall buffer,
// one float per face cell in the passes' order: the fill then reads it
// coalesced instead of three coefficient arrays at the last compute cell.
//
// The second-order condition reads level n-1 at the last two compute layers
// and in the frame.  The stencil overwrites the compute cells' n-1 (the pp
// array becomes n+1), so wall_save copies those two layers of every pass's
// face into the struct's wall buffer before the stencil; the frame's n-1 is
// still in the pp array's frame when the fill reads it, before writing.
// FP32 fields.
#include "dc_internal.hpp"

namespace dcx {

struct WallArgs {
  float *pn, *qn;          // new level: the stencil's output (pp, qp arrays)
  const float *pc, *qc;    // current level (frame from the last wall pass)
  const float *cnn[3];     // cxx, cyy, czz
  const float *v2px, *v2pz;
  float *speed;            // c at each face cell: [pass][side][v][u]
  int64_t soff[3];         // each pass's block in speed
  float *aux;              // order 2: level n-1 of the last two layers
  int64_t aoff[3];         // each pass's block in aux
  int64_t ys, zs;
  int64_t n[3];            // array sizes
  float dt;
  float h[3];              // dx, dy, dz
};

// Face extents of pass `axis`: axes filled before it over their whole
// range, later ones over the compute range.
__host__ __device__ inline void pass_range(const int64_t n[3], int axis,
                                           int a, int64_t &lo, int64_t &cnt) {
  lo = a < axis ? 0 : STENCIL;
  cnt = a < axis ? n[a] : n[a] - 2 * STENCIL;
}

// Cells of one pass's face block in aux: [side][layer][field][v][u].
__host__ __device__ inline int64_t pass_cells(const int64_t n[3], int axis) {
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  int64_t l1, c1, l2, c2;
  pass_range(n, axis, a1, l1, c1);
  pass_range(n, axis, a2, l2, c2);
  return c1 * c2;
}

// The thread's face cell (u, v) of pass `axis`, or false past the face.
struct FaceCell {
  int a1, a2;
  int64_t u, v, iu, iv, nu, nv;  // coordinates, in-face indices, face extents
};
__device__ inline bool face_cell(const WallArgs &w, int axis, FaceCell &f) {
  f.a1 = axis == 0 ? 1 : 0;
  f.a2 = axis == 2 ? 1 : 2;
  int64_t lo1, lo2;
  pass_range(w.n, axis, f.a1, lo1, f.nu);
  pass_range(w.n, axis, f.a2, lo2, f.nv);
  f.iu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  f.iv = blockIdx.y;
  if (f.iu >= f.nu || f.iv >= f.nv) return false;
  f.u = lo1 + f.iu;
  f.v = lo2 + f.iv;
  return true;
}

// aux index of (side, layer, field) at the thread's face cell.
__device__ inline int64_t aux_at(const WallArgs &w, int axis, const FaceCell &f,
                                 int side, int layer, int field) {
  return w.aoff[axis] +
         (((int64_t)(side * 2 + layer) * 2 + field) * f.nv + f.iv) * f.nu + f.iu;
}

// c at each face cell of both faces of pass `axis`, from the coefficients at
// the last compute cell along the normal (u and v clamped into the compute
// range).
__global__ void __launch_bounds__(256) k_wall_speed(const WallArgs w, int axis) {
#pragma clang fp contract(off)
  FaceCell f;
  if (!face_cell(w, axis, f)) return;
  const int side = blockIdx.z;
  const int64_t st[3] = {1, w.ys, w.zs};
  const int64_t uc = min(max(f.u, (int64_t)STENCIL), w.n[f.a1] - STENCIL - 1);
  const int64_t vc = min(max(f.v, (int64_t)STENCIL), w.n[f.a2] - STENCIL - 1);
  const int64_t last = side ? w.n[axis] - STENCIL - 1 : STENCIL;
  const int64_t ic = last * st[axis] + uc * st[f.a1] + vc * st[f.a2];
  const float cn = w.cnn[axis][ic];
  const float c2 = w.v2px[ic] * (1.0f - cn) + w.v2pz[ic] * cn;
  w.speed[w.soff[axis] + (side * f.nv + f.iv) * f.nu + f.iu] = sqrtf(c2);
}

// Order 2: level n-1 of the last two compute layers of both faces of pass
// `axis` (layer 0 the one further in), from the pp/qp arrays before the
// stencil overwrites them.
__global__ void __launch_bounds__(256) k_wall_save(const WallArgs w, int axis) {
  FaceCell f;
  if (!face_cell(w, axis, f)) return;
  const int side = blockIdx.z;
  const int64_t st[3] = {1, w.ys, w.zs};
  const int64_t s = st[axis];
  const int64_t last = side ? w.n[axis] - STENCIL - 1 : STENCIL;
  const int64_t dir = side ? 1 : -1;
  const int64_t base = f.u * st[f.a1] + f.v * st[f.a2];
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const int64_t i = base + (last - dir * (1 - layer)) * s;
    w.aux[aux_at(w, axis, f, side, layer, 0)] = w.pn[i];
    w.aux[aux_at(w, axis, f, side, layer, 1)] = w.qn[i];
  }
}

// k_wall_save for the x faces, four lanes per row (the two layers of p and
// of q), for the reason given at k_wall_x.
__global__ void __launch_bounds__(256) k_wall_save_x(const WallArgs w) {
  const int q = threadIdx.x & 3;
  const int layer = q & 1, fl = q >> 1;
  const int64_t nu = w.n[1] - 2 * STENCIL, nv = w.n[2] - 2 * STENCIL;
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x / 4) + (threadIdx.x >> 2);
  if (row >= 2 * nu * nv) return;
  const int side = row >= nu * nv;
  const int64_t rr = row - (side ? nu * nv : 0);
  const int64_t iv = rr / nu, iu = rr - iv * nu;
  const int64_t last = side ? w.n[0] - STENCIL - 1 : STENCIL;
  const int64_t dir = side ? 1 : -1;
  const int64_t i = (STENCIL + iu) * w.ys + (STENCIL + iv) * w.zs + last -
                    dir * (1 - layer);
  w.aux[w.aoff[0] + (((int64_t)(side * 2 + layer) * 2 + fl) * nv + iv) * nu + iu] =
      (fl ? w.qn : w.pn)[i];
}

template <int ORDER>
__global__ void __launch_bounds__(256) k_wall(const WallArgs w, int axis) {
#pragma clang fp contract(off)
  FaceCell f;
  if (!face_cell(w, axis, f)) return;
  const int side = blockIdx.z;  // 0: low face, 1: high face
  const int64_t st[3] = {1, w.ys, w.zs};
  const int64_t s = st[axis];
  const int64_t last = side ? w.n[axis] - STENCIL - 1 : STENCIL;
  const float cdt = w.speed[w.soff[axis] + (side * f.nv + f.iv) * f.nu + f.iu] * w.dt;
  const float r = (cdt - w.h[axis]) / (cdt + w.h[axis]);
  const int64_t dir = side ? 1 : -1;  // outward along the normal
  const int64_t base = f.u * st[f.a1] + f.v * st[f.a2];
  // Position j = 0..5 along the normal: the last two compute cells, then
  // the four frame cells outward.
  auto at = [&](int j) { return base + (last + dir * (j - 1)) * s; };
  auto pred = [r](float a, float b, float c) {
#pragma clang fp contract(off)
    return a + r * (b - c);
  };
  if constexpr (ORDER == 1) {
    // All loads first, the recurrence in registers, then the stores (the
    // compiler cannot reorder the stores past loads of arrays that might
    // alias them).
    float np_[6], nq_[6], cp_[6], cq_[6];
#pragma unroll
    for (int j = 1; j < 6; ++j) {
      cp_[j] = w.pc[at(j)];
      cq_[j] = w.qc[at(j)];
    }
    np_[1] = w.pn[at(1)];
    nq_[1] = w.qn[at(1)];
#pragma unroll
    for (int j = 2; j < 6; ++j) {
      np_[j] = pred(cp_[j - 1], np_[j - 1], cp_[j]);
      nq_[j] = pred(cq_[j - 1], nq_[j - 1], cq_[j]);
    }
#pragma unroll
    for (int j = 2; j < 6; ++j) {
      w.pn[at(j)] = np_[j];
      w.qn[at(j)] = nq_[j];
    }
  } else {
    float *const nw[2] = {w.pn, w.qn};
    const float *const cu[2] = {w.pc, w.qc};
#pragma unroll
    for (int fl = 0; fl < 2; ++fl) {
      float nv[6], cv[6], ov[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        cv[j] = cu[fl][at(j)];
        // Level n-1: saved for the compute layers, still in the frame.
        ov[j] = j < 2 ? w.aux[aux_at(w, axis, f, side, j, fl)] : nw[fl][at(j)];
        nv[j] = j < 2 ? nw[fl][at(j)] : 0.0f;
      }
#pragma unroll
      for (int j = 2; j < 6; ++j) {
        const float v_in1 = nv[j - 1] - pred(cv[j - 2], nv[j - 2], cv[j - 1]);
        const float v_in = cv[j - 1] - pred(ov[j - 2], cv[j - 2], ov[j - 1]);
        const float v_fn = cv[j] - pred(ov[j - 1], cv[j - 1], ov[j]);
        nv[j] = pred(cv[j - 1], nv[j - 1], cv[j]) + pred(v_in, v_in1, v_fn);
        nw[fl][at(j)] = nv[j];
      }
    }
  }
}

// The x faces with eight lanes per row along x: lane q holds position j = q
// of the fill (two idle lanes), so a wavefront's accesses cover 8 rows, each
// row's six cells in one or two lines, where one lane per row (k_wall, whose
// lanes run along y) made every lane touch lines of its own.  The outward
// recurrence passes each new value to the next lane by a shuffle.  Same
// arithmetic as k_wall, value for value.
template <int ORDER>
__global__ void __launch_bounds__(256) k_wall_x(const WallArgs w) {
#pragma clang fp contract(off)
  const int q = threadIdx.x & 7;
  const int64_t nu = w.n[1] - 2 * STENCIL, nv = w.n[2] - 2 * STENCIL;
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x / 8) + (threadIdx.x >> 3);
  const bool live = row < 2 * nu * nv;
  const int64_t r0 = live ? row : 0;  // rows past the face load row 0, store nothing
  const int side = r0 >= nu * nv;
  const int64_t rr = r0 - (side ? nu * nv : 0);
  const int64_t iv = rr / nu, iu = rr - iv * nu;
  const int64_t last = side ? w.n[0] - STENCIL - 1 : STENCIL;
  const int64_t dir = side ? 1 : -1;
  const int j = q < 6 ? q : 5;  // idle lanes load what lane 5 loads
  const int64_t at = (STENCIL + iu) * w.ys + (STENCIL + iv) * w.zs + last +
                     dir * (j - 1);
  const float cdt = w.speed[w.soff[0] + (side * nv + iv) * nu + iu] * w.dt;
  const float r = (cdt - w.h[0]) / (cdt + w.h[0]);
  auto pred = [r](float a, float b, float c) {
#pragma clang fp contract(off)
    return a + r * (b - c);
  };
  const int q1 = q > 0 ? q - 1 : 0, q2 = q > 1 ? q - 2 : 0;
  float *const nw[2] = {w.pn, w.qn};
  const float *const cu[2] = {w.pc, w.qc};
#pragma unroll
  for (int fl = 0; fl < 2; ++fl) {
    const float c = cu[fl][at];
    float n = j < 2 ? nw[fl][at] : 0.0f;
    float o = 0.0f;
    if constexpr (ORDER == 2) {
      // Level n-1: saved for the compute layers, still in the frame.
      o = j < 2 ? w.aux[w.aoff[0] + (((int64_t)(side * 2 + j) * 2 + fl) * nv + iv) * nu + iu]
                : nw[fl][at];
    }
    const float c1 = __shfl(c, q1, 8), c2 = __shfl(c, q2, 8);
    const float o1 = __shfl(o, q1, 8), o2 = __shfl(o, q2, 8);
#pragma unroll
    for (int k = 2; k < 6; ++k) {
      const float n1 = __shfl(n, k - 1, 8), n2 = __shfl(n, k - 2, 8);
      if (q == k) {
        if constexpr (ORDER == 1) {
          n = pred(c1, n1, c);
        } else {
          const float v_in1 = n1 - pred(c2, n2, c1);
          const float v_in = c1 - pred(o2, c2, o1);
          const float v_fn = c - pred(o1, c1, o);
          n = pred(c1, n1, c) + pred(v_in, v_in1, v_fn);
        }
      }
    }
    if (live && q >= 2 && q < 6) nw[fl][at] = n;
  }
}

static WallArgs wall_args(const State *st, float dx, float dy, float dz,
                          float dt) {
  WallArgs w;
  w.pn = st->abi.pp;
  w.qn = st->abi.qp;
  w.pc = st->abi.pc;
  w.qc = st->abi.qc;
  w.cnn[0] = coef_array(st, CXX);
  w.cnn[1] = coef_array(st, CYY);
  w.cnn[2] = coef_array(st, CZZ);
  w.v2px = coef_array(st, V2PX);
  w.v2pz = coef_array(st, V2PZ);
  w.ys = st->ys;
  w.zs = st->zs;
  for (int a = 0; a < 3; ++a) w.n[a] = st->sizes[a];
  // The wall buffer: the face speeds, then (order 2) the saved layers.
  int64_t off = 0;
  for (int a = 0; a < 3; ++a) {
    w.soff[a] = off;
    off += 2 * pass_cells(w.n, a);
  }
  w.speed = st->wall_aux;
  w.aux = st->wall_aux + off;
  off = 0;
  for (int a = 0; a < 3; ++a) {
    w.aoff[a] = off;
    off += 8 * pass_cells(w.n, a);
  }
  w.dt = dt;
  w.h[0] = dx;
  w.h[1] = dy;
  w.h[2] = dz;
  return w;
}

static dim3 pass_grid(const int64_t n[3], int axis) {
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  int64_t l, n1, n2;
  pass_range(n, axis, a1, l, n1);
  pass_range(n, axis, a2, l, n2);
  return dim3((unsigned)((n1 + 255) / 256), (unsigned)n2, 2);
}

static void check_wall_storage(const State *st) {
  if (st->fstore != DCX_STORE_F32 || st->cstore != DCX_COEF_F32)
    DCX_FAIL("the one-way wall runs on FP32 storage");
}

void wall_save(State *st) {
  if (st->wall != DCX_WALL_ONE_WAY2) return;
  check_wall_storage(st);
  const WallArgs w = wall_args(st, 1.f, 1.f, 1.f, 1.f);
  const int64_t xrows = 2 * (w.n[1] - 2 * STENCIL) * (w.n[2] - 2 * STENCIL);
  hipLaunchKernelGGL(k_wall_save_x, dim3((unsigned)((xrows + 63) / 64)),
                     dim3(256), 0, st->stream, w);
  DCX_CHECK(hipGetLastError());
  for (int axis = 1; axis < 3; ++axis) {
    hipLaunchKernelGGL(k_wall_save, pass_grid(w.n, axis), dim3(256), 0,
                       st->stream, w, axis);
    DCX_CHECK(hipGetLastError());
  }
}

void wall_apply(State *st, float dx, float dy, float dz, float dt) {
  if (st->wall == DCX_WALL_ZERO) return;
  check_wall_storage(st);
  const WallArgs w = wall_args(st, dx, dy, dz, dt);
  const int64_t xrows = 2 * (w.n[1] - 2 * STENCIL) * (w.n[2] - 2 * STENCIL);
  const dim3 xgrid((unsigned)((xrows + 31) / 32));
  if (st->wall == DCX_WALL_ONE_WAY)
    hipLaunchKernelGGL(k_wall_x<1>, xgrid, dim3(256), 0, st->stream, w);
  else
    hipLaunchKernelGGL(k_wall_x<2>, xgrid, dim3(256), 0, st->stream, w);
  DCX_CHECK(hipGetLastError());
  for (int axis = 1; axis < 3; ++axis) {
    if (st->wall == DCX_WALL_ONE_WAY)
      hipLaunchKernelGGL(k_wall<1>, pass_grid(w.n, axis), dim3(256), 0,
                         st->stream, w, axis);
    else
      hipLaunchKernelGGL(k_wall<2>, pass_grid(w.n, axis), dim3(256), 0,
                         st->stream, w, axis);
    DCX_CHECK(hipGetLastError());
  }
}

void wall_free(State *st) {
  if (st->wall_aux) DCX_CHECK(hipFree(st->wall_aux));
  st->wall_aux = nullptr;
  st->wall = DCX_WALL_ZERO;
}

}  // namespace dcx

using namespace dcx;

extern "C" int dcx_set_wall(dc_device_data *data, int mode) {
  State *st = state_of(data);
  if (mode != DCX_WALL_ZERO && mode != DCX_WALL_ONE_WAY &&
      mode != DCX_WALL_ONE_WAY2)
    DCX_FAIL("dcx_set_wall: unknown mode %d", mode);
  if (mode != DCX_WALL_ZERO) {
    if (st->batch) DCX_FAIL("a batch shot cannot carry a one-way wall");
    if (st->halo_nuser > 0)
      DCX_FAIL("dcx_set_wall: the frame holds a registered halo exchange's "
               "receive boxes; a one-way wall is for a whole-domain run");
    check_wall_storage(st);
    for (int a = 0; a < 3; ++a)
      if (st->sizes[a] < 2 * STENCIL + 2)
        DCX_FAIL("dcx_set_wall: fewer than two compute cells along axis %d", a);
  }
  DCX_CHECK(hipStreamSynchronize(st->stream));
  wall_free(st);
  if (mode != DCX_WALL_ZERO) {
    int64_t cells = 0;
    for (int a = 0; a < 3; ++a)
      cells += (mode == DCX_WALL_ONE_WAY2 ? 10 : 2) * pass_cells(st->sizes, a);
    DCX_CHECK(hipMalloc(&st->wall_aux, cells * sizeof(float)));
    const WallArgs w = wall_args(st, 1.f, 1.f, 1.f, 1.f);
    for (int axis = 0; axis < 3; ++axis) {
      hipLaunchKernelGGL(k_wall_speed, pass_grid(w.n, axis), dim3(256), 0,
                         st->stream, w, axis);
      DCX_CHECK(hipGetLastError());
    }
    DCX_CHECK(hipStreamSynchronize(st->stream));
  }
  st->wall = mode;
  return 0;
}
