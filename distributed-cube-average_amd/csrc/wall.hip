// One-way wall for a whole-domain run (dcx_set_wall).
//
// The reference keeps the 4-cell frame beyond the compute region at zero
// (its outer frame of the random-velocity boundary, src/boundary.c:7-104):
// a rigid wall that returns every wave that reaches it.  With the wall on,
// each step fills the frame cells of the new level from the outgoing wave
// instead, by the first-order one-way condition (d/dt + c d/dn) u = 0 in
// Higdon's two-level form, applied outward one frame layer at a time:
//   u_F^{n+1} = u_I^n + r (u_I^{n+1} - u_F^n),  r = (c dt - h) / (c dt + h),
// where I is F's neighbour towards the compute region, h the spacing along
// the face normal and c the qP speed along that normal at the wall,
//   c^2 = v2px (1 - c_nn) + v2pz c_nn,
// c_nn the angle term of the normal's axis (cxx, cyy or czz: the squared
// cosine between the normal and the symmetry axis), read at the last compute
// cell.  The condition is exact for any wave u = f(x - c t) with f linear,
// and absorbs a normally incident wave best at low frequency, where a thin
// damping layer absorbs least.
//
// The faces are filled axis by axis: x faces over the compute range of y and
// z, then y faces over x's whole range (its frames included), then z faces
// over the whole x and y ranges, so that edge and corner frame cells take
// the condition along the last axis whose frame they are in.  FP32 fields.
#include "dc_internal.hpp"

namespace dcx {

struct WallArgs {
  float *pn, *qn;          // new level: the stencil's output
  const float *pc, *qc;    // current level (frame from the last wall pass)
  const float *cnn[3];     // cxx, cyy, czz
  const float *v2px, *v2pz;
  int64_t ys, zs;
  int64_t n[3];            // array sizes
  float dt;
  float h[3];              // dx, dy, dz
};

__global__ void __launch_bounds__(256) k_wall(const WallArgs w, int axis) {
#pragma clang fp contract(off)
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  // Axes filled before this one run over their whole range (frames
  // included); later ones over the compute range.
  const int64_t lo1 = a1 < axis ? 0 : STENCIL;
  const int64_t hi1 = a1 < axis ? w.n[a1] : w.n[a1] - STENCIL;
  const int64_t lo2 = a2 < axis ? 0 : STENCIL;
  const int64_t hi2 = a2 < axis ? w.n[a2] : w.n[a2] - STENCIL;
  const int64_t u = lo1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t v = lo2 + (int64_t)blockIdx.y;
  if (u >= hi1 || v >= hi2) return;
  const int side = blockIdx.z;  // 0: low face, 1: high face
  const int64_t st[3] = {1, w.ys, w.zs};
  const int64_t s = st[axis], s1 = st[a1], s2 = st[a2];
  // The coefficients of the wall: the last compute cell along the normal,
  // u and v clamped into the compute range.
  const int64_t uc = u < STENCIL ? STENCIL
                     : u >= w.n[a1] - STENCIL ? w.n[a1] - STENCIL - 1 : u;
  const int64_t vc = v < STENCIL ? STENCIL
                     : v >= w.n[a2] - STENCIL ? w.n[a2] - STENCIL - 1 : v;
  const int64_t last = side ? w.n[axis] - STENCIL - 1 : STENCIL;
  const int64_t ic = last * s + uc * s1 + vc * s2;
  const float cn = w.cnn[axis][ic];
  const float c2 = w.v2px[ic] * (1.0f - cn) + w.v2pz[ic] * cn;
  const float cdt = sqrtf(c2) * w.dt;
  const float r = (cdt - w.h[axis]) / (cdt + w.h[axis]);
  const int64_t dir = side ? 1 : -1;  // outward along the normal
  const int64_t base = u * s1 + v * s2;
#pragma unroll
  for (int k = 1; k <= STENCIL; ++k) {
    const int64_t F = base + (last + dir * k) * s;
    const int64_t I = F - dir * s;
    w.pn[F] = w.pc[I] + r * (w.pn[I] - w.pc[F]);
    w.qn[F] = w.qc[I] + r * (w.qn[I] - w.qc[F]);
  }
}

void wall_apply(State *st, float dx, float dy, float dz, float dt) {
  if (st->wall == 0) return;
  if (st->fstore != DCX_STORE_F32 || st->cstore != DCX_COEF_F32)
    DCX_FAIL("the one-way wall runs on FP32 storage");
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
  w.dt = dt;
  w.h[0] = dx;
  w.h[1] = dy;
  w.h[2] = dz;
  for (int axis = 0; axis < 3; ++axis) {
    const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
    const int64_t n1 = a1 < axis ? w.n[a1] : w.n[a1] - 2 * STENCIL;
    const int64_t n2 = a2 < axis ? w.n[a2] : w.n[a2] - 2 * STENCIL;
    dim3 grid((unsigned)((n1 + 255) / 256), (unsigned)n2, 2);
    hipLaunchKernelGGL(k_wall, grid, dim3(256), 0, st->stream, w, axis);
    DCX_CHECK(hipGetLastError());
  }
}

}  // namespace dcx

using namespace dcx;

extern "C" int dcx_set_wall(dc_device_data *data, int mode) {
  State *st = state_of(data);
  if (mode != DCX_WALL_ZERO && mode != DCX_WALL_ONE_WAY)
    DCX_FAIL("dcx_set_wall: unknown mode %d", mode);
  if (mode == DCX_WALL_ONE_WAY) {
    if (st->batch) DCX_FAIL("a batch shot cannot carry a one-way wall");
    if (st->halo_nuser > 0)
      DCX_FAIL("dcx_set_wall: the frame holds a registered halo exchange's "
               "receive boxes; a one-way wall is for a whole-domain run");
    if (st->fstore != DCX_STORE_F32 || st->cstore != DCX_COEF_F32)
      DCX_FAIL("the one-way wall runs on FP32 storage");
    for (int a = 0; a < 3; ++a)
      if (st->sizes[a] < 2 * STENCIL + 1)
        DCX_FAIL("dcx_set_wall: no compute cells along axis %d", a);
  }
  st->wall = mode;
  return 0;
}
