// The reference worker's model setup (src/worker.c:49-156) on the device.
//
// The reference builds each rank's model on the host: the default TTI model
// in every local cell (src/worker.c:75-91, src/precomp.c:156-169), the
// random-velocity shell by a walk over the whole global grid that draws one
// rand_r per shell cell (src/boundary.c:25-103, O(global^3) on every rank),
// and the ten coefficient arrays (src/precomp.c:85-105, src/worker.c:130-154);
// dc_device_data_init then uploads the ten arrays.  Here one thread per local
// cell finds its own velocities -- the default model's, or its shell draw in
// closed form (boundary_draw.hpp), or zero outside the shell -- and writes
// its ten coefficients straight into the struct's pitched arrays.
//
// The default model has one (theta, phi) pair for every cell, so its
// trigonometry (FP64 sin/cos rounded to float, then float products) is done
// once on the host by the library's restatement of the reference's
// precompute, and the kernel writes the six resulting angle terms.  The v^2
// terms are formed per cell with the reference's operand types: vsv*vsv and
// vpz*vpz in float, and v2pz*(1 + 2 eps) in double, rounded to float.
#include "boundary_draw.hpp"
#include "dc_internal.hpp"

namespace dcx {

struct ModelArgs {
  float *coef[NCOEF];  // pitched arrays, element x = 0 of row 0
  int64_t ys, zs;
  int64_t size[3];     // local sizes
  int start[3];        // global coordinates of local cell (0, 0, 0)
  ShellGeom geom;
  uint32_t seed;       // generator state before the walk's first draw
  float vp0, vs0;      // the default model's vpz and vsv
  float angle[6];      // ch1dxx .. ch1dxz of the default model
  double k_eps, k_del; // 1 + 2 epsilon, 1 + 2 delta
};

__global__ void __launch_bounds__(256) k_model(const ModelArgs m) {
#pragma clang fp contract(off)
  const int64_t n = m.size[0] * m.size[1] * m.size[2];
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = c / m.size[0];
    const int64_t x = c - row * m.size[0];
    const int64_t z = row / m.size[1];
    const int64_t y = row - z * m.size[1];
    const int g[3] = {m.start[0] + (int)x, m.start[1] + (int)y,
                      m.start[2] + (int)z};
    float vp = m.vp0, vs = m.vs0;
    const int kind = shell_kind(m.geom, g);
    if (kind == 2) {
      vp = 0.0f;
      vs = 0.0f;
    } else if (kind == 1) {
      uint32_t st = lcg_jump(m.seed, 3 * shell_index(m.geom, g));
      const float r = draw_fraction(rand_r_draw(st));
      shell_velocity(m.geom, g, r, &vp, &vs);
    }
    const int64_t i = x + y * m.ys + z * m.zs;
#pragma unroll
    for (int k = 0; k < 6; ++k) m.coef[k][i] = m.angle[k];
    const float v2sz = vs * vs;
    const float v2pz = vp * vp;
    m.coef[V2PX][i] = (float)((double)v2pz * m.k_eps);
    m.coef[V2PZ][i] = v2pz;
    m.coef[V2SZ][i] = v2sz;
    m.coef[V2PN][i] = (float)((double)v2pz * m.k_del);
  }
}

}  // namespace dcx

using namespace dcx;

extern "C" dc_device_data *dcx_device_data_init_model(
    dc_process_t *process, const int global_sizes[3],
    const int start_coords[3], const int problem_sizes[3], int absorption,
    unsigned int *seed) {
  if (process == nullptr || global_sizes == nullptr ||
      start_coords == nullptr || problem_sizes == nullptr || seed == nullptr)
    DCX_FAIL("dcx_device_data_init_model: null argument");
  if (absorption < 0) DCX_FAIL("absorption must not be negative");
  for (int k = 0; k < 3; ++k) {
    if (problem_sizes[k] < 1) DCX_FAIL("empty problem region");
    if (start_coords[k] < 0 ||
        start_coords[k] + (int64_t)process->sizes[k] > global_sizes[k])
      DCX_FAIL("local box [%d, %lld) on axis %d is outside the global grid "
               "of %d cells",
               start_coords[k],
               (long long)(start_coords[k] + (int64_t)process->sizes[k]), k,
               global_sizes[k]);
  }
  // The default model of one cell and its coefficients, on the host (FP64
  // trigonometry as the reference does it).
  float vpz0, vsv0, eps0, del0, phi0, th0;
  dcx_default_model(1, &vpz0, &vsv0, &eps0, &del0, &phi0, &th0);
  float c1[NCOEF];
  float *c1p[NCOEF];
  for (int k = 0; k < NCOEF; ++k) c1p[k] = &c1[k];
  dcx_precompute(1, &vpz0, &vsv0, &eps0, &del0, &phi0, &th0, c1p);

  State *st = create_state(process, nullptr);
  ModelArgs m{};
  for (int k = 0; k < NCOEF; ++k) m.coef[k] = coef_array(st, k);
  m.ys = st->ys;
  m.zs = st->zs;
  for (int k = 0; k < 3; ++k) {
    m.size[k] = st->sizes[k];
    m.start[k] = start_coords[k];
  }
  m.geom = ShellGeom{{problem_sizes[0], problem_sizes[1], problem_sizes[2]},
                     STENCIL, absorption};
  m.seed = *seed;
  m.vp0 = vpz0;
  m.vs0 = vsv0;
  for (int k = 0; k < 6; ++k) m.angle[k] = c1[k];
  m.k_eps = 1.0 + 2.0 * (double)eps0;
  m.k_del = 1.0 + 2.0 * (double)del0;
  int64_t blocks = (st->n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_model, dim3((unsigned)blocks), dim3(256), 0, st->stream,
                     m);
  DCX_CHECK(hipGetLastError());
  // Where the reference's walk over the global grid leaves the generator.
  *seed = lcg_jump(*seed, 3 * shell_count(m.geom));
  return finish_state(st);
}
