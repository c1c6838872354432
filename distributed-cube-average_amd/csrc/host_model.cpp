// Host-side model setup of the reference, in the library: default model,
// TTI coefficient precompute, random-velocity boundary, Ricker amplitude, and
// the output-file writer.  These keep the reference's float/double operand
// types so that the coefficient arrays and source amplitudes are the same
// bits the reference's host code produces.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "boundary_draw.hpp"
#include "dc_hip.h"

extern "C" {

// src/precomp.c:156-169 (SIGMA = 0.75, MAX_SIGMA = 10 from include/precomp.h).
void dcx_default_model(size_t n, float *vpz, float *vsv, float *epsilon,
                       float *delta, float *phi, float *theta) {
  const double sigma = 0.75, max_sigma = 10.0;
  for (size_t i = 0; i < n; i++) {
    vpz[i] = 3000.0;
    epsilon[i] = 0.24;
    delta[i] = 0.1;
    phi[i] = 1.0;
    theta[i] = atanf(1.0);
    vsv[i] = sigma > max_sigma
                 ? 0.0f
                 : (float)(vpz[i] * sqrtf(fabsf(epsilon[i] - delta[i]) / sigma));
  }
}

// src/precomp.c:85-105.  Each cell is independent, so the loop runs on all
// host threads; the per-cell arithmetic is unchanged.
void dcx_precompute(size_t n, const float *vpz, const float *vsv,
                    const float *epsilon, const float *delta, const float *phi,
                    const float *theta, float *const out[10]) {
  float *cxx = out[0], *cyy = out[1], *czz = out[2], *cxy = out[3],
        *cyz = out[4], *cxz = out[5], *v2px = out[6], *v2pz = out[7],
        *v2sz = out[8], *v2pn = out[9];
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < (long long)n; k++) {
    const size_t i = (size_t)k;
    const float st = (float)sin((double)theta[i]);
    const float ct = (float)cos((double)theta[i]);
    const float s2t = (float)sin(2.0 * theta[i]);
    const float sp = (float)sin((double)phi[i]);
    const float cp = (float)cos((double)phi[i]);
    const float s2p = (float)sin(2.0 * phi[i]);
    cxx[i] = st * st * cp * cp;
    cyy[i] = st * st * sp * sp;
    czz[i] = ct * ct;
    cxy[i] = st * st * s2p;
    cyz[i] = s2t * sp;
    cxz[i] = s2t * cp;
    v2sz[i] = vsv[i] * vsv[i];
    v2pz[i] = vpz[i] * vpz[i];
    v2px[i] = (float)(v2pz[i] * (1.0 + 2.0 * epsilon[i]));
    v2pn[i] = (float)(v2pz[i] * (1.0 + 2.0 * delta[i]));
  }
}

// src/boundary.c:7-104, cell by cell: each local cell finds its own draw
// from its index among the global grid's shell cells (boundary_draw.hpp), so
// the work is O(local cells) and parallel instead of a walk over the global
// grid; *seed ends where the reference's walk leaves it.
void dcx_random_boundary(int local_sx, int local_sy, int local_sz,
                         int global_sx, int global_sy, int global_sz,
                         int start_x, int start_y, int start_z, int nx, int ny,
                         int nz, int bord, int absorb, float *vpz, float *vsv,
                         unsigned int *seed) {
  (void)global_sx;
  (void)global_sy;
  (void)global_sz;
  const dcx::ShellGeom s{{nx, ny, nz}, bord, absorb};
  const uint32_t seed0 = *seed;
  const int64_t plane = (int64_t)local_sx * local_sy;
  // The shell cells of one row have consecutive draw indices: one jump per
  // row, then the generator runs on from draw to draw.
#pragma omp parallel for collapse(2) schedule(static)
  for (int lz = 0; lz < local_sz; lz++)
    for (int ly = 0; ly < local_sy; ly++) {
      uint32_t st = 0;
      bool have = false;
      for (int lx = 0; lx < local_sx; lx++) {
        const int g[3] = {start_x + lx, start_y + ly, start_z + lz};
        const int kind = dcx::shell_kind(s, g);
        if (kind == 0) continue;
        const size_t i = (size_t)(lx + (int64_t)ly * local_sx + lz * plane);
        if (kind == 2) {
          vpz[i] = 0.0f;
          vsv[i] = 0.0f;
          continue;
        }
        if (!have) {
          st = dcx::lcg_jump(seed0, 3 * dcx::shell_index(s, g));
          have = true;
        }
        const float r = dcx::draw_fraction(dcx::rand_r_draw(st));
        dcx::shell_velocity(s, g, r, &vpz[i], &vsv[i]);
      }
    }
  *seed = dcx::lcg_jump(seed0, 3 * dcx::shell_count(s));
}

// src/calculate_source.c:5-17; fcut is the reference's FCUT (40).
float dcx_ricker(float dt, int it, double fcut) {
  const double picube = 31.00627668029982017537;
  const double twosqrtpi = 3.54490770181103205458;
  const double threesqrtpi = 5.31736155271654808184;
  const float tf = (float)(twosqrtpi / fcut);
  const float fc = (float)(fcut / threesqrtpi);
  const float fct = fc * (((float)it) * dt - tf);
  const float expo = (float)(picube * fct * fct);
  return (1.0f - 2.0f * expo) * expf(-expo);
}

// src/coordinator.c:116-219 writes pc then qc over the whole global grid.
int dcx_write_dc_file(const char *path, size_t gx, size_t gy, size_t gz,
                      const float *pc, const float *qc) {
  FILE *f = std::fopen(path, "wb");
  if (f == nullptr) {
    std::fprintf(stderr, "[dc_hip] error: failed to open output file: %s\n",
                 path);
    return -1;
  }
  const size_t n = gx * gy * gz;
  const bool ok = std::fwrite(pc, sizeof(float), n, f) == n &&
                  std::fwrite(qc, sizeof(float), n, f) == n;
  return (std::fclose(f) == 0 && ok) ? 0 : -1;
}

}  // extern "C"
