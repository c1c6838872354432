// Host-side model setup of the reference, in the library: default model,
// TTI coefficient precompute, random-velocity boundary, Ricker amplitude, and
// the output-file writer.  These keep the reference's float/double operand
// types so that the coefficient arrays and source amplitudes are the same
// bits the reference's host code produces.
#include <cmath>
#include <cstdio>
#include <cstdlib>

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

// src/boundary.c:7-104.  The draw order (one rand_r per shell cell of the
// whole global grid, z then y then x) is what makes every rank agree.
void dcx_random_boundary(int local_sx, int local_sy, int local_sz,
                         int global_sx, int global_sy, int global_sz,
                         int start_x, int start_y, int start_z, int nx, int ny,
                         int nz, int bord, int absorb, float *vpz, float *vsv,
                         unsigned int *seed) {
  const int last = bord + absorb - 1;  // last shell index before the interior
  const int first = last + 1;          // first interior index
  const float frac = 1.0f / (float)absorb;
  const float maxP = 3000.0f;
  const float maxS = maxP * sqrtf(fabsf(0.24f - 0.1f) / 0.75f);
  const int ex = start_x + local_sx, ey = start_y + local_sy,
            ez = start_z + local_sz;
  auto dist = [](int g, int first_, int last_, int n_) {
    if (g > last_ + n_) return g - last_ - n_;
    if (g < first_) return first_ - g;
    return 0;
  };
  for (int gz = 0; gz < global_sz; gz++) {
    for (int gy = 0; gy < global_sy; gy++) {
      for (int gx = 0; gx < global_sx; gx++) {
        const bool interior = gz >= first && gz <= last + nz && gy >= first &&
                              gy <= last + ny && gx >= first && gx <= last + nx;
        if (interior) continue;
        const bool shell = gz >= bord && gz <= 2 * last + nz && gy >= bord &&
                           gy <= 2 * last + ny && gx >= bord &&
                           gx <= 2 * last + nx;
        const bool mine = gx >= start_x && gx < ex && gy >= start_y &&
                          gy < ey && gz >= start_z && gz < ez;
        if (shell) {
          const float r = (float)rand_r(seed) / (float)RAND_MAX;
          if (!mine) continue;
          const size_t i = (size_t)(gx - start_x) +
                           (size_t)(gy - start_y) * (size_t)local_sx +
                           (size_t)(gz - start_z) * (size_t)local_sx * local_sy;
          int d = dist(gy, first, last, ny);
          const int dz = dist(gz, first, last, nz);
          d = d > dz ? d : dz;
          const int dx = dist(gx, first, last, nx);
          d = d > dx ? d : dx;
          const float bd = (float)d * frac;
          vpz[i] = maxP * (1.0f - bd) + maxP * r * bd;
          vsv[i] = maxS * (1.0f - bd) + maxS * r * bd;
        } else if (mine) {
          const size_t i = (size_t)(gx - start_x) +
                           (size_t)(gy - start_y) * (size_t)local_sx +
                           (size_t)(gz - start_z) * (size_t)local_sx * local_sy;
          vpz[i] = 0.0f;
          vsv[i] = 0.0f;
        }
      }
    }
  }
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
