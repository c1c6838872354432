/*
 * fletcher_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's one-rank Fletcher TTI time step, used as
 * the checker for the HIP library.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this code; the product never does.
 *
 * Every function names the reference file:line it restates.  The FP32 entry
 * points keep the reference's operation order and operand types so that, built
 * with gcc -O3 -fopenmp -ffp-contract=off on x86-64 (no FMA), their results are
 * bit-identical to the reference's OpenMP backend.  That claim is checked
 * against the reference's own compiled code (oracle/_ref) and against committed
 * golden fixtures produced from it (tests/golden).
 *
 * The FP64 entry points evaluate the same update in double precision on
 * double-precision fields; they are the accuracy oracle for error norms.
 */
#include <math.h>
#include <omp.h>
#include <stddef.h>
#include <xmmintrin.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_STENCIL 4

/* Finite-difference weights.  include/derivatives.h:15-38 gives them as float
 * constants; the values are the standard 8th-order central weights:
 *   first derivative  L = 4/5, -1/5, 4/105, -1/280
 *   second derivative K = -205/72, 8/5, -1/5, 8/315, -1/560
 *   cross derivative  L_ab = L_a * L_b (written out as rounded constants).
 * The float spellings below are the decimal expansions, rounded to binary32
 * by the compiler exactly as the reference's literals are. */
static const float oL1 = 0.8f, oL2 = -0.2f, oL3 = 0.0380952380952381f,
                   oL4 = -0.0035714285714285713f;
static const float oL11 = 0.64f, oL12 = -0.16f, oL13 = 0.03047619047619047618f,
                   oL14 = -0.00285714285714285713f, oL22 = 0.04f,
                   oL23 = -0.00761904761904761904f,
                   oL24 = 0.00071428571428571428f,
                   oL33 = 0.00145124716553287981f,
                   oL34 = -0.00013605442176870748f,
                   oL44 = 0.00001275510204081632f;
static const float oK0 = -2.84722222222222222222f, oK1 = 1.6f, oK2 = -0.2f,
                   oK3 = 0.02539682539682539682f,
                   oK4 = -0.00178571428571428571f;

/* Set (on != 0) or clear flush-to-zero and denormals-are-zero in the MXCSR
 * of every OpenMP thread, so a CPU timing can be taken in either FP mode.
 * The reference runs with the default environment (denormals live).
 * Returns the number of threads touched. */
int orc_set_ftz(int on) {
  int n = 0;
#pragma omp parallel reduction(+ : n)
  {
    unsigned int csr = _mm_getcsr();
    csr = on ? (csr | 0x8040u) : (csr & ~0x8040u);
    _mm_setcsr(csr);
    n += 1;
  }
  return n;
}

/* Exposed so tests can pin the constants against their closed forms. */
void orc_weights(float out[19]) {
  const float w[19] = {oL1,  oL2,  oL3,  oL4,  oL11, oL12, oL13,
                       oL14, oL22, oL23, oL24, oL33, oL34, oL44,
                       oK0,  oK1,  oK2,  oK3,  oK4};
  memcpy(out, w, sizeof w);
}

/* ---------------------------------------------------------------- FP32 ---- */

/* 9-point second derivative along stride s; include/derivatives.h:47-54. */
static inline float o_d2(const float *f, ptrdiff_t i, ptrdiff_t s, float inv) {
  return (oK0 * f[i] + oK1 * (f[i + s] + f[i - s]) +
          oK2 * (f[i + 2 * s] + f[i - 2 * s]) +
          oK3 * (f[i + 3 * s] + f[i - 3 * s]) +
          oK4 * (f[i + 4 * s] + f[i - 4 * s])) *
         inv;
}

/* One quadrant-signed 4-point group (+a,+b) - (+a,-b) - (-a,+b) + (-a,-b),
 * a along stride u, b along stride v (used by the FP64 form). */
#define QUAD(f, i, a, u, b, v)                                                \
  (f[(i) + (b) * (v) + (a) * (u)] - f[(i) + (b) * (v) - (a) * (u)] -         \
   f[(i) - (b) * (v) + (a) * (u)] + f[(i) - (b) * (v) - (a) * (u)])

/* The 8 points of weight L_pq (p != q) as one left-to-right chain: the
 * (p along u, q along v) quad, then the (q along u, p along v) quad, with no
 * parenthesis between them -- the evaluation order of the reference's
 * expression, include/derivatives.h:60-87. */
#define PAIR(f, i, p, q, u, v)                                                \
  (f[(i) + (q) * (v) + (p) * (u)] - f[(i) + (q) * (v) - (p) * (u)] -         \
   f[(i) - (q) * (v) + (p) * (u)] + f[(i) - (q) * (v) - (p) * (u)] +         \
   f[(i) + (p) * (v) + (q) * (u)] - f[(i) + (p) * (v) - (q) * (u)] -         \
   f[(i) - (p) * (v) + (q) * (u)] + f[(i) - (p) * (v) - (q) * (u)])

/* 64-point mixed derivative; include/derivatives.h:56-91.  Groups are added
 * in the order 11,12,13,14,22,23,24,33,34,44 and scaled by inv at the end.
 * s1 is the reference's first stride argument (s11), s2 the second (s21). */
static inline float o_dx(const float *f, ptrdiff_t i, ptrdiff_t s1,
                         ptrdiff_t s2, float inv) {
  return (oL11 * QUAD(f, i, 1, s1, 1, s2) + oL12 * PAIR(f, i, 2, 1, s1, s2) +
          oL13 * PAIR(f, i, 3, 1, s1, s2) + oL14 * PAIR(f, i, 4, 1, s1, s2) +
          oL22 * QUAD(f, i, 2, s1, 2, s2) + oL23 * PAIR(f, i, 3, 2, s1, s2) +
          oL24 * PAIR(f, i, 4, 2, s1, s2) + oL33 * QUAD(f, i, 3, s1, 3, s2) +
          oL34 * PAIR(f, i, 4, 3, s1, s2) + oL44 * QUAD(f, i, 4, s1, 4, s2)) *
         inv;
}

/* Coefficient array order everywhere in this file (and in the product's
 * C ABI): ch1dxx ch1dyy ch1dzz ch1dxy ch1dyz ch1dxz v2px v2pz v2sz v2pn, the
 * field order of dc_precomp_vars (include/precomp.h:6-17). */
enum { CXX, CYY, CZZ, CXY, CYZ, CXZ, V2PX, V2PZ, V2SZ, V2PN, NCOEF };

/* Per-cell update; include/sample_compute.h:105-185 (same operand order).
 * Strides: x 1, y sx, z sx*sy (include/indexing.h:15-20). */
static inline void o_cell(ptrdiff_t i, ptrdiff_t sy, ptrdiff_t sz, float *pp,
                          const float *pc, float *qp, const float *qc,
                          const float *const *c, float dxx, float dyy,
                          float dzz, float dxy, float dxz, float dyz,
                          float dt) {
  const float pxx = o_d2(pc, i, 1, dxx);
  const float pyy = o_d2(pc, i, sy, dyy);
  const float pzz = o_d2(pc, i, sz, dzz);
  const float pxy = o_dx(pc, i, 1, sy, dxy);
  const float pyz = o_dx(pc, i, sy, sz, dyz);
  const float pxz = o_dx(pc, i, 1, sz, dxz);
  const float h1p = c[CXX][i] * pxx + c[CYY][i] * pyy + c[CZZ][i] * pzz +
                    c[CXY][i] * pxy + c[CXZ][i] * pxz + c[CYZ][i] * pyz;
  const float h2p = pxx + pyy + pzz - h1p;

  const float qxx = o_d2(qc, i, 1, dxx);
  const float qyy = o_d2(qc, i, sy, dyy);
  const float qzz = o_d2(qc, i, sz, dzz);
  const float qxy = o_dx(qc, i, 1, sy, dxy);
  const float qyz = o_dx(qc, i, sy, sz, dyz);
  const float qxz = o_dx(qc, i, 1, sz, dxz);
  const float h1q = c[CXX][i] * qxx + c[CYY][i] * qyy + c[CZZ][i] * qzz +
                    c[CXY][i] * qxy + c[CXZ][i] * qxz + c[CYZ][i] * qyz;
  const float h2q = qxx + qyy + qzz - h1q;

  const float h1pmq = h1p - h1q;
  const float h2pmq = h2p - h2q;
  const float rhsp = c[V2PX][i] * h2p + c[V2PZ][i] * h1q + c[V2SZ][i] * h1pmq;
  const float rhsq = c[V2PN][i] * h2p + c[V2PZ][i] * h1q - c[V2SZ][i] * h2pmq;
  pp[i] = 2.0f * pc[i] - pp[i] + rhsp * dt * dt;
  qp[i] = 2.0f * qc[i] - qp[i] + rhsq * dt * dt;
}

/* Sweep of the half-open box [start,end); src/openmp_propagate.c:6-32
 * (omp for over z, then y, then x).  The inverse spacings are formed in float
 * exactly as include/sample_compute.h:128-133 forms them. */
void orc_propagate_f32(const size_t start[3], const size_t end[3],
                       const size_t sizes[3], float *pp, const float *pc,
                       float *qp, const float *qc, const float *const coef[10],
                       float dx, float dy, float dz, float dt) {
  const ptrdiff_t sy = (ptrdiff_t)sizes[0];
  const ptrdiff_t sz = (ptrdiff_t)(sizes[0] * sizes[1]);
  const float dxx = 1.0f / (dx * dx), dyy = 1.0f / (dy * dy),
              dzz = 1.0f / (dz * dz), dxy = 1.0f / (dx * dy),
              dxz = 1.0f / (dx * dz), dyz = 1.0f / (dy * dz);
  if (end[0] <= start[0] || end[1] <= start[1] || end[2] <= start[2])
    return;
#pragma omp parallel for schedule(static)
  for (size_t z = start[2]; z < end[2]; z++)
    for (size_t y = start[1]; y < end[1]; y++)
      for (size_t x = start[0]; x < end[0]; x++)
        o_cell((ptrdiff_t)(x + y * sizes[0] + z * sizes[0] * sizes[1]), sy, sz,
               pp, pc, qp, qc, coef, dxx, dyy, dzz, dxy, dxz, dyz, dt);
}

/* Ricker amplitude for step it; src/calculate_source.c:5-17 with the cut-off
 * frequency as a parameter (the reference fixes FCUT = 40).  Operand types
 * follow the reference: the constants are double, tf/fc/fct/expo are float. */
float orc_source_f32(float dt, int it, double fcut) {
  const double picube = 31.00627668029982017537;
  const double twosqrtpi = 3.54490770181103205458;
  const double threesqrtpi = 5.31736155271654808184;
  float tf, fc, fct, expo;
  tf = twosqrtpi / fcut;
  fc = fcut / threesqrtpi;
  fct = fc * (((float)it) * dt - tf);
  expo = picube * fct * fct;
  return ((1.0f - 2.0f * expo) * expf(-expo));
}

/* Default homogeneous model; src/precomp.c:156-169.  Writes n cells of
 * vpz, vsv, epsilon, delta, phi, theta. */
void orc_default_model(size_t n, float *vpz, float *vsv, float *eps,
                       float *del, float *phi, float *theta) {
  const double sigma = 0.75, max_sigma = 10.0;
  for (size_t i = 0; i < n; i++) {
    vpz[i] = 3000.0;
    eps[i] = 0.24;
    del[i] = 0.1;
    phi[i] = 1.0;
    theta[i] = atanf(1.0);
    if (sigma > max_sigma)
      vsv[i] = 0.0;
    else
      vsv[i] = vpz[i] * sqrtf(fabsf(eps[i] - del[i]) / sigma);
  }
}

/* Precomputed TTI coefficients; src/precomp.c:85-105.  Trig in double,
 * stored to float before the products; the (1+2e) factor in double. */
void orc_precomp(size_t n, const float *vpz, const float *vsv,
                 const float *eps, const float *del, const float *phi,
                 const float *theta, float *const out[10]) {
  for (size_t i = 0; i < n; i++) {
    float st = sin(theta[i]);
    float ct = cos(theta[i]);
    float s2t = sin(2.0 * theta[i]);
    float sp = sin(phi[i]);
    float cp = cos(phi[i]);
    float s2p = sin(2.0 * phi[i]);
    out[CXX][i] = st * st * cp * cp;
    out[CYY][i] = st * st * sp * sp;
    out[CZZ][i] = ct * ct;
    out[CXY][i] = st * st * s2p;
    out[CYZ][i] = s2t * sp;
    out[CXZ][i] = s2t * cp;
  }
  for (size_t i = 0; i < n; i++) {
    out[V2SZ][i] = vsv[i] * vsv[i];
    out[V2PZ][i] = vpz[i] * vpz[i];
    out[V2PX][i] = out[V2PZ][i] * (1.0 + 2.0 * eps[i]);
    out[V2PN][i] = out[V2PZ][i] * (1.0 + 2.0 * del[i]);
  }
}

/* Random-velocity boundary over one local box; src/boundary.c:7-104.
 * Walks the whole global grid in z, y, x order, drawing one glibc rand_r
 * value per cell of the absorbing shell, so every rank sees the same draw for
 * the same global cell.  Cells outside the shell's outer edge (the 4-cell
 * frame) get zero velocity. */
void orc_random_boundary(int lsx, int lsy, int lsz, int gsx, int gsy, int gsz,
                         int x0, int y0, int z0, int nx, int ny, int nz,
                         int bord, int absorb, float *vpz, float *vsv,
                         unsigned int *seed) {
  const int blen = bord + absorb - 1;
  const int first = blen + 1;
  const float frac = 1.0f / (float)(absorb);
  const float maxP = 3000.0f;
  const float maxS = maxP * sqrtf(fabsf(0.24f - 0.1f) / 0.75f);
  const int x1 = x0 + lsx, y1 = y0 + lsy, z1 = z0 + lsz;
  for (int gz = 0; gz < gsz; gz++)
    for (int gy = 0; gy < gsy; gy++)
      for (int gx = 0; gx < gsx; gx++) {
        const int inside = gz >= first && gz <= blen + nz && gy >= first &&
                           gy <= blen + ny && gx >= first && gx <= blen + nx;
        const int shell = gz >= bord && gz <= 2 * blen + nz && gy >= bord &&
                          gy <= 2 * blen + ny && gx >= bord &&
                          gx <= 2 * blen + nx;
        const int mine =
            gx >= x0 && gx < x1 && gy >= y0 && gy < y1 && gz >= z0 && gz < z1;
        if (inside)
          continue;
        if (shell) {
          const float r = (float)rand_r(seed) / (float)RAND_MAX;
          if (!mine)
            continue;
          const size_t i = (size_t)(gx - x0) + (size_t)(gy - y0) * lsx +
                           (size_t)(gz - z0) * lsx * lsy;
          int dz = gz > blen + nz ? gz - blen - nz : (gz < first ? first - gz : 0);
          int dy = gy > blen + ny ? gy - blen - ny : (gy < first ? first - gy : 0);
          int dxx = gx > blen + nx ? gx - blen - nx : (gx < first ? first - gx : 0);
          int d = dy > dz ? dy : dz;
          d = d > dxx ? d : dxx;
          const float bd = (float)(d)*frac;
          vpz[i] = maxP * (1.0f - bd) + maxP * r * bd;
          vsv[i] = maxS * (1.0f - bd) + maxS * r * bd;
        } else if (mine) {
          const size_t i = (size_t)(gx - x0) + (size_t)(gy - y0) * lsx +
                           (size_t)(gz - z0) * lsx * lsy;
          vpz[i] = 0.0f;
          vsv[i] = 0.0f;
        }
      }
}

/* Box <-> contiguous buffer, x fastest; src/device_data.c:61-92. */
void orc_extract_box(float *buf, const size_t start[3], const size_t end[3],
                     const size_t sizes[3], const float *from) {
  size_t k = 0;
  for (size_t z = start[2]; z < end[2]; z++)
    for (size_t y = start[1]; y < end[1]; y++)
      for (size_t x = start[0]; x < end[0]; x++)
        buf[k++] = from[x + y * sizes[0] + z * sizes[0] * sizes[1]];
}

void orc_insert_box(const float *buf, const size_t start[3],
                    const size_t end[3], const size_t sizes[3], float *to) {
  size_t k = 0;
  for (size_t z = start[2]; z < end[2]; z++)
    for (size_t y = start[1]; y < end[1]; y++)
      for (size_t x = start[0]; x < end[0]; x++)
        to[x + y * sizes[0] + z * sizes[0] * sizes[1]] = buf[k++];
}

/* One-rank time loop; src/worker.c:451-511 with no neighbours: add the source
 * to pc and qc (src/device_data.c:56-59), update the compute box [4,size-4)^3
 * (the seven region calls of src/worker.c:321-402 cover it exactly once and
 * each cell reads only pc/qc, so one sweep gives the same bits), then swap
 * (src/device_data.c:44-54).  f[] = {pp, pc, qp, qc}; on return f[] holds the
 * swapped pointers, so f[1]/f[3] are the latest fields.  src_amp, when not
 * NULL, supplies the per-step source amplitudes (n_steps entries). */
void orc_run_f32(const size_t sizes[3], float *f[4],
                 const float *const coef[10], float dx, float dy, float dz,
                 float dt, unsigned int n_steps, long long src_index,
                 const float *src_amp) {
  const size_t s0[3] = {ORC_STENCIL, ORC_STENCIL, ORC_STENCIL};
  const size_t e0[3] = {sizes[0] - ORC_STENCIL, sizes[1] - ORC_STENCIL,
                        sizes[2] - ORC_STENCIL};
  for (unsigned int it = 0; it < n_steps; it++) {
    if (src_index >= 0) {
      const float s = src_amp ? src_amp[it] : orc_source_f32(dt, (int)it, 40.0);
      f[1][src_index] += s;
      f[3][src_index] += s;
    }
    orc_propagate_f32(s0, e0, sizes, f[0], f[1], f[2], f[3], coef, dx, dy, dz,
                      dt);
    float *t = f[0];
    f[0] = f[1];
    f[1] = t;
    t = f[2];
    f[2] = f[3];
    f[3] = t;
  }
}

/* orc_run_f32 that also records pc at n_rec receiver cells (linear
 * indices rec[]) after every step into traces[step * n_rec + r]. */
void orc_run_f32_traces(const size_t sizes[3], float *f[4],
                        const float *const coef[10], float dx, float dy,
                        float dz, float dt, unsigned int n_steps,
                        long long src_index, const float *src_amp, int n_rec,
                        const long long *rec, float *traces) {
  for (unsigned int it = 0; it < n_steps; it++) {
    orc_run_f32(sizes, f, coef, dx, dy, dz, dt, 1, src_index,
                src_amp ? src_amp + it : NULL);
    for (int r = 0; r < n_rec; r++) traces[(size_t)it * n_rec + r] = f[1][rec[r]];
  }
}

/* ---------------------------------------------------------------- FP64 ---- */

static inline double o_d2d(const double *f, ptrdiff_t i, ptrdiff_t s,
                           double inv) {
  static const double K[5] = {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0,
                              8.0 / 315.0, -1.0 / 560.0};
  double acc = K[0] * f[i];
  for (int a = 1; a <= 4; a++)
    acc += K[a] * (f[i + a * s] + f[i - a * s]);
  return acc * inv;
}

static inline double o_dxd(const double *f, ptrdiff_t i, ptrdiff_t s1,
                           ptrdiff_t s2, double inv) {
  static const double L[5] = {0.0, 4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0,
                              -1.0 / 280.0};
  double acc = 0.0;
  for (int a = 1; a <= 4; a++)
    for (int b = 1; b <= 4; b++)
      acc += L[a] * L[b] * QUAD(f, i, a, s1, b, s2);
  return acc * inv;
}

/* FP64 restatement of include/sample_compute.h:105-185 on double fields with
 * exact (double) FD weights; coefficients are the float arrays, promoted. */
void orc_propagate_f64(const size_t start[3], const size_t end[3],
                       const size_t sizes[3], double *pp, const double *pc,
                       double *qp, const double *qc,
                       const float *const c[10], double dx, double dy,
                       double dz, double dt) {
  const ptrdiff_t sy = (ptrdiff_t)sizes[0];
  const ptrdiff_t sz = (ptrdiff_t)(sizes[0] * sizes[1]);
  const double dxx = 1.0 / (dx * dx), dyy = 1.0 / (dy * dy),
               dzz = 1.0 / (dz * dz), dxy = 1.0 / (dx * dy),
               dxz = 1.0 / (dx * dz), dyz = 1.0 / (dy * dz);
  if (end[0] <= start[0] || end[1] <= start[1] || end[2] <= start[2])
    return;
#pragma omp parallel for schedule(static)
  for (size_t z = start[2]; z < end[2]; z++)
    for (size_t y = start[1]; y < end[1]; y++)
      for (size_t x = start[0]; x < end[0]; x++) {
        const ptrdiff_t i =
            (ptrdiff_t)(x + y * sizes[0] + z * sizes[0] * sizes[1]);
        const double pxx = o_d2d(pc, i, 1, dxx), pyy = o_d2d(pc, i, sy, dyy),
                     pzz = o_d2d(pc, i, sz, dzz);
        const double pxy = o_dxd(pc, i, 1, sy, dxy),
                     pyz = o_dxd(pc, i, sy, sz, dyz),
                     pxz = o_dxd(pc, i, 1, sz, dxz);
        const double h1p = c[CXX][i] * pxx + c[CYY][i] * pyy +
                           c[CZZ][i] * pzz + c[CXY][i] * pxy +
                           c[CXZ][i] * pxz + c[CYZ][i] * pyz;
        const double h2p = pxx + pyy + pzz - h1p;
        const double qxx = o_d2d(qc, i, 1, dxx), qyy = o_d2d(qc, i, sy, dyy),
                     qzz = o_d2d(qc, i, sz, dzz);
        const double qxy = o_dxd(qc, i, 1, sy, dxy),
                     qyz = o_dxd(qc, i, sy, sz, dyz),
                     qxz = o_dxd(qc, i, 1, sz, dxz);
        const double h1q = c[CXX][i] * qxx + c[CYY][i] * qyy +
                           c[CZZ][i] * qzz + c[CXY][i] * qxy +
                           c[CXZ][i] * qxz + c[CYZ][i] * qyz;
        const double h2q = qxx + qyy + qzz - h1q;
        const double rhsp = c[V2PX][i] * h2p + c[V2PZ][i] * h1q +
                            c[V2SZ][i] * (h1p - h1q);
        const double rhsq = c[V2PN][i] * h2p + c[V2PZ][i] * h1q -
                            c[V2SZ][i] * (h2p - h2q);
        pp[i] = 2.0 * pc[i] - pp[i] + rhsp * dt * dt;
        qp[i] = 2.0 * qc[i] - qp[i] + rhsq * dt * dt;
      }
}

/* FP64 one-rank time loop (same structure as orc_run_f32); the source
 * amplitudes are the float table, promoted. */
void orc_run_f64(const size_t sizes[3], double *f[4],
                 const float *const coef[10], double dx, double dy, double dz,
                 double dt, unsigned int n_steps, long long src_index,
                 const float *src_amp) {
  const size_t s0[3] = {ORC_STENCIL, ORC_STENCIL, ORC_STENCIL};
  const size_t e0[3] = {sizes[0] - ORC_STENCIL, sizes[1] - ORC_STENCIL,
                        sizes[2] - ORC_STENCIL};
  for (unsigned int it = 0; it < n_steps; it++) {
    if (src_index >= 0 && src_amp) {
      f[1][src_index] += (double)src_amp[it];
      f[3][src_index] += (double)src_amp[it];
    }
    orc_propagate_f64(s0, e0, sizes, f[0], f[1], f[2], f[3], coef, dx, dy, dz,
                      dt);
    double *t = f[0];
    f[0] = f[1];
    f[1] = t;
    t = f[2];
    f[2] = f[3];
    f[3] = t;
  }
}

/* orc_run_f64 recording pc at the receivers after every step. */
void orc_run_f64_traces(const size_t sizes[3], double *f[4],
                        const float *const coef[10], double dx, double dy,
                        double dz, double dt, unsigned int n_steps,
                        long long src_index, const float *src_amp, int n_rec,
                        const long long *rec, double *traces) {
  for (unsigned int it = 0; it < n_steps; it++) {
    orc_run_f64(sizes, f, coef, dx, dy, dz, dt, 1, src_index,
                src_amp ? src_amp + it : NULL);
    for (int r = 0; r < n_rec; r++) traces[(size_t)it * n_rec + r] = f[1][rec[r]];
  }
}
