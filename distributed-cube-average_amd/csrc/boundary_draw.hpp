// The reference's random-velocity absorbing shell (src/boundary.c:7-104)
// evaluated cell by cell.
//
// The reference walks the whole global grid in z, y, x order and draws one
// glibc rand_r per shell cell, so every rank sees the same draws.  Here a
// cell's draw is found directly:
//   * its index k among the shell cells in that order has a closed form
//     (shell = box [bord, 2*last + n] per axis minus the interior
//     [last + 1, last + n]);
//   * rand_r (glibc stdlib/rand_r.c, the reference's third-party generator)
//     is three steps of the 32-bit LCG x -> 1103515245 x + 12345 (mod 2^32),
//     keeping 11, 10 and 10 bits of the three states; the state after k
//     draws is the affine map composed 3k times, found by squaring.
// Usable from host and device code; integer and float operations only.
#pragma once

#include <cmath>
#include <cstdint>

// Compiled by hipcc for the device kernel and by the host compiler for the
// host version.
#if defined(__HIPCC__)
#define DCX_HD __host__ __device__
#else
#define DCX_HD
#endif

namespace dcx {

struct AffineU32 {
  uint32_t a, c;  // x -> a x + c (mod 2^32)
};

DCX_HD inline AffineU32 affine_after(AffineU32 f, AffineU32 g) {
  return AffineU32{f.a * g.a, f.a * g.c + f.c};
}

// The generator state after `steps` LCG steps from `seed`.
DCX_HD inline uint32_t lcg_jump(uint32_t seed, uint64_t steps) {
  AffineU32 r{1u, 0u}, p{1103515245u, 12345u};
  while (steps) {
    if (steps & 1) r = affine_after(p, r);
    p = affine_after(p, p);
    steps >>= 1;
  }
  return r.a * seed + r.c;
}

// One rand_r draw from state `next` (advanced in place).
DCX_HD inline int rand_r_draw(uint32_t &next) {
  next = next * 1103515245u + 12345u;
  uint32_t result = (next / 65536u) % 2048u;
  next = next * 1103515245u + 12345u;
  result <<= 10;
  result ^= (next / 65536u) % 1024u;
  next = next * 1103515245u + 12345u;
  result <<= 10;
  result ^= (next / 65536u) % 1024u;
  return (int)result;
}

struct ShellGeom {
  int n[3];      // problem extents (nx, ny, nz)
  int bord, absorb;
};

DCX_HD inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// 0: interior (untouched), 1: shell (a draw), 2: outside the shell box.
DCX_HD inline int shell_kind(const ShellGeom &s, const int g[3]) {
  const int last = s.bord + s.absorb - 1, first = last + 1;
  bool interior = true, box = true;
  for (int a = 0; a < 3; ++a) {
    interior = interior && g[a] >= first && g[a] <= last + s.n[a];
    box = box && g[a] >= s.bord && g[a] <= 2 * last + s.n[a];
  }
  return interior ? 0 : (box ? 1 : 2);
}

// Index of shell cell g among all shell cells in z, y, x order.
DCX_HD inline uint64_t shell_index(const ShellGeom &s, const int g[3]) {
  const int last = s.bord + s.absorb - 1, first = last + 1;
  int64_t B[3], in_lo[3];
  bool inside[3];
  for (int a = 0; a < 3; ++a) {
    B[a] = 2 * last + s.n[a] - s.bord + 1;
    // interior layers of axis a strictly before g[a]
    in_lo[a] = clampi(g[a], first, last + s.n[a] + 1) - first;
    inside[a] = g[a] >= first && g[a] <= last + s.n[a];
  }
  const int64_t nxy = (int64_t)s.n[0] * s.n[1];
  int64_t k = (int64_t)(g[2] - s.bord) * B[0] * B[1] - in_lo[2] * nxy;
  if (inside[2])
    k += (int64_t)(g[1] - s.bord) * B[0] - in_lo[1] * s.n[0];
  else
    k += (int64_t)(g[1] - s.bord) * B[0];
  if (inside[2] && inside[1])
    k += (int64_t)(g[0] - s.bord) - in_lo[0];
  else
    k += (int64_t)(g[0] - s.bord);
  return (uint64_t)k;
}

// Total number of shell cells (draws of one whole walk).
DCX_HD inline uint64_t shell_count(const ShellGeom &s) {
  const int last = s.bord + s.absorb - 1;
  uint64_t box = 1, inner = 1;
  for (int a = 0; a < 3; ++a) {
    box *= (uint64_t)(2 * last + s.n[a] - s.bord + 1);
    inner *= (uint64_t)s.n[a];
  }
  return box - inner;
}

// The reference's velocities of shell cell g given its draw r, in its
// operation order (src/boundary.c:44-95).
// rand_r(...) / RAND_MAX as the reference forms it (glibc RAND_MAX).
DCX_HD inline float draw_fraction(int r) {
  return (float)r / (float)2147483647;
}

DCX_HD inline void shell_velocity(const ShellGeom &s, const int g[3], float r,
                                  float *vp, float *vs) {
#if defined(__clang__)
  // The reference's separate roundings (no fused multiply-add) on the device.
#pragma clang fp contract(off)
#endif
  const int last = s.bord + s.absorb - 1, first = last + 1;
  const float frac = 1.0f / (float)s.absorb;
  const float maxP = 3000.0f;
  const float maxS = maxP * sqrtf(fabsf(0.24f - 0.1f) / 0.75f);
  int d = 0;
  // y, then z, then x: the reference's max order.
  const int order[3] = {1, 2, 0};
  for (int t = 0; t < 3; ++t) {
    const int a = order[t];
    int da = 0;
    if (g[a] > last + s.n[a])
      da = g[a] - last - s.n[a];
    else if (g[a] < first)
      da = first - g[a];
    d = da > d ? da : d;
  }
  const float bd = (float)d * frac;
  *vp = maxP * (1.0f - bd) + maxP * r * bd;
  *vs = maxS * (1.0f - bd) + maxS * r * bd;
}

}  // namespace dcx
