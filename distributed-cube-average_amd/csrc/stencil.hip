// Fletcher TTI time-step kernels for gfx950.
//
// The per-cell update is the reference's include/sample_compute.h:105-185:
//   H1(f) = sum_ij ch1d_ij * D_ij f   (ij in xx,yy,zz,xy,xz,yz)
//   H2(f) = Dxx f + Dyy f + Dzz f - H1(f)
//   rhsp  = v2px*H2(p) + v2pz*H1(q) + v2sz*(H1(p)-H1(q))
//   rhsq  = v2pn*H2(p) + v2pz*H1(q) - v2sz*(H2(p)-H2(q))
//   pp    = 2*pc - pp + rhsp*dt*dt   (same for q)
// with the 8th-order operators of include/derivatives.h:15-91.
//
// Two kernels:
//  * k_exact: one thread per cell, the reference's expression order with FP
//    contraction off, 64-point mixed derivatives read from global memory.
//    Bit-identical to the reference's OpenMP path; used for parity.
//  * k_fast: the production kernel.  A workgroup owns a TX x TY column tile
//    and a chunk of z planes and streams through z.  Each new plane of p and q
//    (tile + 4-cell apron) is staged in LDS; x/y derivatives come from LDS,
//    z derivatives from 8-slot register rings.  Mixed derivatives use the
//    exact factorisation D_ab f = D1_b(D1_a f) (the 64-point weights are the
//    products L_a*L_b), so Dxz and Dyz are first differences over z of the
//    per-plane D1x and D1y rings and Dxy is D1y of an LDS plane of D1x.
//    Every input byte is read once from HBM per step apart from the apron.
#include "dc_internal.hpp"

namespace dcx {

// 8th-order weights, spelled as in include/derivatives.h:15-38.
#define W_L1 0.8f
#define W_L2 -0.2f
#define W_L3 0.0380952380952381f
#define W_L4 -0.0035714285714285713f
#define W_L11 0.64f
#define W_L12 -0.16f
#define W_L13 0.03047619047619047618f
#define W_L14 -0.00285714285714285713f
#define W_L22 0.04f
#define W_L23 -0.00761904761904761904f
#define W_L24 0.00071428571428571428f
#define W_L33 0.00145124716553287981f
#define W_L34 -0.00013605442176870748f
#define W_L44 0.00001275510204081632f
#define W_K0 -2.84722222222222222222f
#define W_K1 1.6f
#define W_K2 -0.2f
#define W_K3 0.02539682539682539682f
#define W_K4 -0.00178571428571428571f

// ------------------------------------------------------------ exact ----

__device__ __forceinline__ float ex_d2(const float *__restrict__ f, int64_t i,
                                       int64_t s, float inv) {
#pragma clang fp contract(off)
  return (W_K0 * f[i] + W_K1 * (f[i + s] + f[i - s]) +
          W_K2 * (f[i + 2 * s] + f[i - 2 * s]) +
          W_K3 * (f[i + 3 * s] + f[i - 3 * s]) +
          W_K4 * (f[i + 4 * s] + f[i - 4 * s])) *
         inv;
}

// (+a along u,+b along v) quad, left to right.
#define EX_QUAD(f, i, a, u, b, v)                                              \
  (f[(i) + (b) * (v) + (a) * (u)] - f[(i) + (b) * (v) - (a) * (u)] -          \
   f[(i) - (b) * (v) + (a) * (u)] + f[(i) - (b) * (v) - (a) * (u)])
// The 8 points of one off-diagonal weight as one left-to-right chain
// (include/derivatives.h:60-87).
#define EX_PAIR(f, i, p, q, u, v)                                              \
  (f[(i) + (q) * (v) + (p) * (u)] - f[(i) + (q) * (v) - (p) * (u)] -          \
   f[(i) - (q) * (v) + (p) * (u)] + f[(i) - (q) * (v) - (p) * (u)] +          \
   f[(i) + (p) * (v) + (q) * (u)] - f[(i) + (p) * (v) - (q) * (u)] -          \
   f[(i) - (p) * (v) + (q) * (u)] + f[(i) - (p) * (v) - (q) * (u)])

__device__ __forceinline__ float ex_dx(const float *__restrict__ f, int64_t i,
                                       int64_t u, int64_t v, float inv) {
#pragma clang fp contract(off)
  return (W_L11 * EX_QUAD(f, i, 1, u, 1, v) + W_L12 * EX_PAIR(f, i, 2, 1, u, v) +
          W_L13 * EX_PAIR(f, i, 3, 1, u, v) + W_L14 * EX_PAIR(f, i, 4, 1, u, v) +
          W_L22 * EX_QUAD(f, i, 2, u, 2, v) + W_L23 * EX_PAIR(f, i, 3, 2, u, v) +
          W_L24 * EX_PAIR(f, i, 4, 2, u, v) + W_L33 * EX_QUAD(f, i, 3, u, 3, v) +
          W_L34 * EX_PAIR(f, i, 4, 3, u, v) + W_L44 * EX_QUAD(f, i, 4, u, 4, v)) *
         inv;
}

__global__ void __launch_bounds__(256) k_exact(const StepArgs a) {
#pragma clang fp contract(off)
  const int64_t x = a.box.x0 + (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t y = a.box.y0 + (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  const int64_t z = a.box.z0 + (int64_t)blockIdx.z;
  if (x >= a.box.x1 || y >= a.box.y1) return;
  const int64_t i = x + y * a.ys + z * a.zs;
  const int64_t sy = a.ys, sz = a.zs;
  const float *__restrict__ pc = a.pc;
  const float *__restrict__ qc = a.qc;
  const float *const *c = a.coef.c;

  const float pxx = ex_d2(pc, i, 1, a.dxx);
  const float pyy = ex_d2(pc, i, sy, a.dyy);
  const float pzz = ex_d2(pc, i, sz, a.dzz);
  const float pxy = ex_dx(pc, i, 1, sy, a.dxy);
  const float pyz = ex_dx(pc, i, sy, sz, a.dyz);
  const float pxz = ex_dx(pc, i, 1, sz, a.dxz);
  const float h1p = c[CXX][i] * pxx + c[CYY][i] * pyy + c[CZZ][i] * pzz +
                    c[CXY][i] * pxy + c[CXZ][i] * pxz + c[CYZ][i] * pyz;
  const float h2p = pxx + pyy + pzz - h1p;

  const float qxx = ex_d2(qc, i, 1, a.dxx);
  const float qyy = ex_d2(qc, i, sy, a.dyy);
  const float qzz = ex_d2(qc, i, sz, a.dzz);
  const float qxy = ex_dx(qc, i, 1, sy, a.dxy);
  const float qyz = ex_dx(qc, i, sy, sz, a.dyz);
  const float qxz = ex_dx(qc, i, 1, sz, a.dxz);
  const float h1q = c[CXX][i] * qxx + c[CYY][i] * qyy + c[CZZ][i] * qzz +
                    c[CXY][i] * qxy + c[CXZ][i] * qxz + c[CYZ][i] * qyz;
  const float h2q = qxx + qyy + qzz - h1q;

  const float h1pmq = h1p - h1q;
  const float h2pmq = h2p - h2q;
  const float rhsp = c[V2PX][i] * h2p + c[V2PZ][i] * h1q + c[V2SZ][i] * h1pmq;
  const float rhsq = c[V2PN][i] * h2p + c[V2PZ][i] * h1q - c[V2SZ][i] * h2pmq;
  float pn = 2.0f * pc[i] - a.pp[i] + rhsp * a.dt * a.dt;
  float qn = 2.0f * qc[i] - a.qp[i] + rhsq * a.dt * a.dt;
  if (z == a.src_z && x + y * a.ys == a.src_xy) {
    pn += a.src_amp;
    qn += a.src_amp;
  }
  a.pp[i] = pn;
  a.qp[i] = qn;
}

void launch_exact(const StepArgs &a, hipStream_t s) {
  const int64_t nx = a.box.x1 - a.box.x0, ny = a.box.y1 - a.box.y0,
                nz = a.box.z1 - a.box.z0;
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  dim3 grid((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)nz);
  hipLaunchKernelGGL(k_exact, grid, dim3(256), 0, s, a);
  DCX_CHECK(hipGetLastError());
}

// ------------------------------------------------------------- fast ----

// Pair sums/differences along one axis from LDS at index l with stride st:
// s_a = f(+a) + f(-a), d_a = f(+a) - f(-a), a = 1..4.
struct Pairs {
  float s1, s2, s3, s4, d1, d2, d3, d4;
};

// The nine points are read from the lowest one up, so that every read is
// one lane address plus a non-negative constant, which ds_read encodes in
// its offset field (a negative offset costs an address add per read).
__device__ __forceinline__ Pairs lds_pairs(const float *L, int l, int st) {
  const float *B = L + (l - 4 * st);
  const float m4 = B[0], m3 = B[st], m2 = B[2 * st], m1 = B[3 * st];
  const float p1 = B[5 * st], p2 = B[6 * st], p3 = B[7 * st], p4 = B[8 * st];
  return Pairs{p1 + m1, p2 + m2, p3 + m3, p4 + m4,
               p1 - m1, p2 - m2, p3 - m3, p4 - m4};
}

// h scales the weights (exactly: 1 or 1/2).
__device__ __forceinline__ float d2_of(float c, const Pairs &q, float h) {
  return (h * W_K0) * c + (h * W_K1) * q.s1 + (h * W_K2) * q.s2 +
         (h * W_K3) * q.s3 + (h * W_K4) * q.s4;
}

__device__ __forceinline__ float d1_of(const Pairs &q) {
  return W_L1 * q.d1 + W_L2 * q.d2 + W_L3 * q.d3 + W_L4 * q.d4;
}

__device__ __forceinline__ float d1_lds(const float *L, int l, int st) {
  const float *B = L + (l - 4 * st);  // as lds_pairs
  return W_L1 * (B[5 * st] - B[3 * st]) + W_L2 * (B[6 * st] - B[2 * st]) +
         W_L3 * (B[7 * st] - B[st]) + W_L4 * (B[8 * st] - B[0]);
}

// Load/store at a wave-uniform base plus a per-lane unsigned 32-bit byte
// offset, the form gfx950's global_load/store take with an SGPR base address
// (one VGPR per offset instead of a 64-bit address pair per load).  The base
// is cast to the global address space: a batched launch reads its array
// pointers from device memory, where the compiler cannot tell global from
// LDS and would emit FLAT accesses, which count against the LDS wait
// counter, so every wait for the LDS plane would also wait for the
// prefetched loads.
#define DCX_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T ldb(const void *base, uint32_t boff) {
  return *reinterpret_cast<const DCX_GLOBAL T *>(
      (const DCX_GLOBAL char *)base + boff);
}
template <class T>
__device__ __forceinline__ void stb(void *base, uint32_t boff, T v) {
  *reinterpret_cast<DCX_GLOBAL T *>((DCX_GLOBAL char *)base + boff) = v;
}

// Field storage policies (DCX_STORE_*): element size, raw load, decode to
// float, store.  Prefetched values stay raw until they are used: a decode
// placed at the load would make the wave wait for the load there, before the
// barrier that ends the iteration, and undo the prefetch.
template <int FS> struct FieldIO;
template <> struct FieldIO<DCX_STORE_F32> {
  static constexpr uint32_t kBytes = 4;
  static __device__ __forceinline__ uint32_t ldr(const void *b, uint32_t o) {
    return ldb<uint32_t>(b, o);
  }
  static __device__ __forceinline__ float dec(uint32_t r) {
    return __uint_as_float(r);
  }
  static __device__ __forceinline__ float ld(const void *b, uint32_t o) {
    return ldb<float>(b, o);
  }
  static __device__ __forceinline__ void st(void *b, uint32_t o, float v) {
    stb<float>(b, o, v);
  }
  static __device__ __forceinline__ bool bad(float) { return false; }
};
template <> struct FieldIO<DCX_STORE_F16> {
  static constexpr uint32_t kBytes = 2;
  static __device__ __forceinline__ uint32_t ldr(const void *b, uint32_t o) {
    return ldb<uint16_t>(b, o);
  }
  static __device__ __forceinline__ float dec(uint32_t r) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)r);
  }
  static __device__ __forceinline__ float ld(const void *b, uint32_t o) {
    return dec(ldr(b, o));
  }
  // Saturating store: a value beyond half's range is written as +-65504
  // (NaN as -65504), never as inf or NaN; `bad` flags it for the
  // struct's overflow counter.
  static __device__ __forceinline__ void st(void *b, uint32_t o, float v) {
    const float c = fminf(fmaxf(v, -65504.0f), 65504.0f);
    stb<uint16_t>(b, o, __builtin_bit_cast(uint16_t, (_Float16)c));
  }
  static __device__ __forceinline__ bool bad(float v) {
    return !(fabsf(v) <= 65504.0f);
  }
};
template <> struct FieldIO<DCX_STORE_BF16> {
  static constexpr uint32_t kBytes = 2;
  static __device__ __forceinline__ uint32_t ldr(const void *b, uint32_t o) {
    return ldb<uint16_t>(b, o);
  }
  static __device__ __forceinline__ float dec(uint32_t r) {
    return __uint_as_float(r << 16);
  }
  static __device__ __forceinline__ float ld(const void *b, uint32_t o) {
    return dec(ldr(b, o));
  }
  static __device__ __forceinline__ void st(void *b, uint32_t o, float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    stb<uint16_t>(b, o, (uint16_t)(u >> 16));
  }
  // bfloat16 has FP32's range: only a non-finite value is flagged.
  static __device__ __forceinline__ bool bad(float v) {
    return !(fabsf(v) <= 3.38953139e38f);
  }
};

// Coefficient storage policies (DCX_COEF_*): element size, raw load of
// array k (a compile-time index after unrolling) and its decode, kept apart
// for the reason given at FieldIO.
template <int CS> struct CoefIO;
template <> struct CoefIO<DCX_COEF_F32> {
  static constexpr uint32_t kBytes = 4;
  static __device__ __forceinline__ uint32_t ldr(const StepArgs &a, int k,
                                                 uint32_t o) {
    return ldb<uint32_t>(a.cf[k], o);
  }
  static __device__ __forceinline__ float dec(const StepArgs &, int,
                                              uint32_t r) {
    return __uint_as_float(r);
  }
};
template <> struct CoefIO<DCX_COEF_F16> {
  static constexpr uint32_t kBytes = 2;
  static __device__ __forceinline__ uint32_t ldr(const StepArgs &a, int k,
                                                 uint32_t o) {
    return ldb<uint16_t>(a.cf[k], o);
  }
  static __device__ __forceinline__ float dec(const StepArgs &a, int k,
                                              uint32_t r) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)r) * a.c_st[k];
  }
};
template <> struct CoefIO<DCX_COEF_U16> {
  static constexpr uint32_t kBytes = 2;
  static __device__ __forceinline__ uint32_t ldr(const StepArgs &a, int k,
                                                 uint32_t o) {
    return ldb<uint16_t>(a.cf[k], o);
  }
  static __device__ __forceinline__ float dec(const StepArgs &a, int k,
                                              uint32_t u) {
    const float t = fmaf((float)u, a.c_st[k], a.c_lo[k]);
    if (k < V2PX) return t;              // angle terms: affine
    return u ? exp2f(t) : 0.0f;          // v^2 terms: affine in log2
  }
};

// Angle terms as 16-bit affine codes, v^2 terms in FP32.  Offsets are formed
// for 2-byte elements and doubled for the FP32 arrays (every array is below
// 4 GiB, checked at launch).
template <> struct CoefIO<DCX_COEF_A16> {
  static constexpr uint32_t kBytes = 2;
  static __device__ __forceinline__ uint32_t ldr(const StepArgs &a, int k,
                                                 uint32_t o) {
    if (k < V2PX) return ldb<uint16_t>(a.cf[k], o);
    return ldb<uint32_t>(a.cf[k], 2u * o);
  }
  static __device__ __forceinline__ float dec(const StepArgs &a, int k,
                                              uint32_t r) {
    if (k < V2PX) return fmaf((float)r, a.c_st[k], a.c_lo[k]);
    return __uint_as_float(r);
  }
};

// TTI axis form: per cell a dword AB = (a, b) codes, a qword CR = (c, rx, rn,
// rs) codes and the FP32 v2pz array.  The in-plane part of H1 needs a and b
// of the staged plane (raw slot cxx = AB), which the kernel then keeps for
// the plane's output 4 planes later; the output plane needs CR (slots cxz,
// cyz) and v2pz (slot vpz).  Offsets are formed for 4-byte elements and
// doubled for CR (every array below 4 GiB, checked at launch).  Decode
// parameters: c_lo/c_st[0..5] for a, b, c, rx, rn, rs.
template <> struct CoefIO<DCX_COEF_N16> {
  static constexpr uint32_t kBytes = 4;
  static __device__ __forceinline__ uint32_t ldr(const StepArgs &a, int k,
                                                 uint32_t o) {
    switch (k) {
      case CXX: return ldb<uint32_t>(a.cf[CXX], o);
      case CXZ: return ldb<uint32_t>(a.cf[CXZ], 2u * o);
      case CYZ: return ldb<uint32_t>(a.cf[CXZ], 2u * o + 4u);
      case V2PZ: return ldb<uint32_t>(a.cf[V2PZ], o);
      default: return 0u;
    }
  }
  // Decode of the low or high code of r; off = c_lo[k], passed in from a
  // VGPR (see fast_tile).
  static __device__ __forceinline__ float lo(const StepArgs &a, int k,
                                             uint32_t r, float off) {
    return fmaf((float)(r & 0xffffu), a.c_st[k], off);
  }
  static __device__ __forceinline__ float hi(const StepArgs &a, int k,
                                             uint32_t r, float off) {
    return fmaf((float)(r >> 16), a.c_st[k], off);
  }
};

// Per-step global inputs, raw: coefficients of the staged plane (in-plane
// part of H1) and of the output plane, and the output plane's old state
// (previous level for leapfrog storage, time difference otherwise).
struct StepLoads {
  uint32_t cxx, cyy, cxy, czz, cxz, cyz, vpx, vpz, vsz, vpn, opp, oqp;
};
// Own and apron values of a staged plane, raw.
struct Staged {
  uint32_t cp, cq, ap, aq, ap2, aq2;
};

// PF = 1: the staged plane is fetched one plane ahead and the step's
// coefficients at the top of the step.  PF = 2: the staged plane two
// planes ahead and the coefficients one step ahead, doubling the bytes each
// wave keeps in flight (more VGPRs; same occupancy at 512 threads).  PF = 3:
// the coefficients and old fields two steps ahead.
// FS/CS select the field and coefficient storage.
template <int TX, int TY, int PF, int FS, int CS, bool OB = false>
__device__ __forceinline__ void fast_tile(const StepArgs &a, const int b,
                                          const float src_amp) {
  using F = FieldIO<FS>;
  using Cf = CoefIO<CS>;
  constexpr bool kTimeDiff = FS != DCX_STORE_F32;
  constexpr bool kAxis = CS == DCX_COEF_N16;
  // Spacing factors of the derivatives.  The axis form leaves them out of
  // the in-plane and mixed derivatives (the factor becomes 1, which the
  // compiler drops): its decoded axis is sqrt(2) (n_x/dx, n_y/dy, n_z/dz),
  // set per launch, so the products of two components are the angle terms
  // over d_a d_b (the squares twice theirs); only the Laplacian's sum scales
  // (Sp, H2).  Its second differences take half weights (kh) to undo the
  // doubled squares.
  auto spf = [](float d) { return kAxis ? 1.0f : d; };
  constexpr float kh = kAxis ? 0.5f : 1.0f;
  constexpr int NT = TX * TY;
  constexpr int PW = TX + 8;       // LDS plane row pitch (4-cell apron)
  constexpr int PH = TY + 8;
  constexpr int PL = PW * PH;      // cells of one staged plane
  constexpr int NAPR = PL - TX * TY;
  constexpr int NDXA = 8 * TX;     // D1x values of the 8 apron rows
  // Apron cells per thread: one for tiles of 16+ rows, two for 8-row tiles.
  constexpr bool kApron2 = NAPR > NT;
  static_assert(NAPR <= 2 * NT, "apron must fit two cells per thread");
  static_assert(NDXA <= NT, "D1x apron must fit one cell per thread");

  // Two barriers per plane (OB false): one buffer each suffices, since
  // every LDS read of plane n happens before the second barrier of step n
  // and every write of plane n+1 after it.  One barrier per plane (OB true,
  // the axis form): two buffers each, plane n in buffer n & 1; plane n+1 is
  // staged into the other buffer before step n's barrier, and every read of
  // a buffer ends before the barrier that precedes its next write.  The
  // compact forms are bound by issue more than by memory and run 5-15%
  // faster with one barrier (the axis form 10%); plain FP32 runs 11% slower.
  static_assert(!OB || PF == 2, "one barrier per plane needs PF 2");
  constexpr int NB = OB ? 2 : 1;
  __shared__ float sP_[NB][PL], sQ_[NB][PL];              // staged planes
  __shared__ float sDP_[NB][TX * PH], sDQ_[NB][TX * PH];  // D1x, rows -4..TY+3

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int64_t zs = a.zs;
  const int ys = (int)a.ys;
  // Tile b of the volume: x fastest, then y, then z chunk.
  const int bx = b % a.gx, by = (b / a.gx) % a.gy, bz = b / (a.gx * a.gy);
  const int tx0 = (int)tile_origin_x(a.box.x0) + bx * TX;
  const int ty0 = (int)a.box.y0 + by * TY;
  // Last coordinate any updated cell reads; loads are clamped to it so that a
  // tile overhanging the box never leaves the array.
  const int xl = (int)a.box.x1 + 3, yl = (int)a.box.y1 + 3;

  const int x = tx0 + tx, y = ty0 + ty;
  const bool act = x >= a.box.x0 && x < a.box.x1 && y < a.box.y1;
  // In-plane offset of this thread's cell (one plane holds < 2^31 cells) and
  // its byte offsets in field and coefficient arrays.
  const int own = (x < xl ? x : xl) + (y < yl ? y : yl) * ys;
  const uint32_t ob = (uint32_t)own * F::kBytes;
  // Coefficient offset: the pitched layout, or a dense x band (cw > 0).
  const int ownc = a.cw > 0 ? min(max(x - a.cx0, 0), a.cw - 1) + (y < yl ? y : yl) * a.cw
                            : own;
  const uint32_t obc = (uint32_t)ownc * Cf::kBytes;
  const int lo = (ty + 4) * PW + tx + 4;  // own cell in the LDS plane
  const int ld = (ty + 4) * TX + tx;      // own cell in the D1x plane

  // Apron cell t of the staged plane (top rows, bottom rows, then side
  // columns): its LDS index (-1 past the apron) and in-plane array offset.
  auto apron_cell = [&](int t, int &la_, int &oa_) {
    la_ = -1;
    oa_ = 0;
    if (t >= NAPR) return;
    int ax, ay;
    if (t < 4 * PW) {
      ay = t / PW;
      ax = t % PW;
    } else if (t < 8 * PW) {
      const int s = t - 4 * PW;
      ay = TY + 4 + s / PW;
      ax = s % PW;
    } else {
      const int s = t - 8 * PW;
      ay = 4 + s / 8;
      const int c8 = s % 8;
      ax = c8 < 4 ? c8 : TX + c8;
    }
    la_ = ay * PW + ax;
    int gx = tx0 + ax - 4, gy = ty0 + ay - 4;
    gx = gx < xl ? gx : xl;
    gy = gy < yl ? gy : yl;
    oa_ = gx + gy * ys;
  };
  int la, oa, la2 = -1, oa2 = 0;
  apron_cell(tid, la, oa);
  if (kApron2) apron_cell(tid + NT, la2, oa2);
  const uint32_t oab = (uint32_t)oa * F::kBytes;
  const uint32_t oab2 = (uint32_t)oa2 * F::kBytes;
  // D1x of the apron rows: rows 0..3 and TY+4..TY+7 of the D1x plane.
  int dr = -1, dc = 0;
  if (tid < NDXA) {
    const int r = tid / TX;
    dc = tid % TX;
    dr = r < 4 ? r : TY + r;
  }

  const float dt2 = a.dt * a.dt;
  // Time-difference storage keeps D in dscale units of the field.
  const float dt2d = dt2 * a.dscale;
  const float inv_d = 1.0f / a.dscale;
  const void *const Gpc = a.sf[0], *const Gqc = a.sf[1];
  void *const Gpp = a.of[0], *const Gqp = a.of[1];
  void *const Gpn = a.nf[0], *const Gqn = a.nf[1];
  // Byte offset of plane k: every array is smaller than 4 GiB (checked at
  // launch), so one 32-bit offset per plane serves all arrays of a kind.
  const uint32_t zsb = (uint32_t)zs * F::kBytes;
  const uint32_t zsbc = (uint32_t)(a.cw > 0 ? a.czs : zs) * Cf::kBytes;
  const int zb = (int)a.box.z0 + bz * a.zchunk;
  const int ze = zb + a.zchunk < (int)a.box.z1 ? zb + a.zchunk : (int)a.box.z1;

  // Register rings (slot = plane index - (zb-4), mod 8 / mod 4).
  float rp[8], rq[8], xp[8], yp[8], xq[8], yq[8];
  float Sp[4], Hp[4], Sq[4], Hq[4];
  // Axis form: a and b of the staged planes' cells, kept for their output.
  float Na[4], Nb[4];
  // Axis form: the six decode offsets, held in VGPRs for the whole tile.  A
  // VOP3 instruction reads at most one SGPR, so a decode fma whose scale and
  // offset both sat in SGPRs would first copy one of them into a VGPR, once
  // per decode; the empty asm makes each offset a VGPR value the compiler
  // cannot re-derive from its SGPR.
  float clo[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    clo[k] = CS == DCX_COEF_N16 ? a.c_lo[k] : 0.0f;
    if constexpr (CS == DCX_COEF_N16) asm volatile("" : "+v"(clo[k]));
  }
  // Whether a 16-bit store of this thread saturated (or met a non-finite
  // value); counted once per wave in a.ovf at the end of the tile.
  bool ovf = false;
#pragma unroll
  for (int k = 0; k < 8; ++k) rp[k] = rq[k] = xp[k] = yp[k] = xq[k] = yq[k] = 0.f;
#pragma unroll
  for (int k = 0; k < 4; ++k) Sp[k] = Hp[k] = Sq[k] = Hq[k] = Na[k] = Nb[k] = 0.f;

  const int nend = ze + 4;  // planes zb-4 .. ze+3 are staged
  auto load_staged = [&](int k) {
    Staged v{0u, 0u, 0u, 0u, 0u, 0u};
    const uint32_t bb = (uint32_t)(k < nend ? k : nend - 1) * zsb;
    v.cp = F::ldr(Gpc, bb + ob);
    v.cq = F::ldr(Gqc, bb + ob);
    if (la >= 0) {
      v.ap = F::ldr(Gpc, bb + oab);
      v.aq = F::ldr(Gqc, bb + oab);
    }
    if (kApron2 && la2 >= 0) {
      v.ap2 = F::ldr(Gpc, bb + oab2);
      v.aq2 = F::ldr(Gqc, bb + oab2);
    }
    return v;
  };
  // Step k's loads: plane k's in-plane coefficients, output plane k-4
  // (clamped into the chunk) for the rest.
  auto load_step = [&](int k) {
    StepLoads v;
    const int kc = k < nend ? k : nend - 1;
    const int kj = k - 4 > zb ? (k - 4 < ze ? k - 4 : ze - 1) : zb;
    const uint32_t bc = (uint32_t)kc * zsbc + obc;
    const uint32_t bjc = (uint32_t)kj * zsbc + obc;
    const uint32_t bjf = (uint32_t)kj * zsb + ob;
    v.cxx = Cf::ldr(a, CXX, bc);
    v.cyy = Cf::ldr(a, CYY, bc);
    v.cxy = Cf::ldr(a, CXY, bc);
    v.czz = Cf::ldr(a, CZZ, bjc);
    v.cxz = Cf::ldr(a, CXZ, bjc);
    v.cyz = Cf::ldr(a, CYZ, bjc);
    v.vpx = Cf::ldr(a, V2PX, bjc);
    v.vpz = Cf::ldr(a, V2PZ, bjc);
    v.vsz = Cf::ldr(a, V2SZ, bjc);
    v.vpn = Cf::ldr(a, V2PN, bjc);
    v.opp = F::ldr(Gpp, bjf);
    v.oqp = F::ldr(Gqp, bjf);
    return v;
  };

  Staged cur = load_staged(zb - 4);
  Staged nx1 = cur;
  StepLoads lcur{}, lnx{};
  // PF == 2 keeps the step values of plane n in stp[u & 1], a slot indexed
  // by the plane's place u in the unrolled block of 8 (a constant in every
  // copy), so that each load writes the registers its consumer reads.
  StepLoads stp[2];
  if (PF >= 2) nx1 = load_staged(zb - 3);
  if (PF == 2) stp[0] = load_step(zb - 4);
  if (PF >= 3) {
    lcur = load_step(zb - 4);
    lnx = load_step(zb - 3);
  }

  // Write a staged plane's own and apron cells into LDS buffer k; returns
  // its own cells' decoded p and q.
  auto stage = [&](const Staged &sc, int k, float &cp_, float &cq_) {
    cp_ = F::dec(sc.cp);
    cq_ = F::dec(sc.cq);
    sP_[k][lo] = cp_;
    sQ_[k][lo] = cq_;
    if (la >= 0) {
      sP_[k][la] = F::dec(sc.ap);
      sQ_[k][la] = F::dec(sc.aq);
    }
    if (kApron2 && la2 >= 0) {
      sP_[k][la2] = F::dec(sc.ap2);
      sQ_[k][la2] = F::dec(sc.aq2);
    }
  };
  // One barrier per plane: plane zb-4 is staged ahead of the loop.
  float cpn = 0.0f, cqn = 0.0f;
  if constexpr (OB) {
    stage(cur, 0, cpn, cqn);
    __syncthreads();
  }

  // Planes zb-4 .. zb+3 (the first block of 8) fill the z rings without
  // output; every later plane n updates output plane j = n - 4.  The output
  // decision is one per block, not one per plane.
  for (int nb = zb - 4; nb < nend; nb += 8) {
    const bool out = nb != zb - 4;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int n = nb + u;
      if (n >= nend) break;
      // This plane's LDS buffer (every block of 8 starts in buffer 0).
      const int bi = OB ? (u & 1) : 0;
      float *const sP = sP_[bi];
      float *const sQ = sQ_[bi];
      float *const sDP = sDP_[bi];
      float *const sDQ = sDQ_[bi];
      float cp, cq;
      if constexpr (OB) {
        cp = cpn;
        cq = cqn;
      } else {
        // Stage plane n.
        stage(cur, 0, cp, cq);
        __syncthreads();
      }

      StepLoads L;
      Staged nxt;
      if (PF >= 3) {
        L = lcur;
        nxt = nx1;
        nx1 = load_staged(n + 2);
        lcur = lnx;
        lnx = load_step(n + 2);
      } else if (PF == 2) {
        L = stp[u & 1];
        nxt = nx1;
        nx1 = load_staged(n + 2);
        stp[(u + 1) & 1] = load_step(n + 1);
      } else {
        L = load_step(n);
        nxt = load_staged(n + 1);
      }
      // L stays raw here; each value is decoded where it is used (the
      // output block, the ring push), after this iteration's barriers.

      // In-plane derivatives of plane n.
      const Pairs pxq = lds_pairs(sP, lo, 1), pyq = lds_pairs(sP, lo, PW);
      const Pairs qxq = lds_pairs(sQ, lo, 1), qyq = lds_pairs(sQ, lo, PW);
      const float pxx = d2_of(cp, pxq, kh) * spf(a.dxx);
      const float pyy = d2_of(cp, pyq, kh) * spf(a.dyy);
      const float qxx = d2_of(cq, qxq, kh) * spf(a.dxx);
      const float qyy = d2_of(cq, qyq, kh) * spf(a.dyy);
      const float dxp = d1_of(pxq), dyp = d1_of(pyq);
      const float dxq = d1_of(qxq), dyq = d1_of(qyq);
      sDP[ld] = dxp;
      sDQ[ld] = dxq;
      if (dr >= 0) {
        const int l = dr * PW + dc + 4;
        sDP[dr * TX + dc] = d1_lds(sP, l, 1);
        sDQ[dr * TX + dc] = d1_lds(sQ, l, 1);
      }
      // One barrier per plane: stage plane n+1 into the other buffer now.
      if constexpr (OB) stage(nxt, (u + 1) & 1, cpn, cqn);
      __syncthreads();
      const float pxy = d1_lds(sDP, ld, TX) * spf(a.dxy);
      const float qxy = d1_lds(sDQ, ld, TX) * spf(a.dxy);

      // Output plane j = n - 4 once the rings hold planes j-4 .. j+4.
      if (out) {
        const int J = (u + 4) & 7;  // slot of plane j
        const int O = u;            // slot of plane j-4 (about to be replaced)
#define RS(r, b) r[(J + (b)) & 7]
        const float pzz = ((kh * W_K0) * rp[J] +
                           (kh * W_K1) * (RS(rp, 1) + RS(rp, 7)) +
                           (kh * W_K2) * (RS(rp, 2) + RS(rp, 6)) +
                           (kh * W_K3) * (RS(rp, 3) + RS(rp, 5)) +
                           (kh * W_K4) * (cp + rp[O])) * spf(a.dzz);
        const float qzz = ((kh * W_K0) * rq[J] +
                           (kh * W_K1) * (RS(rq, 1) + RS(rq, 7)) +
                           (kh * W_K2) * (RS(rq, 2) + RS(rq, 6)) +
                           (kh * W_K3) * (RS(rq, 3) + RS(rq, 5)) +
                           (kh * W_K4) * (cq + rq[O])) * spf(a.dzz);
        const float pxz = (W_L1 * (RS(xp, 1) - RS(xp, 7)) +
                           W_L2 * (RS(xp, 2) - RS(xp, 6)) +
                           W_L3 * (RS(xp, 3) - RS(xp, 5)) +
                           W_L4 * (dxp - xp[O])) * spf(a.dxz);
        const float pyz = (W_L1 * (RS(yp, 1) - RS(yp, 7)) +
                           W_L2 * (RS(yp, 2) - RS(yp, 6)) +
                           W_L3 * (RS(yp, 3) - RS(yp, 5)) +
                           W_L4 * (dyp - yp[O])) * spf(a.dyz);
        const float qxz = (W_L1 * (RS(xq, 1) - RS(xq, 7)) +
                           W_L2 * (RS(xq, 2) - RS(xq, 6)) +
                           W_L3 * (RS(xq, 3) - RS(xq, 5)) +
                           W_L4 * (dxq - xq[O])) * spf(a.dxz);
        const float qyz = (W_L1 * (RS(yq, 1) - RS(yq, 7)) +
                           W_L2 * (RS(yq, 2) - RS(yq, 6)) +
                           W_L3 * (RS(yq, 3) - RS(yq, 5)) +
                           W_L4 * (dyq - yq[O])) * spf(a.dyz);
#undef RS
        const int D = u & 3;
        float rhsp, rhsq;
        // Axis form: v2pz dt^2, the factor of both right-hand sides.
        float vdt = 0.0f;
        if constexpr (kAxis) {
          // n = (na, nb, nc) as decoded (above): czz = nc^2 with the
          // half-weight pzz, cxz = na nc, cyz = nb nc, with na and nb kept
          // from the plane's staging.  Sp and pzz dzz are half the
          // Laplacian's terms.  v2px, v2pn, v2sz are v2pz times the ratios
          // rx, rn, rs, so v2pz factors out (and is applied with dt^2 in the
          // update):
          //   rhsp = v2pz (rx H2p + H1q + rs (H1p - H1q)),
          //   rhsq = v2pz (rn H2p + H1q - rs (H2p - H2q)).
          const float nc = Cf::lo(a, 2, L.cxz, clo[2]);
          const float czz = nc * nc, cxz = nc * Na[D], cyz = nc * Nb[D];
          const float h1p = Hp[D] + czz * pzz + cxz * pxz + cyz * pyz;
          const float h2p = 2.0f * (Sp[D] + pzz * a.dzz) - h1p;
          const float h1q = Hq[D] + czz * qzz + cxz * qxz + cyz * qyz;
          const float h2q = 2.0f * (Sq[D] + qzz * a.dzz) - h1q;
          vdt = __uint_as_float(L.vpz) * dt2;
          const float rx = Cf::hi(a, 3, L.cxz, clo[3]);
          const float rn = Cf::lo(a, 4, L.cyz, clo[4]);
          const float rs = Cf::hi(a, 5, L.cyz, clo[5]);
          rhsp = rx * h2p + h1q + rs * (h1p - h1q);
          rhsq = rn * h2p + h1q - rs * (h2p - h2q);
        } else {
          const float czz = Cf::dec(a, CZZ, L.czz);
          const float cxz = Cf::dec(a, CXZ, L.cxz);
          const float cyz = Cf::dec(a, CYZ, L.cyz);
          const float vpx = Cf::dec(a, V2PX, L.vpx);
          const float vpz = Cf::dec(a, V2PZ, L.vpz);
          const float vsz = Cf::dec(a, V2SZ, L.vsz);
          const float vpn = Cf::dec(a, V2PN, L.vpn);
          const float h1p = Hp[D] + czz * pzz + cxz * pxz + cyz * pyz;
          const float h2p = Sp[D] + pzz - h1p;
          const float h1q = Hq[D] + czz * qzz + cxz * qxz + cyz * qyz;
          const float h2q = Sq[D] + qzz - h1q;
          rhsp = vpx * h2p + vpz * h1q + vsz * (h1p - h1q);
          rhsq = vpn * h2p + vpz * h1q - vsz * (h2p - h2q);
        }
        const float opp = F::dec(L.opp), oqp = F::dec(L.oqp);
        const int j = n - 4;
        const uint32_t bo = (uint32_t)j * zsb + ob;
        if (kTimeDiff) {
          const bool src = j == a.src_z && (int64_t)own == a.src_xy;
          // D_new = D + dt^2 rhs; P_new = P + D_new; the next step's source
          // raises P_new, and with it D_new, by src_amp (field units).
          float dnp = opp + rhsp * dt2d, dnq = oqp + rhsq * dt2d;
          float pn = rp[J] + dnp * inv_d, qn = rq[J] + dnq * inv_d;
          if (src) {
            pn += src_amp;
            qn += src_amp;
            dnp += src_amp * a.dscale;
            dnq += src_amp * a.dscale;
          }
          if (act) {
            F::st(Gpn, bo, pn);
            F::st(Gqn, bo, qn);
            F::st(Gpp, bo, dnp);
            F::st(Gqp, bo, dnq);
            ovf = ovf || F::bad(pn) || F::bad(qn) || F::bad(dnp) || F::bad(dnq);
          }
        } else {
          // (rhs * dt) * dt, the reference's association
          // (include/sample_compute.h:181-184); the axis form multiplies its
          // rhs by v2pz dt^2 instead.
          float pn, qn;
          if constexpr (kAxis) {
            pn = 2.0f * rp[J] - opp + rhsp * vdt;
            qn = 2.0f * rq[J] - oqp + rhsq * vdt;
          } else {
            pn = 2.0f * rp[J] - opp + rhsp * a.dt * a.dt;
            qn = 2.0f * rq[J] - oqp + rhsq * a.dt * a.dt;
          }
          // The source is added after the tile (below).
          if (act) {
            F::st(Gpp, bo, pn);
            F::st(Gqp, bo, qn);
          }
        }
      }
      // Push plane n into the rings.
      rp[u] = cp;
      rq[u] = cq;
      xp[u] = dxp;
      yp[u] = dyp;
      xq[u] = dxq;
      yq[u] = dyq;
      float cxx, cyy, cxy;
      if constexpr (CS == DCX_COEF_N16) {
        const float na = Cf::lo(a, 0, L.cxx, clo[0]);
        const float nb = Cf::hi(a, 1, L.cxx, clo[1]);
        cxx = na * na;
        cyy = nb * nb;
        cxy = na * nb;
        Na[u & 3] = na;
        Nb[u & 3] = nb;
      } else {
        cxx = Cf::dec(a, CXX, L.cxx);
        cyy = Cf::dec(a, CYY, L.cyy);
        cxy = Cf::dec(a, CXY, L.cxy);
      }
      if constexpr (kAxis) {
        Sp[u & 3] = pxx * a.dxx + pyy * a.dyy;
        Sq[u & 3] = qxx * a.dxx + qyy * a.dyy;
      } else {
        Sp[u & 3] = pxx + pyy;
        Sq[u & 3] = qxx + qyy;
      }
      Hp[u & 3] = cxx * pxx + cyy * pyy + cxy * pxy;
      Hq[u & 3] = cxx * qxx + cyy * qyy + cxy * qxy;

      cur = nxt;
    }
  }
  // Only a run that has already left the form's range gets here with ovf
  // set; the flag stays set until the storage is converted again.
  if constexpr (kTimeDiff) {
    if (ovf && a.ovf != nullptr) atomicOr(a.ovf, 1);
  } else {
    // The next step's source, added by the lane that stored the source cell
    // to the value it stored: the same two roundings as adding it before
    // the store (the reference's add-then-swap, src/device_data.c:44-59),
    // with no per-plane test.  No other lane reads the new level's cells.
    if (act && a.src_z >= zb && a.src_z < ze && (int64_t)own == a.src_xy) {
      const uint32_t bs = (uint32_t)a.src_z * zsb + ob;
      F::st(Gpp, bs, F::ld(Gpp, bs) + src_amp);
      F::st(Gqp, bs, F::ld(Gqp, bs) + src_amp);
    }
  }
}

template <int TX, int TY, int PF, int FS, int CS, bool OB = false>
__global__ void __launch_bounds__(TX *TY) k_fast(const StepArgs a) {
  // Workgroups are dealt to the 8 XCDs round-robin by linear id.  Mode 1
  // gives each XCD (and its L2) a contiguous run of tiles, x fastest, so the
  // apron rows and columns a tile shares with its neighbours stay in one L2.
  // Mode 2 gives XCD k the whole tile rows r = k mod 8, in the default row
  // order: x neighbours share an L2, and the tiles in flight at once still
  // cover one contiguous band of rows.
  int b = blockIdx.x;
  const int G = gridDim.x;
  if (a.xcd_remap == 1 && G % 8 == 0) {
    b = (b % 8) * (G / 8) + b / 8;
  } else if (a.xcd_remap == 2 && G % a.gx == 0 && (G / a.gx) % 8 == 0) {
    const int j = b / 8;
    b = j % a.gx + a.gx * (b % 8 + 8 * (j / a.gx));
  }
  fast_tile<TX, TY, PF, FS, CS, OB>(a, b, a.src_amp);
}

// Persistent launch over the tiles of many volumes (shots): workgroup g takes
// work items g, g + G, g + 2G, ... of the concatenated tile lists; every
// workgroup runs the same number of items' worth of barriers on its own.
template <int TX, int TY, int PF, int FS, int CS, bool OB = false>
__global__ void __launch_bounds__(TX *TY)
    k_fast_batch(const StepArgs *__restrict__ shots,
                 const int64_t *__restrict__ prefix, int nshots,
                 int64_t total, int it, int n_steps) {
  for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
    int lo = 0, hi = nshots - 1;  // last shot with prefix <= w
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= w)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int k = __builtin_amdgcn_readfirstlane(lo);
    const StepArgs &a = shots[k];
    const float amp = (a.src_z >= 0 && a.src_tab != nullptr && it + 1 < n_steps)
                          ? a.src_tab[it + 1]
                          : 0.0f;
    fast_tile<TX, TY, PF, FS, CS, OB>(a, (int)(w - prefix[k]), amp);
  }
}

// Several boxes of one array set in one launch (queued dc_propagate calls):
// workgroup b finds its box by the tile prefix sums (uniform per workgroup)
// and runs that box's tile with the box's own bounds, tile grid, z chunk and
// coefficient arrays.
template <int TX, int TY, int PF, int FS, int CS>
__global__ void __launch_bounds__(TX *TY)
    k_fast_multi(const StepArgs a, const MultiBox m) {
  const int b = blockIdx.x;
  int k = 0;
  while (k + 1 < m.n && m.prefix[k + 1] <= b) ++k;
  StepArgs s = a;
  s.box = m.box[k];
  s.gx = m.gx[k];
  s.gy = m.gy[k];
  s.zchunk = m.zchunk[k];
  s.cx0 = m.cx0[k];
  s.cw = m.cw[k];
  s.czs = m.czs[k];
#pragma unroll
  for (int c = 0; c < NCOEF; ++c) s.cf[c] = m.cf[k][c];
  fast_tile<TX, TY, PF, FS, CS>(s, b - (int)m.prefix[k], a.src_amp);
}

void launch_fast_multi(const StepArgs &a, const MultiBox &m, hipStream_t s) {
  if (m.n <= 0 || m.prefix[m.n] <= 0) return;
  if (a.fstore != DCX_STORE_F32 || a.cstore != DCX_COEF_F32)
    DCX_FAIL("multi-box launches run FP32 storage only");
  if (m.prefix[m.n] >= (int64_t(1) << 31)) DCX_FAIL("too many tiles");
  for (int k = 0; k < m.n; ++k)
    if ((m.box[k].z1 + 4) * a.zs * 4 > (int64_t(1) << 32) - 1)
      DCX_FAIL("fast kernel needs arrays below 4 GiB");
  const dim3 grid((unsigned)m.prefix[m.n]), block(32 * 16);
  constexpr int F = DCX_STORE_F32, C = DCX_COEF_F32;
  if (a.deep >= 2)
    hipLaunchKernelGGL((k_fast_multi<32, 16, 3, F, C>), grid, block, 0, s, a, m);
  else if (a.deep == 1)
    hipLaunchKernelGGL((k_fast_multi<32, 16, 2, F, C>), grid, block, 0, s, a, m);
  else
    hipLaunchKernelGGL((k_fast_multi<32, 16, 1, F, C>), grid, block, 0, s, a, m);
  DCX_CHECK(hipGetLastError());
}

// Tile shapes with a kernel instance: 512 threads (32x16, 64x8) or 1024
// threads (32x32, 64x16; at 1024 threads a lane has 128 registers, and these
// spill).
bool fast_tiling_supported(int tx, int ty) {
  return (tx == 32 && (ty == 16 || ty == 32)) || (tx == 64 && (ty == 8 || ty == 16));
}

// Every form but plain FP32 runs one barrier per plane (see fast_tile).
template <int FS, int CS>
constexpr bool kOneBarrier = !(FS == DCX_STORE_F32 && CS == DCX_COEF_F32);

template <int TX, int TY, int FS, int CS>
static void launch_fast_t(const StepArgs &b, dim3 grid, dim3 block,
                          hipStream_t s) {
  if constexpr (kOneBarrier<FS, CS>) {
    if (b.deep) {
      hipLaunchKernelGGL((k_fast<TX, TY, 2, FS, CS, true>), grid, block, 0, s, b);
      return;
    }
  }
  // Three-level prefetch is built for 32x16 FP32 only (it needs the
  // registers the other forms spend on decoding); they take two levels.
  if constexpr (TX == 32 && TY == 16 && FS == DCX_STORE_F32 && CS == DCX_COEF_F32) {
    if (b.deep >= 2) {
      hipLaunchKernelGGL((k_fast<TX, TY, 3, FS, CS>), grid, block, 0, s, b);
      return;
    }
  }
  if (b.deep)
    hipLaunchKernelGGL((k_fast<TX, TY, 2, FS, CS>), grid, block, 0, s, b);
  else
    hipLaunchKernelGGL((k_fast<TX, TY, 1, FS, CS>), grid, block, 0, s, b);
}

template <int TX, int TY>
static void launch_fast_tile(const StepArgs &b, dim3 grid, dim3 block,
                             hipStream_t s) {
#define DCX_FORM(FS, CS)                                                       \
  if (b.fstore == FS && b.cstore == CS)                                        \
    return launch_fast_t<TX, TY, FS, CS>(b, grid, block, s);
  DCX_FORM(DCX_STORE_F32, DCX_COEF_F32)
  DCX_FORM(DCX_STORE_F32, DCX_COEF_F16)
  DCX_FORM(DCX_STORE_F32, DCX_COEF_U16)
  DCX_FORM(DCX_STORE_F32, DCX_COEF_A16)
  DCX_FORM(DCX_STORE_F32, DCX_COEF_N16)
  DCX_FORM(DCX_STORE_F16, DCX_COEF_F32)
  DCX_FORM(DCX_STORE_F16, DCX_COEF_F16)
  DCX_FORM(DCX_STORE_F16, DCX_COEF_U16)
  DCX_FORM(DCX_STORE_BF16, DCX_COEF_F16)
  DCX_FORM(DCX_STORE_BF16, DCX_COEF_U16)
#undef DCX_FORM
  DCX_FAIL("unsupported storage combination %d/%d", b.fstore, b.cstore);
}

int64_t fast_tiles(const StepArgs &a, int tile_x, int tile_y) {
  const int64_t nx = a.box.x1 - tile_origin_x(a.box.x0), ny = a.box.y1 - a.box.y0,
                nz = a.box.z1 - a.box.z0;
  if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
  return ((nx + tile_x - 1) / tile_x) * ((ny + tile_y - 1) / tile_y) *
         ((nz + a.zchunk - 1) / a.zchunk);
}

template <int TX, int TY, int FS, int CS>
static void launch_batch_t(const StepArgs *shots, const int64_t *prefix,
                           int nshots, int64_t total, int deep, dim3 grid,
                           int it, int n_steps, hipStream_t s) {
  if constexpr (kOneBarrier<FS, CS>) {
    if (deep) {
      hipLaunchKernelGGL((k_fast_batch<TX, TY, 2, FS, CS, true>), grid,
                         dim3(TX * TY), 0, s, shots, prefix, nshots, total, it,
                         n_steps);
      return;
    }
  }
  if constexpr (FS == DCX_STORE_F32 && CS == DCX_COEF_F32) {
    if (deep >= 2) {
      hipLaunchKernelGGL((k_fast_batch<TX, TY, 3, FS, CS>), grid,
                         dim3(TX * TY), 0, s, shots, prefix, nshots, total, it,
                         n_steps);
      return;
    }
  }
  if (deep)
    hipLaunchKernelGGL((k_fast_batch<TX, TY, 2, FS, CS>), grid,
                       dim3(TX * TY), 0, s, shots, prefix, nshots, total, it,
                       n_steps);
  else
    hipLaunchKernelGGL((k_fast_batch<TX, TY, 1, FS, CS>), grid,
                       dim3(TX * TY), 0, s, shots, prefix, nshots, total, it,
                       n_steps);
}

void launch_fast_batch(const StepArgs *shots, const int64_t *prefix,
                       int nshots, int64_t total, int fstore, int cstore,
                       int deep, int tile_x, int tile_y, int grid, int it,
                       int n_steps, hipStream_t s) {
  if (total <= 0 || nshots <= 0) return;
  if (tile_x != 32 || tile_y != 16)
    DCX_FAIL("batched launches use 32x16 tiles (asked %dx%d)", tile_x, tile_y);
  const dim3 g((unsigned)(grid > 0 ? grid : 1));
#define DCX_BFORM(FS, CS)                                                      \
  if (fstore == FS && cstore == CS) {                                          \
    launch_batch_t<32, 16, FS, CS>(shots, prefix, nshots, total, deep, g, it,  \
                                   n_steps, s);                                \
    DCX_CHECK(hipGetLastError());                                              \
    return;                                                                    \
  }
  DCX_BFORM(DCX_STORE_F32, DCX_COEF_F32)
  DCX_BFORM(DCX_STORE_F32, DCX_COEF_U16)
  DCX_BFORM(DCX_STORE_F32, DCX_COEF_A16)
  DCX_BFORM(DCX_STORE_F32, DCX_COEF_N16)
  DCX_BFORM(DCX_STORE_F16, DCX_COEF_F16)
  DCX_BFORM(DCX_STORE_F16, DCX_COEF_U16)
#undef DCX_BFORM
  DCX_FAIL("unsupported batched storage combination %d/%d", fstore, cstore);
}

void launch_fast(const StepArgs &a, int tile_x, int tile_y, hipStream_t s) {
  const int64_t nx = a.box.x1 - a.box.x0, ny = a.box.y1 - a.box.y0,
                nz = a.box.z1 - a.box.z0;
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  if (a.zchunk <= 0) DCX_FAIL("z chunk must be positive");
  // The kernel addresses every array with 32-bit byte offsets (the axis
  // form's qword records take 8 bytes per cell).
  const int64_t ebytes = a.cstore == DCX_COEF_N16 ? 8 : 4;
  if ((a.box.z1 + 4) * a.zs * ebytes > (int64_t(1) << 32) - 1)
    DCX_FAIL("fast kernel needs arrays below 4 GiB (%lld planes of %lld cells)",
             (long long)(a.box.z1 + 4), (long long)a.zs);
  const int64_t gz = (nz + a.zchunk - 1) / a.zchunk;
  StepArgs b = a;
  b.gx = (int)((a.box.x1 - tile_origin_x(a.box.x0) + tile_x - 1) / tile_x);
  b.gy = (int)((ny + tile_y - 1) / tile_y);
  const int64_t nblocks = (int64_t)b.gx * b.gy * gz;
  if (nblocks >= (int64_t(1) << 31)) DCX_FAIL("too many tiles");
  dim3 grid((unsigned)nblocks);
  dim3 block(tile_x * tile_y);
  if (tile_x == 32 && tile_y == 16) launch_fast_tile<32, 16>(b, grid, block, s);
  else if (tile_x == 32 && tile_y == 32) launch_fast_tile<32, 32>(b, grid, block, s);
  else if (tile_x == 64 && tile_y == 16) launch_fast_tile<64, 16>(b, grid, block, s);
  else if (tile_x == 64 && tile_y == 8) launch_fast_tile<64, 8>(b, grid, block, s);
  else DCX_FAIL("unsupported tile %dx%d", tile_x, tile_y);
  DCX_CHECK(hipGetLastError());
}

// -------------------------------------------------------------- misc ----

__global__ void k_add_source(float *pc, float *qc, int64_t index, float amp) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    pc[index] += amp;
    qc[index] += amp;
  }
}

void launch_add_source(float *pc, float *qc, int64_t index, float amp,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_add_source, dim3(1), dim3(64), 0, s, pc, qc, index, amp);
  DCX_CHECK(hipGetLastError());
}

// The time step's bytes without its stencil: every cell of the box reads
// pc, qc, pp, qp and the ten coefficients at its own index and writes pp and
// qp (a pointwise leapfrog with the coefficients summed in), 64 B per cell
// over the same pitched arrays.  The rate the memory system gives this
// 16-stream pattern is the ceiling for a step that moves the same bytes.
// One thread per cell, 256-cell rows along x.
__global__ void __launch_bounds__(256) k_stream_probe(const StepArgs a) {
  const int64_t x = a.box.x0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t y = a.box.y0 + blockIdx.y;
  const int64_t z = a.box.z0 + blockIdx.z;
  if (x >= a.box.x1) return;
  const int64_t i = x + y * a.ys + z * a.zs;
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < NCOEF; ++k) s += a.coef.c[k][i];
  a.pp[i] = 2.0f * a.pc[i] - a.pp[i] + s * a.dt;
  a.qp[i] = 2.0f * a.qc[i] - a.qp[i] + s * a.dt;
}

void launch_stream_probe(const StepArgs &a, hipStream_t s) {
  const int64_t nx = a.box.x1 - a.box.x0, ny = a.box.y1 - a.box.y0,
                nz = a.box.z1 - a.box.z0;
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  if (ny > 65535 || nz > 65535) DCX_FAIL("probe box too large");
  dim3 grid((unsigned)((nx + 255) / 256), (unsigned)ny, (unsigned)nz);
  hipLaunchKernelGGL(k_stream_probe, grid, dim3(256), 0, s, a);
  DCX_CHECK(hipGetLastError());
}

}  // namespace dcx
