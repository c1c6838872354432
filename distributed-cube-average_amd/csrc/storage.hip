// Compact storage forms of one rank's arrays (dcx_set_storage).
//
// Fields: the leapfrog pair (p_{n-1}, p_n) of each of p and q is re-expressed
// as P = p_n and D = p_n - p_{n-1} (the time difference) and stored as half or
// bfloat16 in units of a power-of-two field scale (D with an extra power of
// two).  Rounding P each step then perturbs only the displacement; in the
// leapfrog form the same rounding also perturbs the implied velocity, which
// the scheme amplifies by ~1/(omega dt) in the signal band.
// Coefficients: the ten arrays as half (power-of-two scaled per array) or as
// 16-bit codes (affine for the six angle terms, affine in log2 for the four
// v^2 terms).  Encoding and decoding run on the host at conversion time; the
// kernels decode in registers.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dc_internal.hpp"

namespace dcx {
namespace {

__device__ __forceinline__ uint16_t enc_dev(int bf, float v) {
  if (bf) {
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
  }
  // Saturating, as the stencil's stores: never inf or NaN.
  return __builtin_bit_cast(uint16_t,
                            (_Float16)fminf(fmaxf(v, -65504.0f), 65504.0f));
}

// Per block: max |a| and max |a - b| over the sx x nrows logical cells of
// two pitched arrays (element x = 0 of row 0), out[2 block + {0, 1}].
__global__ void __launch_bounds__(256)
    k_absmax(const float *a, const float *b, int64_t sx, int64_t ys,
             int64_t nrows, float *out) {
  float m0 = 0.0f, m1 = 0.0f;
  const int64_t n = sx * nrows;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / sx;
    const int64_t j = (i - row * sx) + row * ys;
    m0 = fmaxf(m0, fabsf(a[j]));
    m1 = fmaxf(m1, fabsf(a[j] - b[j]));
  }
  __shared__ float s[2][256];
  s[0][threadIdx.x] = m0;
  s[1][threadIdx.x] = m1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s[0][threadIdx.x] = fmaxf(s[0][threadIdx.x], s[0][threadIdx.x + w]);
      s[1][threadIdx.x] = fmaxf(s[1][threadIdx.x], s[1][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 2) out[2 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__device__ __forceinline__ float dec_dev(int bf, uint16_t v) {
  if (bf) return __uint_as_float((uint32_t)v << 16);
  return (float)__builtin_bit_cast(_Float16, v);
}

// All conversion kernels walk whole allocations (padding included) with a
// grid-stride loop; the padding stays consistent (zeros encode to zeros).
constexpr int kConvBlocks = 2048;

__global__ void k_encode_fields(const float *pc, const float *pp, uint16_t *P0,
                                uint16_t *P1, uint16_t *D, int64_t n, float fs,
                                float ds, int bf) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint16_t p = enc_dev(bf, pc[i] * fs);
    P0[i] = p;
    P1[i] = p;  // both levels share the ghost frame
    D[i] = enc_dev(bf, (pc[i] - pp[i]) * fs * ds);
  }
}

__global__ void k_decode_fields(const uint16_t *P, const uint16_t *D, float *pc,
                                float *pp, int64_t n, float ifs, float ids,
                                int bf) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float p = dec_dev(bf, P[i]);
    const float d = dec_dev(bf, D[i]) * ids;
    pc[i] = p * ifs;
    pp[i] = (p - d) * ifs;
  }
}

// Per-block (min, max, min positive, max positive) of the sx x nrows logical
// cells of a pitched array (a at x = 0, rows ys apart).  The row pitch, the
// front pad and the tail are left out: no stencil reads them, and counting
// their zeros would widen the code range of every array whose values are
// all positive (and stop a constant array from coding exactly).
__global__ void __launch_bounds__(256)
    k_minmax(const float *a, int64_t sx, int64_t ys, int64_t nrows,
             float *out) {
  float lo = INFINITY, hi = -INFINITY, plo = INFINITY, phi = -INFINITY;
  const int64_t n = sx * nrows;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / sx;
    const float v = a[(i - row * sx) + row * ys];
    lo = fminf(lo, v);
    hi = fmaxf(hi, v);
    if (v > 0.0f) {
      plo = fminf(plo, v);
      phi = fmaxf(phi, v);
    }
  }
  __shared__ float s[4][256];
  s[0][threadIdx.x] = lo;
  s[1][threadIdx.x] = hi;
  s[2][threadIdx.x] = plo;
  s[3][threadIdx.x] = phi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s[0][threadIdx.x] = fminf(s[0][threadIdx.x], s[0][threadIdx.x + w]);
      s[1][threadIdx.x] = fmaxf(s[1][threadIdx.x], s[1][threadIdx.x + w]);
      s[2][threadIdx.x] = fminf(s[2][threadIdx.x], s[2][threadIdx.x + w]);
      s[3][threadIdx.x] = fmaxf(s[3][threadIdx.x], s[3][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) out[4 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// mode 0: half of a*scale; 1: round((a-lo)/step); 2: 0 for a <= 0, else
// 1 + round((log2 a - lo)/step).  Codes saturate to [0, 65535].
__global__ void k_encode_coef(const float *a, uint16_t *out, int64_t n,
                              int mode, double lo, double step, float scale) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float v = a[i];
    uint16_t c;
    if (mode == 0) {
      c = __builtin_bit_cast(uint16_t, (_Float16)(v * scale));
    } else {
      double t;
      if (mode == 1) {
        t = ((double)v - lo) / step;
      } else {
        t = v > 0.0f ? 1.0 + (log2((double)v) - lo) / step : 0.0;
      }
      t = t < 0.0 ? 0.0 : (t > 65535.0 ? 65535.0 : t);
      c = (uint16_t)__double2int_rn(t);
    }
    out[i] = c;
  }
}

// DCX_COEF_N16: the TTI axis n = (a, b, c) and the three v^2 ratios of one
// cell, in double, from its ten FP32 coefficients.  n is found from the
// largest diagonal term (its square root) and the two cross terms beside it;
// n and -n give the same terms, so c >= 0.  v[3..5] = v2px/v2pz,
// v2pn/v2pz, v2sz/v2pz where v2pz > 0 (ratio_ok), else 0.
struct AxisCell {
  double v[6];
  bool ratio_ok;
};

__device__ AxisCell axis_of(const float c10[NCOEF]) {
  const double xx = c10[CXX], yy = c10[CYY], zz = c10[CZZ];
  const double xy = c10[CXY], yz = c10[CYZ], xz = c10[CXZ];
  double a = 0.0, b = 0.0, c = 0.0;
  if (zz >= xx && zz >= yy) {
    c = sqrt(zz);
    if (c > 0.0) {
      a = xz / (2.0 * c);
      b = yz / (2.0 * c);
    }
  } else if (xx >= yy) {
    a = sqrt(xx);
    b = xy / (2.0 * a);
    c = xz / (2.0 * a);
  } else {
    b = sqrt(yy);
    a = xy / (2.0 * b);
    c = yz / (2.0 * b);
  }
  if (c < 0.0) {
    a = -a;
    b = -b;
    c = -c;
  }
  AxisCell r;
  r.v[0] = a;
  r.v[1] = b;
  r.v[2] = c;
  const double vz = c10[V2PZ];
  r.ratio_ok = vz > 0.0;
  r.v[3] = r.ratio_ok ? c10[V2PX] / vz : 0.0;
  r.v[4] = r.ratio_ok ? c10[V2PN] / vz : 0.0;
  r.v[5] = r.ratio_ok ? c10[V2SZ] / vz : 0.0;
  return r;
}

struct Coef10 {
  const float *c[NCOEF];  // pitched, element x = 0 of row 0
};

__device__ void load10(const Coef10 &cs, int64_t i, float c10[NCOEF]) {
#pragma unroll
  for (int k = 0; k < NCOEF; ++k) c10[k] = cs.c[k][i];
}

// Per block: min then max of the six axis-form values over the logical cells
// (ratios over cells with v2pz > 0 only): out[12 * block + 2k + {0, 1}].
__global__ void __launch_bounds__(256)
    k_axis_range(Coef10 cs, int64_t sx, int64_t ys, int64_t nrows,
                 double *out) {
  double lo[6], hi[6];
  for (int k = 0; k < 6; ++k) {
    lo[k] = INFINITY;
    hi[k] = -INFINITY;
  }
  const int64_t n = sx * nrows;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / sx;
    float c10[NCOEF];
    load10(cs, (i - row * sx) + row * ys, c10);
    const AxisCell r = axis_of(c10);
    for (int k = 0; k < 6; ++k) {
      if (k >= 3 && !r.ratio_ok) continue;
      lo[k] = fmin(lo[k], r.v[k]);
      hi[k] = fmax(hi[k], r.v[k]);
    }
  }
  __shared__ double s[12][256];
  for (int k = 0; k < 6; ++k) {
    s[2 * k][threadIdx.x] = lo[k];
    s[2 * k + 1][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 6; ++k) {
        s[2 * k][threadIdx.x] = fmin(s[2 * k][threadIdx.x], s[2 * k][threadIdx.x + w]);
        s[2 * k + 1][threadIdx.x] =
            fmax(s[2 * k + 1][threadIdx.x], s[2 * k + 1][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 12) out[12 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

struct AxisCode {
  float lo[6], st[6];
};

__device__ __forceinline__ uint32_t axis_code(double v, float lo, float st) {
  double t = rint((v - (double)lo) / (double)st);
  t = t < 0.0 ? 0.0 : (t > 65535.0 ? 65535.0 : t);
  return (uint32_t)t;
}

// Codes of every logical cell, and per block the largest error of the terms
// the codes stand for: absolute for the six angle terms, formed by the fast
// kernel's own float operations (out[2 block]), and relative for the three
// v^2 terms as v2pz times the decoded ratio (out[2 block + 1]; the kernel
// factors v2pz out of the right-hand sides instead of forming them).
__global__ void __launch_bounds__(256)
    k_axis_encode(Coef10 cs, int64_t sx, int64_t ys, int64_t nrows,
                  AxisCode q, uint32_t *ab, uint64_t *cr, float *out) {
  float e_ang = 0.0f, e_vel = 0.0f;
  const int64_t n = sx * nrows;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / sx;
    const int64_t j = (i - row * sx) + row * ys;
    float c10[NCOEF];
    load10(cs, j, c10);
    const AxisCell r = axis_of(c10);
    uint32_t u[6];
    for (int k = 0; k < 6; ++k) u[k] = axis_code(r.v[k], q.lo[k], q.st[k]);
    ab[j] = u[0] | (u[1] << 16);
    cr[j] = (uint64_t)(u[2] | (u[3] << 16)) |
            ((uint64_t)(u[4] | (u[5] << 16)) << 32);
    float d[6];
    for (int k = 0; k < 6; ++k) d[k] = fmaf((float)u[k], q.st[k], q.lo[k]);
    const float c2 = d[2] + d[2];
    const float ang[6] = {d[0] * d[0], d[1] * d[1], d[2] * d[2],
                          (d[0] + d[0]) * d[1], c2 * d[1], c2 * d[0]};
    for (int k = 0; k < 6; ++k) e_ang = fmaxf(e_ang, fabsf(ang[k] - c10[k]));
    const float vz = c10[V2PZ];
    const float vel[3] = {vz * d[3], vz * d[4], vz * d[5]};
    const float ref[3] = {c10[V2PX], c10[V2PN], c10[V2SZ]};
    for (int k = 0; k < 3; ++k) {
      const float m = fabsf(ref[k]);
      const float e = fabsf(vel[k] - ref[k]);
      e_vel = fmaxf(e_vel, m > 0.0f ? e / m : (e > 0.0f ? INFINITY : 0.0f));
    }
  }
  __shared__ float s[2][256];
  s[0][threadIdx.x] = e_ang;
  s[1][threadIdx.x] = e_vel;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s[0][threadIdx.x] = fmaxf(s[0][threadIdx.x], s[0][threadIdx.x + w]);
      s[1][threadIdx.x] = fmaxf(s[1][threadIdx.x], s[1][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 2) out[2 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// Build the DCX_COEF_N16 codes of st's FP32 coefficient arrays into st->C16
// (allocated here).  Returns false, with st->C16 freed, when the arrays are
// not of the TTI form within the stated error.
bool encode_axis_form(State *st) {
  const int blocks = kConvBlocks;
  const dc_precomp_vars &pv = st->abi.precomp_vars;
  Coef10 cs{{pv.ch1dxx, pv.ch1dyy, pv.ch1dzz, pv.ch1dxy, pv.ch1dyz, pv.ch1dxz,
             pv.v2px, pv.v2pz, pv.v2sz, pv.v2pn}};
  const int64_t sx = st->sizes[0], nrows = st->sizes[1] * st->sizes[2];
  double *d_rng = nullptr;
  DCX_CHECK(hipMalloc(&d_rng, 12 * blocks * sizeof(double)));
  hipLaunchKernelGGL(k_axis_range, dim3(blocks), dim3(256), 0, st->stream, cs,
                     sx, st->ys, nrows, d_rng);
  DCX_CHECK(hipGetLastError());
  std::vector<double> rng(12 * blocks);
  DCX_CHECK(hipMemcpyAsync(rng.data(), d_rng, rng.size() * sizeof(double),
                           hipMemcpyDeviceToHost, st->stream));
  DCX_CHECK(hipStreamSynchronize(st->stream));
  DCX_CHECK(hipFree(d_rng));
  AxisCode q;
  for (int k = 0; k < 6; ++k) {
    double lo = INFINITY, hi = -INFINITY;
    for (int b = 0; b < blocks; ++b) {
      lo = std::min(lo, rng[12 * b + 2 * k]);
      hi = std::max(hi, rng[12 * b + 2 * k + 1]);
    }
    if (!(lo <= hi)) lo = hi = 0.0;  // no cell with v2pz > 0
    q.lo[k] = (float)lo;
    q.st[k] = hi > lo ? (float)((hi - lo) / 65535.0) : 1.0f;
    st->c_lo[k] = q.lo[k];
    st->c_st[k] = q.st[k];
  }
  const size_t bytes = (size_t)st->alloc_floats * 12;
  uint16_t *raw = nullptr;
  DCX_CHECK(hipMalloc(&raw, bytes));
  DCX_CHECK(hipMemsetAsync(raw, 0, bytes, st->stream));
  st->C16 = raw + kPad;
  float *d_err = nullptr;
  DCX_CHECK(hipMalloc(&d_err, 2 * blocks * sizeof(float)));
  hipLaunchKernelGGL(k_axis_encode, dim3(blocks), dim3(256), 0, st->stream, cs,
                     sx, st->ys, nrows, q, n16_ab(st), n16_cr(st), d_err);
  DCX_CHECK(hipGetLastError());
  std::vector<float> err(2 * blocks);
  DCX_CHECK(hipMemcpyAsync(err.data(), d_err, err.size() * sizeof(float),
                           hipMemcpyDeviceToHost, st->stream));
  DCX_CHECK(hipStreamSynchronize(st->stream));
  DCX_CHECK(hipFree(d_err));
  float e_ang = 0.0f, e_vel = 0.0f;
  for (int b = 0; b < blocks; ++b) {
    e_ang = std::max(e_ang, err[2 * b]);
    e_vel = std::max(e_vel, err[2 * b + 1]);
  }
  // NaN compares false: a NaN error fails too.
  if (!(e_ang <= 1e-4f) || !(e_vel <= 1e-4f)) {
    std::fprintf(stderr,
                 "[dc_hip] DCX_COEF_N16: coefficients are not of the TTI axis "
                 "form (largest angle-term error %.3g, v^2 relative error "
                 "%.3g; bound 1e-4)\n",
                 (double)e_ang, (double)e_vel);
    DCX_CHECK(hipFree(raw));
    st->C16 = nullptr;
    return false;
  }
  return true;
}

uint16_t *alloc16(const State *st) {
  uint16_t *p = nullptr;
  const size_t bytes = (size_t)st->alloc_floats * 2;
  DCX_CHECK(hipMalloc(&p, bytes));
  DCX_CHECK(hipMemsetAsync(p, 0, bytes, st->stream));
  return p + kPad;
}

void free16(uint16_t *&p) {
  if (p) DCX_CHECK(hipFree(p - kPad));
  p = nullptr;
}

// Whole-allocation views (the pointers above are offset by kPad).
template <class T>
T *whole(T *p) {
  return p - kPad;
}

void decode_into(State *st, int f, float *pc, float *pp) {
  hipLaunchKernelGGL(k_decode_fields, dim3(kConvBlocks), dim3(256), 0,
                     st->stream, whole(st->P16[f][st->pcur]), whole(st->D16[f]),
                     whole(pc), whole(pp), st->alloc_floats, 1.0f / st->fscale,
                     1.0f / st->dscale, st->fstore == DCX_STORE_BF16 ? 1 : 0);
  DCX_CHECK(hipGetLastError());
}

}  // namespace

void storage_free(State *st) {
  for (int f = 0; f < 2; ++f) {
    free16(st->P16[f][0]);
    free16(st->P16[f][1]);
    free16(st->D16[f]);
  }
  if (st->C16) DCX_CHECK(hipFree(st->C16 - kPad));
  st->C16 = nullptr;
  if (st->ovf) DCX_CHECK(hipFree(st->ovf));
  st->ovf = nullptr;
}

bool storage_overflowed(State *st) {
  if (st->ovf == nullptr || st->fstore == DCX_STORE_F32) return false;
  int flag = 0;
  DCX_CHECK(hipMemcpyAsync(&flag, st->ovf, sizeof(int), hipMemcpyDeviceToHost,
                           st->stream));
  DCX_CHECK(hipStreamSynchronize(st->stream));
  return flag != 0;
}

void report_overflow(State *st) {
  if (st->ovf_reported || !storage_overflowed(st)) return;
  st->ovf_reported = true;
  std::fprintf(stderr,
               "[dc_hip] warning: %s field storage left its range (a stored "
               "value saturated%s); the fields are not valid. Use FP32 "
               "fields for this run (dcx_storage_overflow)\n",
               st->fstore == DCX_STORE_F16 ? "FP16" : "BF16",
               st->fstore == DCX_STORE_F16 ? " at +-65504 field units"
                                           : " or was not finite");
}

// Host copy of field `which` (0 pp, 1 pc, 2 qp, 3 qc) of a compact struct:
// decoded on the device into scratch arrays, then copied out.
void storage_download(const State *st_c, int which, float *host) {
  State *st = const_cast<State *>(st_c);
  float *tmp = nullptr;
  DCX_CHECK(hipMalloc(&tmp, (size_t)st->alloc_floats * 2 * sizeof(float)));
  float *pc = tmp + kPad, *pp = tmp + st->alloc_floats + kPad;
  decode_into(st, which / 2, pc, pp);
  // The packed staging reuses the other decoded array's memory.
  pitched_to_host(st, host, (which % 2) ? pc : pp, tmp + (which % 2 ? st->alloc_floats : 0));
  DCX_CHECK(hipFree(tmp));
}

// Add a source amplitude (true units) at device index i of a compact struct.
__global__ void k_add_source16(uint16_t *P0, uint16_t *P1, uint16_t *D0,
                               uint16_t *D1, int64_t i, float amp, float ds,
                               int bf) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint16_t *Ps[2] = {P0, P1};
  uint16_t *Ds[2] = {D0, D1};
  for (int f = 0; f < 2; ++f) {
    Ps[f][i] = enc_dev(bf, dec_dev(bf, Ps[f][i]) + amp);
    Ds[f][i] = enc_dev(bf, dec_dev(bf, Ds[f][i]) + amp * ds);
  }
}

void storage_add_source(State *st, int64_t dev_index, float amp) {
  hipLaunchKernelGGL(k_add_source16, dim3(1), dim3(64), 0, st->stream,
                     st->P16[0][st->pcur], st->P16[1][st->pcur], st->D16[0],
                     st->D16[1], dev_index, amp * st->fscale, st->dscale,
                     st->fstore == DCX_STORE_BF16 ? 1 : 0);
  DCX_CHECK(hipGetLastError());
}

}  // namespace dcx

using namespace dcx;

extern "C" int dcx_set_storage(dc_device_data *data, int field_store,
                               int coef_store, int field_scale_log2) {
  State *st = state_of(data);
  const bool fs_ok = field_store == DCX_STORE_F32 ||
                     field_store == DCX_STORE_F16 ||
                     field_store == DCX_STORE_BF16;
  const bool cs_ok = coef_store == DCX_COEF_F32 || coef_store == DCX_COEF_F16 ||
                     coef_store == DCX_COEF_U16 || coef_store == DCX_COEF_A16 ||
                     coef_store == DCX_COEF_N16;
  if (!fs_ok || !cs_ok) return -1;
  if (field_store == DCX_STORE_BF16 && coef_store == DCX_COEF_F32) return -1;
  if ((coef_store == DCX_COEF_A16 || coef_store == DCX_COEF_N16) &&
      field_store != DCX_STORE_F32)
    return -1;
  // A batch's launch tables and kernel follow the storage form the shots had
  // when the batch was made; converting one shot would mix element sizes.
  if (st->batch)
    DCX_FAIL("dcx_set_storage on a batch shot: free the batch "
             "(dcx_batch_free) before converting its shots");
  // The wall runs on FP32 storage and holds speeds computed from the FP32
  // coefficients when it was set.
  if (st->wall != DCX_WALL_ZERO)
    DCX_FAIL("dcx_set_storage with a one-way wall set: dcx_set_wall(data, "
             "DCX_WALL_ZERO) first");
  DCX_CHECK(hipStreamSynchronize(st->stream));

  // Bring the fields back to FP32 first if they are compact.
  if (st->fstore != DCX_STORE_F32) {
    for (int f = 0; f < 2; ++f)
      decode_into(st, f, f == 0 ? st->abi.pc : st->abi.qc,
                  f == 0 ? st->abi.pp : st->abi.qp);
    DCX_CHECK(hipStreamSynchronize(st->stream));
    for (int f = 0; f < 2; ++f) {
      free16(st->P16[f][0]);
      free16(st->P16[f][1]);
      free16(st->D16[f]);
    }
    st->fstore = DCX_STORE_F32;
  }
  if (field_store != DCX_STORE_F32) {
    st->fstore = field_store;
    // Scales: powers of two from the fields themselves, so that the largest
    // |p|, |q| and |p_n - p_{n-1}| land at <= 2^12 field units, 16x below
    // half's largest finite value; the caller's scale (and 2^3 for D) where
    // the fields are still zero.  bfloat16 has FP32's range: scale 1, D x 8.
    float maxp = 0.0f, maxd = 0.0f;
    if (field_store == DCX_STORE_F16) {
      float *d_mx = nullptr;
      DCX_CHECK(hipMalloc(&d_mx, 2 * kConvBlocks * sizeof(float)));
      std::vector<float> mx(2 * kConvBlocks);
      for (int f = 0; f < 2; ++f) {
        hipLaunchKernelGGL(k_absmax, dim3(kConvBlocks), dim3(256), 0,
                           st->stream, f == 0 ? st->abi.pc : st->abi.qc,
                           f == 0 ? st->abi.pp : st->abi.qp, st->sizes[0],
                           st->ys, st->sizes[1] * st->sizes[2], d_mx);
        DCX_CHECK(hipGetLastError());
        DCX_CHECK(hipMemcpyAsync(mx.data(), d_mx, mx.size() * sizeof(float),
                                 hipMemcpyDeviceToHost, st->stream));
        DCX_CHECK(hipStreamSynchronize(st->stream));
        for (int b = 0; b < kConvBlocks; ++b) {
          maxp = std::max(maxp, mx[2 * b]);
          maxd = std::max(maxd, mx[2 * b + 1]);
        }
      }
      DCX_CHECK(hipFree(d_mx));
    }
    constexpr float kHead = 4096.0f;
    st->fscale = 1.0f;
    st->dscale = 8.0f;
    if (field_store == DCX_STORE_F16) {
      st->fscale = maxp > 0.0f
                       ? std::ldexp(1.0f, (int)std::floor(std::log2(kHead / maxp)))
                       : std::ldexp(1.0f, field_scale_log2);
      if (maxd > 0.0f) {
        const int j = (int)std::floor(std::log2(kHead / (maxd * st->fscale)));
        st->dscale = std::ldexp(1.0f, std::min(16, std::max(0, j)));
      }
    }
    if (st->ovf == nullptr) DCX_CHECK(hipMalloc(&st->ovf, sizeof(int)));
    DCX_CHECK(hipMemsetAsync(st->ovf, 0, sizeof(int), st->stream));
    st->ovf_reported = false;
    st->pcur = 0;
    for (int f = 0; f < 2; ++f) {
      st->P16[f][0] = alloc16(st);
      st->P16[f][1] = alloc16(st);
      st->D16[f] = alloc16(st);
      hipLaunchKernelGGL(k_encode_fields, dim3(kConvBlocks), dim3(256), 0,
                         st->stream, whole(f == 0 ? st->abi.pc : st->abi.qc),
                         whole(f == 0 ? st->abi.pp : st->abi.qp),
                         whole(st->P16[f][0]), whole(st->P16[f][1]),
                         whole(st->D16[f]), st->alloc_floats, st->fscale,
                         st->dscale, field_store == DCX_STORE_BF16 ? 1 : 0);
      DCX_CHECK(hipGetLastError());
    }
  }

  if (st->C16) DCX_CHECK(hipFree(st->C16 - kPad));
  st->C16 = nullptr;
  st->cstore = coef_store;
  if (coef_store == DCX_COEF_N16) {
    if (!encode_axis_form(st)) {
      st->cstore = DCX_COEF_F32;
      build_halo_tables(st);
      return -1;
    }
  } else if (coef_store != DCX_COEF_F32) {
    DCX_CHECK(hipMalloc(&st->C16, (size_t)st->alloc_floats * 2 * NCOEF));
    st->C16 += kPad;
    const dc_precomp_vars &pv = st->abi.precomp_vars;
    const float *cs[NCOEF] = {pv.ch1dxx, pv.ch1dyy, pv.ch1dzz, pv.ch1dxy,
                              pv.ch1dyz, pv.ch1dxz, pv.v2px,   pv.v2pz,
                              pv.v2sz,   pv.v2pn};
    float *d_mm = nullptr;
    DCX_CHECK(hipMalloc(&d_mm, 4 * kConvBlocks * sizeof(float)));
    std::vector<float> mm(4 * kConvBlocks);
    // DCX_COEF_A16 codes the angle terms only; its v^2 terms stay in the
    // FP32 arrays.
    const int ncoded = coef_store == DCX_COEF_A16 ? V2PX : NCOEF;
    for (int k = 0; k < ncoded; ++k) {
      hipLaunchKernelGGL(k_minmax, dim3(kConvBlocks), dim3(256), 0, st->stream,
                         cs[k], st->sizes[0], st->ys,
                         st->sizes[1] * st->sizes[2], d_mm);
      DCX_CHECK(hipGetLastError());
      DCX_CHECK(hipMemcpyAsync(mm.data(), d_mm, mm.size() * sizeof(float),
                               hipMemcpyDeviceToHost, st->stream));
      DCX_CHECK(hipStreamSynchronize(st->stream));
      float lo = INFINITY, hi = -INFINITY, plo = INFINITY, phi = -INFINITY;
      for (int b = 0; b < kConvBlocks; ++b) {
        lo = std::min(lo, mm[4 * b]);
        hi = std::max(hi, mm[4 * b + 1]);
        plo = std::min(plo, mm[4 * b + 2]);
        phi = std::max(phi, mm[4 * b + 3]);
      }
      int mode;
      double elo = 0.0, estep = 1.0;
      float scale = 1.0f;
      if (coef_store == DCX_COEF_F16) {
        const float m = std::max(std::fabs(lo), std::fabs(hi));
        // Largest magnitude scaled to [2^13, 2^14): far from half's overflow.
        const int e = m > 0 ? 13 - (int)std::floor(std::log2(m)) : 0;
        scale = std::ldexp(1.0f, e);
        mode = 0;
        st->c_lo[k] = 0.0f;
        st->c_st[k] = std::ldexp(1.0f, -e);
      } else if (k < V2PX) {
        mode = 1;
        elo = lo;
        estep = hi > lo ? ((double)hi - lo) / 65535.0 : 1.0;
        st->c_lo[k] = (float)elo;
        st->c_st[k] = (float)estep;
      } else {
        if (lo < 0) DCX_FAIL("coefficient %d has negative values", k);
        mode = 2;
        double l0 = 0.0, l1 = 0.0;
        if (plo <= phi) {
          l0 = std::log2((double)plo);
          l1 = std::log2((double)phi);
        }
        // Codes 1..65535 span [l0, l1] in log2; code 0 is exactly zero.
        estep = l1 > l0 ? (l1 - l0) / 65534.0 : 1.0;
        elo = l0;
        st->c_lo[k] = (float)(l0 - estep);
        st->c_st[k] = (float)estep;
      }
      hipLaunchKernelGGL(k_encode_coef, dim3(kConvBlocks), dim3(256), 0,
                         st->stream, whole(cs[k]),
                         whole(st->C16 + (size_t)k * st->alloc_floats),
                         st->alloc_floats, mode, elo, estep, scale);
      DCX_CHECK(hipGetLastError());
    }
    DCX_CHECK(hipStreamSynchronize(st->stream));
    DCX_CHECK(hipFree(d_mm));
  }
  DCX_CHECK(hipStreamSynchronize(st->stream));
  build_halo_tables(st);
  return 0;
}

extern "C" int dcx_storage_overflow(dc_device_data *data) {
  return storage_overflowed(state_of(data)) ? 1 : 0;
}

namespace dcx {
namespace {
// Zero n 32-bit words (grid-stride; a fill kernel rather than a memset node,
// so the call can sit in a captured graph).
__global__ void __launch_bounds__(256) k_zero_words(uint32_t *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// `bytes` is a multiple of 4: the arrays hold alloc_floats (a multiple of 64)
// elements of 4 or 2 bytes, the overflow flag is one int.
void zero_bytes(State *st, void *p, size_t bytes) {
  hipLaunchKernelGGL(k_zero_words, dim3(kConvBlocks), dim3(256), 0, st->stream,
                     static_cast<uint32_t *>(p), (int64_t)(bytes / 4));
  DCX_CHECK(hipGetLastError());
}
}  // namespace
}  // namespace dcx

extern "C" void dcx_zero_fields(dc_device_data *data) {
  State *st = state_of(data);
  const size_t n = (size_t)st->alloc_floats;
  if (st->fstore == DCX_STORE_F32) {
    for (int f = 0; f < 4; ++f) zero_bytes(st, st->field_alloc[f], n * 4);
  } else {
    for (int f = 0; f < 2; ++f) {
      zero_bytes(st, whole(st->P16[f][0]), n * 2);
      zero_bytes(st, whole(st->P16[f][1]), n * 2);
      zero_bytes(st, whole(st->D16[f]), n * 2);
    }
    // Fields at rest are in range again: clear the overflow flag.
    zero_bytes(st, st->ovf, sizeof(int));
    st->ovf_reported = false;
  }
}

extern "C" int dcx_bytes_per_cell(const dc_device_data *data) {
  const State *st = state_of(data);
  const int fb = st->fstore == DCX_STORE_F32 ? 4 : 2;
  const int cb = st->cstore == DCX_COEF_F32 ? 4 : 2;
  if (st->cstore == DCX_COEF_A16) {
    // Angle terms 2 B each, v^2 terms 4 B each, FP32 leapfrog fields.
    return 6 * fb + V2PX * 2 + (NCOEF - V2PX) * 4;
  }
  if (st->cstore == DCX_COEF_N16) {
    // FP32 leapfrog fields; v2pz 4 B, six 16-bit codes 12 B.
    return 6 * fb + 4 + 12;
  }
  // Leapfrog: read pc, qc, pp, qp, write pp, qp.  Time difference: read P and
  // D, write P and D, per field.
  const int fields = st->fstore == DCX_STORE_F32 ? 6 * fb : 8 * fb;
  return fields + NCOEF * cb;
}
