// Learned absorbing layer: a small MLP evaluated on every cell of a thin
// layer along the six faces of a rank's compute region, correcting the new
// p and q the stencil wrote there, in place of a thick absorbing shell.
//
// Network (FP16 MFMA operands, FP32 accumulation):
//   x (64) = 3x3x3 patch of the time differences p^n - p^{n-1} and
//            q^n - q^{n-1} around the cell (54 values, in the frame of the
//            nearest face, normalised by their largest magnitude s), a
//            one-hot of the cell's depth in the layer (inputs 54..61), and
//            the cell's Courant numbers sqrt(v2px) dt/dx, sqrt(v2pz) dt/dx
//            (inputs 62, 63)
//   h1 = relu(W1 x + b1) (64), h2 = relu(W2 h1 + b2) (64),
//   y = W3 h2 + b3 (2);  correction (dp, dq) of the new p, q:
//     free form:  s * y                              (0 where s = 0)
//     damping:    -sigma_max * sigmoid(y) * (f^n - f^{n-1}) at the cell
// so the correction scales with the wave: (a * D) -> a * correction, a > 0;
// the damping form only takes energy out (0 <= sigma <= sigma_max < 1).
//
// MFMA orientation: weights are the A operand and activations (features x
// cells) the B operand of v_mfma_f32_16x16x16_f16, so the C tile of one
// layer (rows = units, columns = cells) is, lane for lane, the B operand of
// the next: no data movement between layers.  Lane l of a wave holds
//   A[l%16][4(l/16)+r], B[4(l/16)+r][l%16], C[4(l/16)+r][l%16], r = 0..3.
// Training reduces over samples, which needs the transposed tiles; they go
// through LDS.
#include <cstdint>
#include <cmath>
#include <vector>

#include "dc_internal.hpp"

namespace dcx {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 64, kHid = 64, kOut = 2;
constexpr int kPatch = 54;
constexpr int kMaxWidth = 8;
constexpr int kCourant = kPatch + kMaxWidth;  // inputs 62, 63
// fp32 parameter vector (host-visible order).
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid,
              oB2 = oW2 + kHid * kHid, oW3 = oB2 + kHid,
              oB3 = oW3 + kOut * kHid, kNParam = oB3 + kOut;
// Packed fp16 operands: W1 [64][64], W2 [64][64], W3 [16][64] (rows >= 2
// zero), W2^T [64][64], W3^T [64][16] (columns >= 2 zero).
// W1p [64][64] is W1 with its columns in the layer inference's slot order
// (see slot_feature).
constexpr int hW1 = 0, hW2 = hW1 + 4096, hW3 = hW2 + 4096, hW2t = hW3 + 1024,
              hW3t = hW2t + 4096, hW1p = hW3t + 1024, kNHalf = hW1p + 4096;

// Layer inference reads its inputs in slot order, chosen so that K steps 0-1
// (slots 0..31) touch only the p arrays and K steps 2-3 only the q arrays:
// slots 0..26 p patch, 27..31 inputs 54..58, 32..58 q patch, 59..63 inputs
// 59..63.  slot_feature maps a slot to its input, feature_slot back.
__host__ __device__ constexpr int slot_feature(int s) {
  return s < 27 ? s : s < 32 ? s + 27 : s < 59 ? s - 5 : s;
}
__host__ __device__ constexpr int feature_slot(int k) {
  return k < 27 ? k : k < 54 ? k + 5 : k < 59 ? k - 27 : k;
}

enum { kOutFree = 0, kOutDamping = 1 };

struct Mlp {
  int device;
  int attached;  // structs using the network (dcx_set_learned_layer)
  int out_mode;
  float sigma_max;
  hipStream_t stream;
  float *w32, *m, *v, *grad, *loss;
  _Float16 *wh;
  int64_t t;  // Adam steps taken
};

struct Samples {
  int device;
  int64_t cap, n;
  _Float16 *x;    // [cap][64]
  float *t;       // [cap][2] targets (normalised)
  float *scale;   // [cap]
  int64_t *cell;  // [cap] device offset of the cell in the small struct
  int64_t *src;   // [cap] device offset of the same cell in the padded one
  uint32_t *code; // [cap] packed compute coordinates in the small struct
};

// ---------------------------------------------------------------- features

// Compute region extents of a struct and the layer width.
struct LayerGeom {
  int nx, ny, nz, width;
  int ys;
  int64_t zs;
  const float *v2px, *v2pz;  // coefficient arrays of the struct
  float cdt;                 // dt / dx (as dt * sqrt(1 / dx^2))
};

// Cell codes: compute coordinates packed 10 bits each.
__host__ __device__ inline uint32_t pack_cell(int x, int y, int z) {
  return (uint32_t)x | ((uint32_t)y << 10) | ((uint32_t)z << 20);
}

struct CellFrame {
  int64_t own;          // device offset of the cell
  int64_t sn, s1, s2;   // strides: outward normal, first and second tangent
  int depth;
};

// Nearest face (lowest axis, low side first on ties), its outward normal and
// the two tangent axes in increasing order.
__device__ inline CellFrame cell_frame(uint32_t code, const LayerGeom &g) {
  const int c[3] = {(int)(code & 1023u), (int)((code >> 10) & 1023u),
                    (int)((code >> 20) & 1023u)};
  const int n[3] = {g.nx, g.ny, g.nz};
  const int64_t st[3] = {1, g.ys, g.zs};
  int best = 1 << 30, axis = 0, sign = -1;
  for (int a = 0; a < 3; ++a) {
    const int lo = c[a], hi = n[a] - 1 - c[a];
    if (lo < best) { best = lo; axis = a; sign = -1; }
    if (hi < best) { best = hi; axis = a; sign = 1; }
  }
  CellFrame f;
  f.own = (int64_t)(c[0] + STENCIL) + (int64_t)(c[1] + STENCIL) * g.ys +
          (int64_t)(c[2] + STENCIL) * g.zs;
  f.sn = sign * st[axis];
  f.s1 = st[axis == 0 ? 1 : 0];
  f.s2 = st[axis == 2 ? 1 : 2];
  f.depth = best;
  return f;
}

// Raw feature k of a cell: a time difference (k < 54), the depth one-hot
// (54..61) or a Courant number (62, 63).  Patch order: field (p then q), then
// (i, j, k) in {-1,0,1}^3 along (normal, tangent 1, tangent 2), k fastest.
__device__ inline float raw_feature(int k, const CellFrame &f,
                                    const LayerGeom &g, const float *pc,
                                    const float *pp, const float *qc,
                                    const float *qp) {
  if (k < kPatch) {
    const int q = k >= 27;
    const int idx = k - 27 * q;
    const int i = idx / 9 - 1, j = (idx / 3) % 3 - 1, kk = idx % 3 - 1;
    const int64_t o = f.own + i * f.sn + j * f.s1 + kk * f.s2;
    return q ? qc[o] - qp[o] : pc[o] - pp[o];
  }
  if (k < kCourant) return (k - kPatch) == f.depth ? 1.0f : 0.0f;
  return sqrtf(k == kCourant ? g.v2px[f.own] : g.v2pz[f.own]) * g.cdt;
}

__device__ inline float norm_feature(int k, float raw, float s) {
  if (k >= kPatch) return raw;
  return s > 0.0f ? raw / s : 0.0f;
}

// ------------------------------------------------------------ MLP forward

struct FwdWeights {
  h4 a1[4][4], a2[4][4], a3[4];
  f4 b1[4], b2[4], b3;
};

__device__ inline void load_fwd(FwdWeights &w, const _Float16 *wh,
                                const float *w32, int lane) {
  const int col = lane & 15, kb = lane >> 4;
#pragma unroll
  for (int ot = 0; ot < 4; ++ot) {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      w.a1[ot][ks] = *(const h4 *)(wh + hW1 + (16 * ot + col) * 64 + 16 * ks + 4 * kb);
      w.a2[ot][ks] = *(const h4 *)(wh + hW2 + (16 * ot + col) * 64 + 16 * ks + 4 * kb);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      w.b1[ot][r] = w32[oB1 + 16 * ot + 4 * kb + r];
      w.b2[ot][r] = w32[oB2 + 16 * ot + 4 * kb + r];
    }
  }
#pragma unroll
  for (int ks = 0; ks < 4; ++ks)
    w.a3[ks] = *(const h4 *)(wh + hW3 + col * 64 + 16 * ks + 4 * kb);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = 4 * kb + r;
    w.b3[r] = row < kOut ? w32[oB3 + row] : 0.0f;
  }
}

__device__ inline h4 relu_h(f4 z) {
  h4 h;
#pragma unroll
  for (int r = 0; r < 4; ++r) h[r] = (_Float16)(z[r] > 0.0f ? z[r] : 0.0f);
  return h;
}

// Forward pass of 16 cells (one per lane column); X[ks] is the lane's part of
// the input tile.  Returns the output tile (rows 0, 1 of lanes 0..15 hold y).
__device__ inline f4 mlp_forward(const FwdWeights &w, const h4 X[4]) {
  h4 H1[4], H2[4];
#pragma unroll
  for (int ot = 0; ot < 4; ++ot) {
    f4 acc = w.b1[ot];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      acc = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a1[ot][ks], X[ks], acc, 0, 0, 0);
    H1[ot] = relu_h(acc);
  }
#pragma unroll
  for (int ot = 0; ot < 4; ++ot) {
    f4 acc = w.b2[ot];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      acc = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a2[ot][ks], H1[ks], acc, 0, 0, 0);
    H2[ot] = relu_h(acc);
  }
  f4 y = w.b3;
#pragma unroll
  for (int ks = 0; ks < 4; ++ks)
    y = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a3[ks], H2[ks], y, 0, 0, 0);
  return y;
}

// Features of the lane's cell (column) for its four K slices, normalised,
// and the cell's scale s (the same on the four lanes of a column).
__device__ inline float cell_inputs(h4 X[4], bool valid, uint32_t code,
                                    const LayerGeom &g, const float *pc,
                                    const float *pp, const float *qc,
                                    const float *qp, int lane, int64_t *own) {
  const int kb = lane >> 4;
  float v[4][4];
  float m = 0.0f;
  CellFrame f{};
  if (valid) f = cell_frame(code, g);
  *own = f.own;
#pragma unroll
  for (int ks = 0; ks < 4; ++ks)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = 16 * ks + 4 * kb + r;
      const float raw = valid ? raw_feature(k, f, g, pc, pp, qc, qp) : 0.0f;
      v[ks][r] = raw;
      if (k < kPatch) m = fmaxf(m, fabsf(raw));
    }
  m = fmaxf(m, __shfl_xor(m, 16));
  m = fmaxf(m, __shfl_xor(m, 32));
#pragma unroll
  for (int ks = 0; ks < 4; ++ks)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      X[ks][r] = (_Float16)norm_feature(16 * ks + 4 * kb + r, v[ks][r], m);
  return m;
}

// What slot s of the layer inference holds: bit 0 set for a patch value,
// then (i+1, j+1, k+1) two bits each from bit 1; otherwise the input index
// (54..63) from bit 1.
__device__ inline int slot_desc(int s) {
  const int k = slot_feature(s);
  if (k < kPatch) {
    const int idx = k < 27 ? k : k - 27;
    return 1 | ((idx / 9) << 1) | (((idx / 3) % 3) << 3) | ((idx % 3) << 5);
  }
  return k << 1;
}

// ------------------------------------------------ layer inference by slabs
//
// The layer is cut into tiles of 16 x 16 cells of one face plane, all
// `width` cells deep.  A workgroup stages the time differences D = f^n -
// f^{n-1} of its tile's box, one cell beyond it on every side, in LDS once,
// then gathers every cell's patch from there.  A cell belongs to the face
// cell_frame picks for it; the tiles of other faces skip it.

constexpr int kTile = 16;                         // cells per tile side
// Staged box: x extent padded to an odd row stride (no LDS bank aliasing of
// the rows lanes read together); at most (8 + 2 + 1) x 18 x 18 cells.
constexpr int kBoxMax = (kMaxWidth + 3) * (kTile + 2) * (kTile + 2);

struct SlabTile {
  int face, t1, t2, pad;  // face = 2 * axis + side; tangent origins
};

// The tile's cell e (t1 fastest, then t2, then depth) in compute
// coordinates, and whether it is a layer cell of this face.
__device__ inline bool slab_cell(const SlabTile &T, int e, const LayerGeom &g,
                                 int c[3]) {
  const int axis = T.face >> 1, side = T.face & 1;
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  const int n[3] = {g.nx, g.ny, g.nz};
  const int d = e / (kTile * kTile);
  c[a1] = T.t1 + e % kTile;
  c[a2] = T.t2 + (e / kTile) % kTile;
  c[axis] = side ? n[axis] - 1 - d : d;
  if (d >= g.width || c[a1] >= n[a1] || c[a2] >= n[a2]) return false;
  // Nearest face by cell_frame's rule (lowest axis, low side first).
  int best = 1 << 30, bf = 0;
  for (int a = 0; a < 3; ++a) {
    if (c[a] < best) { best = c[a]; bf = 2 * a; }
    if (n[a] - 1 - c[a] < best) { best = n[a] - 1 - c[a]; bf = 2 * a + 1; }
  }
  return bf == T.face;
}

// Box of a tile in compute coordinates: origin and extent per axis.
__device__ inline void slab_box(const SlabTile &T, const LayerGeom &g, int lo[3],
                                int len[3]) {
  const int axis = T.face >> 1, side = T.face & 1;
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  const int n[3] = {g.nx, g.ny, g.nz};
  lo[axis] = side ? n[axis] - 1 - g.width : -1;
  len[axis] = g.width + 2;
  lo[a1] = T.t1 - 1;
  len[a1] = kTile + 2;
  lo[a2] = T.t2 - 1;
  len[a2] = kTile + 2;
}

__device__ inline int64_t dev_offset(const LayerGeom &g, int x, int y, int z) {
  return (int64_t)(x + STENCIL) + (int64_t)(y + STENCIL) * g.ys +
         (int64_t)(z + STENCIL) * g.zs;
}

// Corrections of the layer cells from the current fields (before the step),
// on v_mfma_f32_16x16x32_f16.  Lane l = 16 kb + col of a wave holds
//   A[col][8kb + j], B[8kb + j][col] (j = 0..7), C[4kb + r][col] (r = 0..3).
// A wave runs 16 cells at a time (one per column).  Inputs: K step s gives
// lane kb the slots 32s + 8kb + j, so K step 0 reads the p arrays and K step
// 1 the q arrays (slot_feature), and only lanes kb = 3 hold non-patch inputs,
// in their slots j >= 3: depth one-hot 0..4 (s = 0), one-hot 5..7 and the two
// Courant numbers (s = 1).  Every input is one LDS read at a per-lane offset
// fixed for the whole tile: the time differences of the tile's box, the
// Courant numbers of its cells and a one-hot table are staged side by side.
// Between layers, C tiles 2s and 2s+1 (units 32s + 4kb + r and 32s + 16 +
// 4kb + r) are, lane for lane, the B operand of K step s; the next layer's
// weight columns are permuted to match when they are staged (inf_unit).
__host__ __device__ constexpr int inf_unit(int s, int p) {  // p = 8kb + j
  return 32 * s + ((p & 7) < 4 ? 4 * (p >> 3) + (p & 7)
                               : 16 + 4 * (p >> 3) + (p & 7) - 4);
}

typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// ReLU after the conversion to half, as packed half maxima (rounding keeps
// the sign, so max(0, round(x)) = round(max(0, x)); a -0 weighs as 0).
__device__ inline h4 relu_pk(f4 z) {
  const h4 h = __builtin_convertvector(z, h4);
  return __builtin_elementwise_max(h, (h4)(_Float16)0.0f);
}

// floor(e / n) for 0 <= e < 4096 and 1 <= n <= 400 from rn = 1.0f / n: the
// float quotient is within 4.2e-4 of e / n and a non-integer e / n is at
// least 1/400 above its floor (checked exhaustively over that range).
__device__ inline int small_div(int e, float rn) {
  return (int)((float)e * rn + 1e-3f);
}

// Eight waves per tile, two workgroups per CU (their LDS fits): four waves
// per SIMD.
constexpr int kInferThreads = 512;
constexpr int kInferCells = kTile * kTile * kMaxWidth;  // tile cells, widest layer

// Static per-cell records of every tile (set once per layer): the cell's
// index in its tile's staged box, -1 if it is not a layer cell of the tile's
// face, and sqrt(v2px), sqrt(v2pz) there.
__global__ void __launch_bounds__(256)
    k_layer_static(LayerGeom g, const SlabTile *__restrict__ tiles,
                   int16_t *__restrict__ lidx, float2 *__restrict__ root) {
  const SlabTile T = tiles[blockIdx.x];
  int lo[3], len[3];
  slab_box(T, g, lo, len);
  const int ls[3] = {1, len[0] | 1, (len[0] | 1) * len[1]};
  const int ncell = kTile * kTile * g.width;
  for (int e = threadIdx.x; e < ncell; e += blockDim.x) {
    int c[3];
    int l = -1;
    float2 r = make_float2(0.0f, 0.0f);
    if (slab_cell(T, e, g, c)) {
      l = (c[0] - lo[0]) + (c[1] - lo[1]) * ls[1] + (c[2] - lo[2]) * ls[2];
      const uint32_t og = (uint32_t)dev_offset(g, c[0], c[1], c[2]);
      r = make_float2(sqrtf(g.v2px[og]), sqrtf(g.v2pz[og]));
    }
    lidx[(size_t)blockIdx.x * ncell + e] = (int16_t)l;
    root[(size_t)blockIdx.x * ncell + e] = r;
  }
}

// One workgroup per tile, two per CU: the weights, the tile's box of time
// differences and its cells' static records are staged in LDS, then the
// tile's cells run through the network.
__global__ void __launch_bounds__(kInferThreads)
    k_layer_infer(const float *pc, const float *pp, const float *qc,
                  const float *qp, LayerGeom g, const SlabTile *__restrict__ tiles,
                  int ntiles, const int16_t *__restrict__ lidx,
                  const float2 *__restrict__ root,
                  const _Float16 *__restrict__ wh, const float *__restrict__ w32,
                  float2 *__restrict__ delta, int mode, float sigma_max) {
  // Weight rows of 64 halves padded to 72 (144 B): the 16-byte A-operand
  // reads of 16 rows by 4 lane groups fall on 64 distinct banks.
  constexpr int kRow = 72;
  constexpr int lW1 = 0, lW2 = 64 * kRow, lW3 = 128 * kRow;
  constexpr int kCells = kInferCells;
  // Staged inputs, one float array: D of p and of q over the box, the two
  // Courant numbers of every tile cell, and the one-hot table [d][i].
  constexpr int xQ = kBoxMax, xCx = 2 * kBoxMax, xCz = xCx + kCells,
                xOne = xCz + kCells, kX = xOne + kMaxWidth * kMaxWidth;
  __shared__ __align__(16) _Float16 sW[144 * kRow];
  __shared__ __align__(16) float sB[kHid + kHid + 16];
  __shared__ float sX[kX];
  __shared__ int16_t sL[kCells];  // box index of tile cell e, -1 if none
  const int tid = threadIdx.x;
  // Weights in runs of 4 halves (8-byte loads): column c..c+3 of a row, c a
  // multiple of 4, is one run in W1p and, permuted, one run of W2 and W3
  // (inf_unit maps 4 consecutive columns to 4 consecutive units).
  typedef _Float16 h4u __attribute__((ext_vector_type(4), aligned(8)));
  for (int t = tid; t < 144 * 16; t += kInferThreads) {
    const int row = t >> 4, c = (t & 15) * 4;
    const int u = inf_unit(c >> 5, c & 31);
    const h4u w = row < 64 ? *reinterpret_cast<const h4u *>(wh + hW1p + row * 64 + c)
                : row < 128 ? *reinterpret_cast<const h4u *>(wh + hW2 + (row - 64) * 64 + u)
                            : *reinterpret_cast<const h4u *>(wh + hW3 + (row - 128) * 64 + u);
    *reinterpret_cast<h4u *>(&sW[row * kRow + c]) = w;
  }
  if (tid < kHid) {
    sB[tid] = w32[oB1 + tid];
    sB[kHid + tid] = w32[oB2 + tid];
  }
  if (tid < 16) sB[2 * kHid + tid] = tid < kOut ? w32[oB3 + tid] : 0.0f;
  if (tid < kMaxWidth * kMaxWidth)
    sX[xOne + tid] = (tid >> 3) == (tid & 7) ? 1.0f : 0.0f;

  // Workgroups are dealt to the 8 XCDs round-robin; giving each XCD a
  // contiguous run of tiles keeps the box lines that neighbouring tiles share
  // in one L2.
  int t = blockIdx.x;
  if (ntiles % 8 == 0) t = (t % 8) * (ntiles / 8) + t / 8;
  if (t >= ntiles) return;
  const SlabTile T = tiles[t];
  const int ncell = kTile * kTile * g.width;
  const int n[3] = {g.nx, g.ny, g.nz};
  int lo[3], len[3];
  slab_box(T, g, lo, len);
  const int ls[3] = {1, len[0] | 1, (len[0] | 1) * len[1]};  // LDS strides
  const int nbox = len[0] * len[1] * len[2];
  // Every load of the tile (each box cell's time differences, the static
  // records of the thread's tile cells) is issued before the first LDS
  // write; the other workgroup on the CU computes meanwhile.
  constexpr int kPer = (kBoxMax + kInferThreads - 1) / kInferThreads;
  constexpr int kCPer = kCells / kInferThreads;
  static_assert(kBoxMax < 4096 && (kTile + 2) * (kTile + 2) <= 400, "small_div range");
  const float r0 = 1.0f / (float)len[0], r01 = 1.0f / (float)(len[0] * len[1]);
  float bp[kPer], bq[kPer];
  int bl[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int e = tid + k * kInferThreads;
    bl[k] = -1;
    if (e < nbox) {
      const int bz = small_div(e, r01), r = e - bz * len[0] * len[1];
      const int by = small_div(r, r0), bx = r - by * len[0];
      const int cx = min(max(lo[0] + bx, -STENCIL), n[0] + STENCIL - 1);
      const int cy = min(max(lo[1] + by, -STENCIL), n[1] + STENCIL - 1);
      const int cz = min(max(lo[2] + bz, -STENCIL), n[2] + STENCIL - 1);
      const uint32_t o = (uint32_t)dev_offset(g, cx, cy, cz);
      bl[k] = bx + by * ls[1] + bz * ls[2];
      bp[k] = pc[o] - pp[o];
      bq[k] = qc[o] - qp[o];
    }
  }
  int16_t cl[kCPer];
  float2 cr[kCPer];
#pragma unroll
  for (int k = 0; k < kCPer; ++k) {
    const int e = tid + k * kInferThreads;
    cl[k] = -1;
    cr[k] = make_float2(0.0f, 0.0f);
    if (e < ncell) {
      cl[k] = lidx[(size_t)t * ncell + e];
      cr[k] = root[(size_t)t * ncell + e];
    }
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k)
    if (bl[k] >= 0) {
      sX[bl[k]] = bp[k];
      sX[xQ + bl[k]] = bq[k];
    }
#pragma unroll
  for (int k = 0; k < kCPer; ++k) {
    const int e = tid + k * kInferThreads;
    if (e < ncell) {
      sL[e] = cl[k];
      sX[xCx + e] = cr[k].x * g.cdt;
      sX[xCz + e] = cr[k].y * g.cdt;
    }
  }
  __syncthreads();

  const int lane = tid & 63, wave = tid >> 6;
  const int col = lane & 15, kb = lane >> 4;
  const bool kb3 = kb == 3;
  // Per-lane input offsets: slot 32s + 8kb + j.  Patch slots: relative to
  // the cell's box index, in the tile's frame (outward normal, tangent 1,
  // tangent 2).  Lanes kb = 3, j >= 3 read the one-hot table or the Courant
  // numbers instead (below).
  const int axis = T.face >> 1, side = T.face & 1;
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  const int sn = side ? ls[axis] : -ls[axis], s1 = ls[a1], s2 = ls[a2];
  int off[2][8];  // bytes
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int dd = slot_desc(32 * s + 8 * kb + j);
      off[s][j] = (dd & 1) ? 4 * ((s ? xQ : 0) + (((dd >> 1) & 3) - 1) * sn +
                                  (((dd >> 3) & 3) - 1) * s1 +
                                  (((dd >> 5) & 3) - 1) * s2)
                           : 0;
    }
  const char *const xb = reinterpret_cast<const char *>(sX);
  // A box cell whose 3x3x3 neighbourhood lies in the box: the stand-in for
  // lanes of cells outside the layer (their results are not stored).
  const int lsafe = ls[0] + ls[1] + ls[2];
  // A wave runs two groups of 16 cells at a time (rows 0-15 and 16-31 of a
  // 32-cell slice of one depth level): every weight operand read from LDS
  // feeds two MFMAs, and the two chains interleave.  Every wave takes the
  // same number of slices (ncell is a multiple of 256).
  constexpr int G = 2;
  for (int g0 = wave * 16 * G; g0 < ncell; g0 += kInferThreads / 4 * G) {
    // An offset the compiler cannot see is zero: it keeps the LDS operand
    // reads inside the loop instead of holding all 18 in registers.
    int z = 0;
    asm volatile("" : "+s"(z));
    const int d8 = (g0 >> 8) * kMaxWidth;  // depth of the slice's cells, x 8
    int e[G], l[G];
    bool valid[G];
    float m[G];
    h8 X[G][2];
#pragma unroll
    for (int q = 0; q < G; ++q) {
      e[q] = g0 + 16 * q + col;
      const int lraw = sL[e[q]];
      valid[q] = lraw >= 0;
      l[q] = valid[q] ? lraw : lsafe;
      const int lb = 4 * l[q];
      float v[2][8];
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          int ad = lb + off[s][j];
          if (j >= 3) {
            const int alt = 4 * (s == 0 ? xOne + d8 + (j - 3)
                               : j < 6  ? xOne + d8 + (j + 2)
                               : (j == 6 ? xCx : xCz) + e[q]);
            ad = kb3 ? alt : ad;
          }
          v[s][j] = *reinterpret_cast<const float *>(xb + ad);
        }
      float mm = 0.0f;
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int j = 0; j < 8; ++j)
          mm = fmaxf(mm, (j >= 3 && kb3) ? 0.0f : fabsf(v[s][j]));
      mm = fmaxf(mm, __shfl_xor(mm, 16));
      mm = fmaxf(mm, __shfl_xor(mm, 32));
      m[q] = mm;
      // One reciprocal per cell instead of 54 divisions; the inputs are
      // rounded to half right after, which hides the last-bit difference.
      const float rm = mm > 0.0f ? __builtin_amdgcn_rcpf(mm) : 0.0f;
      const float rm3 = kb3 ? 1.0f : rm;  // slots j >= 3 of lanes kb = 3 as read
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int j = 0; j < 8; ++j)
          X[q][s][j] = (_Float16)(v[s][j] * (j < 3 ? rm : rm3));
    }
    h4 H1[G][4], H2[G][4];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
      const f4 bias = *reinterpret_cast<const f4 *>(&sB[z + 16 * ot + 4 * kb]);
      f4 acc[G];
#pragma unroll
      for (int q = 0; q < G; ++q) acc[q] = bias;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const h8 w = *reinterpret_cast<const h8 *>(
            &sW[z + lW1 + (16 * ot + col) * kRow + 32 * s + 8 * kb]);
#pragma unroll
        for (int q = 0; q < G; ++q)
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w, X[q][s], acc[q], 0, 0, 0);
      }
#pragma unroll
      for (int q = 0; q < G; ++q) H1[q][ot] = relu_pk(acc[q]);
    }
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
      const f4 bias = *reinterpret_cast<const f4 *>(&sB[z + kHid + 16 * ot + 4 * kb]);
      f4 acc[G];
#pragma unroll
      for (int q = 0; q < G; ++q) acc[q] = bias;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const h8 w = *reinterpret_cast<const h8 *>(
            &sW[z + lW2 + (16 * ot + col) * kRow + 32 * s + 8 * kb]);
#pragma unroll
        for (int q = 0; q < G; ++q)
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
              w, __builtin_shufflevector(H1[q][2 * s], H1[q][2 * s + 1], 0, 1, 2, 3, 4, 5, 6, 7),
              acc[q], 0, 0, 0);
      }
#pragma unroll
      for (int q = 0; q < G; ++q) H2[q][ot] = relu_pk(acc[q]);
    }
    f4 y[G];
    {
      const f4 bias = *reinterpret_cast<const f4 *>(&sB[z + 2 * kHid + 4 * kb]);
#pragma unroll
      for (int q = 0; q < G; ++q) y[q] = bias;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const h8 w = *reinterpret_cast<const h8 *>(&sW[z + lW3 + col * kRow + 32 * s + 8 * kb]);
#pragma unroll
        for (int q = 0; q < G; ++q)
          y[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
              w, __builtin_shufflevector(H2[q][2 * s], H2[q][2 * s + 1], 0, 1, 2, 3, 4, 5, 6, 7),
              y[q], 0, 0, 0);
      }
    }
    if (lane < 16) {
#pragma unroll
      for (int q = 0; q < G; ++q) {
        if (!valid[q]) continue;
        float2 out;
        if (mode == kOutDamping) {
          const float sp = sigma_max * __builtin_amdgcn_rcpf(1.0f + __expf(-y[q][0]));
          const float sq = sigma_max * __builtin_amdgcn_rcpf(1.0f + __expf(-y[q][1]));
          out = make_float2(-sp * sX[l[q]], -sq * sX[xQ + l[q]]);
        } else {
          out = make_float2(m[q] * y[q][0], m[q] * y[q][1]);
        }
        delta[(size_t)t * ncell + e[q]] = out;
      }
    }
  }
}

// Add the corrections to the new fields the stencil wrote.
__global__ void __launch_bounds__(256)
    k_layer_apply(float *pn, float *qn, LayerGeom g,
                  const SlabTile *__restrict__ tiles,
                  const float2 *__restrict__ delta) {
  const SlabTile T = tiles[blockIdx.x];
  const int w = g.width;
  const int ncell = kTile * kTile * w;
  const bool xface = (T.face >> 1) == 0;
  for (int i = threadIdx.x; i < ncell; i += blockDim.x) {
    // x faces: depth (= x) fastest, so that lanes write along memory rows.
    const int e = xface ? (i % w) * kTile * kTile + i / w : i;
    int c[3];
    if (!slab_cell(T, e, g, c)) continue;
    const int64_t o = dev_offset(g, c[0], c[1], c[2]);
    const float2 d = delta[(size_t)blockIdx.x * ncell + e];
    pn[o] += d.x;
    qn[o] += d.y;
  }
}

// MLP output y (unscaled) of stored samples (tests).
__global__ void __launch_bounds__(256)
    k_mlp_forward_samples(const _Float16 *__restrict__ x, int64_t n,
                          const _Float16 *__restrict__ wh,
                          const float *__restrict__ w32, float *__restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int nwave = gridDim.x * (blockDim.x / 64);
  FwdWeights w;
  load_fwd(w, wh, w32, lane);
  const int col = lane & 15, kb = lane >> 4;
  for (int64_t g0 = (int64_t)wave * 16; g0 < n; g0 += (int64_t)nwave * 16) {
    const int64_t c = g0 + col;
    const bool valid = c < n;
    h4 X[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      X[ks] = valid ? *(const h4 *)(x + c * 64 + 16 * ks + 4 * kb) : h4{0, 0, 0, 0};
    const f4 out = mlp_forward(w, X);
    if (lane < 16 && valid) {
      y[2 * c] = out[0];
      y[2 * c + 1] = out[1];
    }
  }
}

// ------------------------------------------------------------ MLP training

__device__ inline uint32_t hash32(uint32_t x) {
  // A 32-bit integer mix (xorshift-multiply, two rounds).
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Sample id j of minibatch `it`: uniform with replacement from [0, n), or the
// identity (j) when seed < 0.
__device__ inline int64_t sample_id(int64_t j, int it, int64_t seed, int64_t n) {
  if (seed < 0) return j;
  const uint32_t h1 = hash32((uint32_t)seed * 0x9e3779b9u ^ hash32((uint32_t)it));
  const uint32_t a = hash32(h1 ^ (uint32_t)j), b = hash32(a ^ 0x68e31da4u);
  const uint64_t r = ((uint64_t)a << 32) | b;
  return (int64_t)(r % (uint64_t)n);
}

// One minibatch: forward, backward and the parameter gradient summed into
// grad (scaled by 1/batch, the gradient of the mean over samples of
// (y0 - t0)^2 + (y1 - t1)^2), and the summed squared error into *loss.
// Every workgroup takes groups of 16 samples, one per wave, in lockstep (so
// that barriers are uniform).
constexpr int kTrainWaves = 4;
__global__ void __launch_bounds__(64 * kTrainWaves)
    k_mlp_train(const _Float16 *__restrict__ xs, const float *__restrict__ ts,
                int64_t n, int batch, int it, int64_t seed,
                const _Float16 *__restrict__ wh, const float *__restrict__ w32,
                float *__restrict__ grad, float *__restrict__ loss, int mode,
                float sigma_max) {
  __shared__ float gsh[kNParam];
  __shared__ _Float16 tile[kTrainWaves][5][64 * 16];  // X, H1, H2, dZ1, dZ2
  __shared__ _Float16 tdy[kTrainWaves][16 * 16];
  __shared__ float lsh[kTrainWaves];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int col = lane & 15, kb = lane >> 4;
  for (int p = tid; p < kNParam; p += blockDim.x) gsh[p] = 0.0f;
  float lacc = 0.0f;
  __syncthreads();
  const int ngroups = (batch + 15) / 16;
  FwdWeights w;
  load_fwd(w, wh, w32, lane);
  _Float16 *tX = tile[wv][0], *tH1 = tile[wv][1], *tH2 = tile[wv][2],
           *tZ1 = tile[wv][3], *tZ2 = tile[wv][4], *tY = tdy[wv];
  for (int g0 = blockIdx.x * kTrainWaves; g0 < ngroups; g0 += gridDim.x * kTrainWaves) {
    const int grp = g0 + wv;
    const int j = grp * 16 + col;
    const bool valid = grp < ngroups && j < batch;
    const int64_t sid = valid ? sample_id(j, it, seed, n) : 0;
    h4 X[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      X[ks] = valid ? *(const h4 *)(xs + sid * 64 + 16 * ks + 4 * kb) : h4{0, 0, 0, 0};
    // Forward, keeping pre-activations for the ReLU masks.
    f4 Z1[4], Z2[4];
    h4 H1[4], H2[4];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
      f4 acc = w.b1[ot];
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
        acc = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a1[ot][ks], X[ks], acc, 0, 0, 0);
      Z1[ot] = acc;
      H1[ot] = relu_h(acc);
    }
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
      f4 acc = w.b2[ot];
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
        acc = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a2[ot][ks], H1[ks], acc, 0, 0, 0);
      Z2[ot] = acc;
      H2[ot] = relu_h(acc);
    }
    f4 Y = w.b3;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      Y = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a3[ks], H2[ks], Y, 0, 0, 0);
    // dL/dy per sample (the 1/batch is applied to the sums at the end).
    // Free form: prediction y.  Damping: prediction -sigma_max sigmoid(y) xc
    // with xc the cell's own normalised time difference (input 13 for p,
    // 40 for q).
    h4 dY;
    float dy32[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 4 * kb + r;
      float d = 0.0f;
      if (valid && row < kOut) {
        if (mode == kOutDamping) {
          const float xc = (float)xs[sid * 64 + (row == 0 ? 13 : 40)];
          const float sg = 1.0f / (1.0f + expf(-Y[r]));
          const float e = -sigma_max * sg * xc - ts[sid * 2 + row];
          lacc += e * e;
          d = 2.0f * e * (-sigma_max * xc) * sg * (1.0f - sg);
        } else {
          const float e = Y[r] - ts[sid * 2 + row];
          lacc += e * e;
          d = 2.0f * e;
        }
      }
      dy32[r] = d;
      dY[r] = (_Float16)d;
    }
    // dH2 = W3^T dY, dZ2 = dH2 * [Z2 > 0].
    h4 dZ2[4];
    f4 dZ2f[4];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
      const h4 a = *(const h4 *)(wh + hW3t + (16 * ot + col) * 16 + 4 * kb);
      f4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, dY, acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        dZ2f[ot][r] = Z2[ot][r] > 0.0f ? acc[r] : 0.0f;
        dZ2[ot][r] = (_Float16)dZ2f[ot][r];
      }
    }
    // dH1 = W2^T dZ2, dZ1 = dH1 * [Z1 > 0].
    h4 dZ1[4];
    f4 dZ1f[4];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const h4 a = *(const h4 *)(wh + hW2t + (16 * ot + col) * 64 + 16 * ks + 4 * kb);
        acc = __builtin_amdgcn_mfma_f32_16x16x16f16(a, dZ2[ks], acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        dZ1f[ot][r] = Z1[ot][r] > 0.0f ? acc[r] : 0.0f;
        dZ1[ot][r] = (_Float16)dZ1f[ot][r];
      }
    }
    // Bias gradients: sums over the samples of the FP32 values.
#pragma unroll
    for (int ot = 0; ot < 4; ++ot)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        atomicAdd(&gsh[oB1 + 16 * ot + 4 * kb + r], dZ1f[ot][r]);
        atomicAdd(&gsh[oB2 + 16 * ot + 4 * kb + r], dZ2f[ot][r]);
      }
    if (kb == 0) {
      atomicAdd(&gsh[oB3 + 0], dy32[0]);
      atomicAdd(&gsh[oB3 + 1], dy32[1]);
    }
    // Tiles to LDS as [unit][sample] for the transposed reads.
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int u = 16 * q + 4 * kb + r;
        tX[u * 16 + col] = X[q][r];
        tH1[u * 16 + col] = H1[q][r];
        tH2[u * 16 + col] = H2[q][r];
        tZ1[u * 16 + col] = dZ1[q][r];
        tZ2[u * 16 + col] = dZ2[q][r];
      }
#pragma unroll
    for (int r = 0; r < 4; ++r) tY[(4 * kb + r) * 16 + col] = dY[r];
    __syncthreads();
    // Weight gradients: dW = dZ H^T, K = the 16 samples.  A[i][k] = dZ[unit
    // 16*ot + i][sample k], B[k][j] = H[unit 16*it + j][sample k].
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
      const h4 a1 = *(const h4 *)(tZ1 + (16 * ot + col) * 16 + 4 * kb);
      const h4 a2 = *(const h4 *)(tZ2 + (16 * ot + col) * 16 + 4 * kb);
#pragma unroll
      for (int it2 = 0; it2 < 4; ++it2) {
        const h4 bx = *(const h4 *)(tX + (16 * it2 + col) * 16 + 4 * kb);
        const h4 bh = *(const h4 *)(tH1 + (16 * it2 + col) * 16 + 4 * kb);
        f4 z = {0.f, 0.f, 0.f, 0.f};
        const f4 g1 = __builtin_amdgcn_mfma_f32_16x16x16f16(a1, bx, z, 0, 0, 0);
        const f4 g2 = __builtin_amdgcn_mfma_f32_16x16x16f16(a2, bh, z, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int o = 16 * ot + 4 * kb + r, i = 16 * it2 + col;
          atomicAdd(&gsh[oW1 + o * 64 + i], g1[r]);
          atomicAdd(&gsh[oW2 + o * 64 + i], g2[r]);
        }
      }
    }
    {
      const h4 a3 = *(const h4 *)(tY + col * 16 + 4 * kb);
#pragma unroll
      for (int it2 = 0; it2 < 4; ++it2) {
        const h4 bh = *(const h4 *)(tH2 + (16 * it2 + col) * 16 + 4 * kb);
        f4 z = {0.f, 0.f, 0.f, 0.f};
        const f4 g3 = __builtin_amdgcn_mfma_f32_16x16x16f16(a3, bh, z, 0, 0, 0);
        if (kb == 0) {
          atomicAdd(&gsh[oW3 + 0 * 64 + 16 * it2 + col], g3[0]);
          atomicAdd(&gsh[oW3 + 1 * 64 + 16 * it2 + col], g3[1]);
        }
      }
    }
    __syncthreads();
  }
  // Wave sums of the loss, then one global add per parameter.
  for (int off = 32; off > 0; off >>= 1) lacc += __shfl_xor(lacc, off);
  if (lane == 0) lsh[wv] = lacc;
  __syncthreads();
  const float inv = 1.0f / (float)batch;
  for (int p = tid; p < kNParam; p += blockDim.x) atomicAdd(&grad[p], gsh[p] * inv);
  if (tid == 0) {
    float l = 0.0f;
    for (int q = 0; q < kTrainWaves; ++q) l += lsh[q];
    atomicAdd(loss, l);
  }
}

// Adam step (update != 0) and/or repacking of the fp16 operand copies; the
// gradient is cleared for the next minibatch.
__global__ void k_mlp_update(float *w32, float *m, float *v, float *grad,
                             _Float16 *wh, int update, float lr, float b1,
                             float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  float wv = w32[p];
  if (update) {
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    wv -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    w32[p] = wv;
  }
  grad[p] = 0.0f;
  const _Float16 h = (_Float16)wv;
  if (p < oB1) {
    const int o = p / 64, i = p % 64;
    wh[hW1 + o * 64 + i] = h;
    wh[hW1p + o * 64 + feature_slot(i)] = h;
  } else if (p >= oW2 && p < oB2) {
    const int q = p - oW2, o = q / 64, i = q % 64;
    wh[hW2 + o * 64 + i] = h;
    wh[hW2t + i * 64 + o] = h;
  } else if (p >= oW3 && p < oB3) {
    const int q = p - oW3, o = q / 64, i = q % 64;
    wh[hW3 + o * 64 + i] = h;
    wh[hW3t + i * 16 + o] = h;
  }
}

// ---------------------------------------------------------------- sampling

struct Fields4 {
  float *f[4];
};
struct CFields4 {
  const float *f[4];
};

// dst compute region <- the box of src starting at src offset `org`.
__global__ void k_window(Fields4 dst, int dys, int64_t dzs, CFields4 src,
                         int sys, int64_t szs, int64_t org, int nx, int ny,
                         int nz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)nx * ny * nz;
  if (i >= total) return;
  const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
  const int64_t od = (x + STENCIL) + (int64_t)(y + STENCIL) * dys + (int64_t)(z + STENCIL) * dzs;
  const int64_t os = org + x + (int64_t)y * sys + (int64_t)z * szs;
#pragma unroll
  for (int k = 0; k < 4; ++k) dst.f[k][od] = src.f[k][os];
}

// Features of n random layer cells into samples [base, base + n).
__global__ void k_sample_features(const float *pc, const float *pp,
                                  const float *qc, const float *qp, LayerGeom g,
                                  const uint32_t *cells, int ncell, int n,
                                  uint32_t seed, Samples s, int64_t base,
                                  int pys, int64_t pzs, int64_t org) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = hash32(seed * 0x9e3779b9u ^ hash32((uint32_t)i));
  const uint32_t code = cells[r % (uint32_t)ncell];
  const CellFrame f = cell_frame(code, g);
  float raw[kIn];
  float m = 0.0f;
#pragma unroll
  for (int k = 0; k < kIn; ++k) {
    raw[k] = raw_feature(k, f, g, pc, pp, qc, qp);
    if (k < kPatch) m = fmaxf(m, fabsf(raw[k]));
  }
  const int64_t j = base + i;
#pragma unroll
  for (int k = 0; k < kIn; ++k) s.x[j * 64 + k] = (_Float16)norm_feature(k, raw[k], m);
  s.scale[j] = m;
  s.cell[j] = f.own;
  s.code[j] = code;
  const int cx = (int)(code & 1023u), cy = (int)((code >> 10) & 1023u),
            cz = (int)((code >> 20) & 1023u);
  s.src[j] = org + cx + (int64_t)cy * pys + (int64_t)cz * pzs;
}

// Taper of the target state at layer depth d: sin^2(pi (d + 1) / (2 (w + 1))),
// rising from near 0 at the edge to near 1 at the layer's inner side.
__device__ inline float layer_taper(int d, int w) {
  const float s = sinpif(0.5f * (float)(d + 1) / (float)(w + 1));
  return s * s;
}

// phase 0: t = -(small's new p, q); phase 1: t = (a * padded new p, q + t) / s
// with a = 1, or the taper at the cell's depth when tapered.
__global__ void k_sample_targets(int phase, const float *pn, const float *qn,
                                 Samples s, int64_t base, int n, LayerGeom g,
                                 int tapered) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t j = base + i;
  if (phase == 0) {
    s.t[2 * j] = -pn[s.cell[j]];
    s.t[2 * j + 1] = -qn[s.cell[j]];
  } else {
    const float sc = s.scale[j];
    const float w = tapered ? layer_taper(cell_frame(s.code[j], g).depth, g.width) : 1.0f;
    const float a = w * pn[s.src[j]] + s.t[2 * j], b = w * qn[s.src[j]] + s.t[2 * j + 1];
    s.t[2 * j] = sc > 0.0f ? a / sc : 0.0f;
    s.t[2 * j + 1] = sc > 0.0f ? b / sc : 0.0f;
  }
}

// Multiply the four fields of every layer cell by the taper at its depth.
__global__ void k_taper_layer(Fields4 f, LayerGeom g, const uint32_t *cells,
                              int ncell) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const CellFrame fr = cell_frame(cells[c], g);
  const float w = layer_taper(fr.depth, g.width);
#pragma unroll
  for (int k = 0; k < 4; ++k) f.f[k][fr.own] *= w;
}

__global__ void k_f32_to_f16(const float *a, _Float16 *b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (_Float16)a[i];
}

// ------------------------------------------------------------ host helpers

// cdt = dt * sqrt(1/dx^2), with 1/dx^2 formed in float as the stencil forms it.
static LayerGeom geom_of(const State *st, int width, float cdt = 0.0f) {
  LayerGeom g;
  g.v2px = st->abi.precomp_vars.v2px;
  g.v2pz = st->abi.precomp_vars.v2pz;
  g.cdt = cdt;
  g.nx = (int)st->sizes[0] - 2 * STENCIL;
  g.ny = (int)st->sizes[1] - 2 * STENCIL;
  g.nz = (int)st->sizes[2] - 2 * STENCIL;
  g.width = width;
  g.ys = (int)st->ys;
  g.zs = st->zs;
  return g;
}

// Layer cells of a struct's compute region, x fastest, then y, then z.
static std::vector<uint32_t> layer_cells(const LayerGeom &g) {
  std::vector<uint32_t> v;
  const int w = g.width;
  for (int z = 0; z < g.nz; ++z)
    for (int y = 0; y < g.ny; ++y) {
      const bool edge_zy = z < w || z >= g.nz - w || y < w || y >= g.ny - w;
      for (int x = 0; x < g.nx; ++x)
        if (edge_zy || x < w || x >= g.nx - w) v.push_back(pack_cell(x, y, z));
    }
  return v;
}

static void check_layer_geom(const State *st, int width) {
  if (width < 1 || width > kMaxWidth)
    DCX_FAIL("learned layer width %d outside [1, %d]", width, kMaxWidth);
  for (int a = 0; a < 3; ++a) {
    const int64_t n = st->sizes[a] - 2 * STENCIL;
    if (n < 2 * width + 1 || n > 1024)
      DCX_FAIL("compute extent %lld unsupported by the learned layer",
               (long long)n);
  }
}

static Mlp *mlp_of(dcx_mlp *p) {
  if (p == nullptr) DCX_FAIL("null dcx_mlp");
  return reinterpret_cast<Mlp *>(p);
}

static Samples *samples_of(dcx_samples *p) {
  if (p == nullptr) DCX_FAIL("null dcx_samples");
  return reinterpret_cast<Samples *>(p);
}

static void repack(Mlp *m) {
  hipLaunchKernelGGL(k_mlp_update, dim3((kNParam + 255) / 256), dim3(256), 0,
                     m->stream, m->w32, m->m, m->v, m->grad, m->wh, 0, 0.f,
                     0.f, 0.f, 0.f, 1.f, 1.f);
  DCX_CHECK(hipGetLastError());
}

// Grid of the layer-cell MLP launches: enough waves to cover the cells, at
// most a few per SIMD.
static int mlp_grid(int64_t n) {
  const int64_t waves = (n + 15) / 16;
  const int64_t wgs = (waves + 3) / 4;
  return (int)std::max<int64_t>(1, std::min<int64_t>(wgs, 4096));
}

void layer_infer(State *st, const StepArgs &a) {
  if (st->layer_n == 0) return;
  if (st->fstore != DCX_STORE_F32 || st->cstore != DCX_COEF_F32)
    DCX_FAIL("the learned layer runs on FP32 storage");
  const Mlp *m = reinterpret_cast<const Mlp *>(st->layer_mlp);
  const LayerGeom g = geom_of(st, st->layer_w, a.dt * sqrtf(a.dxx));
  hipLaunchKernelGGL(k_layer_infer, dim3(st->layer_ntiles), dim3(kInferThreads), 0,
                     st->stream, st->abi.pc, st->abi.pp, st->abi.qc, st->abi.qp,
                     g, (const SlabTile *)st->layer_tiles, st->layer_ntiles,
                     st->layer_lidx, st->layer_root, m->wh, m->w32,
                     (float2 *)st->layer_delta, m->out_mode, m->sigma_max);
  DCX_CHECK(hipGetLastError());
}

void layer_apply(State *st) {
  if (st->layer_n == 0) return;
  const LayerGeom g = geom_of(st, st->layer_w);
  hipLaunchKernelGGL(k_layer_apply, dim3(st->layer_ntiles), dim3(256), 0,
                     st->stream, st->abi.pp, st->abi.qp, g,
                     (const SlabTile *)st->layer_tiles,
                     (const float2 *)st->layer_delta);
  DCX_CHECK(hipGetLastError());
}

void mlp_attach(const void *mlp, int delta) {
  if (mlp) const_cast<Mlp *>(static_cast<const Mlp *>(mlp))->attached += delta;
}

void layer_free(State *st) {
  mlp_attach(st->layer_mlp, -1);
  if (st->layer_tiles) DCX_CHECK(hipFree(st->layer_tiles));
  if (st->layer_delta) DCX_CHECK(hipFree(st->layer_delta));
  if (st->layer_lidx) DCX_CHECK(hipFree(st->layer_lidx));
  if (st->layer_root) DCX_CHECK(hipFree(st->layer_root));
  st->layer_lidx = nullptr;
  st->layer_root = nullptr;
  if (st->sample_cells) DCX_CHECK(hipFree(st->sample_cells));
  st->layer_tiles = nullptr;
  st->layer_delta = nullptr;
  st->sample_cells = nullptr;
  st->layer_n = 0;
  st->sample_n = 0;
  st->layer_mlp = nullptr;
}

}  // namespace dcx

using namespace dcx;

extern "C" {

int dcx_mlp_nparam(void) { return kNParam; }

dcx_mlp *dcx_mlp_create(int device) {
  Mlp *m = new Mlp();
  m->device = device;
  m->out_mode = kOutFree;
  m->sigma_max = 0.0f;
  DCX_CHECK(hipSetDevice(device));
  DCX_CHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  DCX_CHECK(hipMalloc(&m->w32, 4 * kNParam * sizeof(float)));
  m->m = m->w32 + kNParam;
  m->v = m->m + kNParam;
  m->grad = m->v + kNParam;
  DCX_CHECK(hipMalloc(&m->loss, 1 << 20));
  DCX_CHECK(hipMalloc(&m->wh, kNHalf * sizeof(_Float16)));
  DCX_CHECK(hipMemsetAsync(m->w32, 0, 4 * kNParam * sizeof(float), m->stream));
  DCX_CHECK(hipMemsetAsync(m->wh, 0, kNHalf * sizeof(_Float16), m->stream));
  DCX_CHECK(hipStreamSynchronize(m->stream));
  m->t = 0;
  return reinterpret_cast<dcx_mlp *>(m);
}

int dcx_mlp_set_output(dcx_mlp *p, int mode, float sigma_max) {
  Mlp *m = mlp_of(p);
  if (mode == kOutFree) {
    m->out_mode = mode;
    m->sigma_max = 0.0f;
    return 0;
  }
  if (mode == kOutDamping && sigma_max > 0.0f && sigma_max < 1.0f) {
    m->out_mode = mode;
    m->sigma_max = sigma_max;
    return 0;
  }
  return -1;
}

void dcx_mlp_free(dcx_mlp *p) {
  if (p == nullptr) return;
  Mlp *m = mlp_of(p);
  if (m->attached)
    DCX_FAIL("network still attached to %d struct(s): detach it first", m->attached);
  DCX_CHECK(hipStreamSynchronize(m->stream));
  DCX_CHECK(hipFree(m->w32));
  DCX_CHECK(hipFree(m->loss));
  DCX_CHECK(hipFree(m->wh));
  DCX_CHECK(hipStreamDestroy(m->stream));
  delete m;
}

void dcx_mlp_set_params(dcx_mlp *p, const float *params) {
  Mlp *m = mlp_of(p);
  DCX_CHECK(hipMemcpy(m->w32, params, kNParam * sizeof(float), hipMemcpyHostToDevice));
  // Fresh optimiser state.
  DCX_CHECK(hipMemsetAsync(m->m, 0, 3 * kNParam * sizeof(float), m->stream));
  m->t = 0;
  repack(m);
  DCX_CHECK(hipStreamSynchronize(m->stream));
}

void dcx_mlp_get_params(const dcx_mlp *p, float *params) {
  Mlp *m = mlp_of(const_cast<dcx_mlp *>(p));
  DCX_CHECK(hipStreamSynchronize(m->stream));
  DCX_CHECK(hipMemcpy(params, m->w32, kNParam * sizeof(float), hipMemcpyDeviceToHost));
}

dcx_samples *dcx_samples_create(int device, long long capacity) {
  if (capacity <= 0) DCX_FAIL("sample capacity must be positive");
  DCX_CHECK(hipSetDevice(device));
  Samples *s = new Samples();
  s->device = device;
  s->cap = capacity;
  s->n = 0;
  DCX_CHECK(hipMalloc(&s->x, capacity * 64 * sizeof(_Float16)));
  DCX_CHECK(hipMalloc(&s->t, capacity * 2 * sizeof(float)));
  DCX_CHECK(hipMalloc(&s->scale, capacity * sizeof(float)));
  DCX_CHECK(hipMalloc(&s->cell, capacity * sizeof(int64_t)));
  DCX_CHECK(hipMalloc(&s->src, capacity * sizeof(int64_t)));
  DCX_CHECK(hipMalloc(&s->code, capacity * sizeof(uint32_t)));
  return reinterpret_cast<dcx_samples *>(s);
}

void dcx_samples_free(dcx_samples *p) {
  if (p == nullptr) return;
  Samples *s = samples_of(p);
  DCX_CHECK(hipDeviceSynchronize());
  DCX_CHECK(hipFree(s->x));
  DCX_CHECK(hipFree(s->t));
  DCX_CHECK(hipFree(s->scale));
  DCX_CHECK(hipFree(s->cell));
  DCX_CHECK(hipFree(s->src));
  DCX_CHECK(hipFree(s->code));
  delete s;
}

long long dcx_samples_count(const dcx_samples *p) {
  return samples_of(const_cast<dcx_samples *>(p))->n;
}

void dcx_samples_clear(dcx_samples *p) { samples_of(p)->n = 0; }

void dcx_samples_upload(dcx_samples *p, long long n, const float *x,
                        const float *t) {
  Samples *s = samples_of(p);
  if (n < 0 || s->n + n > s->cap) DCX_FAIL("samples capacity exceeded");
  float *tmp = nullptr;
  DCX_CHECK(hipMalloc(&tmp, n * 64 * sizeof(float)));
  DCX_CHECK(hipMemcpy(tmp, x, n * 64 * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_f32_to_f16, dim3((unsigned)((n * 64 + 255) / 256)), dim3(256), 0,
                     0, tmp, s->x + s->n * 64, n * 64);
  DCX_CHECK(hipGetLastError());
  DCX_CHECK(hipMemcpy(s->t + 2 * s->n, t, n * 2 * sizeof(float), hipMemcpyHostToDevice));
  std::vector<float> ones(n, 1.0f);
  DCX_CHECK(hipMemcpy(s->scale + s->n, ones.data(), n * sizeof(float), hipMemcpyHostToDevice));
  DCX_CHECK(hipMemset(s->code + s->n, 0, n * sizeof(uint32_t)));
  DCX_CHECK(hipDeviceSynchronize());
  DCX_CHECK(hipFree(tmp));
  s->n += n;
}

void dcx_samples_download(const dcx_samples *p, long long first, long long n,
                          float *x, float *t, float *scale,
                          unsigned int *codes) {
  Samples *s = samples_of(const_cast<dcx_samples *>(p));
  if (first < 0 || n < 0 || first + n > s->n) DCX_FAIL("sample range outside [0, count)");
  DCX_CHECK(hipDeviceSynchronize());
  if (x) {
    std::vector<_Float16> h(n * 64);
    DCX_CHECK(hipMemcpy(h.data(), s->x + first * 64, n * 64 * sizeof(_Float16),
                        hipMemcpyDeviceToHost));
    for (long long i = 0; i < n * 64; ++i) x[i] = (float)h[i];
  }
  if (t)
    DCX_CHECK(hipMemcpy(t, s->t + 2 * first, n * 2 * sizeof(float), hipMemcpyDeviceToHost));
  if (scale)
    DCX_CHECK(hipMemcpy(scale, s->scale + first, n * sizeof(float), hipMemcpyDeviceToHost));
  if (codes)
    DCX_CHECK(hipMemcpy(codes, s->code + first, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
}

void dcx_mlp_forward(dcx_mlp *p, const dcx_samples *sp, float *y) {
  Mlp *m = mlp_of(p);
  Samples *s = samples_of(const_cast<dcx_samples *>(sp));
  if (s->n == 0) return;
  float *dy = nullptr;
  DCX_CHECK(hipMalloc(&dy, s->n * 2 * sizeof(float)));
  hipLaunchKernelGGL(k_mlp_forward_samples, dim3(mlp_grid(s->n)), dim3(256), 0,
                     m->stream, s->x, s->n, m->wh, m->w32, dy);
  DCX_CHECK(hipGetLastError());
  DCX_CHECK(hipStreamSynchronize(m->stream));
  DCX_CHECK(hipMemcpy(y, dy, s->n * 2 * sizeof(float), hipMemcpyDeviceToHost));
  DCX_CHECK(hipFree(dy));
}

static void train_launch(Mlp *m, Samples *s, int batch, int it, long long seed,
                         float *loss_slot) {
  const int groups = (batch + 15) / 16;
  const int grid = std::max(1, std::min(256, (groups + kTrainWaves - 1) / kTrainWaves));
  hipLaunchKernelGGL(k_mlp_train, dim3(grid), dim3(64 * kTrainWaves), 0,
                     m->stream, s->x, s->t, s->n, batch, it, (int64_t)seed,
                     m->wh, m->w32, m->grad, loss_slot, m->out_mode,
                     m->sigma_max);
  DCX_CHECK(hipGetLastError());
}

double dcx_mlp_gradient(dcx_mlp *p, const dcx_samples *sp, float *grad) {
  Mlp *m = mlp_of(p);
  Samples *s = samples_of(const_cast<dcx_samples *>(sp));
  if (s->n == 0 || s->n > INT32_MAX) DCX_FAIL("gradient needs 1..2^31 samples");
  DCX_CHECK(hipMemsetAsync(m->grad, 0, kNParam * sizeof(float), m->stream));
  DCX_CHECK(hipMemsetAsync(m->loss, 0, sizeof(float), m->stream));
  train_launch(m, s, (int)s->n, 0, -1, m->loss);
  DCX_CHECK(hipStreamSynchronize(m->stream));
  float l = 0.f;
  DCX_CHECK(hipMemcpy(grad, m->grad, kNParam * sizeof(float), hipMemcpyDeviceToHost));
  DCX_CHECK(hipMemcpy(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost));
  DCX_CHECK(hipMemsetAsync(m->grad, 0, kNParam * sizeof(float), m->stream));
  DCX_CHECK(hipStreamSynchronize(m->stream));
  return (double)l / (double)s->n;
}

double dcx_mlp_train(dcx_mlp *p, const dcx_samples *sp, int n_iter, int batch,
                     float lr, long long seed, float *loss_per_iter) {
  Mlp *m = mlp_of(p);
  Samples *s = samples_of(const_cast<dcx_samples *>(sp));
  if (s->n == 0 || batch < 16 || n_iter < 1) DCX_FAIL("training needs samples, batch >= 16");
  if (n_iter > (1 << 18)) DCX_FAIL("at most 2^18 iterations per call");
  DCX_CHECK(hipMemsetAsync(m->loss, 0, n_iter * sizeof(float), m->stream));
  DCX_CHECK(hipMemsetAsync(m->grad, 0, kNParam * sizeof(float), m->stream));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int it = 0; it < n_iter; ++it) {
    train_launch(m, s, batch, (int)(m->t & 0x7fffffff), seed < 0 ? 0 : seed, m->loss + it);
    m->t += 1;
    const float c1 = 1.0f - std::pow(b1, (float)m->t);
    const float c2 = 1.0f - std::pow(b2, (float)m->t);
    hipLaunchKernelGGL(k_mlp_update, dim3((kNParam + 255) / 256), dim3(256), 0,
                       m->stream, m->w32, m->m, m->v, m->grad, m->wh, 1, lr, b1,
                       b2, eps, c1, c2);
    DCX_CHECK(hipGetLastError());
  }
  std::vector<float> l(n_iter);
  DCX_CHECK(hipStreamSynchronize(m->stream));
  DCX_CHECK(hipMemcpy(l.data(), m->loss, n_iter * sizeof(float), hipMemcpyDeviceToHost));
  double mean = 0.0;
  for (int it = 0; it < n_iter; ++it) {
    l[it] /= (float)batch;
    mean += l[it];
    if (loss_per_iter) loss_per_iter[it] = l[it];
  }
  return mean / n_iter;
}

int dcx_set_learned_layer(dc_device_data *data, dcx_mlp *p, int width) {
  State *st = state_of(data);
  DCX_CHECK(hipStreamSynchronize(st->stream));
  if (st->layer_tiles) DCX_CHECK(hipFree(st->layer_tiles));
  if (st->layer_delta) DCX_CHECK(hipFree(st->layer_delta));
  if (st->layer_lidx) DCX_CHECK(hipFree(st->layer_lidx));
  if (st->layer_root) DCX_CHECK(hipFree(st->layer_root));
  st->layer_tiles = nullptr;
  st->layer_delta = nullptr;
  st->layer_lidx = nullptr;
  st->layer_root = nullptr;
  st->layer_n = 0;
  st->layer_ntiles = 0;
  mlp_attach(st->layer_mlp, -1);
  st->layer_mlp = nullptr;
  if (p == nullptr || width == 0) return 0;
  if (st->batch) DCX_FAIL("a batch shot cannot carry a learned layer");
  if (st->fstore != DCX_STORE_F32 || st->cstore != DCX_COEF_F32)
    DCX_FAIL("the learned layer runs on FP32 storage");
  check_layer_geom(st, width);
  Mlp *m = mlp_of(p);
  if (m->device != st->device) DCX_FAIL("network and struct on different devices");
  DCX_CHECK(hipStreamSynchronize(m->stream));
  // Tiles of 16 x 16 cells of each face plane.
  const int64_t n[3] = {st->sizes[0] - 2 * STENCIL, st->sizes[1] - 2 * STENCIL,
                        st->sizes[2] - 2 * STENCIL};
  std::vector<SlabTile> tiles;
  for (int f = 0; f < 6; ++f) {
    const int axis = f >> 1;
    const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
    for (int t2 = 0; t2 < n[a2]; t2 += kTile)
      for (int t1 = 0; t1 < n[a1]; t1 += kTile) tiles.push_back(SlabTile{f, t1, t2, 0});
  }
  DCX_CHECK(hipMalloc(&st->layer_tiles, tiles.size() * sizeof(SlabTile)));
  DCX_CHECK(hipMemcpy(st->layer_tiles, tiles.data(), tiles.size() * sizeof(SlabTile),
                      hipMemcpyHostToDevice));
  const size_t ncells = tiles.size() * (size_t)kTile * kTile * width;
  DCX_CHECK(hipMalloc(&st->layer_delta, ncells * 2 * sizeof(float)));
  DCX_CHECK(hipMalloc(&st->layer_lidx, ncells * sizeof(int16_t)));
  DCX_CHECK(hipMalloc(&st->layer_root, ncells * sizeof(float2)));
  st->layer_w = width;
  {
    const LayerGeom g = geom_of(st, width);
    hipLaunchKernelGGL(k_layer_static, dim3((unsigned)tiles.size()), dim3(256), 0,
                       st->stream, g, (const SlabTile *)st->layer_tiles,
                       st->layer_lidx, st->layer_root);
    DCX_CHECK(hipGetLastError());
    DCX_CHECK(hipStreamSynchronize(st->stream));
  }
  int64_t inner = 1;
  for (int a = 0; a < 3; ++a) inner *= n[a] - 2 * width;
  st->layer_n = (int)(n[0] * n[1] * n[2] - inner);
  st->layer_ntiles = (int)tiles.size();
  st->layer_w = width;
  st->layer_mlp = m;
  mlp_attach(m, +1);
  return st->layer_n;
}

}  // extern "C"

namespace dcx {
struct TargetHook {
  State *st;
  Samples *s;
  int64_t base;
  int n;
};

static void targets_phase0(void *p) {
  const TargetHook *h = static_cast<const TargetHook *>(p);
  hipLaunchKernelGGL(k_sample_targets, dim3((h->n + 255) / 256), dim3(256), 0,
                     h->st->stream, 0, h->st->abi.pp, h->st->abi.qp, *h->s,
                     h->base, h->n, LayerGeom{}, 0);
  DCX_CHECK(hipGetLastError());
}
}  // namespace dcx

extern "C" {

long long dcx_learned_sample_step(dc_device_data *padded, dc_device_data *small,
                                  const long long origin[3], int width,
                                  int n_cells, unsigned int seed,
                                  dcx_samples *sp, float dx, float dy, float dz,
                                  float dt, long long src_index, float src_next,
                                  int add_next, int flags) {
  const int rollout = flags & DCX_SAMPLE_ROLLOUT;
  const int tapered = (flags & DCX_SAMPLE_TAPER) ? 1 : 0;
  State *pd = state_of(padded);
  State *sm = state_of(small);
  Samples *s = samples_of(sp);
  if (pd->fstore != DCX_STORE_F32 || sm->fstore != DCX_STORE_F32 ||
      sm->cstore != DCX_COEF_F32)
    DCX_FAIL("sampling runs on FP32 storage");
  if (pd->device != sm->device || s->device != pd->device)
    DCX_FAIL("padded, small and samples on different devices");
  check_layer_geom(sm, width);
  const LayerGeom g = geom_of(sm, width, dt * sqrtf(1.0f / (dx * dx)));
  for (int a = 0; a < 3; ++a) {
    const int64_t n = sm->sizes[a] - 2 * STENCIL;
    if (origin[a] < STENCIL || origin[a] + n > pd->sizes[a] - STENCIL)
      DCX_FAIL("small domain outside the padded compute region");
  }
  if (n_cells < 1 || s->n + n_cells > s->cap) DCX_FAIL("samples capacity exceeded");
  if (sm->sample_cells == nullptr || sm->sample_w != width) {
    if (sm->sample_cells) DCX_CHECK(hipFree(sm->sample_cells));
    const std::vector<uint32_t> cells = layer_cells(g);
    DCX_CHECK(hipMalloc(&sm->sample_cells, cells.size() * sizeof(uint32_t)));
    DCX_CHECK(hipMemcpy(sm->sample_cells, cells.data(), cells.size() * sizeof(uint32_t),
                        hipMemcpyHostToDevice));
    sm->sample_n = (int)cells.size();
    sm->sample_w = width;
  }
  const int64_t org = origin[0] + origin[1] * pd->ys + origin[2] * pd->zs;
  const int64_t base = s->n;
  const dim3 gs((n_cells + 255) / 256), bs(256);
  const int64_t ncomp = (int64_t)g.nx * g.ny * g.nz;
  if (rollout) {
    // small runs on its own state: the inputs are its fields, the target
    // pulls its step back to the padded one.  Its source is padded's, moved
    // into its coordinates (none when outside its compute region).
    long long ssrc = -1;
    if (src_index >= 0) {
      const int64_t plane = pd->sizes[0] * pd->sizes[1];
      const int64_t z = src_index / plane, r = src_index % plane;
      const int64_t y = r / pd->sizes[0], x = r % pd->sizes[0];
      const int64_t lx = x - (origin[0] - STENCIL), ly = y - (origin[1] - STENCIL),
                    lz = z - (origin[2] - STENCIL);
      if (lx >= STENCIL && lx < sm->sizes[0] - STENCIL && ly >= STENCIL &&
          ly < sm->sizes[1] - STENCIL && lz >= STENCIL && lz < sm->sizes[2] - STENCIL)
        ssrc = lx + ly * sm->sizes[0] + lz * sm->sizes[0] * sm->sizes[1];
    }
    hipLaunchKernelGGL(k_sample_features, gs, bs, 0, sm->stream, sm->abi.pc,
                       sm->abi.pp, sm->abi.qc, sm->abi.qp, g, sm->sample_cells,
                       sm->sample_n, n_cells, seed, *s, base, (int)pd->ys, pd->zs, org);
    DCX_CHECK(hipGetLastError());
    TargetHook hook{sm, s, base, n_cells};
    fused_step(sm, dx, dy, dz, dt, ssrc, src_next, add_next, targets_phase0, &hook);
    DCX_CHECK(hipStreamSynchronize(sm->stream));
  } else {
    // Both structs' work goes on the padded one's stream.
    DCX_CHECK(hipStreamSynchronize(sm->stream));
    hipStream_t str = pd->stream;
    // 1. small's compute region <- the padded fields (frame stays zero).
    Fields4 dst{{sm->abi.pp, sm->abi.pc, sm->abi.qp, sm->abi.qc}};
    CFields4 src{{pd->abi.pp, pd->abi.pc, pd->abi.qp, pd->abi.qc}};
    hipLaunchKernelGGL(k_window, dim3((unsigned)((ncomp + 255) / 256)), dim3(256), 0, str,
                       dst, (int)sm->ys, sm->zs, src, (int)pd->ys, pd->zs, org, g.nx,
                       g.ny, g.nz);
    DCX_CHECK(hipGetLastError());
    if (tapered) {
      hipLaunchKernelGGL(k_taper_layer, dim3((sm->sample_n + 255) / 256), dim3(256), 0,
                         str, dst, g, sm->sample_cells, sm->sample_n);
      DCX_CHECK(hipGetLastError());
    }
    // 2. features of random layer cells.
    hipLaunchKernelGGL(k_sample_features, gs, bs, 0, str, sm->abi.pc, sm->abi.pp,
                       sm->abi.qc, sm->abi.qp, g, sm->sample_cells, sm->sample_n,
                       n_cells, seed, *s, base, (int)pd->ys, pd->zs, org);
    DCX_CHECK(hipGetLastError());
    // 3. one step of the small domain (no source, no swap): its new p, q.
    StepArgs a = make_step_args(sm, Box{STENCIL, STENCIL, STENCIL, sm->sizes[0] - STENCIL,
                                        sm->sizes[1] - STENCIL, sm->sizes[2] - STENCIL},
                                dx, dy, dz, dt);
    if (sm->kernel == DCX_KERNEL_EXACT)
      launch_exact(a, str);
    else
      launch_fast(a, sm->tile_x, sm->tile_y, str);
    hipLaunchKernelGGL(k_sample_targets, gs, bs, 0, str, 0, sm->abi.pp, sm->abi.qp,
                       *s, base, n_cells, g, 0);
    DCX_CHECK(hipGetLastError());
  }
  // The padded step (as dcx_step), then the targets.
  dcx_step(padded, dx, dy, dz, dt, src_index, src_next, add_next);
  hipLaunchKernelGGL(k_sample_targets, gs, bs, 0, pd->stream, 1, pd->abi.pc,
                     pd->abi.qc, *s, base, n_cells, g, tapered);
  DCX_CHECK(hipGetLastError());
  DCX_CHECK(hipStreamSynchronize(pd->stream));
  s->n += n_cells;
  return s->n;
}

}  // extern "C"
