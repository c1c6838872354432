// On-device halo packing: strided boxes of a field <-> contiguous buffers.
//
// A box [x0,x0+nx) x [y0,y0+ny) x [z0,z0+nz) is packed x fastest, then y,
// then z -- the element order of the reference's extract/insert loops
// (src/device_data.c:61-92) and of its MPI messages (src/worker.c:242-246,
// 311-313).  One launch handles any number of boxes (the 26 faces, edges and
// corners of a subdomain, for p and q, for every subdomain on the GPU).
// Consecutive lanes take consecutive packed elements, so the packed side is
// fully contiguous and the strided side is contiguous along each box row.
#include "dc_internal.hpp"

namespace dcx {

// The packed stream of all boxes is one index range; each lane finds the box
// of its element by binary search over the box offsets held in LDS.  Box
// volumes are below 2^32, so in-box coordinates are 32-bit.
constexpr int kMaxBoxesLds = 512;

// Unsigned 32-bit division by a divisor fixed per box, as a multiply-high and
// two shifts (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, fig. 4.1): exact for every 32-bit numerator and
// every divisor 1 <= d <= 2^31.
struct InvDiv {
  uint32_t m;
  uint32_t s1, s2;
};

__device__ __forceinline__ InvDiv inv_div(uint32_t d) {
  const uint32_t l = d > 1 ? 32u - (uint32_t)__clz(d - 1) : 0u;  // ceil(log2 d)
  const uint64_t m = ((uint64_t)1 << 32) * (((uint64_t)1 << l) - d) / d + 1;
  return {(uint32_t)m, l < 1 ? l : 1u, l > 1 ? l - 1 : 0u};
}

__device__ __forceinline__ uint32_t div_by(uint32_t n, const InvDiv &v) {
  const uint32_t t = __umulhi(v.m, n);
  return (t + ((n - t) >> v.s1)) >> v.s2;
}

// What a lane needs of its box, staged in LDS by every block.
struct BoxRec {
  float *arr;
  int64_t base;  // array index of the box's first element
  int64_t ys, zs;
  uint32_t nx, ny;
  InvDiv dx, dy;
  int esize;
};

// Element i of the box's array <-> element e of the packed buffer, in the
// box's element size (the buffer uses the same element size).
__device__ __forceinline__ void copy_elem(const BoxCopy &b, int64_t i,
                                          float *buf, int64_t e, int dir) {
  if (b.esize == 2) {
    uint16_t *arr = reinterpret_cast<uint16_t *>(b.arr);
    uint16_t *pk = reinterpret_cast<uint16_t *>(buf);
    if (dir == 0)
      pk[e] = arr[i];
    else
      arr[i] = pk[e];
  } else {
    if (dir == 0)
      buf[e] = b.arr[i];
    else
      b.arr[i] = buf[e];
  }
}

// Dynamic LDS: nbox + 1 offsets, then nbox records.
static size_t box_copy_lds(int nbox) {
  return (size_t)(nbox + 1) * sizeof(int64_t) + (size_t)nbox * sizeof(BoxRec);
}

__global__ void __launch_bounds__(256)
    k_box_copy(const BoxCopy *__restrict__ boxes, int nbox, int64_t total,
               float *__restrict__ buf, int dir) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  int64_t *soff = reinterpret_cast<int64_t *>(lds_raw);
  BoxRec *srec = reinterpret_cast<BoxRec *>(lds_raw + (size_t)(nbox + 1) *
                                                          sizeof(int64_t));
  for (int k = threadIdx.x; k < nbox; k += blockDim.x) {
    const BoxCopy &b = boxes[k];
    soff[k] = b.off;
    BoxRec r;
    r.arr = b.arr;
    r.base = b.x0 + b.y0 * b.ys + b.z0 * b.zs;
    r.ys = b.ys;
    r.zs = b.zs;
    r.nx = (uint32_t)b.nx;
    r.ny = (uint32_t)b.ny;
    r.dx = inv_div(r.nx > 0 ? r.nx : 1);
    r.dy = inv_div(r.ny > 0 ? r.ny : 1);
    r.esize = b.esize;
    srec[k] = r;
  }
  if (threadIdx.x == 0) soff[nbox] = total;
  __syncthreads();
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = nbox - 1;  // last box with off <= e (binary search)
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (soff[mid] <= e)
        lo = mid;
      else
        hi = mid - 1;
    }
    const BoxRec &b = srec[lo];
    const uint32_t k = (uint32_t)(e - soff[lo]);
    const uint32_t r = div_by(k, b.dx);
    const uint32_t x = k - r * b.nx;
    const uint32_t z = div_by(r, b.dy);
    const uint32_t y = r - z * b.ny;
    const int64_t i = b.base + x + (int64_t)y * b.ys + (int64_t)z * b.zs;
    // The array pointer comes from LDS, where the compiler cannot tell its
    // address space: cast it to global, or the access is FLAT and its wait
    // is also a wait for the LDS reads of the next element's search.
    if (b.esize == 2) {
      auto arr = (__attribute__((address_space(1))) uint16_t *)b.arr;
      uint16_t *pk = reinterpret_cast<uint16_t *>(buf);
      if (dir == 0)
        pk[e] = arr[i];
      else
        arr[i] = pk[e];
    } else {
      auto arr = (__attribute__((address_space(1))) float *)b.arr;
      if (dir == 0)
        buf[e] = arr[i];
      else
        arr[i] = buf[e];
    }
  }
}

// One box of any volume (the whole-allocation copy of dcx_copy_field can pass
// 2^32 elements): in-box coordinates in 64-bit arithmetic.
__global__ void __launch_bounds__(256)
    k_box_copy1(const BoxCopy b, float *__restrict__ buf, int dir) {
  const int64_t total = b.nx * b.ny * b.nz;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / b.nx;
    const int64_t x = e - r * b.nx;
    const int64_t z = r / b.ny;
    const int64_t y = r - z * b.ny;
    const int64_t i = (b.x0 + x) + (b.y0 + y) * b.ys + (b.z0 + z) * b.zs;
    copy_elem(b, i, buf, b.off + e, dir);
  }
}

static unsigned copy_blocks(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 8192) g = 8192;
  return (unsigned)(g < 1 ? 1 : g);
}

void check_multi_box(const BoxCopy &b) {
  if (b.nx * b.ny * b.nz >= (int64_t(1) << 32))
    DCX_FAIL("box of %lld x %lld x %lld cells: a multi-box copy takes boxes "
             "below 2^32 cells",
             (long long)b.nx, (long long)b.ny, (long long)b.nz);
}

void launch_box_copy(const BoxCopy *boxes, int nbox, int64_t total, float *buf,
                     int dir, hipStream_t s) {
  if (nbox <= 0 || total <= 0) return;
  if (nbox > kMaxBoxesLds) DCX_FAIL("too many boxes in one copy: %d", nbox);
  hipLaunchKernelGGL(k_box_copy, dim3(copy_blocks(total)), dim3(256),
                     box_copy_lds(nbox), s, boxes, nbox, total, buf, dir);
  DCX_CHECK(hipGetLastError());
}

void launch_box_copy1(const BoxCopy &box, float *buf, int dir, hipStream_t s) {
  const int64_t total = box.nx * box.ny * box.nz;
  if (total <= 0) return;
  hipLaunchKernelGGL(k_box_copy1, dim3(copy_blocks(total)), dim3(256), 0, s,
                     box, buf, dir);
  DCX_CHECK(hipGetLastError());
}

}  // namespace dcx
