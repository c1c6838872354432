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
// volumes are below 2^32, so in-box coordinates use 32-bit division.
constexpr int kMaxBoxesLds = 512;

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

__global__ void __launch_bounds__(256)
    k_box_copy(const BoxCopy *__restrict__ boxes, int nbox, int64_t total,
               float *__restrict__ buf, int dir) {
  __shared__ int64_t soff[kMaxBoxesLds + 1];
  for (int k = threadIdx.x; k < nbox; k += blockDim.x) soff[k] = boxes[k].off;
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
    const BoxCopy &b = boxes[lo];
    const uint32_t k = (uint32_t)(e - b.off);
    const uint32_t nx = (uint32_t)b.nx, ny = (uint32_t)b.ny;
    const uint32_t r = k / nx;
    const int64_t x = k - r * nx;
    const int64_t z = r / ny;
    const int64_t y = r - (uint32_t)z * ny;
    const int64_t i = (b.x0 + x) + (b.y0 + y) * b.ys + (b.z0 + z) * b.zs;
    copy_elem(b, i, buf, e, dir);
  }
}

__global__ void __launch_bounds__(256)
    k_box_copy1(const BoxCopy b, float *__restrict__ buf, int dir) {
  const int64_t total = b.nx * b.ny * b.nz;
  const uint32_t nx = (uint32_t)b.nx, ny = (uint32_t)b.ny;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t k = (uint32_t)e;
    const uint32_t r = k / nx;
    const int64_t x = k - r * nx;
    const int64_t z = r / ny;
    const int64_t y = r - (uint32_t)z * ny;
    const int64_t i = (b.x0 + x) + (b.y0 + y) * b.ys + (b.z0 + z) * b.zs;
    copy_elem(b, i, buf, b.off + e, dir);
  }
}

static unsigned copy_blocks(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > 8192) g = 8192;
  return (unsigned)(g < 1 ? 1 : g);
}

void launch_box_copy(const BoxCopy *boxes, int nbox, int64_t total, float *buf,
                     int dir, hipStream_t s) {
  if (nbox <= 0 || total <= 0) return;
  if (nbox > kMaxBoxesLds) DCX_FAIL("too many boxes in one copy: %d", nbox);
  hipLaunchKernelGGL(k_box_copy, dim3(copy_blocks(total)), dim3(256), 0, s,
                     boxes, nbox, total, buf, dir);
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
