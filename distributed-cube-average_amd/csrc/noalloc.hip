// dcx_halo_face_noalloc: the reference ABI's halo-face extract and insert
// (dc_device_extract_halo_face / dc_device_insert_halo_face) for callers that
// cannot allocate device memory.  Those two entry points stage the packed box
// in a per-struct device buffer that they allocate on first use and grow on
// demand; this variant stages it in a fixed array of every device's copy of
// the code object instead and calls no allocating routine.
#include <mutex>

#include "dc_internal.hpp"

namespace dcx {

// The staging array, held by each device the library runs on until the
// process ends.  Host code never passes the array's name to a kernel or a
// copy (on the host it is a stand-in, not a device address): it uses the
// address hipGetSymbolAddress returns for each device, resolved once per
// device by dc_device_data_init through noalloc_init.
constexpr size_t kNoallocFloats = size_t(1) << 22;  // 16 MiB per device
static __device__ __align__(256) float scratch[kNoallocFloats];

constexpr int kMaxDevices = 64;
static float *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

void noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    DCX_FAIL("device ordinal %d beyond the %d the library tracks", device,
             kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device] != nullptr) return;
  void *ptr = nullptr;
  DCX_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (ptr == nullptr) DCX_FAIL("no device address for the staging array");
  g_scratch_ptr[device] = static_cast<float *>(ptr);
}

}  // namespace dcx

using namespace dcx;

extern "C" void dcx_halo_face_noalloc(dc_device_data *data, float *buffer,
                                      const size_t start_coords[DIMENSIONS],
                                      const size_t end_coords[DIMENSIONS],
                                      const size_t sizes[DIMENSIONS],
                                      float *array, int dir) {
  State *st = state_of(data);
  check_sizes(st, sizes);
  if (dir != 0 && dir != 1) DCX_FAIL("dir must be 0 (extract) or 1 (insert)");
  const BoxCopy b = box_of(st, start_coords, end_coords, field_of(st, array));
  const size_t n = (size_t)(b.nx * b.ny * b.nz);
  if (n > kNoallocFloats)
    DCX_FAIL("box of %zu cells exceeds the %zu-cell staging array of "
             "dcx_halo_face_noalloc",
             n, kNoallocFloats);
  if (n == 0) return;
  float *stage = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    stage = g_scratch_ptr[st->device];
  }
  if (stage == nullptr)
    DCX_FAIL("staging array not resolved for device %d", st->device);
  if (dir == 0) {
    launch_box_copy1(b, stage, 0, st->stream);
    DCX_CHECK(hipMemcpyAsync(buffer, stage, n * sizeof(float),
                             hipMemcpyDeviceToHost, st->stream));
  } else {
    DCX_CHECK(hipMemcpyAsync(stage, buffer, n * sizeof(float),
                             hipMemcpyHostToDevice, st->stream));
    launch_box_copy1(b, stage, 1, st->stream);
  }
  // Every call on the device shares the staging array: it is free again
  // once this returns.
  DCX_CHECK(hipStreamSynchronize(st->stream));
}
