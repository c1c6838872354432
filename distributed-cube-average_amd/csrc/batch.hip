// Batched shots: many independent one-rank volumes (shots of different
// apertures) advanced together, one persistent launch per time step.
//
// Every shot is a dc_device_data of its own (arrays sized to its aperture);
// the batch puts them all on one stream, builds per-shot launch arguments for
// both pointer parities, the prefix sums of their tile counts and a device
// table of each shot's source amplitudes, and runs k_fast_batch: a grid sized
// to the GPU (workgroups per CU x CUs) that loops over the concatenated tile
// lists.  A time loop can be captured into one HIP graph.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "dc_internal.hpp"

namespace dcx {
struct Batch {
  std::vector<State *> shots;
  hipStream_t stream;
  int zchunk = 128;
  int deep = 1;
  int blocks_per_cu = 1;
  int n_cu = 0;
  bool autotune = false;
  double last_ms_per_step = 0.0;
  hipEvent_t ev[2];
};
}  // namespace dcx

using namespace dcx;

struct dcx_batch {
  Batch b;
};

extern "C" {

dcx_batch *dcx_batch_create(dc_device_data **shots, int n) {
  if (n <= 0 || shots == nullptr) DCX_FAIL("empty batch");
  dcx_batch *bb = new dcx_batch();
  Batch &b = bb->b;
  State *s0 = state_of(shots[0]);
  DCX_CHECK(hipSetDevice(s0->device));
  DCX_CHECK(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
  DCX_CHECK(hipEventCreate(&b.ev[0]));
  DCX_CHECK(hipEventCreate(&b.ev[1]));
  hipDeviceProp_t prop;
  DCX_CHECK(hipGetDeviceProperties(&prop, s0->device));
  b.n_cu = prop.multiProcessorCount;
  for (int k = 0; k < n; ++k) {
    State *st = state_of(shots[k]);
    if (st->device != s0->device) DCX_FAIL("batch shots must share a device");
    if (st->fstore != s0->fstore || st->cstore != s0->cstore)
      DCX_FAIL("batch shots must share one storage form");
    if (st->halo_nuser) DCX_FAIL("batch shots cannot carry a halo loopback");
    if (st->batch) DCX_FAIL("shot %d already belongs to a batch", k);
    if (st->layer_n) DCX_FAIL("batch shots cannot carry a learned layer");
    adopt_stream(st, b.stream);
    st->batch = bb;
    b.shots.push_back(st);
  }
  return bb;
}

void dcx_batch_free(dcx_batch *bb) {
  if (bb == nullptr) return;
  Batch &b = bb->b;
  DCX_CHECK(hipStreamSynchronize(b.stream));
  // Give every shot a stream of its own again so it outlives the batch.
  for (State *st : b.shots) {
    DCX_CHECK(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
    st->own_stream = true;
    st->batch = nullptr;
  }
  DCX_CHECK(hipEventDestroy(b.ev[0]));
  DCX_CHECK(hipEventDestroy(b.ev[1]));
  DCX_CHECK(hipStreamDestroy(b.stream));
  delete bb;
}

int dcx_batch_set_tiling(dcx_batch *bb, int z_chunk, int prefetch_planes,
                         int blocks_per_cu) {
  if (bb == nullptr || z_chunk < 1 || prefetch_planes < 1 || prefetch_planes > 3 ||
      blocks_per_cu < 0 || blocks_per_cu > 64)
    return -1;
  bb->b.zchunk = z_chunk;
  bb->b.deep = prefetch_planes - 1;
  bb->b.blocks_per_cu = blocks_per_cu;
  return 0;
}

int dcx_batch_set_autotune(dcx_batch *bb, int enabled) {
  if (bb == nullptr) return -1;
  bb->b.autotune = enabled != 0;
  return 0;
}

void dcx_batch_get_tiling(const dcx_batch *bb, int out[3]) {
  if (bb == nullptr) DCX_FAIL("null batch");
  out[0] = bb->b.zchunk;
  out[1] = bb->b.deep + 1;
  out[2] = bb->b.blocks_per_cu;
}

}  // extern "C"

namespace dcx {

// Launch tables of one tiling: per-shot arguments for both pointer parities
// and the prefix sums of the shots' tile counts, on the device.
struct Plan {
  const Batch &b;
  const float *d4;
  const long long *src_index;
  int n;
  StepArgs *d_args = nullptr;
  int64_t *d_prefix = nullptr;
  float *d_amp = nullptr;
  int64_t total = 0;

  Plan(const Batch &bt, const float *q, const long long *si)
      : b(bt), d4(q), src_index(si), n((int)bt.shots.size()) {
    DCX_CHECK(hipMalloc(&d_args, 2 * n * sizeof(StepArgs)));
    DCX_CHECK(hipMalloc(&d_prefix, (n + 1) * sizeof(int64_t)));
  }
  ~Plan() {
    DCX_CHECK(hipFree(d_args));
    DCX_CHECK(hipFree(d_prefix));
    if (d_amp) DCX_CHECK(hipFree(d_amp));
  }

  // Per-shot source tables (field units) for n_steps steps.
  void set_sources(const float *src_amp, unsigned n_steps) {
    std::vector<float> amp((size_t)n * n_steps, 0.0f);
    for (int k = 0; k < n; ++k) {
      const State *st = b.shots[k];
      const float fs = st->fstore == DCX_STORE_F32 ? 1.0f : st->fscale;
      if (src_amp && src_index && src_index[k] >= 0)
        for (unsigned it = 0; it < n_steps; ++it)
          amp[(size_t)k * n_steps + it] = src_amp[(size_t)k * n_steps + it] * fs;
    }
    DCX_CHECK(hipMalloc(&d_amp, amp.size() * sizeof(float)));
    DCX_CHECK(hipMemcpy(d_amp, amp.data(), amp.size() * sizeof(float),
                        hipMemcpyHostToDevice));
    steps = n_steps;
  }
  unsigned steps = 0;

  // Fill the device tables for one tiling (z chunk, prefetch depth).
  void upload(int zchunk, int deep) {
    std::vector<StepArgs> args[2];
    std::vector<int64_t> prefix(n + 1, 0);
    const State *s0 = b.shots[0];
    // Tiles are dealt to the compute units in table order, so the shots
    // with the longest tiles (most planes per z chunk) go first: the last
    // tiles of a step are then short ones and the step ends evenly.
    std::vector<int> ord(n);
    for (int k = 0; k < n; ++k) ord[k] = k;
    auto planes = [&](int k) {
      return std::min<int64_t>(zchunk, b.shots[k]->sizes[2] - 2 * STENCIL);
    };
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int i, int j) { return planes(i) > planes(j); });
    for (int j = 0; j < n; ++j) {
      const int k = ord[j];
      State *st = b.shots[k];
      // The entry points that could change these refuse batch shots; the
      // launch picks one kernel for all shots from shot 0's form.
      if (st->fstore != s0->fstore || st->cstore != s0->cstore ||
          st->kernel != DCX_KERNEL_FAST || st->halo_nuser != 0)
        DCX_FAIL("batch shot %d no longer matches the batch's storage form, "
                 "kernel or halo setting", k);
      const Box box{STENCIL, STENCIL, STENCIL, st->sizes[0] - STENCIL,
                    st->sizes[1] - STENCIL, st->sizes[2] - STENCIL};
      const float *q = d4 + 4 * k;
      for (int par = 0; par < 2; ++par) {
        StepArgs a = make_step_args(st, box, q[0], q[1], q[2], q[3]);
        a.zchunk = zchunk;
        a.deep = deep;
        a.xcd_remap = 0;
        a.gx = (int)((box.x1 - tile_origin_x(box.x0) + 31) / 32);
        a.gy = (int)((box.y1 - box.y0 + 15) / 16);
        a.src_z = -1;
        a.src_xy = -1;
        a.src_tab = d_amp ? d_amp + (size_t)k * steps : nullptr;
        if (src_index && src_index[k] >= 0) {
          if (src_index[k] >= st->n) DCX_FAIL("shot %d source out of range", k);
          const int64_t i = device_index(st, src_index[k]);
          a.src_z = i / st->zs;
          a.src_xy = i % st->zs;
        }
        args[par].push_back(a);
        // Step the host-side pointer state to the other parity and back.
        if (st->fstore == DCX_STORE_F32)
          dc_device_swap_arrays(&st->abi);
        else
          st->pcur ^= 1;
      }
      const int64_t ebytes = st->cstore == DCX_COEF_N16 ? 8 : 4;
      if ((int64_t)(box.z1 + 4) * st->zs * ebytes > (int64_t(1) << 32) - 1)
        DCX_FAIL("shot %d arrays exceed 4 GiB", k);
      prefix[j + 1] = prefix[j] + fast_tiles(args[0][j], 32, 16);
    }
    DCX_CHECK(hipMemcpy(d_args, args[0].data(), n * sizeof(StepArgs),
                        hipMemcpyHostToDevice));
    DCX_CHECK(hipMemcpy(d_args + n, args[1].data(), n * sizeof(StepArgs),
                        hipMemcpyHostToDevice));
    DCX_CHECK(hipMemcpy(d_prefix, prefix.data(), (n + 1) * sizeof(int64_t),
                        hipMemcpyHostToDevice));
    total = prefix[n];
  }

  // Grid of a persistent launch: blocks_per_cu workgroups per CU, 0 = one
  // workgroup per tile.
  int grid(int blocks_per_cu) const {
    const int64_t g = blocks_per_cu > 0 ? (int64_t)b.n_cu * blocks_per_cu : total;
    return (int)std::min<int64_t>(std::min<int64_t>(g, total), INT32_MAX);
  }

  void launch(unsigned it, int deep, int blocks_per_cu) const {
    const State *s0 = b.shots[0];
    launch_fast_batch(d_args + (it & 1) * n, d_prefix, n, total, s0->fstore,
                      s0->cstore, deep, 32, 16, grid(blocks_per_cu), (int)it,
                      (int)steps, b.stream);
  }
};

}  // namespace dcx

extern "C" {

double dcx_batch_run(dcx_batch *bb, unsigned int n_steps, const float *d4,
                     const long long *src_index, const float *src_amp,
                     int use_graph) {
  if (bb == nullptr) DCX_FAIL("null batch");
  Batch &b = bb->b;
  const int n = (int)b.shots.size();
  if (n_steps == 0) return 0.0;
  Plan plan(b, d4, src_index);
  plan.set_sources(src_amp, n_steps);

  float ms_total = 0.0f;
  // Step 0's source goes in before the first launch, as dcx_run does.
  auto add_first_source = [&]() {
    if (!(src_amp && src_index)) return;
    for (int k = 0; k < n; ++k) {
      if (src_index[k] < 0) continue;
      State *st = b.shots[k];
      const int64_t i = device_index(st, src_index[k]);
      const float a0 = src_amp[(size_t)k * n_steps];
      if (st->fstore == DCX_STORE_F32)
        launch_add_source(st->abi.pc, st->abi.qc, i, a0, b.stream);
      else
        storage_add_source(st, i, a0);
    }
  };

  unsigned it0 = 0;
  if (b.autotune) {
    // The run's first steps are the trials: every tiling computes the same
    // values, so trying one on a real step costs only its time.  Coordinate
    // search: z chunk, then prefetch depth, then workgroups per CU; each
    // candidate is timed on kTrial steps.
    constexpr unsigned kTrial = 2;
    DCX_CHECK(hipEventRecord(b.ev[0], b.stream));
    add_first_source();
    DCX_CHECK(hipEventRecord(b.ev[1], b.stream));
    DCX_CHECK(hipEventSynchronize(b.ev[1]));
    float t = 0.f;
    DCX_CHECK(hipEventElapsedTime(&t, b.ev[0], b.ev[1]));
    ms_total += t;
    int best[3] = {b.zchunk, b.deep, b.blocks_per_cu};
    float best_ms = 1e30f;
    auto trial = [&](int zc, int dp, int bpc) {
      if (it0 + kTrial > n_steps) return;
      plan.upload(zc, dp);
      DCX_CHECK(hipEventRecord(b.ev[0], b.stream));
      for (unsigned s = 0; s < kTrial; ++s, ++it0) plan.launch(it0, dp, bpc);
      DCX_CHECK(hipEventRecord(b.ev[1], b.stream));
      DCX_CHECK(hipEventSynchronize(b.ev[1]));
      float ms = 0.f;
      DCX_CHECK(hipEventElapsedTime(&ms, b.ev[0], b.ev[1]));
      ms_total += ms;
      if (ms < best_ms) {
        best_ms = ms;
        best[0] = zc;
        best[1] = dp;
        best[2] = bpc;
      }
    };
    for (int zc : {32, 64, 128, 256, 512}) trial(zc, 1, 1);
    const int zc = best[0];
    trial(zc, 0, 1);
    if (b.shots[0]->fstore == DCX_STORE_F32 && b.shots[0]->cstore == DCX_COEF_F32)
      trial(zc, 2, 1);
    const int dp = best[1];
    for (int bpc : {2, 4, 0}) trial(zc, dp, bpc);
    b.zchunk = best[0];
    b.deep = best[1];
    b.blocks_per_cu = best[2];
  }
  plan.upload(b.zchunk, b.deep);

  auto enqueue = [&]() {
    if (it0 == 0) add_first_source();
    for (unsigned it = it0; it < n_steps; ++it)
      plan.launch(it, b.deep, b.blocks_per_cu);
  };
  float ms = 0.0f;
  if (it0 < n_steps) {
    if (use_graph) {
      hipGraph_t g;
      hipGraphExec_t ge;
      DCX_CHECK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
      enqueue();
      DCX_CHECK(hipStreamEndCapture(b.stream, &g));
      DCX_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
      DCX_CHECK(hipEventRecord(b.ev[0], b.stream));
      DCX_CHECK(hipGraphLaunch(ge, b.stream));
      DCX_CHECK(hipEventRecord(b.ev[1], b.stream));
      DCX_CHECK(hipEventSynchronize(b.ev[1]));
      DCX_CHECK(hipGraphExecDestroy(ge));
      DCX_CHECK(hipGraphDestroy(g));
    } else {
      DCX_CHECK(hipEventRecord(b.ev[0], b.stream));
      enqueue();
      DCX_CHECK(hipEventRecord(b.ev[1], b.stream));
      DCX_CHECK(hipEventSynchronize(b.ev[1]));
    }
    DCX_CHECK(hipEventElapsedTime(&ms, b.ev[0], b.ev[1]));
  }
  b.last_ms_per_step = n_steps > it0 ? ms / (n_steps - it0) : 0.0;
  // Host-side pointer state after n_steps.
  if (n_steps & 1)
    for (State *st : b.shots) {
      if (st->fstore == DCX_STORE_F32)
        dc_device_swap_arrays(&st->abi);
      else
        st->pcur ^= 1;
    }
  for (State *st : b.shots) report_overflow(st);
  return (double)(ms + ms_total);
}

double dcx_batch_last_step_ms(const dcx_batch *bb) {
  if (bb == nullptr) DCX_FAIL("null batch");
  return bb->b.last_ms_per_step;
}

}  // extern "C"
