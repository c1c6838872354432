// The reference program (src/main.c, src/coordinator.c, src/worker.c) on one
// GPU.  The global grid is cut into subdomains exactly as the coordinator cuts
// it for a given rank topology; each subdomain is a dc_device_data created
// through the drop-in ABI, and all of them live on the one GPU, ordered on one
// stream.  Halo exchange is on-device: every send box of every subdomain is
// packed into one buffer by one launch and unpacked into the neighbours'
// receive boxes by a second launch -- the packed buffer is what a multi-GPU
// run would hand to RCCL or peer stores.
#include <cmath>
#include <cstring>
#include <vector>

#include "dc_internal.hpp"

namespace dcx {
namespace {

struct Sub {
  int coord[3];
  int64_t size[3];   // local sizes incl. the 4-cell ghost frame
  int64_t start[3];  // global coordinate of local (0,0,0)
  long long src;     // local source index or -1
  dc_device_data *d;
};

// Send/receive box of displacement d along one axis (src/worker.c:196-228,
// 679-711): d = -1 -> [4,8) / [0,4), d = +1 -> [s-8,s-4) / [s-4,s),
// d = 0 -> [4,s-4) for both.
void axis_box(int d, int64_t s, bool recv, int64_t &a, int64_t &b) {
  if (d == -1) {
    a = recv ? 0 : STENCIL;
    b = recv ? STENCIL : 2 * STENCIL;
  } else if (d == 1) {
    a = recv ? s - STENCIL : s - 2 * STENCIL;
    b = recv ? s : s - STENCIL;
  } else {
    a = STENCIL;
    b = s - STENCIL;
  }
}

}  // namespace
}  // namespace dcx

using namespace dcx;

// MPI_Dims_create(n, 3, dims) with dims zeroed, as src/main.c:89-90 calls it:
// of the factorisations n = a*b*c with a >= b >= c, the one with the least
// spread a - c, and of equal spreads the one with the larger c (MPICH's
// balanced decomposition; n = 360 gives 10,6,6, not 9,8,5).
extern "C" int dcx_dims_create(int n, int dims[3]) {
  if (n < 1 || dims == nullptr) return -1;
  int best_spread = -1, best_c = 0;
  for (int a = 1; a <= n; ++a) {
    if (n % a) continue;
    const int m = n / a;
    for (int b = 1; b <= a; ++b) {
      if (m % b) continue;
      const int c = m / b;
      if (c > b) continue;
      const int spread = a - c;
      if (best_spread < 0 || spread < best_spread ||
          (spread == best_spread && c > best_c)) {
        best_spread = spread;
        best_c = c;
        dims[0] = a;
        dims[1] = b;
        dims[2] = c;
      }
    }
  }
  return 0;
}

extern "C" int dcx_program_run(const dcx_program_args *args, float *out_pc,
                               float *out_qc, double stats[3]) {
  if (args == nullptr) return -1;
  const int T[3] = {args->topology[0], args->topology[1], args->topology[2]};
  for (int k = 0; k < 3; ++k)
    if (T[k] < 1) DCX_FAIL("topology entries must be >= 1");
  const size_t prob[3] = {args->size_x, args->size_y, args->size_z};
  const int64_t ab = (int64_t)args->absorption;
  int64_t g[3];
  for (int k = 0; k < 3; ++k) g[k] = (int64_t)prob[k] + 2 * ab + 2 * STENCIL;
  // src/main.c:132 -- float division, then ceil.
  const unsigned iterations =
      (unsigned)std::ceil(args->time_max / args->dt);

  if (g[0] > INT32_MAX || g[1] > INT32_MAX || g[2] > INT32_MAX)
    DCX_FAIL("global grid too large");
  const int gsz[3] = {(int)g[0], (int)g[1], (int)g[2]};
  const int psz[3] = {(int)prob[0], (int)prob[1], (int)prob[2]};
  // Source at the global centre (src/coordinator.c:11-16).
  const int64_t sc[3] = {g[0] / 2, g[1] / 2, g[2] / 2};

  // Partition (src/coordinator.c:29-91, src/main.c:110-129).
  int64_t part[3], rem[3];
  for (int k = 0; k < 3; ++k) {
    part[k] = (g[k] - 2 * STENCIL) / T[k];
    rem[k] = (g[k] - 2 * STENCIL) % T[k];
    if (part[k] < 1) DCX_FAIL("topology too fine for the grid");
  }
  hipStream_t stream;
  DCX_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  std::vector<Sub> subs;
  for (int cz = 0; cz < T[2]; ++cz)
    for (int cy = 0; cy < T[1]; ++cy)
      for (int cx = 0; cx < T[0]; ++cx) {
        Sub s{};
        const int c[3] = {cx, cy, cz};
        for (int k = 0; k < 3; ++k) {
          s.coord[k] = c[k];
          s.size[k] = part[k] + (c[k] == T[k] - 1 ? rem[k] : 0) + 2 * STENCIL;
          s.start[k] = c[k] * part[k];
        }
        bool has = true;
        for (int k = 0; k < 3; ++k)
          has = has && sc[k] >= s.start[k] && sc[k] < s.start[k] + s.size[k];
        s.src = has ? (long long)((sc[0] - s.start[0]) +
                                  (sc[1] - s.start[1]) * s.size[0] +
                                  (sc[2] - s.start[2]) * s.size[0] * s.size[1])
                    : -1;
        dc_process_t p{};
        for (int k = 0; k < 3; ++k) {
          p.coordinates[k] = c[k];
          p.topology[k] = T[k];
          p.sizes[k] = (size_t)s.size[k];
        }
        p.iterations = iterations;
        p.source_index = (int)s.src;
        p.dx = args->dx;
        p.dy = args->dy;
        p.dz = args->dz;
        p.dt = args->dt;
        // The rank's model, built on the device as the reference's worker
        // builds it on the host (src/worker.c:49-156); every rank's shell
        // walk starts from seed 0 (src/worker.c:93).  Fields start at zero.
        const int st0[3] = {(int)s.start[0], (int)s.start[1], (int)s.start[2]};
        unsigned int seed = 0;
        s.d = dcx_device_data_init_model(&p, gsz, st0, psz, (int)ab, &seed);
        State *st = state_of(s.d);
        adopt_stream(st, stream);
        st->kernel = args->kernel;
        subs.push_back(s);
      }
  auto index_of = [&](int cx, int cy, int cz) {
    return (cx * 1) + cy * T[0] + cz * T[0] * T[1];
  };

  // Halo tables for the two pointer parities (all subdomains swap together).
  std::vector<BoxCopy> pack[2], unpack[2];
  int64_t total = 0;
  for (int par = 0; par < 2; ++par) {
    total = 0;
    for (const Sub &s : subs) {
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            if (dx == 0 && dy == 0 && dz == 0) continue;
            const int d[3] = {dx, dy, dz};
            int nc[3];
            bool ok = true;
            for (int k = 0; k < 3; ++k) {
              nc[k] = s.coord[k] + d[k];
              ok = ok && nc[k] >= 0 && nc[k] < T[k];
            }
            if (!ok) continue;
            const Sub &r = subs[index_of(nc[0], nc[1], nc[2])];
            // After the update the new fields are in pp/qp; with the
            // initial pointers that is field_mem[0]/[2] at even steps and
            // field_mem[1]/[3] at odd steps.
            const State *ss = state_of(s.d);
            const State *rs = state_of(r.d);
            for (int f = 0; f < 2; ++f) {
              BoxCopy pb{}, ub{};
              pb.arr = ss->field_mem[2 * f + par];
              ub.arr = rs->field_mem[2 * f + par];
              pb.ys = ss->ys;
              pb.zs = ss->zs;
              ub.ys = rs->ys;
              ub.zs = rs->zs;
              int64_t a0[3], a1[3], b0[3], b1[3];
              for (int k = 0; k < 3; ++k) {
                axis_box(d[k], s.size[k], false, a0[k], a1[k]);
                axis_box(-d[k], r.size[k], true, b0[k], b1[k]);
                if (a1[k] - a0[k] != b1[k] - b0[k])
                  DCX_FAIL("halo box mismatch");
              }
              pb.x0 = a0[0];
              pb.y0 = a0[1];
              pb.z0 = a0[2];
              ub.x0 = b0[0];
              ub.y0 = b0[1];
              ub.z0 = b0[2];
              pb.nx = ub.nx = a1[0] - a0[0];
              pb.ny = ub.ny = a1[1] - a0[1];
              pb.nz = ub.nz = a1[2] - a0[2];
              pb.off = ub.off = total;
              check_multi_box(pb);
              total += pb.nx * pb.ny * pb.nz;
              pack[par].push_back(pb);
              unpack[par].push_back(ub);
            }
          }
    }
  }
  const int nbox = (int)pack[0].size();
  BoxCopy *d_tab = nullptr;
  float *d_buf = nullptr;
  if (nbox > 0) {
    DCX_CHECK(hipMalloc(&d_tab, 4 * (size_t)nbox * sizeof(BoxCopy)));
    DCX_CHECK(hipMalloc(&d_buf, (size_t)total * sizeof(float)));
    for (int par = 0; par < 2; ++par) {
      DCX_CHECK(hipMemcpy(d_tab + (2 * par) * nbox, pack[par].data(),
                          nbox * sizeof(BoxCopy), hipMemcpyHostToDevice));
      DCX_CHECK(hipMemcpy(d_tab + (2 * par + 1) * nbox, unpack[par].data(),
                          nbox * sizeof(BoxCopy), hipMemcpyHostToDevice));
    }
  }

  hipEvent_t e0, e1;
  DCX_CHECK(hipEventCreate(&e0));
  DCX_CHECK(hipEventCreate(&e1));
  DCX_CHECK(hipEventRecord(e0, stream));
  for (unsigned it = 0; it < iterations; ++it) {
    const float s = dcx_ricker(args->dt, (int)it, 40.0);
    for (const Sub &sb : subs)
      if (sb.src >= 0) dc_device_add_source(sb.d, (size_t)sb.src, s);
    for (const Sub &sb : subs) {
      const State *st = state_of(sb.d);
      const Box box{STENCIL, STENCIL, STENCIL, sb.size[0] - STENCIL,
                    sb.size[1] - STENCIL, sb.size[2] - STENCIL};
      launch_stencil(st, make_step_args(st, box, args->dx, args->dy, args->dz,
                                        args->dt));
    }
    if (nbox > 0) {
      const int par = it & 1;
      launch_box_copy(d_tab + (2 * par) * nbox, nbox, total, d_buf, 0, stream);
      launch_box_copy(d_tab + (2 * par + 1) * nbox, nbox, total, d_buf, 1,
                      stream);
    }
    for (const Sub &sb : subs) dc_device_swap_arrays(sb.d);
  }
  DCX_CHECK(hipEventRecord(e1, stream));
  DCX_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  DCX_CHECK(hipEventElapsedTime(&ms, e0, e1));

  // Gather the compute regions into the global layout.
  if (out_pc || out_qc) {
    const int64_t gn = g[0] * g[1] * g[2];
    if (out_pc) std::memset(out_pc, 0, (size_t)gn * sizeof(float));
    if (out_qc) std::memset(out_qc, 0, (size_t)gn * sizeof(float));
    for (const Sub &sb : subs) {
      const int64_t ln = sb.size[0] * sb.size[1] * sb.size[2];
      std::vector<float> pc(ln), qc(ln);
      dcx_download(sb.d, 1, pc.data());
      dcx_download(sb.d, 3, qc.data());
      for (int64_t z = STENCIL; z < sb.size[2] - STENCIL; ++z)
        for (int64_t y = STENCIL; y < sb.size[1] - STENCIL; ++y) {
          const int64_t li = STENCIL + y * sb.size[0] + z * sb.size[0] * sb.size[1];
          const int64_t gi = (STENCIL + sb.coord[0] * part[0]) +
                             (sb.coord[1] * part[1] + y) * g[0] +
                             (sb.coord[2] * part[2] + z) * g[0] * g[1];
          const size_t cnt = (size_t)(sb.size[0] - 2 * STENCIL);
          if (out_pc) std::memcpy(out_pc + gi, pc.data() + li, cnt * sizeof(float));
          if (out_qc) std::memcpy(out_qc + gi, qc.data() + li, cnt * sizeof(float));
        }
    }
  }
  if (stats) {
    const double secs = ms * 1e-3;
    const double cells = (double)(g[0] - 2 * STENCIL) * (g[1] - 2 * STENCIL) *
                         (g[2] - 2 * STENCIL);
    stats[0] = secs;
    stats[1] = iterations;
    stats[2] = secs > 0 ? cells * iterations / 1e6 / secs : 0.0;
  }
  for (Sub &sb : subs) dc_device_data_free(sb.d);
  if (d_tab) DCX_CHECK(hipFree(d_tab));
  if (d_buf) DCX_CHECK(hipFree(d_buf));
  DCX_CHECK(hipEventDestroy(e0));
  DCX_CHECK(hipEventDestroy(e1));
  DCX_CHECK(hipStreamDestroy(stream));
  return 0;
}
