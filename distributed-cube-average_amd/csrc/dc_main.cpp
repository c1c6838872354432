// dc_hip: the reference program's command line (src/main.c:16-76) on one GPU.
//
//   dc_hip --size-x N --size-y N --size-z N --dx F --dy F --dz F --dt F
//          -t TIME_MAX -a ABSORPTION -o OUTPUT [--ranks R | --topology X,Y,Z]
//          [--kernel fast|exact]
//
// All reference flags are required and must be non-zero (src/main.c:64-71);
// otherwise the usage is printed and the program exits with status 64, as
// argp_usage does.  --ranks R cuts the grid for R ranks the way MPI_Dims_create
// and the reference's coordinator would (default 1); all of them run on the one
// GPU with on-device halo exchange.  Output: the .dc file (pc then qc, float32,
// src/coordinator.c:116-219) and the reference's CSV report
// (src/main.c:200-215): "rank,total_time,msamples_per_s", a row per rank and a
// "*" row for the whole grid.
#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dc_hip.h"

static void usage(const char *prog) {
  std::fprintf(stderr,
               "Usage: %s --size-x=INTEGER --size-y=INTEGER --size-z=INTEGER "
               "--dx=FLOAT --dy=FLOAT --dz=FLOAT --dt=FLOAT -t TIME_MAX "
               "-a ABSORPTION -o PATH [--ranks=R | --topology=X,Y,Z] "
               "[--kernel=fast|exact]\n"
               "A program that solves Fletcher equations in a distributed setup\n",
               prog);
  std::exit(64);
}

int main(int argc, char **argv) {
  dcx_program_args a{};
  a.kernel = DCX_KERNEL_FAST;
  const char *out = nullptr;
  int ranks = 1;
  int topo[3] = {0, 0, 0};
  static const option opts[] = {
      {"size-x", required_argument, nullptr, 128},
      {"size-y", required_argument, nullptr, 129},
      {"size-z", required_argument, nullptr, 130},
      {"dx", required_argument, nullptr, 131},
      {"dy", required_argument, nullptr, 132},
      {"dz", required_argument, nullptr, 133},
      {"dt", required_argument, nullptr, 134},
      {"time-max", required_argument, nullptr, 't'},
      {"absorption", required_argument, nullptr, 'a'},
      {"output-file", required_argument, nullptr, 'o'},
      {"ranks", required_argument, nullptr, 200},
      {"topology", required_argument, nullptr, 201},
      {"kernel", required_argument, nullptr, 202},
      {"help", no_argument, nullptr, 'h'},
      {nullptr, 0, nullptr, 0}};
  int c;
  while ((c = getopt_long(argc, argv, "t:a:o:h", opts, nullptr)) != -1) {
    switch (c) {
      case 128: a.size_x = (size_t)std::atoi(optarg); break;
      case 129: a.size_y = (size_t)std::atoi(optarg); break;
      case 130: a.size_z = (size_t)std::atoi(optarg); break;
      case 131: a.dx = (float)std::atof(optarg); break;
      case 132: a.dy = (float)std::atof(optarg); break;
      case 133: a.dz = (float)std::atof(optarg); break;
      case 134: a.dt = (float)std::atof(optarg); break;
      case 't': a.time_max = (float)std::atof(optarg); break;
      case 'a': a.absorption = (size_t)std::atoi(optarg); break;
      case 'o': out = optarg; break;
      case 200: ranks = std::atoi(optarg); break;
      case 201:
        if (std::sscanf(optarg, "%d,%d,%d", &topo[0], &topo[1], &topo[2]) != 3)
          usage(argv[0]);
        break;
      case 202:
        if (std::strcmp(optarg, "exact") == 0)
          a.kernel = DCX_KERNEL_EXACT;
        else if (std::strcmp(optarg, "fast") == 0)
          a.kernel = DCX_KERNEL_FAST;
        else
          usage(argv[0]);
        break;
      default: usage(argv[0]);
    }
  }
  if (a.size_x == 0 || a.size_y == 0 || a.size_z == 0 || a.dx == 0 ||
      a.dy == 0 || a.dz == 0 || a.time_max == 0 || a.dt == 0 ||
      a.absorption == 0 || out == nullptr || ranks < 1)
    usage(argv[0]);
  if (topo[0] > 0) {
    for (int k = 0; k < 3; ++k) a.topology[k] = topo[k];
  } else {
    dcx_dims_create(ranks, a.topology);
  }
  const size_t gx = a.size_x + 2 * a.absorption + 8;
  const size_t gy = a.size_y + 2 * a.absorption + 8;
  const size_t gz = a.size_z + 2 * a.absorption + 8;
  std::vector<float> pc(gx * gy * gz), qc(gx * gy * gz);
  double stats[3] = {0, 0, 0};
  if (dcx_program_run(&a, pc.data(), qc.data(), stats) != 0) return 1;
  if (dcx_write_dc_file(out, gx, gy, gz, pc.data(), qc.data()) != 0) return 1;
  const int n = a.topology[0] * a.topology[1] * a.topology[2];
  std::printf("rank,total_time,msamples_per_s\n");
  // The subdomains share one GPU and one timed loop: each row reports that
  // loop, per-rank throughput being the global figure divided by the ranks.
  for (int r = 0; r < n; ++r)
    std::printf("%d,%lf,%lf\n", r, stats[0], stats[2] / n);
  std::printf("*,%lf,%lf\n", stats[0], stats[2]);
  return 0;
}
