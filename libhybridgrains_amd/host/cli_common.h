// cli_common.h -- option parsing shared by the three command-line drivers.
#pragma once

#include <getopt.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "hg.h"

namespace hgcli {

struct Options {
  std::string scene;
  double end_time = -1.0;  // -e: overrides the scene's end_time
  long steps = -1;         // -n: step count instead of an end time
  int precision = HG_FP64; // -p fp32|fp64
  long report_every = 0;   // -f: status line every f steps (0: start and end only)
};

inline void usage(const char *prog, const char *what) {
  std::printf(
      "Usage: %s [options] scene.xml\n"
      "Steps a %s scene on the GPU.\n"
      "  -h, --help           this message\n"
      "  -e, --end T          simulate until time T (overrides the scene's end_time)\n"
      "  -n, --steps N        take N steps instead\n"
      "  -p, --precision P    fp64 (default) or fp32\n"
      "  -f, --frequency F    print a status line every F steps\n",
      prog, what);
}

// Returns 0 to run, 1 after --help, -1 on a bad command line.
inline int parse(int argc, char **argv, const char *what, Options &o) {
  static const option longopts[] = {{"help", no_argument, nullptr, 'h'},
                                    {"end", required_argument, nullptr, 'e'},
                                    {"steps", required_argument, nullptr, 'n'},
                                    {"precision", required_argument, nullptr, 'p'},
                                    {"frequency", required_argument, nullptr, 'f'},
                                    {nullptr, 0, nullptr, 0}};
  int c;
  while ((c = getopt_long(argc, argv, "he:n:p:f:", longopts, nullptr)) != -1) {
    switch (c) {
      case 'h':
        usage(argv[0], what);
        return 1;
      case 'e':
        o.end_time = std::atof(optarg);
        if (!(o.end_time > 0.0)) {
          std::fprintf(stderr, "end time must be a positive scalar\n");
          return -1;
        }
        break;
      case 'n':
        o.steps = std::atol(optarg);
        if (o.steps < 0) {
          std::fprintf(stderr, "step count must be non-negative\n");
          return -1;
        }
        break;
      case 'p':
        if (std::strcmp(optarg, "fp64") == 0) o.precision = HG_FP64;
        else if (std::strcmp(optarg, "fp32") == 0) o.precision = HG_FP32;
        else {
          std::fprintf(stderr, "precision must be fp64 or fp32\n");
          return -1;
        }
        break;
      case 'f':
        o.report_every = std::atol(optarg);
        break;
      default:
        return -1;
    }
  }
  if (optind != argc - 1) {
    usage(argv[0], what);
    return -1;
  }
  o.scene = argv[optind];
  return 0;
}

// Steps until iteration * dt reaches the end time, as the reference drivers
// stop (mpmgrains2d_cli.cpp:296-303).
inline long step_count(const Options &o, double dt, double scene_end) {
  if (o.steps >= 0) return o.steps;
  const double end = o.end_time > 0.0 ? o.end_time : scene_end;
  if (!std::isfinite(end)) return -1;
  long k = static_cast<long>(std::ceil(end / dt));
  while (k > 0 && double(k - 1) * dt >= end) --k;
  while (double(k) * dt < end) ++k;
  return k;
}

inline double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace hgcli
