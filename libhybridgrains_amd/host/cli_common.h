// cli_common.h -- option parsing shared by the three command-line drivers.
#pragma once

#include <getopt.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hg.h"

namespace hgcli {

struct Options {
  std::string scene;
  double end_time = -1.0;  // -e: overrides the scene's end_time
  long steps = -1;         // -n: step count instead of an end time
  int precision = HG_FP64; // -p fp32|fp64
  long report_every = 0;   // -f: status line every f steps (0: start and end only)
  int serialize = -1;      // -s: resumable snapshots (0: one overwritten file, 1: one per snapshot)
  std::string resume;      // -r: snapshot to resume from
};

inline void usage(const char *prog, const char *what) {
  std::printf(
      "Usage: %s [options] scene.xml\n"
      "Steps a %s scene on the GPU.\n"
      "  -h, --help           this message\n"
      "  -e, --end T          simulate until time T (overrides the scene's end_time)\n"
      "  -n, --steps N        take N steps instead\n"
      "  -p, --precision P    fp64 (default) or fp32\n"
      "  -f, --frequency F    print a status line every F steps\n"
      "  -s, --serialize_snapshots B\n"
      "                       save a bit-identical resumable snapshot at every status\n"
      "                       line and at the end: B = 0 overwrites serial.bin, B = 1\n"
      "                       writes serial_<step>.bin each time\n"
      "  -r, --resume FILE    resume from a snapshot of this scene\n",
      prog, what);
}

// Returns 0 to run, 1 after --help, -1 on a bad command line.
inline int parse(int argc, char **argv, const char *what, Options &o) {
  static const option longopts[] = {{"help", no_argument, nullptr, 'h'},
                                    {"end", required_argument, nullptr, 'e'},
                                    {"steps", required_argument, nullptr, 'n'},
                                    {"precision", required_argument, nullptr, 'p'},
                                    {"frequency", required_argument, nullptr, 'f'},
                                    {"serialize_snapshots", required_argument, nullptr, 's'},
                                    {"resume", required_argument, nullptr, 'r'},
                                    {nullptr, 0, nullptr, 0}};
  int c;
  while ((c = getopt_long(argc, argv, "he:n:p:f:s:r:", longopts, nullptr)) != -1) {
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
      case 's':
        if (std::strcmp(optarg, "0") == 0) o.serialize = 0;
        else if (std::strcmp(optarg, "1") == 0) o.serialize = 1;
        else {
          std::fprintf(stderr, "-s/--serialize_snapshots takes 0 or 1\n");
          return -1;
        }
        break;
      case 'r':
        o.resume = optarg;
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

// Snapshot files of the drivers: magic, scene kind, step index, dt, then the
// engine's snapshot blob (hg_*_snapshot).
constexpr char kSnapMagic[8] = {'H', 'G', 'C', 'L', 'I', 'S', 'N', '1'};
enum : std::uint32_t { kSnapMpm = 1, kSnapDem = 2, kSnapHybrid = 3 };

inline std::string snapshot_path(int serialize, long it) {
  if (serialize != 1) return "serial.bin";
  char buf[64];
  std::snprintf(buf, sizeof(buf), "serial_%010ld.bin", it);
  return buf;
}

inline bool write_snapshot(const std::string &path, std::uint32_t kind, long it, double dt,
                           const std::vector<std::uint8_t> &blob) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::fprintf(stderr, "Failed to open serialization file: %s\n", path.c_str());
    return false;
  }
  const std::int64_t step = it;
  const std::uint64_t n = blob.size();
  bool ok = std::fwrite(kSnapMagic, 1, 8, f) == 8 && std::fwrite(&kind, 4, 1, f) == 1 &&
            std::fwrite(&step, 8, 1, f) == 1 && std::fwrite(&dt, 8, 1, f) == 1 &&
            std::fwrite(&n, 8, 1, f) == 1 && std::fwrite(blob.data(), 1, n, f) == n;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) std::fprintf(stderr, "Failed to write serialization file: %s\n", path.c_str());
  else std::printf("Serialized step %ld to %s\n", it, path.c_str());
  return ok;
}

inline bool read_snapshot(const std::string &path, std::uint32_t kind, double dt, long &it,
                          std::vector<std::uint8_t> &blob) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) {
    std::fprintf(stderr, "Failed to open serialized state in file: %s\n", path.c_str());
    return false;
  }
  char magic[8];
  std::uint32_t k = 0;
  std::int64_t step = 0;
  double sdt = 0.0;
  std::uint64_t n = 0;
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kSnapMagic, 8) == 0 &&
            std::fread(&k, 4, 1, f) == 1 && std::fread(&step, 8, 1, f) == 1 &&
            std::fread(&sdt, 8, 1, f) == 1 && std::fread(&n, 8, 1, f) == 1 && k == kind &&
            step >= 0 && n < (std::uint64_t(1) << 40);
  if (ok) {
    blob.resize(n);
    ok = std::fread(blob.data(), 1, n, f) == n;
  }
  std::fclose(f);
  if (!ok) {
    std::fprintf(stderr, "File %s does not appear to be a snapshot of this kind of scene.\n",
                 path.c_str());
    return false;
  }
  if (sdt != dt) std::fprintf(stderr, "Warning: the snapshot was taken with another dt.\n");
  it = long(step);
  return true;
}

inline double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace hgcli
