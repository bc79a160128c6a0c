// grains_gpu <file.xml>: the standalone application of Grains3D
// (GRAINS_V1/Main/src/main.cpp:9-76) over libgrains_gpu.so: builds the
// system from the input file and runs Grains::Simulation's loop to
// <TimeInterval End> (loop guard t_end - t > 0.01 dt, Grains.cpp:286),
// printing the time and the maximum and mean particle speeds
// (AllComponents::ComputeMaxMeanVelocity, AllComponents.cpp:1781-1821) every
// --every steps.  Output writers are not part of this program.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/grains_gpu.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <file.xml> [--mixed] [--every N]\n", argv[0]);
    return 1;
  }
  std::string path = argv[1];
  if (path.size() < 4 || path.compare(path.size() - 4, 4, ".xml") != 0) {
    std::fprintf(stderr, "the input file needs the .xml suffix\n");
    return 1;
  }
  int precision = GRAINS_FP64;
  long every = 1000;
  for (int i = 2; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--mixed")) precision = GRAINS_MIXED;
    else if (!std::strcmp(argv[i], "--every") && i + 1 < argc) every = std::atol(argv[++i]);
  }
  if (every < 1) every = 1;
  grains_gpu_system* sys = nullptr;
  int rc = grains_gpu_create_from_xml(path.c_str(), precision, &sys);
  if (rc) {
    std::fprintf(stderr, "%s\n", sys ? grains_gpu_last_error(sys) : "cannot build the system");
    grains_gpu_destroy(sys);
    return 1;
  }
  double t0, tend, dt;
  grains_gpu_time_interval(sys, &t0, &tend, &dt);
  const size_t n = grains_gpu_num_particles(sys);
  std::vector<double> v(3 * n);
  std::printf("Grains3D on the GPU: %zu particles, t = %g .. %g, dt = %g\n", n, t0, tend, dt);
  double t = grains_gpu_time(sys);
  int status = 0;
  while (tend - t > 0.01 * dt) {
    long left = (long)std::ceil((tend - t - 0.01 * dt) / dt);
    long k = left < every ? left : every;
    if (k < 1) k = 1;
    rc = grains_gpu_step(sys, (int)k);
    if (rc) {
      int a = 0, b = 0;
      grains_gpu_error_ids(sys, &a, &b);
      std::fprintf(stderr, "%s (components %d and %d)\n", grains_gpu_last_error(sys), a, b);
      status = 1;
      break;
    }
    t = grains_gpu_time(sys);
    grains_gpu_get_state(sys, nullptr, v.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
    double vmax = 0., vmean = 0.;
    for (size_t i = 0; i < n; ++i) {
      double s = std::sqrt(v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
      vmax = s > vmax ? s : vmax;
      vmean += s;
    }
    if (n) vmean /= double(n);
    std::printf("t = %.6e  max|v| = %.6e  mean|v| = %.6e\n", t, vmax, vmean);
  }
  grains_gpu_destroy(sys);
  return status;
}
