// grains_gpu <file.xml>: the standalone application of Grains3D
// (GRAINS_V1/Main/src/main.cpp:9-76) over libgrains_gpu.so: builds the
// system from the input file and runs Grains::Simulation's loop to
// <TimeInterval End> with the outputs the file asks for (grains_gpu_run):
// <RestartFile Name>_VelocityMaxMean.dat and the RawData files at the
// <TimeSave> times.
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/grains_gpu.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <file.xml> [--mixed]\n", argv[0]);
    return 1;
  }
  std::string path = argv[1];
  if (path.size() < 4 || path.compare(path.size() - 4, 4, ".xml") != 0) {
    std::fprintf(stderr, "the input file needs the .xml suffix\n");
    return 1;
  }
  int precision = GRAINS_FP64;
  for (int i = 2; i < argc; ++i)
    if (!std::strcmp(argv[i], "--mixed")) precision = GRAINS_MIXED;
  grains_gpu_system* sys = nullptr;
  int rc = grains_gpu_create_from_xml(path.c_str(), precision, &sys);
  if (rc) {
    std::fprintf(stderr, "%s\n", sys ? grains_gpu_last_error(sys) : "cannot build the system");
    grains_gpu_destroy(sys);
    return 1;
  }
  double t0, tend, dt;
  grains_gpu_time_interval(sys, &t0, &tend, &dt);
  std::printf("Grains3D on the GPU: %zu particles, t = %g .. %g, dt = %g\n", grains_gpu_num_particles(sys), t0, tend,
              dt);
  rc = grains_gpu_run(sys, 1);
  if (rc) {
    int a = 0, b = 0;
    grains_gpu_error_ids(sys, &a, &b);
    std::fprintf(stderr, "%s (components %d and %d)\n", grains_gpu_last_error(sys), a, b);
  }
  grains_gpu_destroy(sys);
  return rc ? 1 : 0;
}
