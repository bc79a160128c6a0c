// Output of the time loop over the C ABI: the RawData per-field files
// (RawDataPostProcessingWriter.cpp:152-600), the _VelocityMaxMean.dat
// monitor (Grains.cpp:136-146, 398-405, 199-211; AllComponents.cpp:1781-1821),
// coordination numbers (Particle.cpp:1408-1415, 1702-1706) and the time loop
// with its output times (Grains.cpp:279-452, <TimeSave> Grains.cpp:1224-1235).
//
// Everything here runs on the host over the library's own entry points: the
// state is copied down at output times only.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/grains_gpu.h"
#include "dem_system.hpp"

namespace {

int fail(grains_gpu_system* s, int code, const std::string& msg) {
  if (s) s->last_error = msg;
  return code;
}

// GrainsExec::doubleToString(ios::scientific, digits, x) (GrainsExec.cpp:
// 212-224): scientific with `digits` decimals, except that zero keeps the
// stream's default format ("0", or "-0" for a negative zero).
std::string sci(double x, int digits) {
  char buf[64];
  if (x != 0.) std::snprintf(buf, sizeof buf, "%.*e", digits, x);
  else std::snprintf(buf, sizeof buf, "%g", x);
  return buf;
}

// Steps per call of the run loop between outputs (bounds the int argument
// of grains_gpu_step on long runs without output times).
constexpr long kRunChunk = 1L << 20;

// Field files of the RawData writer, in the order the reference opens them
// (RawDataPostProcessingWriter.cpp:327-358).
const char* const kFields[10] = {"_position_x.dat",
                                 "_position_y.dat",
                                 "_position_z.dat",
                                 "_translational_velocity_x.dat",
                                 "_translational_velocity_y.dat",
                                 "_translational_velocity_z.dat",
                                 "_angular_velocity_x.dat",
                                 "_angular_velocity_y.dat",
                                 "_angular_velocity_z.dat",
                                 "_coordinationNumber.dat"};

}  // namespace

extern "C" {

int grains_gpu_set_particle_class(grains_gpu_system* s, int cls) {
  if (!s || cls < 0) return GRAINS_ARGUMENT_ERROR;
  if (s->initialized) return fail(s, GRAINS_ARGUMENT_ERROR, "particle classes are set before initialisation");
  s->next_class = cls;
  return GRAINS_OK;
}

int grains_gpu_velocity_max_mean(grains_gpu_system* s, double* vmax, double* vmean) {
  if (!s || !s->initialized || !vmax || !vmean) return GRAINS_ARGUMENT_ERROR;
  const size_t n = (size_t)s->nparticles;
  std::vector<double> v(3 * n);
  int e = grains_gpu_get_state(s, nullptr, v.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
  if (e) return e;
  // particles in id order (the active list), then the obstacles that move:
  // none here (obstacles are fixed)
  double mx = 0., mean = 0.;
  for (size_t i = 0; i < n; ++i) {
    const double* u = &v[3 * i];
    const double vit = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);  // Norm (Vector3.cpp:147-151)
    mx = vit > mx ? vit : mx;
    mean += vit;
  }
  if (n) mean /= double(n);
  *vmax = mx;
  *vmean = mean;
  return GRAINS_OK;
}

int grains_gpu_coordination_numbers(grains_gpu_system* s, int* out) {
  if (!s || !s->initialized || !out) return GRAINS_ARGUMENT_ERROR;
  if (!s->record_pairs)
    return fail(s, GRAINS_ARGUMENT_ERROR,
                "coordination numbers come from the contact records of the last force evaluation: "
                "switch recording on (grains_gpu_set_record_pairs) before it");
  size_t np = 0;
  int e = grains_gpu_get_pairs(s, nullptr, 0, &np);
  if (e) return e;
  if (s->last_status.pair_overflow)
    return fail(s, GRAINS_DEVICE_ERROR, "the contact records overflowed: coordination numbers incomplete");
  std::vector<int> buf(2 * np);
  if (np && (e = grains_gpu_get_pairs(s, buf.data(), np, &np))) return e;
  // the set of contacting component ids of each particle (Particle::
  // addContactingComponentID): a partner met through several periodic
  // images, or recorded twice, counts once
  std::vector<std::pair<int, int>> rec(np);
  for (size_t i = 0; i < np; ++i) rec[i] = {buf[2 * i], buf[2 * i + 1]};
  std::sort(rec.begin(), rec.end());
  rec.erase(std::unique(rec.begin(), rec.end()), rec.end());
  const int n = s->nparticles;
  std::fill(out, out + n, 0);
  for (auto& r : rec)
    if (r.first >= 1 && r.first <= n) ++out[r.first - 1];
  return GRAINS_OK;
}

// Coordination numbers of the current state: a force evaluation with
// contact records, then the torsor the time loop holds is put back, so that
// the next step is the one it would have been.
int grains_gpu_coordination_at_state(grains_gpu_system* s, int* out) {
  if (!s || !s->initialized || !out) return GRAINS_ARGUMENT_ERROR;
  const size_t n = (size_t)s->nparticles;
  std::vector<double> ft(6 * n);
  int e = hipSetDevice(s->device) == hipSuccess ? GRAINS_OK : GRAINS_DEVICE_ERROR;
  if (e) return fail(s, e, "hipSetDevice failed");
  if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(s, GRAINS_DEVICE_ERROR, "stream sync failed");
  for (int c = 0; c < 6 && n; ++c) {
    double* src = c < 3 ? s->P.F[c] : s->P.T[c - 3];
    if (hipMemcpy(ft.data() + c * n, src, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
      return fail(s, GRAINS_DEVICE_ERROR, "torsor copy failed");
  }
  const int rec0 = s->record_pairs;
  e = grains_gpu_set_record_pairs(s, 1);
  if (!e) e = grains_gpu_compute_forces(s);
  if (!e) e = grains_gpu_coordination_numbers(s, out);
  grains_gpu_set_record_pairs(s, rec0);
  for (int c = 0; c < 6 && n; ++c) {
    double* dst = c < 3 ? s->P.F[c] : s->P.T[c - 3];
    if (hipMemcpy(dst, ft.data() + c * n, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
      return fail(s, GRAINS_DEVICE_ERROR, "torsor copy failed");
  }
  return e;
}

// The writers of the input file's <PostProcessing> at the current state
// (AllComponents::PostProcessing_start / PostProcessing, called by
// GrainsCoupledWithFluid::InitialPostProcessing / doPostProcessing,
// GrainsCoupledWithFluid.cpp:870-893): start != 0 starts the files.
int grains_gpu_post_processing(grains_gpu_system* s, int start) {
  if (!s || !s->initialized) return GRAINS_ARGUMENT_ERROR;
  if (s->rawdata_root.empty()) return GRAINS_OK;
  std::vector<int> coord((size_t)s->nparticles, 0);
  int e = grains_gpu_coordination_at_state(s, coord.data());
  if (e) return e;
  return grains_gpu_write_rawdata(s, s->rawdata_root.c_str(), s->rawdata_binary, start ? 1 : 0, coord.data());
}

int grains_gpu_write_rawdata(grains_gpu_system* s, const char* root, int binary, int truncate, const int* coord) {
  if (!s || !s->initialized || !root || !*root) return GRAINS_ARGUMENT_ERROR;
  const size_t n = (size_t)s->nparticles;
  std::vector<double> x(3 * n), v(3 * n), w(3 * n);
  int e = grains_gpu_get_state(s, x.data(), v.data(), w.data(), nullptr, nullptr, nullptr, nullptr);
  if (e) return e;
  std::vector<int> cn(n, 0);
  if (coord) std::copy(coord, coord + n, cn.begin());
  const std::string r(root);
  // ReloadType "new": the files start empty, later frames are appended
  // (PostProcessing_start, RawDataPostProcessingWriter.cpp:55-77)
  const char* mode = truncate ? (binary ? "wb" : "w") : (binary ? "ab" : "a");
  FILE* f[10] = {};
  for (int k = 0; k < 10; ++k) {
    f[k] = std::fopen((r + kFields[k]).c_str(), mode);
    if (!f[k]) {
      for (int j = 0; j < k; ++j) std::fclose(f[j]);
      return fail(s, GRAINS_ARGUMENT_ERROR, "cannot open " + r + kFields[k]);
    }
  }
  // the particle class file holds the last frame only
  FILE* fc = std::fopen((r + "_particleType.dat").c_str(), binary ? "wb" : "w");
  if (!fc) {
    for (int k = 0; k < 10; ++k) std::fclose(f[k]);
    return fail(s, GRAINS_ARGUMENT_ERROR, "cannot open " + r + "_particleType.dat");
  }
  const double t = s->time;
  const double* src[9];
  for (int a = 0; a < 3; ++a) {
    src[a] = &x[a];
    src[3 + a] = &v[a];
    src[6 + a] = &w[a];
  }
  auto cls = [&](size_t i) { return i < s->h_class.size() ? s->h_class[i] : 0; };
  if (binary) {
    // the time leads the nine vector-field files; the reference writes it a
    // second time into the z angular velocity file and not into the
    // coordination file (RawDataPostProcessingWriter.cpp:373-393)
    for (int k = 0; k < 9; ++k) std::fwrite(&t, sizeof t, 1, f[k]);
    std::fwrite(&t, sizeof t, 1, f[8]);
    for (size_t i = 0; i < n; ++i) {
      for (int k = 0; k < 9; ++k) std::fwrite(src[k] + 3 * i, sizeof(double), 1, f[k]);
      const int c = cn[i], ty = cls(i);
      std::fwrite(&c, sizeof c, 1, f[9]);
      std::fwrite(&ty, sizeof ty, 1, fc);
    }
  } else {
    const std::string st = sci(t, 6);
    for (int k = 0; k < 9; ++k) std::fputs(st.c_str(), f[k]);
    for (size_t i = 0; i < n; ++i) {
      for (int k = 0; k < 9; ++k) {
        std::fputc(' ', f[k]);
        std::fputs(sci(src[k][3 * i], 6).c_str(), f[k]);
      }
      std::fprintf(f[9], " %d", cn[i]);
      std::fprintf(fc, "%d ", cls(i));
    }
    for (int k = 0; k < 10; ++k) std::fputc('\n', f[k]);
    std::fputc('\n', fc);
  }
  bool ok = true;
  for (int k = 0; k < 10; ++k) ok &= std::fclose(f[k]) == 0;
  ok &= std::fclose(fc) == 0;
  return ok ? GRAINS_OK : fail(s, GRAINS_ARGUMENT_ERROR, "cannot write the RawData files of " + r);
}

int grains_gpu_set_outputs(grains_gpu_system* s, const double* save_times, size_t nsave, const char* save_root,
                           const char* rawdata_root, int rawdata_binary) {
  if (!s || (nsave && !save_times)) return GRAINS_ARGUMENT_ERROR;
  s->save_times.assign(save_times, save_times + nsave);
  s->save_root = save_root ? save_root : "";
  s->rawdata_root = rawdata_root ? rawdata_root : "";
  s->rawdata_binary = rawdata_binary != 0;
  return GRAINS_OK;
}

int grains_gpu_run(grains_gpu_system* s, int verbose) {
  if (!s || !s->initialized) return GRAINS_ARGUMENT_ERROR;
  const double dt = s->cfg.dt, tend = s->t_end;
  const bool raw = !s->rawdata_root.empty();
  const bool vit = !s->save_root.empty();
  std::vector<int> coord((size_t)s->nparticles, 0);
  long waiting = 0;
  auto pending = [&]() { return grains_gpu_pending_insertions(s, &waiting) == GRAINS_OK && waiting > 0; };
  FILE* fv = nullptr;
  int e = GRAINS_OK;
  double vmax = 0., vmean = 0.;
  auto vit_line = [&]() -> int {
    if (!vit) return GRAINS_OK;
    int r = grains_gpu_velocity_max_mean(s, &vmax, &vmean);
    if (r) return r;
    std::fprintf(fv, "%s\t%s\t%s\n", sci(s->time, 6).c_str(), sci(vmax, 6).c_str(), sci(vmean, 6).c_str());
    std::fflush(fv);
    if (verbose) std::printf("Component velocity : max = %g average = %g\n", vmax, vmean);
    return GRAINS_OK;
  };
  // initial outputs (Grains.cpp:130-146): the particles have no contact yet
  if (raw && (e = grains_gpu_write_rawdata(s, s->rawdata_root.c_str(), s->rawdata_binary, 1, coord.data())))
    return e;
  if (vit) {
    fv = std::fopen((s->save_root + "_VelocityMaxMean.dat").c_str(), "w");
    if (!fv) return fail(s, GRAINS_ARGUMENT_ERROR, "cannot open " + s->save_root + "_VelocityMaxMean.dat");
  }
  if ((e = vit_line())) {
    if (fv) std::fclose(fv);
    return e;
  }
  // first output time after the start (Grains.cpp:156-159)
  size_t next = 0;
  while (next < s->save_times.size() && s->save_times[next] - s->time < 0.01 * dt) ++next;
  bool last_saved = false;
  const int rec0 = s->record_pairs;
  while (tend - s->time > 0.01 * dt) {
    // steps up to the next output time or the end, with the reference's
    // time accumulation and tests (Grains.cpp:286-306)
    double t = s->time;
    long k = 0;
    bool out = false;
    while (tend - t > 0.01 * dt) {
      t += dt;
      ++k;
      if (next < s->save_times.size() && s->save_times[next] - t < 0.01 * dt) {
        out = true;
        break;
      }
    }
    // the contact records of the step that ends at an output time give the
    // coordination numbers; the other steps run without records
    // while particles wait for insertion over time, one step at a time with
    // the insertion call before it (Grains.cpp:322-327)
    e = grains_gpu_set_record_pairs(s, rec0);
    long left = k - 1;
    int in = 0;
    while (!e && left > 0) {
      if (pending()) {
        e = grains_gpu_try_insert(s, &in);
        if (!e) e = grains_gpu_step(s, 1);
        --left;
      } else {
        const long c = std::min(left, kRunChunk);
        e = grains_gpu_step(s, (int)c);
        left -= c;
      }
    }
    if (!e && out && raw) e = grains_gpu_set_record_pairs(s, 1);
    if (!e && pending()) e = grains_gpu_try_insert(s, &in);
    if (!e) e = grains_gpu_step(s, 1);
    coord.assign((size_t)s->nparticles, 0);
    if (!e && out && raw) e = grains_gpu_coordination_numbers(s, coord.data());
    grains_gpu_set_record_pairs(s, rec0);
    if (e) break;
    last_saved = out;
    if (out) {
      ++next;
      if (verbose) std::printf("Time = %.6e\n", s->time);
      if ((e = vit_line())) break;
      if (raw && (e = grains_gpu_write_rawdata(s, s->rawdata_root.c_str(), s->rawdata_binary, 0, coord.data())))
        break;
    }
  }
  // end of the run (Grains.cpp:199-211)
  if (!e && !last_saved) e = vit_line();
  if (fv) std::fclose(fv);
  return e;
}

}  // extern "C"
