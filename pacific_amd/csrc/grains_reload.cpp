// Reload (restart) files in the __Format2024__ layout of Grains3D:
//
//   Grains::saveReload ..................... Base/src/Grains.cpp:2543-2584
//   ContactBuilderFactory::save ............ Component/src/ContactBuilderFactory.cpp:270-329
//   AllComponents::write ................... Component/src/AllComponents.cpp:1299-1415
//   Particle::write / write2014(_binary) ... Component/src/Particle.cpp:1076-1193
//   SecondOrderLeapFrog 2014 fields ........ TimeIntegration/src/SecondOrderLeapFrog.cpp:98-150
//   Component::writeContactMemory2014 ...... Component/src/Component.cpp:776-835
//   reading: Grains.cpp:784-844, AllComponents::read / read_particles
//   (AllComponents.cpp:1024-1290), Particle::read2014(_binary) (Particle.cpp:968-1060)
//
// A save writes the main file <root>A or <root>B (alternating, A first), the
// contact models <root>_MatContact.xml (first save only), the particle
// records <root>X_Particles (text) or <root>X_Particles.bin (Hybrid) and one
// line "time <root>X" appended to <root>_RFTable.txt.  A load reads the time,
// and for every particle id its position, basis, velocities, quaternion,
// torsor (LeapFrog: the force of the last step, so that a restarted run needs
// no new force evaluation, Grains.cpp:169-174) and contact memory, into a
// system that holds the same particles (ids, classes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/grains_gpu.h"
#include "dem_math.hpp"
#include "dem_system.hpp"

using dem::d3;
using dem::quat;

namespace {

int fail(grains_gpu_system* s, int code, const std::string& msg) {
  if (s) s->last_error = msg;
  return code;
}

// GrainsExec::doubleToString(ios::scientific, FORMAT16DIGITS, x): %.16e,
// zero in the stream's default format
std::string s16(double x) {
  char buf[64];
  if (x != 0.) std::snprintf(buf, sizeof buf, "%.16e", x);
  else std::snprintf(buf, sizeof buf, "%g", x);
  return buf;
}
// a double through an ostream at its default precision (6 significant)
std::string dflt(double x) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%g", x);
  return buf;
}
std::string g3(const double* v) { return s16(v[0]) + " " + s16(v[1]) + " " + s16(v[2]); }

// file part of a path (GrainsExec::extractFileName)
std::string file_part(const std::string& p) {
  size_t k = p.find_last_of('/');
  return k == std::string::npos ? p : p.substr(k + 1);
}
std::string dir_part(const std::string& p) {
  size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string(".") : p.substr(0, k);
}
// the particle file of a reload file: its name up to the first "." of the
// file part, + "_Particles" (AllComponents.cpp:1373-1378 cut the whole path
// at its first "."; only the file part is cut here, so that directories with
// dots in their names stay intact)
std::string particle_file(const std::string& reload, bool hybrid) {
  const std::string dir = dir_part(reload), f = file_part(reload);
  size_t pos = f.find('.');
  std::string base = pos == std::string::npos ? f : f.substr(0, pos);
  return (reload.find('/') == std::string::npos ? std::string() : dir + "/") + base + "_Particles" +
         (hybrid ? ".bin" : "");
}

// rotation matrix rows of a unit quaternion (x, y, z, w)
void quat_matrix(const double* q, double* B) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  B[0] = 1. - 2. * (y * y + z * z);
  B[1] = 2. * (x * y - z * w);
  B[2] = 2. * (x * z + y * w);
  B[3] = 2. * (x * y + z * w);
  B[4] = 1. - 2. * (x * x + z * z);
  B[5] = 2. * (y * z - x * w);
  B[6] = 2. * (x * z - y * w);
  B[7] = 2. * (y * z + x * w);
  B[8] = 1. - 2. * (x * x + y * y);
}

struct HostState {
  std::vector<double> x, v, w, wb, q, F, T, B;
  std::vector<double> ab[4];  // AB2: v_nm2, wb_nm2, dUdt_nm2, dWdt_nm2 (3 per particle)
  // contact memory: per particle, (partner id, kd, pn, sr)
  std::vector<std::vector<std::tuple<int, d3, d3, d3>>> mem;
};

// state of the particles in id order (storage index = id - 1 on the host
// side of every download below)
int download(grains_gpu_system* s, HostState& h) {
  const size_t n = (size_t)s->nparticles;
  h.x.resize(3 * n);
  h.v.resize(3 * n);
  h.w.resize(3 * n);
  h.wb.resize(3 * n);
  h.q.resize(4 * n);
  h.F.resize(3 * n);
  h.T.resize(3 * n);
  int e = grains_gpu_get_state(s, h.x.data(), h.v.data(), h.w.data(), h.wb.data(), h.q.data(), h.F.data(),
                               h.T.data());
  if (e) return e;
  h.B.resize(9 * n);
  if (s->has_shapes()) {
    if ((e = grains_gpu_get_basis(s, h.B.data()))) return e;
  } else {
    for (size_t i = 0; i < n; ++i) quat_matrix(&h.q[4 * i], &h.B[9 * i]);
  }
  std::vector<int> ids(n);
  if (n && hipMemcpy(ids.data(), s->P.id, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(s, GRAINS_DEVICE_ERROR, "reload: id download failed");
  std::vector<double> buf(n);
  auto down3 = [&](double* const* src, std::vector<double>& dst) -> bool {
    dst.assign(3 * n, 0.);
    for (int a = 0; a < 3; ++a) {
      if (hipMemcpy(buf.data(), src[a], sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) return false;
      for (size_t i = 0; i < n; ++i) dst[3 * (size_t)(ids[i] - 1) + a] = buf[i];
    }
    return true;
  };
  if (s->cfg.scheme == GRAINS_ADAMS_BASHFORTH2 && n) {
    double* const* src[4] = {s->P.v_nm2, s->P.wb_nm2, s->P.dUdt_nm2, s->P.dWdt_nm2};
    for (int k = 0; k < 4; ++k)
      if (!down3(src[k], h.ab[k])) return fail(s, GRAINS_DEVICE_ERROR, "reload: AB2 download failed");
  }
  h.mem.assign(n, {});
  const dem::ContactMemory& cm = s->mem;
  if (cm.cap && n) {
    const int c = cm.cur;
    std::vector<int> cnt(n), key((size_t)cm.cap * n);
    std::vector<double> f[9];
    bool ok = hipMemcpy(cnt.data(), cm.n[c], sizeof(int) * n, hipMemcpyDeviceToHost) == hipSuccess &&
              hipMemcpy(key.data(), cm.key[c], sizeof(int) * key.size(), hipMemcpyDeviceToHost) == hipSuccess;
    for (int a = 0; a < 3 && ok; ++a) {
      const double* srcs[3] = {cm.kd[c][a], cm.pn[c][a], cm.sr[c][a]};
      for (int k = 0; k < 3 && ok; ++k) {
        f[3 * k + a].resize(key.size());
        ok = hipMemcpy(f[3 * k + a].data(), srcs[k], sizeof(double) * key.size(), hipMemcpyDeviceToHost) ==
             hipSuccess;
      }
    }
    if (!ok) return fail(s, GRAINS_DEVICE_ERROR, "reload: contact memory download failed");
    for (size_t si = 0; si < n; ++si) {
      auto& m = h.mem[(size_t)(ids[si] - 1)];
      for (int j = 0; j < std::min(cnt[si], cm.cap); ++j) {
        const size_t e2 = (size_t)j * n + si;
        m.emplace_back(key[e2], d3{f[0][e2], f[1][e2], f[2][e2]}, d3{f[3][e2], f[4][e2], f[5][e2]},
                       d3{f[6][e2], f[7][e2], f[8][e2]});
      }
      // Component::m_contactMap is ordered by (1, partner id, 1)
      std::sort(m.begin(), m.end(),
                [](const std::tuple<int, d3, d3, d3>& a, const std::tuple<int, d3, d3, d3>& b) {
                  return std::get<0>(a) < std::get<0>(b);
                });
    }
  }
  return GRAINS_OK;
}

// shape description of a class as ConvexBuilderFactory writes it
// (Sphere.cpp:162-165, Box.cpp:286-289, Polyhedron.cpp:300-304)
std::string convex_text(grains_gpu_system* s, int cls_shape, double radius) {
  if (cls_shape < 0) return "*Sphere " + dflt(radius) + " *END";
  const dem::ShapeDesc& d = s->shapes[(size_t)cls_shape];
  if (d.type == dem::kShapeSphere) return "*Sphere " + dflt(d.radius) + " *END";
  if (d.type == dem::kShapeBox)
    return "*Box " + dflt(d.extent[0]) + " " + dflt(d.extent[1]) + " " + dflt(d.extent[2]) + " *END";
  std::string name = (size_t)cls_shape < s->shape_files.size() && !s->shape_files[(size_t)cls_shape].empty()
                         ? s->shape_files[(size_t)cls_shape]
                         : "polyhedron" + std::to_string(cls_shape) + ".insert";
  return "*Polyhedron " + file_part(name) + " *END";
}

std::string mat_name(grains_gpu_system* s, int m) {
  if ((size_t)m < s->mat_names.size() && !s->mat_names[(size_t)m].empty()) return s->mat_names[(size_t)m];
  return "mat" + std::to_string(m);
}

void write_le(std::ofstream& f, const void* p, size_t n) { f.write(reinterpret_cast<const char*>(p), (std::streamsize)n); }

}  // namespace

extern "C" {

int grains_gpu_save_reload(grains_gpu_system* s, const char* root_c, int hybrid) {
  if (!s || !s->initialized || !root_c || !*root_c) return GRAINS_ARGUMENT_ERROR;
  const std::string root(root_c);
  HostState h;
  int e = download(s, h);
  if (e) return e;
  const size_t n = (size_t)s->nparticles;
  const std::string suffix = (s->reload_count % 2 == 0) ? "A" : "B";
  const std::string reload = root + suffix;

  // contact models, first save only (ContactBuilderFactory::save)
  const std::string matfile = root + "_MatContact.xml";
  if (s->reload_count == 0) {
    std::ofstream mx(matfile);
    if (!mx) return fail(s, GRAINS_ARGUMENT_ERROR, "cannot write " + matfile);
    mx << "<?xml version=\"1.0\" encoding=\"UTF-8\" standalone=\"no\" ?>\n<GRAINS>\n\n  <Materials>\n";
    std::vector<int> used(dem::kMaxMaterials, 0);
    for (int m : s->h_mat) used[(size_t)m] = 1;
    for (auto& o : s->obstacles) used[(size_t)o.mat] = 1;
    for (int m = 0; m < dem::kMaxMaterials; ++m)
      if (used[(size_t)m]) mx << "    <Material value=\"" << (1 << m) << "\">" << mat_name(s, m) << "</Material>\n";
    mx << "  </Materials>\n\n  <ContactForceModels>\n";
    for (int a = 0; a < dem::kMaxMaterials; ++a)
      for (int b = a; b < dem::kMaxMaterials; ++b) {
        const dem::HookeModel& m = s->models[a * dem::kMaxMaterials + b];
        if (!m.defined || !used[(size_t)a] || !used[(size_t)b]) continue;
        mx << "    <ContactForceModel type=\"" << (m.memory ? 1 : 0) << "\" value=\"" << ((1 << a) | (1 << b))
           << "\">\n      <Parameters>\n";
        std::map<std::string, double> p;  // ordered by name, as the reference's map
        p["kn"] = m.kn;
        p["en"] = m.en;
        p["etat"] = m.etat;
        p["muc"] = m.muc;
        if (m.memory) {
          p["kt"] = m.kt;
          p["mur"] = m.mur;
          p["etarpf"] = m.etarpf;
          p["epst"] = m.epst;
        } else {
          p["kr"] = m.kr;
        }
        for (auto& kv : p)
          mx << "        <Parameter name=\"" << kv.first << "\">" << dflt(kv.second) << "</Parameter>\n";
        mx << "      </Parameters>\n    </ContactForceModel>\n";
      }
    mx << "  </ContactForceModels>\n\n</GRAINS>\n";
  }

  // particle types: the class of the input file (the RawData particle type)
  // while each class holds one convex; a system built particle by particle
  // with several convexes in one class gets a type per (class, convex) pair
  std::vector<int> cls(n, 0);
  {
    auto class_of = [&](size_t i) { return i < s->h_class.size() ? s->h_class[i] : 0; };
    auto shape_of = [&](size_t i) { return s->h_shape.empty() ? -1 : s->h_shape[i]; };
    std::map<std::pair<int, int>, int> key;
    std::map<int, int> shape_of_class;
    bool one_convex_per_class = true;
    for (size_t i = 0; i < n; ++i) {
      key[{class_of(i), shape_of(i)}] = 0;
      auto it = shape_of_class.emplace(class_of(i), shape_of(i)).first;
      if (it->second != shape_of(i)) one_convex_per_class = false;
    }
    int t = 0;
    for (auto& kv : key) kv.second = t++;
    for (size_t i = 0; i < n; ++i) cls[i] = one_convex_per_class ? class_of(i) : key[{class_of(i), shape_of(i)}];
  }
  int ntypes = 0;
  for (int c : cls) ntypes = std::max(ntypes, c + 1);
  std::vector<size_t> first(ntypes, n);
  for (size_t i = 0; i < n; ++i)
    if (first[(size_t)cls[i]] == n) first[(size_t)cls[i]] = i;

  std::ofstream f(reload);
  if (!f) return fail(s, GRAINS_ARGUMENT_ERROR, "cannot write " + reload);
  f << "__Format2024__\n#Time " << dflt(s->time) << "\n";
  f << "<ContactForceModels> " << file_part(matfile) << " </ContactForceModels>\n";
  f << "\nNumberOfParticleTypes " << ntypes << "\n";
  for (int t = 0; t < ntypes; ++t) {
    const size_t i = first[(size_t)t];
    f << "\n<Particle>\n";
    if (i == n) {  // a class without particles
      f << "*IDnumber -1\n*Material " << mat_name(s, 0) << "\n*Sphere 0 *END\n*CrustThickness 0\n\n*Type " << t
        << "\n*Tag 0\n</Particle>\n";
      continue;
    }
    const int sh = s->h_shape.empty() ? -1 : s->h_shape[i];
    const double I0 = s->h_inertia[i];
    double I[6] = {I0, 0., 0., I0, 0., I0}, Iinv[6] = {1. / I0, 0., 0., 1. / I0, 0., 1. / I0};
    if (sh >= 0 && !s->h_I3.empty() && s->shapes[(size_t)sh].type != dem::kShapeSphere) {
      I[0] = s->h_I3[3 * i];
      I[3] = s->h_I3[3 * i + 1];
      I[5] = s->h_I3[3 * i + 2];
      Iinv[0] = 1. / I[0];
      Iinv[3] = 1. / I[3];
      Iinv[5] = 1. / I[5];
    }
    f << "*IDnumber -1\n*Material " << mat_name(s, s->h_mat[i]) << "\n"
      << convex_text(s, sh, s->h_rad[i]) << "\n*CrustThickness " << dflt(s->h_crust[i]) << "\n\n";
    f << "*Type " << t << "\n*Tag 0\n*Mass " << s16(s->h_mass[i]) << "\n*Density "
      << s16(i < s->h_density.size() ? s->h_density[i] : 0.) << "\n";
    f << "*MomentOfInertiaTensor\n";
    for (int k = 0; k < 6; ++k) f << s16(I[k]) << (k < 5 ? " " : "\n");
    f << "*InverseOfMOIT\n";
    for (int k = 0; k < 6; ++k) f << s16(Iinv[k]) << (k < 5 ? " " : "\n");
    f << "*Position 0\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n*Activity 1\nParticleKinematics3D\n0 0 0\n1 0 0 0\n0 0 0\n";
    f << "</Particle>\n";
  }
  f << "\nRemainingNbParticlesToInsertPerType";
  for (int t = 0; t < ntypes; ++t) f << " 0";
  f << "\n\nRemainingKnownInsertionPositions 0\n";
  // obstacles (CompositeObstacle::write, SimpleObstacle::write)
  f << "\n<Obstacle>\n<Composite> obstacles CompositeObstacle\n";
  for (size_t k = 0; k < s->obstacles.size(); ++k) {
    const dem::BoxObstacle& o = s->obstacles[k];
    f << "<Simple> " << (k < s->obs_names.size() ? s->obs_names[k] : "obstacle" + std::to_string(k)) << " " << o.id
      << "\n*Torsor " << g3(o.centre) << " 0 0 0 0 0 0\n*Material " << mat_name(s, o.mat) << "\n*IDnumber " << o.id
      << "\n*Material " << mat_name(s, o.mat) << "\n*Box " << dflt(o.extent[0]) << " " << dflt(o.extent[1]) << " "
      << dflt(o.extent[2]) << " *END\n*CrustThickness " << dflt(o.crust) << "\n\n*ToFluid "
      << (k < s->obs_to_fluid.size() ? s->obs_to_fluid[k] : 0) << "\n*Position 0\n"
      << g3(o.centre) << "\n" << g3(o.B) << "\n" << g3(o.B + 3) << "\n" << g3(o.B + 6) << "\n\n 0\n</Simple>\n";
  }
  f << "</Composite>\n</Obstacle>\n\n";
  f << "NbFiles 1\n" << (hybrid ? "Hybrid " : "Text ") << n << "\n#EndTime\n";
  f.close();
  if (!f) return fail(s, GRAINS_ARGUMENT_ERROR, "error writing " + reload);

  // particle records (Particle::write2014 / write2014_binary)
  const std::string pfile = particle_file(reload, hybrid != 0);
  const bool ab2 = s->cfg.scheme == GRAINS_ADAMS_BASHFORTH2, lf = s->cfg.scheme == GRAINS_LEAPFROG;
  if (hybrid) {
    std::ofstream b(pfile, std::ios::binary);
    if (!b) return fail(s, GRAINS_ARGUMENT_ERROR, "cannot write " + pfile);
    for (size_t i = 0; i < n; ++i) {
      const size_t nm = h.mem[i].size();
      size_t nb = 3 * sizeof(int) + sizeof(unsigned) + (12 * sizeof(double) + sizeof(unsigned)) +
                  10 * sizeof(double) + (lf ? 6 : ab2 ? 12 : 0) * sizeof(double) + sizeof(int) +
                  nm * (4 * sizeof(int) + 9 * sizeof(double));
      write_le(b, &nb, sizeof nb);
      int gt = cls[i], id = (int)i + 1, tag = 0;
      unsigned act = 1, ttype = 0;
      write_le(b, &gt, 4);
      write_le(b, &id, 4);
      write_le(b, &tag, 4);
      write_le(b, &act, 4);
      write_le(b, &h.x[3 * i], 24);
      write_le(b, &h.B[9 * i], 72);
      write_le(b, &ttype, 4);
      write_le(b, &h.v[3 * i], 24);
      write_le(b, &h.q[4 * i + 3], 8);  // w first (Quaternion::writeQuaternion_binary)
      write_le(b, &h.q[4 * i], 24);
      write_le(b, &h.w[3 * i], 24);
      if (lf) {
        write_le(b, &h.F[3 * i], 24);
        write_le(b, &h.T[3 * i], 24);
      } else if (ab2) {
        for (int k = 0; k < 4; ++k) write_le(b, &h.ab[k][3 * i], 24);
      }
      int ms = (int)nm;
      write_le(b, &ms, 4);
      for (auto& m : h.mem[i]) {
        int ids3[4] = {1, std::get<0>(m), 1, 1};
        write_le(b, ids3, 16);
        const d3 vs[3] = {std::get<1>(m), std::get<2>(m), std::get<3>(m)};
        for (const d3& v : vs) {
          double t3[3] = {v.x, v.y, v.z};
          write_le(b, t3, 24);
        }
      }
    }
    b.close();
    if (!b) return fail(s, GRAINS_ARGUMENT_ERROR, "error writing " + pfile);
  } else {
    std::ofstream t(pfile);
    if (!t) return fail(s, GRAINS_ARGUMENT_ERROR, "cannot write " + pfile);
    for (size_t i = 0; i < n; ++i) {
      t << cls[i] << " " << i + 1 << " 0 1 " << g3(&h.x[3 * i]) << " " << g3(&h.B[9 * i]) << " "
        << g3(&h.B[9 * i + 3]) << " " << g3(&h.B[9 * i + 6]) << " 0 " << g3(&h.v[3 * i]) << " " << s16(h.q[4 * i + 3])
        << " " << g3(&h.q[4 * i]) << " " << g3(&h.w[3 * i]);
      if (lf) t << " " << g3(&h.F[3 * i]) << " " << g3(&h.T[3 * i]);
      else if (ab2)
        for (int k = 0; k < 4; ++k) t << " " << g3(&h.ab[k][3 * i]);
      t << " " << h.mem[i].size();
      for (auto& m : h.mem[i]) {
        const d3 vs[3] = {std::get<1>(m), std::get<2>(m), std::get<3>(m)};
        t << " 1 " << std::get<0>(m) << " 1 1";
        for (const d3& v : vs) {
          double t3[3] = {v.x, v.y, v.z};
          t << " " << g3(t3);
        }
      }
      t << "\n";
    }
    t.close();
    if (!t) return fail(s, GRAINS_ARGUMENT_ERROR, "error writing " + pfile);
  }
  std::ofstream rft(root + "_RFTable.txt", std::ios::app);
  rft << dflt(s->time) << " " << reload << "\n";
  ++s->reload_count;
  return GRAINS_OK;
}

int grains_gpu_load_reload(grains_gpu_system* s, const char* file_c) {
  if (!s || !s->initialized || !file_c) return GRAINS_ARGUMENT_ERROR;
  const std::string file(file_c);
  std::ifstream f(file);
  if (!f) return fail(s, GRAINS_ARGUMENT_ERROR, "cannot open reload file " + file);
  std::string tok;
  f >> tok;
  if (tok != "__Format2024__") return fail(s, GRAINS_CONFIG_ERROR, file + " is not a __Format2024__ reload file");
  double t = 0.;
  f >> tok >> t;
  // NbFiles / Text|Hybrid n
  std::string mode;
  size_t nrec = 0;
  while (f >> tok)
    if (tok == "NbFiles") {
      int nf = 0;
      f >> nf >> mode >> nrec;
      if (nf != 1) return fail(s, GRAINS_CONFIG_ERROR, "reload files written by several processes are not supported");
      break;
    }
  if (mode != "Text" && mode != "Hybrid") return fail(s, GRAINS_CONFIG_ERROR, file + ": no NbFiles record");
  f >> tok;
  if (tok != "#EndTime") return fail(s, GRAINS_CONFIG_ERROR, file + ": missing #EndTime");
  const size_t n = (size_t)s->nparticles;
  if (nrec != n)
    return fail(s, GRAINS_CONFIG_ERROR,
                file + " holds " + std::to_string(nrec) + " particles, the system " + std::to_string(n));
  HostState h;
  int e = download(s, h);  // current values of what a record may not carry
  if (e) return e;
  std::vector<int> seen(n, 0);
  const bool lf = s->cfg.scheme == GRAINS_LEAPFROG, ab2 = s->cfg.scheme == GRAINS_ADAMS_BASHFORTH2;
  const std::string pfile = particle_file(file, mode == "Hybrid");
  auto bad = [&](const std::string& why) { return fail(s, GRAINS_CONFIG_ERROR, pfile + ": " + why); };
  if (mode == "Hybrid") {
    std::ifstream b(pfile, std::ios::binary);
    if (!b) return bad("cannot open");
    for (size_t r = 0; r < n; ++r) {
      size_t nb = 0;
      int gt, id, tag, ms;
      unsigned act, ttype;
      double x[3], B[9], v[3], qw, qv[3], w[3];
      b.read(reinterpret_cast<char*>(&nb), 8);
      b.read(reinterpret_cast<char*>(&gt), 4);
      b.read(reinterpret_cast<char*>(&id), 4);
      b.read(reinterpret_cast<char*>(&tag), 4);
      b.read(reinterpret_cast<char*>(&act), 4);
      b.read(reinterpret_cast<char*>(x), 24);
      b.read(reinterpret_cast<char*>(B), 72);
      b.read(reinterpret_cast<char*>(&ttype), 4);
      b.read(reinterpret_cast<char*>(v), 24);
      b.read(reinterpret_cast<char*>(&qw), 8);
      b.read(reinterpret_cast<char*>(qv), 24);
      b.read(reinterpret_cast<char*>(w), 24);
      if (!b || id < 1 || (size_t)id > n || seen[(size_t)id - 1]) return bad("bad record " + std::to_string(r));
      const size_t i = (size_t)id - 1;
      seen[i] = 1;
      std::copy(x, x + 3, &h.x[3 * i]);
      std::copy(B, B + 9, &h.B[9 * i]);
      std::copy(v, v + 3, &h.v[3 * i]);
      std::copy(qv, qv + 3, &h.q[4 * i]);
      h.q[4 * i + 3] = qw;
      std::copy(w, w + 3, &h.w[3 * i]);
      if (lf) {
        b.read(reinterpret_cast<char*>(&h.F[3 * i]), 24);
        b.read(reinterpret_cast<char*>(&h.T[3 * i]), 24);
      } else if (ab2) {
        for (int k = 0; k < 4; ++k) b.read(reinterpret_cast<char*>(&h.ab[k][3 * i]), 24);
      }
      b.read(reinterpret_cast<char*>(&ms), 4);
      if (!b || ms < 0) return bad("bad record " + std::to_string(r));
      h.mem[i].clear();
      for (int k = 0; k < ms; ++k) {
        int ids3[4];
        double kd[3], pn[3], sr[3];
        b.read(reinterpret_cast<char*>(ids3), 16);
        b.read(reinterpret_cast<char*>(kd), 24);
        b.read(reinterpret_cast<char*>(pn), 24);
        b.read(reinterpret_cast<char*>(sr), 24);
        h.mem[i].emplace_back(ids3[1], d3{kd[0], kd[1], kd[2]}, d3{pn[0], pn[1], pn[2]}, d3{sr[0], sr[1], sr[2]});
      }
      if (!b) return bad("truncated record " + std::to_string(r));
    }
  } else {
    std::ifstream tf(pfile);
    if (!tf) return bad("cannot open");
    for (size_t r = 0; r < n; ++r) {
      int gt, id, tag, act, ttype, ms;
      double x[3], B[9], v[3], qw, qv[3], w[3];
      tf >> gt >> id >> tag >> act;
      for (double& a : x) tf >> a;
      for (double& a : B) tf >> a;
      tf >> ttype;
      for (double& a : v) tf >> a;
      tf >> qw;
      for (double& a : qv) tf >> a;
      for (double& a : w) tf >> a;
      if (!tf || id < 1 || (size_t)id > n || seen[(size_t)id - 1]) return bad("bad record " + std::to_string(r));
      const size_t i = (size_t)id - 1;
      seen[i] = 1;
      std::copy(x, x + 3, &h.x[3 * i]);
      std::copy(B, B + 9, &h.B[9 * i]);
      std::copy(v, v + 3, &h.v[3 * i]);
      std::copy(qv, qv + 3, &h.q[4 * i]);
      h.q[4 * i + 3] = qw;
      std::copy(w, w + 3, &h.w[3 * i]);
      if (lf) {
        for (int a = 0; a < 3; ++a) tf >> h.F[3 * i + a];
        for (int a = 0; a < 3; ++a) tf >> h.T[3 * i + a];
      } else if (ab2) {
        for (int k = 0; k < 4; ++k)
          for (int a = 0; a < 3; ++a) tf >> h.ab[k][3 * i + a];
      }
      tf >> ms;
      if (!tf || ms < 0) return bad("bad record " + std::to_string(r));
      h.mem[i].clear();
      for (int k = 0; k < ms; ++k) {
        int i0, pid, i2, on;
        double a9[9];
        tf >> i0 >> pid >> i2 >> on;
        for (double& a : a9) tf >> a;
        h.mem[i].emplace_back(pid, d3{a9[0], a9[1], a9[2]}, d3{a9[3], a9[4], a9[5]}, d3{a9[6], a9[7], a9[8]});
      }
      if (!tf) return bad("truncated record " + std::to_string(r));
    }
  }
  // upload: bases first (set_state bins the positions), the kinematics, then
  // the torsor and the scheme data the records carry
  if (s->has_shapes() && (e = grains_gpu_set_basis(s, h.B.data()))) return e;
  if ((e = grains_gpu_set_state(s, h.x.data(), h.v.data(), h.w.data(), h.q.data(), 0))) return e;
  std::vector<int> ids(n);
  if (n && hipMemcpy(ids.data(), s->P.id, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(s, GRAINS_DEVICE_ERROR, "reload: id download failed");
  std::vector<double> buf(n);
  auto up3 = [&](double* const* dst, const std::vector<double>& src) -> bool {
    for (int a = 0; a < 3; ++a) {
      for (size_t si = 0; si < n; ++si) buf[si] = src[3 * (size_t)(ids[si] - 1) + a];
      if (hipMemcpy(dst[a], buf.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) return false;
    }
    return true;
  };
  if (n && !(up3(s->P.F, h.F) && up3(s->P.T, h.T)))
    return fail(s, GRAINS_DEVICE_ERROR, "reload: torsor upload failed");
  if (ab2 && n) {
    double* const* dst[4] = {s->P.v_nm2, s->P.wb_nm2, s->P.dUdt_nm2, s->P.dWdt_nm2};
    for (int k = 0; k < 4; ++k)
      if (!up3(dst[k], h.ab[k])) return fail(s, GRAINS_DEVICE_ERROR, "reload: AB2 upload failed");
  }
  dem::ContactMemory& cm = s->mem;
  if (cm.cap && n) {
    const int c = cm.cur;
    std::vector<int> cnt(n, 0), key((size_t)cm.cap * n, 0);
    std::vector<double> fv[9];
    for (auto& a : fv) a.assign(key.size(), 0.);
    for (size_t si = 0; si < n; ++si) {
      const auto& m = h.mem[(size_t)(ids[si] - 1)];
      if ((int)m.size() > cm.cap) return fail(s, GRAINS_DEVICE_ERROR, "reload: more contacts with memory than kept");
      cnt[si] = (int)m.size();
      for (size_t j = 0; j < m.size(); ++j) {
        const size_t e2 = j * n + si;
        key[e2] = std::get<0>(m[j]);
        const d3 vs[3] = {std::get<1>(m[j]), std::get<2>(m[j]), std::get<3>(m[j])};
        for (int k = 0; k < 3; ++k) {
          fv[3 * k][e2] = vs[k].x;
          fv[3 * k + 1][e2] = vs[k].y;
          fv[3 * k + 2][e2] = vs[k].z;
        }
      }
    }
    bool ok = hipMemcpy(cm.n[c], cnt.data(), sizeof(int) * n, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(cm.key[c], key.data(), sizeof(int) * key.size(), hipMemcpyHostToDevice) == hipSuccess;
    for (int a = 0; a < 3 && ok; ++a) {
      double* dsts[3] = {cm.kd[c][a], cm.pn[c][a], cm.sr[c][a]};
      for (int k = 0; k < 3 && ok; ++k)
        ok = hipMemcpy(dsts[k], fv[3 * k + a].data(), sizeof(double) * key.size(), hipMemcpyHostToDevice) ==
             hipSuccess;
    }
    if (!ok) return fail(s, GRAINS_DEVICE_ERROR, "reload: contact memory upload failed");
  }
  grains_gpu_set_time(s, t);
  s->drop_graphs();
  return GRAINS_OK;
}

// GrainsCoupledWithFluid::setReloadSame / <Reload Type="same">: the last
// reload file listed in <RestartFile Name>_RFTable.txt
// (GrainsExec::restartFileName_AorB, GrainsExec.cpp:455-469).
int grains_gpu_reload_same(grains_gpu_system* s) {
  if (!s || !s->initialized) return GRAINS_ARGUMENT_ERROR;
  std::ifstream rft(s->save_root + "_RFTable.txt");
  std::string t, rf, file;
  while (rft >> t >> rf) file = rf;
  if (file.empty()) return fail(s, GRAINS_CONFIG_ERROR, "no reload file listed in " + s->save_root + "_RFTable.txt");
  return grains_gpu_load_reload(s, file.c_str());
}

}  // extern "C"
