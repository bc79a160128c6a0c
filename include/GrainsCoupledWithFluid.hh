/* GrainsCoupledWithFluid: the calls the Cartesian solver's plug-in makes on a
 * Grains3D coupled with a fluid (Cartesian/FLUID/FluidSolidInterface/src/
 * FS_Grains3DPlugIn.cc:10-205 on GRAINS_V1/Grains/Fluid/include/
 * GrainsCoupledWithFluid.hh), as a header-only C++ class over the C ABI of
 * libgrains_gpu.so (grains_gpu.h).  The plug-in keeps its call sequence:
 *
 *   GrainsCoupledWithFluid* g = GrainsCoupledWithFluid::create(file, rho_f);
 *   g->setFluidCorrectedAcceleration(correct);      // before the next call
 *   g->do_before_time_stepping();
 *   g->Simulation(dt_fluid);                        // each fluid step
 *   g->GrainsToFluid(is);                           // istringstream
 *   g->updateParticlesHydroFT(&hydroFT);            // vector<vector<double>>
 *   g->InitialPostProcessing(); g->doPostProcessing();
 *
 * Differences from the reference class: the XML input is read by the library
 * from the file name (no Xerces DOM node is passed).  Errors in the input
 * file end the process with exit(1) after a message, as Grains::grainsAbort
 * does (Grains.cpp:2533-2536); errors raised
 * during a fluid step print a message and stop the sub-cycling, as the
 * reference's Simulation does (GrainsCoupledWithFluid.cpp:201-230). */
#ifndef GRAINS_COUPLED_WITH_FLUID_HH
#define GRAINS_COUPLED_WITH_FLUID_HH

#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include "grains_gpu.h"

class GrainsCoupledWithFluid {
 public:
  /* GrainsBuilderFactory::createCoupledWithFluid (GrainsBuilderFactory.cpp) */
  static GrainsCoupledWithFluid* create(std::string const& input_file, double fluid_density) {
    return new GrainsCoupledWithFluid(input_file, fluid_density);
  }
  ~GrainsCoupledWithFluid() {
    if (sys_) grains_gpu_destroy(sys_);
  }
  GrainsCoupledWithFluid(GrainsCoupledWithFluid const&) = delete;
  GrainsCoupledWithFluid& operator=(GrainsCoupledWithFluid const&) = delete;

  void setFluidCorrectedAcceleration(bool correct) { corrected_ = correct; }
  /* continue from the last reload file of <RestartFile Name>_RFTable.txt */
  void setReloadSame() { reload_same_ = true; }

  /* Builds the system from the input file (Grains::do_before_time_stepping) */
  void do_before_time_stepping() {
    if (sys_) return;
    grains_gpu_system* s = nullptr;
    int e = grains_gpu_create_coupled(file_.c_str(), rho_, corrected_ ? 1 : 0, &s);
    if (e) {
      std::string msg = s ? grains_gpu_last_error(s) : "cannot build the system";
      if (s) grains_gpu_destroy(s);
      abort_with(msg);
    }
    sys_ = s;
    if (reload_same_ && grains_gpu_reload_same(sys_)) abort_with(grains_gpu_last_error(sys_));
  }

  /* One fluid time step of sub-cycled DEM steps (GrainsCoupledWithFluid::
   * Simulation, :123-232) */
  void Simulation(double time_interval) {
    if (!sys_ || error_) return;
    if (grains_gpu_fluid_step(sys_, time_interval, nullptr)) {
      std::fprintf(stderr, "Grains: %s\n", grains_gpu_last_error(sys_));
      error_ = true;
    }
  }

  /* GrainsToFluid text into the stream (:912-1120) */
  void GrainsToFluid(std::istringstream& is) {
    char* buf = nullptr;
    int size = 0;
    if (!sys_ || grains_gpu_to_fluid(sys_, &buf, &size)) abort_with(sys_ ? grains_gpu_last_error(sys_) : "no system");
    is.str(std::string(buf));
    is.clear();
    std::free(buf);
  }

  /* Rows (Fx, Fy, Fz, Tx, Ty, Tz) by particle id - 1 (:1187-1191) */
  void updateParticlesHydroFT(std::vector<std::vector<double> > const* hydroFT) {
    std::vector<double> flat;
    flatten(*hydroFT, flat);
    if (grains_gpu_set_hydro_ft(sys_, flat.data(), hydroFT->size())) abort_with(grains_gpu_last_error(sys_));
  }

  /* Rows (vx, vy, vz, wx, wy, wz) by particle id - 1 (:1126-1181) */
  void updateParticlesVelocity(std::vector<std::vector<double> > const& velocities, bool b_set_velocity_nm1_and_diff) {
    (void)b_set_velocity_nm1_and_diff;
    std::vector<double> flat;
    flatten(velocities, flat);
    if (grains_gpu_set_velocities(sys_, flat.data(), velocities.size())) abort_with(grains_gpu_last_error(sys_));
  }

  /* Post-processing writers of the input file (:870-893) */
  void InitialPostProcessing(size_t indent_width = 0) {
    (void)indent_width;
    if (grains_gpu_post_processing(sys_, 1)) abort_with(grains_gpu_last_error(sys_));
  }
  void doPostProcessing(size_t indent_width = 0) {
    (void)indent_width;
    if (grains_gpu_post_processing(sys_, 0)) abort_with(grains_gpu_last_error(sys_));
  }
  /* The Paraview writer's directory check: there is no Paraview writer */
  void checkParaviewPostProcessing(std::string const& name, std::string const& root, bool const& is_binary) {
    (void)name;
    (void)root;
    (void)is_binary;
  }

  void setInitialCycleNumber(int cycle0) { cycle0_ = cycle0; }
  void setInitialTime(double tinit) { grains_gpu_set_time(sys_, tinit); }
  void numberOfRBToFluid(size_t* nparticles, size_t* nobstacles) const {
    grains_gpu_count_rigid_bodies(sys_, nparticles, nobstacles);
  }
  /* true once a fluid step stopped on an error raised on the device */
  bool failed() const { return error_; }
  grains_gpu_system* handle() const { return sys_; }

 private:
  GrainsCoupledWithFluid(std::string const& file, double rho) : file_(file), rho_(rho) {}
  static void flatten(std::vector<std::vector<double> > const& rows, std::vector<double>& flat) {
    flat.assign(6 * rows.size(), 0.);
    for (size_t i = 0; i < rows.size(); ++i)
      for (size_t j = 0; j < 6 && j < rows[i].size(); ++j) flat[6 * i + j] = rows[i][j];
  }
  [[noreturn]] static void abort_with(std::string const& msg) {
    std::fprintf(stderr, "Grains: %s\n", msg.c_str());
    std::exit(1);
  }
  std::string file_;
  double rho_;
  bool corrected_ = true;
  bool reload_same_ = false;
  bool error_ = false;
  int cycle0_ = 0;
  grains_gpu_system* sys_ = nullptr;
};

#endif
