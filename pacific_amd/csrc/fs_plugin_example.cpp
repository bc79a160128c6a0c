// A fluid solver's side of the Cartesian coupling, reduced to what
// FS_Grains3DPlugIn.cc:10-205 does with Grains3D: build the coupled system,
// start the post-processing, then per fluid step hand over the fluid's
// force and torque on every particle, advance the particles by one fluid
// step, read them back as GrainsToFluid text; a last post-processing frame.
// Every call goes through include/GrainsCoupledWithFluid.hh.
//
//   fs_plugin_example <coupled.xml> <fluid density> <fluid steps> <dt_fluid> <Fx> <Fy> <Fz>
//
// The force (Fx, Fy, Fz) acts on particle 1 and no other; the program prints
// the GrainsToFluid text of the last step on stdout.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/GrainsCoupledWithFluid.hh"

int main(int argc, char** argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: %s <coupled.xml> <fluid density> <fluid steps> <dt_fluid> <Fx> <Fy> <Fz>\n", argv[0]);
    return 2;
  }
  const double rho = std::atof(argv[2]), dt = std::atof(argv[4]);
  const int nsteps = std::atoi(argv[3]);
  GrainsCoupledWithFluid* g = GrainsCoupledWithFluid::create(argv[1], rho);
  g->setFluidCorrectedAcceleration(true);
  g->do_before_time_stepping();
  g->InitialPostProcessing();
  size_t np = 0, nobs = 0;
  g->numberOfRBToFluid(&np, &nobs);
  std::vector<std::vector<double> > hydro(np, std::vector<double>(6, 0.));
  for (int a = 0; a < 3; ++a) hydro[0][a] = std::atof(argv[5 + a]);
  std::istringstream is;
  for (int k = 0; k < nsteps; ++k) {
    g->updateParticlesHydroFT(&hydro);
    g->Simulation(dt);
    g->GrainsToFluid(is);
  }
  g->doPostProcessing();
  const bool failed = g->failed();
  std::cout << is.str();
  delete g;
  return failed ? 1 : 0;
}
