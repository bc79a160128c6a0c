// XML front-end entry used by the C ABI and by the fluid-coupling interface.
#pragma once
#include "../../include/grains_gpu.h"

namespace dem {

struct XmlOptions {
  bool coupled = false;          // GrainsCoupledWithFluid input (TimeStep mindt maxdt ndt)
  double fluid_density = 0.;     // fluid density of a coupled run
  bool track_basis = false;      // keep every particle's Transform basis on the device
  bool fluid_corrected = true;   // (1 - rho_f / rho) acceleration correction
  bool no_init = false;          // stop before initialisation (the construction data only)
};

int create_from_xml(const char* xml_path, int precision, const XmlOptions& opt, grains_gpu_system** out);

}  // namespace dem
