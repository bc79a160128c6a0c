/*
 * InterfaceGrains3DBasilisk.h -- the C interface through which Basilisk's
 * DLMFD solver drives the granular solver, provided by libgrains_gpu.so
 * with the signatures of GRAINS_V1/Grains/InterfaceCBasilisk/include/
 * InterfaceGrains3DBasilisk.h:12-32 (implementation replaced:
 * InterfaceGrains3DBasilisk.cpp:15-139; caller Octree/DLMFD/DLMFD_Grains3D.h:
 * 34-169).  One system per process; not re-entrant; called from one thread.
 */
#ifndef INTERFACE_GRAINS3D_BASILISK_H
#define INTERFACE_GRAINS3D_BASILISK_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Builds the system from a coupled Grains3D input file (<TimeStep mindt
 * maxdt ndt>) with fluid density fluid_rho; b_fluidcorrectedacc = false
 * drops the (1 - rho_f / rho) acceleration correction
 * (GrainsCoupledWithFluid::setFluidCorrectedAcceleration).  Reload files
 * (b_restart) are not supported: the process exits with status 1, as on any
 * configuration error (Grains::grainsAbort). */
void Init_Grains(char const* inputfile, double fluid_rho, const bool b_restart, const bool b_fluidcorrectedacc);

/* Advances the granular system over one fluid time step in
 * max(min(maxdt, dt_fluid / ndt), mindt) sub-steps
 * (GrainsCoupledWithFluid::Simulation, GrainsCoupledWithFluid.cpp:123-232) */
void Simu_Grains(const double dt_fluid);

/* Rigid-body features for the fluid solver in GrainsToFluid's text layout
 * (GrainsCoupledWithFluid.cpp:912-1120), tokens separated by single spaces;
 * *pstrsize receives the buffer size.  The caller releases the buffer with
 * free(). */
char* GrainsToBasilisk(int* pstrsize);

void SetInitialTime(double tinit);

/* Post-processing writers are not part of this library: prints a notice
 * once (use grains_gpu_get_state for output). */
void SaveResults_Grains(void);
void checkParaviewPostProcessing_Grains(char* solid_resDir);

/* Sets translational and angular velocities from m rows of (vx, vy, vz, wx,
 * wy, wz) indexed by particle id - 1 (updateParticlesVelocity,
 * GrainsCoupledWithFluid.cpp:1126-1181). */
void UpdateVelocityGrains(double** arrayv, const int m);

void SetInitialCycleNumber(int cycle0);

/* Number of particles and of obstacles sent to the fluid (<Status ToFluid="1"/>) */
void NumberOfRigidBodiesInBasilisk(size_t* nparticles, size_t* nobstacles);

#ifdef __cplusplus
}
#endif

#endif /* INTERFACE_GRAINS3D_BASILISK_H */
