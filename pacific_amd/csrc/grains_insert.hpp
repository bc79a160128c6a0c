// Particle insertion of the XML front-end (Grains::insertParticle,
// Grains.cpp:2070-2140): positions from a list or random points of insertion
// windows, random or ordered choice of the particle class, random initial
// angular positions, initial velocities, and the test that the new particle
// touches no particle or obstacle (LinkedCell::insertParticleSerial,
// LinkedCell.cpp:1765-1860).  Host code; at the initial time it fills the
// construction data, over time it grows an initialised system.
#pragma once
#include <cstddef>
#include <unordered_map>
#include <vector>

struct grains_gpu_system;

namespace grains_insert {

// Window::readWindow / getInsertionPoint (Window.cpp:53-320)
enum WindowType { kBox = 0, kCylinder = 1, kAnnulus = 2, kLine = 3 };
struct Window {
  int type = kBox;
  double A[3] = {0., 0., 0.}, B[3] = {0., 0., 0.};  // corners / bottom centre, line ends
  double radius = 0., radius_int = 0., height = 0.;
  int axis = 2;  // cylinder and annulus axis: 0 X, 1 Y, 2 Z
};

// A <Particle> class waiting for insertion
struct Class {
  int shape = -1;  // shape class of the system (-1: sphere of a sphere-only system)
  int xml_class = 0;  // index of its <Particle> (RawData type)
  double radius = 0., density = 0., crust = 0.;
  int mat = 0;
  double B[9] = {1., 0., 0., 0., 1., 0., 0., 0., 1.};  // <AngularPosition> of the class
  long remaining = 0;
};

struct Plan {
  bool over_time = false;  // <Mode Type="OverTime">
  bool random_order = false;  // <Order Type="Random">
  bool random_angle = false;  // <InitialAngularPosition Type="Random">
  // <InitialAngularPosition Type="File">: one matrix (9, by rows) per
  // particle, taken in file order
  std::vector<double> angles;
  size_t angle_next = 0;
  bool force = false;  // <ForceInsertion Value="True">
  bool bv_only = false;  // <GeometricOverlap Value="BV">
  size_t frequency = 1;  // <Frequency TryEvery>
  std::vector<Window> windows;
  // insertion positions; ordered insertion takes the last one first (the
  // list is stored reversed)
  std::vector<double> positions;
  int vmode = 0;  // 0 zero, 1 constant, 2 random (Grains::computeInitVit)
  double vconst[3] = {0., 0., 0.}, wconst[3] = {0., 0., 0.}, amp_t = 0., amp_r = 0.;
  std::vector<Class> classes;
  // the particle waiting for insertion: its class (-1: none picked yet) and
  // its basis, to which every attempt composes a new random rotation
  int wait = -1;
  double wait_B[9];
  size_t counter = 0;  // calls since the last attempt (insert_counter)
  // the attempt of the next step was made before initialisation
  bool preattempted = false;
  long attempts = 0, inserted = 0;
  // particles in place before initialisation, binned on cells of edge at
  // least the largest circumscribed diameter (per axis: h[a], n[a] cells
  // on a periodic axis), so that the overlap test of an initial-time fill
  // visits 27 cells
  double h[3] = {0., 0., 0.};
  long ncell[3] = {0, 0, 0};
  std::unordered_map<long long, std::vector<int>> cells;

  long pending() const {
    long n = 0;
    for (const Class& c : classes) n += c.remaining;
    return n;
  }
};

// One call of Grains::insertParticle.  Before initialisation the particle is
// appended to the construction data; afterwards the system grows by one
// particle.  *inserted tells whether a particle went in.
int attempt(grains_gpu_system* s, Plan& p, bool* inserted);

// The insertion call of one time step (Grains.cpp:322-327), unless the
// attempt of that step was made before initialisation.
int step_attempt(grains_gpu_system* s, Plan& p, bool* inserted);

// Insertion at the initial time: attempts until every particle is in
// (Grains.cpp:2004-2017), at most max_attempts in a row without success.
int insert_all(grains_gpu_system* s, Plan& p, long max_attempts);

}  // namespace grains_insert
