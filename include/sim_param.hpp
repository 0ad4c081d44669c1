// SimParam and CalculationMethod: the reference's parameter interface
// (src/sim_param.hpp:8-39), kept name for name and field for field so that a
// driver written against the reference compiles unchanged.
#pragma once

#include <cstddef>
#include <cstdlib>
#include <string>

enum class CalculationMethod { BRANCH, PREDICATED };

// Parses "BRANCH" / "PREDICATED"; anything else throws std::invalid_argument
// (src/sim_param.cpp:25-38).
CalculationMethod getCalculationMethod(const std::string &method);

class SimParam {
public:
  // Defaults of src/sim_param.cpp:12-22.
  SimParam();

  // Positional arguments, src/sim_param.cpp:40-67:
  //   numParticles/256 simIterationsPerFrame damping dt distEps G numFrames
  //   gwSize calcMethod
  void parseArgs(int argc, char **argv);

  float G;                   ///< gravitational parameter
  float dt;                  ///< time step
  size_t numParticles;       ///< number of particles
  size_t numFrames;          ///< frames to simulate
  int simIterationsPerFrame; ///< kernel launches per frame
  float damping;             ///< velocity drag (1.0 = none)
  float distEps;             ///< softening added to the squared distance
  int gwSize;                ///< work-group size (a hint: shapes are tuned)
  CalculationMethod calcMethod; ///< BRANCH or PREDICATED
};
