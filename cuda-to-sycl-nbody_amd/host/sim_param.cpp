// SimParam defaults and positional-argument parsing: the behaviour of
// src/sim_param.cpp:12-67 (atoi/atof semantics, the batch-of-256 particle
// count, std::invalid_argument for an unknown method).

#include "sim_param.hpp"

#include "nbody_simparam.h"

#include <cstdint>
#include <stdexcept>

SimParam::SimParam() {
  G = 2.0f;
  dt = 0.005f;
  numParticles = 50 * 256;
  numFrames = SIZE_MAX;
  simIterationsPerFrame = 4;
  damping = 0.999998f;
  distEps = 1.0e-7f;
  gwSize = 64;
  calcMethod = CalculationMethod::BRANCH;
}

CalculationMethod getCalculationMethod(const std::string &method) {
  if (method == "BRANCH") return CalculationMethod::BRANCH;
  if (method == "PREDICATED") return CalculationMethod::PREDICATED;
  throw std::invalid_argument(
      "Valid calculation methods are BRANCH or PREDICATED");
}

void SimParam::parseArgs(int argc, char **argv) {
  // argv[1]: batches of 256 particles
  if (argc >= 2) numParticles = 256 * atoi(argv[1]);
  if (argc >= 3) simIterationsPerFrame = atoi(argv[2]);
  if (argc >= 4) damping = atof(argv[3]);
  if (argc >= 5) dt = atof(argv[4]);
  if (argc >= 6) distEps = atof(argv[5]);
  if (argc >= 7) G = atof(argv[6]);
  if (argc >= 8) numFrames = atoi(argv[7]);
  if (argc >= 9) gwSize = atoi(argv[8]);
  if (argc >= 10) calcMethod = getCalculationMethod(argv[9]);
}

extern "C" int nb_simparam_parse(int argc, char **argv, nb_params *out) {
  if (!out) return NB_ERR_INVALID_ARG;
  SimParam p;
  try {
    p.parseArgs(argc, argv);
  } catch (const std::invalid_argument &) {
    return NB_ERR_INVALID_ARG;
  }
  out->G = p.G;
  out->dt = p.dt;
  out->num_particles = p.numParticles;
  out->num_frames = p.numFrames;
  out->sim_iterations_per_frame = p.simIterationsPerFrame;
  out->damping = p.damping;
  out->dist_eps = p.distEps;
  out->gw_size = p.gwSize;
  out->calc_method = p.calcMethod == CalculationMethod::PREDICATED
                         ? NB_CALC_PREDICATED
                         : NB_CALC_BRANCH;
  return NB_OK;
}
