// C entry point over the reference's own SimParam (src/sim_param.cpp), so the
// tests can compare this project's defaults and argv parsing with the
// reference's compiled behaviour.  Test infrastructure only.
#include <stdexcept>

#include "sim_param.hpp" // the reference's header (-I<reference>/src)

extern "C" int ref_simparam(int argc, char **argv, float *G, float *dt,
                            unsigned long long *num_particles,
                            unsigned long long *num_frames, int *iters,
                            float *damping, float *dist_eps, int *gw_size,
                            int *calc_method) {
  try {
    SimParam p;
    p.parseArgs(argc, argv);
    *G = p.G;
    *dt = p.dt;
    *num_particles = p.numParticles;
    *num_frames = p.numFrames;
    *iters = p.simIterationsPerFrame;
    *damping = p.damping;
    *dist_eps = p.distEps;
    *gw_size = p.gwSize;
    *calc_method = p.calcMethod == CalculationMethod::PREDICATED ? 1 : 0;
    return 0;
  } catch (const std::invalid_argument &) {
    return 1;
  }
}
