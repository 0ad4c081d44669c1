// nbody_amd: the headless driver of the reference (src/nbody.cpp:31-141 with
// DISABLE_GL) over this project's DiskGalaxySimulator.
//
//   nbody_amd numParticles/256 simIterationsPerFrame damping dt distEps G
//             numFrames gwSize calcMethod
//
// Every frame calls stepSim(); after the first two (warm-up) frames it prints
//   At step <k> kernel time is <ms> and mean is <ms> and stddev is: <ms>
// where the statistics cover all frames after the warm-up (population
// standard deviation), as the reference prints them.

#include <cmath>
#include <cstdio>
#include <exception>
#include <iostream>
#include <vector>

#include "sim_param.hpp"
#include "simulator.hpp"

int main(int argc, char **argv) {
  SimParam params;
  try {
    params.parseArgs(argc, argv);
  } catch (const std::exception &e) {
    std::cerr << e.what() << "\n";
    return 1;
  }
  try {
    simulation::DiskGalaxySimulator sim(params);
    const size_t warm = 2;
    std::vector<float> times;
    size_t step = 0;
    while (step < params.numFrames) {
      sim.stepSim();
      ++step;
      if (step <= warm) continue;
      times.push_back(sim.getLastStepTime());
      double sum = 0.0;
      for (float t : times) sum += t;
      const float mean = (float)(sum / times.size());
      double var = 0.0;
      for (float t : times) var += (double)(t - mean) * (double)(t - mean);
      const float sd = (float)std::sqrt(var / times.size());
      std::cout << "At step " << step << " kernel time is " << times.back()
                << " and mean is " << mean << " and stddev is: " << sd
                << "\n";
    }
  } catch (const std::exception &e) {
    std::fprintf(stderr, "GPUassert: %s\n", e.what());
    return 1;
  }
  return 0;
}
