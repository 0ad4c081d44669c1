// DiskGalaxySimulator over the nbody_amd C ABI (host C++ only).
//
// Member-for-member counterpart of src/simulator.cu:24-181: the constructor
// generates the disk and uploads it, stepSim() runs one frame and reads the
// state back, getDeviceName() queries device 0 once.  The device work is in
// libnbody_amd (nb_sim_*).

#include "simulator.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <utility>

namespace simulation {

namespace {

void check(int status, const char *what) {
  if (status != NB_OK)
    throw std::runtime_error(std::string(what) + ": " + nb_last_error());
}

nb_params to_c(const SimParam &p) {
  nb_params c;
  c.G = p.G;
  c.dt = p.dt;
  c.num_particles = p.numParticles;
  c.num_frames = p.numFrames;
  c.sim_iterations_per_frame = p.simIterationsPerFrame;
  c.damping = p.damping;
  c.dist_eps = p.distEps;
  c.gw_size = p.gwSize;
  c.calc_method = p.calcMethod == CalculationMethod::PREDICATED
                      ? NB_CALC_PREDICATED
                      : NB_CALC_BRANCH;
  return c;
}

} // namespace

// Host helpers of src/simulator.cu:165-181.
vec3 cross(const vec3 v0, const vec3 v1) {
  return vec3(v0.y * v1.z - v0.z * v1.y, v0.z * v1.x - v0.x * v1.z,
              v0.x * v1.y - v0.y * v1.x);
}

coords_t length(const vec3 v) {
  const double xx = (double)v.x * v.x, yy = (double)v.y * v.y,
               zz = (double)v.z * v.z;
  return (coords_t)std::sqrt((xx + yy) + zz);
}

vec3 normalize(const vec3 v) {
  const coords_t len = length(v);
  return vec3(v.x / len, v.y / len, v.z / len);
}

DiskGalaxySimulator::DiskGalaxySimulator(SimParam params_)
    : params(params_), pos(params_.numParticles), vel(params_.numParticles) {
  randomParticlePos();
  initialParticleVel();
  sendToDevice();
}

DiskGalaxySimulator::DiskGalaxySimulator(SimParam params_, ParticleData pos_,
                                         ParticleData vel_)
    : params(params_), pos(std::move(pos_)), vel(std::move(vel_)) {
  const size_t n = params.numParticles;
  if (pos.x.size() != n || pos.y.size() != n || pos.z.size() != n ||
      vel.x.size() != n || vel.y.size() != n || vel.z.size() != n)
    throw std::invalid_argument(
        "particle arrays must hold SimParam::numParticles values each");
  sendToDevice();
}

DiskGalaxySimulator::~DiskGalaxySimulator() {
  if (handle) nb_sim_destroy(handle);
}

void DiskGalaxySimulator::randomParticlePos() {
  // Positions and velocities come from one generator call; the velocity half
  // is applied by initialParticleVel's counterpart below.
  const size_t n = params.numParticles;
  check(nb_ic_disk(n, 5489u, 0, pos.x.data(), pos.y.data(), pos.z.data(),
                   vel.x.data(), vel.y.data(), vel.z.data()),
        "nb_ic_disk");
}

void DiskGalaxySimulator::initialParticleVel() {
  // Filled together with the positions (same engine, same order).
}

void DiskGalaxySimulator::sendToDevice() {
  const nb_params c = to_c(params);
  if (handle) {
    nb_sim_destroy(handle);
    handle = nullptr;
  }
  check(nb_sim_create(&c, pos.x.data(), pos.y.data(), pos.z.data(),
                      vel.x.data(), vel.y.data(), vel.z.data(), &handle),
        "nb_sim_create");
}

void DiskGalaxySimulator::recvFromDevice() {
  // one asynchronous pinned copy of all six arrays, one wait
  check(nb_sim_stage_readback(handle), "nb_sim_stage_readback");
  const float *a[6];
  check(nb_sim_staged(handle, a), "nb_sim_staged");
  const size_t n = params.numParticles;
  if (n == 0) return;
  std::vector<coords_t> *dst[6] = {&pos.x, &pos.y, &pos.z,
                                   &vel.x, &vel.y, &vel.z};
  for (int c = 0; c < 6; ++c) std::copy(a[c], a[c] + n, dst[c]->begin());
}

void DiskGalaxySimulator::stepSim() {
  check(nb_sim_step(handle), "nb_sim_step");
  lastStepTime = nb_sim_last_step_ms(handle);
  recvFromDevice();
}

const ParticleData &DiskGalaxySimulator::getParticlePos() { return pos; }

const ParticleData &DiskGalaxySimulator::getParticleVel() { return vel; }

const std::string *DiskGalaxySimulator::getDeviceName() {
  if (devName.empty()) {
    char buf[256];
    if (nb_device_name(buf, sizeof buf) != NB_OK) devName = "Unknown Device";
    else devName = buf;
  }
  return &devName;
}

float DiskGalaxySimulator::getLastKernelTime() {
  return nb_sim_last_kernel_ms(handle);
}

void DiskGalaxySimulator::setAccumulation(int mode) {
  check(nb_sim_set_accumulation(handle, mode), "nb_sim_set_accumulation");
}

} // namespace simulation
