// simulation::Simulator / DiskGalaxySimulator over the nbody_amd C ABI.
//
// Same namespace, names and member signatures as the reference's
// src/simulator.cuh:33-160 (the Simulator interface :107-116 and the
// DiskGalaxySimulator class :129-160), so that src/nbody.cpp-style drivers and
// renderers compile against it.  One constructor is added: it takes caller
// particle arrays instead of generating the disk.
//
// This header is plain C++: the HIP side is reached only through
// nbody_amd.h.  Errors from the C ABI are thrown as std::runtime_error.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "nbody_amd.h"
#include "sim_param.hpp"

namespace simulation {

const float PI = 3.14159265358979323846f;

typedef float coords_t;

// Host vector helper (src/simulator.cuh:39-73).
struct vec3 {
  coords_t x = 0.0f;
  coords_t y = 0.0f;
  coords_t z = 0.0f;

  vec3() {}
  vec3(coords_t x_, coords_t y_, coords_t z_) : x{x_}, y{y_}, z{z_} {}

  vec3 &operator+=(const vec3 &rhs) {
    x += rhs.x;
    y += rhs.y;
    z += rhs.z;
    return *this;
  }
  vec3 &operator*=(const coords_t &scale) {
    x *= scale;
    y *= scale;
    z *= scale;
    return *this;
  }
};

inline const vec3 operator*(const vec3 &v, const coords_t &s) {
  return {v.x * s, v.y * s, v.z * s};
}
inline const vec3 operator-(const vec3 &a, const vec3 &b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
inline coords_t dot(const vec3 &a, const vec3 &b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}

// Host SoA particle arrays (src/simulator.cuh:75-84).
struct ParticleData {
  std::vector<coords_t> x;
  std::vector<coords_t> y;
  std::vector<coords_t> z;

  ParticleData(std::vector<coords_t> x_, std::vector<coords_t> y_,
               std::vector<coords_t> z_)
      : x(std::move(x_)), y(std::move(y_)), z(std::move(z_)) {}
  ParticleData(size_t n) : x(n, 0.0f), y(n, 0.0f), z(n, 0.0f) {}
};

coords_t length(const vec3 v);
vec3 cross(const vec3 v0, const vec3 v1);
vec3 normalize(const vec3 v);

// Interface of src/simulator.cuh:107-116: the seven pure virtuals of the
// CUDA version, nothing more, so a class written against the reference
// interface compiles against this one.  getCM() is a non-virtual member of
// DiskGalaxySimulator, as at src/simulator.cuh:140 (the SYCL twin's
// interface has it in place of getGwSize, src_sycl/simulator.dp.hpp:116-125).
class Simulator {
public:
  virtual void stepSim() = 0;
  virtual size_t getNumParticles() = 0;
  virtual const ParticleData &getParticlePos() = 0;
  virtual const ParticleData &getParticleVel() = 0;
  virtual float getLastStepTime() = 0;
  virtual const std::string *getDeviceName() = 0;
  virtual int getGwSize() = 0;
};

// src/simulator.cuh:129-160 on MI355X.  State stays on the device between
// steps; stepSim() advances simIterationsPerFrame steps, synchronises, times
// the frame and refreshes the host copies returned by getParticlePos/Vel.
class DiskGalaxySimulator : public Simulator {
public:
  // Disk initial conditions of src/simulator.cu:131-158 (default engine).
  DiskGalaxySimulator(SimParam params_);
  // Caller-supplied initial conditions; each array holds numParticles values.
  DiskGalaxySimulator(SimParam params_, ParticleData pos_, ParticleData vel_);
  ~DiskGalaxySimulator();

  DiskGalaxySimulator(const DiskGalaxySimulator &) = delete;
  DiskGalaxySimulator &operator=(const DiskGalaxySimulator &) = delete;

  void stepSim();
  float getLastStepTime() { return lastStepTime; }
  size_t getNumParticles() { return params.numParticles; }
  const ParticleData &getParticlePos();
  const ParticleData &getParticleVel();
  const std::string *getDeviceName();
  int getGwSize() { return params.gwSize; }
  CalculationMethod getCM() { return params.calcMethod; }

  // Extensions: device time of the last frame (HIP events, ms) and the
  // compensated force sum (NB_ACC_KAHAN).
  float getLastKernelTime();
  void setAccumulation(int mode);

private:
  SimParam params;
  std::string devName;
  float lastStepTime{0.0f};

  ParticleData pos;
  ParticleData vel;

  nb_sim *handle{nullptr};

  void randomParticlePos();
  void initialParticleVel();
  void sendToDevice();
  void recvFromDevice();
};

} // namespace simulation
