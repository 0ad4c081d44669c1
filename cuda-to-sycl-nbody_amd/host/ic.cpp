// Initial-condition generators exported through the C ABI (host code).
//
// nb_ic_disk follows DiskGalaxySimulator::randomParticlePos and
// initialParticleVel (src/simulator.cu:131-158, helpers :165-181): the same
// engine (std::mt19937), the same distribution (uniform_real_distribution
// <double> on [0,1)), the same draw order and the same float/double rounding
// points, so a given seed reproduces the reference's arrays.  The draw order
// and rounding points have to match for that; the code is this project's own.
//
// nb_ic_disk_gen follows gen.cpp's randomParticlePos/randomParticleVel
// (src/gen.cpp:15-34), whose draw order interleaves t, s, z per particle.
//
// nb_ic_plummer has no reference counterpart (Aarseth, Henon & Wielen 1974).
//
// Built with -ffp-contract=off so that no product/sum pair is fused.

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <random>
#include <vector>

#include "nbody_amd.h"

namespace {

// simulation::PI is a float constant (src/simulator.cuh:35).
const float kPiF = 3.14159265358979323846f;

// |v| the way the CUDA host helper computes it: squares and sum in double,
// double sqrt, rounded to float on return (src/simulator.cu:170-172).
float length_double_form(float x, float y, float z) {
  const double xx = (double)x * (double)x; // == pow(x, 2): exact in double
  const double yy = (double)y * (double)y;
  const double zz = (double)z * (double)z;
  return (float)std::sqrt((xx + yy) + zz);
}

// |v| in float, as the SYCL twin computes it (src_sycl/simulator.dp.cpp:299-301).
float length_float_form(float x, float y, float z) {
  return std::sqrt((x * x + y * y) + z * z);
}

// Tangential orbital velocity of src/simulator.cu:149-158 for one particle.
void orbital_velocity(float x, float y, float z, bool len_f32, float *vx,
                      float *vy, float *vz) {
  // cross(p, (0, 0, 1)) in float (src/simulator.cu:165-168)
  const float cx = y * 1.0f - z * 0.0f;
  const float cy = z * 0.0f - x * 1.0f;
  const float cz = x * 0.0f - y * 0.0f;
  const float len = len_f32 ? length_float_form(cx, cy, cz)
                            : length_double_form(cx, cy, cz);
  const float orbital = (float)std::sqrt(2.0 * (double)len);
  // normalize(v) divides each component by the (same) length, then scales
  *vx = (cx / len) * orbital;
  *vy = (cy / len) * orbital;
  *vz = (cz / len) * orbital;
}

bool any_null(const float *a, const float *b, const float *c, const float *d,
              const float *e, const float *f) {
  return !a || !b || !c || !d || !e || !f;
}

} // namespace

extern "C" {

int nb_ic_disk(size_t n, uint32_t seed, int flags, float *px, float *py,
               float *pz, float *vx, float *vy, float *vz) {
  if (n == 0) return NB_OK;
  if (any_null(px, py, pz, vx, vy, vz)) return NB_ERR_INVALID_ARG;
  std::mt19937 gen(seed);
  std::uniform_real_distribution<double> dis(0.0, 1.0);
  const bool trig_double = (flags & NB_IC_TRIG_DOUBLE) != 0;
  for (size_t i = 0; i < n; ++i) {
    const float t = (float)(dis(gen) * 2 * (double)kPiF);
    const float s = (float)(dis(gen) * 100);
    if (trig_double) {
      px[i] = (float)(std::cos((double)t) * (double)s);
      py[i] = (float)(std::sin((double)t) * (double)s);
    } else {
      px[i] = std::cos(t) * s; // float overloads
      py[i] = std::sin(t) * s;
    }
  }
  // z is drawn after all of x, y: uniform on [0, 4) (src/simulator.cu:144-146)
  for (size_t i = 0; i < n; ++i) pz[i] = (float)(4.0 * dis(gen));
  const bool len_f32 = (flags & NB_IC_LENGTH_F32) != 0;
  for (size_t i = 0; i < n; ++i)
    orbital_velocity(px[i], py[i], pz[i], len_f32, &vx[i], &vy[i], &vz[i]);
  return NB_OK;
}

int nb_ic_disk_gen(size_t n, uint32_t seed, float *px, float *py, float *pz,
                   float *vx, float *vy, float *vz) {
  if (n == 0) return NB_OK;
  if (any_null(px, py, pz, vx, vy, vz)) return NB_ERR_INVALID_ARG;
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> dis(0.0, 1.0);
  for (size_t i = 0; i < n; ++i) {
    const float t = (float)(dis(rng) * 2 * (double)kPiF);
    const float s = (float)(dis(rng) * 100);
    px[i] = std::cos(t) * s; // `using namespace std`: float overloads
    py[i] = std::sin(t) * s;
    pz[i] = (float)(dis(rng) * 4);
    // glm::cross / glm::length / glm::normalize, all in float
    const float cx = py[i] * 1.0f - pz[i] * 0.0f;
    const float cy = pz[i] * 0.0f - px[i] * 1.0f;
    const float cz = px[i] * 0.0f - py[i] * 0.0f;
    const float dot = (cx * cx + cy * cy) + cz * cz;
    const float orbital = (float)std::sqrt(2.0 * (double)std::sqrt(dot));
    const float inv = 1.0f / std::sqrt(dot);
    vx[i] = (cx * inv) * orbital;
    vy[i] = (cy * inv) * orbital;
    vz[i] = (cz * inv) * orbital;
  }
  return NB_OK;
}

int nb_ic_plummer(size_t n, uint32_t seed, float a, float g_mass, float *px,
                  float *py, float *pz, float *vx, float *vy, float *vz) {
  if (n == 0) return NB_OK;
  if (any_null(px, py, pz, vx, vy, vz)) return NB_ERR_INVALID_ARG;
  if (!(a > 0.0f) || !(g_mass >= 0.0f)) return NB_ERR_INVALID_ARG;
  std::mt19937 gen(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const double two_pi = 6.283185307179586476925286766559;
  std::vector<double> buf(6 * n);
  double cm[6] = {0, 0, 0, 0, 0, 0};
  for (size_t i = 0; i < n; ++i) {
    // radius: invert M(<r)/M = r^3 / (r^2 + a^2)^{3/2}; drop the outer 0.1 %
    double m;
    do {
      m = U(gen);
    } while (m < 1e-12 || m > 0.999);
    const double r = a / std::sqrt(std::pow(m, -2.0 / 3.0) - 1.0);
    const double ct = 1.0 - 2.0 * U(gen);
    const double ph = two_pi * U(gen);
    const double st = std::sqrt(std::fmax(0.0, 1.0 - ct * ct));
    double *b = &buf[6 * i];
    b[0] = r * st * std::cos(ph);
    b[1] = r * st * std::sin(ph);
    b[2] = r * ct;
    // speed: q = v / v_esc has density ~ q^2 (1 - q^2)^{7/2} (von Neumann
    // rejection with bound 0.1 > max 0.0923)
    double q, y;
    do {
      q = U(gen);
      y = 0.1 * U(gen);
    } while (y > q * q * std::pow(1.0 - q * q, 3.5));
    const double vesc =
        std::sqrt(2.0 * (double)g_mass) * std::pow(r * r + (double)a * a, -0.25);
    const double v = q * vesc;
    const double cv = 1.0 - 2.0 * U(gen);
    const double pv = two_pi * U(gen);
    const double sv = std::sqrt(std::fmax(0.0, 1.0 - cv * cv));
    b[3] = v * sv * std::cos(pv);
    b[4] = v * sv * std::sin(pv);
    b[5] = v * cv;
    for (int k = 0; k < 6; ++k) cm[k] += b[k];
  }
  for (int k = 0; k < 6; ++k) cm[k] /= (double)n;
  for (size_t i = 0; i < n; ++i) {
    const double *b = &buf[6 * i];
    px[i] = (float)(b[0] - cm[0]);
    py[i] = (float)(b[1] - cm[1]);
    pz[i] = (float)(b[2] - cm[2]);
    vx[i] = (float)(b[3] - cm[3]);
    vy[i] = (float)(b[4] - cm[4]);
    vz[i] = (float)(b[5] - cm[5]);
  }
  return NB_OK;
}

} // extern "C"
