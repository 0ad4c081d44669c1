/*
 * nbody_oracle.c - CPU restatement of the reference's all-pairs gravity step.
 *
 * TEST INFRASTRUCTURE ONLY.  The product (libnbody_amd) never links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline use it, as the checker.
 *
 * What each function restates (paths relative to the reference checkout):
 *   orc_mt_*          std::mt19937 as specified by the C++ standard
 *                     [rand.eng.mers] (engine of src/simulator.cu:133)
 *   orc_canonical     std::generate_canonical<double, 53> over mt19937, as
 *                     uniform_real_distribution<double>(0,1) draws it
 *                     (src/simulator.cu:134; libstdc++ 11)
 *   orc_disk          DiskGalaxySimulator::randomParticlePos +
 *                     initialParticleVel (src/simulator.cu:131-181)
 *   orc_disk_gen      randomParticlePos/randomParticleVel of src/gen.cpp:15-34
 *   orc_step_f32      particle_interaction<BRANCH|PREDICATED>
 *                     (src/simulator.cu:186-229; SYCL twin
 *                     src_sycl/simulator.dp.cpp:315-360), fp32, in source
 *                     order, rsqrt as 1/sqrtf, no contraction
 *   orc_forces_f64    the BRANCH force sum in fp64 with Neumaier compensation
 *   orc_step_f64      the whole step in fp64
 *   orc_energy_momentum_f64  E = sum |v|^2/2 - G sum_{i<j} (|r|^2+eps)^-1/2,
 *                     P = sum v (the potential whose gradient is the force)
 *   orc_adjoint_step_f64  reverse-mode derivative of one step (no reference
 *                     counterpart; derivation in the function comment)
 *
 * How each is pinned (tests/test_oracle.py):
 *   mt19937: the standard's known answer (10000th output of the default
 *            engine = 4123659995) and libstdc++'s std::mt19937 via the
 *            product's generator;  canonical/disk: bit-equality with the
 *            libstdc++-based generator, closed-form invariants;
 *   step_f32: two-body closed form, N=1 drift/damping closed form, fp64
 *            brute force, momentum antisymmetry, self-skip with eps = 0;
 *   forces_f64/step_f64: the same closed forms;
 *   energy: two-body closed form;  adjoint: central finite differences of
 *            orc_step_f64 / of an fp64 trajectory.
 *
 * Build: see oracle/Makefile (-O3 -fopenmp -ffp-contract=off, no fast-math).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Same layout as nb_params in include/nbody_amd.h. */
typedef struct orc_params {
  float G;
  float dt;
  size_t num_particles;
  size_t num_frames;
  int sim_iterations_per_frame;
  float damping;
  float dist_eps;
  int gw_size;
  int calc_method; /* 0 BRANCH, 1 PREDICATED */
} orc_params;

/* ---------------------------------------------------------------- mt19937 */
typedef struct orc_mt {
  uint32_t s[624];
  int i;
} orc_mt;

void orc_mt_seed(orc_mt *g, uint32_t seed) {
  g->s[0] = seed;
  for (int k = 1; k < 624; ++k)
    g->s[k] = 1812433253u * (g->s[k - 1] ^ (g->s[k - 1] >> 30)) + (uint32_t)k;
  g->i = 624;
}

static void orc_mt_twist(orc_mt *g) {
  for (int k = 0; k < 624; ++k) {
    const uint32_t y = (g->s[k] & 0x80000000u) | (g->s[(k + 1) % 624] & 0x7fffffffu);
    g->s[k] = g->s[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
  }
  g->i = 0;
}

uint32_t orc_mt_next(orc_mt *g) {
  if (g->i >= 624) orc_mt_twist(g);
  uint32_t y = g->s[g->i++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* n-th output (1-based) of an engine seeded with `seed`. */
uint32_t orc_mt_nth(uint32_t seed, uint64_t n) {
  orc_mt g;
  uint32_t v = 0;
  orc_mt_seed(&g, seed);
  for (uint64_t k = 0; k < n; ++k) v = orc_mt_next(&g);
  return v;
}

/* generate_canonical<double,53>: k = 2 draws, S = u0 + u1 * 2^32, S / 2^64,
 * with libstdc++'s clamp of a rounded-up 1.0 to the largest double below 1. */
double orc_canonical(orc_mt *g) {
  const double r = 4294967296.0;
  double sum = 0.0, tmp = 1.0;
  for (int k = 0; k < 2; ++k) {
    sum += (double)orc_mt_next(g) * tmp;
    tmp *= r;
  }
  double ret = sum / tmp;
  if (ret >= 1.0) ret = nextafter(1.0, 0.0);
  return ret;
}

/* ------------------------------------------------------------ disk ICs */
static const float kPiF = 3.14159265358979323846f;

static float len_double_form(float x, float y, float z) {
  const double xx = (double)x * (double)x;
  const double yy = (double)y * (double)y;
  const double zz = (double)z * (double)z;
  return (float)sqrt((xx + yy) + zz);
}

static float len_float_form(float x, float y, float z) {
  return sqrtf((x * x + y * y) + z * z);
}

/* flags: 1 = cos/sin in double, 2 = fp32 length (SYCL twin). */
int orc_disk(size_t n, uint32_t seed, int flags, float *px, float *py,
             float *pz, float *vx, float *vy, float *vz) {
  orc_mt g;
  orc_mt_seed(&g, seed);
  for (size_t i = 0; i < n; ++i) {
    const float t = (float)(orc_canonical(&g) * 2 * (double)kPiF);
    const float s = (float)(orc_canonical(&g) * 100);
    if (flags & 1) {
      px[i] = (float)(cos((double)t) * (double)s);
      py[i] = (float)(sin((double)t) * (double)s);
    } else {
      px[i] = cosf(t) * s;
      py[i] = sinf(t) * s;
    }
  }
  for (size_t i = 0; i < n; ++i) pz[i] = (float)(4.0 * orc_canonical(&g));
  for (size_t i = 0; i < n; ++i) {
    /* cross((x,y,z), (0,0,1)) */
    const float cx = py[i] * 1.0f - pz[i] * 0.0f;
    const float cy = pz[i] * 0.0f - px[i] * 1.0f;
    const float cz = px[i] * 0.0f - py[i] * 0.0f;
    const float len = (flags & 2) ? len_float_form(cx, cy, cz)
                                  : len_double_form(cx, cy, cz);
    const float orbital = (float)sqrt(2.0 * (double)len);
    vx[i] = (cx / len) * orbital;
    vy[i] = (cy / len) * orbital;
    vz[i] = (cz / len) * orbital;
  }
  return 0;
}

int orc_disk_gen(size_t n, uint32_t seed, float *px, float *py, float *pz,
                 float *vx, float *vy, float *vz) {
  orc_mt g;
  orc_mt_seed(&g, seed);
  for (size_t i = 0; i < n; ++i) {
    const float t = (float)(orc_canonical(&g) * 2 * (double)kPiF);
    const float s = (float)(orc_canonical(&g) * 100);
    px[i] = cosf(t) * s;
    py[i] = sinf(t) * s;
    pz[i] = (float)(orc_canonical(&g) * 4);
    const float cx = py[i] * 1.0f - pz[i] * 0.0f;
    const float cy = pz[i] * 0.0f - px[i] * 1.0f;
    const float cz = px[i] * 0.0f - py[i] * 0.0f;
    const float dot = (cx * cx + cy * cy) + cz * cz;
    const float orbital = (float)sqrt(2.0 * (double)sqrtf(dot));
    const float inv = 1.0f / sqrtf(dot);
    vx[i] = (cx * inv) * orbital;
    vy[i] = (cy * inv) * orbital;
    vz[i] = (cz * inv) * orbital;
  }
  return 0;
}

/* ------------------------------------------------------------ fp32 step */
#define ORC_LANES 16

/* One step for targets [i0, i1): reads pos (all n), updates vel[i] and
 * writes next[i] for the targets only.  Each target's sum runs over j in
 * increasing order, like the reference's sequential loop. */
void orc_step_f32(const orc_params *p, const float *px, const float *py,
                  const float *pz, float *vx, float *vy, float *vz,
                  float *nx, float *ny, float *nz, size_t n, size_t i0,
                  size_t i1) {
  const float eps = p->dist_eps;
  const int pred = (p->calc_method == 1);
  if (i1 > n) i1 = n;
  if (i0 >= i1) return;
  const long nb = (long)((i1 - i0 + ORC_LANES - 1) / ORC_LANES);
#pragma omp parallel for schedule(dynamic, 1)
  for (long b = 0; b < nb; ++b) {
    const size_t base = i0 + (size_t)b * ORC_LANES;
    float xi[ORC_LANES], yi[ORC_LANES], zi[ORC_LANES];
    float fx[ORC_LANES], fy[ORC_LANES], fz[ORC_LANES];
    long id[ORC_LANES];
    for (int k = 0; k < ORC_LANES; ++k) {
      const size_t i = base + (size_t)k < i1 ? base + (size_t)k : base;
      id[k] = (long)(base + (size_t)k);
      xi[k] = px[i];
      yi[k] = py[i];
      zi[k] = pz[i];
      fx[k] = fy[k] = fz[k] = 0.0f;
    }
    if (!pred) {
      for (size_t j = 0; j < n; ++j) {
        const float xj = px[j], yj = py[j], zj = pz[j];
#pragma omp simd
        for (int k = 0; k < ORC_LANES; ++k) {
          const float rx = xj - xi[k];
          const float ry = yj - yi[k];
          const float rz = zj - zi[k];
          const float d = rx * rx + ry * ry + rz * rz + eps;
          const float inv = 1.0f / sqrtf(d * d * d);
          /* if (i == id) continue;  force += r * inv */
          const int self = ((long)j == id[k]);
          fx[k] = self ? fx[k] : fx[k] + rx * inv;
          fy[k] = self ? fy[k] : fy[k] + ry * inv;
          fz[k] = self ? fz[k] : fz[k] + rz * inv;
        }
      }
    } else {
      for (size_t j = 0; j < n; ++j) {
        const float xj = px[j], yj = py[j], zj = pz[j];
#pragma omp simd
        for (int k = 0; k < ORC_LANES; ++k) {
          const float rx = xj - xi[k];
          const float ry = yj - yi[k];
          const float rz = zj - zi[k];
          const float d = rx * rx + ry * ry + rz * rz + eps;
          const float inv = 1.0f / sqrtf(d * d * d);
          /* force += r * inv * (i == id) */
          const float m = ((long)j == id[k]) ? 1.0f : 0.0f;
          fx[k] = fx[k] + rx * inv * m;
          fy[k] = fy[k] + ry * inv * m;
          fz[k] = fz[k] + rz * inv * m;
        }
      }
    }
    for (int k = 0; k < ORC_LANES; ++k) {
      const size_t i = base + (size_t)k;
      if (i >= i1) break;
      float ux = vx[i] * p->damping;
      float uy = vy[i] * p->damping;
      float uz = vz[i] * p->damping;
      ux = ux + fx[k] * p->dt * p->G;
      uy = uy + fy[k] * p->dt * p->G;
      uz = uz + fz[k] * p->dt * p->G;
      vx[i] = ux;
      vy[i] = uy;
      vz[i] = uz;
      nx[i] = px[i] + ux * p->dt;
      ny[i] = py[i] + uy * p->dt;
      nz[i] = pz[i] + uz * p->dt;
    }
  }
}

/* ------------------------------------------------------------ fp64 parts */
static void neumaier(double *s, double *c, double v) {
  const double t = *s + v;
  if (fabs(*s) >= fabs(v)) *c += (*s - t) + v;
  else *c += (v - t) + *s;
  *s = t;
}

/* fp64 forces on the targets idx[0..m) from all n float sources.
 * skip_self: 1 = BRANCH semantics (j == i excluded). */
void orc_forces_f64(const float *px, const float *py, const float *pz,
                    size_t n, const int64_t *idx, size_t m, double eps,
                    int skip_self, double *fx, double *fy, double *fz) {
#pragma omp parallel for schedule(dynamic, 16)
  for (long t = 0; t < (long)m; ++t) {
    const int64_t i = idx[t];
    const double xi = px[i], yi = py[i], zi = pz[i];
    double sx = 0, cx = 0, sy = 0, cy = 0, sz = 0, cz = 0;
    for (size_t j = 0; j < n; ++j) {
      if (skip_self && (int64_t)j == i) continue;
      const double rx = (double)px[j] - xi;
      const double ry = (double)py[j] - yi;
      const double rz = (double)pz[j] - zi;
      const double d = rx * rx + ry * ry + rz * rz + eps;
      const double inv = 1.0 / (d * sqrt(d));
      neumaier(&sx, &cx, rx * inv);
      neumaier(&sy, &cy, ry * inv);
      neumaier(&sz, &cz, rz * inv);
    }
    fx[t] = sx + cx;
    fy[t] = sy + cy;
    fz[t] = sz + cz;
  }
}

/* fp64 force on every particle of a double-precision state (BRANCH). */
static void forces_all_f64(const double *x, const double *y, const double *z,
                           size_t n, double eps, double *fx, double *fy,
                           double *fz) {
#pragma omp parallel for schedule(dynamic, 16)
  for (long i = 0; i < (long)n; ++i) {
    double sx = 0, cx = 0, sy = 0, cy = 0, sz = 0, cz = 0;
    for (size_t j = 0; j < n; ++j) {
      if ((long)j == i) continue;
      const double rx = x[j] - x[i], ry = y[j] - y[i], rz = z[j] - z[i];
      const double d = rx * rx + ry * ry + rz * rz + eps;
      const double inv = 1.0 / (d * sqrt(d));
      neumaier(&sx, &cx, rx * inv);
      neumaier(&sy, &cy, ry * inv);
      neumaier(&sz, &cz, rz * inv);
    }
    fx[i] = sx + cx;
    fy[i] = sy + cy;
    fz[i] = sz + cz;
  }
}

/* One fp64 step (BRANCH semantics): x, v updated in place. */
void orc_step_f64(const orc_params *p, double *x, double *y, double *z,
                  double *vx, double *vy, double *vz, size_t n) {
  double *f = (double *)malloc(sizeof(double) * 3 * (n ? n : 1));
  if (!f) return;
  forces_all_f64(x, y, z, n, (double)p->dist_eps, f, f + n, f + 2 * n);
  const double a = p->damping, c = (double)p->dt * (double)p->G,
               dt = p->dt;
  for (size_t i = 0; i < n; ++i) {
    vx[i] = a * vx[i] + c * f[i];
    vy[i] = a * vy[i] + c * f[n + i];
    vz[i] = a * vz[i] + c * f[2 * n + i];
    x[i] += dt * vx[i];
    y[i] += dt * vy[i];
    z[i] += dt * vz[i];
  }
  free(f);
}

/* out[0] kinetic, out[1] potential, out[2..4] momentum. */
void orc_energy_momentum_f64(const orc_params *p, const float *px,
                             const float *py, const float *pz,
                             const float *vx, const float *vy,
                             const float *vz, size_t n, double *out) {
  double ke = 0, pe = 0, mx = 0, my = 0, mz = 0;
  const double eps = p->dist_eps;
  for (size_t i = 0; i < n; ++i) {
    ke += 0.5 * ((double)vx[i] * vx[i] + (double)vy[i] * vy[i] +
                 (double)vz[i] * vz[i]);
    mx += vx[i];
    my += vy[i];
    mz += vz[i];
  }
#pragma omp parallel for reduction(+ : pe) schedule(dynamic, 16)
  for (long i = 0; i < (long)n; ++i) {
    double s = 0;
    for (size_t j = (size_t)i + 1; j < n; ++j) {
      const double rx = (double)px[j] - px[i];
      const double ry = (double)py[j] - py[i];
      const double rz = (double)pz[j] - pz[i];
      s += 1.0 / sqrt(rx * rx + ry * ry + rz * rz + eps);
    }
    pe += s;
  }
  out[0] = ke;
  out[1] = -(double)p->G * pe;
  out[2] = mx;
  out[3] = my;
  out[4] = mz;
}

/*
 * Reverse-mode derivative of one step (x, v) -> (x', v'):
 *   v' = a v + c F(x),   x' = x + dt v',   a = damping, c = dt G,
 *   F_i = sum_{j != i} f(x_j - x_i),  f(r) = r K,  K = (|r|^2 + eps)^{-3/2}.
 * Given lx' = dL/dx', lv' = dL/dv':
 *   g   = lv' + dt lx'
 *   lv  = a g
 *   lx  = lx' + c J^T g,  (J^T g)_k = sum_{j != k} M_kj (g_j - g_k),
 *   M(r) = df/dr = K (I - 3 r r^T / (|r|^2 + eps))   (symmetric, even in r).
 */
void orc_adjoint_step_f64(const orc_params *p, const double *x,
                          const double *y, const double *z, const double *lxn,
                          const double *lyn, const double *lzn,
                          const double *lvxn, const double *lvyn,
                          const double *lvzn, double *lx, double *ly,
                          double *lz, double *lvx, double *lvy, double *lvz,
                          size_t n) {
  const double a = p->damping, dt = p->dt, c = (double)p->dt * (double)p->G;
  const double eps = p->dist_eps;
  double *g = (double *)malloc(sizeof(double) * 3 * (n ? n : 1));
  if (!g) return;
  for (size_t i = 0; i < n; ++i) {
    g[i] = lvxn[i] + dt * lxn[i];
    g[n + i] = lvyn[i] + dt * lyn[i];
    g[2 * n + i] = lvzn[i] + dt * lzn[i];
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (long k = 0; k < (long)n; ++k) {
    double ax = 0, ay = 0, az = 0;
    for (size_t j = 0; j < n; ++j) {
      if ((long)j == k) continue;
      const double rx = x[j] - x[k], ry = y[j] - y[k], rz = z[j] - z[k];
      const double d = rx * rx + ry * ry + rz * rz + eps;
      const double K = 1.0 / (d * sqrt(d));
      const double hx = g[j] - g[k], hy = g[n + j] - g[n + k],
                   hz = g[2 * n + j] - g[2 * n + k];
      const double rh = rx * hx + ry * hy + rz * hz;
      const double w = 3.0 * rh / d;
      ax += K * (hx - w * rx);
      ay += K * (hy - w * ry);
      az += K * (hz - w * rz);
    }
    lx[k] = lxn[k] + c * ax;
    ly[k] = lyn[k] + c * ay;
    lz[k] = lzn[k] + c * az;
  }
  for (size_t i = 0; i < n; ++i) {
    lvx[i] = a * g[i];
    lvy[i] = a * g[n + i];
    lvz[i] = a * g[2 * n + i];
  }
  free(g);
}

/*
 * fp64 trajectory of T steps from (x0, v0) and the gradient of
 *   L = (1/n) sum_i |x_T,i - target_i|^2
 * with respect to x0 and v0 (every state stored; n * T doubles * 3).
 * Returns L; writes gx, gv (3n each, SoA x|y|z) when non-NULL.
 */
double orc_traj_grad_f64(const orc_params *p, const double *x0,
                         const double *v0, const double *target, size_t n,
                         int T, double *gx, double *gv) {
  const size_t n3 = 3 * n;
  double *xs = (double *)malloc(sizeof(double) * n3 * (size_t)(T + 1));
  double *v = (double *)malloc(sizeof(double) * n3);
  if (!xs || !v) {
    free(xs);
    free(v);
    return NAN;
  }
  memcpy(xs, x0, sizeof(double) * n3);
  memcpy(v, v0, sizeof(double) * n3);
  for (int t = 0; t < T; ++t) {
    double *cur = xs + n3 * (size_t)(t + 1);
    memcpy(cur, xs + n3 * (size_t)t, sizeof(double) * n3);
    orc_step_f64(p, cur, cur + n, cur + 2 * n, v, v + n, v + 2 * n, n);
  }
  const double *xT = xs + n3 * (size_t)T;
  double L = 0;
  for (size_t i = 0; i < n3; ++i) {
    const double dlt = xT[i] - target[i];
    L += dlt * dlt;
  }
  L /= (double)n;
  if (gx && gv) {
    double *lx = (double *)malloc(sizeof(double) * n3);
    double *lv = (double *)calloc(n3, sizeof(double));
    double *lx2 = (double *)malloc(sizeof(double) * n3);
    double *lv2 = (double *)malloc(sizeof(double) * n3);
    for (size_t i = 0; i < n3; ++i) lx[i] = 2.0 * (xT[i] - target[i]) / (double)n;
    for (int t = T - 1; t >= 0; --t) {
      const double *xt = xs + n3 * (size_t)t;
      orc_adjoint_step_f64(p, xt, xt + n, xt + 2 * n, lx, lx + n, lx + 2 * n,
                           lv, lv + n, lv + 2 * n, lx2, lx2 + n, lx2 + 2 * n,
                           lv2, lv2 + n, lv2 + 2 * n, n);
      double *tmp = lx;
      lx = lx2;
      lx2 = tmp;
      tmp = lv;
      lv = lv2;
      lv2 = tmp;
    }
    memcpy(gx, lx, sizeof(double) * n3);
    memcpy(gv, lv, sizeof(double) * n3);
    free(lx);
    free(lv);
    free(lx2);
    free(lv2);
  }
  free(xs);
  free(v);
  return L;
}
