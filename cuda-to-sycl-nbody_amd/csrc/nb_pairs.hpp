// Device building blocks of the all-pairs force sum, shared by the single-
// system kernel (nb_forward.hip) and the batched persistent kernel
// (nb_batch.hip).
//
//   pair<MODE>()        one (i, j) interaction
//   Targets<U, MODE>    U target particles per thread held in VGPRs
//   accumulate_sources  streams sources [jbeg, jend) through LDS tiles and
//                       adds their force on the thread's targets
//   integrate_ref       the reference's velocity/position update
//
// Per interaction: 11 VALU ops + 1 transcendental, the reference's
// arithmetic (rsqrt(d*d*d), src/simulator.cu:201-202):
//   3 sub, 3 fma (|r|^2 + eps), 2 mul (d^3), v_rsq_f32, 3 fma (F += r * inv).
// On gfx950 the v_rsq_f32 issues beside the VALU stream (measured: 11.6
// issue slots per interaction at 2.4 GHz), so the 11 VALU ops set the
// ceiling.  Trading a multiply for a second transcendental
// (rsq(d) * rcp(d)) measured 15 % slower: the transcendental pipe, not the
// VALU, then limits.
//
// Translation units including this header are built with
// -ffp-contract=off: every fused operation below is an explicit fmaf.
#pragma once

#include "nb_internal.hpp"

namespace nb {

// Integrator of src/simulator.cu:214-228 in source order (no contraction):
//   v = v * damping;  v += (F * dt) * G;  x_next = x + v * dt
__device__ __forceinline__ void integrate_ref(const StepConsts &c, int i,
                                              float fx, float fy, float fz,
                                              const float *px, const float *py,
                                              const float *pz, float *vx,
                                              float *vy, float *vz, float *nx,
                                              float *ny, float *nz) {
  float ux = vx[i] * c.damping;
  float uy = vy[i] * c.damping;
  float uz = vz[i] * c.damping;
  ux = ux + (fx * c.dt) * c.G;
  uy = uy + (fy * c.dt) * c.G;
  uz = uz + (fz * c.dt) * c.G;
  vx[i] = ux;
  vy[i] = uy;
  vz[i] = uz;
  nx[i] = px[i] + ux * c.dt;
  ny[i] = py[i] + uy * c.dt;
  nz[i] = pz[i] + uz * c.dt;
}

// Kahan step: s += r * inv with running compensation c (the product is
// fused into the first subtraction).
__device__ __forceinline__ void kahan_add(float r, float inv, float &s,
                                          float &c) {
  const float y = __builtin_fmaf(r, inv, -c);
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// One (i, j) interaction.  j and i are global indices (used by the masked
// variants only).
template <int MODE>
__device__ __forceinline__ void pair(float xj, float yj, float zj, int j,
                                     float xi, float yi, float zi, int i,
                                     float eps, float &ax, float &ay,
                                     float &az, float &cx, float &cy,
                                     float &cz) {
  const float rx = xj - xi;
  const float ry = yj - yi;
  const float rz = zj - zi;
  float d;
  if constexpr (MODE == MODE_PREDICATED) {
    // dot(r, r) + distEps, as written (src/simulator.cu:201)
    d = ((rx * rx + ry * ry) + rz * rz) + eps;
  } else {
    d = __builtin_fmaf(rx, rx, eps);
    d = __builtin_fmaf(ry, ry, d);
    d = __builtin_fmaf(rz, rz, d);
  }
  float inv = __builtin_amdgcn_rsqf((d * d) * d); // src/simulator.cu:202
  if constexpr (MODE == MODE_PLAIN_SKIP || MODE == MODE_KAHAN_SKIP) {
    inv = (j == i) ? 0.0f : inv;
  }
  if constexpr (MODE == MODE_PLAIN || MODE == MODE_PLAIN_SKIP) {
    ax = __builtin_fmaf(rx, inv, ax);
    ay = __builtin_fmaf(ry, inv, ay);
    az = __builtin_fmaf(rz, inv, az);
  } else if constexpr (MODE == MODE_KAHAN || MODE == MODE_KAHAN_SKIP) {
    kahan_add(rx, inv, ax, cx);
    kahan_add(ry, inv, ay, cy);
    kahan_add(rz, inv, az, cz);
  } else {
    // force += r * inv_dist_cube * (i == id)  (src/simulator.cu:209)
    const float m = (j == i) ? 1.0f : 0.0f;
    ax = ax + (rx * inv) * m;
    ay = ay + (ry * inv) * m;
    az = az + (rz * inv) * m;
  }
}

template <int U, int MODE>
struct Targets {
  float x[U], y[U], z[U];
  float ax[U], ay[U], az[U];
  float cx[U], cy[U], cz[U];
  int idx[U];

  // Targets base + k * block + lane (k < U); lanes past `end` compute on a
  // copy of target end - 1 and are never stored.
  __device__ __forceinline__ void load(const float *px, const float *py,
                                       const float *pz, int base, int block,
                                       int lane, int end) {
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int i = base + k * block + lane;
      const int ic = i < end ? i : end - 1;
      idx[k] = i;
      x[k] = px[ic];
      y[k] = py[ic];
      z[k] = pz[ic];
      ax[k] = ay[k] = az[k] = 0.0f;
      cx[k] = cy[k] = cz[k] = 0.0f;
    }
  }

  // Accumulated force of target k (the Kahan modes apply the compensation).
  __device__ __forceinline__ void force(int k, float &fx, float &fy,
                                        float &fz) const {
    fx = ax[k];
    fy = ay[k];
    fz = az[k];
    if constexpr (MODE == MODE_KAHAN || MODE == MODE_KAHAN_SKIP) {
      fx = fx - cx[k];
      fy = fy - cy[k];
      fz = fz - cz[k];
    }
  }

  __device__ __forceinline__ void one(float xj, float yj, float zj, int j,
                                      float eps) {
#pragma unroll
    for (int k = 0; k < U; ++k)
      pair<MODE>(xj, yj, zj, j, x[k], y[k], z[k], idx[k], eps, ax[k], ay[k],
                 az[k], cx[k], cy[k], cz[k]);
  }

  // Four consecutive sources j0..j0+3 (one b128 LDS read per coordinate).
  __device__ __forceinline__ void four(float4 X, float4 Y, float4 Z, int j0,
                                       float eps) {
    one(X.x, Y.x, Z.x, j0 + 0, eps);
    one(X.y, Y.y, Z.y, j0 + 1, eps);
    one(X.z, Y.z, Z.z, j0 + 2, eps);
    one(X.w, Y.w, Z.w, j0 + 3, eps);
  }
};

// LDS source tile: three SoA float arrays of TILE particles.
template <int TILE>
struct SourceTile {
  float4 x[TILE / 4];
  float4 y[TILE / 4];
  float4 z[TILE / 4];
};

// Adds the force of sources [jbeg, jend) on the thread's targets.  Every
// thread of the block calls it with the same range (it contains barriers).
//   * sources are fetched from HBM with coalesced loads into registers one
//     tile ahead (the fetch of tile t+1 overlaps the pair loop of tile t);
//   * each lane reads 4 sources per coordinate with one ds_read_b128 at a
//     wave-uniform address (an LDS broadcast: no bank conflicts);
//   * the group loop is unrolled twice; the compiler issues the six b128
//     reads of an iteration before its 8 x U interactions (an alternating
//     two-register-set prefetch across iterations measured slower).
template <int BLOCK, int TILE, int U, int MODE>
__device__ __forceinline__ void accumulate_sources(
    SourceTile<TILE> &sm, Targets<U, MODE> &T, const float *px,
    const float *py, const float *pz, int jbeg, int jend, float eps) {
  static_assert(TILE % BLOCK == 0, "tile must be a multiple of the block");
  constexpr int L = TILE / BLOCK; // source loads per thread per coordinate
  constexpr int G = TILE / 4;     // float4 groups per tile
  float *fsx = reinterpret_cast<float *>(sm.x);
  float *fsy = reinterpret_cast<float *>(sm.y);
  float *fsz = reinterpret_cast<float *>(sm.z);
  const int lane = threadIdx.x;

  float bx[L], by[L], bz[L];
  auto fetch = [&](int j0) {
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int j = j0 + l * BLOCK + lane;
      const bool ok = j < jend;
      bx[l] = ok ? px[j] : 0.0f;
      by[l] = ok ? py[j] : 0.0f;
      bz[l] = ok ? pz[j] : 0.0f;
    }
  };

  if (jbeg < jend) fetch(jbeg);
  for (int j0 = jbeg; j0 < jend; j0 += TILE) {
    __syncthreads(); // the previous tile is no longer read
#pragma unroll
    for (int l = 0; l < L; ++l) {
      fsx[l * BLOCK + lane] = bx[l];
      fsy[l * BLOCK + lane] = by[l];
      fsz[l * BLOCK + lane] = bz[l];
    }
    __syncthreads();
    if (j0 + TILE < jend) fetch(j0 + TILE); // overlaps the pair loop below

    const int cnt = min(TILE, jend - j0);
    if (cnt == TILE) {
#pragma unroll 2
      for (int q = 0; q < G; ++q)
        T.four(sm.x[q], sm.y[q], sm.z[q], j0 + 4 * q, eps);
    } else {
      const int q4 = cnt >> 2;
      for (int q = 0; q < q4; ++q)
        T.four(sm.x[q], sm.y[q], sm.z[q], j0 + 4 * q, eps);
      for (int jj = q4 * 4; jj < cnt; ++jj)
        T.one(fsx[jj], fsy[jj], fsz[jj], j0 + jj, eps);
    }
  }
}

} // namespace nb
