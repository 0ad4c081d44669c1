// fp64 diagnostics on the GPU (no reference counterpart): the oracle force
// on a sampled subset of targets and the total energy / momentum.  These are
// the only fp64 computations on the device; the step itself stays fp32.
//
//   forces:  F_i = sum_{j != i} r (|r|^2 + eps)^{-3/2} in fp64 from the float
//            positions (r formed in double), for targets idx[0..m).  The
//            sources are split over grid.y chunks; a second kernel adds the
//            chunk partials in chunk order (deterministic).
//   energy:  K = sum |v|^2 / 2,  W = -G sum_{i<j} (|r|^2 + eps)^{-1/2}
//            (the potential whose negative gradient is the step's force),
//            P = sum v (unit masses).  W is summed over all ordered pairs
//            and halved; every sum is fp64, per-workgroup partials are
//            added in workgroup order (deterministic).
// 1/sqrt and the divisions use correctly rounded fp64 sqrt and division.

#include <algorithm>
#include <vector>

#include "nb_internal.hpp"

namespace nb {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 1024;

struct ForceArgs {
  const float *px, *py, *pz;
  const int *idx;
  int m, n, chunk;
  double eps;
  double *part; // [chunks][3][m]
};

__global__ __launch_bounds__(kBlock) void f64_force_partial(ForceArgs a) {
  __shared__ float sx[kTile], sy[kTile], sz[kTile];
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int raw = a.idx[t < a.m ? t : a.m - 1];
  const bool valid = raw >= 0 && raw < a.n; // an invalid index yields NaN
  const int i = valid ? raw : 0;
  const double xi = a.px[i], yi = a.py[i], zi = a.pz[i];
  const int jbeg = blockIdx.y * a.chunk;
  const int jend = min(jbeg + a.chunk, a.n);
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int j0 = jbeg; j0 < jend; j0 += kTile) {
    __syncthreads();
    for (int l = threadIdx.x; l < kTile; l += kBlock) {
      const int j = j0 + l;
      sx[l] = j < jend ? a.px[j] : 0.0f;
      sy[l] = j < jend ? a.py[j] : 0.0f;
      sz[l] = j < jend ? a.pz[j] : 0.0f;
    }
    __syncthreads();
    const int cnt = min(kTile, jend - j0);
    for (int l = 0; l < cnt; ++l) {
      if (j0 + l == i) continue;
      const double rx = (double)sx[l] - xi;
      const double ry = (double)sy[l] - yi;
      const double rz = (double)sz[l] - zi;
      const double d = rx * rx + ry * ry + rz * rz + a.eps;
      const double inv = 1.0 / (d * sqrt(d));
      fx += rx * inv;
      fy += ry * inv;
      fz += rz * inv;
    }
  }
  if (!valid) fx = fy = fz = __builtin_nan("");
  if (t < a.m) {
    const size_t base = (size_t)blockIdx.y * 3 * a.m;
    a.part[base + t] = fx;
    a.part[base + a.m + t] = fy;
    a.part[base + 2 * a.m + t] = fz;
  }
}

__global__ void f64_force_finish(const double *part, int m, int chunks,
                                 double *fx, double *fy, double *fz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < chunks; ++c)
    for (int k = 0; k < 3; ++k) s[k] += part[((size_t)c * 3 + k) * m + t];
  fx[t] = s[0];
  fy[t] = s[1];
  fz[t] = s[2];
}

// Block-wide fp64 sum of `v` (all threads of the block call it).
__device__ double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

struct EnergyArgs {
  const float *px, *py, *pz, *vx, *vy, *vz;
  int n;
  double eps;
  double *part; // [blocks][5]
};

// Per workgroup: 256 targets' potential against all sources plus their
// kinetic energy and momentum.
__global__ __launch_bounds__(kBlock) void f64_energy_partial(EnergyArgs a) {
  __shared__ float sx[kTile], sy[kTile], sz[kTile];
  __shared__ double red[kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool live = i < a.n;
  const int ic = live ? i : a.n - 1;
  const double xi = a.px[ic], yi = a.py[ic], zi = a.pz[ic];
  double w = 0.0;
  for (int j0 = 0; j0 < a.n; j0 += kTile) {
    __syncthreads();
    for (int l = threadIdx.x; l < kTile; l += kBlock) {
      const int j = j0 + l;
      sx[l] = j < a.n ? a.px[j] : 0.0f;
      sy[l] = j < a.n ? a.py[j] : 0.0f;
      sz[l] = j < a.n ? a.pz[j] : 0.0f;
    }
    __syncthreads();
    const int cnt = min(kTile, a.n - j0);
    for (int l = 0; l < cnt; ++l) {
      if (j0 + l == ic) continue;
      const double rx = (double)sx[l] - xi;
      const double ry = (double)sy[l] - yi;
      const double rz = (double)sz[l] - zi;
      w += 1.0 / sqrt(rx * rx + ry * ry + rz * rz + a.eps);
    }
  }
  double k = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
  if (live) {
    const double ux = a.vx[i], uy = a.vy[i], uz = a.vz[i];
    k = 0.5 * (ux * ux + uy * uy + uz * uz);
    mx = ux;
    my = uy;
    mz = uz;
  } else {
    w = 0.0;
  }
  const double s[5] = {k, w, mx, my, mz};
  for (int c = 0; c < 5; ++c) {
    const double r = block_sum(s[c], red);
    if (threadIdx.x == 0) a.part[(size_t)blockIdx.x * 5 + c] = r;
  }
}

__global__ __launch_bounds__(kBlock) void f64_energy_finish(
    const double *part, int blocks, double G, double *out) {
  __shared__ double red[kBlock];
  for (int c = 0; c < 5; ++c) {
    double s = 0.0;
    for (int b = threadIdx.x; b < blocks; b += kBlock) s += part[(size_t)b * 5 + c];
    const double r = block_sum(s, red);
    if (threadIdx.x == 0) out[c] = (c == 1) ? -0.5 * G * r : r;
  }
}

} // namespace
} // namespace nb

using namespace nb;

extern "C" {

int nb_forces_f64_dev(const nb_params *p, const nb_vec3 *pos,
                      const int32_t *idx, size_t m, double *fx, double *fy,
                      double *fz, size_t n, void *stream) {
  if (!p || !pos) return set_error(NB_ERR_INVALID_ARG, "NULL argument");
  if (n > ((size_t)1 << 28) || m > ((size_t)1 << 28))
    return set_error(NB_ERR_INVALID_ARG, "size exceeds 2^28");
  if (m == 0) return NB_OK;
  if (n == 0) return set_error(NB_ERR_INVALID_ARG, "targets need sources");
  if (!pos->x || !pos->y || !pos->z || !idx || !fx || !fy || !fz)
    return set_error(NB_ERR_INVALID_ARG, "device pointer is NULL");
  // Enough source chunks that the grid holds several workgroups per CU.
  const int tblocks = (int)((m + kBlock - 1) / kBlock);
  int chunks = std::max(1, std::min(1024 / tblocks, (int)((n + kTile - 1) / kTile)));
  ForceArgs a;
  a.px = pos->x;
  a.py = pos->y;
  a.pz = pos->z;
  a.idx = idx;
  a.m = (int)m;
  a.n = (int)n;
  a.chunk = (int)(((n + chunks - 1) / chunks + kTile - 1) / kTile * kTile);
  chunks = (int)((n + a.chunk - 1) / a.chunk);
  a.eps = (double)p->dist_eps;
  hipStream_t s = static_cast<hipStream_t>(stream);
  ScratchLease part; // held until both launches are enqueued
  int st = acquire_scratch(sizeof(double) * 3 * m * (size_t)chunks, s, part);
  if (st) return st;
  a.part = reinterpret_cast<double *>(part.get());
  void *args[] = {&a};
  st = hip_check(hipLaunchKernel(reinterpret_cast<const void *>(f64_force_partial),
                                 dim3(tblocks, chunks), dim3(kBlock), args, 0, s),
                 "fp64 force launch");
  if (st) return st;
  hipLaunchKernelGGL(f64_force_finish, dim3(tblocks), dim3(kBlock), 0, s,
                     a.part, a.m, chunks, fx, fy, fz);
  return hip_check(hipGetLastError(), "fp64 force finish launch");
}

int nb_energy_momentum_dev(const nb_params *p, const nb_vec3 *pos,
                           const nb_vec3 *vel, size_t n, double *out,
                           void *stream) {
  if (!p || !pos || !vel || !out)
    return set_error(NB_ERR_INVALID_ARG, "NULL argument");
  if (n > ((size_t)1 << 28)) return set_error(NB_ERR_INVALID_ARG, "n exceeds 2^28");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (n == 0) {
    hipLaunchKernelGGL(f64_energy_finish, dim3(1), dim3(kBlock), 0, s,
                       static_cast<const double *>(nullptr), 0, 0.0, out);
    return hip_check(hipGetLastError(), "energy finish launch");
  }
  if (!pos->x || !pos->y || !pos->z || !vel->x || !vel->y || !vel->z)
    return set_error(NB_ERR_INVALID_ARG, "device pointer is NULL");
  const int blocks = (int)((n + kBlock - 1) / kBlock);
  EnergyArgs a;
  a.px = pos->x;
  a.py = pos->y;
  a.pz = pos->z;
  a.vx = vel->x;
  a.vy = vel->y;
  a.vz = vel->z;
  a.n = (int)n;
  a.eps = (double)p->dist_eps;
  ScratchLease part; // held until both launches are enqueued
  int st = acquire_scratch(sizeof(double) * 5 * (size_t)blocks, s, part);
  if (st) return st;
  a.part = reinterpret_cast<double *>(part.get());
  void *args[] = {&a};
  st = hip_check(hipLaunchKernel(reinterpret_cast<const void *>(f64_energy_partial),
                                 dim3(blocks), dim3(kBlock), args, 0, s),
                 "energy launch");
  if (st) return st;
  hipLaunchKernelGGL(f64_energy_finish, dim3(1), dim3(kBlock), 0, s,
                     static_cast<const double *>(a.part), blocks,
                     (double)p->G, out);
  return hip_check(hipGetLastError(), "energy finish launch");
}

} // extern "C"
