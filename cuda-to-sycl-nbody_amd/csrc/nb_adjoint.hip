// Adjoint (reverse-mode derivative) of the all-pairs step for gfx950.
//
// One step (x, v) -> (x', v') of the reference integrator
// (src/simulator.cu:214-228):
//   v' = a v + c F(x),  x' = x + dt v',  a = damping, c = dt * G,
//   F_i = sum_{j != i} f(x_j - x_i),  f(r) = r K,  K = (|r|^2 + eps)^{-3/2}.
// Given the cotangents lx' = dL/dx', lv' = dL/dv':
//   g_j  = lv'_j + dt lx'_j
//   lv_k = a g_k
//   lx_k = lx'_k + c sum_{j != k} M_kj (g_j - g_k),
//   M(r) = df/dr = K (I - 3 r r^T / d),  d = |r|^2 + eps   (symmetric, even).
// The pair term is evaluated as K (h + t3 r) with h = g_j - g_k and
// t3 = -3 (r . h) / d: per interaction 22 VALU ops + one transcendental
// (s = rsqrt(d), q = s^2 = 1/d, K = s^3), the pair quantities r, d, K
// recomputed in registers (no stored forward state beyond the positions).
//
// The fused variant also accumulates F(x) = sum K r in the same pass (3 more
// FMAs) and applies the forward step, so one launch advances the state and
// back-propagates the cotangents at the same positions.  It forms K exactly
// as the forward kernel does (rsqrt(d^3), and q = 1/d as K^2 d^2), so its
// forward outputs are bit-identical to the forward kernel's; its adjoint
// outputs differ from the plain adjoint's by the rounding of K and q.
//
// Mapping: as the forward kernel (U targets per thread in VGPRs, sources in
// LDS tiles of TILE particles, 6 SoA arrays: x, y, z, g_x, g_y, g_z; g is
// formed while the tile is staged).  Optional source split with a finalize
// kernel.  Built with -ffp-contract=off; fused ops are explicit.

#include "nb_internal.hpp"

namespace nb {
namespace {

struct AdjArgs {
  const float *px, *py, *pz;    // x_n
  const float *lxn[3];          // lx' (cotangent of x_{n+1})
  const float *lvn[3];          // lv' (cotangent of v_{n+1})
  float *lx[3], *lv[3];         // outputs: cotangents of x_n, v_n
  float *vx, *vy, *vz;          // fused: velocity in/out
  float *nx, *ny, *nz;          // fused: x_{n+1}
  float *part;                  // split partials [split][6][n]
  int n;
  int jchunk;
  int split;
  StepConsts c;
};

template <int U, bool SKIP, bool FUSED>
struct AdjTargets {
  float x[U], y[U], z[U];
  float gx[U], gy[U], gz[U];
  float ax[U], ay[U], az[U]; // sum M (g_j - g_k)
  float fx[U], fy[U], fz[U]; // fused: sum K r
  int idx[U];

  __device__ __forceinline__ void one(float xj, float yj, float zj, float gjx,
                                      float gjy, float gjz, int j, float eps) {
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const float rx = xj - x[k];
      const float ry = yj - y[k];
      const float rz = zj - z[k];
      float d = __builtin_fmaf(rx, rx, eps);
      d = __builtin_fmaf(ry, ry, d);
      d = __builtin_fmaf(rz, rz, d);
      float K, q;
      if constexpr (FUSED) {
        // K exactly as the forward kernel forms it (rsqrt(d^3)), so the
        // fused variant's forward outputs are bit-identical to its
        const float dd = d * d;
        K = __builtin_amdgcn_rsqf(dd * d);
        q = (K * K) * dd; // 1/d without a second transcendental
      } else {
        // one transcendental (the pipe that bounds this loop's issue):
        // s = d^-1/2, q = s^2 = 1/d, K = s^3
        const float sr = __builtin_amdgcn_rsqf(d);
        q = sr * sr;
        K = q * sr;
      }
      if constexpr (SKIP) K = (j == idx[k]) ? 0.0f : K;
      const float hx = gjx - gx[k];
      const float hy = gjy - gy[k];
      const float hz = gjz - gz[k];
      float rh = rx * hx;
      rh = __builtin_fmaf(ry, hy, rh);
      rh = __builtin_fmaf(rz, hz, rh);
      float t3 = (rh * q) * -3.0f;
      if constexpr (SKIP) t3 = (j == idx[k]) ? 0.0f : t3;
      const float ux = __builtin_fmaf(t3, rx, hx);
      const float uy = __builtin_fmaf(t3, ry, hy);
      const float uz = __builtin_fmaf(t3, rz, hz);
      ax[k] = __builtin_fmaf(K, ux, ax[k]);
      ay[k] = __builtin_fmaf(K, uy, ay[k]);
      az[k] = __builtin_fmaf(K, uz, az[k]);
      if constexpr (FUSED) {
        fx[k] = __builtin_fmaf(rx, K, fx[k]);
        fy[k] = __builtin_fmaf(ry, K, fy[k]);
        fz[k] = __builtin_fmaf(rz, K, fz[k]);
      }
    }
  }
};

__device__ __forceinline__ void load_g(const AdjArgs &a, int j, float &gx,
                                       float &gy, float &gz) {
  const float dt = a.c.dt;
  gx = __builtin_fmaf(dt, a.lxn[0][j], a.lvn[0][j]);
  gy = __builtin_fmaf(dt, a.lxn[1][j], a.lvn[1][j]);
  gz = __builtin_fmaf(dt, a.lxn[2][j], a.lvn[2][j]);
}

// Writes the outputs of target i from its accumulated sums.
template <bool FUSED>
__device__ __forceinline__ void adj_epilogue(const AdjArgs &a, int i, float gx,
                                             float gy, float gz, float ax,
                                             float ay, float az, float fx,
                                             float fy, float fz) {
  const float cG = a.c.dt * a.c.G;
  const float acc[3] = {ax, ay, az};
  const float g[3] = {gx, gy, gz};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    a.lx[c][i] = __builtin_fmaf(cG, acc[c], a.lxn[c][i]);
    a.lv[c][i] = a.c.damping * g[c];
  }
  if constexpr (FUSED) {
    // forward step in the reference's rounding order (src/simulator.cu:214-228)
    float vx = a.vx[i] * a.c.damping;
    float vy = a.vy[i] * a.c.damping;
    float vz = a.vz[i] * a.c.damping;
    vx = vx + (fx * a.c.dt) * a.c.G;
    vy = vy + (fy * a.c.dt) * a.c.G;
    vz = vz + (fz * a.c.dt) * a.c.G;
    a.vx[i] = vx;
    a.vy[i] = vy;
    a.vz[i] = vz;
    a.nx[i] = a.px[i] + vx * a.c.dt;
    a.ny[i] = a.py[i] + vy * a.c.dt;
    a.nz[i] = a.pz[i] + vz * a.c.dt;
  }
}

template <int BLOCK, int U, int TILE, bool SKIP, bool FUSED, bool SPLIT>
__global__ __launch_bounds__(BLOCK) void adj_kernel(AdjArgs a) {
  static_assert(TILE % BLOCK == 0 && TILE % 8 == 0, "tile shape");
  constexpr int L = TILE / BLOCK;
  constexpr int G = TILE / 4;

  __shared__ float4 s4[6][G];
  float *sf[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) sf[c] = reinterpret_cast<float *>(s4[c]);

  const int lane = threadIdx.x;
  const int n = a.n;
  const int base = blockIdx.x * (BLOCK * U);

  AdjTargets<U, SKIP, FUSED> T;
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int i = base + k * BLOCK + lane;
    const int ic = i < n ? i : n - 1;
    T.idx[k] = i;
    T.x[k] = a.px[ic];
    T.y[k] = a.py[ic];
    T.z[k] = a.pz[ic];
    load_g(a, ic, T.gx[k], T.gy[k], T.gz[k]);
    T.ax[k] = T.ay[k] = T.az[k] = 0.0f;
    T.fx[k] = T.fy[k] = T.fz[k] = 0.0f;
  }

  int jbeg = 0, jend = n;
  if constexpr (SPLIT) {
    jbeg = blockIdx.y * a.jchunk;
    jend = min(jbeg + a.jchunk, n);
  }
  const float eps = a.c.eps;

  float b[6][L];
  auto fetch = [&](int j0) {
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int j = j0 + l * BLOCK + lane;
      if (j < jend) {
        b[0][l] = a.px[j];
        b[1][l] = a.py[j];
        b[2][l] = a.pz[j];
        load_g(a, j, b[3][l], b[4][l], b[5][l]);
      } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) b[c][l] = 0.0f;
      }
    }
  };

  if (jbeg < jend) fetch(jbeg);
  for (int j0 = jbeg; j0 < jend; j0 += TILE) {
    __syncthreads();
#pragma unroll
    for (int l = 0; l < L; ++l)
#pragma unroll
      for (int c = 0; c < 6; ++c) sf[c][l * BLOCK + lane] = b[c][l];
    __syncthreads();
    if (j0 + TILE < jend) fetch(j0 + TILE);

    const int cnt = min(TILE, jend - j0);
    const int q4 = cnt >> 2;
#pragma unroll 1
    for (int q = 0; q < q4; ++q) {
      const float4 X = s4[0][q], Y = s4[1][q], Z = s4[2][q];
      const float4 GX = s4[3][q], GY = s4[4][q], GZ = s4[5][q];
      const int jq = j0 + 4 * q;
      T.one(X.x, Y.x, Z.x, GX.x, GY.x, GZ.x, jq + 0, eps);
      T.one(X.y, Y.y, Z.y, GX.y, GY.y, GZ.y, jq + 1, eps);
      T.one(X.z, Y.z, Z.z, GX.z, GY.z, GZ.z, jq + 2, eps);
      T.one(X.w, Y.w, Z.w, GX.w, GY.w, GZ.w, jq + 3, eps);
    }
    for (int jj = q4 * 4; jj < cnt; ++jj)
      T.one(sf[0][jj], sf[1][jj], sf[2][jj], sf[3][jj], sf[4][jj], sf[5][jj],
            j0 + jj, eps);
  }

#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int i = T.idx[k];
    if (i >= n) continue;
    if constexpr (SPLIT) {
      const size_t s = blockIdx.y;
      const size_t nn = n;
      float *p = a.part + s * 6 * nn;
      p[0 * nn + i] = T.ax[k];
      p[1 * nn + i] = T.ay[k];
      p[2 * nn + i] = T.az[k];
      p[3 * nn + i] = T.fx[k];
      p[4 * nn + i] = T.fy[k];
      p[5 * nn + i] = T.fz[k];
    } else {
      adj_epilogue<FUSED>(a, i, T.gx[k], T.gy[k], T.gz[k], T.ax[k], T.ay[k],
                          T.az[k], T.fx[k], T.fy[k], T.fz[k]);
    }
  }
}

template <bool FUSED>
__global__ __launch_bounds__(256) void adj_finalize(AdjArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n) return;
  const size_t nn = a.n;
  float s[6] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < a.split; ++k) {
    const float *p = a.part + (size_t)k * 6 * nn;
#pragma unroll
    for (int c = 0; c < 6; ++c) s[c] = s[c] + p[c * nn + i];
  }
  float gx, gy, gz;
  load_g(a, i, gx, gy, gz);
  adj_epilogue<FUSED>(a, i, gx, gy, gz, s[0], s[1], s[2], s[3], s[4], s[5]);
}

using AdjFn = void (*)(AdjArgs);

template <int B, int U, int T>
AdjFn adj_shape(bool skip, bool fused, bool split) {
  if (skip) {
    if (fused) return split ? adj_kernel<B, U, T, true, true, true>
                            : adj_kernel<B, U, T, true, true, false>;
    return split ? adj_kernel<B, U, T, true, false, true>
                 : adj_kernel<B, U, T, true, false, false>;
  }
  if (fused) return split ? adj_kernel<B, U, T, false, true, true>
                          : adj_kernel<B, U, T, false, true, false>;
  return split ? adj_kernel<B, U, T, false, false, true>
               : adj_kernel<B, U, T, false, false, false>;
}

struct AdjShape {
  int block, unroll, tile;
  AdjFn (*get)(bool, bool, bool);
};

const AdjShape kAdjShapes[] = {
    {256, 1, 1024, adj_shape<256, 1, 1024>},
    {256, 2, 1024, adj_shape<256, 2, 1024>},
    {256, 4, 1024, adj_shape<256, 4, 1024>},
    {128, 4, 512, adj_shape<128, 4, 512>},
};

// {block, unroll, tile, split[, scheme]}; the symmetric shapes (nb_sym.hip)
// are {256, particles per lane, super-tile, super-tiles per work item}.
const FwdConfig kAdjConfigs[] = {
    {256, 1, 1024, 1}, {256, 2, 1024, 1}, {256, 2, 1024, 2},
    {256, 4, 1024, 1}, {256, 4, 1024, 2}, {256, 4, 1024, 4},
    {128, 4, 512, 2},  {128, 4, 512, 4},  {256, 1, 1024, 2},
    {256, 2, 1024, 4}, {256, 1, 1024, 4}, {256, 1, 1024, 8},
    {256, 2, 1024, 8}, {256, 4, 1024, 8},
    {256, 4, 1024, 1, SCHEME_SYMMETRIC}, {256, 4, 1024, 2, SCHEME_SYMMETRIC},
    {256, 8, 2048, 1, SCHEME_SYMMETRIC},
    {256, 4, 256, 1, SCHEME_SYMMETRIC_TILES},
    {256, 2, 128, 1, SCHEME_SYMMETRIC_TILES},
    {256, 2, 512, 1, SCHEME_SYMMETRIC}, {256, 2, 512, 2, SCHEME_SYMMETRIC},
    {256, 4, 1024, 4, SCHEME_SYMMETRIC}, {256, 2, 512, 4, SCHEME_SYMMETRIC},
};
constexpr int kNumAdj = sizeof(kAdjConfigs) / sizeof(kAdjConfigs[0]);

AdjFn find_adj(const FwdConfig &c, bool skip, bool fused) {
  for (const AdjShape &s : kAdjShapes)
    if (s.block == c.block && s.unroll == c.unroll && s.tile == c.tile)
      return s.get(skip, fused, c.split > 1);
  return nullptr;
}

} // namespace

int adjoint_config_count() { return kNumAdj; }
const FwdConfig &adjoint_config(int cfg) { return kAdjConfigs[cfg]; }

int adjoint_default_config(size_t n, bool skip) {
  // Measured on MI355X (disk ICs), adjoint / fused us.  Tile pairs, 2 per
  // lane: n = 4096 28.7 / 35.7, 16384 124.5 / 146.3 (super-tile pairs, 2 per
  // lane: 141.7 / 181.0), 25600 245.8 / 289.7 (217.0 / 282.2), 32768 379.9 /
  // 449.9.  Super-tile pairs, 2 per lane, 1 / 2 per item: 32768 354.7 / 471.7 vs
  // 403.2 / 542.8, 65536 1219 / 1644 (tile pairs 1447 / 1771), 131072
  // 4780 / 6511 vs 4779 / 6478, 262144 19280 / 25892 vs 19168 / 25601
  // (4 per lane 23660 / 31135, 8 per lane 24283 / 34040).  Best
  // target-sharded: 8192 65.8 / 74.2, 262144 38.7 / 43.5 ms.  Past 2^18 two
  // per lane need four super-tiles per item and bands of source super-tiles
  // (4 per lane at 2^21: 1501 / 1986 ms).
  if (n <= kSymBandedMaxParticles && n > kSymSmallMaxParticles)
    return 22; // super-tile pairs, 2 per lane, 4 per item, in bands
  if (n <= kSymMaxParticles) { // skip or not: self pairs masked in-kernel
    if (n >= 98304) return 20; // super-tile pairs, 2 per lane, 2 per item
    if (n >= 20480) return 19; // super-tile pairs, 2 per lane, 1 per item
    if (n >= 1024) return 18;  // tile pairs, 2 per lane
  }
  if (n >= (size_t)1 << 17) return 8; // 256 x 1, split 2
  if (n < 4096) return 8;
  return 9;                           // 256 x 2, split 4
}

size_t adjoint_scratch_bytes(int cfg, size_t n) {
  const FwdConfig &c = kAdjConfigs[cfg];
  if (c.scheme != SCHEME_TARGETS) { // either variant (3 or 6 sums)
    const bool tiles = c.scheme == SCHEME_SYMMETRIC_TILES;
    const size_t fused = sym_scratch_bytes(c.unroll, c.split, n, 6, tiles, true);
    const size_t plain = sym_scratch_bytes(c.unroll, c.split, n, 3, tiles, true);
    return fused && plain ? std::max(fused, plain) : 0;
  }
  return c.split > 1 ? sizeof(float) * 6 * (size_t)c.split * n : 0;
}

int launch_adjoint(const StepConsts &c, bool skip, bool fused, int cfg,
                   const float *const pos[3], const float *const lxn[3],
                   const float *const lvn[3], float *const lx[3],
                   float *const lv[3], float *const vel[3],
                   float *const nxt[3], int n, float *scratch,
                   hipStream_t stream) {
  if (cfg < 0 || cfg >= kNumAdj)
    return set_error(NB_ERR_INVALID_ARG, "adjoint config out of range");
  if (n <= 0) return NB_OK;
  const FwdConfig &fc = kAdjConfigs[cfg];
  if (fc.scheme != SCHEME_TARGETS) { // self pairs masked in-kernel if skip
    return launch_adjoint_sym(c, fused, fc.unroll, fc.split,
                              fc.scheme == SCHEME_SYMMETRIC_TILES, pos, lxn, lvn,
                              lx, lv, vel, nxt, n, scratch, stream);
  }
  AdjFn fn = find_adj(fc, skip, fused);
  if (!fn) return set_error(NB_ERR_UNSUPPORTED, "no adjoint kernel for shape");
  if (fc.split > 1 && !scratch)
    return set_error(NB_ERR_INVALID_ARG, "split launch needs scratch");
  AdjArgs a = {};
  a.px = pos[0];
  a.py = pos[1];
  a.pz = pos[2];
  for (int d = 0; d < 3; ++d) {
    a.lxn[d] = lxn[d];
    a.lvn[d] = lvn[d];
    a.lx[d] = lx[d];
    a.lv[d] = lv[d];
  }
  if (fused) {
    a.vx = vel[0];
    a.vy = vel[1];
    a.vz = vel[2];
    a.nx = nxt[0];
    a.ny = nxt[1];
    a.nz = nxt[2];
  }
  a.part = scratch;
  a.n = n;
  a.split = fc.split;
  const int per_chunk = (n + fc.split - 1) / fc.split;
  a.jchunk = ((per_chunk + fc.tile - 1) / fc.tile) * fc.tile;
  a.c = c;
  const int targets = fc.block * fc.unroll;
  dim3 grid((n + targets - 1) / targets, fc.split);
  void *args[] = {&a};
  int st = hip_check(hipLaunchKernel(reinterpret_cast<const void *>(fn), grid,
                                     dim3(fc.block), args, 0, stream),
                     "adjoint launch");
  if (st) return st;
  if (fc.split > 1) {
    AdjFn fin = fused ? adj_finalize<true> : adj_finalize<false>;
    st = hip_check(hipLaunchKernel(reinterpret_cast<const void *>(fin),
                                   dim3((n + 255) / 256), dim3(256), args, 0,
                                   stream),
                   "adjoint finalize launch");
  }
  return st;
}

} // namespace nb
