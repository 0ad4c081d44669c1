// Forward all-pairs gravity step for gfx950 (MI355X).
//
// Semantics follow particle_interaction<BRANCH|PREDICATED>
// (src/simulator.cu:186-229):
//   F_i = sum_{j != i} r_ij * rsqrt(d^3),  r_ij = x_j - x_i,  d = |r_ij|^2 + eps
//   v_i <- v_i * damping;  v_i <- v_i + (F_i * dt) * G;  x'_i = x_i + v_i * dt
//
// Mapping (one workgroup = BLOCK threads x U targets):
//   * target i's position and force accumulator live in VGPRs, U per thread
//     (i = base + k * BLOCK + lane, so global loads/stores stay coalesced);
//   * sources are streamed through LDS in tiles of TILE particles stored as
//     three SoA float arrays; each lane reads 4 sources per coordinate with
//     one ds_read_b128 at a wave-uniform address (an LDS broadcast: no bank
//     conflicts);
//   * the next tile is fetched from HBM into registers while the current
//     one is consumed (register double-buffering, 2 barriers per tile);
//   * optional source split (grid.y = S chunks) writes partial forces and a
//     second kernel sums the S partials in chunk order and integrates.
// Per interaction the pair loop issues 11 VALU ops + 1 v_rsq_f32:
//   3 sub, 3 fma (|r|^2 + eps), 2 mul (d^3), rsq, 3 fma (F += r * inv).
// No matrix cores: the pair loop is not a dense contraction.
//
// Build with -ffp-contract=off: the integrator must round each product and
// sum separately, as written in the reference; the pair loop uses explicit
// fmaf where a fused operation is intended.

#include "nb_internal.hpp"

namespace nb {
namespace {

struct FwdArgs {
  const float *px, *py, *pz;
  float *vx, *vy, *vz;
  float *nx, *ny, *nz;
  float *part; // [split][3][n] partial forces when split > 1
  int n;
  int jchunk; // sources per grid.y chunk (a multiple of the tile length)
  int split;
  StepConsts c;
};

// Integrator of src/simulator.cu:214-228 in source order (no contraction).
__device__ __forceinline__ void integrate(const FwdArgs &a, int i, float fx,
                                          float fy, float fz) {
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

// Kahan step: s += term with running compensation c (term = r * inv fused
// into the first subtraction).
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
  float inv = __builtin_amdgcn_rsqf((d * d) * d);
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

template <int BLOCK, int U, int TILE, int MODE, bool SPLIT>
__global__ __launch_bounds__(BLOCK) void fwd_kernel(FwdArgs a) {
  static_assert(TILE % BLOCK == 0, "tile must be a multiple of the block");
  static_assert(TILE % 4 == 0, "tile must hold whole float4 groups");
  constexpr int L = TILE / BLOCK; // source loads per thread per coordinate

  __shared__ float4 sx[TILE / 4];
  __shared__ float4 sy[TILE / 4];
  __shared__ float4 sz[TILE / 4];
  float *fsx = reinterpret_cast<float *>(sx);
  float *fsy = reinterpret_cast<float *>(sy);
  float *fsz = reinterpret_cast<float *>(sz);

  const int lane = threadIdx.x;
  const int n = a.n;
  const int base = blockIdx.x * (BLOCK * U);

  Targets<U, MODE> T;
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int i = base + k * BLOCK + lane;
    const int ic = i < n ? i : n - 1; // tail lanes compute on a valid copy
    T.idx[k] = i;
    T.x[k] = a.px[ic];
    T.y[k] = a.py[ic];
    T.z[k] = a.pz[ic];
    T.ax[k] = T.ay[k] = T.az[k] = 0.0f;
    T.cx[k] = T.cy[k] = T.cz[k] = 0.0f;
  }

  int jbeg = 0, jend = n;
  if constexpr (SPLIT) {
    jbeg = blockIdx.y * a.jchunk;
    jend = min(jbeg + a.jchunk, n);
  }
  const float eps = a.c.eps;

  float bx[L], by[L], bz[L];
  auto fetch = [&](int j0) {
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int j = j0 + l * BLOCK + lane;
      const bool ok = j < jend;
      bx[l] = ok ? a.px[j] : 0.0f;
      by[l] = ok ? a.py[j] : 0.0f;
      bz[l] = ok ? a.pz[j] : 0.0f;
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
      for (int q = 0; q < TILE / 4; ++q)
        T.four(sx[q], sy[q], sz[q], j0 + 4 * q, eps);
    } else {
      const int q4 = cnt >> 2;
      for (int q = 0; q < q4; ++q)
        T.four(sx[q], sy[q], sz[q], j0 + 4 * q, eps);
      for (int jj = q4 * 4; jj < cnt; ++jj)
        T.one(fsx[jj], fsy[jj], fsz[jj], j0 + jj, eps);
    }
  }

  constexpr bool kKahan = (MODE == MODE_KAHAN || MODE == MODE_KAHAN_SKIP);
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int i = T.idx[k];
    if (i >= n) continue;
    float fx = T.ax[k], fy = T.ay[k], fz = T.az[k];
    if constexpr (kKahan) {
      fx = fx - T.cx[k];
      fy = fy - T.cy[k];
      fz = fz - T.cz[k];
    }
    if constexpr (SPLIT) {
      const size_t s = blockIdx.y;
      a.part[(s * 3 + 0) * n + i] = fx;
      a.part[(s * 3 + 1) * n + i] = fy;
      a.part[(s * 3 + 2) * n + i] = fz;
    } else {
      integrate(a, i, fx, fy, fz);
    }
  }
}

// Sums the split partial forces in chunk order and integrates.
template <bool KAHAN>
__global__ __launch_bounds__(256) void fwd_finalize(FwdArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n) return;
  const size_t n = a.n;
  float f[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float s = 0.0f, comp = 0.0f;
    for (int k = 0; k < a.split; ++k) {
      const float v = a.part[(k * 3 + c) * n + i];
      if constexpr (KAHAN) {
        const float y = v - comp;
        const float t = s + y;
        comp = (t - s) - y;
        s = t;
      } else {
        s = s + v;
      }
    }
    f[c] = s;
  }
  integrate(a, i, f[0], f[1], f[2]);
}

using KernelFn = void (*)(FwdArgs);

template <int BLOCK, int U, int TILE>
KernelFn shape_kernel(int mode, bool split) {
#define NB_MODE_CASE(M)                                                        \
  case M:                                                                      \
    return split ? fwd_kernel<BLOCK, U, TILE, M, true>                         \
                 : fwd_kernel<BLOCK, U, TILE, M, false>;
  switch (mode) {
    NB_MODE_CASE(MODE_PLAIN)
    NB_MODE_CASE(MODE_PLAIN_SKIP)
    NB_MODE_CASE(MODE_KAHAN)
    NB_MODE_CASE(MODE_KAHAN_SKIP)
    NB_MODE_CASE(MODE_PREDICATED)
  default:
    return nullptr;
  }
#undef NB_MODE_CASE
}

struct ShapeEntry {
  int block, unroll, tile;
  KernelFn (*get)(int, bool);
};

const ShapeEntry kShapes[] = {
    {256, 1, 1024, shape_kernel<256, 1, 1024>},
    {256, 2, 1024, shape_kernel<256, 2, 1024>},
    {256, 4, 1024, shape_kernel<256, 4, 1024>},
    {256, 8, 1024, shape_kernel<256, 8, 1024>},
    {128, 4, 512, shape_kernel<128, 4, 512>},
    {512, 2, 2048, shape_kernel<512, 2, 2048>},
};

// Compiled launch shapes: {block, unroll, tile, split}.
const FwdConfig kConfigs[] = {
    {256, 1, 1024, 1}, {256, 2, 1024, 1}, {256, 2, 1024, 2},
    {256, 4, 1024, 1}, {256, 4, 1024, 2}, {256, 4, 1024, 4},
    {256, 8, 1024, 2}, {256, 8, 1024, 4}, {256, 8, 1024, 8},
    {128, 4, 512, 1},  {128, 4, 512, 2},  {128, 4, 512, 4},
    {512, 2, 2048, 1}, {512, 2, 2048, 2},
};
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);

KernelFn find_kernel(const FwdConfig &c, int mode) {
  for (const ShapeEntry &s : kShapes)
    if (s.block == c.block && s.unroll == c.unroll && s.tile == c.tile)
      return s.get(mode, c.split > 1);
  return nullptr;
}

} // namespace

int forward_config_count() { return kNumConfigs; }

const FwdConfig &forward_config(int cfg) { return kConfigs[cfg]; }

int forward_default_config(size_t n, int acc) {
  (void)acc;
  // Enough workgroups to fill 256 CUs several times over.
  if (n >= (size_t)1 << 17) return 4; // 256 x 4, split 2
  if (n >= (size_t)1 << 14) return 2; // 256 x 2, split 2
  return 0;                           // 256 x 1
}

size_t forward_scratch_bytes(int cfg, size_t n) {
  const FwdConfig &c = kConfigs[cfg];
  return c.split > 1 ? sizeof(float) * 3 * (size_t)c.split * n : 0;
}

int launch_forward(const StepConsts &c, int mode, int cfg, const float *px,
                   const float *py, const float *pz, float *vx, float *vy,
                   float *vz, float *nx, float *ny, float *nz, int n,
                   float *scratch, hipStream_t stream) {
  if (cfg < 0 || cfg >= kNumConfigs)
    return set_error(NB_ERR_INVALID_ARG, "forward config out of range");
  if (n <= 0) return NB_OK;
  const FwdConfig &fc = kConfigs[cfg];
  KernelFn fn = find_kernel(fc, mode);
  if (!fn) return set_error(NB_ERR_UNSUPPORTED, "no kernel for this mode");
  if (fc.split > 1 && !scratch)
    return set_error(NB_ERR_INVALID_ARG, "split launch needs scratch");

  FwdArgs a;
  a.px = px;
  a.py = py;
  a.pz = pz;
  a.vx = vx;
  a.vy = vy;
  a.vz = vz;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.part = scratch;
  a.n = n;
  a.split = fc.split;
  const int per_chunk = (n + fc.split - 1) / fc.split;
  a.jchunk = ((per_chunk + fc.tile - 1) / fc.tile) * fc.tile;
  a.c = c;

  const int targets = fc.block * fc.unroll;
  dim3 grid((n + targets - 1) / targets, fc.split);
  dim3 block(fc.block);
  void *args[] = {&a};
  int st = hip_check(
      hipLaunchKernel(reinterpret_cast<const void *>(fn), grid, block, args,
                      0, stream),
      "forward launch");
  if (st) return st;
  if (fc.split > 1) {
    const bool kahan = (mode == MODE_KAHAN || mode == MODE_KAHAN_SKIP);
    KernelFn fin = kahan ? fwd_finalize<true> : fwd_finalize<false>;
    dim3 g2((n + 255) / 256);
    st = hip_check(hipLaunchKernel(reinterpret_cast<const void *>(fin), g2,
                                   dim3(256), args, 0, stream),
                   "forward finalize launch");
    if (st) return st;
  }
  return NB_OK;
}

} // namespace nb
