// Symmetric all-pairs forward step for gfx950 (MI355X): each pair term is
// evaluated once and applied to both particles (Newton's third law).
//
// Same step as particle_interaction<BRANCH> (src/simulator.cu:186-229) in the
// parameter range where the self term vanishes (dist_eps^3 a normal float):
//   F_i = sum_j r_ij * rsqrt(d^3),  r_ij = x_j - x_i,  d = |r_ij|^2 + eps,
// with the reference's per-pair arithmetic (nb_pairs.hpp: pair<MODE_PLAIN>).
// The pair term t_ij = r_ij * K_ij is added to F_i with fma(r, K, F) and to
// F_j with fma(-r, K, S): both are the correctly rounded sum with the exact
// product, so t_ji = -t_ij holds bit for bit.  Only the order in which the
// terms of F_i are summed differs from the target-sharded kernel.
//
// Decomposition (U particles per lane, P = 64 U per wave, ST = 4 P):
//   * particles are grouped in super-tiles of ST; a work item is one
//     workgroup (4 waves) and a chunk of one row of the super-tile pair
//     triangle: super-tile GA against super-tiles GB in [GB0, GB1), GB >= GA;
//   * wave w holds targets 4 GA + w (tile of P) in VGPRs, the chunk's source
//     super-tile sits in LDS;
//   * for GB > GA, in pass b wave w meets source tile j = (w + b) & 3; in
//     each of 64 steps lane l takes the U sources of set (l + s) & 63 from
//     LDS and evaluates the U x U pair terms, adding the reaction -t to U
//     per-source sums that rotate one lane down the wave (DPP wave_rol:1)
//     after each step, so that after 64 steps lane l again holds set l's
//     sums.  Between passes the reaction sums of the four tiles live in LDS
//     (each tile is owned by one wave per pass), and after pass 3 the chunk
//     writes them as the (GA, GB) reaction partial;
//   * for GB == GA (first super-tile of a row) every wave sums the forces of
//     all four tiles on its targets without reactions (self terms are +0);
//   * the targets' forces over the chunk are one force partial;
//   * a finalize kernel sums, per particle, its row's force partials in
//     chunk order and then the reaction partials of its column in GA order,
//     and integrates (integrate_ref).  The order is fixed, so results are
//     deterministic.
// Padding particles (index >= n) sit at 1e30 in every coordinate: any pair
// with a real particle has |r|^2 = +inf, rsqrt(inf) = 0 and r * 0 = 0, so
// they add exact zeros without a mask in the pair loop.
//
// Per pair term: 14 VALU (3 sub, 3 fma, 2 mul, 6 fma) + v_rsq_f32 + 3/U DPP
// moves, against 11 VALU per interaction for the target-sharded kernel that
// evaluates each pair twice.
//
// Build with -ffp-contract=off (explicit fmaf only).

#include "nb_pairs.hpp"

namespace nb {
namespace {

constexpr int kWaves = 4;        // waves per workgroup = tiles per super-tile
constexpr float kFar = 1.0e30f;  // coordinate of padding particles
constexpr int kDppWaveRol1 = 0x134;

// Lane l receives lane (l + 1) mod 64's value.
__device__ __forceinline__ float wave_rol1(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                         kDppWaveRol1, 0xf, 0xf, false));
}

// sum_{l=1..m} ceil(l / L): work items of the rows of lengths 1..m.
__host__ __device__ inline long long items_upto(long long m, int L) {
  const long long q = m / L, r = m % L;
  return (long long)L * q * (q + 1) / 2 + r * (q + 1);
}

// First work item (= force-partial slot) of row ga.
__host__ __device__ inline long long row_first(int ga, int ts, int L) {
  return items_upto(ts, L) - items_upto(ts - ga, L);
}

// Reaction-partial slot of super-tile pair (ga, gb), ga < gb.
__host__ __device__ inline long long pair_slot(int ga, int gb, int ts) {
  return (long long)ga * ts - (long long)ga * (ga + 1) / 2 + (gb - ga - 1);
}

struct SymArgs {
  const float *px, *py, *pz;
  float *vx, *vy, *vz;
  float *nx, *ny, *nz;
  float *fpart; // [items][3][ST] force partials
  float *spart; // [ts (ts - 1) / 2][3][ST] reaction partials
  int n;
  int ts;    // super-tiles
  int chunk; // super-tiles per work item
  StepConsts c;
};

template <int U>
struct Lane {
  float x[U], y[U], z[U];    // targets
  float fx[U], fy[U], fz[U]; // their forces
  float sx[U], sy[U], sz[U]; // reaction sums of the current source set
};

// 64 steps of one wave against one LDS source tile (3 x P floats at
// tx/ty/tz).  With SYM the reaction sums rotate through the wave.
template <int U, bool SYM>
__device__ __forceinline__ void tile_pass(Lane<U> &L, const float4 *tx,
                                          const float4 *ty, const float4 *tz,
                                          int lane, float eps) {
  constexpr int Q = U / 4;
  for (int s = 0; s < 64; ++s) {
    const int m = (lane + s) & 63;
    float xs[U], ys[U], zs[U];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const float4 X = tx[m * Q + q], Y = ty[m * Q + q], Z = tz[m * Q + q];
      xs[4 * q + 0] = X.x; xs[4 * q + 1] = X.y;
      xs[4 * q + 2] = X.z; xs[4 * q + 3] = X.w;
      ys[4 * q + 0] = Y.x; ys[4 * q + 1] = Y.y;
      ys[4 * q + 2] = Y.z; ys[4 * q + 3] = Y.w;
      zs[4 * q + 0] = Z.x; zs[4 * q + 1] = Z.y;
      zs[4 * q + 2] = Z.z; zs[4 * q + 3] = Z.w;
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const float rx = xs[q] - L.x[k];
        const float ry = ys[q] - L.y[k];
        const float rz = zs[q] - L.z[k];
        float d = __builtin_fmaf(rx, rx, eps);
        d = __builtin_fmaf(ry, ry, d);
        d = __builtin_fmaf(rz, rz, d);
        const float inv = __builtin_amdgcn_rsqf((d * d) * d);
        L.fx[k] = __builtin_fmaf(rx, inv, L.fx[k]);
        L.fy[k] = __builtin_fmaf(ry, inv, L.fy[k]);
        L.fz[k] = __builtin_fmaf(rz, inv, L.fz[k]);
        if constexpr (SYM) {
          L.sx[q] = __builtin_fmaf(-rx, inv, L.sx[q]);
          L.sy[q] = __builtin_fmaf(-ry, inv, L.sy[q]);
          L.sz[q] = __builtin_fmaf(-rz, inv, L.sz[q]);
        }
      }
    }
    if constexpr (SYM) {
#pragma unroll
      for (int q = 0; q < U; ++q) {
        L.sx[q] = wave_rol1(L.sx[q]);
        L.sy[q] = wave_rol1(L.sy[q]);
        L.sz[q] = wave_rol1(L.sz[q]);
      }
    }
  }
}

template <int U>
__global__ __launch_bounds__(256) void sym_kernel(SymArgs a) {
  constexpr int P = 64 * U;
  constexpr int ST = kWaves * P;
  __shared__ float4 src[3][ST / 4];  // source super-tile, SoA
  __shared__ float4 sacc[3][ST / 4]; // reaction sums of its four tiles
  float *fsrc = reinterpret_cast<float *>(src);
  float *fsacc = reinterpret_cast<float *>(sacc);
  const int tid = threadIdx.x;
  const int w = tid >> 6, lane = tid & 63;
  const int n = a.n, ts = a.ts, L = a.chunk;

  // Work item -> (row ga, chunk c): rows are enumerated in order, row ga
  // holding ceil((ts - ga) / L) items.
  const long long it = blockIdx.x;
  int lo = 0, hi = ts - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (row_first(mid, ts, L) <= it) lo = mid;
    else hi = mid - 1;
  }
  const int ga = lo;
  const int c = (int)(it - row_first(ga, ts, L));
  const int gb0 = ga + c * L;
  const int gb1 = min(gb0 + L, ts);

  Lane<U> T;
  const int t0 = ga * ST + w * P + lane * U;
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int i = t0 + k;
    const bool ok = i < n;
    T.x[k] = ok ? a.px[i] : kFar;
    T.y[k] = ok ? a.py[i] : kFar;
    T.z[k] = ok ? a.pz[i] : kFar;
    T.fx[k] = T.fy[k] = T.fz[k] = 0.0f;
  }

  for (int gb = gb0; gb < gb1; ++gb) {
    __syncthreads(); // the previous super-tile and its sums are written out
#pragma unroll
    for (int l = 0; l < ST / 256; ++l) {
      const int j = gb * ST + l * 256 + tid;
      const bool ok = j < n;
      fsrc[0 * ST + l * 256 + tid] = ok ? a.px[j] : kFar;
      fsrc[1 * ST + l * 256 + tid] = ok ? a.py[j] : kFar;
      fsrc[2 * ST + l * 256 + tid] = ok ? a.pz[j] : kFar;
    }
    __syncthreads();

    if (gb == ga) {
#pragma unroll 1
      for (int b = 0; b < kWaves; ++b) {
        const int j = (w + b) & (kWaves - 1);
        tile_pass<U, false>(T, src[0] + j * (P / 4), src[1] + j * (P / 4),
                            src[2] + j * (P / 4), lane, a.c.eps);
      }
      continue;
    }

#pragma unroll 1
    for (int b = 0; b < kWaves; ++b) {
      const int j = (w + b) & (kWaves - 1);
      // tile j's sums start at zero with its first visitor (b = 0, w = j)
      const int o = j * P + lane * U;
#pragma unroll
      for (int q = 0; q < U; ++q) {
        T.sx[q] = b ? fsacc[0 * ST + o + q] : 0.0f;
        T.sy[q] = b ? fsacc[1 * ST + o + q] : 0.0f;
        T.sz[q] = b ? fsacc[2 * ST + o + q] : 0.0f;
      }
      tile_pass<U, true>(T, src[0] + j * (P / 4), src[1] + j * (P / 4),
                         src[2] + j * (P / 4), lane, a.c.eps);
#pragma unroll
      for (int q = 0; q < U; ++q) {
        fsacc[0 * ST + o + q] = T.sx[q];
        fsacc[1 * ST + o + q] = T.sy[q];
        fsacc[2 * ST + o + q] = T.sz[q];
      }
      __syncthreads(); // tile j passes to the next wave
    }
    float *sp = a.spart + (size_t)pair_slot(ga, gb, ts) * (3 * ST);
#pragma unroll
    for (int l = 0; l < 3 * ST / 256; ++l) sp[l * 256 + tid] = fsacc[l * 256 + tid];
  }

  float *fp = a.fpart + (size_t)it * (3 * ST) + w * P + lane * U;
#pragma unroll
  for (int k = 0; k < U; ++k) {
    fp[0 * ST + k] = T.fx[k];
    fp[1 * ST + k] = T.fy[k];
    fp[2 * ST + k] = T.fz[k];
  }
}

// Per particle: force partials of its row (chunk order), then reaction
// partials of its column (ga order), then the reference's integrator.
template <int U>
__global__ __launch_bounds__(256) void sym_finalize(SymArgs a) {
  constexpr int ST = kWaves * 64 * U;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n) return;
  const int g = i / ST, o = i % ST;
  const int ts = a.ts, L = a.chunk;
  const long long first = row_first(g, ts, L);
  const int nc = (ts - g + L - 1) / L;
  float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
  for (int c = 0; c < nc; ++c) {
    const float *p = a.fpart + (size_t)(first + c) * (3 * ST) + o;
    f0 = f0 + p[0];
    f1 = f1 + p[ST];
    f2 = f2 + p[2 * ST];
  }
  for (int ga = 0; ga < g; ++ga) {
    const float *p = a.spart + (size_t)pair_slot(ga, g, ts) * (3 * ST) + o;
    f0 = f0 + p[0];
    f1 = f1 + p[ST];
    f2 = f2 + p[2 * ST];
  }
  integrate_ref(a.c, i, f0, f1, f2, a.px, a.py, a.pz, a.vx, a.vy, a.vz, a.nx,
                a.ny, a.nz);
}

using KernelFn = void (*)(SymArgs);

struct SymShape {
  int unroll;
  KernelFn pairs, finalize;
};

const SymShape kSymShapes[] = {
    {4, sym_kernel<4>, sym_finalize<4>},
    {8, sym_kernel<8>, sym_finalize<8>},
};

const SymShape *find_shape(int unroll) {
  for (const SymShape &s : kSymShapes)
    if (s.unroll == unroll) return &s;
  return nullptr;
}

} // namespace

size_t sym_scratch_bytes(int unroll, int chunk, size_t n) {
  if (n == 0 || n > kSymMaxParticles || chunk < 1) return 0;
  const long long st = (long long)kWaves * 64 * unroll;
  const long long ts = ((long long)n + st - 1) / st;
  const long long slots = items_upto(ts, chunk) + ts * (ts - 1) / 2;
  return sizeof(float) * (size_t)slots * 3 * (size_t)st;
}

int launch_forward_sym(const StepConsts &c, int unroll, int chunk,
                       const float *px, const float *py, const float *pz,
                       float *vx, float *vy, float *vz, float *nx, float *ny,
                       float *nz, int n, float *scratch, hipStream_t stream) {
  if (n <= 0) return NB_OK;
  if ((size_t)n > kSymMaxParticles)
    return set_error(NB_ERR_UNSUPPORTED,
                     "symmetric launch shapes take at most 2^20 particles");
  const SymShape *sh = find_shape(unroll);
  if (!sh || chunk < 1)
    return set_error(NB_ERR_UNSUPPORTED, "no symmetric kernel for this shape");
  if (!scratch)
    return set_error(NB_ERR_INVALID_ARG, "symmetric launch needs scratch");
  const int st = kWaves * 64 * unroll;
  const int ts = (n + st - 1) / st;
  const long long items = items_upto(ts, chunk);

  SymArgs a;
  a.px = px;
  a.py = py;
  a.pz = pz;
  a.vx = vx;
  a.vy = vy;
  a.vz = vz;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.fpart = scratch;
  a.spart = scratch + (size_t)items * 3 * st;
  a.n = n;
  a.ts = ts;
  a.chunk = chunk;
  a.c = c;
  void *args[] = {&a};
  int r = hip_check(hipLaunchKernel(reinterpret_cast<const void *>(sh->pairs),
                                    dim3((unsigned)items), dim3(256), args, 0,
                                    stream),
                    "symmetric forward launch");
  if (r) return r;
  return hip_check(
      hipLaunchKernel(reinterpret_cast<const void *>(sh->finalize),
                      dim3((n + 255) / 256), dim3(256), args, 0, stream),
      "symmetric forward finalize launch");
}

} // namespace nb
