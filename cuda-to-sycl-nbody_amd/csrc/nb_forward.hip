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
//   * sources stream through LDS tiles of TILE particles (nb_pairs.hpp:
//     accumulate_sources), read back as wave-uniform b128 broadcasts;
//   * optional source split (grid.y = S chunks) writes partial forces and a
//     second kernel sums the S partials in chunk order and integrates.
// The pair arithmetic (11 VALU + v_rsq_f32 per interaction) is in
// nb_pairs.hpp.  No matrix cores: the pair loop is not a dense contraction.
//
// Build with -ffp-contract=off: the integrator must round each product and
// sum separately, as written in the reference.

#include "nb_pairs.hpp"


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

template <int BLOCK, int U, int TILE, int MODE, bool SPLIT>
__global__ __launch_bounds__(BLOCK) void fwd_kernel(FwdArgs a) {
  __shared__ SourceTile<TILE> sm;
  const int n = a.n;
  Targets<U, MODE> T;
  T.load(a.px, a.py, a.pz, blockIdx.x * (BLOCK * U), BLOCK, threadIdx.x, n);

  int jbeg = 0, jend = n;
  if constexpr (SPLIT) {
    jbeg = blockIdx.y * a.jchunk;
    jend = min(jbeg + a.jchunk, n);
  }
  accumulate_sources<BLOCK, TILE>(sm, T, a.px, a.py, a.pz, jbeg, jend,
                                  a.c.eps);

#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int i = T.idx[k];
    if (i >= n) continue;
    float fx, fy, fz;
    T.force(k, fx, fy, fz);
    if constexpr (SPLIT) {
      const size_t s = blockIdx.y;
      a.part[(s * 3 + 0) * n + i] = fx;
      a.part[(s * 3 + 1) * n + i] = fy;
      a.part[(s * 3 + 2) * n + i] = fz;
    } else {
      integrate_ref(a.c, i, fx, fy, fz, a.px, a.py, a.pz, a.vx, a.vy, a.vz,
                    a.nx, a.ny, a.nz);
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
  integrate_ref(a.c, i, f[0], f[1], f[2], a.px, a.py, a.pz, a.vx, a.vy, a.vz,
                a.nx, a.ny, a.nz);
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
    {128, 1, 512, shape_kernel<128, 1, 512>},
    {128, 2, 512, shape_kernel<128, 2, 512>},
    {512, 1, 2048, shape_kernel<512, 1, 2048>},
    {512, 2, 2048, shape_kernel<512, 2, 2048>},
};

// Compiled launch shapes: {block, unroll, tile, split[, scheme]}.  Index 0
// must stay the unsplit 256 x 1 shape (a valid choice for every n).  The
// symmetric shapes (nb_sym.hip) are {256, particles per lane, super-tile,
// super-tiles per work item}.
const FwdConfig kConfigs[] = {
    {256, 1, 1024, 1}, {256, 1, 1024, 2}, {256, 1, 1024, 4},
    {256, 1, 1024, 8}, {256, 2, 1024, 1}, {256, 2, 1024, 2},
    {256, 2, 1024, 4}, {256, 2, 1024, 8}, {256, 4, 1024, 1},
    {256, 4, 1024, 2}, {256, 4, 1024, 4}, {256, 4, 1024, 8},
    {128, 1, 512, 1},  {128, 1, 512, 2},  {128, 1, 512, 4},
    {128, 2, 512, 2},  {128, 2, 512, 4},  {512, 1, 2048, 1},
    {512, 1, 2048, 2}, {512, 2, 2048, 1}, {512, 2, 2048, 2},
    {512, 2, 2048, 4},
    {256, 8, 2048, 1, SCHEME_SYMMETRIC}, {256, 8, 2048, 2, SCHEME_SYMMETRIC},
    {256, 8, 2048, 4, SCHEME_SYMMETRIC}, {256, 4, 1024, 1, SCHEME_SYMMETRIC},
    {256, 4, 1024, 2, SCHEME_SYMMETRIC}, {256, 4, 1024, 4, SCHEME_SYMMETRIC},
    {256, 1, 1024, 16}, {256, 2, 1024, 16}, {128, 1, 512, 8},
    {128, 1, 512, 16},
    {256, 4, 256, 1, SCHEME_SYMMETRIC_TILES},
    {256, 2, 128, 1, SCHEME_SYMMETRIC_TILES},
    {256, 2, 512, 1, SCHEME_SYMMETRIC}, {256, 2, 512, 2, SCHEME_SYMMETRIC},
    {256, 2, 512, 4, SCHEME_SYMMETRIC},
    {256, 2, 128, 2, SCHEME_SYMMETRIC_TILES}, // two items per tile pair
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

int forward_default_config(size_t n, int mode) {
  // the symmetric shapes mask the self pairs where eps^3 is not a normal
  // float at the cost of the diagonal passes only: same choice
  if (mode == MODE_PLAIN_SKIP) mode = MODE_PLAIN;
  if (mode == MODE_KAHAN_SKIP) mode = MODE_KAHAN;
  // Measured on MI355X (disk ICs), us per step.  Tile pairs, 2 per lane
  // (two items per tile pair below 6656: n = 4096 15.5, 6144 20.9; one item
  // with priority steps from there: 7168 23.1, 8192 27.2; see
  // launch_forward_sym): n = 4096 22.6 (target-sharded 128 x 1 split 8: 29), 8192 38.1 (38.7),
  // 12800 57.5 (82.5), 32768 238.4 (323.6; 4 per lane 266.3).  Super-tile
  // pairs, 2 per lane, 1 / 2 per item: 49152 481.8 / 532.7 (tile pairs
  // 499.0), 65536 810.7 / 842.7 (867.9), 98304 1764 / 1780 (4 per lane
  // 1860, 8 per lane 1915), 131072 3118 / 3104 (4 per lane 3176, 8 per lane
  // 3299), 262144 12236 / 12133 (8 per lane 11804).  Past 2^19 two
  // super-tiles of 8 per lane per item run as fast as one and need 25 %
  // less scratch (6042 Gint/s at 2^20); past 2^20 four per item keep the
  // partials at 16 GB for 2^21 particles, and past 2^21 the step runs in
  // bands of source super-tiles with at most 8 GB of partials each (the
  // 4-per-lane shape with 4 per item is banded past 2^20).
  if (mode == MODE_PLAIN && n <= kSymBandedMaxParticles) {
    if (n > kSymMaxParticles) return 24; // 8 per lane, 4 per item (banded past 2^21)
    if (n > ((size_t)1 << 19)) return 23; // symmetric, 8 per lane, 2 per item
    if (n >= ((size_t)1 << 18)) return 22; // symmetric, 8 per lane, 1 per item
    if (n >= 98304) return 35;             // symmetric, 2 per lane, 2 per item
    if (n >= 49152) return 34;             // symmetric, 2 per lane, 1 per item
    if (n >= 6656) return 33;              // symmetric tile pairs, 2 per lane
    if (n >= 4096) return 37;              // the same, two items per tile pair
  }
  // Kahan, Gint/s: tile pairs, 2 per lane: n = 16384 1961 (best
  // target-sharded 1664), 32768 2408 (1847); super-tile pairs, 2 per lane,
  // 1 / 2 per item: 65536 2774 / 2674, 131072 2873 / 2867, 262144 2916 /
  // 2929 (4 per lane 2411, target-sharded 2373).  The compensated sums'
  // dependency chains bound this mode, so the shapes with the most waves
  // win; below 12288 the target-sharded ones (8192: 1229 vs 1128).  Above
  // 2^18 four super-tiles per item run in bands of source super-tiles (16 GB
  // of (sum, compensation) partials per band).
  if (mode == MODE_KAHAN) {
    if (n > ((size_t)1 << 18) && n <= kSymBandedMaxParticles)
      return 36;               // symmetric, 2 per lane, 4 per item, banded
    if (n >= ((size_t)1 << 17) && n <= ((size_t)1 << 18))
      return 35;               // symmetric, 2 per lane, 2 per item
    if (n >= 49152 && n < ((size_t)1 << 17))
      return 34;               // symmetric, 2 per lane, 1 per item
    if (n >= 12288 && n < 49152)
      return 33;               // symmetric tile pairs, 2 per lane
  }
  // Target-sharded, Gint/s: n = 4096: 128 x 1 split 8 629 (256 x 1 split 8
  // 480); 8192: 256 x 1 split 8 1746 (split 16 1629); 16384 .. 32768:
  // 256 x 1 split 16 2843 .. 3432 (split 8 2792 .. 3292); 65536: 256 x 4
  // split 8 3402 (256 x 1 split 8 3279).
  if (n >= (size_t)1 << 16) return 11; // 256 x 4, tile 1024, split 8
  if (n >= 12288) return 28;           // 256 x 1, tile 1024, split 16
  if (n >= 6144) return 3;             // 256 x 1, tile 1024, split 8
  return 30;                           // 128 x 1, tile 512, split 8
}

size_t forward_scratch_bytes(int cfg, size_t n, int mode) {
  const FwdConfig &c = kConfigs[cfg];
  if (c.scheme != SCHEME_TARGETS) // (sum, compensation) pairs for Kahan
  {
    const bool kahan = mode == MODE_KAHAN || mode == MODE_KAHAN_SKIP;
    return sym_scratch_bytes(c.unroll, c.split, n, kahan ? 6 : 3,
                             c.scheme == SCHEME_SYMMETRIC_TILES,
                             !kahan || c.unroll <= 4);
  }
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
  if (fc.scheme != SCHEME_TARGETS) {
    // the kernels mask the self pairs themselves when eps^3 is not a normal
    // float (the SKIP modes)
    if (mode == MODE_PREDICATED)
      return set_error(NB_ERR_UNSUPPORTED,
                       "symmetric launch shapes take BRANCH only");
    const bool kahan = mode == MODE_KAHAN || mode == MODE_KAHAN_SKIP;
    if (kahan && fc.unroll > 4)
      return set_error(NB_ERR_UNSUPPORTED,
                       "symmetric Kahan shapes have 2 or 4 particles per lane");
    return launch_forward_sym(c, fc.unroll, fc.split,
                              fc.scheme == SCHEME_SYMMETRIC_TILES, kahan, px,
                              py, pz, vx, vy, vz, nx, ny, nz, n, scratch,
                              stream);
  }
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
