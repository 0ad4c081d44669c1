// Symmetric all-pairs kernels for gfx950 (MI355X): each pair term is
// evaluated once and applied to both particles of the pair.
//
// Two pair rules share one decomposition:
//   ForceRule    the forward step, particle_interaction<BRANCH>
//                (src/simulator.cu:186-229):
//                  F_i = sum_{j != i} r_ij K_ij,  r_ij = x_j - x_i,
//                  K = rsqrt(d^3),  d = |r|^2 + eps  (nb_pairs.hpp arithmetic);
//   AdjointRule  the reverse-mode step of the same integrator (nb_adjoint.hip
//                derives it): A_k = sum_j K (h + t3 r),  h = g_j - g_k,
//                t3 = -3 (r . h) / d, optionally fused with the forward F.
// Both pair terms are odd under exchanging the two particles (r -> -r,
// h -> -h, d, K, t3 unchanged).  The reaction on particle j is accumulated
// negated, S' = fma(u, K, S'), which is bit for bit -fma(-u, K, S) (rounding
// is symmetric under negation), and subtracted in the finalize: particle j
// receives exactly the negative of the term added to particle i.  Only the
// order in which each particle's terms are summed differs from the
// target-sharded kernels.
//
// Decomposition (U particles per lane, P = 64 U per wave, ST = 4 P):
//   * particles are grouped in super-tiles of ST; a work item is one
//     workgroup (4 waves) and a chunk of one row of the super-tile pair
//     triangle: super-tile GA against super-tiles GB in [GB0, GB1), GB >= GA;
//   * wave w holds the inputs and sums of targets 4 GA + w (tile of P) in
//     VGPRs; the source super-tile sits in LDS;
//   * for GB > GA, in pass b wave w meets source tile j = (w + b) & 3; in
//     each of 64 steps lane l takes the U sources of set (l + s) & 63 from
//     LDS and evaluates the U x U pair terms, adding the (negated) reaction
//     to U per-source sums that rotate one lane down the wave (wave_rol1 or
//     lds_rol1) after each step, so that after 64 steps lane l again holds
//     set l's sums.  Between passes the reaction sums of the four tiles live in LDS
//     (each tile is owned by one wave per pass); after pass 3 the chunk
//     writes them as the (GA, GB) reaction partial;
//   * for GB == GA (first super-tile of a row) each wave meets its own tile
//     without reactions (self terms are +0) and then, with 2 or 4 per lane,
//     the tiles above it with reactions (each of the six tile pairs once),
//     subtracting at the end the reactions its tile received; with 8 per
//     lane it meets the other three tiles without reactions as well;
//   * the targets' sums over the chunk are one row partial;
//   * a finalize kernel sums, per particle, its row partials in chunk order
//     and then subtracts the reaction partials of its column in GA order,
//     and applies the rule's epilogue.  The order is fixed: results are
//     deterministic.
// Two more forms reuse the same passes: one tile pair per wave for small n
// (sym_tile_kernel) and the super-tile pairs of many independent systems on
// one persistent queue (sym_batch_kernel).
// Padding (index >= n): ForceRule puts padding particles at 1e30 in every
// coordinate, so any pair with a real particle has |r|^2 = +inf,
// rsqrt(inf) = 0 and r * 0 = 0 (exact zeros, no mask; a pair of two padding
// particles only touches padding sums).  AdjointRule masks K and t3 of pairs
// with a padding particle in the items that contain the last, partial
// (super-)tile (a product r . h of a far padding particle could overflow).
// Self pairs: with dist_eps^3 a normal float the self term is exactly +0
// (r = 0, finite K); otherwise (eps = 0 or tiny) the pass of each wave over
// its own tile, the only one holding self pairs, masks them as the
// reference's `continue` skips them.
//
// Per pair term (two interactions): forward 14 VALU + v_rsq_f32 + 3/U
// reaction moves through the LDS crossbar (ds_bpermute, no VALU slot);
// adjoint 25 VALU + v_rsq_f32 (K and 1/d from rsqrt(d)) + 3/U DPP moves;
// fused adjoint 33 VALU + v_rsq_f32 (K as the forward rule forms it, 1/d =
// K^2 d^2) + 6/U DPP moves.  The target-sharded kernels spend 11 (forward)
// and 22 (adjoint) VALU per interaction.
//
// Build with -ffp-contract=off (explicit fmaf only).

#include "nb_pairs.hpp"

#include <vector>

namespace nb {
namespace {

constexpr int kWaves = 4;        // waves per workgroup = tiles per super-tile
constexpr size_t kTilePrioMaxParticles = 32768; // see launch_forward_sym
constexpr size_t kKahanTilePrioMaxParticles = 16384; // the same, Kahan sums
constexpr float kFar = 1.0e30f;  // coordinate of ForceRule padding particles
constexpr int kDppWaveRol1 = 0x134;

// Lane l receives lane (l + 1) mod 64's value.  Two forms of the same move:
//   wave_rol1  a DPP wave_rol:1 move on the VALU.  Every lane is written
//              (full row and bank masks), so the "old" operand is the source
//              itself: with a constant there the compiler materialises it
//              with an extra v_mov per call;
//   lds_rol1   ds_bpermute_b32 through the LDS crossbar (no LDS memory): it
//              takes no VALU issue slot.  Inside the pair loop's mix a DPP
//              move costs about as much VALU time as a v_rsq_f32 (10-14 SIMD
//              cycles, measured with probe/issue_rates.hip), while the
//              permute's latency is hidden (each sum is rotated as soon as
//              it is complete and read again only in the next step).  The
//              forward rules rotate with lds_rol1 (N = 262144: 11.3 vs 11.9
//              ms with plain sums, 22.5 vs 23.8 ms with Kahan sums; the
//              64-system batch 4.43 vs 4.52 ms); the adjoint rules, whose
//              loops read twice the LDS per step already, keep DPP for the
//              adjoint sums (lds_rol1 for all of them measured 1.5 % slower
//              for the adjoint, 4 % for the fused step) and the fused step
//              rotates its force sums with lds_rol1.
__device__ __forceinline__ float wave_rol1(float v) {
  const int b = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(b, b, kDppWaveRol1, 0xf, 0xf, false));
}

__device__ __forceinline__ float lds_rol1(float v) {
  const int lane =
      (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  return __builtin_bit_cast(
      float, __builtin_amdgcn_ds_bpermute(((lane + 1) & 63) << 2,
                                          __builtin_bit_cast(int, v)));
}

// sum_{l=1..m} ceil(l / L): work items of the rows of lengths 1..m.
__host__ __device__ inline long long items_upto(long long m, int L) {
  const long long q = m / L, r = m % L;
  return (long long)L * q * (q + 1) / 2 + r * (q + 1);
}

// First work item (= row-partial slot) of row ga.
__host__ __device__ inline long long row_first(int ga, int ts, int L) {
  return items_upto(ts, L) - items_upto(ts - ga, L);
}

// Reaction-partial slot of super-tile pair (ga, gb), ga < gb.
__host__ __device__ inline long long pair_slot(int ga, int gb, int ts) {
  return (long long)ga * ts - (long long)ga * (ga + 1) / 2 + (gb - ga - 1);
}

__host__ __device__ inline long long tri_slot(int ga, int gb, int ts) {
  // row-major index of (ga, gb), ga <= gb, in the upper triangle with diagonal
  return (long long)ga * ts - (long long)ga * (ga - 1) / 2 + (gb - ga);
}

struct SymArgs {
  const float *px, *py, *pz;
  const float *lxn[3], *lvn[3]; // adjoint: cotangents of the step's outputs
  float *lx[3], *lv[3];         // adjoint: cotangents of its inputs
  float *vx, *vy, *vz;          // velocity in/out (forward, fused adjoint)
  float *nx, *ny, *nz;          // next positions (forward, fused adjoint)
  float *fpart;                 // [items][ACC][ST] row partials
  float *spart;                 // [ts (ts - 1) / 2][ACC][ST] reaction partials
  int n;
  int ts;    // super-tiles
  int chunk; // super-tiles per work item
  StepConsts c;
  // banded launches (sym_band_*): source super-tiles [b0, b0 + bw) of this
  // band, first band flag, running per-particle sums [3][n]
  int b0, bw, first;
  float *acc3;
};

// Finalize arithmetic of the plain rules: row partials are added, the
// negated reaction partials subtracted, in the order the caller visits them.
// fetch reads one partial ([ACC] floats, stride apart); add_v / sub_v fold it
// into the running sums.
template <int ACC>
struct PlainPartials {
  static __device__ __forceinline__ void fetch(float (&v)[ACC], const float *p,
                                               int stride) {
#pragma unroll
    for (int e = 0; e < ACC; ++e) v[e] = p[e * stride];
  }
  static __device__ __forceinline__ void add_v(float (&f)[ACC],
                                               const float (&v)[ACC]) {
#pragma unroll
    for (int e = 0; e < ACC; ++e) f[e] = f[e] + v[e];
  }
  static __device__ __forceinline__ void sub_v(float (&f)[ACC],
                                               const float (&v)[ACC]) {
#pragma unroll
    for (int e = 0; e < ACC; ++e) f[e] = f[e] - v[e];
  }
  static __device__ __forceinline__ void sub_reaction(float (&f)[ACC],
                                                      const float *p,
                                                      int stride) {
    float v[ACC];
    fetch(v, p, stride);
    sub_v(f, v);
  }
};

struct ForceRule : PlainPartials<3> {
  static constexpr int SRC = 3; // x, y, z
  static constexpr int ACC = 3; // F
  static constexpr bool MASKED_TAIL = false;
  // every sum rotates with lds_rol1 (see wave_rol1)
  static constexpr bool rotate_lds(int) { return true; }

  static __device__ __forceinline__ void load(const SymArgs &a, int j,
                                              float (&v)[SRC]) {
    const bool ok = j < a.n;
    v[0] = ok ? a.px[j] : kFar;
    v[1] = ok ? a.py[j] : kFar;
    v[2] = ok ? a.pz[j] : kFar;
  }

  // Source s on target t: a += r K.  With SYM, rs accumulates +r K too: rs
  // holds the negated reaction sum of the source (the finalize subtracts it),
  // which keeps every accumulation a two-operand v_fmac_f32.  With MASK the
  // term of a pair that is not `valid` is zero (the self pair when eps^3 is
  // not a normal float, as the reference's `continue` skips it).
  template <bool SYM, bool MASK>
  static __device__ __forceinline__ void pair(const float (&s)[SRC],
                                              const float (&t)[SRC],
                                              float (&a)[ACC],
                                              float (&rs)[ACC], float eps,
                                              bool valid) {
    const float rx = s[0] - t[0];
    const float ry = s[1] - t[1];
    const float rz = s[2] - t[2];
    float d = __builtin_fmaf(rx, rx, eps);
    d = __builtin_fmaf(ry, ry, d);
    d = __builtin_fmaf(rz, rz, d);
    float K = __builtin_amdgcn_rsqf((d * d) * d);
    if constexpr (MASK) K = valid ? K : 0.0f;
    a[0] = __builtin_fmaf(rx, K, a[0]);
    a[1] = __builtin_fmaf(ry, K, a[1]);
    a[2] = __builtin_fmaf(rz, K, a[2]);
    if constexpr (SYM) {
      rs[0] = __builtin_fmaf(rx, K, rs[0]);
      rs[1] = __builtin_fmaf(ry, K, rs[1]);
      rs[2] = __builtin_fmaf(rz, K, rs[2]);
    }
  }

  static __device__ __forceinline__ void finish(const SymArgs &a, int i,
                                                const float (&f)[ACC]) {
    integrate_ref(a.c, i, f[0], f[1], f[2], a.px, a.py, a.pz, a.vx, a.vy,
                  a.vz, a.nx, a.ny, a.nz);
  }
};

// Kahan-compensated forward rule (NB_ACC_KAHAN): the target sums and the
// (negated) reaction sums each carry a running compensation c (kahan_add,
// nb_pairs.hpp: y = fma(r, K, -c), t = s + y, c = (t - s) - y, s = t; the
// sum is s - c).  A partial is (s_x, s_y, s_z, c_x, c_y, c_z); the finalize
// Kahan-adds the partials' values s - c.
struct KahanForceRule {
  static constexpr int SRC = 3;
  static constexpr int ACC = 6;
  static constexpr bool MASKED_TAIL = false;
  // every sum rotates with lds_rol1 (see wave_rol1)
  static constexpr bool rotate_lds(int) { return true; }

  static __device__ __forceinline__ void load(const SymArgs &a, int j,
                                              float (&v)[SRC]) {
    ForceRule::load(a, j, v);
  }

  template <bool SYM, bool MASK>
  static __device__ __forceinline__ void pair(const float (&s)[SRC],
                                              const float (&t)[SRC],
                                              float (&a)[ACC],
                                              float (&rs)[ACC], float eps,
                                              bool valid) {
    const float rx = s[0] - t[0];
    const float ry = s[1] - t[1];
    const float rz = s[2] - t[2];
    float d = __builtin_fmaf(rx, rx, eps);
    d = __builtin_fmaf(ry, ry, d);
    d = __builtin_fmaf(rz, rz, d);
    float K = __builtin_amdgcn_rsqf((d * d) * d);
    if constexpr (MASK) K = valid ? K : 0.0f; // as ForceRule
    kahan_add(rx, K, a[0], a[3]);
    kahan_add(ry, K, a[1], a[4]);
    kahan_add(rz, K, a[2], a[5]);
    if constexpr (SYM) {
      kahan_add(rx, K, rs[0], rs[3]);
      kahan_add(ry, K, rs[1], rs[4]);
      kahan_add(rz, K, rs[2], rs[5]);
    }
  }

  // f: running (sum, compensation) of each coordinate over the partials
  static __device__ __forceinline__ void kahan_value(float (&f)[ACC], int e,
                                                     float v) {
    const float y = v - f[3 + e];
    const float t = f[e] + y;
    f[3 + e] = (t - f[e]) - y;
    f[e] = t;
  }
  static __device__ __forceinline__ void fetch(float (&v)[ACC], const float *p,
                                               int stride) {
#pragma unroll
    for (int e = 0; e < ACC; ++e) v[e] = p[e * stride];
  }
  static __device__ __forceinline__ void add_v(float (&f)[ACC],
                                               const float (&v)[ACC]) {
#pragma unroll
    for (int e = 0; e < 3; ++e) kahan_value(f, e, v[e] - v[3 + e]);
  }
  static __device__ __forceinline__ void sub_v(float (&f)[ACC],
                                               const float (&v)[ACC]) {
#pragma unroll
    for (int e = 0; e < 3; ++e) kahan_value(f, e, v[3 + e] - v[e]);
  }
  static __device__ __forceinline__ void sub_reaction(float (&f)[ACC],
                                                      const float *p,
                                                      int stride) {
    float v[ACC];
    fetch(v, p, stride);
    sub_v(f, v);
  }

  static __device__ __forceinline__ void finish(const SymArgs &a, int i,
                                                const float (&f)[ACC]) {
    integrate_ref(a.c, i, f[0] - f[3], f[1] - f[4], f[2] - f[5], a.px, a.py,
                  a.pz, a.vx, a.vy, a.vz, a.nx, a.ny, a.nz);
  }
};

template <bool FUSED>
struct AdjointRule : PlainPartials<FUSED ? 6 : 3> {
  static constexpr int SRC = 6; // x, y, z, g_x, g_y, g_z
  static constexpr int ACC = FUSED ? 6 : 3; // sum M (g_j - g_k) [, F]
  static constexpr bool MASKED_TAIL = true;
  // the adjoint sums rotate by DPP; the fused step's force sums (3..5) with
  // lds_rol1 (N = 262144: 24.4 vs 25.1 ms; all six through the crossbar: 26.1
  // ms; any of the plain adjoint's through it: no gain)
  static constexpr bool rotate_lds(int c) { return FUSED && c >= 3; }

  // g = lv' + dt lx' (the cotangent the pair sum is taken of)
  static __device__ __forceinline__ void load(const SymArgs &a, int j,
                                              float (&v)[SRC]) {
    if (j < a.n) {
      v[0] = a.px[j];
      v[1] = a.py[j];
      v[2] = a.pz[j];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        v[3 + c] = __builtin_fmaf(a.c.dt, a.lxn[c][j], a.lvn[c][j]);
    } else {
#pragma unroll
      for (int c = 0; c < SRC; ++c) v[c] = 0.0f;
    }
  }

  template <bool SYM, bool MASK>
  static __device__ __forceinline__ void pair(const float (&s)[SRC],
                                              const float (&t)[SRC],
                                              float (&a)[ACC],
                                              float (&rs)[ACC], float eps,
                                              bool valid) {
    const float rx = s[0] - t[0];
    const float ry = s[1] - t[1];
    const float rz = s[2] - t[2];
    float d = __builtin_fmaf(rx, rx, eps);
    d = __builtin_fmaf(ry, ry, d);
    d = __builtin_fmaf(rz, rz, d);
    float K, q;
    if constexpr (FUSED) { // K exactly as the forward kernels form it
      // 1/d = K^2 d^2: two multiplies instead of v_rcp_f32 (2 % faster at
      // N = 262144, 8 % with tile pairs at N = 16384)
      const float dd = d * d;
      K = __builtin_amdgcn_rsqf(dd * d);
      q = (K * K) * dd;
    } else { // one transcendental (8 % faster): s = d^-1/2, q = s^2, K = s^3
      const float sr = __builtin_amdgcn_rsqf(d);
      q = sr * sr;
      K = q * sr;
    }
    const float hx = s[3] - t[3];
    const float hy = s[4] - t[4];
    const float hz = s[5] - t[5];
    float rh = rx * hx;
    rh = __builtin_fmaf(ry, hy, rh);
    rh = __builtin_fmaf(rz, hz, rh);
    float t3 = (rh * q) * -3.0f;
    if constexpr (MASK) {
      K = valid ? K : 0.0f;
      t3 = valid ? t3 : 0.0f;
    }
    const float ux = __builtin_fmaf(t3, rx, hx);
    const float uy = __builtin_fmaf(t3, ry, hy);
    const float uz = __builtin_fmaf(t3, rz, hz);
    a[0] = __builtin_fmaf(K, ux, a[0]);
    a[1] = __builtin_fmaf(K, uy, a[1]);
    a[2] = __builtin_fmaf(K, uz, a[2]);
    if constexpr (SYM) { // negated reaction sums, as in ForceRule
      rs[0] = __builtin_fmaf(K, ux, rs[0]);
      rs[1] = __builtin_fmaf(K, uy, rs[1]);
      rs[2] = __builtin_fmaf(K, uz, rs[2]);
    }
    if constexpr (FUSED) {
      a[3] = __builtin_fmaf(rx, K, a[3]);
      a[4] = __builtin_fmaf(ry, K, a[4]);
      a[5] = __builtin_fmaf(rz, K, a[5]);
      if constexpr (SYM) {
        rs[3] = __builtin_fmaf(rx, K, rs[3]);
        rs[4] = __builtin_fmaf(ry, K, rs[4]);
        rs[5] = __builtin_fmaf(rz, K, rs[5]);
      }
    }
  }

  // lx = lx' + dt G sum,  lv = damping g; fused: the forward step too.
  static __device__ __forceinline__ void finish(const SymArgs &a, int i,
                                                const float (&f)[ACC]) {
    const float cG = a.c.dt * a.c.G;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float g = __builtin_fmaf(a.c.dt, a.lxn[c][i], a.lvn[c][i]);
      a.lx[c][i] = __builtin_fmaf(cG, f[c], a.lxn[c][i]);
      a.lv[c][i] = a.c.damping * g;
    }
    if constexpr (FUSED)
      integrate_ref(a.c, i, f[3], f[4], f[5], a.px, a.py, a.pz, a.vx, a.vy,
                    a.vz, a.nx, a.ny, a.nz);
  }
};

// LDS position of super-tile particle o (particle order o = j P + m U + q:
// tile j, source set m, member q).  Within a tile, float4 slot qq * 64 + m
// holds members 4 qq .. 4 qq + 3 of set m, so the 64 lanes of a step (sets
// (l + s) mod 64, one quad each) read one contiguous 1 KiB block: no bank
// conflicts.
// With U = 2 (tile-pair form only) set m's two members sit in float2 slot m.
template <int U>
__device__ __forceinline__ int lds_slot(int o) {
  constexpr int P = 64 * U;
  const int j = o / P, r = o % P;
  const int m = r / U, q = r % U;
  if constexpr (U == 2) return j * P + m * 2 + q;
  return j * P + (q >> 2) * 256 + m * 4 + (q & 3);
}

template <class R, int U>
struct Lane {
  float t[U][R::SRC];  // target inputs
  float a[U][R::ACC];  // target sums
  float rs[U][R::ACC]; // reaction sums of the current source set
};

// 64 steps of one wave against source tile j of the LDS super-tile (src:
// R::SRC arrays of ST floats in lds_slot order; src0 = global index of the
// tile's first particle).  With SYM the reaction sums rotate through the
// wave.  With MASK, pairs with a padding particle contribute nothing (vt:
// target validity; sources are valid below index n).  With SELF (a wave's
// pass over its own tile, with MASK) the self pairs contribute nothing
// either: needed when eps^3 is not a normal float.
// [s0, s1): the steps of the pass this call runs (a half pass leaves lane l
// with the reaction sums of set (l + s1) mod 64).
// PRIO: the wave lowers its issue priority at each quarter of the pass
// (s_setprio 3, 2, 1, 0), so that the SIMD's arbiter, which otherwise
// favours the oldest wave, lets the waves that lag behind catch up.  Used by
// the tile-pair forms (forward up to n = 32768, Kahan sums up to 16384,
// adjoint and two-item forms at every size they run), where most waves of the launch start at once
// and, without it, finish between 10 and 37 us into a 40 us kernel (n =
// 12800, per-wave clock stamps): the last quarter of the launch ran on one
// or two waves per SIMD.
template <class R, int U, bool SYM, bool MASK, bool SELF = false,
          bool PRIO = false>
__device__ __forceinline__ void tile_pass(Lane<R, U> &L, const float *src,
                                          int j, int lane, float eps,
                                          const bool (&vt)[U], int src0,
                                          int n, int s0 = 0, int s1 = 64) {
  static_assert(!SELF || (MASK && !SYM), "self pairs: masked own-tile pass");
  constexpr int P = 64 * U;
  constexpr int ST = kWaves * P;
  constexpr int Q = U / 4;
  static_assert(U == 2 || U % 4 == 0, "particles per lane: 2 or a multiple of 4");
  auto step = [&](int s) {
    const int m = (lane + s) & 63;
    float sv[U][R::SRC];
#pragma unroll
    for (int c = 0; c < R::SRC; ++c) {
      if constexpr (U == 2) {
        const float2 v = reinterpret_cast<const float2 *>(src + c * ST + j * P)[m];
        sv[0][c] = v.x;
        sv[1][c] = v.y;
      } else {
        const float4 *p4 =
            reinterpret_cast<const float4 *>(src + c * ST + j * P);
#pragma unroll
        for (int qq = 0; qq < Q; ++qq) {
          const float4 v = p4[qq * 64 + m];
          sv[4 * qq + 0][c] = v.x;
          sv[4 * qq + 1][c] = v.y;
          sv[4 * qq + 2][c] = v.z;
          sv[4 * qq + 3][c] = v.w;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const bool vs = !MASK || (src0 + m * U + q < n);
#pragma unroll
      for (int k = 0; k < U; ++k) {
        // SELF (own tile): at step 0 member q of the lane's own set is its
        // target q, the pair the reference skips
        const bool self = SELF && s == 0 && q == k;
        R::template pair<SYM, MASK>(sv[q], L.t[k], L.a[k], L.rs[q], eps,
                                    vs && vt[k] && !self);
      }
    }
    if constexpr (SYM) {
#pragma unroll
      for (int q = 0; q < U; ++q)
#pragma unroll
        for (int c = 0; c < R::ACC; ++c)
          L.rs[q][c] =
              R::rotate_lds(c) ? lds_rol1(L.rs[q][c]) : wave_rol1(L.rs[q][c]);
    }
  };
  if constexpr (PRIO) {
    // the priority is set once per quarter, outside the step loop
    const int qlen = (s1 - s0) >> 2; // 16 or 8 steps
    for (int quarter = 0; quarter < 4; ++quarter) {
      if (quarter == 0) __builtin_amdgcn_s_setprio(3);
      else if (quarter == 1) __builtin_amdgcn_s_setprio(2);
      else if (quarter == 2) __builtin_amdgcn_s_setprio(1);
      else __builtin_amdgcn_s_setprio(0);
      const int e = s0 + (quarter + 1) * qlen;
      for (int s = s0 + quarter * qlen; s < e; ++s) step(s);
    }
  } else {
    for (int s = s0; s < s1; ++s) step(s);
  }
}

template <class R, int U, bool MASK>
__device__ __forceinline__ void super_pair(Lane<R, U> &L, const float *src,
                                           float *sacc, bool diag, int w,
                                           int lane, float eps,
                                           const bool (&vt)[U], int gb0src,
                                           int n) {
  constexpr int P = 64 * U;
  constexpr int ST = kWaves * P;
  // Diagonal super-tile, two measured strategies: with 8 per lane every wave
  // meets all four tiles without reactions (16 tile passes, all waves equally
  // long: 11.77 ms at N = 262144 against 12.01 ms the other way); with 2 or
  // 4 per lane the six tile pairs with reactions (4 per lane: 13.13 -> 12.30
  // ms).
  constexpr bool kDiagReactions = U <= 4;
  // the pass of a wave over its own tile holds the self pairs: +0 when eps^3
  // is a normal float, masked otherwise
  const bool self_masked = diag && !self_term_vanishes(eps);
  if (diag && !kDiagReactions) {
    if (self_masked)
      tile_pass<R, U, false, true, true>(L, src, w, lane, eps, vt,
                                         gb0src + w * P, n);
    else
      tile_pass<R, U, false, MASK>(L, src, w, lane, eps, vt, gb0src + w * P,
                                   n);
#pragma unroll 1
    for (int b = 1; b < kWaves; ++b) {
      const int j = (w + b) & (kWaves - 1);
      tile_pass<R, U, false, MASK>(L, src, j, lane, eps, vt, gb0src + j * P,
                                   n);
    }
    return;
  }
  if (diag) {
    // Pass 0: each wave against its own tile, without reactions (self terms
    // +0 or masked), and zeroes that tile's reaction sums.  Pass b > 0: wave w
    // meets tile j = w + b only if j < 4, with reactions, so each of the six
    // tile pairs is evaluated once; tile j is visited by one wave per pass.
    // Last, each wave subtracts the negated reactions its own tile received.
    {
      const int o = w * P + lane * U;
#pragma unroll
      for (int q = 0; q < U; ++q)
#pragma unroll
        for (int c = 0; c < R::ACC; ++c) sacc[c * ST + o + q] = 0.0f;
    }
    if (self_masked)
      tile_pass<R, U, false, true, true>(L, src, w, lane, eps, vt,
                                         gb0src + w * P, n);
    else
      tile_pass<R, U, false, MASK>(L, src, w, lane, eps, vt, gb0src + w * P,
                                   n);
    __syncthreads();
#pragma unroll 1
    for (int b = 1; b < kWaves; ++b) {
      const int j = w + b;
      if (j < kWaves) {
        const int o = j * P + lane * U;
#pragma unroll
        for (int q = 0; q < U; ++q)
#pragma unroll
          for (int c = 0; c < R::ACC; ++c) L.rs[q][c] = sacc[c * ST + o + q];
        tile_pass<R, U, true, MASK>(L, src, j, lane, eps, vt, gb0src + j * P,
                                    n);
#pragma unroll
        for (int q = 0; q < U; ++q)
#pragma unroll
          for (int c = 0; c < R::ACC; ++c) sacc[c * ST + o + q] = L.rs[q][c];
      }
      __syncthreads(); // tile j passes to the next wave
    }
    const int o = w * P + lane * U;
#pragma unroll
    for (int k = 0; k < U; ++k) R::sub_reaction(L.a[k], sacc + o + k, ST);
    return;
  }
#pragma unroll 1
  for (int b = 0; b < kWaves; ++b) {
    const int j = (w + b) & (kWaves - 1);
    // tile j's sums start at zero with its first visitor (b = 0, w = j)
    const int o = j * P + lane * U;
#pragma unroll
    for (int q = 0; q < U; ++q)
#pragma unroll
      for (int c = 0; c < R::ACC; ++c)
        L.rs[q][c] = b ? sacc[c * ST + o + q] : 0.0f;
    tile_pass<R, U, true, MASK>(L, src, j, lane, eps, vt, gb0src + j * P, n);
#pragma unroll
    for (int q = 0; q < U; ++q)
#pragma unroll
      for (int c = 0; c < R::ACC; ++c) sacc[c * ST + o + q] = L.rs[q][c];
    __syncthreads(); // tile j passes to the next wave
  }
}

// Targets of super-tile ga held by this lane (wave w): inputs, zeroed sums
// and validity (index < n).
template <class R, int U>
__device__ __forceinline__ void load_targets(Lane<R, U> &T, bool (&vt)[U],
                                             const SymArgs &a, int ga, int w,
                                             int lane) {
  constexpr int P = 64 * U;
  constexpr int ST = kWaves * P;
  const int t0 = ga * ST + w * P + lane * U;
#pragma unroll
  for (int k = 0; k < U; ++k) {
    R::load(a, t0 + k, T.t[k]);
    vt[k] = t0 + k < a.n;
#pragma unroll
    for (int e = 0; e < R::ACC; ++e) T.a[k][e] = 0.0f;
  }
}

// Super-tile pair (ga, gb), gb >= ga: stages source super-tile gb in LDS,
// adds its terms to the targets' sums and, for gb > ga, writes the negated
// reaction sums of its particles to sp ([ACC][ST] floats).  Called by every
// thread of the workgroup (it contains barriers).
template <class R, int U>
__device__ __forceinline__ void pair_step(Lane<R, U> &T, const bool (&vt)[U],
                                          const SymArgs &a, float *src,
                                          float *sacc, int ga, int gb,
                                          float *sp) {
  constexpr int P = 64 * U;
  constexpr int ST = kWaves * P;
  const int tid = threadIdx.x;
  const int w = tid >> 6, lane = tid & 63;
  const int n = a.n;
  __syncthreads(); // the previous super-tile and its sums are written out
#pragma unroll
  for (int l = 0; l < ST / 256; ++l) {
    float v[R::SRC];
    R::load(a, gb * ST + l * 256 + tid, v);
    const int o = lds_slot<U>(l * 256 + tid);
#pragma unroll
    for (int e = 0; e < R::SRC; ++e) src[e * ST + o] = v[e];
  }
  __syncthreads();

  // only the last super-tile holds padding particles (gb >= ga)
  bool masked = false;
  if constexpr (R::MASKED_TAIL) masked = (n % ST) != 0 && (gb + 1) * ST > n;
  if (masked)
    super_pair<R, U, R::MASKED_TAIL>(T, src, sacc, gb == ga, w, lane, a.c.eps,
                                     vt, gb * ST, n);
  else
    super_pair<R, U, false>(T, src, sacc, gb == ga, w, lane, a.c.eps, vt,
                            gb * ST, n);
  if (gb == ga) return;
#pragma unroll
  for (int l = 0; l < R::ACC * ST / 256; ++l)
    sp[l * 256 + tid] = sacc[l * 256 + tid];
}

// The targets' sums as one row partial ([ACC][ST] floats at fp).
template <class R, int U>
__device__ __forceinline__ void store_row(const Lane<R, U> &T, float *fp,
                                          int w, int lane) {
  constexpr int P = 64 * U;
  constexpr int ST = kWaves * P;
  float *q = fp + w * P + lane * U;
#pragma unroll
  for (int k = 0; k < U; ++k)
#pragma unroll
    for (int e = 0; e < R::ACC; ++e) q[e * ST + k] = T.a[k][e];
}

// f += partial(0), partial(1), ... in that order (SUB: f -= ...), with
// kInFlight partials loaded ahead of the running sum: the loads do not
// depend on it, so a finalize thread waits about once per kInFlight
// partials instead of once per partial.  The order, and so the result, is
// that of the plain loop.
constexpr int kInFlight = 8;

template <class R, bool SUB, class At>
__device__ __forceinline__ void sum_partials(float (&f)[R::ACC], int count,
                                             At at, int stride) {
  int k = 0;
  for (; k + kInFlight <= count; k += kInFlight) {
    float v[kInFlight][R::ACC];
#pragma unroll
    for (int u = 0; u < kInFlight; ++u) R::fetch(v[u], at(k + u), stride);
#pragma unroll
    for (int u = 0; u < kInFlight; ++u) {
      if constexpr (SUB) R::sub_v(f, v[u]);
      else R::add_v(f, v[u]);
    }
  }
  if (k < count) { // the last count - k < kInFlight partials, loaded together
    float v[kInFlight][R::ACC];
#pragma unroll
    for (int u = 0; u < kInFlight; ++u)
      if (k + u < count) R::fetch(v[u], at(k + u), stride);
#pragma unroll
    for (int u = 0; u < kInFlight; ++u) {
      if (k + u < count) {
        if constexpr (SUB) R::sub_v(f, v[u]);
        else R::add_v(f, v[u]);
      }
    }
  }
}

// A finalize workgroup of 256 threads serves 256 / RUNS particles: with
// RUNS > 1, wave group r sums run r of each particle's partial sequence (its
// row partials, then its negated reaction partials: `rows + reacs` partials
// cut into RUNS contiguous runs), and the first group adds the runs' sums in
// run order.  The order is fixed either way, so results are deterministic.
// RUNS = 4 pays where one thread per particle leaves most CUs idle (the
// tile-pair form: 5.3 vs 10.7 us at n = 12800; 8 and 16 runs made the whole
// step 0-1 % and 3-7 % slower at n = 4096 .. 40960); with enough particles to
// fill the chip one run per particle is faster (super-tile finalize at
// n = 262144: 82 vs 91 us).
template <class R, int RUNS, class RowAt, class ReacAt>
__device__ __forceinline__ void run_sum(float (&f)[R::ACC], int rows,
                                        int reacs, int run, RowAt row_at,
                                        ReacAt reac_at, int stride) {
  const int len = rows + reacs;
  const int k0 = run * len / RUNS, k1 = (run + 1) * len / RUNS;
#pragma unroll
  for (int e = 0; e < R::ACC; ++e) f[e] = 0.0f;
  if (k0 < rows)
    sum_partials<R, false>(
        f, min(k1, rows) - k0, [&](int k) { return row_at(k0 + k); }, stride);
  const int r0 = max(k0, rows) - rows, r1 = k1 - rows;
  if (r1 > r0)
    sum_partials<R, true>(
        f, r1 - r0, [&](int k) { return reac_at(r0 + k); }, stride);
}

// First group: f (run 0's sum) plus runs 1 .. RUNS-1 in order.  A run's
// running sum has the layout and value of a row partial (Kahan: value = sum
// - compensation), so it is added as one.  Called by every thread (barrier);
// returns true in the threads that hold the total.
template <class R, int RUNS>
__device__ __forceinline__ bool combine_runs(float (&f)[R::ACC],
                                             float (*red)[256]) {
  if constexpr (RUNS == 1) {
    return true;
  } else {
    constexpr int PW = 256 / RUNS;
    const int t = threadIdx.x;
#pragma unroll
    for (int e = 0; e < R::ACC; ++e) red[e][t] = f[e];
    __syncthreads();
    if (t >= PW) return false;
#pragma unroll
    for (int q = 1; q < RUNS; ++q) {
      float v[R::ACC];
#pragma unroll
      for (int e = 0; e < R::ACC; ++e) v[e] = red[e][t + q * PW];
      R::add_v(f, v);
    }
    return true;
  }
}

constexpr int kTileRuns = 4; // runs per particle, tile-pair finalize

template <class R, int U>
__global__ __launch_bounds__(256) void sym_kernel(SymArgs a) {
  constexpr int P = 64 * U;
  constexpr int ST = kWaves * P;
  __shared__ float4 src4[R::SRC][ST / 4];  // source super-tile, SoA
  __shared__ float4 sacc4[R::ACC][ST / 4]; // reaction sums of its 4 tiles
  float *src = reinterpret_cast<float *>(src4);
  float *sacc = reinterpret_cast<float *>(sacc4);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ts = a.ts, L = a.chunk;

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

  Lane<R, U> T;
  bool vt[U];
  load_targets(T, vt, a, ga, w, lane);
  for (int gb = gb0; gb < gb1; ++gb)
    pair_step(T, vt, a, src, sacc, ga, gb,
              gb > ga ? a.spart + (size_t)pair_slot(ga, gb, ts) * (R::ACC * ST)
                      : nullptr);
  store_row(T, a.fpart + (size_t)it * (R::ACC * ST), w, lane);
}

// Per particle: row partials (chunk order), minus the negated reaction
// partials of its column (ga order); then the rule's epilogue.
template <class R, int U>
__global__ __launch_bounds__(256) void sym_finalize(SymArgs a) {
  constexpr int ST = kWaves * 64 * U;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n) return;
  const int g = i / ST, o = i % ST;
  const int ts = a.ts, L = a.chunk;
  const long long first = row_first(g, ts, L);
  float f[R::ACC];
  run_sum<R, 1>(
      f, (ts - g + L - 1) / L, g, 0,
      [&](int c) { return a.fpart + (size_t)(first + c) * (R::ACC * ST) + o; },
      [&](int ga) {
        return a.spart + (size_t)pair_slot(ga, g, ts) * (R::ACC * ST) + o;
      },
      ST);
  R::finish(a, i, f);
}

// ---- banded launches (n past the single-launch limit) ----------------------
//
// The super-tile pair triangle is cut by source super-tile into bands
// [B0, B0 + W): band rows GA < B0 + W, each with the sources GB in
// [max(GA, B0), B0 + W) in chunks of L.  Per band: the pair kernel writes the
// band's row partials (one per item) and reaction partials (one per pair
// with GB > GA), a fold kernel adds, per particle, its band rows minus its
// band reactions to a running sum (band order), and after the last band a
// finish kernel applies the rule's epilogue (the integrator; the adjoint's
// cotangent update).  The partials of one band stay within a fixed
// budget whatever n is; the order of every sum is fixed.

// Reaction slot of pair (ga, gb), gb > ga, gb in the band [b0, b0 + w).
__host__ __device__ inline long long band_pair_slot(int ga, int gb, int b0,
                                                    int w) {
  return ga < b0 ? (long long)ga * w + (gb - b0)
                 : (long long)b0 * w + pair_slot(ga - b0, gb - b0, w);
}

__host__ __device__ inline long long band_items(int b0, int w, int L) {
  return (long long)b0 * ((w + L - 1) / L) + items_upto(w, L);
}

__host__ __device__ inline long long band_reactions(int b0, int w) {
  return (long long)b0 * w + (long long)w * (w - 1) / 2;
}

template <class R, int U>
__global__ __launch_bounds__(256) void sym_band_kernel(SymArgs a) {
  constexpr int P = 64 * U;
  constexpr int ST = kWaves * P;
  __shared__ float4 src4[R::SRC][ST / 4];
  __shared__ float4 sacc4[R::ACC][ST / 4];
  float *src = reinterpret_cast<float *>(src4);
  float *sacc = reinterpret_cast<float *>(sacc4);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L = a.chunk, b0 = a.b0, W = a.bw, b1 = b0 + W;
  const int c0 = (W + L - 1) / L; // items of a row above the band
  const long long head = (long long)b0 * c0;
  const long long it = blockIdx.x;
  int ga, gb0;
  if (it < head) {
    ga = (int)(it / c0);
    gb0 = b0 + (int)(it % c0) * L;
  } else { // the band's own triangle, rows enumerated in order
    const long long j = it - head;
    int lo = 0, hi = W - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (row_first(mid, W, L) <= j) lo = mid;
      else hi = mid - 1;
    }
    ga = b0 + lo;
    gb0 = ga + (int)(j - row_first(lo, W, L)) * L;
  }
  const int gb1 = min(gb0 + L, b1);
  Lane<R, U> T;
  bool vt[U];
  load_targets(T, vt, a, ga, w, lane);
  for (int gb = gb0; gb < gb1; ++gb)
    pair_step(T, vt, a, src, sacc, ga, gb,
              gb > ga ? a.spart + (size_t)band_pair_slot(ga, gb, b0, W) *
                                      (R::ACC * ST)
                      : nullptr);
  store_row(T, a.fpart + (size_t)it * (R::ACC * ST), w, lane);
}

// Per particle: its band rows (chunk order) minus its band reactions (GA
// order), added to the running sum (band order; Kahan sums carry their
// compensation from band to band).
template <class R, int U>
__global__ __launch_bounds__(256) void sym_band_fold(SymArgs a) {
  constexpr int ST = kWaves * 64 * U;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n) return;
  const int g = i / ST, o = i % ST;
  const int L = a.chunk, b0 = a.b0, W = a.bw, b1 = b0 + W;
  if (g >= b1 && !a.first) return; // no band term; its sum stays
  float f[R::ACC];
#pragma unroll
  for (int e = 0; e < R::ACC; ++e) f[e] = 0.0f;
  if (g < b1) {
    const int c0 = (W + L - 1) / L;
    const long long first =
        g < b0 ? (long long)g * c0
               : (long long)b0 * c0 + row_first(g - b0, W, L);
    const int rows = g < b0 ? c0 : (b1 - g + L - 1) / L;
    run_sum<R, 1>(
        f, rows, g >= b0 ? g : 0, 0,
        [&](int c) { return a.fpart + (size_t)(first + c) * (R::ACC * ST) + o; },
        [&](int gs) {
          return a.spart +
                 (size_t)band_pair_slot(gs, g, b0, W) * (R::ACC * ST) + o;
        },
        ST);
  }
  float *acc = a.acc3 + i;
  if (!a.first) {
    float r[R::ACC];
    R::fetch(r, acc, a.n);
    R::add_v(r, f);
#pragma unroll
    for (int e = 0; e < R::ACC; ++e) f[e] = r[e];
  }
#pragma unroll
  for (int e = 0; e < R::ACC; ++e) acc[(size_t)e * a.n] = f[e];
}

template <class R, int U>
__global__ __launch_bounds__(256) void sym_band_finish(SymArgs a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.n) return;
  float f[R::ACC];
#pragma unroll
  for (int e = 0; e < R::ACC; ++e) f[e] = a.acc3[(size_t)e * a.n + i];
  R::finish(a, i, f);
}

// ---- tile-pair items (small n) ----------------------------------------------
//
// Work item = one tile pair (A, B), A <= B, of P = 64 U particles, owned by
// one wave; a workgroup holds four independent waves (one LDS tile each) and
// has a single barrier, after staging.  The wave keeps the reaction sums in
// registers for the whole pair.  Slot tri_slot(A, B, T) holds the row
// partial [ACC][P] and, for A < B, the negated reaction partial [ACC][P].
// T (T + 1) / 2 wave items fill the chip at sizes where the super-tile items
// (T / 4)^2 / 2 workgroups of four waves) do not.

// HALVES = 2 splits each tile pair into two work items of 32 steps (the
// first and second half of the pass; item = 2 x pair + half), each with its
// own row and reaction partial: twice the items, so that at small n the
// launch holds more items than resident waves and the dispatcher keeps every
// SIMD fed until a short tail, for twice the partials in the finalize.
template <class R, int U, bool PRIO = false, int HALVES = 1>
__global__ __launch_bounds__(256) void sym_tile_kernel(SymArgs a) {
  static_assert(HALVES == 1 || HALVES == 2, "one or two items per tile pair");
  constexpr int P = 64 * U;
  constexpr int ST = kWaves * P;
  __shared__ float4 src4[R::SRC][ST / 4]; // wave w's source tile is tile w
  float *src = reinterpret_cast<float *>(src4);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = a.n, T = a.ts;
  const long long items = (long long)HALVES * T * (T + 1) / 2;
  const long long item = (long long)blockIdx.x * kWaves + w;
  const bool live = item < items;
  const long long pair = item / HALVES;
  const int s0 = (int)(item % HALVES) * (64 / HALVES), s1 = s0 + 64 / HALVES;
  int A = 0, B = 0;
  if (live) { // row A: the last row whose first slot is <= pair
    int lo = 0, hi = T - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tri_slot(mid, mid, T) <= pair) lo = mid;
      else hi = mid - 1;
    }
    A = lo;
    B = A + (int)(pair - tri_slot(A, A, T));
  }
  // targets first: their loads are in flight while the source tile is staged
  Lane<R, U> L;
  bool vt[U];
  const int t0 = A * P + lane * U;
#pragma unroll
  for (int k = 0; k < U; ++k) {
    R::load(a, t0 + k, L.t[k]);
    vt[k] = t0 + k < n;
#pragma unroll
    for (int e = 0; e < R::ACC; ++e) L.a[k][e] = 0.0f;
  }
#pragma unroll
  for (int r = 0; r < U; ++r) {
    const int o = r * 64 + lane;
    float v[R::SRC];
    R::load(a, B * P + o, v);
    const int d = lds_slot<U>(w * P + o);
#pragma unroll
    for (int e = 0; e < R::SRC; ++e) src[e * ST + d] = v[e];
  }
  __syncthreads(); // the only barrier: every wave has staged its tile
  if (!live) return;
  float *slot = a.fpart + (size_t)item * (2 * R::ACC * P);
  bool masked = false;
  if constexpr (R::MASKED_TAIL) masked = (n % P) != 0 && B == T - 1;
  if (A == B) {
    if (!self_term_vanishes(a.c.eps)) // self pairs masked, as the tail
      tile_pass<R, U, false, true, true, PRIO>(L, src, w, lane, a.c.eps, vt, B * P,
                                         n, s0, s1);
    else if (masked)
      tile_pass<R, U, false, R::MASKED_TAIL, false, PRIO>(L, src, w, lane,
                                                          a.c.eps, vt,
                                             B * P, n, s0, s1);
    else
      tile_pass<R, U, false, false, false, PRIO>(L, src, w, lane, a.c.eps,
                                                 vt, B * P, n, s0, s1);
  } else {
#pragma unroll
    for (int q = 0; q < U; ++q)
#pragma unroll
      for (int e = 0; e < R::ACC; ++e) L.rs[q][e] = 0.0f;
    if (masked)
      tile_pass<R, U, true, R::MASKED_TAIL, false, PRIO>(L, src, w, lane,
                                                         a.c.eps, vt,
                                            B * P, n, s0, s1);
    else
      tile_pass<R, U, true, false, false, PRIO>(L, src, w, lane, a.c.eps, vt,
                                                B * P, n, s0, s1);
    // after the rotations of steps [s0, s1) lane l holds the sums of source
    // set (l + s1) mod 64 (set l again after a whole pass)
    float *sp = slot + R::ACC * P + ((lane + s1) & 63) * U;
#pragma unroll
    for (int q = 0; q < U; ++q)
#pragma unroll
      for (int e = 0; e < R::ACC; ++e) sp[e * P + q] = L.rs[q][e];
  }
  float *fp = slot + lane * U;
#pragma unroll
  for (int k = 0; k < U; ++k)
#pragma unroll
    for (int e = 0; e < R::ACC; ++e) fp[e * P + k] = L.a[k][e];
}

// Per particle: row partials of its tile (B order) minus the negated
// reaction partials of its column (A order), in kTileRuns runs; then the
// rule's epilogue.
template <class R, int U, int HALVES = 1>
__global__ __launch_bounds__(256) void sym_tile_finalize(SymArgs a) {
  constexpr int P = 64 * U;
  constexpr int PW = 256 / kTileRuns; // particles per workgroup
  __shared__ float red[R::ACC][256];
  const int i = blockIdx.x * PW + threadIdx.x % PW;
  const int run = threadIdx.x / PW;
  const bool live = i < a.n;
  float f[R::ACC];
  if (live) {
    const int g = i / P, o = i % P, T = a.ts;
    // rows of pairs (g, g + k / HALVES), then the reactions of (c / HALVES,
    // g), each pair's halves in order
    run_sum<R, kTileRuns>(
        f, HALVES * (T - g), HALVES * g, run,
        [&](int k) {
          const long long it = tri_slot(g, g + k / HALVES, T) * HALVES + k % HALVES;
          return a.fpart + (size_t)it * (2 * R::ACC * P) + o;
        },
        [&](int c) {
          const long long it = tri_slot(c / HALVES, g, T) * HALVES + c % HALVES;
          return a.fpart + (size_t)it * (2 * R::ACC * P) + R::ACC * P + o;
        },
        P);
  } else {
#pragma unroll
    for (int e = 0; e < R::ACC; ++e) f[e] = 0.0f;
  }
  if (combine_runs<R, kTileRuns>(f, red) && live) R::finish(a, i, f);
}

// ---- many independent systems (batched persistent launch) ----------------
//
// Work item = (system s, super-tile pair ga <= gb of s, slot); the items of
// all systems are handed out longest first through a device counter to a
// grid sized to the GPU's resident capacity.  Slot k owns a row partial
// part[k][0] and, for ga < gb, a reaction partial part[k][1] (ACC x ST
// floats each); the slots of system s are sys_slot[s] + the row-major index
// of (ga, gb) in its upper triangle.  The finalize sums, per particle, the
// row partials of its row in gb order minus the reaction partials of its
// column in ga order: the same values in the same order as the
// single-system kernel with one super-tile per item, so a batched step
// equals the systems' single-system steps bit for bit.

struct SymBatchArgs {
  const float *px, *py, *pz; // concatenated systems
  float *vx, *vy, *vz;
  float *nx, *ny, *nz;
  const int4 *items;     // (system, ga, gb, slot)
  const int *sys_off;    // [S + 1] first particle of each system
  const int *sys_slot;   // [S + 1] first slot of each system
  const float4 *sys_par; // (eps, damping, dt, G)
  float *part;           // [slots][2][3][ST]
  int num_items;
  int num_systems;
  int total; // particles in all systems
  int *counters; // [0]: next item, [1]: workgroups finished
};

// The single-system view of system s (arrays shifted to its first particle).
__device__ __forceinline__ SymArgs system_view(const SymBatchArgs &b, int s,
                                               int ST) {
  const int off0 = b.sys_off[s];
  const float4 pr = b.sys_par[s];
  SymArgs a = {};
  a.px = b.px + off0;
  a.py = b.py + off0;
  a.pz = b.pz + off0;
  a.vx = b.vx + off0;
  a.vy = b.vy + off0;
  a.vz = b.vz + off0;
  a.nx = b.nx + off0;
  a.ny = b.ny + off0;
  a.nz = b.nz + off0;
  a.n = b.sys_off[s + 1] - off0;
  a.ts = (a.n + ST - 1) / ST;
  a.chunk = 1;
  a.c.eps = pr.x;
  a.c.damping = pr.y;
  a.c.dt = pr.z;
  a.c.G = pr.w;
  return a;
}

template <int U>
__global__ __launch_bounds__(256) void sym_batch_kernel(SymBatchArgs b) {
  using R = ForceRule;
  constexpr int P = 64 * U;
  constexpr int ST = kWaves * P;
  __shared__ float4 src4[R::SRC][ST / 4];
  __shared__ float4 sacc4[R::ACC][ST / 4];
  __shared__ int s_item;
  float *src = reinterpret_cast<float *>(src4);
  float *sacc = reinterpret_cast<float *>(sacc4);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(b.counters, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads(); // every lane has read s_item before it is rewritten
    if (item >= b.num_items) break;
    const int4 it = b.items[item];
    const SymArgs a = system_view(b, it.x, ST);
    float *slot = b.part + (size_t)it.w * (2 * R::ACC * ST);
    Lane<R, U> T;
    bool vt[U];
    load_targets(T, vt, a, it.y, w, lane);
    pair_step(T, vt, a, src, sacc, it.y, it.z, slot + R::ACC * ST);
    store_row(T, slot, w, lane);
  }
  if (threadIdx.x == 0) {
    __threadfence(); // this workgroup's last queue access precedes the count
    const int done = atomicAdd(b.counters + 1, 1);
    if (done == (int)gridDim.x - 1) {
      atomicExch(b.counters, 0);
      atomicExch(b.counters + 1, 0);
    }
  }
}

template <int U>
__global__ __launch_bounds__(256) void sym_batch_finalize(SymBatchArgs b) {
  using R = ForceRule;
  constexpr int ST = kWaves * 64 * U;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= b.total) return;
  // system of particle i: the last s with sys_off[s] <= i (empty systems
  // share their offset with the next one and are skipped)
  int lo = 0, hi = b.num_systems - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.sys_off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  const SymArgs a = system_view(b, lo, ST);
  const int li = i - b.sys_off[lo];
  const int g = li / ST, o = li % ST, ts = a.ts;
  const float *base = b.part + (size_t)b.sys_slot[lo] * (2 * R::ACC * ST) + o;
  float f[R::ACC];
  run_sum<R, 1>(
      f, ts - g, g, 0,
      [&](int k) {
        return base + (size_t)tri_slot(g, g + k, ts) * (2 * R::ACC * ST);
      },
      [&](int ga) {
        return base + (size_t)tri_slot(ga, g, ts) * (2 * R::ACC * ST) +
               R::ACC * ST;
      },
      ST);
  R::finish(a, li, f);
}

using KernelFn = void (*)(SymArgs);

struct SymShape {
  int unroll;
  KernelFn pairs, finalize;
};

template <class R>
const SymShape *find_shape(int unroll, bool tiles) {
  static const SymShape shapes[] = {
      {2, sym_kernel<R, 2>, sym_finalize<R, 2>},
      {4, sym_kernel<R, 4>, sym_finalize<R, 4>},
      {8, sym_kernel<R, 8>, sym_finalize<R, 8>},
  };
  static const SymShape tile_shapes[] = {
      {2, sym_tile_kernel<R, 2>, sym_tile_finalize<R, 2>},
      {4, sym_tile_kernel<R, 4>, sym_tile_finalize<R, 4>},
  };
  if (tiles) {
    for (const SymShape &s : tile_shapes)
      if (s.unroll == unroll) return &s;
    return nullptr;
  }
  for (const SymShape &s : shapes)
    if (s.unroll == unroll) return &s;
  return nullptr;
}

// Kahan shapes: 2 or 4 per lane (the compensations double the sums).
const SymShape *find_kahan_shape(int unroll, bool tiles) {
  using R = KahanForceRule;
  static const SymShape shapes[] = {
      {2, sym_kernel<R, 2>, sym_finalize<R, 2>},
      {4, sym_kernel<R, 4>, sym_finalize<R, 4>},
  };
  static const SymShape tile_shapes[] = {
      {2, sym_tile_kernel<R, 2>, sym_tile_finalize<R, 2>},
      {4, sym_tile_kernel<R, 4>, sym_tile_finalize<R, 4>},
  };
  for (int k = 0; k < 2; ++k) {
    const SymShape &s = tiles ? tile_shapes[k] : shapes[k];
    if (s.unroll == unroll) return &s;
  }
  return nullptr;
}

// Banded launches: super-tile shapes with at least four super-tiles per
// item, past their single-launch limit.
bool sym_banded(int unroll, bool tiles, int chunk, int acc, size_t n) {
  return !tiles && (unroll == 2 || unroll == 4 || unroll == 8) && chunk >= 4 &&
         n > sym_max_particles(unroll, tiles, chunk, acc) &&
         n <= kSymBandedMaxParticles;
}

struct BandShape {
  int unroll;
  KernelFn pairs, fold, finish;
};

template <class R, int U>
BandShape band_shape() {
  return {U, sym_band_kernel<R, U>, sym_band_fold<R, U>, sym_band_finish<R, U>};
}

// Band kernels of rule R: 2, 4 or 8 particles per lane (Kahan: 2 or 4).
template <class R>
const BandShape *find_band_shape(int unroll) {
  static const BandShape shapes[] = {band_shape<R, 2>(), band_shape<R, 4>(),
                                     band_shape<R, 8>()};
  for (const BandShape &s : shapes)
    if (s.unroll == unroll) return &s;
  return nullptr;
}

template <>
const BandShape *find_band_shape<KahanForceRule>(int unroll) {
  static const BandShape shapes[] = {band_shape<KahanForceRule, 2>(),
                                     band_shape<KahanForceRule, 4>()};
  for (const BandShape &s : shapes)
    if (s.unroll == unroll) return &s;
  return nullptr;
}

// Band width (the widest whose partials fit kSymBandBytes at 3 sums per
// particle, whatever the rule's count) and the largest row-item and reaction
// slot counts over the bands.
struct BandPlan {
  int ts, w;
  long long rows, reacs;
};

BandPlan band_plan(size_t n, int unroll, int chunk) {
  const int st = kWaves * 64 * unroll;
  const int ts = (int)((n + st - 1) / st);
  const size_t per_slot = sizeof(float) * 3 * st;
  for (int w = ts; w >= 1; --w) {
    long long rows = 0, reacs = 0;
    for (int b0 = 0; b0 < ts; b0 += w) {
      const int ww = std::min(w, ts - b0);
      rows = std::max(rows, band_items(b0, ww, chunk));
      reacs = std::max(reacs, band_reactions(b0, ww));
    }
    if ((size_t)(rows + reacs) * per_slot <= kSymBandBytes)
      return {ts, w, rows, reacs};
  }
  return {ts, 0, 0, 0};
}

int launch_sym_banded(const BandShape *band, SymArgs &a, int chunk, int n,
                      float *scratch, int acc, hipStream_t stream) {
  const int unroll = band->unroll;
  const int st = kWaves * 64 * unroll;
  const BandPlan bp = band_plan((size_t)n, unroll, chunk);
  if (bp.w < 1)
    return set_error(NB_ERR_UNSUPPORTED, "no band plan for this size");
  a.fpart = scratch;
  a.spart = scratch + (size_t)bp.rows * acc * st;
  a.acc3 = a.spart + (size_t)bp.reacs * acc * st;
  a.ts = bp.ts;
  a.chunk = chunk;
  a.n = n;
  void *args[] = {&a};
  const dim3 block(256), part_grid((n + 255) / 256);
  for (int b0 = 0; b0 < bp.ts; b0 += bp.w) {
    a.b0 = b0;
    a.bw = std::min(bp.w, bp.ts - b0);
    a.first = b0 == 0;
    const long long items = band_items(a.b0, a.bw, chunk);
    int r = hip_check(
        hipLaunchKernel(reinterpret_cast<const void *>(band->pairs),
                        dim3((unsigned)items), block, args, 0, stream),
        "symmetric band launch");
    if (!r)
      r = hip_check(hipLaunchKernel(reinterpret_cast<const void *>(band->fold),
                                    part_grid, block, args, 0, stream),
                    "symmetric band fold launch");
    if (r) return r;
  }
  return hip_check(
      hipLaunchKernel(reinterpret_cast<const void *>(band->finish), part_grid,
                      block, args, 0, stream),
      "symmetric band finish launch");
}

int launch_sym(const SymShape *sh, SymArgs &a, int unroll, int chunk, int n,
               float *scratch, int acc, bool tiles, hipStream_t stream,
               const BandShape *band = nullptr) {
  if (band && sym_banded(unroll, tiles, chunk, acc, (size_t)n)) {
    if (!scratch)
      return set_error(NB_ERR_INVALID_ARG, "symmetric launch needs scratch");
    return launch_sym_banded(band, a, chunk, n, scratch, acc, stream);
  }
  const size_t cap = sym_max_particles(unroll, tiles, chunk, acc);
  if ((size_t)n > cap)
    return set_error(
        NB_ERR_UNSUPPORTED,
        tiles ? "symmetric tile-pair shapes take at most 2^17 particles"
        : chunk >= 4 && band
            ? "symmetric launch shapes take at most 2^24 particles"
        : unroll <= 2
            ? "symmetric shapes with 2 per lane take at most 2^18 particles "
              "(2^24 with 4 super-tiles per item)"
            : "symmetric launch shapes take at most 2^20 particles (2^24 with "
              "4 super-tiles per item; Kahan sums: 2 or 4 per lane)");
  if (!sh || chunk < 1)
    return set_error(NB_ERR_UNSUPPORTED, "no symmetric kernel for this shape");
  if (!scratch)
    return set_error(NB_ERR_INVALID_ARG, "symmetric launch needs scratch");
  long long blocks;
  if (tiles) {
    if (chunk > 2)
      return set_error(NB_ERR_UNSUPPORTED,
                       "tile-pair shapes take one or two items per tile pair");
    const int P = 64 * unroll;
    const int T = (n + P - 1) / P;
    a.fpart = scratch;
    a.spart = nullptr;
    a.ts = T;
    a.chunk = chunk; // items per tile pair
    blocks = ((long long)chunk * T * (T + 1) / 2 + kWaves - 1) / kWaves;
  } else {
    const int st = kWaves * 64 * unroll;
    const int ts = (n + st - 1) / st;
    blocks = items_upto(ts, chunk);
    a.fpart = scratch;
    a.spart = scratch + (size_t)blocks * acc * st;
    a.ts = ts;
    a.chunk = chunk;
  }
  a.n = n;
  void *args[] = {&a};
  int r = hip_check(hipLaunchKernel(reinterpret_cast<const void *>(sh->pairs),
                                    dim3((unsigned)blocks), dim3(256), args, 0,
                                    stream),
                    "symmetric pair launch");
  if (r) return r;
  const int per_block = tiles ? 256 / kTileRuns : 256;
  return hip_check(
      hipLaunchKernel(reinterpret_cast<const void *>(sh->finalize),
                      dim3((n + per_block - 1) / per_block), dim3(256), args,
                      0, stream),
      "symmetric finalize launch");
}

} // namespace

size_t sym_scratch_bytes(int unroll, int chunk, size_t n, int acc,
                         bool tiles, bool bandable) {
  if (tiles) {
    if (n == 0 || n > sym_max_particles(unroll, tiles, chunk, acc)) return 0;
    const long long P = 64LL * unroll;
    const long long T = ((long long)n + P - 1) / P;
    return sizeof(float) * (size_t)(T * (T + 1) / 2) * 2 * acc * (size_t)P *
           (size_t)std::max(chunk, 1);
  }
  if (bandable && sym_banded(unroll, tiles, chunk, acc, n)) {
    const BandPlan bp = band_plan(n, unroll, chunk);
    if (bp.w < 1) return 0;
    return sizeof(float) * ((size_t)(bp.rows + bp.reacs) * acc *
                                (size_t)(kWaves * 64 * unroll) +
                            (size_t)acc * n);
  }
  if (n == 0 || n > sym_max_particles(unroll, tiles, chunk, acc) || chunk < 1)
    return 0;
  const long long st = (long long)kWaves * 64 * unroll;
  const long long ts = ((long long)n + st - 1) / st;
  const long long slots = items_upto(ts, chunk) + ts * (ts - 1) / 2;
  return sizeof(float) * (size_t)slots * acc * (size_t)st;
}

int sym_batch_layout(int unroll, const int *sys_off, int num_systems,
                     std::vector<int4> &items, std::vector<int> &sys_slot,
                     size_t &part_bytes) {
  if (unroll != 2 && unroll != 4 && unroll != 8)
    return set_error(NB_ERR_UNSUPPORTED, "no symmetric batch kernel for shape");
  const int ST = kWaves * 64 * unroll;
  std::vector<int4> off_diag, diag;
  sys_slot.assign(num_systems + 1, 0);
  long long slots = 0;
  for (int s = 0; s < num_systems; ++s) {
    sys_slot[s] = (int)slots;
    const int n = sys_off[s + 1] - sys_off[s];
    const int ts = (n + ST - 1) / ST;
    for (int ga = 0; ga < ts; ++ga)
      for (int gb = ga; gb < ts; ++gb) {
        const int4 it = make_int4(s, ga, gb, (int)(slots + tri_slot(ga, gb, ts)));
        (gb == ga ? diag : off_diag).push_back(it);
      }
    slots += (long long)ts * (ts + 1) / 2;
    if (slots > (1LL << 30))
      return set_error(NB_ERR_UNSUPPORTED, "too many symmetric work items");
  }
  sys_slot[num_systems] = (int)slots;
  // longest first: pairs with reactions, then the diagonal super-tiles
  items = off_diag;
  items.insert(items.end(), diag.begin(), diag.end());
  part_bytes = sizeof(float) * (size_t)slots * 2 * 3 * ST;
  return NB_OK;
}

namespace {
template <int U> const void *batch_pairs_fn() {
  return reinterpret_cast<const void *>(sym_batch_kernel<U>);
}
template <int U> const void *batch_finalize_fn() {
  return reinterpret_cast<const void *>(sym_batch_finalize<U>);
}
} // namespace

const void *sym_batch_kernel_fn(int unroll) {
  switch (unroll) {
  case 2: return batch_pairs_fn<2>();
  case 4: return batch_pairs_fn<4>();
  case 8: return batch_pairs_fn<8>();
  default: return nullptr;
  }
}

int launch_sym_batch(int unroll, const SymBatchLaunch &L, hipStream_t stream) {
  const void *pairs = sym_batch_kernel_fn(unroll);
  const void *fin = unroll == 2   ? batch_finalize_fn<2>()
                    : unroll == 4 ? batch_finalize_fn<4>()
                    : unroll == 8 ? batch_finalize_fn<8>()
                                  : nullptr;
  if (!pairs || !fin)
    return set_error(NB_ERR_UNSUPPORTED, "no symmetric batch kernel for shape");
  if (L.num_items == 0) return NB_OK;
  SymBatchArgs b;
  b.px = L.pos[0];
  b.py = L.pos[1];
  b.pz = L.pos[2];
  b.vx = L.vel[0];
  b.vy = L.vel[1];
  b.vz = L.vel[2];
  b.nx = L.nxt[0];
  b.ny = L.nxt[1];
  b.nz = L.nxt[2];
  b.items = L.items;
  b.sys_off = L.sys_off;
  b.sys_slot = L.sys_slot;
  b.sys_par = L.sys_par;
  b.part = L.part;
  b.num_items = L.num_items;
  b.num_systems = L.num_systems;
  b.total = L.total;
  b.counters = L.counters;
  void *args[] = {&b};
  int r = hip_check(hipLaunchKernel(pairs, dim3(L.grid), dim3(256), args, 0,
                                    stream),
                    "symmetric batch launch");
  if (r) return r;
  return hip_check(hipLaunchKernel(fin, dim3((L.total + 255) / 256),
                                   dim3(256), args, 0, stream),
                   "symmetric batch finalize launch");
}

int launch_forward_sym(const StepConsts &c, int unroll, int chunk, bool tiles,
                       bool kahan, const float *px, const float *py,
                       const float *pz, float *vx, float *vy, float *vz,
                       float *nx, float *ny, float *nz, int n, float *scratch,
                       hipStream_t stream) {
  if (n <= 0) return NB_OK;
  SymArgs a = {};
  a.px = px;
  a.py = py;
  a.pz = pz;
  a.vx = vx;
  a.vy = vy;
  a.vz = vz;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.c = c;
  if (tiles && chunk != 1 && (kahan || unroll != 2))
    return set_error(NB_ERR_UNSUPPORTED,
                     "two items per tile pair: plain sums, 2 per lane only");
  // Kahan tile pairs with the priority steps up to n = 16384 (same results
  // bit for bit): 12288 76.2-76.5 vs 79.8-80.0 us, 16384 118.8-118.9 vs
  // 119.4-119.5 us; from 20480 on 2-5 % slower (more work per wave).
  if (kahan && tiles && unroll == 2 && (size_t)n <= kKahanTilePrioMaxParticles) {
    static const SymShape kprio = {2, sym_tile_kernel<KahanForceRule, 2, true>,
                                   sym_tile_finalize<KahanForceRule, 2>};
    return launch_sym(&kprio, a, unroll, chunk, n, scratch, KahanForceRule::ACC,
                      tiles, stream, nullptr);
  }
  if (kahan)
    return launch_sym(
        find_kahan_shape(unroll, tiles), a, unroll, chunk, n, scratch,
        KahanForceRule::ACC, tiles, stream,
        tiles ? nullptr : find_band_shape<KahanForceRule>(unroll));
  // Tile pairs of 2 per lane up to n = 32768 with the quarter-pass priority
  // steps (PRIO in tile_pass; same results bit for bit): n = 4096 20.6 vs
  // 23.8 us, 8192 30.6 vs 31.4 us, 12800 47.4 vs 49.2 us; with the priority
  // set outside the step loop also 14336 56.5-57.5 vs 59.6-60.3 us, 20480
  // 94.5-96.8 vs 97.2-98.5 us, 25600 and 32768 about 1 % faster, 16384 the
  // same, 40960 0.3 % slower.
  // Two items per tile pair (forward config 37, the default at 4096 <= n <
  // 6656) against one item with priority steps: n = 4096 15.5 vs 20.0 us,
  // 6144 20.9 vs 22.2 us; from 7168 on the doubled partials cost more than
  // the balance gains (7168 24.6 vs 23.1 us, 8192 28.6 vs 27.2 us, 12800
  // 47.2 vs 42.5 us).  The half items take the priority steps as well (per
  // quarter of their 32 steps): 4096 15.1-15.4 vs 15.4-15.5 us, 5120
  // 16.5-16.8 vs 17.0-17.1 us, 6144 19.8-19.9 vs 20.9-21.0 us.
  if (tiles && unroll == 2 && chunk == 2) {
    static const SymShape halves = {2, sym_tile_kernel<ForceRule, 2, true, 2>,
                                    sym_tile_finalize<ForceRule, 2, 2>};
    return launch_sym(&halves, a, unroll, chunk, n, scratch, ForceRule::ACC,
                      tiles, stream, nullptr);
  }
  if (tiles && unroll == 2 && (size_t)n <= kTilePrioMaxParticles) {
    static const SymShape prio = {2, sym_tile_kernel<ForceRule, 2, true>,
                                  sym_tile_finalize<ForceRule, 2>};
    return launch_sym(&prio, a, unroll, chunk, n, scratch, ForceRule::ACC,
                      tiles, stream, nullptr);
  }
  return launch_sym(find_shape<ForceRule>(unroll, tiles), a, unroll, chunk, n,
                    scratch, ForceRule::ACC, tiles, stream,
                    tiles ? nullptr : find_band_shape<ForceRule>(unroll));
}

// The adjoint's tile pairs of 2 per lane take the quarter-pass priority
// steps too (same results bit for bit): adjoint / fused n = 8192 39.9-40.3 vs
// 43.3-44.1 us / 47.8-48.2 vs 51.1-51.4 us, 12800 59.8-60.3 vs 63.6-64.1 /
// 75.8-76.2 vs 80.3-80.4, 16384 95.7-96.1 vs 99.3-99.9 / 120.4-120.6 vs
// 125.3-125.5, 20000 135.4-135.9 vs 137.3-139.2 / 173.3-173.5 vs 177.8-177.9;
// 2048 and 4096 the same.
template <bool F>
const SymShape *adjoint_shape(int unroll, bool tiles, int n) {
  if (tiles && unroll == 2 && (size_t)n <= kTilePrioMaxParticles) {
    static const SymShape prio = {2, sym_tile_kernel<AdjointRule<F>, 2, true>,
                                  sym_tile_finalize<AdjointRule<F>, 2>};
    return &prio;
  }
  return find_shape<AdjointRule<F>>(unroll, tiles);
}

int launch_adjoint_sym(const StepConsts &c, bool fused, int unroll, int chunk,
                       bool tiles, const float *const pos[3], const float *const lxn[3],
                       const float *const lvn[3], float *const lx[3],
                       float *const lv[3], float *const vel[3],
                       float *const nxt[3], int n, float *scratch,
                       hipStream_t stream) {
  if (n <= 0) return NB_OK;
  SymArgs a = {};
  a.c = c;
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
    return launch_sym(
        adjoint_shape<true>(unroll, tiles, n), a, unroll, chunk, n,
        scratch, AdjointRule<true>::ACC, tiles, stream,
        tiles ? nullptr : find_band_shape<AdjointRule<true>>(unroll));
  }
  return launch_sym(adjoint_shape<false>(unroll, tiles, n), a, unroll,
                    chunk, n, scratch, AdjointRule<false>::ACC, tiles, stream,
                    tiles ? nullptr : find_band_shape<AdjointRule<false>>(unroll));
}

} // namespace nb
