// Internal declarations shared by the HIP translation units of libnbody_amd.
// Not part of the public ABI (include/nbody_amd.h is).
#pragma once

#include <hip/hip_runtime.h>

#include "nbody_amd.h"

#include <algorithm>
#include <cfloat>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

namespace nb {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string &msg);
// Checks a HIP status; on failure records "<what>: <hip message>".
int hip_check(hipError_t e, const char *what);

// Physical step constants, in the order the kernels consume them.
struct StepConsts {
  float eps;     // SimParam::distEps
  float damping; // SimParam::damping
  float dt;      // SimParam::dt
  float G;       // SimParam::G
};

// Pair-loop variants.  The self term of the reference's BRANCH kernel is
// skipped with `continue` (src/simulator.cu:206); when dist_eps > 0 and
// dist_eps^3 is a normal float the self term is exactly +0 (r = 0, finite
// 1/d^1.5), so the fast variants omit the test.  The *_SKIP variants mask
// j == i for the remaining parameter values.
enum PairMode : int {
  MODE_PLAIN = 0,
  MODE_PLAIN_SKIP = 1,
  MODE_KAHAN = 2,
  MODE_KAHAN_SKIP = 3,
  MODE_PREDICATED = 4,
};

// Forward decomposition schemes.
enum FwdScheme : int {
  SCHEME_TARGETS = 0,   // each thread sums all sources for its targets
  SCHEME_SYMMETRIC = 1, // each pair term evaluated once, applied to both;
                        // work item = super-tile pair (4 waves)
  SCHEME_SYMMETRIC_TILES = 2, // the same, work item = tile pair (one wave)
};

// One compiled launch shape of the forward kernel.  For SCHEME_SYMMETRIC
// (nb_sym.hip): block = 256 (4 waves), unroll = particles per lane,
// tile = super-tile length (4 * 64 * unroll), split = super-tiles per work
// item (row chunk).  For SCHEME_SYMMETRIC_TILES: tile = tile length
// (64 * unroll), split = 1.
struct FwdConfig {
  int block;  // threads per workgroup
  int unroll; // target particles per thread
  int tile;   // source particles staged in LDS per pass
  int split;  // source-range chunks (grid.y); >1 writes partial forces
  int scheme = SCHEME_TARGETS;
};

int forward_config_count();
const FwdConfig &forward_config(int cfg);
// Library's launch shape for n particles in pair mode `mode` (MODE_*).
int forward_default_config(size_t n, int mode);

// Largest particle count of the symmetric scheme (its partial buffers grow
// as n^2: 4.0 GB at 2^20 with 8 particles per lane and 4-tile chunks).
constexpr size_t kSymMaxParticles = (size_t)1 << 20;
// Largest particle count of the tile-pair form (0.8 GB of partials at 2^17).
constexpr size_t kSymTileMaxParticles = (size_t)1 << 17;
// Largest particle count of the super-tile shapes with 2 particles per lane
// (1.6 GB of partials at 2^18).
constexpr size_t kSymSmallMaxParticles = (size_t)1 << 18;

// Largest particle count of the plain-sum shapes with 8 particles per lane
// and at least four super-tiles per work item (16 GB of partials at 2^21).
constexpr size_t kSymLargeMaxParticles = (size_t)1 << 21;

// Past their single-launch limit (2^18 for 2 per lane, 2^21 for 8 per lane
// with plain sums, 2^20 otherwise) the super-tile shapes with at least four
// super-tiles per item run the forward step (plain or Kahan sums) and both
// adjoint steps in bands of source super-tiles, up to
// kSymBandedMaxParticles.  A band's partials take
// at most kSymBandBytes at 3 sums per particle (twice that for the fused
// adjoint's 6): the band width depends on the slot count only, so the fused
// adjoint cuts the bands the forward step cuts and its forward part keeps
// the forward step's bits.
constexpr size_t kSymBandedMaxParticles = (size_t)1 << 24;
constexpr size_t kSymBandBytes = (size_t)8 << 30;

// Largest particle count of a single symmetric launch (acc: sums per
// particle, 3 plain, 6 for Kahan and the fused adjoint).
constexpr size_t sym_max_particles(int unroll, bool tiles, int chunk = 1,
                                   int acc = 3) {
  return tiles         ? kSymTileMaxParticles
         : unroll <= 2 ? kSymSmallMaxParticles
         : (unroll == 8 && chunk >= 4 && acc == 3) ? kSymLargeMaxParticles
                                                   : kSymMaxParticles;
}

// Symmetric scheme (nb_sym.hip; BRANCH, any eps): scratch bytes
// for `acc` sums per particle (3 forward and adjoint, 6 fused adjoint or
// Kahan forward) and stream-ordered launches of one forward step / one
// adjoint step.  `bandable`: the rule has banded kernels for the shape
// (every rule with 2, 4 or 8 per lane, Kahan with 2 or 4), which take the
// super-tile shapes with at least 4 super-tiles per item past their
// single-launch limit.
size_t sym_scratch_bytes(int unroll, int chunk, size_t n, int acc,
                         bool tiles, bool bandable);
int launch_adjoint_sym(const StepConsts &c, bool fused, int unroll, int chunk,
                       bool tiles, const float *const pos[3], const float *const lxn[3],
                       const float *const lvn[3], float *const lx[3],
                       float *const lv[3], float *const vel[3],
                       float *const nxt[3], int n, float *scratch,
                       hipStream_t stream);
// Batched symmetric step (nb_sym.hip), plain mode only.  Layout of the work
// items (system, ga, gb, slot), longest first, the first slot of each system
// ([S + 1]) and the bytes of the partial buffer, for systems given by their
// offsets (sys_off[S + 1]).
int sym_batch_layout(int unroll, const int *sys_off, int num_systems,
                     std::vector<int4> &items, std::vector<int> &sys_slot,
                     size_t &part_bytes);
struct SymBatchLaunch {
  const float *pos[3];
  float *vel[3];
  float *nxt[3];
  const int4 *items;
  const int *sys_off, *sys_slot;
  const float4 *sys_par;
  float *part;
  int num_items, num_systems, total;
  int *counters; // two self-resetting queue counters (zero before first use)
  int grid;      // persistent workgroups
};
// Pair kernel of the batched symmetric step (for occupancy queries).
const void *sym_batch_kernel_fn(int unroll);
int launch_sym_batch(int unroll, const SymBatchLaunch &L, hipStream_t stream);

int launch_forward_sym(const StepConsts &c, int unroll, int chunk, bool tiles,
                       bool kahan, const float *px, const float *py, const float *pz,
                       float *vx, float *vy, float *vz, float *nx, float *ny,
                       float *nz, int n, float *scratch, hipStream_t stream);

// Stream-ordered launch of one step on device SoA buffers.  `scratch` must
// hold 3 * split * n floats when the shape has split > 1 (else unused).
int launch_forward(const StepConsts &c, int mode, int cfg, const float *px,
                   const float *py, const float *pz, float *vx, float *vy,
                   float *vz, float *nx, float *ny, float *nz, int n,
                   float *scratch, hipStream_t stream);

// Bytes of scratch launch_forward needs for (cfg, n, pair mode).
size_t forward_scratch_bytes(int cfg, size_t n, int mode);

// Scratch owned by the library: one buffer per (device, stream), grown on
// demand and freed by nb_scratch_release.  A lease holds the buffer's lock:
// keep it alive until the launches that use the buffer are enqueued on the
// stream it was acquired for.
class ScratchLease {
 public:
  ScratchLease() = default;
  ScratchLease(ScratchLease &&) = default;
  ScratchLease &operator=(ScratchLease &&) = default;
  float *get() const { return ptr_; }

 private:
  friend int acquire_scratch(size_t, hipStream_t, ScratchLease &);
  std::unique_lock<std::mutex> lock_;
  float *ptr_ = nullptr;
};
// `bytes` (0: none) of scratch for launches on `stream`.  Refuses to grow a
// buffer while the stream is being captured into a graph.
int acquire_scratch(size_t bytes, hipStream_t stream, ScratchLease &lease);

// SoA -> AoS vec4 pack on the device (out: 4 n floats, 16 B aligned).
// Copies six device arrays of n floats to dst + c * stride (c = 0..5), a
// device pointer to host-mapped pinned memory, in one launch on `stream`.
int launch_stage_six(const float *const src[6], float *dst, int n,
                     size_t stride, hipStream_t stream);
int launch_pack_vec4(const float *x, const float *y, const float *z, float w,
                     float *out, int n, hipStream_t stream);

// Pair mode for a parameter set.
int pair_mode(int calc_method, int acc, float eps);

// True when the self term of the pair sum is exactly zero (no j == i test
// needed): the self term r * d^-3/2 (r = 0, d = eps) vanishes when eps^3 is
// a positive normal (or infinite) float; false for NaN, 0, < 0, subnormal.
__host__ __device__ inline bool self_term_vanishes(float eps) {
  const float e3 = (eps * eps) * eps;
  return e3 >= FLT_MIN;
}

// Adjoint launch shapes (same fields as the forward ones).  skip: the self
// term needs masking (dist_eps^3 not a normal float).
int adjoint_config_count();
const FwdConfig &adjoint_config(int cfg);
int adjoint_default_config(size_t n, bool skip);
size_t adjoint_scratch_bytes(int cfg, size_t n);

// Adjoint of one step at positions pos: reads the cotangents of the step's
// outputs (lxn, lvn), writes those of its inputs (lx, lv).  With fused set,
// the same launch also applies the forward step (vel in/out, nxt written).
int launch_adjoint(const StepConsts &c, bool skip, bool fused, int cfg,
                   const float *const pos[3], const float *const lxn[3],
                   const float *const lvn[3], float *const lx[3],
                   float *const lv[3], float *const vel[3],
                   float *const nxt[3], int n, float *scratch,
                   hipStream_t stream);

} // namespace nb
