// Batched persistent kernel: many independent systems of uneven size
// advanced by one launch per step (no reference counterpart; the reference
// simulates one system, src/simulator.cu:47-75).
//
// Layout: the systems' SoA arrays are concatenated (system s owns
// [off[s], off[s+1])); per-system constants (eps, damping, dt, G) sit in a
// small device table.  Work item = (system, BLOCK*U consecutive targets)
// against all of that system's sources; items are ordered by decreasing
// cost (longest first) and handed out through a device counter, so a grid
// sized to the resident capacity of the GPU drains one queue for all
// systems.  Each item is processed by one workgroup from start to finish,
// so every target's force is summed in source order whatever the schedule
// (results are deterministic).
//
// The counters reset themselves: the last workgroup to leave zeroes them,
// so consecutive launches (and replays of a captured hipGraph) need no
// memset between them.  Two steps are captured per graph (A -> B, B -> A).
//
// The pair loop and integrator are those of the single-system kernel
// (nb_pairs.hpp).  The symmetric shapes run the batched form of nb_sym.hip
// instead: work item = one super-tile pair of one system, each pair term
// evaluated once for both particles, then a finalize launch (results equal
// the single-system symmetric step bit for bit).  A project-written
// autotuner times every compiled shape on the batch itself and keeps the
// fastest.

#include <algorithm>
#include <chrono>
#include <new>
#include <numeric>
#include <vector>

#include "nb_pairs.hpp"

namespace nb {
namespace {

struct BatchArgs {
  const float *px, *py, *pz;
  float *vx, *vy, *vz;
  float *nx, *ny, *nz;
  const int2 *items;     // (system, first target)
  const int *sys_off;    // [S + 1]
  const float4 *sys_par; // (eps, damping, dt, G) per system
  int num_items;
  int *counters; // [0]: next item, [1]: workgroups finished
};

template <int BLOCK, int U, int TILE, int MODE>
__global__ __launch_bounds__(BLOCK) void batch_kernel(BatchArgs a) {
  __shared__ SourceTile<TILE> sm;
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(a.counters, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads(); // every lane has read s_item before it is rewritten
    if (item >= a.num_items) break;
    const int2 it = a.items[item];
    const int off0 = a.sys_off[it.x];
    const int off1 = a.sys_off[it.x + 1];
    const float4 pr = a.sys_par[it.x];
    StepConsts c;
    c.eps = pr.x;
    c.damping = pr.y;
    c.dt = pr.z;
    c.G = pr.w;
    Targets<U, MODE> T;
    T.load(a.px, a.py, a.pz, it.y, BLOCK, threadIdx.x, off1);
    accumulate_sources<BLOCK, TILE>(sm, T, a.px, a.py, a.pz, off0, off1,
                                    c.eps);
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int i = T.idx[k];
      if (i >= off1) continue;
      float fx, fy, fz;
      T.force(k, fx, fy, fz);
      integrate_ref(c, i, fx, fy, fz, a.px, a.py, a.pz, a.vx, a.vy, a.vz,
                    a.nx, a.ny, a.nz);
    }
  }
  if (threadIdx.x == 0) {
    __threadfence(); // this workgroup's last queue access precedes the count
    const int done = atomicAdd(a.counters + 1, 1);
    if (done == (int)gridDim.x - 1) {
      atomicExch(a.counters, 0);
      atomicExch(a.counters + 1, 0);
    }
  }
}

// Zeroes the two queue counters (used once, before the first launch).
__global__ void zero_counters(int *c) {
  if (threadIdx.x < 2) c[threadIdx.x] = 0;
}

using BatchFn = void (*)(BatchArgs);

template <int B, int U, int T>
BatchFn batch_shape(int mode) {
  switch (mode) {
  case MODE_PLAIN:
    return batch_kernel<B, U, T, MODE_PLAIN>;
  case MODE_PLAIN_SKIP:
    return batch_kernel<B, U, T, MODE_PLAIN_SKIP>;
  case MODE_KAHAN:
    return batch_kernel<B, U, T, MODE_KAHAN>;
  case MODE_KAHAN_SKIP:
    return batch_kernel<B, U, T, MODE_KAHAN_SKIP>;
  default:
    return nullptr;
  }
}

struct BatchShape {
  int block, unroll, tile;
  BatchFn (*get)(int);            // target-sharded kernels (nullptr: symmetric)
  int scheme = SCHEME_TARGETS;
};

const BatchShape kBatchShapes[] = {
    {256, 1, 1024, batch_shape<256, 1, 1024>},
    {256, 2, 1024, batch_shape<256, 2, 1024>},
    {256, 4, 1024, batch_shape<256, 4, 1024>},
    {128, 1, 512, batch_shape<128, 1, 512>},
    {128, 2, 512, batch_shape<128, 2, 512>},
    {512, 1, 2048, batch_shape<512, 1, 2048>},
    {64, 1, 256, batch_shape<64, 1, 256>},
    {64, 2, 256, batch_shape<64, 2, 256>},
    // symmetric (nb_sym.hip): 4 waves, particles per lane, super-tile
    {256, 4, 1024, nullptr, SCHEME_SYMMETRIC},
    {256, 8, 2048, nullptr, SCHEME_SYMMETRIC},
    {256, 2, 512, nullptr, SCHEME_SYMMETRIC},
};
constexpr int kNumBatchShapes = sizeof(kBatchShapes) / sizeof(kBatchShapes[0]);

} // namespace
} // namespace nb

using namespace nb;

struct nb_batch {
  int S = 0;
  size_t total = 0, stride = 0;
  std::vector<int> off;
  std::vector<nb_params> params;
  int mode = MODE_PLAIN;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int device = 0, num_cu = 0;
  float *mem = nullptr; // P0 (3), P1 (3), V (3), backup (6)
  float *P[2][3] = {}, *V[3] = {}, *backup = nullptr;
  int cur = 0;
  int *d_off = nullptr;
  float4 *d_par = nullptr;
  int2 *d_items = nullptr;
  int4 *d_sym_items = nullptr; // symmetric shapes: (system, ga, gb, slot)
  int *d_sym_slot = nullptr;   // first partial slot of each system
  float *sym_part = nullptr;   // symmetric partial sums
  size_t sym_part_bytes = 0;
  int *d_counters = nullptr;
  float *scratch = nullptr; // per-system split launches
  int cfg = -1, num_items = 0, grid = 0;
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  float last_ms = 0.0f;
};

namespace {

void drop_graphs(nb_batch *b) {
  for (auto &e : b->exec)
    if (e) {
      (void)hipGraphExecDestroy(e);
      e = nullptr;
    }
}

void destroy_batch(nb_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  for (void *p : {(void *)b->mem, (void *)b->d_off, (void *)b->d_par,
                  (void *)b->d_items, (void *)b->d_counters,
                  (void *)b->scratch, (void *)b->d_sym_items,
                  (void *)b->d_sym_slot, (void *)b->sym_part})
    if (p) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

template <class T>
int upload(T **dst, const std::vector<T> &host, const char *what) {
  if (*dst) {
    (void)hipFree(*dst);
    *dst = nullptr;
  }
  if (host.empty()) return NB_OK;
  int st = hip_check(hipMalloc(dst, sizeof(T) * host.size()), what);
  if (st) return st;
  return hip_check(hipMemcpy(*dst, host.data(), sizeof(T) * host.size(),
                             hipMemcpyHostToDevice),
                   what);
}

// Symmetric shape: work items of every system's super-tile pairs, slot
// table, partial buffer and the persistent grid.
int configure_sym(nb_batch *b, int cfg) {
  const BatchShape &sh = kBatchShapes[cfg];
  if (b->mode != MODE_PLAIN && b->mode != MODE_PLAIN_SKIP)
    return set_error(NB_ERR_UNSUPPORTED,
                     "symmetric batch shapes take BRANCH with plain sums");
  std::vector<int4> items;
  std::vector<int> slots;
  size_t bytes = 0;
  int st = sym_batch_layout(sh.unroll, b->off.data(), b->S, items, slots, bytes);
  if (st) return st;
  if ((st = hip_check(hipStreamSynchronize(b->stream), "stream synchronize")) ||
      (st = upload(&b->d_sym_items, items, "symmetric item upload")) ||
      (st = upload(&b->d_sym_slot, slots, "symmetric slot upload")))
    return st;
  if (bytes > b->sym_part_bytes) {
    if (b->sym_part) (void)hipFree(b->sym_part);
    b->sym_part = nullptr;
    b->sym_part_bytes = 0;
    if ((st = hip_check(hipMalloc(&b->sym_part, bytes), "symmetric partials")))
      return st;
    b->sym_part_bytes = bytes;
  }
  b->num_items = (int)items.size();
  int per_cu = 0;
  if ((st = hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &per_cu, sym_batch_kernel_fn(sh.unroll), 256, 0),
                      "occupancy query")))
    return st;
  b->grid = std::max(1, std::min(b->num_items, std::max(1, per_cu) * b->num_cu));
  b->cfg = cfg;
  drop_graphs(b);
  return NB_OK;
}

// Builds the longest-first item list and the grid for shape `cfg`.
int configure(nb_batch *b, int cfg) {
  const BatchShape &sh = kBatchShapes[cfg];
  if (sh.scheme == SCHEME_SYMMETRIC) return configure_sym(b, cfg);
  const int per = sh.block * sh.unroll;
  struct Item {
    int sys, first;
    long long cost;
  };
  std::vector<Item> items;
  for (int s = 0; s < b->S; ++s) {
    const int n = b->off[s + 1] - b->off[s];
    for (int t = b->off[s]; t < b->off[s + 1]; t += per)
      items.push_back({s, t, (long long)n * std::min(per, b->off[s + 1] - t)});
  }
  std::stable_sort(items.begin(), items.end(),
                   [](const Item &x, const Item &y) { return x.cost > y.cost; });
  std::vector<int2> host(items.size());
  for (size_t k = 0; k < items.size(); ++k)
    host[k] = make_int2(items[k].sys, items[k].first);
  int st = hip_check(hipStreamSynchronize(b->stream), "stream synchronize");
  if (st) return st;
  if ((st = upload(&b->d_items, host, "item upload"))) return st;
  b->num_items = (int)host.size();
  BatchFn fn = sh.get(b->mode);
  int per_cu = 0;
  if ((st = hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &per_cu, reinterpret_cast<const void *>(fn),
                          sh.block, 0),
                      "occupancy query")))
    return st;
  b->grid = std::max(1, std::min(b->num_items, std::max(1, per_cu) * b->num_cu));
  b->cfg = cfg;
  drop_graphs(b);
  return NB_OK;
}

int launch_batch(nb_batch *b, int from) {
  if (b->num_items == 0) return NB_OK;
  const BatchShape &sh = kBatchShapes[b->cfg];
  if (sh.scheme == SCHEME_SYMMETRIC) {
    SymBatchLaunch L;
    for (int c = 0; c < 3; ++c) {
      L.pos[c] = b->P[from][c];
      L.vel[c] = b->V[c];
      L.nxt[c] = b->P[from ^ 1][c];
    }
    L.items = b->d_sym_items;
    L.sys_off = b->d_off;
    L.sys_slot = b->d_sym_slot;
    L.sys_par = b->d_par;
    L.part = b->sym_part;
    L.num_items = b->num_items;
    L.num_systems = b->S;
    L.total = (int)b->total;
    L.counters = b->d_counters;
    L.grid = b->grid;
    return launch_sym_batch(sh.unroll, L, b->stream);
  }
  BatchArgs a;
  a.px = b->P[from][0];
  a.py = b->P[from][1];
  a.pz = b->P[from][2];
  a.vx = b->V[0];
  a.vy = b->V[1];
  a.vz = b->V[2];
  a.nx = b->P[from ^ 1][0];
  a.ny = b->P[from ^ 1][1];
  a.nz = b->P[from ^ 1][2];
  a.items = b->d_items;
  a.sys_off = b->d_off;
  a.sys_par = b->d_par;
  a.num_items = b->num_items;
  a.counters = b->d_counters;
  void *args[] = {&a};
  return hip_check(hipLaunchKernel(reinterpret_cast<const void *>(
                                       sh.get(b->mode)),
                                   dim3(b->grid), dim3(sh.block), args, 0,
                                   b->stream),
                   "batch launch");
}

// Two steps starting from buffer `from` captured as one graph.
int graph_for(nb_batch *b, int from, hipGraphExec_t *out) {
  if (b->exec[from]) {
    *out = b->exec[from];
    return NB_OK;
  }
  int st = hip_check(hipStreamBeginCapture(b->stream,
                                           hipStreamCaptureModeThreadLocal),
                     "hipStreamBeginCapture");
  if (st) return st;
  const int st1 = launch_batch(b, from);
  const int st2 = st1 ? st1 : launch_batch(b, from ^ 1);
  hipGraph_t g = nullptr;
  st = hip_check(hipStreamEndCapture(b->stream, &g), "hipStreamEndCapture");
  if (st2) {
    if (g) (void)hipGraphDestroy(g);
    return st2;
  }
  if (st) return st;
  st = hip_check(hipGraphInstantiate(&b->exec[from], g, nullptr, nullptr, 0),
                 "hipGraphInstantiate");
  (void)hipGraphDestroy(g);
  if (st) return st;
  *out = b->exec[from];
  return NB_OK;
}

int run_steps(nb_batch *b, int steps, bool graph) {
  if (b->num_items == 0) return NB_OK; // no particles: nothing to advance
  if (graph) {
    // pairs of steps replay the captured two-step graph (state returns to
    // the same buffer); an odd last step is launched directly
    for (int k = 0; k + 1 < steps; k += 2) {
      hipGraphExec_t e = nullptr;
      int st = graph_for(b, b->cur, &e);
      if (!st) st = hip_check(hipGraphLaunch(e, b->stream), "hipGraphLaunch");
      if (st) return st;
    }
  }
  const int direct = graph ? (steps & 1) : steps;
  for (int k = 0; k < direct; ++k) {
    int st = launch_batch(b, b->cur);
    if (st) return st;
    b->cur ^= 1;
  }
  return NB_OK;
}

int timed(nb_batch *b, int steps, bool graph, bool separate) {
  int st = hip_check(hipEventRecord(b->ev0, b->stream), "hipEventRecord");
  if (st) return st;
  if (separate) {
    for (int t = 0; t < steps && !st; ++t) {
      for (int s = 0; s < b->S && !st; ++s) {
        const int o = b->off[s];
        const int n = b->off[s + 1] - o;
        const nb_params &p = b->params[s];
        StepConsts c;
        c.eps = p.dist_eps;
        c.damping = p.damping;
        c.dt = p.dt;
        c.G = p.G;
        float *const *src = b->P[b->cur];
        float *const *dst = b->P[b->cur ^ 1];
        st = launch_forward(c, b->mode, forward_default_config(n, b->mode),
                            src[0] + o, src[1] + o, src[2] + o, b->V[0] + o,
                            b->V[1] + o, b->V[2] + o, dst[0] + o, dst[1] + o,
                            dst[2] + o, n, b->scratch, b->stream);
      }
      b->cur ^= 1;
    }
  } else {
    st = run_steps(b, steps, graph);
  }
  if (st) return st;
  if ((st = hip_check(hipEventRecord(b->ev1, b->stream), "hipEventRecord")) ||
      (st = hip_check(hipEventSynchronize(b->ev1), "hipEventSynchronize")))
    return st;
  return hip_check(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1),
                   "hipEventElapsedTime");
}

int copy_state(nb_batch *b, bool save) {
  const size_t bytes = sizeof(float) * b->total;
  float *state[6] = {b->P[b->cur][0], b->P[b->cur][1], b->P[b->cur][2],
                     b->V[0], b->V[1], b->V[2]};
  for (int c = 0; c < 6; ++c) {
    float *keep = b->backup + c * b->stride;
    int st = hip_check(hipMemcpyAsync(save ? keep : state[c],
                                      save ? state[c] : keep, bytes,
                                      hipMemcpyDeviceToDevice, b->stream),
                       "state copy");
    if (st) return st;
  }
  return NB_OK;
}

} // namespace

extern "C" {

int nb_batch_config_count(void) { return kNumBatchShapes; }

int nb_batch_config_info(int cfg, int *block, int *unroll, int *tile) {
  if (cfg < 0 || cfg >= kNumBatchShapes)
    return set_error(NB_ERR_INVALID_ARG, "batch config out of range");
  if (block) *block = kBatchShapes[cfg].block;
  if (unroll) *unroll = kBatchShapes[cfg].unroll;
  if (tile) *tile = kBatchShapes[cfg].tile;
  return NB_OK;
}

int nb_batch_config_scheme(int cfg) {
  if (cfg < 0 || cfg >= kNumBatchShapes) {
    set_error(NB_ERR_INVALID_ARG, "batch config out of range");
    return -1;
  }
  return kBatchShapes[cfg].scheme;
}

int nb_batch_create(int num_systems, const nb_params *params, const float *px,
                    const float *py, const float *pz, const float *vx,
                    const float *vy, const float *vz, nb_batch **out) {
  if (!out) return set_error(NB_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (num_systems < 0 || (num_systems > 0 && !params))
    return set_error(NB_ERR_INVALID_ARG, "bad system list");
  size_t total = 0;
  bool skip = false;
  for (int s = 0; s < num_systems; ++s) {
    if (params[s].calc_method != NB_CALC_BRANCH)
      return set_error(NB_ERR_UNSUPPORTED,
                       "batched systems use CalculationMethod BRANCH");
    total += params[s].num_particles;
    skip = skip || !self_term_vanishes(params[s].dist_eps);
  }
  if (total > ((size_t)1 << 28))
    return set_error(NB_ERR_INVALID_ARG, "more than 2^28 particles in total");
  if (total > 0 && (!px || !py || !pz || !vx || !vy || !vz))
    return set_error(NB_ERR_INVALID_ARG, "particle array is NULL");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return set_error(NB_ERR_NO_DEVICE, "no HIP device visible");
  }
  nb_batch *b = new (std::nothrow) nb_batch;
  if (!b) return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed");
  b->S = num_systems;
  b->total = total;
  b->stride = (total + 63) / 64 * 64;
  b->params.assign(params, params + num_systems);
  b->off.resize(num_systems + 1, 0);
  for (int s = 0; s < num_systems; ++s)
    b->off[s + 1] = b->off[s] + (int)params[s].num_particles;
  b->mode = skip ? MODE_PLAIN_SKIP : MODE_PLAIN;
  int st = 0;
  hipDeviceProp_t prop;
  if ((st = hip_check(hipGetDevice(&b->device), "hipGetDevice")) ||
      (st = hip_check(hipGetDeviceProperties(&prop, b->device),
                      "hipGetDeviceProperties")) ||
      (st = hip_check(hipStreamCreateWithFlags(&b->stream,
                                               hipStreamNonBlocking),
                      "hipStreamCreate")) ||
      (st = hip_check(hipEventCreate(&b->ev0), "hipEventCreate")) ||
      (st = hip_check(hipEventCreate(&b->ev1), "hipEventCreate"))) {
    destroy_batch(b);
    return st;
  }
  b->num_cu = prop.multiProcessorCount;
  std::vector<float4> par(std::max(num_systems, 1));
  for (int s = 0; s < num_systems; ++s)
    par[s] = make_float4(params[s].dist_eps, params[s].damping, params[s].dt,
                         params[s].G);
  size_t scratch_bytes = 0;
  for (int s = 0; s < num_systems; ++s) {
    const size_t n = params[s].num_particles;
    scratch_bytes = std::max(
        scratch_bytes,
        forward_scratch_bytes(forward_default_config(n, b->mode), n, b->mode));
  }
  const size_t floats = 15 * std::max<size_t>(b->stride, 64);
  if ((st = hip_check(hipMalloc(&b->mem, sizeof(float) * floats),
                      "batch hipMalloc")) ||
      (st = hip_check(hipMalloc(&b->d_off, sizeof(int) * (num_systems + 1)),
                      "batch hipMalloc")) ||
      (st = hip_check(hipMalloc(&b->d_par, sizeof(float4) * par.size()),
                      "batch hipMalloc")) ||
      (st = hip_check(hipMalloc(&b->d_counters, sizeof(int) * 2),
                      "batch hipMalloc")) ||
      (scratch_bytes &&
       (st = hip_check(hipMalloc(&b->scratch, scratch_bytes),
                       "batch hipMalloc")))) {
    destroy_batch(b);
    return st;
  }
  for (int c = 0; c < 3; ++c) {
    b->P[0][c] = b->mem + (0 + c) * b->stride;
    b->P[1][c] = b->mem + (3 + c) * b->stride;
    b->V[c] = b->mem + (6 + c) * b->stride;
  }
  b->backup = b->mem + 9 * b->stride;
  const float *src[6] = {px, py, pz, vx, vy, vz};
  float *dst[6] = {b->P[0][0], b->P[0][1], b->P[0][2],
                   b->V[0],    b->V[1],    b->V[2]};
  for (int c = 0; c < 6 && !st && total; ++c)
    st = hip_check(hipMemcpy(dst[c], src[c], sizeof(float) * total,
                             hipMemcpyHostToDevice),
                   "host-to-device copy");
  if (!st)
    st = hip_check(hipMemcpy(b->d_off, b->off.data(),
                             sizeof(int) * (num_systems + 1),
                             hipMemcpyHostToDevice),
                   "offset copy");
  if (!st)
    st = hip_check(hipMemcpy(b->d_par, par.data(),
                             sizeof(float4) * par.size(),
                             hipMemcpyHostToDevice),
                   "parameter copy");
  if (!st) {
    hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, b->stream,
                       b->d_counters);
    st = hip_check(hipGetLastError(), "counter reset launch");
  }
  if (!st) st = configure(b, 0);
  if (st) {
    destroy_batch(b);
    return st;
  }
  *out = b;
  return NB_OK;
}

int nb_batch_set_config(nb_batch *b, int cfg) {
  if (!b) return set_error(NB_ERR_INVALID_ARG, "batch is NULL");
  if (cfg < 0 || cfg >= kNumBatchShapes)
    return set_error(NB_ERR_INVALID_ARG, "batch config out of range");
  return configure(b, cfg);
}

int nb_batch_config(const nb_batch *b) { return b ? b->cfg : -1; }

int nb_batch_step(nb_batch *b, int steps) {
  if (!b) return set_error(NB_ERR_INVALID_ARG, "batch is NULL");
  if (steps < 0) return set_error(NB_ERR_INVALID_ARG, "steps < 0");
  return timed(b, steps, true, false);
}

int nb_batch_step_direct(nb_batch *b, int steps) {
  if (!b) return set_error(NB_ERR_INVALID_ARG, "batch is NULL");
  if (steps < 0) return set_error(NB_ERR_INVALID_ARG, "steps < 0");
  return timed(b, steps, false, false);
}

int nb_batch_step_separate(nb_batch *b, int steps) {
  if (!b) return set_error(NB_ERR_INVALID_ARG, "batch is NULL");
  if (steps < 0) return set_error(NB_ERR_INVALID_ARG, "steps < 0");
  return timed(b, steps, false, true);
}

int nb_batch_autotune(nb_batch *b, int trial_steps, int *best_cfg) {
  if (!b) return set_error(NB_ERR_INVALID_ARG, "batch is NULL");
  if (trial_steps < 1) return set_error(NB_ERR_INVALID_ARG, "trial_steps < 1");
  int st = copy_state(b, true);
  if (st) return st;
  const int cur0 = b->cur;
  int best = -1;
  float best_ms = 0.0f;
  for (int cfg = 0; cfg < kNumBatchShapes; ++cfg) {
    if (kBatchShapes[cfg].scheme == SCHEME_SYMMETRIC && b->mode != MODE_PLAIN &&
        b->mode != MODE_PLAIN_SKIP)
      continue; // BRANCH with plain sums only
    if ((st = configure(b, cfg)) || (st = timed(b, 1, false, false)))
      return st; // warm-up step
    if ((st = copy_state(b, false))) return st;
    b->cur = cur0;
    if ((st = timed(b, trial_steps, false, false))) return st;
    if (best < 0 || b->last_ms < best_ms) {
      best = cfg;
      best_ms = b->last_ms;
    }
    b->cur = cur0;
    if ((st = copy_state(b, false))) return st;
  }
  if ((st = configure(b, best))) return st;
  if ((st = hip_check(hipStreamSynchronize(b->stream), "stream synchronize")))
    return st;
  if (best_cfg) *best_cfg = best;
  b->last_ms = best_ms;
  return NB_OK;
}

float nb_batch_last_ms(const nb_batch *b) { return b ? b->last_ms : 0.0f; }

size_t nb_batch_num_particles(const nb_batch *b) { return b ? b->total : 0; }

static int batch_read(nb_batch *b, float *const src[3], float *x, float *y,
                      float *z) {
  if (b->total == 0) return NB_OK;
  if (!x || !y || !z) return set_error(NB_ERR_INVALID_ARG, "output is NULL");
  float *dst[3] = {x, y, z};
  for (int c = 0; c < 3; ++c) {
    int st = hip_check(hipMemcpyAsync(dst[c], src[c], sizeof(float) * b->total,
                                      hipMemcpyDeviceToHost, b->stream),
                       "device-to-host copy");
    if (st) return st;
  }
  return hip_check(hipStreamSynchronize(b->stream), "stream synchronize");
}

int nb_batch_read_pos(nb_batch *b, float *x, float *y, float *z) {
  if (!b) return set_error(NB_ERR_INVALID_ARG, "batch is NULL");
  return batch_read(b, b->P[b->cur], x, y, z);
}

int nb_batch_read_vel(nb_batch *b, float *x, float *y, float *z) {
  if (!b) return set_error(NB_ERR_INVALID_ARG, "batch is NULL");
  return batch_read(b, b->V, x, y, z);
}

int nb_batch_destroy(nb_batch *b) {
  destroy_batch(b);
  return NB_OK;
}

} // extern "C"
