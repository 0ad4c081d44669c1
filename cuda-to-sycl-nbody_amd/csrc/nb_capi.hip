// C ABI of libnbody_amd: the simulator handle (DiskGalaxySimulator's device
// side, src/simulator.cu:24-129) and the stream-ordered device entry points.
// Nothing here throws across the boundary; HIP failures become status codes
// with a per-thread message (the reference's gpuErrchk prints and exits
// instead, src/simulator.cuh:22-31).

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <utility>

#include "nb_internal.hpp"

// Scratch of nb_forward_step_noalloc_dev: a file-scope device array, present
// on every device the library runs on for the life of the process.  In host
// code `scratch` names a host-side stand-in; the device address comes from
// hipGetSymbolAddress (static_scratch below), once per device.
constexpr size_t kStaticScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) float scratch[kStaticScratchBytes / sizeof(float)];

namespace nb {

namespace {
thread_local std::string g_last_error;
} // namespace

int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return NB_OK;
  const int code =
      (e == hipErrorOutOfMemory) ? NB_ERR_OUT_OF_MEMORY
      : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? NB_ERR_NO_DEVICE
                                                               : NB_ERR_HIP;
  return set_error(code, std::string(what) + ": " + hipGetErrorString(e));
}

int pair_mode(int calc_method, int acc, float eps) {
  if (calc_method == NB_CALC_PREDICATED) return MODE_PREDICATED;
  const bool self_zero = self_term_vanishes(eps);
  if (acc == NB_ACC_KAHAN) return self_zero ? MODE_KAHAN : MODE_KAHAN_SKIP;
  return self_zero ? MODE_PLAIN : MODE_PLAIN_SKIP;
}

namespace {
// One scratch buffer per (device, stream) pair, grown on demand.  Launches on
// one stream run in order, so successive calls on a stream share its buffer
// safely; a call on another stream gets another buffer.  The entry's mutex
// is held from lookup until the caller has enqueued its launches (the
// ScratchLease's lifetime), so a second host thread on the same stream waits
// instead of growing the buffer under the first one's launch.  Growing first
// waits for the stream (every earlier use of the buffer was on it), then
// frees.  Entries are never destroyed; nb_scratch_release frees their
// buffers.
struct ScratchEntry {
  int device = 0;
  hipStream_t stream = nullptr;
  float *ptr = nullptr;
  size_t bytes = 0;
  std::mutex mu;
};
std::mutex g_scratch_mu; // guards g_scratch (the list, not the entries)
std::vector<std::unique_ptr<ScratchEntry>> g_scratch;

ScratchEntry *scratch_entry(int dev, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  for (auto &e : g_scratch)
    if (e->device == dev && e->stream == stream) return e.get();
  g_scratch.push_back(std::make_unique<ScratchEntry>());
  g_scratch.back()->device = dev;
  g_scratch.back()->stream = stream;
  return g_scratch.back().get();
}
} // namespace

int acquire_scratch(size_t bytes, hipStream_t stream, ScratchLease &lease) {
  lease = ScratchLease();
  if (bytes == 0) return NB_OK;
  int dev = 0;
  int st = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (st) return st;
  ScratchEntry *e = scratch_entry(dev, stream);
  std::unique_lock<std::mutex> lock(e->mu);
  if (e->bytes < bytes) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if ((st = hip_check(hipStreamIsCapturing(stream, &cap), "hipStreamIsCapturing")))
      return st;
    if (cap != hipStreamCaptureStatusNone)
      return set_error(NB_ERR_UNSUPPORTED,
                       "the call needs " + std::to_string(bytes) +
                           " bytes of scratch on this stream, more than it holds: "
                           "make the same call once before capturing the stream");
    if (e->ptr) {
      // every earlier use of the buffer was enqueued on this stream
      if ((st = hip_check(hipStreamSynchronize(stream), "stream synchronize")))
        return st;
      (void)hipFree(e->ptr);
      e->ptr = nullptr;
      e->bytes = 0;
    }
    void *p = nullptr;
    if ((st = hip_check(hipMalloc(&p, bytes), "scratch hipMalloc"))) return st;
    e->ptr = static_cast<float *>(p);
    e->bytes = bytes;
  }
  lease.ptr_ = e->ptr;
  lease.lock_ = std::move(lock);
  return NB_OK;
}

} // namespace nb

extern "C" int nb_scratch_release(void) {
  using namespace nb;
  int dev = 0;
  int st = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (st) return st;
  // Every launch that uses a buffer is enqueued under its entry's lock: with
  // all of the device's entry locks held no new one can be, and the device
  // synchronisation then waits for those already enqueued.
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  std::vector<std::unique_lock<std::mutex>> held;
  for (auto &e : g_scratch)
    if (e->device == dev) held.emplace_back(e->mu);
  if ((st = hip_check(hipDeviceSynchronize(), "device synchronize"))) return st;
  for (auto &e : g_scratch) {
    if (e->device != dev) continue;
    if (e->ptr) (void)hipFree(e->ptr);
    e->ptr = nullptr;
    e->bytes = 0;
  }
  return NB_OK;
}

extern "C" int nb_scratch_held(size_t *bytes) {
  using namespace nb;
  if (!bytes) return set_error(NB_ERR_INVALID_ARG, "bytes is NULL");
  *bytes = 0;
  int dev = 0;
  int st = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (st) return st;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  for (auto &e : g_scratch) {
    if (e->device != dev) continue;
    std::lock_guard<std::mutex> entry_lock(e->mu);
    *bytes += e->bytes;
  }
  return NB_OK;
}

namespace nb {
} // namespace nb

using namespace nb;

struct nb_sim {
  nb_params p{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  size_t n = 0, stride = 0;
  float *mem = nullptr; // 9 SoA arrays: pos, pos_next, vel
  float *pos[3] = {}, *nxt[3] = {}, *vel[3] = {};
  float *scratch = nullptr;
  size_t scratch_bytes = 0;
  float *vec4 = nullptr; // nb_sim_read_vec4's packing buffer (4 n floats)
  int acc = NB_ACC_PLAIN;
  int cfg = 0;      // launch shape in use
  int cfg_req = -1; // shape asked for with nb_sim_set_config (-1: default)
  float last_ms = 0.0f, last_kernel_ms = 0.0f;
  float *staging = nullptr; // pinned host copy of pos and vel (6 arrays)
  float *staging_dev = nullptr; // the same memory, device address
  hipEvent_t ev_stage = nullptr;
  bool staged = false;
  // Frames of two or more steps replay a captured hipGraph, one per start
  // buffer, valid for the (steps, shape, pair mode, scratch) they hold.
  bool use_graph = true;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_iters = 0, graph_cfg = -1, graph_mode = -1;
  const float *graph_scratch = nullptr;
};

namespace {

constexpr size_t kMaxParticles = (size_t)1 << 28; // int indices, 3 KiB-scale margins

int check_params(const nb_params *p) {
  if (!p) return set_error(NB_ERR_INVALID_ARG, "params is NULL");
  if (p->calc_method != NB_CALC_BRANCH && p->calc_method != NB_CALC_PREDICATED)
    return set_error(NB_ERR_INVALID_ARG,
                     "Valid calculation methods are BRANCH or PREDICATED");
  if (p->num_particles > kMaxParticles)
    return set_error(NB_ERR_INVALID_ARG, "num_particles exceeds 2^28");
  if (p->sim_iterations_per_frame < 0)
    return set_error(NB_ERR_INVALID_ARG, "sim_iterations_per_frame < 0");
  return NB_OK;
}

StepConsts consts_of(const nb_params *p) {
  StepConsts c;
  c.eps = p->dist_eps;
  c.damping = p->damping;
  c.dt = p->dt;
  c.G = p->G;
  return c;
}

void drop_sim_graphs(nb_sim *s) {
  for (hipGraphExec_t &g : s->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

void destroy_sim(nb_sim *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  drop_sim_graphs(s);
  if (s->mem) (void)hipFree(s->mem);
  if (s->scratch) (void)hipFree(s->scratch);
  if (s->vec4) (void)hipFree(s->vec4);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->ev_stage) (void)hipEventDestroy(s->ev_stage);
  if (s->staging) (void)hipHostFree(s->staging);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int ensure_sim_scratch(nb_sim *s) {
  const size_t bytes =
      s->n ? forward_scratch_bytes(
                 s->cfg, s->n, pair_mode(s->p.calc_method, s->acc, s->p.dist_eps))
           : 0;
  if (bytes == s->scratch_bytes) return NB_OK;
  if (s->scratch) {
    (void)hipFree(s->scratch);
    s->scratch = nullptr;
    s->scratch_bytes = 0;
  }
  if (bytes == 0) return NB_OK;
  void *p = nullptr;
  int st = hip_check(hipMalloc(&p, bytes), "scratch hipMalloc");
  if (st) return st;
  s->scratch = static_cast<float *>(p);
  s->scratch_bytes = bytes;
  return NB_OK;
}

// `iters` stream-ordered steps launched one by one, ping-ponging the
// position buffers (src/simulator.cu:65).
int launch_steps(nb_sim *s, int iters, int mode) {
  const StepConsts c = consts_of(&s->p);
  for (int k = 0; k < iters; ++k) {
    int st = launch_forward(c, mode, s->cfg, s->pos[0], s->pos[1], s->pos[2],
                            s->vel[0], s->vel[1], s->vel[2], s->nxt[0],
                            s->nxt[1], s->nxt[2], (int)s->n, s->scratch,
                            s->stream);
    if (st) return st;
    for (int d = 0; d < 3; ++d) std::swap(s->pos[d], s->nxt[d]);
  }
  return NB_OK;
}

// Launch shape for the current accumulation mode and parameters: the one
// asked for with nb_sim_set_config, else the library's choice.
int refresh_config(nb_sim *s) {
  const int mode = pair_mode(s->p.calc_method, s->acc, s->p.dist_eps);
  s->cfg = s->cfg_req < 0 ? forward_default_config(s->n, mode) : s->cfg_req;
  return ensure_sim_scratch(s);
}

} // namespace

extern "C" {

void nb_params_default(nb_params *p) {
  if (!p) return;
  p->G = 2.0f;
  p->dt = 0.005f;
  p->num_particles = 50 * 256;
  p->num_frames = SIZE_MAX;
  p->sim_iterations_per_frame = 4;
  p->damping = 0.999998f;
  p->dist_eps = 1.0e-7f;
  p->gw_size = 64;
  p->calc_method = NB_CALC_BRANCH;
}

const char *nb_last_error(void) { return g_last_error.c_str(); }

int nb_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int nb_device_name(char *buf, size_t len) {
  if (!buf || len == 0) return set_error(NB_ERR_INVALID_ARG, "empty buffer");
  hipDeviceProp_t prop;
  std::string name = "Unknown Device";
  if (hipGetDeviceProperties(&prop, 0) == hipSuccess) {
    name = prop.name;
    // Without the marketing-name table the runtime reports a generic name;
    // the ISA name identifies the part.
    const std::string arch = prop.gcnArchName;
    const std::string isa = arch.substr(0, arch.find(':'));
    if (!isa.empty() && name.find(isa) == std::string::npos)
      name += " [" + isa + "]";
  } else {
    (void)hipGetLastError();
  }
  std::strncpy(buf, name.c_str(), len - 1);
  buf[len - 1] = '\0';
  return NB_OK;
}

int nb_sim_create(const nb_params *p, const float *px, const float *py,
                  const float *pz, const float *vx, const float *vy,
                  const float *vz, nb_sim **out) {
  if (!out) return set_error(NB_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  int st = check_params(p);
  if (st) return st;
  const size_t n = p->num_particles;
  if (n > 0 && (!px || !py || !pz || !vx || !vy || !vz))
    return set_error(NB_ERR_INVALID_ARG, "particle array is NULL");
  if (nb_device_count() == 0)
    return set_error(NB_ERR_NO_DEVICE, "no HIP device visible");

  nb_sim *s = new (std::nothrow) nb_sim;
  if (!s) return set_error(NB_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->p = *p;
  s->n = n;
  s->stride = (n + 63) / 64 * 64; // every array starts 256 B aligned
  s->cfg = forward_default_config(
      n, pair_mode(s->p.calc_method, s->acc, s->p.dist_eps));
  if ((st = hip_check(hipGetDevice(&s->device), "hipGetDevice")) ||
      (st = hip_check(hipStreamCreateWithFlags(&s->stream,
                                               hipStreamNonBlocking),
                      "hipStreamCreate")) ||
      (st = hip_check(hipEventCreate(&s->ev0), "hipEventCreate")) ||
      (st = hip_check(hipEventCreate(&s->ev1), "hipEventCreate"))) {
    destroy_sim(s);
    return st;
  }
  if (n > 0) {
    void *mem = nullptr;
    st = hip_check(hipMalloc(&mem, sizeof(float) * 9 * s->stride),
                   "particle hipMalloc");
    if (st) {
      destroy_sim(s);
      return st;
    }
    s->mem = static_cast<float *>(mem);
    for (int c = 0; c < 3; ++c) {
      s->pos[c] = s->mem + (0 + c) * s->stride;
      s->nxt[c] = s->mem + (3 + c) * s->stride;
      s->vel[c] = s->mem + (6 + c) * s->stride;
    }
    const float *src[6] = {px, py, pz, vx, vy, vz};
    float *dst[6] = {s->pos[0], s->pos[1], s->pos[2],
                     s->vel[0], s->vel[1], s->vel[2]};
    for (int k = 0; k < 6 && !st; ++k)
      st = hip_check(hipMemcpyAsync(dst[k], src[k], sizeof(float) * n,
                                    hipMemcpyHostToDevice, s->stream),
                     "host-to-device copy");
    if (!st) st = hip_check(hipStreamSynchronize(s->stream), "sync");
    if (!st) st = ensure_sim_scratch(s);
    if (st) {
      destroy_sim(s);
      return st;
    }
  }
  *out = s;
  return NB_OK;
}

int nb_sim_advance(nb_sim *s, int iters) {
  if (!s) return set_error(NB_ERR_INVALID_ARG, "sim is NULL");
  if (iters < 0) return set_error(NB_ERR_INVALID_ARG, "iters < 0");
  if (s->n == 0) return NB_OK;
  const int mode = pair_mode(s->p.calc_method, s->acc, s->p.dist_eps);
  if (!s->use_graph || iters < 2) return launch_steps(s, iters, mode);
  if (s->graph_iters != iters || s->graph_cfg != s->cfg ||
      s->graph_mode != mode || s->graph_scratch != s->scratch) {
    int st = hip_check(hipStreamSynchronize(s->stream), "stream synchronize");
    if (st) return st;
    drop_sim_graphs(s);
    s->graph_iters = iters;
    s->graph_cfg = s->cfg;
    s->graph_mode = mode;
    s->graph_scratch = s->scratch;
  }
  const int from = s->pos[0] == s->mem ? 0 : 1;
  if (!s->graph[from]) {
    float *pos[3], *nxt[3];
    for (int d = 0; d < 3; ++d) {
      pos[d] = s->pos[d];
      nxt[d] = s->nxt[d];
    }
    int st = hip_check(hipStreamBeginCapture(s->stream,
                                             hipStreamCaptureModeThreadLocal),
                       "hipStreamBeginCapture");
    if (st) return st;
    const int st_launch = launch_steps(s, iters, mode);
    for (int d = 0; d < 3; ++d) { // capturing advances nothing
      s->pos[d] = pos[d];
      s->nxt[d] = nxt[d];
    }
    hipGraph_t g = nullptr;
    st = hip_check(hipStreamEndCapture(s->stream, &g), "hipStreamEndCapture");
    if (st_launch || st) {
      if (g) (void)hipGraphDestroy(g);
      return st_launch ? st_launch : st;
    }
    st = hip_check(hipGraphInstantiate(&s->graph[from], g, nullptr, nullptr, 0),
                   "hipGraphInstantiate");
    (void)hipGraphDestroy(g);
    if (st) return st;
  }
  int st = hip_check(hipGraphLaunch(s->graph[from], s->stream), "hipGraphLaunch");
  if (st) return st;
  if (iters & 1)
    for (int d = 0; d < 3; ++d) std::swap(s->pos[d], s->nxt[d]);
  return NB_OK;
}

int nb_sim_set_graph(nb_sim *s, int enable) {
  if (!s) return set_error(NB_ERR_INVALID_ARG, "sim is NULL");
  int st = hip_check(hipStreamSynchronize(s->stream), "stream synchronize");
  if (st) return st;
  s->use_graph = enable != 0;
  if (!s->use_graph) drop_sim_graphs(s);
  return NB_OK;
}

int nb_sim_sync(nb_sim *s) {
  if (!s) return set_error(NB_ERR_INVALID_ARG, "sim is NULL");
  return hip_check(hipStreamSynchronize(s->stream), "stream synchronize");
}

int nb_sim_step(nb_sim *s) {
  if (!s) return set_error(NB_ERR_INVALID_ARG, "sim is NULL");
  const auto t0 = std::chrono::steady_clock::now();
  int st = hip_check(hipEventRecord(s->ev0, s->stream), "hipEventRecord");
  if (!st) st = nb_sim_advance(s, s->p.sim_iterations_per_frame);
  if (!st) st = hip_check(hipEventRecord(s->ev1, s->stream), "hipEventRecord");
  if (!st) st = hip_check(hipStreamSynchronize(s->stream), "stream synchronize");
  const auto t1 = std::chrono::steady_clock::now();
  if (st) return st;
  s->last_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, s->ev0, s->ev1) == hipSuccess)
    s->last_kernel_ms = ms;
  else
    (void)hipGetLastError();
  return NB_OK;
}

static int read3(nb_sim *s, float *const src[3], float *x, float *y,
                 float *z) {
  if (!s) return set_error(NB_ERR_INVALID_ARG, "sim is NULL");
  if (s->n == 0) return NB_OK;
  if (!x || !y || !z) return set_error(NB_ERR_INVALID_ARG, "output is NULL");
  float *dst[3] = {x, y, z};
  for (int c = 0; c < 3; ++c) {
    int st = hip_check(hipMemcpyAsync(dst[c], src[c], sizeof(float) * s->n,
                                      hipMemcpyDeviceToHost, s->stream),
                       "device-to-host copy");
    if (st) return st;
  }
  return hip_check(hipStreamSynchronize(s->stream), "stream synchronize");
}

int nb_sim_stage_readback(nb_sim *s) {
  if (!s) return set_error(NB_ERR_INVALID_ARG, "sim is NULL");
  if (s->n == 0) {
    s->staged = true;
    return NB_OK;
  }
  int st = NB_OK;
  if (!s->staging_dev) {
    // pinned, coherent and mapped into the device's address space: the
    // staging kernel stores straight into it
    void *p = nullptr, *d = nullptr;
    if ((st = hip_check(hipHostMalloc(&p, sizeof(float) * 6 * s->stride,
                                      hipHostMallocMapped),
                        "pinned hipHostMalloc")))
      return st;
    if ((st = hip_check(hipHostGetDevicePointer(&d, p, 0),
                        "hipHostGetDevicePointer"))) {
      (void)hipHostFree(p); // no half-set-up state is kept
      return st;
    }
    s->staging = static_cast<float *>(p);
    s->staging_dev = static_cast<float *>(d);
  }
  if (!s->ev_stage &&
      (st = hip_check(hipEventCreateWithFlags(&s->ev_stage,
                                              hipEventDisableTiming),
                      "hipEventCreate")))
    return st;
  const float *const src[6] = {s->pos[0], s->pos[1], s->pos[2],
                               s->vel[0], s->vel[1], s->vel[2]};
  if ((st = launch_stage_six(src, s->staging_dev, (int)s->n, s->stride,
                             s->stream)))
    return st;
  if ((st = hip_check(hipEventRecord(s->ev_stage, s->stream), "hipEventRecord")))
    return st;
  s->staged = true;
  return NB_OK;
}

int nb_sim_staged(nb_sim *s, const float **arrays) {
  if (!s || !arrays) return set_error(NB_ERR_INVALID_ARG, "NULL argument");
  if (!s->staged)
    return set_error(NB_ERR_INVALID_ARG, "nb_sim_stage_readback was not called");
  for (int c = 0; c < 6; ++c)
    arrays[c] = s->staging ? s->staging + c * s->stride : nullptr;
  if (s->n == 0) return NB_OK;
  return hip_check(hipEventSynchronize(s->ev_stage), "hipEventSynchronize");
}

int nb_sim_read_vec4(nb_sim *s, int which, float w, float *out) {
  if (!s) return set_error(NB_ERR_INVALID_ARG, "sim is NULL");
  if (which != 0 && which != 1)
    return set_error(NB_ERR_INVALID_ARG, "which must be 0 (pos) or 1 (vel)");
  if (s->n == 0) return NB_OK;
  if (!out) return set_error(NB_ERR_INVALID_ARG, "output is NULL");
  int st = NB_OK;
  if (!s->vec4) { // the handle's own packing buffer, freed with the handle
    void *p = nullptr;
    if ((st = hip_check(hipMalloc(&p, sizeof(float) * 4 * s->n), "vec4 hipMalloc")))
      return st;
    s->vec4 = static_cast<float *>(p);
  }
  float *const buf = s->vec4;
  float *const *src = which == 0 ? s->pos : s->vel;
  if ((st = launch_pack_vec4(src[0], src[1], src[2], w, buf, (int)s->n,
                             s->stream)) ||
      (st = hip_check(hipMemcpyAsync(out, buf, sizeof(float) * 4 * s->n,
                                     hipMemcpyDeviceToHost, s->stream),
                      "device-to-host copy")))
    return st;
  return hip_check(hipStreamSynchronize(s->stream), "stream synchronize");
}

int nb_sim_read_pos(nb_sim *s, float *x, float *y, float *z) {
  return s ? read3(s, s->pos, x, y, z)
           : set_error(NB_ERR_INVALID_ARG, "sim is NULL");
}

int nb_sim_read_vel(nb_sim *s, float *x, float *y, float *z) {
  return s ? read3(s, s->vel, x, y, z)
           : set_error(NB_ERR_INVALID_ARG, "sim is NULL");
}

float nb_sim_last_step_ms(const nb_sim *s) { return s ? s->last_ms : 0.0f; }
float nb_sim_last_kernel_ms(const nb_sim *s) {
  return s ? s->last_kernel_ms : 0.0f;
}
size_t nb_sim_num_particles(const nb_sim *s) { return s ? s->n : 0; }
int nb_sim_calc_method(const nb_sim *s) { return s ? s->p.calc_method : 0; }
int nb_sim_gw_size(const nb_sim *s) { return s ? s->p.gw_size : 0; }

int nb_sim_set_accumulation(nb_sim *s, int mode) {
  if (!s) return set_error(NB_ERR_INVALID_ARG, "sim is NULL");
  if (mode != NB_ACC_PLAIN && mode != NB_ACC_KAHAN)
    return set_error(NB_ERR_INVALID_ARG, "unknown accumulation mode");
  int st = hip_check(hipStreamSynchronize(s->stream), "stream synchronize");
  if (st) return st;
  s->acc = mode;
  return refresh_config(s);
}

int nb_sim_set_frame_iterations(nb_sim *s, int iters) {
  if (!s) return set_error(NB_ERR_INVALID_ARG, "sim is NULL");
  if (iters < 0) return set_error(NB_ERR_INVALID_ARG, "iters < 0");
  s->p.sim_iterations_per_frame = iters;
  return NB_OK;
}

int nb_sim_set_config(nb_sim *s, int cfg) {
  if (!s) return set_error(NB_ERR_INVALID_ARG, "sim is NULL");
  if (cfg < -1 || cfg >= forward_config_count())
    return set_error(NB_ERR_INVALID_ARG, "forward config out of range");
  int st = hip_check(hipStreamSynchronize(s->stream), "stream synchronize");
  if (st) return st;
  s->cfg_req = cfg;
  return refresh_config(s);
}

int nb_sim_config(const nb_sim *s) { return s ? s->cfg : -1; }

int nb_sim_destroy(nb_sim *s) {
  destroy_sim(s);
  return NB_OK;
}

} // extern "C"

namespace {

// Argument checks of the stream-ordered forward entry points; on success
// *mode and *cfg hold the pair mode and the launch shape (the library's
// choice for cfg = -1).  n == 0 passes (nothing to launch).
int check_forward_dev(const nb_params *p, const float *px, const float *py,
                      const float *pz, const float *vx, const float *vy,
                      const float *vz, const float *nx, const float *ny,
                      const float *nz, size_t n, int acc, int *cfg,
                      int *mode) {
  int st = check_params(p);
  if (st) return st;
  if (n > kMaxParticles)
    return set_error(NB_ERR_INVALID_ARG, "n exceeds 2^28");
  if (acc != NB_ACC_PLAIN && acc != NB_ACC_KAHAN)
    return set_error(NB_ERR_INVALID_ARG, "unknown accumulation mode");
  if (*cfg < -1 || *cfg >= forward_config_count())
    return set_error(NB_ERR_INVALID_ARG, "forward config out of range");
  if (n == 0) return NB_OK;
  if (!px || !py || !pz || !vx || !vy || !vz || !nx || !ny || !nz)
    return set_error(NB_ERR_INVALID_ARG, "device pointer is NULL");
  if (px == nx || py == ny || pz == nz)
    return set_error(NB_ERR_INVALID_ARG,
                     "positions and next positions must not alias");
  *mode = pair_mode(p->calc_method, acc, p->dist_eps);
  if (*cfg < 0) *cfg = forward_default_config(n, *mode);
  return NB_OK;
}

} // namespace

extern "C" {

int nb_forward_step_dev(const nb_params *p, const float *px, const float *py,
                        const float *pz, float *vx, float *vy, float *vz,
                        float *nx, float *ny, float *nz, size_t n, int cfg,
                        int acc, void *stream) {
  int mode = 0;
  int st = check_forward_dev(p, px, py, pz, vx, vy, vz, nx, ny, nz, n, acc,
                             &cfg, &mode);
  if (st || n == 0) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  ScratchLease lease; // held until the launches are enqueued
  if ((st = acquire_scratch(forward_scratch_bytes(cfg, n, mode), s, lease)))
    return st;
  return launch_forward(consts_of(p), mode, cfg, px, py, pz, vx, vy, vz, nx,
                        ny, nz, (int)n, lease.get(), s);
}

} // extern "C"

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_static_mu;
float *g_static_ptr[kMaxDevices] = {}; // device address of `scratch`, per device

// Device address of the static scratch array on the current device, looked
// up on the first call on that device.
int static_scratch(float **out) {
  *out = nullptr;
  int dev = 0;
  int st = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (st) return st;
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(NB_ERR_UNSUPPORTED, "device index too large");
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[dev]) {
    void *ptr = nullptr;
    if ((st = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)),
                        "hipGetSymbolAddress(scratch)")))
      return st;
    g_static_ptr[dev] = static_cast<float *>(ptr);
  }
  *out = g_static_ptr[dev];
  return NB_OK;
}

} // namespace

extern "C" {

int nb_forward_step_noalloc_dev(const nb_params *p, const float *px,
                                const float *py, const float *pz, float *vx,
                                float *vy, float *vz, float *nx, float *ny,
                                float *nz, size_t n, int cfg, int acc,
                                void *stream) {
  int mode = 0;
  int st = check_forward_dev(p, px, py, pz, vx, vy, vz, nx, ny, nz, n, acc,
                             &cfg, &mode);
  if (st || n == 0) return st;
  const size_t bytes = forward_scratch_bytes(cfg, n, mode);
  if (bytes > kStaticScratchBytes)
    return set_error(NB_ERR_UNSUPPORTED,
                     "the launch shape needs " + std::to_string(bytes) +
                         " bytes of scratch at n = " + std::to_string(n) +
                         ", more than the 64 MiB static buffer of the "
                         "allocation-free step: use nb_forward_step_dev or a "
                         "shape with fewer partial sums");
  float *buf = nullptr;
  if (bytes && (st = static_scratch(&buf))) return st;
  return launch_forward(consts_of(p), mode, cfg, px, py, pz, vx, vy, vz, nx,
                        ny, nz, (int)n, buf, static_cast<hipStream_t>(stream));
}

int nb_forward_config_count(void) { return forward_config_count(); }

int nb_forward_default_config(const nb_params *p, size_t n, int acc) {
  if (check_params(p)) return -1;
  if (acc != NB_ACC_PLAIN && acc != NB_ACC_KAHAN) {
    set_error(NB_ERR_INVALID_ARG, "unknown accumulation mode");
    return -1;
  }
  return forward_default_config(n, pair_mode(p->calc_method, acc, p->dist_eps));
}

int nb_forward_config_info(int cfg, int *block, int *unroll, int *tile,
                           int *split) {
  if (cfg < 0 || cfg >= forward_config_count())
    return set_error(NB_ERR_INVALID_ARG, "forward config out of range");
  const FwdConfig &c = forward_config(cfg);
  if (block) *block = c.block;
  if (unroll) *unroll = c.unroll;
  if (tile) *tile = c.tile;
  if (split) *split = c.split;
  return NB_OK;
}

int nb_forward_config_scheme(int cfg) {
  if (cfg < 0 || cfg >= forward_config_count()) {
    set_error(NB_ERR_INVALID_ARG, "forward config out of range");
    return -1;
  }
  return forward_config(cfg).scheme;
}

} // extern "C"
