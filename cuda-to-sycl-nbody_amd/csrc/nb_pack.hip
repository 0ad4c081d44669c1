// SoA -> AoS vec4(x, y, z, w) pack: the renderer's particle buffer format
// (RendererGL::setParticleData writes glm::vec4(x, y, z, 1.0f) per particle
// from the host SoA arrays, src/renderer_gl.cpp:156-172).  Here the pack runs
// on the device: 12 B read and 16 B written per particle, one 16 B vector
// store per lane, coalesced; bound by HBM bandwidth.

#include <cstdint>

#include "nb_internal.hpp"

namespace nb {
namespace {

struct Six {
  const float *p[6];
};

__global__ __launch_bounds__(256) void pack_vec4(const float *__restrict__ x,
                                                 const float *__restrict__ y,
                                                 const float *__restrict__ z,
                                                 float w,
                                                 float4 *__restrict__ out,
                                                 int n) {
  const int stride = gridDim.x * 256;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    out[i] = make_float4(x[i], y[i], z[i], w);
}

// Six device arrays into one host-mapped pinned buffer (row c at dst + c *
// stride): the readback of a frame's positions and velocities as one launch
// whose stores cross the bus, instead of six DMA copies whose set-up time
// dominates at small n.
__global__ __launch_bounds__(256) void stage_six(Six src,
                                                 float *__restrict__ dst,
                                                 int n, long long stride) {
  const int c = blockIdx.y;
  const float *__restrict__ a = src.p[c];
  float *__restrict__ d = dst + c * stride;
  const int step = gridDim.x * 256;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += step) d[i] = a[i];
}

} // namespace

int launch_stage_six(const float *const src[6], float *dst, int n,
                     size_t stride, hipStream_t stream) {
  if (n <= 0) return NB_OK;
  if (!dst) return set_error(NB_ERR_INVALID_ARG, "staging destination is NULL");
  Six s;
  for (int c = 0; c < 6; ++c) {
    if (!src[c]) return set_error(NB_ERR_INVALID_ARG, "staging source is NULL");
    s.p[c] = src[c];
  }
  const int blocks = std::min((n + 255) / 256, 256 * 2);
  hipLaunchKernelGGL(stage_six, dim3(blocks, 6), dim3(256), 0, stream, s, dst,
                     n, (long long)stride);
  return hip_check(hipGetLastError(), "staging launch");
}

int launch_pack_vec4(const float *x, const float *y, const float *z, float w,
                     float *out, int n, hipStream_t stream) {
  if (n <= 0) return NB_OK;
  if (reinterpret_cast<uintptr_t>(out) % 16)
    return set_error(NB_ERR_INVALID_ARG, "vec4 output must be 16 B aligned");
  // a few waves per CU; the loop covers any n
  const int blocks = std::min((n + 255) / 256, 256 * 8);
  hipLaunchKernelGGL(pack_vec4, dim3(blocks), dim3(256), 0, stream, x, y, z,
                     w, reinterpret_cast<float4 *>(out), n);
  return hip_check(hipGetLastError(), "pack launch");
}

} // namespace nb

using namespace nb;

extern "C" int nb_pack_vec4_dev(const nb_vec3 *soa, float w, float *out,
                                size_t n, void *stream) {
  if (!soa) return set_error(NB_ERR_INVALID_ARG, "NULL argument");
  if (n > ((size_t)1 << 28)) return set_error(NB_ERR_INVALID_ARG, "n exceeds 2^28");
  if (n == 0) return NB_OK;
  if (!soa->x || !soa->y || !soa->z || !out)
    return set_error(NB_ERR_INVALID_ARG, "device pointer is NULL");
  const char *o = reinterpret_cast<const char *>(out);
  const size_t bytes_out = 16 * n, bytes_in = 4 * n;
  for (const float *a : {soa->x, soa->y, soa->z}) {
    const char *p = reinterpret_cast<const char *>(a);
    if (p < o + bytes_out && o < p + bytes_in)
      return set_error(NB_ERR_INVALID_ARG, "input and output overlap");
  }
  return launch_pack_vec4(soa->x, soa->y, soa->z, w, out, (int)n,
                          static_cast<hipStream_t>(stream));
}
