// Issue-rate microbenchmark of the gfx950 instructions that make up the
// all-pairs pair loop: v_fmac_f32 (FP32 multiply-add), v_rsq_f32 (the
// transcendental inverse square root), the DPP wave rotation
// (v_mov_b32_dpp wave_rol:1) and the forward pair term of the symmetric
// kernels (3 sub, 3 fma, 2 mul, v_rsq_f32, 6 fma), each as one long
// dependent chain per lane (latency) and as several independent chains
// (throughput), at 1..4 waves per SIMD.
//
// Every kernel works in registers only and writes one float per thread
// (keeps the chains alive) and, from lane 0 of each wave, the wave's shader
// clock cycles (s_memtime) around its loop.  Reported per case:
//   cyc/inst    cycles of one wave per instruction it issues (latency bound
//               for one chain);
//   simd_cpi    SIMD cycles per wave-instruction = cyc/inst / waves per SIMD
//               (throughput bound with enough chains and waves);
//   per_cu_clk  wave-instructions issued per CU per clock = 4 / simd_cpi
//               (MI355X: 4 SIMDs per CU);
//   ghz         shader clock (s_memtime cycles / s_memrealtime at 100 MHz).
//
//   issue_rates [--iters N]
//
// Prints one line per case and a JSON summary line.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int kBlock = 256; // 4 waves: one per SIMD of a CU
constexpr int kUnroll = 16; // chain steps per loop trip

#define CHECK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));           \
      std::exit(1);                                                           \
    }                                                                         \
  } while (0)

__device__ __forceinline__ float wave_rol1(float v) {
  const int b = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(b, b, 0x134, 0xf, 0xf, false));
}

enum Op : int {
  OP_FMA = 0,     // v_fmaak_f32 (literal addend)
  OP_RSQ = 1,     // v_rsq_f32
  OP_DPP = 2,     // v_mov_b32_dpp wave_rol:1
  OP_PAIR = 3,    // the forward pair term, one dependent chain per c
  OP_FMAC_VV = 4, // v_fmac_f32_e32, VGPR operands only
  OP_FMA_S = 5,   // v_fma_f32 (VOP3) with an SGPR addend
  OP_SUB = 6,     // v_sub_f32_e32
  OP_MUL = 7,     // v_mul_f32_e32
  OP_MIX_RSQ = 8, // per step: CHAINS v_fmac_f32 + one v_rsq_f32 every 2 steps
  OP_MIX_DPP = 9, // per step: CHAINS v_fmac_f32 + one DPP move every 2 steps
  OP_ADD_S = 10,   // v_add_f32_e32 with an SGPR first operand
  OP_FMA_VVV = 11, // v_fma_f32 (VOP3), VGPR operands only
  OP_FMAC_SQ = 12, // v_fmac_f32_e32 d, r, r (two distinct VGPRs read)
};

// Instructions one chain step issues (the loop counter adds a few SALU/VALU
// per trip, not counted: < 1 % at kUnroll x CHAINS >= 16).  The mixed cases
// count their extra instruction per 2 steps in steps_insts below.
constexpr int insts_per_step(int op) {
  return op == OP_PAIR ? 16 : 1; // pair: 14 VALU + v_rsq_f32 + the link
}

struct Out {
  float *val;          // [threads]
  long long *cycles;   // [waves]
  long long *realtime; // [waves] (100 MHz ticks)
};

template <int OP, int CHAINS>
__global__ __launch_bounds__(kBlock) void chain_kernel(Out o, int iters,
                                                       float x) {
  const int tid = blockIdx.x * kBlock + threadIdx.x;
  const float xl = x + 1.0e-9f * (float)threadIdx.x;
  const float yl = 1.0e-3f + 1.0e-9f * (float)threadIdx.x;
  float a[CHAINS], b[CHAINS][3], t[CHAINS][3], s[CHAINS][3];
  float r[4] = {1.5f + yl, 1.6f + yl, 1.7f + yl, 1.8f + yl}; // side chains
#pragma unroll
  for (int c = 0; c < CHAINS; ++c) {
    a[c] = 1.0f + 0.01f * (float)c + 1.0e-6f * (float)threadIdx.x;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      b[c][e] = 0.0f;
      t[c][e] = 0.1f * (float)(c + e) + 1.0e-4f * (float)threadIdx.x;
      s[c][e] = 0.0f;
    }
  }
  const long long c0 = __builtin_amdgcn_s_memtime();
  const long long r0 = __builtin_amdgcn_s_memrealtime();
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) {
        if constexpr (OP == OP_FMA) {
          a[c] = __builtin_fmaf(a[c], xl, 1.0e-3f);
        } else if constexpr (OP == OP_RSQ) {
          a[c] = __builtin_amdgcn_rsqf(a[c]);
        } else if constexpr (OP == OP_DPP) {
          a[c] = wave_rol1(a[c]);
        } else if constexpr (OP == OP_FMAC_VV) {
          a[c] = __builtin_fmaf(xl, t[c][0], a[c]);
        } else if constexpr (OP == OP_FMA_S) {
          a[c] = __builtin_fmaf(a[c], xl, x);
        } else if constexpr (OP == OP_SUB) {
          a[c] = a[c] - t[c][0];
        } else if constexpr (OP == OP_MUL) {
          a[c] = a[c] * xl;
        } else if constexpr (OP == OP_ADD_S) {
          a[c] = x + a[c];
        } else if constexpr (OP == OP_FMA_VVV) {
          a[c] = __builtin_fmaf(a[c], xl, t[c][1]);
        } else if constexpr (OP == OP_FMAC_SQ) {
          a[c] = __builtin_fmaf(t[c][2], t[c][2], a[c]);
        } else if constexpr (OP == OP_MIX_RSQ || OP == OP_MIX_DPP) {
          a[c] = __builtin_fmaf(xl, t[c][0], a[c]);
          if (c == CHAINS - 1 && (u & 1)) {
            float &v = r[(u >> 1) & 3];
            if constexpr (OP == OP_MIX_RSQ) v = __builtin_amdgcn_rsqf(v);
            else v = wave_rol1(v);
          }
        } else {
          // forward pair term: source a[c] (+ offsets) against target t[c];
          // target sums b[c], reaction sums s[c]; the next step's source
          // depends on this step's K (one chain per c)
          const float rx = a[c] - t[c][0];
          const float ry = a[c] - t[c][1];
          const float rz = a[c] - t[c][2];
          float d = __builtin_fmaf(rx, rx, 1.0e-7f);
          d = __builtin_fmaf(ry, ry, d);
          d = __builtin_fmaf(rz, rz, d);
          const float K = __builtin_amdgcn_rsqf((d * d) * d);
          b[c][0] = __builtin_fmaf(rx, K, b[c][0]);
          b[c][1] = __builtin_fmaf(ry, K, b[c][1]);
          b[c][2] = __builtin_fmaf(rz, K, b[c][2]);
          s[c][0] = __builtin_fmaf(rx, K, s[c][0]);
          s[c][1] = __builtin_fmaf(ry, K, s[c][1]);
          s[c][2] = __builtin_fmaf(rz, K, s[c][2]);
          a[c] = __builtin_fmaf(K, 1.0e-9f, a[c]); // the chain's link
        }
      }
    }
  }
  const long long c1 = __builtin_amdgcn_s_memtime();
  const long long r1 = __builtin_amdgcn_s_memrealtime();
  float acc = r[0] + r[1] + r[2] + r[3];
#pragma unroll
  for (int c = 0; c < CHAINS; ++c)
    acc += a[c] + b[c][0] + b[c][1] + b[c][2] + s[c][0] + s[c][1] + s[c][2];
  o.val[tid] = acc;
  if ((threadIdx.x & 63) == 0) {
    const int wave = tid >> 6;
    o.cycles[wave] = c1 - c0;
    o.realtime[wave] = r1 - r0;
  }
}

struct Case {
  const char *name;
  int op, chains;
  const void *fn;
};

template <int OP, int C>
Case make(const char *name) {
  return Case{name, OP, C, reinterpret_cast<const void *>(&chain_kernel<OP, C>)};
}

} // namespace

int main(int argc, char **argv) {
  int iters = 4096;
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], "--iters")) iters = std::atoi(argv[i + 1]);
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  const int cus = prop.multiProcessorCount;
  std::printf("device %s (%s), %d CUs\n", prop.name, prop.gcnArchName, cus);

  const Case cases[] = {
      make<OP_FMA, 1>("fma dep"),   make<OP_FMA, 2>("fma x2"),
      make<OP_FMA, 4>("fma x4"),    make<OP_FMA, 8>("fma x8"),
      make<OP_RSQ, 1>("rsq dep"),   make<OP_RSQ, 4>("rsq x4"),
      make<OP_RSQ, 8>("rsq x8"),    make<OP_DPP, 1>("dpp dep"),
      make<OP_DPP, 8>("dpp x8"),    make<OP_PAIR, 1>("pair dep"),
      make<OP_PAIR, 2>("pair x2"),  make<OP_PAIR, 4>("pair x4"),
      make<OP_FMAC_VV, 8>("fmac_vv x8"), make<OP_FMA_S, 8>("fma_sgpr x8"),
      make<OP_SUB, 8>("sub x8"),    make<OP_MUL, 8>("mul x8"),
      make<OP_MIX_RSQ, 7>("fmac14+rsq"), make<OP_MIX_DPP, 8>("fmac16+dpp"),
      make<OP_ADD_S, 8>("add_sgpr x8"), make<OP_FMA_VVV, 8>("fma_vop3 x8"),
      make<OP_FMAC_SQ, 8>("fmac_sq x8"),
  };
  const int wps_list[] = {1, 2, 3, 4}; // waves per SIMD
  const int max_groups = cus * 4;
  float *val = nullptr;
  long long *cyc = nullptr, *rt = nullptr;
  CHECK(hipMalloc(&val, sizeof(float) * max_groups * kBlock));
  CHECK(hipMalloc(&cyc, sizeof(long long) * max_groups * (kBlock / 64)));
  CHECK(hipMalloc(&rt, sizeof(long long) * max_groups * (kBlock / 64)));
  std::vector<long long> hc(max_groups * (kBlock / 64)),
      hr(max_groups * (kBlock / 64));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));

  std::string json = "{\"issue_rates\": [";
  bool first = true;
  for (const Case &c : cases) {
    for (int wps : wps_list) {
      const int groups = cus * wps; // one 4-wave group per CU per wave/SIMD
      const int waves = groups * (kBlock / 64);
      // pair cases carry 15 instructions per step: fewer trips
      const int it = c.op == OP_PAIR ? std::max(1, iters / 8) : iters;
      Out o{val, cyc, rt};
      void *args[] = {&o, const_cast<int *>(&it), nullptr};
      float x = 0.999f;
      args[2] = &x;
      // warm-up launch, then the timed one
      CHECK(hipLaunchKernel(c.fn, dim3(groups), dim3(kBlock), args, 0, 0));
      CHECK(hipEventRecord(e0, 0));
      CHECK(hipLaunchKernel(c.fn, dim3(groups), dim3(kBlock), args, 0, 0));
      CHECK(hipEventRecord(e1, 0));
      CHECK(hipEventSynchronize(e1));
      float ms = 0.0f;
      CHECK(hipEventElapsedTime(&ms, e0, e1));
      CHECK(hipMemcpy(hc.data(), cyc, sizeof(long long) * waves,
                      hipMemcpyDeviceToHost));
      CHECK(hipMemcpy(hr.data(), rt, sizeof(long long) * waves,
                      hipMemcpyDeviceToHost));
      std::vector<long long> sc(hc.begin(), hc.begin() + waves);
      std::sort(sc.begin(), sc.end());
      const double med = (double)sc[waves / 2];
      double ghz = 0.0;
      for (int w = 0; w < waves; ++w) ghz += (double)hc[w] / (double)hr[w];
      ghz = ghz / waves * 0.1; // realtime ticks at 100 MHz
      double insts = (double)it * kUnroll * c.chains * insts_per_step(c.op);
      if (c.op == OP_MIX_RSQ || c.op == OP_MIX_DPP)
        insts += (double)it * kUnroll / 2; // the side instruction, every 2 steps
      const double cpi = med / insts;     // one wave
      const double simd_cpi = cpi / wps;  // SIMD cycles per wave-instruction
      const double per_cu = 4.0 / simd_cpi;
      // the same from the event time, at the measured clock
      const double ev_per_cu =
          insts * waves / (ms * 1e-3 * ghz * 1e9) / cus;
      std::printf("%-9s waves/SIMD %d: cyc/inst %6.2f  simd_cpi %6.2f  "
                  "per_cu_clk %5.3f  (event: %5.3f)  %.3f GHz  %.3f ms\n",
                  c.name, wps, cpi, simd_cpi, per_cu, ev_per_cu, ghz, ms);
      char buf[320];
      std::snprintf(buf, sizeof buf,
                    "%s{\"case\": \"%s\", \"waves_per_simd\": %d, "
                    "\"cyc_per_inst\": %.3f, \"simd_cpi\": %.3f, "
                    "\"per_cu_clk\": %.4f, \"event_per_cu_clk\": %.4f, "
                    "\"ghz\": %.3f}",
                    first ? "" : ", ", c.name, wps, cpi, simd_cpi, per_cu,
                    ev_per_cu, ghz);
      json += buf;
      first = false;
    }
  }
  json += "]}";
  std::printf("%s\n", json.c_str());
  CHECK(hipFree(val));
  CHECK(hipFree(cyc));
  CHECK(hipFree(rt));
  return 0;
}
