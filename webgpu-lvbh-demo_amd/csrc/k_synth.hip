// k_synth.hip -- device-side synthetic meshes for the serving loop.
//
// The same counter-based generator as the host workloads library
// (include/lbvh_workloads.h, csrc/synth_math.h): SplitMix64 of (seed,
// element, draw) -> f64 in [0,1).  Uniform meshes reproduce lbvh_wl_uniform
// bit for bit (only f64 adds and multiplies, each rounded once), and so
// does the jitter: its Gaussians come from synth_math.h's Box-Muller, whose
// ln, cos and sqrt are plain f64 arithmetic, so lbvh_wl_jitter on the host
// replays a jittered mesh exactly.
#define LBVH_HD __host__ __device__
#include "device_util.h"
#include "synth_math.h"

namespace lbvh {

using lbvh_synth::splitmix64;
using lbvh_synth::u01;

__global__ __launch_bounds__(kMapThreads) void k_synth_uniform(
    uint64_t seed, uint32_t n, double h, float4 *tri) {
  const uint64_t hseed = splitmix64(seed);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    double c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = u01(hseed, i, k);
    float v[9];
#pragma unroll
    for (int j = 0; j < 9; ++j)
      v[j] = (float)__dadd_rn(c[j % 3],
                              __dmul_rn(__dsub_rn(__dmul_rn(2.0, u01(hseed, i, 3 + j)), 1.0), h));
    tri[3 * i] = make_float4(v[0], v[1], v[2], 3.1415f);
    tri[3 * i + 1] = make_float4(v[3], v[4], v[5], 3.1415f);
    tri[3 * i + 2] = make_float4(v[6], v[7], v[8], 3.1415f);
  }
}

__global__ __launch_bounds__(kMapThreads) void k_synth_jitter(
    uint64_t seed, uint32_t n, double sigma, float4 *tri) {
  const uint64_t hseed = splitmix64(seed);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float4 v = tri[3 * i + k];
      float *c = &v.x;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double g = lbvh_synth::gauss(hseed, i, 2 * (3 * k + a));
        c[a] = (float)((double)c[a] + sigma * g);
      }
      tri[3 * i + k] = v;
    }
  }
}

static uint32_t grid_for(uint32_t n) {
  uint32_t b = (n + kMapThreads - 1) / kMapThreads;
  if (b > 4096) b = 4096;
  return b ? b : 1;
}

void launch_synth_uniform(uint64_t seed, uint32_t n, double half_extent,
                          float *tri, hipStream_t s) {
  hipLaunchKernelGGL(k_synth_uniform, dim3(grid_for(n)), dim3(kMapThreads), 0,
                     s, seed, n, half_extent, reinterpret_cast<float4 *>(tri));
}

void launch_synth_jitter(uint64_t seed, uint32_t n, double sigma, float *tri,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_synth_jitter, dim3(grid_for(n)), dim3(kMapThreads), 0,
                     s, seed, n, sigma, reinterpret_cast<float4 *>(tri));
}

}  // namespace lbvh
