// k_raycast.hip -- closest-hit ray casting over the built BVH, following the
// demo path tracer's traversal, as a consumer check of the output layout.
//
// Reference: scripts/pathtracer/pathtracer.js intersect_bvh (:495-584),
// tri_intersect (Moller-Trumbore, :587-619), aabb_intersect (:621-628).
// Root = node 0; child < 0 is leaf -(child+1) in O_TRI; the nearer valid
// child is visited first and the other pushed.  Rays are 8 floats
// (origin.xyz, pad, dir.xyz, pad); hits are the path tracer's RayHit
// (normal.xyz, dist), dist = -1 for a miss.  Every operation is explicitly
// rounded so the CPU restatement in the oracle reproduces hits bit for bit.
#include "device_util.h"

namespace lbvh {

struct V3 {
  float x, y, z;
};

__device__ __forceinline__ V3 sub(V3 a, V3 b) {
  return {__fsub_rn(a.x, b.x), __fsub_rn(a.y, b.y), __fsub_rn(a.z, b.z)};
}
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {__fsub_rn(__fmul_rn(a.y, b.z), __fmul_rn(a.z, b.y)),
          __fsub_rn(__fmul_rn(a.z, b.x), __fmul_rn(a.x, b.z)),
          __fsub_rn(__fmul_rn(a.x, b.y), __fmul_rn(a.y, b.x))};
}
__device__ __forceinline__ float dot(V3 a, V3 b) {
  return __fadd_rn(__fadd_rn(__fmul_rn(a.x, b.x), __fmul_rn(a.y, b.y)),
                   __fmul_rn(a.z, b.z));
}
__device__ __forceinline__ V3 v3(int4 v) {
  return {__int_as_float(v.x), __int_as_float(v.y), __int_as_float(v.z)};
}

// tri_intersect (pathtracer.js:587-619): (normal, t) or t = -1.
__device__ __forceinline__ float4 tri_intersect(V3 o, V3 d, V3 v0, V3 v1,
                                                V3 v2) {
  const V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
  const V3 pvec = cross(d, e2);
  const float det = dot(e1, pvec);
  if (fabsf(det) < 1e-10f) return make_float4(-1.f, -1.f, -1.f, -1.f);
  const float idet = __fdiv_rn(1.0f, det);
  const V3 tvec = sub(o, v0);
  const float u = __fmul_rn(dot(tvec, pvec), idet);
  if (u < 0.f || u > 1.f) return make_float4(-1.f, -1.f, -1.f, -1.f);
  const V3 qvec = cross(tvec, e1);
  const float v = __fmul_rn(dot(d, qvec), idet);
  if (v < 0.f || __fadd_rn(u, v) > 1.f) return make_float4(-1.f, -1.f, -1.f, -1.f);
  const V3 nrm = cross(e1, e2);
  // __builtin_sqrtf is correctly rounded (the default fp32 mode); HIP's
  // __fsqrt_rn maps to the hardware approximation, which is not.
  const float len = __builtin_sqrtf(dot(nrm, nrm));
  return make_float4(__fdiv_rn(nrm.x, len), __fdiv_rn(nrm.y, len),
                     __fdiv_rn(nrm.z, len), __fmul_rn(dot(e2, qvec), idet));
}

// aabb_intersect (pathtracer.js:621-628): entry distance or -1e30.
__device__ __forceinline__ float aabb_intersect(V3 lo, V3 hi, V3 o, V3 inv) {
  const float fx = __fmul_rn(__fsub_rn(hi.x, o.x), inv.x);
  const float fy = __fmul_rn(__fsub_rn(hi.y, o.y), inv.y);
  const float fz = __fmul_rn(__fsub_rn(hi.z, o.z), inv.z);
  const float nx = __fmul_rn(__fsub_rn(lo.x, o.x), inv.x);
  const float ny = __fmul_rn(__fsub_rn(lo.y, o.y), inv.y);
  const float nz = __fmul_rn(__fsub_rn(lo.z, o.z), inv.z);
  const float t0 = fmaxf(fminf(fx, nx), fmaxf(fminf(fy, ny), fminf(fz, nz)));
  const float t1 = fminf(fmaxf(fx, nx), fminf(fmaxf(fy, ny), fmaxf(fz, nz)));
  return (t1 >= t0) ? ((t1 < 0.f) ? -1e30f : t0) : -1e30f;
}

constexpr int kRayStack = 64;  // the path tracer's is 32 (pathtracer.js:493)

__global__ __launch_bounds__(kMapThreads) void k_raycast(
    const int4 *__restrict__ nodes, const int4 *__restrict__ tri, uint32_t n,
    const float4 *__restrict__ rays, uint32_t count, float4 *__restrict__ hits,
    uint32_t max_steps) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const float4 ro = rays[2 * (size_t)q], rd = rays[2 * (size_t)q + 1];
  const V3 o = {ro.x, ro.y, ro.z};
  V3 d = {rd.x, rd.y, rd.z};
  // d += (abs(d) < 1e-5) * 1e-5  (pathtracer.js:502-503)
  if (fabsf(d.x) < 1e-5f) d.x = __fadd_rn(d.x, 1e-5f);
  if (fabsf(d.y) < 1e-5f) d.y = __fadd_rn(d.y, 1e-5f);
  if (fabsf(d.z) < 1e-5f) d.z = __fadd_rn(d.z, 1e-5f);
  const V3 inv = {__fdiv_rn(1.0f, d.x), __fdiv_rn(1.0f, d.y),
                  __fdiv_rn(1.0f, d.z)};
  float dist = 1e30f;
  V3 nrm = {0.f, 0.f, 0.f};
  int stack[kRayStack];
  int sp = 0;
  int node = 0;
  for (uint32_t step = 0; sp >= 0 && step < max_steps; ++step) {
    if (node < 0) {
      const uint32_t t = (uint32_t)(-(node + 1));
      if (t < n) {
        const float4 h = tri_intersect(o, d, v3(tri[3 * (size_t)t]),
                                       v3(tri[3 * (size_t)t + 1]),
                                       v3(tri[3 * (size_t)t + 2]));
        if (h.w > 0.f && h.w < dist) {
          nrm = {h.x, h.y, h.z};
          dist = fminf(h.w, dist);
        }
      }
      sp -= 1;
      node = sp >= 0 ? stack[sp] : 0;
    } else {
      if ((uint32_t)node >= n) break;
      const int4 *nd = nodes + 4 * (size_t)node;
      const int4 a = nd[0], b = nd[1], c = nd[2], e = nd[3];
      const float ld = aabb_intersect(v3(a), v3(b), o, inv);
      const float rdist = aabb_intersect(v3(c), v3(e), o, inv);
      const bool lv = ld != -1e30f && ld < dist;
      const bool rv = rdist != -1e30f && rdist < dist;
      if (lv && rv) {
        const int near = ld < rdist ? a.w : c.w;
        const int far = ld < rdist ? c.w : a.w;
        if (sp < kRayStack) stack[sp] = far;
        sp = sp + 1 < kRayStack ? sp + 1 : kRayStack;
        node = near;
      } else if (lv) {
        node = a.w;
      } else if (rv) {
        node = c.w;
      } else {
        sp -= 1;
        node = sp >= 0 ? stack[sp] : 0;
      }
    }
  }
  // RayHit: normal facing the ray, dist -1 for a miss (pathtracer.js:572-583)
  const float s = dot(d, {-nrm.x, -nrm.y, -nrm.z}) > 0.f ? 1.f : -1.f;
  hits[q] = make_float4(s * nrm.x, s * nrm.y, s * nrm.z,
                        dist == 1e30f ? -1.f : dist);
}

void launch_raycast(const void *nodes, const float *tri, uint32_t n,
                    const float *rays, uint32_t count, float *hits,
                    hipStream_t s) {
  const uint32_t blocks = (count + kMapThreads - 1) / kMapThreads;
  const uint64_t cap = 4ull * n + 256;
  hipLaunchKernelGGL(k_raycast, dim3(blocks ? blocks : 1), dim3(kMapThreads),
                     0, s, static_cast<const int4 *>(nodes),
                     reinterpret_cast<const int4 *>(tri), n,
                     reinterpret_cast<const float4 *>(rays), count,
                     reinterpret_cast<float4 *>(hits),
                     (uint32_t)(cap > 0xFFFFFFFFull ? 0xFFFFFFFFull : cap));
}

}  // namespace lbvh
