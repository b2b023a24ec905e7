/*
 * lbvh_oracle.h -- CPU restatement of the reference LBVH build pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle that the HIP
 * product path is compared against.  Only tests/, the smoke check and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * Each function restates one stage of the reference's WGSL/JS pipeline
 * (AddisonPrairie/WebGPU-LVBH-demo, scripts/bvh/) and cites the file:line it
 * follows.  Buffer layouts are the reference's:
 *   triangle  48 B : v0.xyz,pad  v1.xyz,pad  v2.xyz,pad   (aabb-z-idx.js:153-157)
 *   AABB      32 B : min.xyz,pad max.xyz,pad               (aabb-z-idx.js:159-162)
 *   BVH node  64 B : l_min.xyz l_child l_max.xyz f_1
 *                    r_min.xyz r_child r_max.xyz f_2       (bvh-up-pass.js:161-178)
 *   parents   i32[2N]: [0,N-1) internal parents, [N,2N) leaf parents
 *                                                          (radix-tree.js:146-156)
 */
#ifndef LBVH_ORACLE_H
#define LBVH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Morton division modes: how `(cen - min) / (max - min)` is evaluated. */
#define ORACLE_DIV_MUL_RCP 0 /* d * fl(1/e): fdiv lowered to fmul(frcp) */
#define ORACLE_DIV_IEEE 1    /* correctly rounded d / e */

uint32_t oracle_split_3(uint32_t u);
uint32_t oracle_morton_code(uint32_t x, uint32_t y, uint32_t z);
uint32_t oracle_f32_to_u32(float v);
int32_t oracle_dif(uint32_t a, uint32_t b, int faithful);

void oracle_bounds(const float *tri, uint32_t n, float out_min_max[6]);
void oracle_aabb_z_idx(const float *tri, uint32_t n, const float bounds[6],
                       int div_mode, float *aabb_out, uint32_t *keys_out);
void oracle_radix_sort(uint32_t *keys, int32_t *idx_out, uint32_t n);
void oracle_radix_tree(const uint32_t *sorted_keys, uint32_t n,
                       int32_t *parents_out, int faithful);
void oracle_up_pass(const int32_t *idx, const float *aabb,
                    const int32_t *parents, uint32_t n, void *bvh_out);
void oracle_rearrange(const float *tri, const int32_t *idx, uint32_t n,
                      float *tri_out);
void oracle_canonicalize_bvh(void *bvh, uint32_t n);
void oracle_build(const float *tri, uint32_t n, const float *bounds_or_null,
                  int div_mode, void *bvh_out, float *tri_out,
                  uint32_t *sorted_keys_out, int32_t *perm_out);

/* The same build, multi-threaded (lbvh_cpu_mt.c, OpenMP): byte-identical
 * outputs; returns the number of threads it used.  The CPU baseline. */
int oracle_build_mt(const float *tri, uint32_t n, const float *bounds_or_null,
                    int div_mode, void *bvh_out, float *tri_out,
                    uint32_t *sorted_keys_out, int32_t *perm_out);
void oracle_refit(const float *tri, uint32_t n, const uint32_t *sorted_keys,
                  const int32_t *perm, void *bvh_out, float *tri_out);
/* The refit, multi-threaded (lbvh_cpu_mt.c): byte-identical outputs. */
void oracle_refit_mt(const float *tri, uint32_t n, const uint32_t *sorted_keys,
                     const int32_t *perm, void *bvh_out, float *tri_out);
/* Digests of a build's BVH and O_TRI (the serving loop's per-request
 * digests): out[0] = sum_j w_j * (splitmix64(j) | 1) mod 2^64 over the BVH's
 * u32 words, out[1] the same over O_TRI's (lbvh_cpu_mt.c). */
void oracle_digest(const void *bvh, const float *tri_out, uint32_t n,
                   uint64_t out[2]);

/* Closest-hit ray casting as the demo's path tracer does it
 * (pathtracer.js:495-628).  rays: 8 floats each (origin.xyz, pad, dir.xyz,
 * pad); hits: 4 floats each (normal.xyz, dist; dist = -1 for a miss). */
void oracle_raycast(const void *bvh, const float *tri, uint32_t n,
                    const float *rays, uint32_t count, float *hits);
/* Same rays against every triangle, no BVH. */
void oracle_raycast_bruteforce(const float *tri, uint32_t n, const float *rays,
                               uint32_t count, float *hits);

/* Independent brute-force references (no shared code with the restatement). */
void oracle_bruteforce_tree(const uint32_t *sorted_keys, uint32_t n,
                            int32_t *parents_out);
void oracle_bruteforce_bvh(const float *tri, const int32_t *perm,
                           const uint32_t *sorted_keys, uint32_t n,
                           void *bvh_out);

#ifdef __cplusplus
}
#endif

#endif
