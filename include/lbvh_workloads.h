/*
 * lbvh_workloads.h -- seeded synthetic triangle meshes (liblbvh_workloads.so).
 *
 * No counterpart in the reference, which only loads OBJ files
 * (scripts/obj/parse-obj.js).  These generators produce the benchmark and
 * test meshes in the reference's TRI_ARR layout (parse-obj.js:55-59):
 * 12 floats per triangle, v0.xyz 3.1415 | v1.xyz 3.1415 | v2.xyz 3.1415.
 * All randomness is counter-based (SplitMix64 of seed and element index),
 * so results are independent of thread count and reproducible anywhere.
 */
#ifndef LBVH_WORKLOADS_H
#define LBVH_WORKLOADS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Uniform meshes: centroid c ~ U[0,1]^3, each vertex = c + U[-h,h]^3 with
 * h = half_extent (0.01 for the 65,536-triangle parity mesh), computed in
 * f64 and rounded once to fp32. */
void lbvh_wl_uniform(uint64_t seed, uint32_t n, double half_extent,
                     float *tri);

/* Clustered meshes: centroids from a 64-component Gaussian mixture (means
 * U[0,1]^3, sigma 0.02, clamped to [0,1]^3), offsets U[-0.005,0.005]^3;
 * a dup_fraction share of triangles copy all three vertices of an earlier
 * triangle exactly (duplicate Morton keys). */
void lbvh_wl_clustered(uint64_t seed, uint32_t n, double dup_fraction,
                       float *tri);

/* Heightfield: nx*ny grid cells over [0,1] x [0,0.5], 2 triangles per cell
 * (row-major cells, cell-major triangles), height z = 0.1 * 4-octave value
 * noise (base lattice 8 cells per unit, lacunarity 2, gain 0.5).
 * Writes 2*nx*ny triangles. */
void lbvh_wl_heightfield(uint64_t seed, uint32_t nx, uint32_t ny, float *tri);

/* count sizes, log-uniform in [lo, hi]: floor(exp(U[ln lo, ln hi])). */
void lbvh_wl_log_uniform_sizes(uint64_t seed, uint32_t count, uint32_t lo,
                               uint32_t hi, uint32_t *sizes);

/* out = in + N(0, sigma^2) on every vertex coordinate (pads copied). */
void lbvh_wl_jitter(uint64_t seed, uint32_t n, double sigma, const float *in,
                    float *out);

/* Caller-side scene bounds as the reference's OBJ loader computes them
 * (parse-obj.js:10-53): f64 min/max over all vertices, rounded to fp32.
 * out = min.xyz, max.xyz. */
void lbvh_wl_bounds(const float *tri, uint32_t n, float *out);

/* Uniform doubles in [0,1) of the counter-based generator, for callers that
 * need more draws from the same stream (request streams of the server). */
void lbvh_wl_uniform01(uint64_t seed, uint64_t first, uint32_t count,
                       double *out);

#ifdef __cplusplus
}
#endif

#endif
