/*
 * lbvh.h -- C ABI of the MI355X LBVH builder (liblbvh.so).
 *
 * Drop-in for the scripts/bvh build path of AddisonPrairie/WebGPU-LVBH-demo.
 * Each entry point below names the reference interface it replaces.  All
 * arguments are plain pointers and sizes: `const float *tri` etc. are DEVICE
 * pointers (from lbvh_alloc or any hipMalloc on the context's device) unless
 * the comment says "host".  Work is enqueued on the context's HIP stream and
 * returns immediately; call lbvh_synchronize() (or lbvh_download, which
 * synchronises) before reading results on the host.
 *
 * Buffer layouts are the reference's, so the demo's path tracer reads the
 * outputs unchanged:
 *   triangle   48 B  v0.xyz pad | v1.xyz pad | v2.xyz pad
 *                    (aabb-z-idx.js:153-157; filled at build-bvh.js:14-21)
 *   AABB       32 B  min.xyz pad | max.xyz pad          (aabb-z-idx.js:159-162)
 *   key        u32   30-bit Morton code                  (aabb-z-idx.js:197-208)
 *   index      i32   sorted position -> input triangle  (radix-sort.js:499)
 *   parents    i32[2N]  [0,N-1) internal, [N,2N) leaves (radix-tree.js:146-156)
 *   BVH node   64 B  l_min.xyz l_child | l_max.xyz f_1 |
 *                    r_min.xyz r_child | r_max.xyz f_2   (bvh-up-pass.js:161-178,
 *                    pathtracer.js:441-450).  Root = node 0; child >= 0 is an
 *                    internal node, child < 0 is leaf -(child+1) in O_TRI;
 *                    node N-1 is all zero.
 * Children are written in Karras order (l = the split's left child, f_1 = 2,
 * f_2 = 0); the reference's slot order is thread-arrival order
 * (bvh-up-pass.js:230-252) and the path tracer reads both slots alike
 * (pathtracer.js:526-565).
 *
 * Every function returns an lbvh_status; details via lbvh_last_error().
 * Nothing throws across this boundary.
 *
 * Extents: a device pointer that lies inside a buffer from lbvh_alloc must
 * have room, from that address to the buffer's end, for every byte the
 * call reads or writes through it (e.g. 48*N for triangles, 64*N for BVH
 * nodes); otherwise the call fails with LBVH_ERR_SIZE_MISMATCH before any
 * copy or launch.  WebGPU gives the reference the same guarantee by
 * bounds-checking buffer accesses, and Float32Array.set throws RangeError
 * on an oversize upload (build-bvh.js:14-21).  Pointers from the caller's
 * own hipMalloc are not checked.
 */
#ifndef LBVH_H
#define LBVH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  LBVH_OK = 0,
  LBVH_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad mode, n too large */
  LBVH_ERR_SIZE_MISMATCH = 2,    /* buffer size does not match n */
  LBVH_ERR_OUT_OF_BUDGET = 3,    /* request exceeds the context's HBM budget */
  LBVH_ERR_OUT_OF_MEMORY = 4,    /* the device allocator failed */
  LBVH_ERR_DEVICE = 5,           /* a HIP runtime call failed */
  LBVH_ERR_NO_DEVICE = 6,        /* no such device ordinal */
  LBVH_ERR_BUSY = 7,             /* serving queue full / no free slot */
  LBVH_ERR_NOT_FOUND = 8,        /* unknown handle / request id */
  LBVH_ERR_CANCELLED = 9,        /* serving: cancelled before it ran */
  LBVH_ERR_TIMEOUT = 10          /* serving: not admitted within its wait */
} lbvh_status;

/* Morton division mode for rel = (cen - min) / (max - min)
 * (aabb-z-idx.js:192).  The reference's result depends on how the WebGPU
 * stack lowers fp32 division; both lowerings are provided. */
typedef enum {
  LBVH_DIV_MUL_RCP = 0, /* d * fl(1/e)  (fdiv lowered to fmul + frcp) */
  LBVH_DIV_IEEE = 1     /* correctly rounded d / e */
} lbvh_div_mode;

/* Largest triangle count of one build: keys carry 30 bits, indices are i32
 * and the sort's look-back counts are 32-bit. */
#define LBVH_MAX_TRIS (1u << 30)

typedef struct lbvh_ctx lbvh_ctx;

/* ---- context ----------------------------------------------------------- */

/* Replaces initBVHBuild(device) (build-bvh.js:1-10) and the per-stage
 * init*Kernel(device) calls: one context owns a HIP stream, the scratch
 * arena the stages share, and an HBM budget (bytes; 0 = no limit) that
 * caps lbvh_alloc plus scratch. */
int lbvh_create(int device_ordinal, uint64_t hbm_budget_bytes,
                lbvh_ctx **out_ctx);
int lbvh_destroy(lbvh_ctx *ctx);
const char *lbvh_last_error(const lbvh_ctx *ctx);
const char *lbvh_status_string(int status);
int lbvh_set_div_mode(lbvh_ctx *ctx, int div_mode);
int lbvh_get_div_mode(const lbvh_ctx *ctx);
/* Launch-shape tuning (no counterpart in the reference, whose workgroup and
 * digit sizes are constants, e.g. WG_SIZE = 64 at aabb-z-idx.js:4).  Values
 * are validated; results do not depend on them. */
typedef enum {
  LBVH_TUNE_TREE_TILE = 0, /* leaves per block of the fused up pass: 256|512|1024 */
  LBVH_TUNE_SORT_BITS = 1  /* digit width of the builds' sort: 8 (4 passes
                              over the 30-bit Morton keys) or 10 (3 passes);
                              the reference uses 2-bit digits,
                              radix-sort.js:486 */
} lbvh_tunable;
/* set_tuning fixes a value for every size (value 0: no fixed value, builds
 * use the autotuned table, else the size-based default: tree tile 1024, and
 * 10-bit digits up to 2^17 triangles and from 2^20 + 1 to 2^21, 8-bit
 * elsewhere); get_tuning_for returns
 * the value a build of n triangles would use (n = 0: the fixed/default
 * value). */
int lbvh_set_tuning(lbvh_ctx *ctx, int param, int value);
int lbvh_get_tuning(const lbvh_ctx *ctx, int param);
int lbvh_get_tuning_for(const lbvh_ctx *ctx, int param, uint32_t n);
/* Autotuner: for every power-of-two size bucket 2^k in [n_min, n_max] it
 * builds a device-synthesised uniform mesh of 2^k triangles with each
 * variant of every tunable, times them with HIP events (median of `reps`
 * after a warm-up) and keeps the fastest per bucket if it beats the
 * size-based default by more than 2 %, else the default; later builds pick the
 * entry of their size's bucket (or the nearest tuned one).  A fixed value
 * from lbvh_set_tuning is cleared.  ms_out (optional) receives, per bucket
 * and variant in order, the measured milliseconds: buckets x 6 variants,
 * tree tile major (256, 512, 1024) and digit width minor (8, 10); NaN for a
 * variant that did not run.  Synchronous. */
int lbvh_autotune(lbvh_ctx *ctx, uint32_t n_min, uint32_t n_max, int reps,
                  float *ms_out);
/* The context's hipStream_t, for callers that record events or chain work. */
void *lbvh_stream(lbvh_ctx *ctx);
int lbvh_synchronize(lbvh_ctx *ctx);
/* Bytes currently charged to the budget (allocations + scratch). */
uint64_t lbvh_bytes_in_use(const lbvh_ctx *ctx);
/* Releases the scratch arena (it regrows on the next call that needs it). */
int lbvh_trim(lbvh_ctx *ctx);

/* ---- device buffers (GPUDevice.createBuffer / GPUBuffer.destroy) -------- */

/* Replaces device.createBuffer({size}) (e.g. build-bvh.js:14-18): device
 * memory charged to the context's budget. */
int lbvh_alloc(lbvh_ctx *ctx, uint64_t bytes, void **out_dptr);
/* Replaces GPUBuffer.destroy() (build-bvh.js:59-63). */
int lbvh_free(lbvh_ctx *ctx, void *dptr);
/* Host -> device copy (the mappedAtCreation upload, build-bvh.js:20-21).
 * Synchronous with respect to the host buffer.  Copies of 64 MB or more go
 * through the context's pinned staging buffers (128 MB, allocated on the
 * first such copy, freed by lbvh_destroy), filled by four host threads. */
int lbvh_upload(lbvh_ctx *ctx, void *dst_dev, const void *src_host,
                uint64_t bytes);
/* Device -> host copy after all queued work (the reference never reads
 * back; tests and callers that want the bytes do).  Synchronises.  Large
 * copies are staged as in lbvh_upload. */
int lbvh_download(lbvh_ctx *ctx, void *dst_host, const void *src_dev,
                  uint64_t bytes);

/* ---- per-stage entry points (the execute() shims) ----------------------- */

/* Scene bounds: fp32 min/max over every vertex coordinate, written to
 * bounds_out[0..5] (device) as min.xyz, max.xyz.  The reference takes
 * bounds from the caller (parse-obj.js:10-53 -> aabb-z-idx.js:113-124);
 * lbvh_build uses this reduction when the caller passes none.  An empty
 * mesh has no bounds: n = 0 fails with LBVH_ERR_INVALID_ARGUMENT. */
int lbvh_bounds(lbvh_ctx *ctx, const float *tri, uint32_t n,
                float *bounds_out);

/* Replaces initAABB_ZidxKernel(device).execute(TRI, N, bounds)
 * (aabb-z-idx.js:57-148, WGSL :177-208): per-triangle AABB (32 B) and
 * 30-bit Morton key of the AABB centre.  bounds_host = 6 host floats
 * (min.xyz, max.xyz) or NULL to use the GPU bounds reduction.  Caller bounds
 * must be finite (else LBVH_ERR_INVALID_ARGUMENT; the reference would
 * quantise against them into implementation-defined keys).  A flat axis
 * (min == max, a planar mesh, from the caller or the reduction) is
 * deterministic: 0/0 converts to coordinate 0, and centres off the plane
 * saturate (0 below, 1023's bits above). */
int lbvh_aabb_morton(lbvh_ctx *ctx, const float *tri, uint32_t n,
                     const float *bounds_host, float *aabb_out,
                     uint32_t *keys_out);

/* Replaces initRadixSortKernel(device).execute(keyBuf, N)
 * (radix-sort.js:161-500): stable LSD sort of u32 keys over all 32 bits, as
 * the reference's 16 passes of 2 bits sort them: 4 passes of 8 bits, or of
 * 10 bits (10, 10, 10, 2) when LBVH_TUNE_SORT_BITS is fixed at 10 (the
 * autotuned table is for the builds' 30-bit keys).  Destructive like the
 * sorted keys replace `keys`; idx_out[j] is the input position of the key
 * now at j. */
int lbvh_sort(lbvh_ctx *ctx, uint32_t *keys, int32_t *idx_out, uint32_t n);

/* Replaces initRadixTreeKernel(device).execute(sortedKeyBuf, N)
 * (radix-tree.js:50-120, WGSL :137-238): Karras radix tree over the sorted
 * keys, ties broken by sorted position; writes the 2N-entry parents array
 * (slots 0 and N-1 are 0). */
int lbvh_radix_tree(lbvh_ctx *ctx, const uint32_t *sorted_keys, uint32_t n,
                    int32_t *parents_out);

/* Replaces initBVHUpPassKernel(device).execute(IDX, AABB, PARENT, N)
 * (bvh-up-pass.js:70-156, WGSL :200-280): bottom-up node boxes; writes
 * N 64-byte nodes.  Scratch: about 7 bytes per triangle.  A parents array
 * that is not a radix tree makes the next lbvh_synchronize return
 * LBVH_ERR_DEVICE. */
int lbvh_up_pass(lbvh_ctx *ctx, const int32_t *idx, const float *aabb,
                 const int32_t *parents, uint32_t n, void *bvh_out);

/* Replaces initRearrangeKernel(device).execute(I_TRI, IDX, N)
 * (rearrange.js:57-139, WGSL :162-169): tri_out[i] = tri_in[idx[i]]. */
int lbvh_rearrange(lbvh_ctx *ctx, const float *tri_in, const int32_t *idx,
                   uint32_t n, float *tri_out);

/* ---- fused build and refit --------------------------------------------- */

/* Replaces build(TRI_ARRAY, NUM_TRIS, MODEL_BOUNDS) (build-bvh.js:12-66)
 * once the triangles are on the device: bvh_out (64*N bytes) and
 * tri_out (48*N bytes) receive BVH_BUFFER and O_TRIANGLE_BUFFER.
 * bounds_host = 6 host floats (MODEL_BOUNDS rounded to fp32) or NULL for
 * the GPU bounds reduction.  The reference's intermediates (AABB, keys,
 * index, parents; freed at build-bvh.js:59-63) live in the context's
 * scratch; pass sorted_keys_out / perm_out (4*N bytes each, or NULL) to keep
 * the sorted keys and permutation a later lbvh_refit needs. */
int lbvh_build(lbvh_ctx *ctx, const float *tri, uint32_t n,
               const float *bounds_host, void *bvh_out, float *tri_out,
               uint32_t *sorted_keys_out, int32_t *perm_out);

/* lbvh_build with no allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as lbvh_build, for at most 1,048,576 triangles (larger n fails
 * with LBVH_ERR_INVALID_ARGUMENT).  Its scratch is one static 64 MiB array
 * per device in the library's code object, held from the library's load on
 * that device until the process ends, not charged to the context's budget;
 * the call enqueues kernels only (no hipMalloc, hipFree, copy or memset, no
 * profiling events).  Every call on one device shares that array: calls
 * that overlap on a device (two contexts or threads, streams not ordered)
 * race and give undefined results; calls ordered on one stream are safe.
 * Nothing in the array needs to start at zero. */
int lbvh_build_static(lbvh_ctx *ctx, const float *tri, uint32_t n,
                      const float *bounds_host, void *bvh_out, float *tri_out,
                      uint32_t *sorted_keys_out, int32_t *perm_out);

/* Refit: no counterpart in the reference, which rebuilds from scratch; it
 * re-runs the up pass (bvh-up-pass.js:200-280) and rearrange
 * (rearrange.js:162-169) over new vertex positions `tri` (input order,
 * same N) with the topology of an earlier lbvh_build kept fixed through its
 * sorted_keys and perm. */
int lbvh_refit(lbvh_ctx *ctx, const float *tri, uint32_t n,
               const uint32_t *sorted_keys, const int32_t *perm,
               void *bvh_out, float *tri_out);

/* ---- batched build ------------------------------------------------------- */

/* Builds num_meshes independent LBVHs in one captured graph (no counterpart
 * in the reference, which builds one mesh per call to build(),
 * build-bvh.js:12-66).  The meshes' triangles are concatenated in `tri`;
 * offsets (host, num_meshes + 1 entries, offsets[0] = 0, non-decreasing)
 * gives mesh k's triangles [offsets[k], offsets[k+1]).  Every output keeps
 * that layout: mesh k's BVH_BUFFER is nodes [offsets[k], offsets[k+1]) of
 * bvh_out with node ids and leaf ids relative to the mesh (its root is its
 * first node), its O_TRIANGLE_BUFFER the same range of tri_out, and
 * sorted_keys_out / perm_out (optional) hold per-mesh sorted keys and
 * mesh-relative permutations.  Each mesh's result is byte-identical to
 * lbvh_build on that mesh alone.  bounds_host: 6 host floats per mesh, or
 * NULL for per-mesh GPU bounds.  Repeated calls with the same offsets and
 * buffers replay the captured graph. */
int lbvh_build_batched(lbvh_ctx *ctx, const float *tri,
                       const uint32_t *offsets, uint32_t num_meshes,
                       const float *bounds_host, void *bvh_out, float *tri_out,
                       uint32_t *sorted_keys_out, int32_t *perm_out);

/* ---- serving ------------------------------------------------------------- */

/* A device-resident build/refit service (no counterpart in the reference,
 * which builds once per dropped file, main.js:88).  Requests are admitted as
 * execution slots (each with its own HIP stream and scratch) and HBM budget
 * free up, so work from different requests overlaps on the GPU; requests on
 * the same mesh run in submission order.  Built trees stay resident (the
 * triangles, BVH_BUFFER, O_TRIANGLE_BUFFER, sorted keys and permutation,
 * 168 bytes per triangle) until evicted: least recently used idle meshes
 * are evicted when a new build needs room.  A request that can never fit
 * the budget completes with LBVH_ERR_OUT_OF_BUDGET; a refit of a mesh that
 * is not resident, or with a different triangle count, completes with
 * LBVH_ERR_NOT_FOUND / LBVH_ERR_SIZE_MISMATCH.  A request whose tree
 * kernels flag inconsistent input (lbvh_synchronize's LBVH_ERR_DEVICE)
 * completes with LBVH_ERR_DEVICE and its mesh stops being resident for
 * refits.  Resident results are byte-identical to lbvh_build / lbvh_refit
 * on the same input.
 * The budget bounds all device memory the server holds: its mesh arena
 * (device chunks of up to 4 GiB, allocated as the resident set grows and
 * carved into mesh ranges) plus every slot's scratch; bytes_in_use in the
 * stats reports that sum, resident_bytes the part holding meshes. */
typedef struct lbvh_server lbvh_server;

typedef enum {
  LBVH_REQ_BUILD = 0, /* (re)build mesh_id from new triangles */
  LBVH_REQ_REFIT = 1  /* refit resident mesh_id to new vertex positions */
} lbvh_request_kind;

/* Where a request's triangles come from. */
typedef enum {
  LBVH_SRC_HOST = 0,    /* `tri` points to 12*n host floats (copied at submit) */
  LBVH_SRC_DEVICE = 1,  /* `tri` is a device pointer, read when the request runs */
  LBVH_SRC_UNIFORM = 2, /* synthesised on the device: the uniform mesh of
                           lbvh_wl_uniform(seed, n, param) (bit-identical) */
  LBVH_SRC_JITTER = 3   /* the resident mesh's triangles + N(0, param^2) per
                           coordinate, drawn on the device from seed (refits):
                           bit-identical to lbvh_wl_jitter(seed, n, param) of
                           the triangles in input order */
} lbvh_source;

typedef struct {
  uint64_t request_id;
  uint64_t mesh_id;
  int kind;
  int status;        /* lbvh_status of the request */
  uint32_t n;
  uint32_t slot;
  double submit_s;   /* host steady-clock seconds: submitted (arrival) */
  double start_s;    /* admitted to a slot */
  double done_s;     /* device work observed complete */
  /* With lbvh_server_set_digest on, digests of the request's results as
   * they were when its device work ended, else 0: the sum over the u32
   * words w_j of BVH_BUFFER (16 per triangle, j from 0) of
   * w_j * (splitmix64(j) | 1) modulo 2^64, and the same over
   * O_TRIANGLE_BUFFER (12 words per triangle).  A checker can compare
   * every request's results with a host replay even after the mesh has
   * been refitted again or evicted. */
  uint64_t digest_bvh;
  uint64_t digest_tri;
} lbvh_completion;

typedef struct {
  uint64_t submitted, completed, failed, evictions;
  uint64_t resident_meshes, resident_bytes, bytes_in_use, peak_bytes;
  uint32_t queued, in_flight;
} lbvh_server_stats;

int lbvh_server_create(int device_ordinal, uint64_t hbm_budget_bytes,
                       uint32_t slots, lbvh_server **out);
/* Waits for in-flight work, then frees every resident mesh and slot. */
int lbvh_server_destroy(lbvh_server *srv);
int lbvh_server_set_div_mode(lbvh_server *srv, int div_mode);
/* Requests admitted from now on also digest their results (one extra pass
 * over the mesh's BVH and O_TRI, about 112 bytes per triangle); no
 * counterpart in the reference: for checkers of the serving loop. */
int lbvh_server_set_digest(lbvh_server *srv, int enable);
int lbvh_server_submit(lbvh_server *srv, int kind, uint64_t mesh_id,
                       int source, const float *tri, uint32_t n,
                       uint64_t seed, double param, uint64_t *out_request_id);
/* Cancels a request that is still waiting for admission: it completes with
 * LBVH_ERR_CANCELLED and nothing of it runs.  LBVH_ERR_BUSY if it is already
 * admitted (its device work runs to the end), LBVH_ERR_NOT_FOUND if it has
 * completed or never existed.  No counterpart in the reference. */
int lbvh_server_cancel(lbvh_server *srv, uint64_t request_id);
/* Requests submitted from now on that are not admitted within max_wait_s
 * seconds of their submission complete with LBVH_ERR_TIMEOUT without
 * running (0: wait as long as it takes, the default).  Checked as the
 * dispatcher scans the queue's head, within about a millisecond of the
 * deadline.  No counterpart in the reference. */
int lbvh_server_set_max_wait(lbvh_server *srv, double max_wait_s);
/* Up to `max` completions not yet returned (oldest first). */
int lbvh_server_poll(lbvh_server *srv, lbvh_completion *out, uint32_t max,
                     uint32_t *out_count);
/* Blocks until nothing is queued or in flight. */
int lbvh_server_drain(lbvh_server *srv);
/* Device buffers of a resident mesh (valid until it is rebuilt or evicted;
 * read them only when no request on the mesh is in flight). */
int lbvh_server_mesh(lbvh_server *srv, uint64_t mesh_id, void **bvh,
                     float **tri_out, uint32_t *n);
int lbvh_server_evict(lbvh_server *srv, uint64_t mesh_id);
int lbvh_server_stats_get(lbvh_server *srv, lbvh_server_stats *out);
const char *lbvh_server_last_error(lbvh_server *srv);

/* ---- consumer check ------------------------------------------------------ */

/* Closest-hit ray casting through a built BVH the way the demo's path tracer
 * traverses it (pathtracer.js:495-584: stack traversal from node 0, nearer
 * child first; Moller-Trumbore at :587-619; slab test at :621-628), on the
 * device.  rays: count x 8 floats (origin.xyz, pad, dir.xyz, pad); hits:
 * count x 4 floats (normal.xyz facing the ray, dist; dist = -1 for a miss),
 * the path tracer's RayHit.  bvh / tri are lbvh_build's outputs for n
 * triangles.  The stack holds 64 entries (the path tracer's 32 can overflow
 * on deeper trees). */
int lbvh_raycast(lbvh_ctx *ctx, const void *bvh, const float *tri, uint32_t n,
                 const float *rays, uint32_t count, float *hits);

/* ---- profiling ------------------------------------------------------------
 * No counterpart in the reference (which times build() with Date.now() at
 * main.js:87-90).  When enabled, lbvh_build and lbvh_refit record a HIP event
 * pair around each stage on the context's stream; lbvh_profile_read waits
 * for them and returns the accumulated device time per stage. */
typedef enum {
  LBVH_STAGE_BOUNDS = 0,
  LBVH_STAGE_MORTON = 1,   /* leaf AABB + Morton key (+ per-tile digit counts) */
  LBVH_STAGE_SORT0 = 2,    /* radix sort digit passes 0..3 */
  LBVH_STAGE_SORT1 = 3,
  LBVH_STAGE_SORT2 = 4,
  LBVH_STAGE_SORT3 = 5,
  LBVH_STAGE_TREE = 6,     /* leaf tiles: O_TRI gather, leaf boxes, the radix
                              tree and node boxes inside each tile */
  LBVH_STAGE_TREE_TOP = 7, /* the levels above the leaf tiles */
  LBVH_STAGE_COUNT = 8
} lbvh_stage;

typedef struct {
  double ms[LBVH_STAGE_COUNT];        /* summed device milliseconds */
  uint64_t launches[LBVH_STAGE_COUNT];
} lbvh_profile;

int lbvh_profile_enable(lbvh_ctx *ctx, int enable);
int lbvh_profile_read(lbvh_ctx *ctx, lbvh_profile *out, int reset);

#ifdef __cplusplus
}
#endif

#endif /* LBVH_H */
