/*
 * lbvh_oracle.c -- CPU restatement of the reference LBVH build pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle, never the product).  See the header
 * for the buffer layouts.  Compiled with -ffp-contract=off so every fp32
 * operation rounds exactly once, as in the WGSL source.
 *
 * Reference files (AddisonPrairie/WebGPU-LVBH-demo):
 *   scripts/bvh/build-bvh.js              pipeline order, which buffers survive
 *   scripts/bvh/kernels/aabb-z-idx.js     leaf AABB + 30-bit Morton key
 *   scripts/bvh/kernels/radix-sort.js     stable LSD radix sort (2-bit digits)
 *   scripts/bvh/kernels/radix-tree.js     Karras radix tree -> parents array
 *   scripts/bvh/kernels/bvh-up-pass.js    bottom-up node boxes
 *   scripts/bvh/kernels/rearrange.js      triangle gather into leaf order
 */
#include "lbvh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* WGSL min()/max() on f32 (aabb-z-idx.js:186-187, bvh-up-pass.js:270-271). */
static float fmin32(float a, float b) { return (b < a) ? b : a; }
static float fmax32(float a, float b) { return (b > a) ? b : a; }

/* ------------------------------------------------------------------------ */
/* Morton code: aabb-z-idx.js:197-208                                        */
/* ------------------------------------------------------------------------ */

/* split_3 (aabb-z-idx.js:202-208): spreads the bits of u so that bit k lands
 * at 3k.  The masks keep only the low 10 bits of u plus bits 24..25, which
 * fold onto bits 8..9 (the reference applies it to the full u32). */
uint32_t oracle_split_3(uint32_t u) {
  uint32_t x = u;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

/* morton_code (aabb-z-idx.js:197-199): x at bits 3k, y at 3k+1, z at 3k+2. */
uint32_t oracle_morton_code(uint32_t x, uint32_t y, uint32_t z) {
  return oracle_split_3(x) | (oracle_split_3(y) << 1) |
         (oracle_split_3(z) << 2);
}

/* WGSL u32(f32) conversion (aabb-z-idx.js:194): truncation toward zero,
 * saturating to [0, 2^32-1].  NaN maps to 0 (WGSL leaves it
 * implementation-defined; AMD's v_cvt_u32_f32 also gives 0). */
uint32_t oracle_f32_to_u32(float v) {
  if (!(v > 0.0f)) return 0u; /* negative, zero, NaN */
  if (v >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)v;
}

/* Scene bounds: fp32 min/max over every vertex coordinate.  The reference
 * takes bounds from the caller (parse-obj.js:10-53, uploaded as a uniform at
 * aabb-z-idx.js:113-124); for synthetic inputs the caller's f64 min/max
 * rounded to fp32 equals this fp32 min/max because rounding is monotonic. */
void oracle_bounds(const float *tri, uint32_t n, float out[6]) {
  float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
  if (n == 0) {
    for (int c = 0; c < 3; ++c) { out[c] = 0.0f; out[3 + c] = 0.0f; }
    return;
  }
  for (uint32_t i = 0; i < n; ++i)
    for (int v = 0; v < 3; ++v)
      for (int c = 0; c < 3; ++c) {
        float x = tri[12 * (size_t)i + 4 * v + c];
        mn[c] = fmin32(mn[c], x);
        mx[c] = fmax32(mx[c], x);
      }
  for (int c = 0; c < 3; ++c) { out[c] = mn[c]; out[3 + c] = mx[c]; }
}

/* compute_aabb_z_idx (aabb-z-idx.js:177-195), one iteration per thread. */
void oracle_aabb_z_idx(const float *tri, uint32_t n, const float bounds[6],
                       int div_mode, float *aabb, uint32_t *keys) {
  float ext[3], rcp[3];
  for (int c = 0; c < 3; ++c) {
    ext[c] = bounds[3 + c] - bounds[c]; /* uniforms.max - uniforms.min */
    rcp[c] = 1.0f / ext[c];
  }
  for (uint32_t i = 0; i < n; ++i) {
    const float *t = tri + 12 * (size_t)i;
    float bmin[3], bmax[3];
    uint32_t q[3];
    for (int c = 0; c < 3; ++c) {
      /* box.min = min(v0, min(v1, v2)); box.max likewise (:186-187) */
      bmin[c] = fmin32(t[c], fmin32(t[4 + c], t[8 + c]));
      bmax[c] = fmax32(t[c], fmax32(t[4 + c], t[8 + c]));
      /* cen = (max + min) * .5f (:191) */
      float cen = (bmax[c] + bmin[c]) * 0.5f;
      /* rel = (cen - min) / (max - min) (:192) */
      float d = cen - bounds[c];
      float rel = (div_mode == ORACLE_DIV_IEEE) ? d / ext[c] : d * rcp[c];
      /* vec3u(rel * 1023.99f) (:194) */
      q[c] = oracle_f32_to_u32(rel * 1023.99f);
    }
    if (aabb) {
      float *a = aabb + 8 * (size_t)i;
      a[0] = bmin[0]; a[1] = bmin[1]; a[2] = bmin[2]; a[3] = 0.0f;
      a[4] = bmax[0]; a[5] = bmax[1]; a[6] = bmax[2]; a[7] = 0.0f;
    }
    keys[i] = oracle_morton_code(q[0], q[1], q[2]);
  }
}

/* ------------------------------------------------------------------------ */
/* Radix sort: radix-sort.js:162-500 (host), :527-629 (WGSL)                 */
/* ------------------------------------------------------------------------ */

/* 16 passes of a stable 2-bit LSD sort (radix-sort.js:486-488).  Each pass
 * counts digit (v >> 2k) & 3 (init_off :540-570), takes exclusive prefixes
 * over digits and over earlier elements (scan_up :764-803 and scan_and_sort
 * :576-629), then scatters key and index to their ranks.  init_idx
 * (:527-534) seeds idx[i] = i.  The sort is destructive: sorted keys end in
 * `keys` (:499, even pass count). */
void oracle_radix_sort(uint32_t *keys, int32_t *idx, uint32_t n) {
  if (n == 0) return;
  uint32_t *k2 = (uint32_t *)malloc(sizeof(uint32_t) * n);
  int32_t *i2 = (int32_t *)malloc(sizeof(int32_t) * n);
  for (uint32_t i = 0; i < n; ++i) idx[i] = (int32_t)i;
  uint32_t *src_k = keys, *dst_k = k2;
  int32_t *src_i = idx, *dst_i = i2;
  for (uint32_t win = 0; win < 16; ++win) {
    uint32_t count[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < n; ++i) count[(src_k[i] >> (2u * win)) & 3u]++;
    uint32_t base[4] = {0, count[0], count[0] + count[1],
                        count[0] + count[1] + count[2]};
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t d = (src_k[i] >> (2u * win)) & 3u;
      uint32_t pos = base[d]++;
      dst_k[pos] = src_k[i];
      dst_i[pos] = src_i[i];
    }
    uint32_t *tk = src_k; src_k = dst_k; dst_k = tk;
    int32_t *ti = src_i; src_i = dst_i; dst_i = ti;
  }
  /* 16 passes: the data is back in keys/idx. */
  free(k2);
  free(i2);
}

/* ------------------------------------------------------------------------ */
/* Radix tree: radix-tree.js:137-238                                         */
/* ------------------------------------------------------------------------ */

static int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

/* dif (radix-tree.js:160-169): index, counted from the MSB, of the first bit
 * where the keys differ; -1 if equal.  `faithful` runs the reference's
 * 32-step bit loop; otherwise count-leading-zeros of a^b (same value). */
int32_t oracle_dif(uint32_t a, uint32_t b, int faithful) {
  if (faithful) {
    for (uint32_t i = 0; i < 32; ++i) {
      uint32_t mask = 1u << (31u - i);
      if ((a & mask) != (b & mask)) return (int32_t)i;
    }
    return -1;
  }
  return (a == b) ? -1 : clz32(a ^ b);
}

typedef struct {
  const uint32_t *keys;
  int32_t n;
  int faithful;
} tree_ctx;

/* del (radix-tree.js:172-185): common-prefix length of the keys at two sorted
 * positions, -1 out of range, ties broken by 32 + dif(i, j). */
static int32_t del(const tree_ctx *t, int32_t i, int32_t j) {
  if (i >= t->n || j >= t->n || i < 0 || j < 0) return -1;
  int32_t kd = oracle_dif(t->keys[i], t->keys[j], t->faithful);
  if (kd == -1) kd = 32 + oracle_dif((uint32_t)i, (uint32_t)j, t->faithful);
  return kd;
}

/* WGSL i32 arithmetic wraps; these emulate it without C overflow UB. */
static int32_t wrap_i32(int64_t v) { return (int32_t)(uint32_t)(uint64_t)v; }
static int32_t wgsl_idiv(int32_t a, int32_t b) {
  if (b == 0) return a; /* WGSL: integer division by zero yields e1 */
  if (a == INT32_MIN && b == -1) return a;
  return a / b;
}

/* compute_child_index (radix-tree.js:188-238). */
static void child_index(const tree_ctx *t, int32_t i, int32_t out[2]) {
  int32_t dp = del(t, i, i + 1), dm = del(t, i, i - 1);
  int32_t d = (dp - dm > 0) ? 1 : ((dp - dm < 0) ? -1 : 0); /* sign() */
  int32_t del_min = del(t, i, i - d);
  int64_t l_max = 2;
  while (del(t, i, (int32_t)(i + l_max * d)) > del_min) l_max *= 2;
  int32_t l = 0;
  for (int64_t s = l_max / 2; s > 0; s /= 2)
    if (del(t, i, (int32_t)(i + (l + s) * d)) > del_min) l += (int32_t)s;
  int32_t j = i + l * d;
  int32_t del_node = del(t, i, j);
  int32_t s = 0;
  if (t->faithful) {
    /* The reference's loop (:216-224): t = (l-1+v)/v with v doubling in
     * i32.  It sticks at t = 1 until v wraps to INT32_MIN and then 0; the
     * extra t = 1 probes are idempotent. */
    int32_t v = 2;
    int32_t tt = wgsl_idiv(wrap_i32((int64_t)l - 1 + v), v);
    while (tt > 0) {
      if (del(t, i, wrap_i32((int64_t)i + ((int64_t)s + tt) * d)) > del_node)
        s += tt;
      v = wrap_i32((int64_t)v * 2);
      tt = wgsl_idiv(wrap_i32((int64_t)l - 1 + v), v);
    }
  } else {
    /* Textbook Karras: t = ceil(l/2), ceil(l/4), ..., 1. */
    for (int64_t v = 2;; v *= 2) {
      int32_t tt = (int32_t)((l - 1 + v) / v);
      if (del(t, i, i + (s + tt) * d) > del_node) s += tt;
      if (tt == 1) break;
    }
  }
  int32_t gamma = i + s * d + (d < 0 ? d : 0); /* min(d, 0) */
  out[0] = gamma;
  out[1] = gamma + 1;
  if ((i < j ? i : j) == gamma) out[0] = -out[0] - 1;
  if ((i > j ? i : j) == gamma + 1) out[1] = -out[1] - 1;
}

/* compute_radix_tree_pointers (radix-tree.js:137-157).  The buffer is
 * zero-initialised like a fresh WebGPU buffer, so slots 0 (root) and N-1
 * (no such internal node) stay 0. */
void oracle_radix_tree(const uint32_t *keys, uint32_t n, int32_t *parents,
                       int faithful) {
  memset(parents, 0, sizeof(int32_t) * 2 * (size_t)n);
  tree_ctx t = {keys, (int32_t)n, faithful};
  for (int32_t i = 0; i + 1 < (int32_t)n; ++i) {
    int32_t c[2];
    child_index(&t, i, c);
    for (int k = 0; k < 2; ++k) {
      if (c[k] >= 0) parents[c[k]] = i;
      else parents[n + (uint32_t)(-(c[k] + 1))] = i;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Up pass: bvh-up-pass.js:200-280                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
  float l_min[3]; int32_t l_child;
  float l_max[3]; int32_t f_1;
  float r_min[3]; int32_t r_child;
  float r_max[3]; int32_t f_2;
} bvh_node; /* bvh-up-pass.js:161-178, pathtracer.js:441-450 */

/* Zero components are moved off 0 because 0 is the reference's "not yet
 * written" sentinel (bvh-up-pass.js:208-210, checked at :266). */
static void perturb(float mn[3], float mx[3]) {
  for (int c = 0; c < 3; ++c) {
    mn[c] = mn[c] - ((mn[c] == 0.0f) ? 1.0f : 0.0f) * 1e-8f;
    mx[c] = mx[c] + ((mx[c] == 0.0f) ? 1.0f : 0.0f) * 1e-8f;
  }
}

/* bvh_upward_pass run one thread at a time, each to completion: a valid
 * interleaving of the reference's threads, in which the sibling box is
 * always visible, so the spin at :265-268 never triggers and f_1 ends at 2
 * on every internal node.  Slot order is arrival order (first arriver in
 * l_*), like the reference. */
void oracle_up_pass(const int32_t *idx, const float *aabb,
                    const int32_t *parents, uint32_t n, void *out) {
  bvh_node *bvh = (bvh_node *)out;
  memset(bvh, 0, sizeof(bvh_node) * (size_t)n);
  for (uint32_t g = 0; g < n; ++g) {
    const float *a = aabb + 8 * (size_t)idx[g];
    float mn[3] = {a[0], a[1], a[2]}, mx[3] = {a[4], a[5], a[6]};
    perturb(mn, mx);
    int32_t c_idx = (int32_t)g, w_idx = -((int32_t)g + 1), level = 0;
    while (w_idx != 0 || level == 0) {
      int32_t p = (level == 0) ? parents[c_idx + (int32_t)n] : parents[c_idx];
      int32_t sib = bvh[p].f_1++;
      if (sib == 0) {
        memcpy(bvh[p].l_min, mn, 12);
        memcpy(bvh[p].l_max, mx, 12);
        bvh[p].l_child = w_idx;
        break;
      }
      memcpy(bvh[p].r_min, mn, 12);
      memcpy(bvh[p].r_max, mx, 12);
      bvh[p].r_child = w_idx;
      for (int c = 0; c < 3; ++c) {
        mn[c] = fmin32(mn[c], bvh[p].l_min[c]);
        mx[c] = fmax32(mx[c], bvh[p].l_max[c]);
      }
      c_idx = p;
      w_idx = p;
      level += 1;
    }
  }
}

/* Child "number" in Karras order: internal node c -> c, leaf -(p+1) -> p.
 * The left child of a node split at gamma has number gamma, the right one
 * gamma + 1 (radix-tree.js:226-235). */
static int32_t child_number(int32_t c) { return c >= 0 ? c : -(c + 1); }

/* Puts each internal node's children in Karras order (left = split child
 * gamma) and sets f_1 = 2, f_2 = 0.  The reference's slot order and f_1 are
 * arrival-order artefacts (bvh-up-pass.js:230-252); the path tracer reads
 * both slots symmetrically (pathtracer.js:526-565), so this canonical form
 * is what the GPU path writes and what parity compares. */
void oracle_canonicalize_bvh(void *p, uint32_t n) {
  bvh_node *bvh = (bvh_node *)p;
  if (n < 2) return; /* N = 1: node 0 holds a single child, f_1 = 1 */
  for (uint32_t i = 0; i + 1 < n; ++i) {
    bvh_node *b = &bvh[i];
    if (child_number(b->l_child) > child_number(b->r_child)) {
      bvh_node t = *b;
      memcpy(b->l_min, t.r_min, 12); memcpy(b->l_max, t.r_max, 12);
      memcpy(b->r_min, t.l_min, 12); memcpy(b->r_max, t.l_max, 12);
      b->l_child = t.r_child; b->r_child = t.l_child;
    }
    b->f_1 = 2;
    b->f_2 = 0;
  }
}

/* rearrange_triangles (rearrange.js:162-169): o[i] = in[idx[i]], whole
 * 48-byte records (the padding lane is carried along). */
void oracle_rearrange(const float *tri, const int32_t *idx, uint32_t n,
                      float *out) {
  for (uint32_t i = 0; i < n; ++i)
    memcpy(out + 12 * (size_t)i, tri + 12 * (size_t)idx[i], 48);
}

/* build (build-bvh.js:12-66): AABB/Morton -> sort -> radix tree -> up pass
 * -> rearrange.  `bounds_or_null` plays MODEL_BOUNDS; when absent the scene
 * bounds are the fp32 min/max over all vertices.  The BVH is returned in
 * canonical child order. */
void oracle_build(const float *tri, uint32_t n, const float *bounds_or_null,
                  int div_mode, void *bvh_out, float *tri_out,
                  uint32_t *sorted_keys_out, int32_t *perm_out) {
  if (n == 0) return;
  float b[6];
  if (bounds_or_null) memcpy(b, bounds_or_null, sizeof b);
  else oracle_bounds(tri, n, b);
  float *aabb = (float *)malloc(32 * (size_t)n);
  uint32_t *keys = (uint32_t *)malloc(4 * (size_t)n);
  int32_t *idx = (int32_t *)malloc(4 * (size_t)n);
  int32_t *par = (int32_t *)malloc(8 * (size_t)n);
  oracle_aabb_z_idx(tri, n, b, div_mode, aabb, keys);
  oracle_radix_sort(keys, idx, n);
  oracle_radix_tree(keys, n, par, 0);
  oracle_up_pass(idx, aabb, par, n, bvh_out);
  oracle_canonicalize_bvh(bvh_out, n);
  if (tri_out) oracle_rearrange(tri, idx, n, tri_out);
  if (sorted_keys_out) memcpy(sorted_keys_out, keys, 4 * (size_t)n);
  if (perm_out) memcpy(perm_out, idx, 4 * (size_t)n);
  free(aabb); free(keys); free(idx); free(par);
}

/* Refit: the reference has no refit entry point; a refit is its up pass
 * (bvh-up-pass.js:200-280) and rearrange (rearrange.js:162-169) re-run on
 * new vertex positions with the topology (sorted keys, permutation) of an
 * earlier build kept fixed. */
void oracle_refit(const float *tri, uint32_t n, const uint32_t *sorted_keys,
                  const int32_t *perm, void *bvh_out, float *tri_out) {
  if (n == 0) return;
  float *aabb = (float *)malloc(32 * (size_t)n);
  int32_t *par = (int32_t *)malloc(8 * (size_t)n);
  uint32_t *dummy = (uint32_t *)malloc(4 * (size_t)n);
  float b[6] = {0, 0, 0, 1, 1, 1};
  oracle_aabb_z_idx(tri, n, b, ORACLE_DIV_IEEE, aabb, dummy);
  oracle_radix_tree(sorted_keys, n, par, 0);
  oracle_up_pass(perm, aabb, par, n, bvh_out);
  oracle_canonicalize_bvh(bvh_out, n);
  if (tri_out) oracle_rearrange(tri, perm, n, tri_out);
  free(aabb); free(par); free(dummy);
}

/* ------------------------------------------------------------------------ */
/* Independent brute force (shares no code with the restatement above).     */
/* ------------------------------------------------------------------------ */

/* Augmented key: the Morton key with the sorted position appended, so all
 * keys are distinct; the common-prefix length of two augmented keys is the
 * reference's del() (radix-tree.js:172-185). */
static uint64_t akey(const uint32_t *k, int32_t i) {
  return ((uint64_t)k[i] << 32) | (uint32_t)i;
}

/* Split of the range [l, r] (l < r) in a binary radix tree: the last
 * position whose augmented key has a 0 at the first bit where the range's
 * end keys differ. */
static int32_t bf_split(const uint32_t *k, int32_t l, int32_t r) {
  uint64_t x = akey(k, l) ^ akey(k, r);
  int b = 63 - __builtin_clzll(x);
  int32_t g = l;
  while (g + 1 <= r && !((akey(k, g + 1) >> b) & 1)) ++g;
  return g;
}

/* Top-down construction with Karras numbering: root = 0, the child holding
 * the left half of a node split at g is numbered g, the right one g + 1. */
static void bf_tree(const uint32_t *k, int32_t n, int32_t l, int32_t r,
                    int32_t id, int32_t *par) {
  int32_t g = bf_split(k, l, r);
  if (l == g) par[n + g] = id;
  else { par[g] = id; bf_tree(k, n, l, g, g, par); }
  if (g + 1 == r) par[n + r] = id;
  else { par[g + 1] = id; bf_tree(k, n, g + 1, r, g + 1, par); }
}

void oracle_bruteforce_tree(const uint32_t *sorted_keys, uint32_t n,
                            int32_t *parents_out) {
  memset(parents_out, 0, 8 * (size_t)n);
  if (n >= 2) bf_tree(sorted_keys, (int32_t)n, 0, (int32_t)n - 1, 0,
                      parents_out);
}

static void bf_leaf_box(const float *tri, int32_t t, float mn[3],
                        float mx[3]) {
  const float *p = tri + 12 * (size_t)t;
  for (int c = 0; c < 3; ++c) {
    float a = p[c], b = p[4 + c], d = p[8 + c];
    mn[c] = a; if (b < mn[c]) mn[c] = b; if (d < mn[c]) mn[c] = d;
    mx[c] = a; if (b > mx[c]) mx[c] = b; if (d > mx[c]) mx[c] = d;
    if (mn[c] == 0.0f) mn[c] = -1e-8f;
    if (mx[c] == 0.0f) mx[c] = 1e-8f;
  }
}

/* Box of a child = min/max over its leaf range, computed directly. */
static void bf_range_box(const float *tri, const int32_t *perm, int32_t l,
                         int32_t r, float mn[3], float mx[3]) {
  bf_leaf_box(tri, perm[l], mn, mx);
  for (int32_t i = l + 1; i <= r; ++i) {
    float a[3], b[3];
    bf_leaf_box(tri, perm[i], a, b);
    for (int c = 0; c < 3; ++c) {
      if (a[c] < mn[c]) mn[c] = a[c];
      if (b[c] > mx[c]) mx[c] = b[c];
    }
  }
}

static void bf_nodes(const float *tri, const int32_t *perm, const uint32_t *k,
                     int32_t l, int32_t r, int32_t id, bvh_node *bvh) {
  int32_t g = bf_split(k, l, r);
  bvh_node *b = &bvh[id];
  bf_range_box(tri, perm, l, g, b->l_min, b->l_max);
  bf_range_box(tri, perm, g + 1, r, b->r_min, b->r_max);
  b->l_child = (l == g) ? -(g + 1) : g;
  b->r_child = (g + 1 == r) ? -(g + 2) : g + 1;
  b->f_1 = 2;
  b->f_2 = 0;
  if (l != g) bf_nodes(tri, perm, k, l, g, g, bvh);
  if (g + 1 != r) bf_nodes(tri, perm, k, g + 1, r, g + 1, bvh);
}

void oracle_bruteforce_bvh(const float *tri, const int32_t *perm,
                           const uint32_t *sorted_keys, uint32_t n,
                           void *out) {
  bvh_node *bvh = (bvh_node *)out;
  memset(bvh, 0, sizeof(bvh_node) * (size_t)n);
  if (n == 1) {
    bf_leaf_box(tri, perm[0], bvh[0].l_min, bvh[0].l_max);
    bvh[0].l_child = -1;
    bvh[0].f_1 = 1;
    return;
  }
  if (n >= 2)
    bf_nodes(tri, perm, sorted_keys, 0, (int32_t)n - 1, 0, bvh);
}

/* ------------------------------------------------------------------------ */
/* Ray casting: pathtracer.js:495-628                                        */
/* ------------------------------------------------------------------------ */

typedef struct { float x, y, z; } v3;
static v3 v3sub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static v3 v3cross(v3 a, v3 b) {
  v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  return r;
}
static float v3dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 v3at(const float *p) { v3 r = {p[0], p[1], p[2]}; return r; }

/* tri_intersect (pathtracer.js:587-619), Moller-Trumbore; out[3] = t or -1 */
static void tri_intersect(v3 o, v3 d, const float *t, float out[4]) {
  v3 v0 = v3at(t), v1 = v3at(t + 4), v2 = v3at(t + 8);
  v3 e1 = v3sub(v1, v0), e2 = v3sub(v2, v0);
  v3 pvec = v3cross(d, e2);
  float det = v3dot(e1, pvec);
  out[0] = out[1] = out[2] = out[3] = -1.0f;
  if (fabsf(det) < 1e-10f) return;
  float idet = 1.0f / det;
  v3 tvec = v3sub(o, v0);
  float u = v3dot(tvec, pvec) * idet;
  if (u < 0.f || u > 1.f) return;
  v3 qvec = v3cross(tvec, e1);
  float v = v3dot(d, qvec) * idet;
  if (v < 0.f || u + v > 1.f) return;
  v3 nrm = v3cross(e1, e2);
  float len = sqrtf(v3dot(nrm, nrm));
  out[0] = nrm.x / len; out[1] = nrm.y / len; out[2] = nrm.z / len;
  out[3] = v3dot(e2, qvec) * idet;
}

/* aabb_intersect (pathtracer.js:621-628) */
static float aabb_intersect(const float *lo, const float *hi, v3 o, v3 inv) {
  float fx = (hi[0] - o.x) * inv.x, fy = (hi[1] - o.y) * inv.y,
        fz = (hi[2] - o.z) * inv.z;
  float nx = (lo[0] - o.x) * inv.x, ny = (lo[1] - o.y) * inv.y,
        nz = (lo[2] - o.z) * inv.z;
  float t0 = fmax32(fmin32(fx, nx), fmax32(fmin32(fy, ny), fmin32(fz, nz)));
  float t1 = fmin32(fmax32(fx, nx), fmin32(fmax32(fy, ny), fmax32(fz, nz)));
  return (t1 >= t0) ? ((t1 < 0.f) ? -1e30f : t0) : -1e30f;
}

static v3 ray_dir(const float *r) {
  v3 d = {r[4], r[5], r[6]};
  /* d += (abs(d) < 1e-5) * 1e-5 (pathtracer.js:502-503) */
  if (fabsf(d.x) < 1e-5f) d.x = d.x + 1e-5f;
  if (fabsf(d.y) < 1e-5f) d.y = d.y + 1e-5f;
  if (fabsf(d.z) < 1e-5f) d.z = d.z + 1e-5f;
  return d;
}

static void finish_hit(v3 d, v3 nrm, float dist, float *h) {
  /* RayHit: normal facing the ray, dist -1 for a miss (:572-583) */
  v3 neg = {-nrm.x, -nrm.y, -nrm.z};
  float s = v3dot(d, neg) > 0.f ? 1.f : -1.f;
  h[0] = s * nrm.x; h[1] = s * nrm.y; h[2] = s * nrm.z;
  h[3] = (dist == 1e30f) ? -1.f : dist;
}

/* intersect_bvh (pathtracer.js:495-584): stack traversal from node 0,
 * nearer valid child first.  The stack holds 64 entries (the path tracer's
 * 32 can overflow on trees deeper than 32) and the loop is capped, since a
 * single-triangle BVH's empty r-slot (child 0, zero box) can loop. */
void oracle_raycast(const void *bvhp, const float *tri, uint32_t n,
                    const float *rays, uint32_t count, float *hits) {
  const bvh_node *bvh = (const bvh_node *)bvhp;
  const uint64_t cap = 4ull * n + 256;
  for (uint32_t q = 0; q < count; ++q) {
    const float *r = rays + 8 * (size_t)q;
    v3 o = {r[0], r[1], r[2]};
    v3 d = ray_dir(r);
    v3 inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float dist = 1e30f;
    v3 nrm = {0.f, 0.f, 0.f};
    int32_t stack[64];
    int sp = 0;
    int32_t node = 0;
    for (uint64_t step = 0; sp >= 0 && step < cap; ++step) {
      if (node < 0) {
        uint32_t t = (uint32_t)(-(node + 1));
        if (t < n) {
          float h[4];
          tri_intersect(o, d, tri + 12 * (size_t)t, h);
          if (h[3] > 0.f && h[3] < dist) {
            nrm.x = h[0]; nrm.y = h[1]; nrm.z = h[2];
            dist = fmin32(h[3], dist);
          }
        }
        sp -= 1;
        node = sp >= 0 ? stack[sp] : 0;
      } else {
        if ((uint32_t)node >= n) break;
        const bvh_node *b = &bvh[node];
        float ld = aabb_intersect(b->l_min, b->l_max, o, inv);
        float rd = aabb_intersect(b->r_min, b->r_max, o, inv);
        int lv = ld != -1e30f && ld < dist, rv = rd != -1e30f && rd < dist;
        if (lv && rv) {
          int32_t near = ld < rd ? b->l_child : b->r_child;
          int32_t far = ld < rd ? b->r_child : b->l_child;
          if (sp < 64) stack[sp] = far;
          sp = sp + 1 < 64 ? sp + 1 : 64;
          node = near;
        } else if (lv) {
          node = b->l_child;
        } else if (rv) {
          node = b->r_child;
        } else {
          sp -= 1;
          node = sp >= 0 ? stack[sp] : 0;
        }
      }
    }
    finish_hit(d, nrm, dist, hits + 4 * (size_t)q);
  }
}

void oracle_raycast_bruteforce(const float *tri, uint32_t n, const float *rays,
                               uint32_t count, float *hits) {
  for (uint32_t q = 0; q < count; ++q) {
    const float *r = rays + 8 * (size_t)q;
    v3 o = {r[0], r[1], r[2]};
    v3 d = ray_dir(r);
    float dist = 1e30f;
    v3 nrm = {0.f, 0.f, 0.f};
    for (uint32_t t = 0; t < n; ++t) {
      float h[4];
      tri_intersect(o, d, tri + 12 * (size_t)t, h);
      if (h[3] > 0.f && h[3] < dist) {
        nrm.x = h[0]; nrm.y = h[1]; nrm.z = h[2];
        dist = h[3];
      }
    }
    finish_hit(d, nrm, dist, hits + 4 * (size_t)q);
  }
}
