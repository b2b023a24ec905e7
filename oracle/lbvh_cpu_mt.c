/*
 * lbvh_cpu_mt.c -- multi-threaded CPU restatement of the reference LBVH
 * build (OpenMP), for the CPU baseline timed beside the GPU.
 *
 * TEST INFRASTRUCTURE ONLY (a baseline and an oracle, never the product).
 * Same results as oracle_build (lbvh_oracle.c), byte for byte, with the
 * stages parallelised the way a CPU implementation would be:
 *   bounds        per-thread min/max, then a fold    (parse-obj.js:10-53)
 *   AABB + Morton parallel loop                     (aabb-z-idx.js:177-208)
 *   sort          4 stable LSD passes of 8-bit digits with per-thread
 *                 histograms; the stable permutation is unique, so it is
 *                 the reference's 16 x 2-bit result  (radix-sort.js:162-500)
 *   radix tree    Karras per internal node, clz for dif()
 *                                                   (radix-tree.js:137-238)
 *   up pass       one climber per leaf, atomic arrival counters, the
 *                 second arrival merges              (bvh-up-pass.js:200-280);
 *                 children are written in Karras order with f_1 = 2,
 *                 f_2 = 0, the canonical form oracle_build returns
 *   rearrange     parallel gather                   (rearrange.js:162-169)
 * Leaf boxes carry the +-1e-8 perturbation of exact zeros, so no box has a
 * signed zero and the min/max merges give the same bits in any order.
 */
#include "lbvh_oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

static float mt_fmin(float a, float b) { return (b < a) ? b : a; }
static float mt_fmax(float a, float b) { return (b > a) ? b : a; }

typedef struct {
  float l_min[3]; int32_t l_child;
  float l_max[3]; int32_t f_1;
  float r_min[3]; int32_t r_child;
  float r_max[3]; int32_t f_2;
} mt_node;

static void mt_bounds(const float *tri, uint32_t n, float out[6]) {
  float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
#pragma omp parallel
  {
    float lm[3] = {1e30f, 1e30f, 1e30f}, lx[3] = {-1e30f, -1e30f, -1e30f};
#pragma omp for schedule(static) nowait
    for (int64_t i = 0; i < (int64_t)n; ++i)
      for (int v = 0; v < 3; ++v)
        for (int c = 0; c < 3; ++c) {
          const float x = tri[12 * i + 4 * v + c];
          lm[c] = mt_fmin(lm[c], x);
          lx[c] = mt_fmax(lx[c], x);
        }
#pragma omp critical
    for (int c = 0; c < 3; ++c) {
      mn[c] = mt_fmin(mn[c], lm[c]);
      mx[c] = mt_fmax(mx[c], lx[c]);
    }
  }
  for (int c = 0; c < 3; ++c) { out[c] = mn[c]; out[3 + c] = mx[c]; }
}

static void mt_morton(const float *tri, uint32_t n, const float b[6],
                      int div_mode, float *box, uint32_t *keys) {
  float ext[3], rcp[3];
  for (int c = 0; c < 3; ++c) {
    ext[c] = b[3 + c] - b[c];
    rcp[c] = 1.0f / ext[c];
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const float *t = tri + 12 * i;
    uint32_t q[3];
    for (int c = 0; c < 3; ++c) {
      const float lo = mt_fmin(t[c], mt_fmin(t[4 + c], t[8 + c]));
      const float hi = mt_fmax(t[c], mt_fmax(t[4 + c], t[8 + c]));
      box[6 * i + c] = lo;
      box[6 * i + 3 + c] = hi;
      const float cen = (hi + lo) * 0.5f;
      const float d = cen - b[c];
      const float rel = (div_mode == ORACLE_DIV_IEEE) ? d / ext[c] : d * rcp[c];
      q[c] = oracle_f32_to_u32(rel * 1023.99f);
    }
    keys[i] = oracle_morton_code(q[0], q[1], q[2]);
  }
}

/* Stable LSD sort of (key, index): 4 passes of 8 bits.  Thread t owns a
 * contiguous chunk; its digit-d keys go after every earlier thread's. */
static void mt_sort(uint32_t *keys, int32_t *idx, uint32_t n) {
  const int nt = omp_get_max_threads();
  uint32_t *k2 = (uint32_t *)malloc(4 * (size_t)n);
  int32_t *i2 = (int32_t *)malloc(4 * (size_t)n);
  uint32_t *hist = (uint32_t *)calloc((size_t)nt * 256, 4);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) idx[i] = (int32_t)i;
  uint32_t *sk = keys, *dk = k2;
  int32_t *si = idx, *di = i2;
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 8 * pass;
#pragma omp parallel num_threads(nt)
    {
      const int t = omp_get_thread_num(), T = omp_get_num_threads();
      const uint64_t lo = (uint64_t)n * t / T, hi = (uint64_t)n * (t + 1) / T;
      uint32_t *h = hist + 256 * (size_t)t;
      memset(h, 0, 256 * 4);
      for (uint64_t i = lo; i < hi; ++i) h[(sk[i] >> shift) & 255u]++;
#pragma omp barrier
#pragma omp single
      {
        uint32_t run = 0;
        for (int d = 0; d < 256; ++d)
          for (int u = 0; u < T; ++u) {
            const uint32_t c = hist[256 * (size_t)u + d];
            hist[256 * (size_t)u + d] = run;
            run += c;
          }
      }
      for (uint64_t i = lo; i < hi; ++i) {
        const uint32_t pos = h[(sk[i] >> shift) & 255u]++;
        dk[pos] = sk[i];
        di[pos] = si[i];
      }
    }
    uint32_t *tk = sk; sk = dk; dk = tk;
    int32_t *ti = si; si = di; di = ti;
  }
  /* four passes: the result is back in keys / idx */
  free(k2); free(i2); free(hist);
}

static int mt_delta(const uint32_t *k, int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = k[i], b = k[j];
  if (a != b) return __builtin_clz(a ^ b);
  return 32 + __builtin_clz((uint32_t)i ^ (uint32_t)j);
}

/* Karras: children of internal node i (leaves as -(p+1)) and the parents of
 * every node and leaf. */
static void mt_tree(const uint32_t *k, uint32_t n, int32_t *child,
                    int32_t *par_node, int32_t *par_leaf) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n - 1; ++i) {
    const int64_t N = n;
    const int d = (mt_delta(k, N, i, i + 1) - mt_delta(k, N, i, i - 1)) > 0 ? 1 : -1;
    const int dmin = mt_delta(k, N, i, i - d);
    int64_t lmax = 2;
    while (mt_delta(k, N, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
      if (mt_delta(k, N, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = mt_delta(k, N, i, j);
    int64_t s = 0;
    for (int64_t t = (l + 1) / 2;; t = (t + 1) / 2) {
      if (mt_delta(k, N, i, i + (s + t) * d) > dnode) s += t;
      if (t == 1) break;
    }
    const int64_t g = i + s * d + (d < 0 ? -1 : 0);
    const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    int32_t c0 = (int32_t)g, c1 = (int32_t)(g + 1);
    if (lo == g) { par_leaf[g] = (int32_t)i; c0 = -(int32_t)g - 1; }
    else par_node[g] = (int32_t)i;
    if (hi == g + 1) { par_leaf[g + 1] = (int32_t)i; c1 = -(int32_t)(g + 1) - 1; }
    else par_node[g + 1] = (int32_t)i;
    child[2 * i] = c0;
    child[2 * i + 1] = c1;
  }
}

static void mt_up(const int32_t *idx, const float *box, const int32_t *child,
                  const int32_t *par_node, const int32_t *par_leaf, uint32_t n,
                  mt_node *bvh, uint32_t *arrive) {
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < (int64_t)n; ++g) {
    const float *b = box + 6 * (size_t)idx[g];
    float mn[3], mx[3];
    for (int c = 0; c < 3; ++c) {
      mn[c] = b[c] == 0.0f ? -1e-8f : b[c];
      mx[c] = b[3 + c] == 0.0f ? 1e-8f : b[3 + c];
    }
    int32_t self = -(int32_t)g - 1;
    int32_t p = par_leaf[g];
    while (1) {
      mt_node *nd = &bvh[p];
      const int right = child[2 * p + 1] == self;
      float *dmn = right ? nd->r_min : nd->l_min;
      float *dmx = right ? nd->r_max : nd->l_max;
      memcpy(dmn, mn, 12);
      memcpy(dmx, mx, 12);
      if (right) nd->r_child = self; else nd->l_child = self;
      if (__atomic_fetch_add(&arrive[p], 1u, __ATOMIC_ACQ_REL) == 0) break;
      const float *smn = right ? nd->l_min : nd->r_min;
      const float *smx = right ? nd->l_max : nd->r_max;
      for (int c = 0; c < 3; ++c) {
        mn[c] = mt_fmin(mn[c], smn[c]);
        mx[c] = mt_fmax(mx[c], smx[c]);
      }
      nd->f_1 = 2;
      nd->f_2 = 0;
      if (p == 0) break;
      self = p;
      p = par_node[p];
    }
  }
}

int oracle_build_mt(const float *tri, uint32_t n, const float *bounds_or_null,
                    int div_mode, void *bvh_out, float *tri_out,
                    uint32_t *sorted_keys_out, int32_t *perm_out) {
  const int threads = omp_get_max_threads();
  if (n == 0) return threads;
  float b[6];
  if (bounds_or_null) memcpy(b, bounds_or_null, sizeof b);
  else mt_bounds(tri, n, b);
  float *box = (float *)malloc(24 * (size_t)n);
  uint32_t *keys = (uint32_t *)malloc(4 * (size_t)n);
  int32_t *idx = (int32_t *)malloc(4 * (size_t)n);
  mt_morton(tri, n, b, div_mode, box, keys);
  mt_sort(keys, idx, n);
  mt_node *bvh = (mt_node *)bvh_out;
  memset(bvh, 0, sizeof(mt_node) * (size_t)n);
  if (n == 1) {
    /* no internal node: the lone leaf fills node 0's l_* with f_1 = 1 */
    float mn[3], mx[3];
    for (int c = 0; c < 3; ++c) {
      mn[c] = box[c] == 0.0f ? -1e-8f : box[c];
      mx[c] = box[3 + c] == 0.0f ? 1e-8f : box[3 + c];
    }
    memcpy(bvh[0].l_min, mn, 12);
    memcpy(bvh[0].l_max, mx, 12);
    bvh[0].l_child = -1;
    bvh[0].f_1 = 1;
  } else {
    int32_t *child = (int32_t *)malloc(8 * (size_t)n);
    int32_t *par_node = (int32_t *)calloc(n, 4);
    int32_t *par_leaf = (int32_t *)malloc(4 * (size_t)n);
    uint32_t *arrive = (uint32_t *)calloc(n, 4);
    mt_tree(keys, n, child, par_node, par_leaf);
    mt_up(idx, box, child, par_node, par_leaf, n, bvh, arrive);
    free(child); free(par_node); free(par_leaf); free(arrive);
  }
  if (tri_out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
      memcpy(tri_out + 12 * i, tri + 12 * (size_t)idx[i], 48);
  }
  if (sorted_keys_out) memcpy(sorted_keys_out, keys, 4 * (size_t)n);
  if (perm_out) memcpy(perm_out, idx, 4 * (size_t)n);
  free(box); free(keys); free(idx);
  return threads;
}

/* Refit: the up pass over new vertices with the topology of an earlier
 * build (its sorted keys and permutation), multi-threaded.  Byte-identical
 * to oracle_refit (lbvh_oracle.c). */
void oracle_refit_mt(const float *tri, uint32_t n, const uint32_t *sorted_keys,
                     const int32_t *perm, void *bvh_out, float *tri_out) {
  if (n == 0) return;
  float *box = (float *)malloc(24 * (size_t)n);
  const int64_t N = n;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const float *t = tri + 12 * i;
    for (int c = 0; c < 3; ++c) {
      box[6 * i + c] = mt_fmin(t[c], mt_fmin(t[4 + c], t[8 + c]));
      box[6 * i + 3 + c] = mt_fmax(t[c], mt_fmax(t[4 + c], t[8 + c]));
    }
  }
  mt_node *bvh = (mt_node *)bvh_out;
  memset(bvh, 0, sizeof(mt_node) * (size_t)n);
  if (n == 1) {
    for (int c = 0; c < 3; ++c) {
      bvh[0].l_min[c] = box[c] == 0.0f ? -1e-8f : box[c];
      bvh[0].l_max[c] = box[3 + c] == 0.0f ? 1e-8f : box[3 + c];
    }
    bvh[0].l_child = -1;
    bvh[0].f_1 = 1;
  } else {
    int32_t *child = (int32_t *)malloc(8 * (size_t)n);
    int32_t *par_node = (int32_t *)calloc(n, 4);
    int32_t *par_leaf = (int32_t *)malloc(4 * (size_t)n);
    uint32_t *arrive = (uint32_t *)calloc(n, 4);
    mt_tree(sorted_keys, n, child, par_node, par_leaf);
    mt_up(perm, box, child, par_node, par_leaf, n, bvh, arrive);
    free(child); free(par_node); free(par_leaf); free(arrive);
  }
  if (tri_out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i)
      memcpy(tri_out + 12 * i, tri + 12 * (size_t)perm[i], 48);
  }
  free(box);
}

static uint64_t mt_splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* Digests of a build's outputs: out[0] = sum over the u32 words w_j of the
 * BVH (16 per triangle) of w_j * (splitmix64(j) | 1), out[1] the same over
 * O_TRI (12 per triangle), both modulo 2^64.  The serving loop reports the
 * same digests of its results (lbvh_completion.digest_bvh / digest_tri). */
void oracle_digest(const void *bvh, const float *tri_out, uint32_t n,
                   uint64_t out[2]) {
  const uint32_t *wb = (const uint32_t *)bvh;
  const uint32_t *wt = (const uint32_t *)tri_out;
  const int64_t nb = 16 * (int64_t)n, nt = 12 * (int64_t)n;
  uint64_t a = 0, b = 0;
#pragma omp parallel for schedule(static) reduction(+ : a)
  for (int64_t j = 0; j < nb; ++j)
    a += (uint64_t)wb[j] * (mt_splitmix64((uint64_t)j) | 1ull);
#pragma omp parallel for schedule(static) reduction(+ : b)
  for (int64_t j = 0; j < nt; ++j)
    b += (uint64_t)wt[j] * (mt_splitmix64((uint64_t)j) | 1ull);
  out[0] = a;
  out[1] = b;
}
