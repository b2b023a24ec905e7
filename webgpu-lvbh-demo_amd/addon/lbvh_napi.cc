// lbvh_napi.cc -- N-API addon binding the C ABI of liblbvh.so for Node.js.
//
// The JavaScript module bvh/index.js builds the reference's interface
// (initBVHBuild, init*Kernel(device).execute, GPUBuffer-like handles) on top
// of these functions.  Device pointers cross into JavaScript as BigInts, the
// context as an external with a finalizer.  Every enqueue call returns at
// once; synchronize() waits for the context's stream on a worker thread and
// settles a Promise, which is how `await device.queue.onSubmittedWorkDone()`
// (e.g. aabb-z-idx.js:145) maps onto HIP.  Failures throw (or reject with)
// an Error whose `code` is the lbvh_status.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/lbvh.h"

namespace {

#define NAPI_OK(env, call)                                        \
  do {                                                            \
    if ((call) != napi_ok) {                                      \
      napi_throw_error((env), "EINTERNAL", "N-API call failed: " #call); \
      return nullptr;                                             \
    }                                                             \
  } while (0)

napi_value make_error(napi_env env, int status, const std::string &msg) {
  napi_value m, err, code;
  napi_create_string_utf8(env, msg.c_str(), msg.size(), &m);
  napi_create_error(env, nullptr, m, &err);
  napi_create_int32(env, status, &code);
  napi_set_named_property(env, err, "code", code);
  return err;
}

// Throws for a failed status; returns true when the caller must bail out.
bool throw_status(napi_env env, lbvh_ctx *ctx, int st, const char *what) {
  if (st == LBVH_OK) return false;
  std::string msg = std::string(what) + ": " + lbvh_status_string(st);
  if (ctx) msg += std::string(" (") + lbvh_last_error(ctx) + ")";
  napi_throw(env, make_error(env, st, msg));
  return true;
}

bool get_args(napi_env env, napi_callback_info info, size_t want,
              napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok ||
      argc < want) {
    napi_throw_type_error(env, "EARGS", "wrong number of arguments");
    return false;
  }
  return true;
}

bool get_ctx(napi_env env, napi_value v, lbvh_ctx **out) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, "EARGS", "expected an LBVH context");
    return false;
  }
  *out = static_cast<lbvh_ctx *>(p);
  return true;
}

bool is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_null || t == napi_undefined;
}

// Device pointer from a BigInt (null/undefined -> nullptr).
bool get_ptr(napi_env env, napi_value v, void **out) {
  if (is_nullish(env, v)) {
    *out = nullptr;
    return true;
  }
  uint64_t x = 0;
  bool lossless = false;
  if (napi_get_value_bigint_uint64(env, v, &x, &lossless) != napi_ok) {
    napi_throw_type_error(env, "EARGS", "expected a device pointer (BigInt)");
    return false;
  }
  *out = reinterpret_cast<void *>(static_cast<uintptr_t>(x));
  return true;
}

bool get_u32(napi_env env, napi_value v, uint32_t *out) {
  double d = 0;
  if (napi_get_value_double(env, v, &d) != napi_ok || d < 0 ||
      d > 4294967295.0 || d != (double)(uint64_t)d) {
    napi_throw_type_error(env, "EARGS", "expected an unsigned 32-bit count");
    return false;
  }
  *out = static_cast<uint32_t>(d);
  return true;
}

bool get_u64(napi_env env, napi_value v, uint64_t *out) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_bigint) {
    bool lossless = false;
    return napi_get_value_bigint_uint64(env, v, out, &lossless) == napi_ok;
  }
  double d = 0;
  if (napi_get_value_double(env, v, &d) != napi_ok || d < 0) {
    napi_throw_type_error(env, "EARGS", "expected a byte count");
    return false;
  }
  *out = static_cast<uint64_t>(d);
  return true;
}

// Host bytes of a TypedArray / ArrayBuffer / DataView (zero copy).
bool get_bytes(napi_env env, napi_value v, void **data, size_t *len) {
  bool is;
  if (napi_is_typedarray(env, v, &is) == napi_ok && is) {
    napi_typedarray_type type;
    size_t count, offset;
    napi_value ab;
    void *base;
    napi_get_typedarray_info(env, v, &type, &count, &base, &ab, &offset);
    size_t elem = 1;
    switch (type) {
      case napi_int16_array: case napi_uint16_array: elem = 2; break;
      case napi_int32_array: case napi_uint32_array: case napi_float32_array:
        elem = 4; break;
      case napi_float64_array: case napi_bigint64_array:
      case napi_biguint64_array: elem = 8; break;
      default: elem = 1;
    }
    *data = base;  // already offset by N-API
    *len = count * elem;
    return true;
  }
  if (napi_is_arraybuffer(env, v, &is) == napi_ok && is) {
    napi_get_arraybuffer_info(env, v, data, len);
    return true;
  }
  if (napi_is_dataview(env, v, &is) == napi_ok && is) {
    napi_value ab;
    size_t off;
    napi_get_dataview_info(env, v, len, data, &ab, &off);
    return true;
  }
  napi_throw_type_error(env, "EARGS", "expected a TypedArray or ArrayBuffer");
  return false;
}

// Six host floats (min.xyz, max.xyz) from a Float32Array, or null.
bool get_bounds(napi_env env, napi_value v, const float **out, float *buf) {
  if (is_nullish(env, v)) {
    *out = nullptr;
    return true;
  }
  void *d;
  size_t len;
  if (!get_bytes(env, v, &d, &len)) return false;
  if (len != 6 * sizeof(float)) {
    napi_throw_type_error(env, "EARGS", "bounds must be a Float32Array(6)");
    return false;
  }
  std::memcpy(buf, d, 6 * sizeof(float));
  *out = buf;
  return true;
}

napi_value ptr_value(napi_env env, void *p) {
  napi_value r;
  napi_create_bigint_uint64(env, reinterpret_cast<uintptr_t>(p), &r);
  return r;
}

napi_value undefined(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

// ---- context ----------------------------------------------------------------

void finalize_ctx(napi_env, void *data, void *) {
  lbvh_destroy(static_cast<lbvh_ctx *>(data));
}

napi_value CreateContext(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  int32_t ordinal = 0;
  uint64_t budget = 0;
  napi_get_value_int32(env, argv[0], &ordinal);
  if (!get_u64(env, argv[1], &budget)) return nullptr;
  lbvh_ctx *ctx = nullptr;
  if (throw_status(env, nullptr, lbvh_create(ordinal, budget, &ctx),
                   "lbvh_create"))
    return nullptr;
  napi_value ext;
  NAPI_OK(env, napi_create_external(env, ctx, finalize_ctx, nullptr, &ext));
  return ext;
}

napi_value SetDivMode(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  lbvh_ctx *ctx;
  if (!get_args(env, info, 2, argv) || !get_ctx(env, argv[0], &ctx))
    return nullptr;
  int32_t mode = 0;
  napi_get_value_int32(env, argv[1], &mode);
  if (throw_status(env, ctx, lbvh_set_div_mode(ctx, mode), "lbvh_set_div_mode"))
    return nullptr;
  return undefined(env);
}

napi_value BytesInUse(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  lbvh_ctx *ctx;
  if (!get_args(env, info, 1, argv) || !get_ctx(env, argv[0], &ctx))
    return nullptr;
  napi_value r;
  napi_create_double(env, (double)lbvh_bytes_in_use(ctx), &r);
  return r;
}

napi_value Alloc(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  lbvh_ctx *ctx;
  uint64_t bytes;
  if (!get_args(env, info, 2, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_u64(env, argv[1], &bytes))
    return nullptr;
  void *p = nullptr;
  if (throw_status(env, ctx, lbvh_alloc(ctx, bytes, &p), "lbvh_alloc"))
    return nullptr;
  return ptr_value(env, p);
}

napi_value Free(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  lbvh_ctx *ctx;
  void *p;
  if (!get_args(env, info, 2, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &p))
    return nullptr;
  if (throw_status(env, ctx, lbvh_free(ctx, p), "lbvh_free")) return nullptr;
  return undefined(env);
}

napi_value Upload(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  lbvh_ctx *ctx;
  void *p, *host;
  size_t len;
  if (!get_args(env, info, 3, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &p) || !get_bytes(env, argv[2], &host, &len))
    return nullptr;
  if (throw_status(env, ctx, lbvh_upload(ctx, p, host, len), "lbvh_upload"))
    return nullptr;
  return undefined(env);
}

napi_value Download(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  lbvh_ctx *ctx;
  void *p;
  uint64_t bytes;
  if (!get_args(env, info, 3, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &p) || !get_u64(env, argv[2], &bytes))
    return nullptr;
  void *data = nullptr;
  napi_value ab;
  NAPI_OK(env, napi_create_arraybuffer(env, bytes, &data, &ab));
  if (throw_status(env, ctx, lbvh_download(ctx, data, p, bytes),
                   "lbvh_download"))
    return nullptr;
  return ab;
}

// ---- stage entry points (device pointers as BigInt) ---------------------------

napi_value Bounds(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  lbvh_ctx *ctx;
  void *tri, *out;
  uint32_t n;
  if (!get_args(env, info, 4, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &tri) || !get_u32(env, argv[2], &n) ||
      !get_ptr(env, argv[3], &out))
    return nullptr;
  if (throw_status(env, ctx,
                   lbvh_bounds(ctx, (const float *)tri, n, (float *)out),
                   "lbvh_bounds"))
    return nullptr;
  return undefined(env);
}

napi_value AabbMorton(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  lbvh_ctx *ctx;
  void *tri, *aabb, *keys;
  uint32_t n;
  float bbuf[6];
  const float *b;
  if (!get_args(env, info, 6, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &tri) || !get_u32(env, argv[2], &n) ||
      !get_bounds(env, argv[3], &b, bbuf) || !get_ptr(env, argv[4], &aabb) ||
      !get_ptr(env, argv[5], &keys))
    return nullptr;
  if (throw_status(env, ctx,
                   lbvh_aabb_morton(ctx, (const float *)tri, n, b,
                                    (float *)aabb, (uint32_t *)keys),
                   "lbvh_aabb_morton"))
    return nullptr;
  return undefined(env);
}

napi_value Sort(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  lbvh_ctx *ctx;
  void *keys, *idx;
  uint32_t n;
  if (!get_args(env, info, 4, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &keys) || !get_ptr(env, argv[2], &idx) ||
      !get_u32(env, argv[3], &n))
    return nullptr;
  if (throw_status(env, ctx,
                   lbvh_sort(ctx, (uint32_t *)keys, (int32_t *)idx, n),
                   "lbvh_sort"))
    return nullptr;
  return undefined(env);
}

napi_value RadixTree(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  lbvh_ctx *ctx;
  void *keys, *par;
  uint32_t n;
  if (!get_args(env, info, 4, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &keys) || !get_u32(env, argv[2], &n) ||
      !get_ptr(env, argv[3], &par))
    return nullptr;
  if (throw_status(env, ctx,
                   lbvh_radix_tree(ctx, (const uint32_t *)keys, n,
                                   (int32_t *)par),
                   "lbvh_radix_tree"))
    return nullptr;
  return undefined(env);
}

napi_value UpPass(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  lbvh_ctx *ctx;
  void *idx, *aabb, *par, *bvh;
  uint32_t n;
  if (!get_args(env, info, 6, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &idx) || !get_ptr(env, argv[2], &aabb) ||
      !get_ptr(env, argv[3], &par) || !get_u32(env, argv[4], &n) ||
      !get_ptr(env, argv[5], &bvh))
    return nullptr;
  if (throw_status(env, ctx,
                   lbvh_up_pass(ctx, (const int32_t *)idx, (const float *)aabb,
                                (const int32_t *)par, n, bvh),
                   "lbvh_up_pass"))
    return nullptr;
  return undefined(env);
}

napi_value Rearrange(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  lbvh_ctx *ctx;
  void *tri, *idx, *out;
  uint32_t n;
  if (!get_args(env, info, 5, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &tri) || !get_ptr(env, argv[2], &idx) ||
      !get_u32(env, argv[3], &n) || !get_ptr(env, argv[4], &out))
    return nullptr;
  if (throw_status(env, ctx,
                   lbvh_rearrange(ctx, (const float *)tri,
                                  (const int32_t *)idx, n, (float *)out),
                   "lbvh_rearrange"))
    return nullptr;
  return undefined(env);
}

napi_value Build(napi_env env, napi_callback_info info) {
  napi_value argv[8];
  lbvh_ctx *ctx;
  void *tri, *bvh, *otri, *sk, *pm;
  uint32_t n;
  float bbuf[6];
  const float *b;
  if (!get_args(env, info, 8, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &tri) || !get_u32(env, argv[2], &n) ||
      !get_bounds(env, argv[3], &b, bbuf) || !get_ptr(env, argv[4], &bvh) ||
      !get_ptr(env, argv[5], &otri) || !get_ptr(env, argv[6], &sk) ||
      !get_ptr(env, argv[7], &pm))
    return nullptr;
  if (throw_status(env, ctx,
                   lbvh_build(ctx, (const float *)tri, n, b, bvh,
                              (float *)otri, (uint32_t *)sk, (int32_t *)pm),
                   "lbvh_build"))
    return nullptr;
  return undefined(env);
}

napi_value Refit(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  lbvh_ctx *ctx;
  void *tri, *sk, *pm, *bvh, *otri;
  uint32_t n;
  if (!get_args(env, info, 7, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &tri) || !get_u32(env, argv[2], &n) ||
      !get_ptr(env, argv[3], &sk) || !get_ptr(env, argv[4], &pm) ||
      !get_ptr(env, argv[5], &bvh) || !get_ptr(env, argv[6], &otri))
    return nullptr;
  if (throw_status(env, ctx,
                   lbvh_refit(ctx, (const float *)tri, n,
                              (const uint32_t *)sk, (const int32_t *)pm, bvh,
                              (float *)otri),
                   "lbvh_refit"))
    return nullptr;
  return undefined(env);
}

// ---- batched builds, autotune, ray casting --------------------------------------

// buildBatched(ctx, tri, offsets: Uint32Array(S+1), bounds: Float32Array(6S) |
// null, bvh, otri, keys | null, perm | null)
napi_value BuildBatched(napi_env env, napi_callback_info info) {
  napi_value argv[8];
  lbvh_ctx *ctx;
  void *tri, *bvh, *otri, *sk, *pm, *off, *bnd = nullptr;
  size_t off_len = 0, bnd_len = 0;
  if (!get_args(env, info, 8, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &tri) || !get_bytes(env, argv[2], &off, &off_len) ||
      !get_ptr(env, argv[4], &bvh) || !get_ptr(env, argv[5], &otri) ||
      !get_ptr(env, argv[6], &sk) || !get_ptr(env, argv[7], &pm))
    return nullptr;
  if (off_len < 4 || off_len % 4) {
    napi_throw_type_error(env, "EARGS", "offsets must be a Uint32Array(S+1)");
    return nullptr;
  }
  const uint32_t S = (uint32_t)(off_len / 4) - 1;
  if (!is_nullish(env, argv[3])) {
    if (!get_bytes(env, argv[3], &bnd, &bnd_len)) return nullptr;
    if (bnd_len != 24ull * S) {
      napi_throw_type_error(env, "EARGS", "bounds must be a Float32Array(6 * S)");
      return nullptr;
    }
  }
  if (throw_status(env, ctx,
                   lbvh_build_batched(ctx, (const float *)tri,
                                      (const uint32_t *)off, S,
                                      (const float *)bnd, bvh, (float *)otri,
                                      (uint32_t *)sk, (int32_t *)pm),
                   "lbvh_build_batched"))
    return nullptr;
  return undefined(env);
}

// autotune(ctx, nMin, nMax, reps) -> Float32Array(buckets x 3 tree tiles) ms
napi_value Autotune(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  lbvh_ctx *ctx;
  uint32_t nmin, nmax, reps;
  if (!get_args(env, info, 4, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_u32(env, argv[1], &nmin) || !get_u32(env, argv[2], &nmax) ||
      !get_u32(env, argv[3], &reps))
    return nullptr;
  int b0 = 0, b1 = 0;
  while (nmin > 1 && (2ull << b0) <= nmin) ++b0;
  while (nmax > 1 && (2ull << b1) <= nmax) ++b1;
  // per bucket: tree tiles 256, 512, 1024 x digit widths 8, 10
  const size_t count = nmax >= nmin ? (size_t)(b1 - b0 + 1) * 6 : 0;
  void *data = nullptr;
  napi_value ab, arr;
  NAPI_OK(env, napi_create_arraybuffer(env, 4 * (count ? count : 1), &data, &ab));
  if (throw_status(env, ctx,
                   lbvh_autotune(ctx, nmin, nmax, (int)reps, (float *)data),
                   "lbvh_autotune"))
    return nullptr;
  NAPI_OK(env, napi_create_typedarray(env, napi_float32_array, count, ab, 0, &arr));
  return arr;
}

// raycast(ctx, bvh, tri, n, rays, count, hits): device pointers
napi_value Raycast(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  lbvh_ctx *ctx;
  void *bvh, *tri, *rays, *hits;
  uint32_t n, count;
  if (!get_args(env, info, 7, argv) || !get_ctx(env, argv[0], &ctx) ||
      !get_ptr(env, argv[1], &bvh) || !get_ptr(env, argv[2], &tri) ||
      !get_u32(env, argv[3], &n) || !get_ptr(env, argv[4], &rays) ||
      !get_u32(env, argv[5], &count) || !get_ptr(env, argv[6], &hits))
    return nullptr;
  if (throw_status(env, ctx,
                   lbvh_raycast(ctx, bvh, (const float *)tri, n,
                                (const float *)rays, count, (float *)hits),
                   "lbvh_raycast"))
    return nullptr;
  return undefined(env);
}

// ---- serving ----------------------------------------------------------------------

void finalize_server(napi_env, void *data, void *) {
  lbvh_server_destroy(static_cast<lbvh_server *>(data));
}

bool get_server(napi_env env, napi_value v, lbvh_server **out) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, "EARGS", "expected an LBVH server");
    return false;
  }
  *out = static_cast<lbvh_server *>(p);
  return true;
}

bool throw_server(napi_env env, lbvh_server *srv, int st, const char *what) {
  if (st == LBVH_OK) return false;
  std::string msg = std::string(what) + ": " + lbvh_status_string(st) + " (" +
                    lbvh_server_last_error(srv) + ")";
  napi_throw(env, make_error(env, st, msg));
  return true;
}

napi_value u64_value(napi_env env, uint64_t x) {
  napi_value r;
  napi_create_bigint_uint64(env, x, &r);
  return r;
}

napi_value num_value(napi_env env, double x) {
  napi_value r;
  napi_create_double(env, x, &r);
  return r;
}

// createServer(ordinal, hbmBudget, slots) -> external
napi_value CreateServer(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  int32_t ordinal = 0;
  uint64_t budget = 0;
  uint32_t slots = 0;
  napi_get_value_int32(env, argv[0], &ordinal);
  if (!get_u64(env, argv[1], &budget) || !get_u32(env, argv[2], &slots))
    return nullptr;
  lbvh_server *srv = nullptr;
  const int st = lbvh_server_create(ordinal, budget, slots, &srv);
  if (st != LBVH_OK) {
    napi_throw(env, make_error(env, st, std::string("lbvh_server_create: ") +
                                            lbvh_status_string(st)));
    return nullptr;
  }
  napi_value ext;
  NAPI_OK(env, napi_create_external(env, srv, finalize_server, nullptr, &ext));
  return ext;
}

napi_value ServerSetDivMode(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  lbvh_server *srv;
  if (!get_args(env, info, 2, argv) || !get_server(env, argv[0], &srv))
    return nullptr;
  int32_t mode = 0;
  napi_get_value_int32(env, argv[1], &mode);
  if (throw_server(env, srv, lbvh_server_set_div_mode(srv, mode),
                   "lbvh_server_set_div_mode"))
    return nullptr;
  return undefined(env);
}

// serverSetDigest(srv, enable): later requests digest their results.
napi_value ServerSetDigest(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  lbvh_server *srv;
  if (!get_args(env, info, 2, argv) || !get_server(env, argv[0], &srv))
    return nullptr;
  bool on = false;
  napi_get_value_bool(env, argv[1], &on);
  if (throw_server(env, srv, lbvh_server_set_digest(srv, on ? 1 : 0),
                   "lbvh_server_set_digest"))
    return nullptr;
  return undefined(env);
}

// serverCancel(srv, requestId): cancels a request still waiting.
napi_value ServerCancel(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  lbvh_server *srv;
  uint64_t id = 0;
  if (!get_args(env, info, 2, argv) || !get_server(env, argv[0], &srv) ||
      !get_u64(env, argv[1], &id))
    return nullptr;
  if (throw_server(env, srv, lbvh_server_cancel(srv, id), "lbvh_server_cancel"))
    return nullptr;
  return undefined(env);
}

// serverSetMaxWait(srv, seconds): queueing limit of later requests.
napi_value ServerSetMaxWait(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  lbvh_server *srv;
  if (!get_args(env, info, 2, argv) || !get_server(env, argv[0], &srv))
    return nullptr;
  double w = 0.0;
  napi_get_value_double(env, argv[1], &w);
  if (throw_server(env, srv, lbvh_server_set_max_wait(srv, w),
                   "lbvh_server_set_max_wait"))
    return nullptr;
  return undefined(env);
}

// serverSubmit(srv, kind, meshId, source, tri, n, seed, param) -> request id.
// tri: a TypedArray (host source, copied at submit), a BigInt device pointer
// (device source) or null (synthesised sources).
napi_value ServerSubmit(napi_env env, napi_callback_info info) {
  napi_value argv[8];
  lbvh_server *srv;
  int32_t kind = 0, source = 0;
  uint64_t mesh = 0, seed = 0;
  uint32_t n = 0;
  double param = 0.0;
  if (!get_args(env, info, 8, argv) || !get_server(env, argv[0], &srv) ||
      !get_u64(env, argv[2], &mesh) || !get_u32(env, argv[5], &n) ||
      !get_u64(env, argv[6], &seed))
    return nullptr;
  napi_get_value_int32(env, argv[1], &kind);
  napi_get_value_int32(env, argv[3], &source);
  napi_get_value_double(env, argv[7], &param);
  const float *tri = nullptr;
  if (!is_nullish(env, argv[4])) {
    napi_valuetype t;
    napi_typeof(env, argv[4], &t);
    void *p = nullptr;
    size_t len = 0;
    if (t == napi_bigint) {
      if (!get_ptr(env, argv[4], &p)) return nullptr;
    } else {
      if (!get_bytes(env, argv[4], &p, &len)) return nullptr;
      if (len < 48ull * n) {
        napi_throw_type_error(env, "EARGS", "triangle array shorter than 48 * n bytes");
        return nullptr;
      }
    }
    tri = static_cast<const float *>(p);
  }
  uint64_t id = 0;
  if (throw_server(env, srv,
                   lbvh_server_submit(srv, kind, mesh, source, tri, n, seed,
                                      param, &id),
                   "lbvh_server_submit"))
    return nullptr;
  return u64_value(env, id);
}

// serverPoll(srv) -> [{requestId, meshId, kind, status, n, slot, submitS,
// startS, doneS, digestBvh, digestTri}] (digests as BigInt)
napi_value ServerPoll(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  lbvh_server *srv;
  if (!get_args(env, info, 1, argv) || !get_server(env, argv[0], &srv))
    return nullptr;
  napi_value arr;
  NAPI_OK(env, napi_create_array(env, &arr));
  lbvh_completion c[256];
  uint32_t got = 0, at = 0;
  do {
    if (throw_server(env, srv, lbvh_server_poll(srv, c, 256, &got),
                     "lbvh_server_poll"))
      return nullptr;
    for (uint32_t i = 0; i < got; ++i) {
      napi_value o;
      napi_create_object(env, &o);
      napi_set_named_property(env, o, "requestId", u64_value(env, c[i].request_id));
      napi_set_named_property(env, o, "meshId", u64_value(env, c[i].mesh_id));
      napi_set_named_property(env, o, "kind", num_value(env, c[i].kind));
      napi_set_named_property(env, o, "status", num_value(env, c[i].status));
      napi_set_named_property(env, o, "n", num_value(env, c[i].n));
      napi_set_named_property(env, o, "slot", num_value(env, c[i].slot));
      napi_set_named_property(env, o, "submitS", num_value(env, c[i].submit_s));
      napi_set_named_property(env, o, "startS", num_value(env, c[i].start_s));
      napi_set_named_property(env, o, "doneS", num_value(env, c[i].done_s));
      napi_set_named_property(env, o, "digestBvh", u64_value(env, c[i].digest_bvh));
      napi_set_named_property(env, o, "digestTri", u64_value(env, c[i].digest_tri));
      napi_set_element(env, arr, at++, o);
    }
  } while (got == 256);
  return arr;
}

struct DrainWork {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  napi_ref srv_ref = nullptr;
  lbvh_server *srv = nullptr;
  int status = LBVH_OK;
  std::string error;
};

void drain_execute(napi_env, void *data) {
  DrainWork *w = static_cast<DrainWork *>(data);
  w->status = lbvh_server_drain(w->srv);
  if (w->status != LBVH_OK) w->error = lbvh_server_last_error(w->srv);
}

void drain_complete(napi_env env, napi_status, void *data) {
  DrainWork *w = static_cast<DrainWork *>(data);
  if (w->status == LBVH_OK) {
    napi_resolve_deferred(env, w->deferred, undefined(env));
  } else {
    napi_reject_deferred(
        env, w->deferred,
        make_error(env, w->status,
                   std::string("lbvh_server_drain: ") +
                       lbvh_status_string(w->status) + " (" + w->error + ")"));
  }
  napi_delete_reference(env, w->srv_ref);
  napi_delete_async_work(env, w->work);
  delete w;
}

// serverDrain(srv) -> Promise, settled off the JS thread
napi_value ServerDrain(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  lbvh_server *srv;
  if (!get_args(env, info, 1, argv) || !get_server(env, argv[0], &srv))
    return nullptr;
  DrainWork *w = new DrainWork();
  w->srv = srv;
  napi_value promise, name;
  NAPI_OK(env, napi_create_promise(env, &w->deferred, &promise));
  NAPI_OK(env, napi_create_reference(env, argv[0], 1, &w->srv_ref));
  napi_create_string_utf8(env, "lbvh_server_drain", NAPI_AUTO_LENGTH, &name);
  NAPI_OK(env, napi_create_async_work(env, nullptr, name, drain_execute,
                                      drain_complete, w, &w->work));
  NAPI_OK(env, napi_queue_async_work(env, w->work));
  return promise;
}

// serverMesh(srv, meshId) -> {bvh: BigInt, tri: BigInt, n}
napi_value ServerMesh(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  lbvh_server *srv;
  uint64_t mesh = 0;
  if (!get_args(env, info, 2, argv) || !get_server(env, argv[0], &srv) ||
      !get_u64(env, argv[1], &mesh))
    return nullptr;
  void *bvh = nullptr;
  float *tri = nullptr;
  uint32_t n = 0;
  if (throw_server(env, srv, lbvh_server_mesh(srv, mesh, &bvh, &tri, &n),
                   "lbvh_server_mesh"))
    return nullptr;
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "bvh", ptr_value(env, bvh));
  napi_set_named_property(env, o, "tri", ptr_value(env, tri));
  napi_set_named_property(env, o, "n", num_value(env, n));
  return o;
}

napi_value ServerEvict(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  lbvh_server *srv;
  uint64_t mesh = 0;
  if (!get_args(env, info, 2, argv) || !get_server(env, argv[0], &srv) ||
      !get_u64(env, argv[1], &mesh))
    return nullptr;
  if (throw_server(env, srv, lbvh_server_evict(srv, mesh), "lbvh_server_evict"))
    return nullptr;
  return undefined(env);
}

napi_value ServerStats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  lbvh_server *srv;
  if (!get_args(env, info, 1, argv) || !get_server(env, argv[0], &srv))
    return nullptr;
  lbvh_server_stats st;
  if (throw_server(env, srv, lbvh_server_stats_get(srv, &st),
                   "lbvh_server_stats_get"))
    return nullptr;
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "submitted", num_value(env, (double)st.submitted));
  napi_set_named_property(env, o, "completed", num_value(env, (double)st.completed));
  napi_set_named_property(env, o, "failed", num_value(env, (double)st.failed));
  napi_set_named_property(env, o, "evictions", num_value(env, (double)st.evictions));
  napi_set_named_property(env, o, "residentMeshes", num_value(env, (double)st.resident_meshes));
  napi_set_named_property(env, o, "residentBytes", num_value(env, (double)st.resident_bytes));
  napi_set_named_property(env, o, "bytesInUse", num_value(env, (double)st.bytes_in_use));
  napi_set_named_property(env, o, "peakBytes", num_value(env, (double)st.peak_bytes));
  napi_set_named_property(env, o, "queued", num_value(env, st.queued));
  napi_set_named_property(env, o, "inFlight", num_value(env, st.in_flight));
  return o;
}

// ---- synchronize: Promise settled after the stream drains ---------------------

struct SyncWork {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  napi_ref ctx_ref = nullptr;  // keeps the context alive while waiting
  lbvh_ctx *ctx = nullptr;
  int status = LBVH_OK;
  std::string error;
};

void sync_execute(napi_env, void *data) {
  SyncWork *w = static_cast<SyncWork *>(data);
  w->status = lbvh_synchronize(w->ctx);
  if (w->status != LBVH_OK) w->error = lbvh_last_error(w->ctx);
}

void sync_complete(napi_env env, napi_status, void *data) {
  SyncWork *w = static_cast<SyncWork *>(data);
  if (w->status == LBVH_OK) {
    napi_resolve_deferred(env, w->deferred, undefined(env));
  } else {
    napi_reject_deferred(
        env, w->deferred,
        make_error(env, w->status,
                   std::string("lbvh_synchronize: ") +
                       lbvh_status_string(w->status) + " (" + w->error + ")"));
  }
  napi_delete_reference(env, w->ctx_ref);
  napi_delete_async_work(env, w->work);
  delete w;
}

napi_value Synchronize(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  lbvh_ctx *ctx;
  if (!get_args(env, info, 1, argv) || !get_ctx(env, argv[0], &ctx))
    return nullptr;
  SyncWork *w = new SyncWork();
  w->ctx = ctx;
  napi_value promise, name;
  NAPI_OK(env, napi_create_promise(env, &w->deferred, &promise));
  NAPI_OK(env, napi_create_reference(env, argv[0], 1, &w->ctx_ref));
  napi_create_string_utf8(env, "lbvh_synchronize", NAPI_AUTO_LENGTH, &name);
  NAPI_OK(env, napi_create_async_work(env, nullptr, name, sync_execute,
                                      sync_complete, w, &w->work));
  NAPI_OK(env, napi_queue_async_work(env, w->work));
  return promise;
}

napi_value ProfileEnable(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  lbvh_ctx *ctx;
  if (!get_args(env, info, 2, argv) || !get_ctx(env, argv[0], &ctx))
    return nullptr;
  bool on = false;
  napi_get_value_bool(env, argv[1], &on);
  if (throw_status(env, ctx, lbvh_profile_enable(ctx, on ? 1 : 0),
                   "lbvh_profile_enable"))
    return nullptr;
  return undefined(env);
}

napi_value ProfileRead(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  lbvh_ctx *ctx;
  if (!get_args(env, info, 1, argv) || !get_ctx(env, argv[0], &ctx))
    return nullptr;
  lbvh_profile p;
  if (throw_status(env, ctx, lbvh_profile_read(ctx, &p, 1),
                   "lbvh_profile_read"))
    return nullptr;
  static const char *names[LBVH_STAGE_COUNT] = {
      "bounds", "morton", "sort0", "sort1", "sort2", "sort3", "tree",
      "treeTop"};
  napi_value obj;
  napi_create_object(env, &obj);
  for (int s = 0; s < LBVH_STAGE_COUNT; ++s) {
    napi_value v;
    napi_create_double(env, p.ms[s], &v);
    napi_set_named_property(env, obj, names[s], v);
  }
  return obj;
}

napi_value Init(napi_env env, napi_value exports) {
  struct Fn {
    const char *name;
    napi_callback cb;
  } fns[] = {
      {"createContext", CreateContext}, {"setDivMode", SetDivMode},
      {"bytesInUse", BytesInUse},       {"alloc", Alloc},
      {"free", Free},                   {"upload", Upload},
      {"download", Download},           {"bounds", Bounds},
      {"aabbMorton", AabbMorton},       {"sort", Sort},
      {"radixTree", RadixTree},         {"upPass", UpPass},
      {"rearrange", Rearrange},         {"build", Build},
      {"refit", Refit},                 {"synchronize", Synchronize},
      {"profileEnable", ProfileEnable}, {"profileRead", ProfileRead},
      {"buildBatched", BuildBatched},   {"autotune", Autotune},
      {"raycast", Raycast},             {"createServer", CreateServer},
      {"serverSetDivMode", ServerSetDivMode},
      {"serverSetDigest", ServerSetDigest},
      {"serverCancel", ServerCancel},
      {"serverSetMaxWait", ServerSetMaxWait},
      {"serverSubmit", ServerSubmit},   {"serverPoll", ServerPoll},
      {"serverDrain", ServerDrain},     {"serverMesh", ServerMesh},
      {"serverEvict", ServerEvict},     {"serverStats", ServerStats},
  };
  for (const Fn &f : fns) {
    napi_value v;
    NAPI_OK(env, napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb,
                                      nullptr, &v));
    NAPI_OK(env, napi_set_named_property(env, exports, f.name, v));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
