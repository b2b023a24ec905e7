"""ctypes wrapper of liblbvh_oracle.so -- the CPU restatement of the
reference pipeline.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the oracle results are compared with.  The product library
(webgpu-lvbh-demo_amd/lib/liblbvh.so) never loads it.

Parity status of each function (how it is pinned, see tests/test_oracle.py):
  split_3 / morton_code  closed-form known answers from the masks
                         (aabb-z-idx.js:197-208), exhaustive 10-bit sweep
                         against an independent bit-interleave
  f32_to_u32             special cases of WGSL's saturating conversion
  aabb_z_idx             boxes: numpy min/max; keys: an independent numpy
                         float32 restatement in both division modes.  Which
                         division mode the reference's WebGPU stack uses is
                         PARITY UNPINNED (no WebGPU implementation exists in
                         this environment); both are provided.
  radix_sort             numpy stable argsort
  radix_tree             faithful WGSL loop (i32 wrap) == textbook loop ==
                         independent top-down brute force; structural
                         invariants
  up_pass                brute-force per-node range min/max, canonical order
  rearrange              numpy gather
  raycast                against raycast_bruteforce (every triangle, no BVH)
                         on hits/misses and closest distances
  build_mt               the multi-threaded CPU build (the CPU baseline):
                         byte-identical to build() on every output
  refit / refit_mt       brute force on jittered vertices; refit_mt
                         byte-identical to refit
  digest                 an independent numpy restatement of the weighted
                         word sums; a single changed word changes it
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbvh_oracle.so")

DIV_MUL_RCP = 0
DIV_IEEE = 1

_lib = None


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        L = C.CDLL(LIB_PATH)
        vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
        sig = {
            "oracle_split_3": (u32, [u32]),
            "oracle_morton_code": (u32, [u32, u32, u32]),
            "oracle_f32_to_u32": (u32, [C.c_float]),
            "oracle_dif": (C.c_int32, [u32, u32, i32]),
            "oracle_bounds": (None, [vp, u32, vp]),
            "oracle_aabb_z_idx": (None, [vp, u32, vp, i32, vp, vp]),
            "oracle_radix_sort": (None, [vp, vp, u32]),
            "oracle_radix_tree": (None, [vp, u32, vp, i32]),
            "oracle_up_pass": (None, [vp, vp, vp, u32, vp]),
            "oracle_rearrange": (None, [vp, vp, u32, vp]),
            "oracle_canonicalize_bvh": (None, [vp, u32]),
            "oracle_build": (None, [vp, u32, vp, i32, vp, vp, vp, vp]),
            "oracle_build_mt": (i32, [vp, u32, vp, i32, vp, vp, vp, vp]),
            "oracle_refit": (None, [vp, u32, vp, vp, vp, vp]),
            "oracle_refit_mt": (None, [vp, u32, vp, vp, vp, vp]),
            "oracle_digest": (None, [vp, vp, u32, vp]),
            "oracle_raycast": (None, [vp, vp, u32, vp, u32, vp]),
            "oracle_raycast_bruteforce": (None, [vp, u32, vp, u32, vp]),
            "oracle_bruteforce_tree": (None, [vp, u32, vp]),
            "oracle_bruteforce_bvh": (None, [vp, vp, vp, u32, vp]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _lib = L
    return _lib


def _tri(tri):
    return np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 12)


def split_3(u: int) -> int:
    return lib().oracle_split_3(u)


def morton_code(x: int, y: int, z: int) -> int:
    return lib().oracle_morton_code(x, y, z)


def f32_to_u32(v: float) -> int:
    return lib().oracle_f32_to_u32(v)


def dif(a: int, b: int, faithful: bool) -> int:
    return lib().oracle_dif(a, b, int(faithful))


def bounds(tri) -> np.ndarray:
    tri = _tri(tri)
    out = np.empty(6, np.float32)
    lib().oracle_bounds(_p(tri), tri.shape[0], _p(out))
    return out


def aabb_z_idx(tri, bnds, div_mode=DIV_MUL_RCP):
    tri = _tri(tri)
    n = tri.shape[0]
    b = np.ascontiguousarray(bnds, dtype=np.float32)
    aabb = np.empty((n, 8), np.float32)
    keys = np.empty(n, np.uint32)
    lib().oracle_aabb_z_idx(_p(tri), n, _p(b), div_mode, _p(aabb), _p(keys))
    return aabb, keys


def radix_sort(keys):
    k = np.array(keys, dtype=np.uint32, copy=True)
    idx = np.empty(k.shape[0], np.int32)
    lib().oracle_radix_sort(_p(k), _p(idx), k.shape[0])
    return k, idx


def radix_tree(skeys, faithful=False):
    k = np.ascontiguousarray(skeys, dtype=np.uint32)
    par = np.empty(2 * k.shape[0], np.int32)
    lib().oracle_radix_tree(_p(k), k.shape[0], _p(par), int(faithful))
    return par


def up_pass(idx, aabb, parents, canonical=True):
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    aabb = np.ascontiguousarray(aabb, dtype=np.float32)
    par = np.ascontiguousarray(parents, dtype=np.int32)
    n = idx.shape[0]
    out = np.zeros(64 * n, np.uint8)
    lib().oracle_up_pass(_p(idx), _p(aabb), _p(par), n, _p(out))
    if canonical:
        lib().oracle_canonicalize_bvh(_p(out), n)
    return out


def canonicalize(bvh: np.ndarray, n: int) -> np.ndarray:
    out = np.array(bvh, dtype=np.uint8, copy=True)
    lib().oracle_canonicalize_bvh(_p(out), n)
    return out


def rearrange(tri, idx):
    tri = _tri(tri)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    out = np.empty((idx.shape[0], 12), np.float32)
    lib().oracle_rearrange(_p(tri), _p(idx), idx.shape[0], _p(out))
    return out


def build(tri, bnds=None, div_mode=DIV_MUL_RCP):
    """Returns (bvh uint8[64n] canonical, otri float32[n,12], sorted keys,
    permutation)."""
    tri = _tri(tri)
    n = tri.shape[0]
    b = None if bnds is None else np.ascontiguousarray(bnds, dtype=np.float32)
    bvh = np.zeros(64 * n, np.uint8)
    otri = np.empty((n, 12), np.float32)
    sk = np.empty(n, np.uint32)
    pm = np.empty(n, np.int32)
    lib().oracle_build(_p(tri), n, _p(b), div_mode, _p(bvh), _p(otri), _p(sk), _p(pm))
    return bvh, otri, sk, pm


def build_mt(tri, bnds=None, div_mode=DIV_MUL_RCP):
    """The multi-threaded CPU build (same outputs as build()); returns
    (bvh, otri, sorted keys, permutation, threads used)."""
    tri = _tri(tri)
    n = tri.shape[0]
    b = None if bnds is None else np.ascontiguousarray(bnds, dtype=np.float32)
    bvh = np.zeros(64 * n, np.uint8)
    otri = np.empty((n, 12), np.float32)
    sk = np.empty(n, np.uint32)
    pm = np.empty(n, np.int32)
    t = lib().oracle_build_mt(_p(tri), n, _p(b), div_mode, _p(bvh), _p(otri), _p(sk), _p(pm))
    return bvh, otri, sk, pm, int(t)


def refit(tri, skeys, perm):
    tri = _tri(tri)
    n = tri.shape[0]
    sk = np.ascontiguousarray(skeys, dtype=np.uint32)
    pm = np.ascontiguousarray(perm, dtype=np.int32)
    bvh = np.zeros(64 * n, np.uint8)
    otri = np.empty((n, 12), np.float32)
    lib().oracle_refit(_p(tri), n, _p(sk), _p(pm), _p(bvh), _p(otri))
    return bvh, otri


def refit_mt(tri, skeys, perm):
    """refit() on all host cores (OpenMP), byte-identical."""
    tri = _tri(tri)
    n = tri.shape[0]
    sk = np.ascontiguousarray(skeys, dtype=np.uint32)
    pm = np.ascontiguousarray(perm, dtype=np.int32)
    bvh = np.zeros(64 * n, np.uint8)
    otri = np.empty((n, 12), np.float32)
    lib().oracle_refit_mt(_p(tri), n, _p(sk), _p(pm), _p(bvh), _p(otri))
    return bvh, otri


def digest(bvh, otri):
    """(BVH digest, O_TRI digest) of a build's outputs, as the serving loop
    reports them per request."""
    b = np.ascontiguousarray(bvh).view(np.uint8)
    t = np.ascontiguousarray(otri, dtype=np.float32).reshape(-1, 12)
    n = t.shape[0]
    assert b.size == 64 * n
    out = np.zeros(2, np.uint64)
    lib().oracle_digest(_p(b), _p(t), n, _p(out))
    return int(out[0]), int(out[1])


def bruteforce_tree(skeys):
    k = np.ascontiguousarray(skeys, dtype=np.uint32)
    par = np.empty(2 * k.shape[0], np.int32)
    lib().oracle_bruteforce_tree(_p(k), k.shape[0], _p(par))
    return par


def bruteforce_bvh(tri, perm, skeys):
    tri = _tri(tri)
    pm = np.ascontiguousarray(perm, dtype=np.int32)
    k = np.ascontiguousarray(skeys, dtype=np.uint32)
    n = k.shape[0]
    out = np.zeros(64 * n, np.uint8)
    lib().oracle_bruteforce_bvh(_p(tri), _p(pm), _p(k), n, _p(out))
    return out


def raycast(bvh, otri, rays):
    """Closest hits through the BVH, path-tracer traversal (n x 4 floats)."""
    b = np.ascontiguousarray(bvh, dtype=np.uint8)
    t = _tri(otri)
    r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    out = np.empty((r.shape[0], 4), np.float32)
    lib().oracle_raycast(_p(b), _p(t), t.shape[0], _p(r), r.shape[0], _p(out))
    return out


def raycast_bruteforce(tri, rays):
    t = _tri(tri)
    r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    out = np.empty((r.shape[0], 4), np.float32)
    lib().oracle_raycast_bruteforce(_p(t), t.shape[0], _p(r), r.shape[0], _p(out))
    return out


def bvh_fields(bvh: np.ndarray, n: int):
    """Views a BVH byte buffer as (float32 [n,16], int32 [n,16])."""
    b = np.ascontiguousarray(bvh, dtype=np.uint8)
    return b.view(np.float32).reshape(n, 16), b.view(np.int32).reshape(n, 16)
