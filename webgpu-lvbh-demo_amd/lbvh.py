"""ctypes binding of liblbvh.so (include/lbvh.h) and liblbvh_workloads.so.

Used by the Python tests and bench.py; the Node.js entry point
(bvh/index.js) binds the same C ABI through the N-API addon.

The binding is a thin mirror of the C ABI: every call goes to the HIP
library; if the library is missing, loading raises instead of falling back
to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_PATH = os.path.join(LIB_DIR, "liblbvh.so")
WL_PATH = os.path.join(LIB_DIR, "liblbvh_workloads.so")

DIV_MUL_RCP = 0
DIV_IEEE = 1

TUNE_TREE_TILE = 0
TUNE_SORT_BITS = 1

# lbvh_status (include/lbvh.h)
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_SIZE_MISMATCH = 2
ERR_OUT_OF_BUDGET = 3
ERR_OUT_OF_MEMORY = 4
ERR_DEVICE = 5
ERR_NO_DEVICE = 6
ERR_BUSY = 7
ERR_NOT_FOUND = 8
ERR_CANCELLED = 9
ERR_TIMEOUT = 10

STATUS = {
    0: "ok",
    1: "invalid argument",
    2: "size mismatch",
    3: "out of HBM budget",
    4: "out of device memory",
    5: "device error",
    6: "no such device",
    7: "busy",
    8: "not found",
    9: "cancelled",
    10: "timed out",
}

TRI_BYTES = 48
NODE_BYTES = 64
AABB_BYTES = 32

# Every symbol include/lbvh.h declares, with (restype, argtypes).
_vp, _u32, _u64, _i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
SIGNATURES = {
    "lbvh_create": (_i32, [_i32, _u64, C.POINTER(_vp)]),
    "lbvh_destroy": (_i32, [_vp]),
    "lbvh_last_error": (C.c_char_p, [_vp]),
    "lbvh_status_string": (C.c_char_p, [_i32]),
    "lbvh_set_div_mode": (_i32, [_vp, _i32]),
    "lbvh_get_div_mode": (_i32, [_vp]),
    "lbvh_set_tuning": (_i32, [_vp, _i32, _i32]),
    "lbvh_get_tuning": (_i32, [_vp, _i32]),
    "lbvh_get_tuning_for": (_i32, [_vp, _i32, _u32]),
    "lbvh_autotune": (_i32, [_vp, _u32, _u32, _i32, _vp]),
    "lbvh_stream": (_vp, [_vp]),
    "lbvh_synchronize": (_i32, [_vp]),
    "lbvh_bytes_in_use": (_u64, [_vp]),
    "lbvh_trim": (_i32, [_vp]),
    "lbvh_alloc": (_i32, [_vp, _u64, C.POINTER(_vp)]),
    "lbvh_free": (_i32, [_vp, _vp]),
    "lbvh_upload": (_i32, [_vp, _vp, _vp, _u64]),
    "lbvh_download": (_i32, [_vp, _vp, _vp, _u64]),
    "lbvh_bounds": (_i32, [_vp, _vp, _u32, _vp]),
    "lbvh_aabb_morton": (_i32, [_vp, _vp, _u32, _vp, _vp, _vp]),
    "lbvh_sort": (_i32, [_vp, _vp, _vp, _u32]),
    "lbvh_radix_tree": (_i32, [_vp, _vp, _u32, _vp]),
    "lbvh_up_pass": (_i32, [_vp, _vp, _vp, _vp, _u32, _vp]),
    "lbvh_rearrange": (_i32, [_vp, _vp, _vp, _u32, _vp]),
    "lbvh_build": (_i32, [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp]),
    "lbvh_build_static": (_i32, [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp]),
    "lbvh_refit": (_i32, [_vp, _vp, _u32, _vp, _vp, _vp, _vp]),
    "lbvh_raycast": (_i32, [_vp, _vp, _vp, _u32, _vp, _u32, _vp]),
    "lbvh_build_batched": (_i32, [_vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp]),
    "lbvh_server_create": (_i32, [_i32, _u64, _u32, C.POINTER(_vp)]),
    "lbvh_server_destroy": (_i32, [_vp]),
    "lbvh_server_set_div_mode": (_i32, [_vp, _i32]),
    "lbvh_server_set_digest": (_i32, [_vp, _i32]),
    "lbvh_server_cancel": (_i32, [_vp, _u64]),
    "lbvh_server_set_max_wait": (_i32, [_vp, C.c_double]),
    "lbvh_server_submit": (_i32, [_vp, _i32, _u64, _i32, _vp, _u32, _u64, C.c_double,
                                  C.POINTER(_u64)]),
    "lbvh_server_poll": (_i32, [_vp, _vp, _u32, C.POINTER(_u32)]),
    "lbvh_server_drain": (_i32, [_vp]),
    "lbvh_server_mesh": (_i32, [_vp, _u64, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_u32)]),
    "lbvh_server_evict": (_i32, [_vp, _u64]),
    "lbvh_server_stats_get": (_i32, [_vp, _vp]),
    "lbvh_server_last_error": (C.c_char_p, [_vp]),
    "lbvh_profile_enable": (_i32, [_vp, _i32]),
    "lbvh_profile_read": (_i32, [_vp, _vp, _i32]),
}

STAGES = ["bounds", "morton", "sort0", "sort1", "sort2", "sort3", "tree",
          "tree_top"]


class Profile(C.Structure):
    _fields_ = [("ms", C.c_double * len(STAGES)),
                ("launches", C.c_uint64 * len(STAGES))]

WL_SIGNATURES = {
    "lbvh_wl_uniform": (None, [_u64, _u32, C.c_double, _vp]),
    "lbvh_wl_clustered": (None, [_u64, _u32, C.c_double, _vp]),
    "lbvh_wl_heightfield": (None, [_u64, _u32, _u32, _vp]),
    "lbvh_wl_log_uniform_sizes": (None, [_u64, _u32, _u32, _u32, _vp]),
    "lbvh_wl_jitter": (None, [_u64, _u32, C.c_double, _vp, _vp]),
    "lbvh_wl_bounds": (None, [_vp, _u32, _vp]),
    "lbvh_wl_uniform01": (None, [_u64, _u64, _u32, _vp]),
}

_lib = None
_wl = None


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load_library() -> C.CDLL:
    """Loads liblbvh.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it first (python -c "
                f"'import __graft_entry__ as g; g.build()')")
        _lib = _bind(C.CDLL(LIB_PATH), SIGNATURES)
    return _lib


def load_workloads() -> C.CDLL:
    global _wl
    if _wl is None:
        if not os.path.exists(WL_PATH):
            raise RuntimeError(f"{WL_PATH} is missing: build it first")
        _wl = _bind(C.CDLL(WL_PATH), WL_SIGNATURES)
    return _wl


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


class LBVHError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class DeviceBuffer:
    """Device allocation from lbvh_alloc (GPUBuffer stand-in)."""

    def __init__(self, ctx: "Context", ptr: int, size: int):
        self.ctx, self.ptr, self.size = ctx, ptr, size

    def destroy(self):
        if self.ptr:
            self.ctx._check(self.ctx.lib.lbvh_free(self.ctx.handle, self.ptr))
            self.ptr = 0

    def read(self, dtype=np.uint8, count: Optional[int] = None) -> np.ndarray:
        dt = np.dtype(dtype)
        if count is None:
            count = self.size // dt.itemsize
        out = np.empty(count, dtype=dt)
        self.ctx.download(out, self)
        return out

    def write(self, arr: np.ndarray):
        self.ctx.upload(self, arr)


def _dp(b) -> Optional[int]:
    if b is None:
        return None
    return b.ptr if isinstance(b, DeviceBuffer) else int(b)


class Context:
    """One lbvh_ctx: a HIP stream, scratch arena and HBM budget."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.lbvh_create(device, hbm_budget, C.byref(h))
        if st != 0:
            raise LBVHError(st, f"lbvh_create(device={device})")
        self.handle = h.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.lbvh_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- plumbing --------------------------------------------------------
    def _check(self, st: int):
        if st != 0:
            raise LBVHError(st, self.lib.lbvh_last_error(self.handle).decode())

    def last_error(self) -> str:
        return self.lib.lbvh_last_error(self.handle).decode()

    @property
    def stream(self) -> int:
        return self.lib.lbvh_stream(self.handle)

    def synchronize(self):
        self._check(self.lib.lbvh_synchronize(self.handle))

    def bytes_in_use(self) -> int:
        return self.lib.lbvh_bytes_in_use(self.handle)

    def trim(self):
        self._check(self.lib.lbvh_trim(self.handle))

    def set_div_mode(self, mode: int):
        self._check(self.lib.lbvh_set_div_mode(self.handle, mode))

    def set_tuning(self, param: int, value: int):
        self._check(self.lib.lbvh_set_tuning(self.handle, param, value))

    def get_tuning(self, param: int) -> int:
        return self.lib.lbvh_get_tuning(self.handle, param)

    def get_tuning_for(self, param: int, n: int) -> int:
        return self.lib.lbvh_get_tuning_for(self.handle, param, n)

    TUNE_TILES = (256, 512, 1024)
    TUNE_BITS = (8, 10)

    def autotune(self, n_min: int, n_max: int, reps: int = 3) -> dict:
        """Runs lbvh_autotune; returns {bucket size: {(tile, digit bits): ms}}."""
        b0, b1 = n_min.bit_length() - 1, n_max.bit_length() - 1
        variants = [(t, b) for t in self.TUNE_TILES for b in self.TUNE_BITS]
        ms = np.zeros((b1 - b0 + 1) * len(variants), np.float32)
        self._check(self.lib.lbvh_autotune(self.handle, n_min, n_max, reps, _ptr(ms)))
        ms = ms.reshape(-1, len(variants))
        return {1 << (b0 + i): dict(zip(variants, map(float, ms[i])))
                for i in range(ms.shape[0])}

    def profile_enable(self, on: bool = True):
        self._check(self.lib.lbvh_profile_enable(self.handle, int(on)))

    def profile_read(self, reset: bool = True) -> dict:
        """Per-stage device time: {stage: (total_ms, launches)}."""
        p = Profile()
        self._check(self.lib.lbvh_profile_read(self.handle, C.byref(p), int(reset)))
        return {s: (p.ms[i], int(p.launches[i])) for i, s in enumerate(STAGES)}

    def alloc(self, nbytes: int) -> DeviceBuffer:
        p = C.c_void_p()
        self._check(self.lib.lbvh_alloc(self.handle, nbytes, C.byref(p)))
        return DeviceBuffer(self, p.value or 0, nbytes)

    def upload(self, buf, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        self._check(self.lib.lbvh_upload(self.handle, _dp(buf), _ptr(arr), arr.nbytes))

    def download(self, out: np.ndarray, buf):
        self._check(self.lib.lbvh_download(self.handle, _ptr(out), _dp(buf), out.nbytes))

    def to_device(self, arr: np.ndarray) -> DeviceBuffer:
        arr = np.ascontiguousarray(arr)
        b = self.alloc(arr.nbytes)
        if arr.nbytes:
            self.upload(b, arr)
        return b

    # ---- per-stage entry points (device buffers) ---------------------------
    def bounds(self, tri, n: int, out):
        self._check(self.lib.lbvh_bounds(self.handle, _dp(tri), n, _dp(out)))

    def aabb_morton(self, tri, n: int, bounds: Optional[Sequence[float]], aabb, keys):
        b = None if bounds is None else np.asarray(bounds, dtype=np.float32)
        self._check(self.lib.lbvh_aabb_morton(self.handle, _dp(tri), n, _ptr(b),
                                              _dp(aabb), _dp(keys)))

    def sort(self, keys, idx, n: int):
        self._check(self.lib.lbvh_sort(self.handle, _dp(keys), _dp(idx), n))

    def radix_tree(self, skeys, n: int, parents):
        self._check(self.lib.lbvh_radix_tree(self.handle, _dp(skeys), n, _dp(parents)))

    def up_pass(self, idx, aabb, parents, n: int, bvh):
        self._check(self.lib.lbvh_up_pass(self.handle, _dp(idx), _dp(aabb),
                                          _dp(parents), n, _dp(bvh)))

    def rearrange(self, tri, idx, n: int, out):
        self._check(self.lib.lbvh_rearrange(self.handle, _dp(tri), _dp(idx), n, _dp(out)))

    def build(self, tri, n: int, bounds: Optional[Sequence[float]], bvh, otri,
              skeys=None, perm=None):
        b = None if bounds is None else np.asarray(bounds, dtype=np.float32)
        self._check(self.lib.lbvh_build(self.handle, _dp(tri), n, _ptr(b), _dp(bvh),
                                        _dp(otri), _dp(skeys), _dp(perm)))

    def build_static(self, tri, n: int, bounds: Optional[Sequence[float]], bvh, otri,
                     skeys=None, perm=None):
        """lbvh_build_static: lbvh_build without allocating (n <= 2^20)."""
        b = None if bounds is None else np.asarray(bounds, dtype=np.float32)
        self._check(self.lib.lbvh_build_static(self.handle, _dp(tri), n, _ptr(b), _dp(bvh),
                                               _dp(otri), _dp(skeys), _dp(perm)))

    def build_batched(self, tri, offsets: np.ndarray, bounds, bvh, otri,
                      skeys=None, perm=None):
        off = np.ascontiguousarray(offsets, dtype=np.uint32)
        b = None if bounds is None else np.ascontiguousarray(bounds, dtype=np.float32)
        self._check(self.lib.lbvh_build_batched(
            self.handle, _dp(tri), _ptr(off), len(off) - 1, _ptr(b), _dp(bvh),
            _dp(otri), _dp(skeys), _dp(perm)))

    def refit(self, tri, n: int, skeys, perm, bvh, otri):
        self._check(self.lib.lbvh_refit(self.handle, _dp(tri), n, _dp(skeys),
                                        _dp(perm), _dp(bvh), _dp(otri)))

    def raycast(self, bvh, tri, n: int, rays, count: int, hits):
        self._check(self.lib.lbvh_raycast(self.handle, _dp(bvh), _dp(tri), n,
                                          _dp(rays), count, _dp(hits)))

    def raycast_arrays(self, bvh: np.ndarray, otri: np.ndarray, rays: np.ndarray) -> np.ndarray:
        """Host arrays in, closest hits (count x 4 float32) out."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        otri = np.ascontiguousarray(otri, dtype=np.float32).reshape(-1, 12)
        n, count = otri.shape[0], rays.shape[0]
        db, dt, dr = self.to_device(bvh), self.to_device(otri), self.to_device(rays)
        dh = self.alloc(16 * count)
        try:
            self.raycast(db, dt, n, dr, count, dh)
            return dh.read(np.float32, 4 * count).reshape(count, 4)
        finally:
            for b in (db, dt, dr, dh):
                b.destroy()

    # ---- host-array conveniences (tests) -----------------------------------
    def build_arrays(self, tri: np.ndarray, bounds=None, keep_topology=False):
        """Uploads `tri` (n x 12 float32), runs lbvh_build, returns host arrays
        (bvh bytes as uint8 [n*64], otri float32 [n,12], and optionally the
        sorted keys and permutation)."""
        tri = np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 12)
        n = tri.shape[0]
        dtri = self.to_device(tri)
        bvh = self.alloc(NODE_BYTES * n)
        otri = self.alloc(TRI_BYTES * n)
        sk = self.alloc(4 * n) if keep_topology else None
        pm = self.alloc(4 * n) if keep_topology else None
        try:
            self.build(dtri, n, bounds, bvh, otri, sk, pm)
            out = [bvh.read(np.uint8, NODE_BYTES * n),
                   otri.read(np.float32, 12 * n).reshape(n, 12)]
            if keep_topology:
                out += [sk.read(np.uint32, n), pm.read(np.int32, n)]
            return tuple(out)
        finally:
            for b in (dtri, bvh, otri, sk, pm):
                if b is not None:
                    b.destroy()


# ---- serving -----------------------------------------------------------------

REQ_BUILD, REQ_REFIT = 0, 1
SRC_HOST, SRC_DEVICE, SRC_UNIFORM, SRC_JITTER = 0, 1, 2, 3


class Completion(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("mesh_id", C.c_uint64),
                ("kind", C.c_int), ("status", C.c_int), ("n", C.c_uint32),
                ("slot", C.c_uint32), ("submit_s", C.c_double),
                ("start_s", C.c_double), ("done_s", C.c_double),
                ("digest_bvh", C.c_uint64), ("digest_tri", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64),
                ("failed", C.c_uint64), ("evictions", C.c_uint64),
                ("resident_meshes", C.c_uint64), ("resident_bytes", C.c_uint64),
                ("bytes_in_use", C.c_uint64), ("peak_bytes", C.c_uint64),
                ("queued", C.c_uint32), ("in_flight", C.c_uint32)]


class Server:
    """lbvh_server: device-resident build/refit service."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, slots: int = 4):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.lbvh_server_create(device, hbm_budget, slots, C.byref(h))
        if st != 0:
            raise LBVHError(st, "lbvh_server_create")
        self.handle = h.value
        self._buf = (Completion * 1024)()

    def close(self):
        if getattr(self, "handle", None):
            self.lib.lbvh_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st != 0:
            raise LBVHError(st, self.lib.lbvh_server_last_error(self.handle).decode())

    def set_div_mode(self, mode: int):
        self._check(self.lib.lbvh_server_set_div_mode(self.handle, mode))

    def cancel(self, request_id: int):
        """Cancels a request still waiting for admission (LBVHError BUSY if
        admitted, NOT_FOUND if completed)."""
        self._check(self.lib.lbvh_server_cancel(self.handle, request_id))

    def set_max_wait(self, seconds: float):
        """Later requests not admitted within `seconds` time out (0: none)."""
        self._check(self.lib.lbvh_server_set_max_wait(self.handle, seconds))

    def set_digest(self, enable: bool = True):
        """Requests admitted from now on report digests of their results."""
        self._check(self.lib.lbvh_server_set_digest(self.handle, 1 if enable else 0))

    def submit(self, kind: int, mesh_id: int, source: int, n: int, tri=None,
               seed: int = 0, param: float = 0.0) -> int:
        rid = C.c_uint64()
        ptr = None
        if source == SRC_HOST:
            tri = np.ascontiguousarray(tri, dtype=np.float32)
            ptr = _ptr(tri)
        elif source == SRC_DEVICE:
            ptr = _dp(tri)
        self._check(self.lib.lbvh_server_submit(self.handle, kind, mesh_id, source, ptr,
                                                n, seed, param, C.byref(rid)))
        return rid.value

    def poll(self):
        got = C.c_uint32()
        out = []
        while True:
            self._check(self.lib.lbvh_server_poll(self.handle, self._buf, 1024, C.byref(got)))
            out += [dict((f, getattr(self._buf[i], f)) for f, _ in Completion._fields_)
                    for i in range(got.value)]
            if got.value < 1024:
                return out

    def drain(self):
        self._check(self.lib.lbvh_server_drain(self.handle))

    def mesh(self, mesh_id: int):
        """(bvh device ptr, O_TRI device ptr, n) of a resident mesh."""
        b, t, n = C.c_void_p(), C.c_void_p(), C.c_uint32()
        self._check(self.lib.lbvh_server_mesh(self.handle, mesh_id, C.byref(b),
                                              C.byref(t), C.byref(n)))
        return b.value, t.value, n.value

    def evict(self, mesh_id: int):
        self._check(self.lib.lbvh_server_evict(self.handle, mesh_id))

    def stats(self) -> dict:
        st = ServerStats()
        self._check(self.lib.lbvh_server_stats_get(self.handle, C.byref(st)))
        return dict((f, getattr(st, f)) for f, _ in ServerStats._fields_)

    def read_mesh(self, ctx: "Context", mesh_id: int):
        """Host copies of a resident mesh's BVH bytes and O_TRI."""
        b, t, n = self.mesh(mesh_id)
        bvh = np.empty(64 * n, np.uint8)
        otri = np.empty((n, 12), np.float32)
        ctx.download(bvh, b)
        ctx.download(otri, t)
        return bvh, otri


# ---- workloads -------------------------------------------------------------

def wl_uniform(seed: int, n: int, half_extent: float = 0.01) -> np.ndarray:
    out = np.empty((n, 12), dtype=np.float32)
    load_workloads().lbvh_wl_uniform(seed, n, half_extent, _ptr(out))
    return out


def wl_clustered(seed: int, n: int, dup_fraction: float = 0.05) -> np.ndarray:
    out = np.empty((n, 12), dtype=np.float32)
    load_workloads().lbvh_wl_clustered(seed, n, dup_fraction, _ptr(out))
    return out


def wl_heightfield(seed: int, nx: int, ny: int) -> np.ndarray:
    out = np.empty((2 * nx * ny, 12), dtype=np.float32)
    load_workloads().lbvh_wl_heightfield(seed, nx, ny, _ptr(out))
    return out


def wl_log_uniform_sizes(seed: int, count: int, lo: int, hi: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    load_workloads().lbvh_wl_log_uniform_sizes(seed, count, lo, hi, _ptr(out))
    return out


def wl_jitter(seed: int, tri: np.ndarray, sigma: float) -> np.ndarray:
    tri = np.ascontiguousarray(tri, dtype=np.float32)
    out = np.empty_like(tri)
    load_workloads().lbvh_wl_jitter(seed, tri.shape[0], sigma, _ptr(tri), _ptr(out))
    return out


def wl_bounds(tri: np.ndarray) -> np.ndarray:
    tri = np.ascontiguousarray(tri, dtype=np.float32)
    out = np.empty(6, dtype=np.float32)
    load_workloads().lbvh_wl_bounds(_ptr(tri), tri.shape[0], _ptr(out))
    return out


def wl_uniform01(seed: int, first: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    load_workloads().lbvh_wl_uniform01(seed, first, count, _ptr(out))
    return out


# Named workloads of the benchmark configurations.
PARITY_SEED = 1
CLUSTERED_SEED = 2
HEIGHTFIELD_SEED = 3
BATCH_SEED = 4


def wl_serving_plan(rng, requests: int, size_seed: int = 6, lo: int = 1024,
                    hi: int = 4194304):
    """The serving workload's request plan: sizes log-uniform in [lo, hi]
    (seeded), 70% refits of one of the 32 most recently built meshes
    (vertices + N(0, 0.001) drawn on the device), 30% builds of a new
    uniform mesh synthesised on the device.  Returns a list of (kind,
    mesh id, source, n, seed, param); `rng` (numpy Generator) draws the mix
    and the refit targets."""
    sizes = wl_log_uniform_sizes(size_seed, requests, lo, hi)
    is_refit = rng.random(requests) < 0.7
    plan, built, next_mesh = [], [], 0
    for i in range(requests):
        if is_refit[i] and built:
            m, n = built[int(rng.integers(max(0, len(built) - 32), len(built)))]
            plan.append((REQ_REFIT, m, SRC_JITTER, n, 1000 + i, 0.001))
        else:
            n = int(sizes[i])
            plan.append((REQ_BUILD, next_mesh, SRC_UNIFORM, n, 2000 + i, 0.01))
            built.append((next_mesh, n))
            next_mesh += 1
    return plan


def workload(name: str) -> np.ndarray:
    if name == "uniform-64k":
        return wl_uniform(PARITY_SEED, 65536, 0.01)
    if name == "clustered-1m":
        return wl_clustered(CLUSTERED_SEED, 1 << 20, 0.05)
    if name == "heightfield-16m":
        return wl_heightfield(HEIGHTFIELD_SEED, 4096, 2048)
    raise KeyError(name)
