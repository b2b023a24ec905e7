"""The Morton division's freedom, enumerated (aabb-z-idx.js:191-194).

The reference computes rel = (cen - min) / (max - min) in WGSL, where f32
division is only accurate to 2.5 ULP (WGSL floating-point accuracy rules);
the additions, the multiplication by 0.5 and by 1023.99f, and the min/max
are exact or correctly rounded.  So the reference's quantised coordinate
q = u32(rel * 1023.99f) is any value reachable from an f32 rel within
2.5 ULP of the exact quotient, and q is monotonic in rel: the two f32 ends
of that interval bound it.  A key whose three coordinates have equal ends
is fixed by the WGSL rules whatever the division's lowering; the others are
the keys on which parity with the reference is unpinned.

This file computes the interval in numpy (f64, independent of the oracle's
C code) and checks, per workload:
  * both oracle division modes give exactly the fixed key wherever the key
    is fixed (the GPU equals the oracle bit for bit in both modes, so the
    GPU equals the reference on every fixed key);
  * the two modes differ only on ambiguous keys;
  * the ambiguous keys are few: printed, and bounded by the interval's
    width (at most ~9.2e-4 of the keys).
"""
import numpy as np
import pytest

C1023 = np.float32(1023.99)


def f32_to_u32(v):
    """WGSL u32(f32): truncation, saturating; NaN to 0."""
    v = v.astype(np.float64)
    out = np.zeros(v.shape, np.uint64)
    pos = v > 0
    out[pos] = np.floor(np.minimum(v[pos], 4294967295.0)).astype(np.uint64)
    return out.astype(np.uint32)


def split_3(u):
    x = u.astype(np.uint32)
    x = (x | (x << np.uint32(16))) & np.uint32(0x030000FF)
    x = (x | (x << np.uint32(8))) & np.uint32(0x0300F00F)
    x = (x | (x << np.uint32(4))) & np.uint32(0x030C30C3)
    x = (x | (x << np.uint32(2))) & np.uint32(0x09249249)
    return x


def morton(q):
    return split_3(q[:, 0]) | (split_3(q[:, 1]) << np.uint32(1)) | (split_3(q[:, 2]) << np.uint32(2))


def f32_ceil(x):
    """Smallest f32 >= x (x f64)."""
    f = x.astype(np.float32)
    low = f.astype(np.float64) < x
    f[low] = np.nextafter(f[low], np.float32(np.inf))
    return f


def f32_floor(x):
    f = x.astype(np.float32)
    high = f.astype(np.float64) > x
    f[high] = np.nextafter(f[high], np.float32(-np.inf))
    return f


def key_interval(tri, bounds):
    """(fixed mask, key at the interval's low end, key at its high end)."""
    v = np.ascontiguousarray(tri, np.float32).reshape(-1, 3, 4)[:, :, :3]
    lo, hi = v.min(axis=1), v.max(axis=1)
    b = np.asarray(bounds, np.float32)
    cen = (hi + lo) * np.float32(0.5)       # f32, correctly rounded
    d = cen - b[:3]                          # f32 subtraction
    e = b[3:] - b[:3]
    assert np.all(e > 0)
    x = d.astype(np.float64) / e.astype(np.float64)   # the exact quotient to ~2^-53
    ax = np.abs(x)
    _, k = np.frexp(ax)                      # ax = m 2^k, m in [0.5, 1)
    ulp = np.ldexp(1.0, np.maximum(k - 24, -149))
    ulp[ax < 2.0 ** -126] = 2.0 ** -149
    # 2.5 ULP, widened by a hair so the f64 approximation of the quotient can
    # only make the interval larger (more keys counted as ambiguous)
    half = 2.5 * ulp * (1 + 1e-9) + 1e-300
    r_lo, r_hi = f32_ceil(x - half), f32_floor(x + half)
    q_lo = f32_to_u32(r_lo * C1023)
    q_hi = f32_to_u32(r_hi * C1023)
    fixed = np.all(q_lo == q_hi, axis=1)
    return fixed, morton(q_lo), morton(q_hi)


def workloads(lb):
    floor = lb.wl_uniform(77, 20000, 0.01)
    # the OBJ-floor case (parse-obj.js:65-78): two large triangles outside
    # the caller's bounds, so their coordinates saturate
    floor[:2, [0, 4, 8]] = [[-5, 5, 5], [-5, -5, 5]]
    floor[:2, [2, 6, 10]] = [[-5, -5, 5], [-5, 5, 5]]
    floor[:2, [1, 5, 9]] = -0.3
    return [
        ("uniform-64k", lambda: lb.workload("uniform-64k"), None),
        ("clustered-1m", lambda: lb.workload("clustered-1m"), None),
        ("heightfield-16m", lambda: lb.workload("heightfield-16m"), None),
        ("obj-floor", lambda: floor, lambda t: lb.wl_bounds(t[2:])),
    ]


@pytest.mark.parametrize("name", ["uniform-64k", "clustered-1m", "heightfield-16m", "obj-floor"])
def test_division_freedom_is_confined_to_ambiguous_keys(lb, ora, name):
    (make, bnd), = [(m, b) for w, m, b in workloads(lb) if w == name]
    tri = make()
    bounds = lb.wl_bounds(tri) if bnd is None else bnd(tri)
    fixed, k_lo, k_hi = key_interval(tri, bounds)
    _, keys_rcp = ora.aabb_z_idx(tri, bounds, ora.DIV_MUL_RCP)
    _, keys_ieee = ora.aabb_z_idx(tri, bounds, ora.DIV_IEEE)
    assert np.array_equal(k_lo[fixed], k_hi[fixed])
    assert np.array_equal(keys_rcp[fixed], k_lo[fixed])
    assert np.array_equal(keys_ieee[fixed], k_lo[fixed])
    differ = keys_rcp != keys_ieee
    assert not np.any(differ & fixed)
    amb = int((~fixed).sum())
    print(f"{name}: {len(tri)} keys, {amb} ambiguous under the 2.5-ULP division rule "
          f"({amb / len(tri):.2e}), {int(differ.sum())} where the two lowerings differ",
          flush=True)
    # per coordinate the interval is at most 5 * 2^-24 wide (rel < 1), so it
    # straddles a quantum boundary with probability <= 5 * 2^-24 * 1023.99
    # = 3.1e-4 for keys spread over the quanta: <= 9.2e-4 for three
    assert amb <= len(tri) // 1000 + 16
