"""lbvh_build_static, the build entry point that allocates nothing: its
output equals lbvh_build's on the same input, through the C ABI, on two
calls in a row (the second smaller, so stale scratch from the first must
not leak into it), with and without caller bounds, and it refuses sizes
beyond its static scratch."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def build_both(ctx, lb, tri, bounds):
    n = len(tri)
    d = ctx.to_device(tri)
    out = []
    try:
        for static in (False, True):
            bvh, otri = ctx.alloc(64 * n), ctx.alloc(48 * n)
            sk, pm = ctx.alloc(4 * n), ctx.alloc(4 * n)
            try:
                (ctx.build_static if static else ctx.build)(d, n, bounds, bvh, otri, sk, pm)
                ctx.synchronize()
                out.append((bvh.read(np.uint8, 64 * n), otri.read(np.uint32, 12 * n),
                            sk.read(np.uint32, n), pm.read(np.int32, n)))
            finally:
                for b in (bvh, otri, sk, pm):
                    b.destroy()
    finally:
        d.destroy()
    return out


@pytest.mark.parametrize("with_bounds", [False, True])
def test_static_build_equals_build_twice(lb, with_bounds):
    with lb.Context(0) as ctx:
        for n, seed in ((1 << 20, 11), (70001, 12), (1, 13)):
            say(f"static build n={n} bounds={with_bounds}")
            tri = lb.wl_clustered(seed, n, 0.05) if n > 1 else lb.wl_uniform(seed, n, 0.01)
            bounds = lb.wl_bounds(tri) if with_bounds else None
            (a, b) = build_both(ctx, lb, tri, bounds)
            for x, y in zip(a, b):
                assert np.array_equal(x, y), n


def test_static_build_refuses_sizes_beyond_its_scratch(lb):
    with lb.Context(0) as ctx:
        n = (1 << 20) + 1
        say(f"static build refusal n={n}")
        d = ctx.alloc(48 * n)
        bvh, otri = ctx.alloc(64 * n), ctx.alloc(48 * n)
        try:
            with pytest.raises(lb.LBVHError) as err:
                ctx.build_static(d, n, None, bvh, otri)
            assert err.value.code == lb.ERR_INVALID_ARGUMENT
            assert "1048576" in str(err.value)
        finally:
            for b in (d, bvh, otri):
                b.destroy()
