#!/usr/bin/env python3
"""LBVH build benchmark on one MI355X.

A step is one full lbvh_build (bounds -> Morton -> 4 onesweep passes -> fused
tree/up pass/rearrange) of the 16,777,216-triangle heightfield with the
triangles already resident in HBM.  Prints ONE JSON line (the last line of
stdout); progress goes to stderr.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--workload W]

Device time comes from HIP events recorded on the library's own stream.  The
per-stage breakdown (and the dominant kernel's roofline) comes from a second,
profiled timed region in which the library brackets each stage with events.
The cpu_baseline leg times the CPU oracle (the restatement of the reference
pipeline, test infrastructure) on the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "webgpu-lvbh-demo_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0

# Compulsory HBM bytes per triangle of each stage (N triangles per launch).
#   bounds   R triangles 48
#   morton   R triangles 48 + W key 4 (digit histograms are fused in)
#   sort0    R key 4 + W key, index 8 (the index is generated, not read)
#   sort1-3  R key, index 8 + W key, index 8
#   tree     R sorted key 4 + R index 4 + gather triangle 48 + W O_TRI 48
#            + W node 64 (topology, leaf boxes and node boxes in one pass)
STAGE_BYTES_PER_TRI = {"bounds": 48, "morton": 52, "sort0": 12, "sort1": 16,
                       "sort2": 16, "sort3": 16, "tree": 168}


def build_bytes(n: int) -> int:
    """Algorithmic bytes of one whole build: 380 N - 100 (minimal dataflow of
    the 8-bit-digit pipeline with bounds, counted per stage: bounds 48,
    Morton 52, sort 64, tree 12, leaf stage 104, internal nodes 100 each)."""
    return 380 * n - 100


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class Hip:
    """hipEvent_* through the HIP runtime the library itself links."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so", mode=C.RTLD_GLOBAL)
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def copy_d2d(self, dst, src, nbytes, stream):
        f = self.h.hipMemcpyAsync
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        assert f(C.c_void_p(dst), C.c_void_p(src), nbytes, 3, C.c_void_p(stream)) == 0

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def timed_builds(lb, ctx, hip, d_tri, n, bounds, bvh, otri, steps):
    """Times `steps` builds back to back on the library stream."""
    a, b = hip.event(), hip.event()
    ctx.synchronize()
    t0 = time.perf_counter()
    hip.record(a, ctx.stream)
    for _ in range(steps):
        ctx.build(d_tri, n, bounds, bvh, otri)
    hip.record(b, ctx.stream)
    ms = hip.elapsed_ms(a, b)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    return ms, wall


def cpu_baseline(tri, reps=2):
    """The CPU oracle (single-threaded C restatement) on the same mesh."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as ora
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        ora.build(tri, None, ora.DIV_MUL_RCP)
        times.append(time.perf_counter() - t)
    return min(times), times


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="heightfield-16m")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tree-tile", type=int, default=0,
                    help="leaves per block of the fused up pass (256/512/1024)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run device-to-device copies of MIB MiB, the "
                         "known-size reference for HBM traffic counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus "
              f"{args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import lbvh as lb
    hip = Hip()
    log(f"generating {args.workload}")
    tri = lb.workload(args.workload)
    n = tri.shape[0]
    ctx = lb.Context(0)
    if args.tree_tile:
        ctx.set_tuning(lb.TUNE_TREE_TILE, args.tree_tile)
    d_tri = ctx.to_device(tri)
    bvh = ctx.alloc(64 * n)
    otri = ctx.alloc(48 * n)
    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
        for _ in range(3):
            hip.copy_d2d(dst.ptr, src.ptr, nbytes, ctx.stream)
        ctx.synchronize()
        src.destroy()
        dst.destroy()
        log(f"copy calibration: 3 x {nbytes} bytes")
    log(f"warmup x{args.warmup}")
    for _ in range(args.warmup):
        ctx.build(d_tri, n, None, bvh, otri)
    ctx.synchronize()

    log(f"timed x{args.steps}")
    ms, wall = timed_builds(lb, ctx, hip, d_tri, n, None, bvh, otri, args.steps)
    ms_per_step = ms / args.steps

    log("profiled pass")
    ctx.profile_enable(True)
    ctx.profile_read(reset=True)
    pms, _ = timed_builds(lb, ctx, hip, d_tri, n, None, bvh, otri, args.steps)
    prof = ctx.profile_read(reset=True)
    ctx.profile_enable(False)
    stages = {}
    for s, (tot, cnt) in prof.items():
        if cnt:
            avg = tot / cnt
            gbs = STAGE_BYTES_PER_TRI[s] * n / (avg * 1e-3) / 1e9
            stages[s] = {"ms": round(avg, 4), "GB/s": round(gbs, 1),
                         "frac": round(gbs / HBM_PEAK_GBS, 4)}
    dom = max(stages, key=lambda s: stages[s]["ms"])
    dom_bytes = STAGE_BYTES_PER_TRI[dom] * n
    dom_gbs = dom_bytes / (stages[dom]["ms"] * 1e-3) / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(args.workload, {}).get(dom)

    pipe_bytes = build_bytes(n)
    pipe_gbs = pipe_bytes / (ms_per_step * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle)")
        best, times = cpu_baseline(tri)
        cpu = {"value": round(n / best / 1e6, 3), "unit": "Mtri/s", "cores": 1,
               "kind": "port",
               "sample": f"{len(times)} full single-threaded builds of {args.workload} "
                         f"({n} triangles) by the C restatement, best of "
                         f"{[round(t, 2) for t in times]} s"}

    result = {
        "metric": "LBVH build Mtri/s",
        "value": round(n / (ms_per_step * 1e-3) / 1e6, 1),
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+u32",
        "data": "synthetic",
        "config": {"workload": args.workload, "triangles": n,
                   "tree_tile": ctx.get_tuning(lb.TUNE_TREE_TILE),
                   "bounds": "GPU reduction", "div_mode": "mul_rcp",
                   "inputs": "resident in HBM"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(dom_gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom_gbs / HBM_PEAK_GBS, 4), "traffic": traffic},
        "pipeline_roofline": {"algorithmic_bytes": pipe_bytes,
                              "achieved": round(pipe_gbs, 1), "unit": "GB/s",
                              "frac_of_hbm_peak": round(pipe_gbs / HBM_PEAK_GBS, 4)},
        "stages": stages,
        "profiled_ms_per_step": round(pms / args.steps, 4),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
        "cpu_baseline": cpu,
    }
    for b in (d_tri, bvh, otri):
        b.destroy()
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
