"""Autotuner for the sweep's tile shape.

Times the level sweep of a representative level for each candidate pair of
tile caps (txmax, tymax) and records the fastest in tuning.json, keyed by
solver and precision.  The tile caps fix the LDS footprint of a workgroup
and with it how many workgroups (waves) share a CU, so this also picks the
waves per CU.

    python -m amrclaw_amd.autotune [--precision mixed|fp32|fp64] [--npatch N] [--out PATH]
    python -m amrclaw_amd.autotune --workload euler3d   # plane block of the 3-D step
"""
import argparse
import json
import os
import time

import numpy as np

from . import capi, workload

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuning.json")

CANDIDATES = {
    "fp32": [(12, 12), (28, 12), (12, 28), (28, 20), (20, 28), (28, 28), (44, 12),
             (44, 20), (60, 12), (60, 20)],
    "fp64": [(12, 12), (28, 12), (12, 28), (20, 20), (28, 20), (44, 12), (60, 12)],
}
# the mixed path's tiles hold FP32 copies, like the FP32 path's
CANDIDATES["mixed"] = CANDIDATES["fp32"]
PLANE_BLOCKS = [2, 3, 4, 6, 8]
PRECISIONS = {"fp32": capi.FP32, "fp64": capi.FP64, "mixed": capi.MIXED}


def load(solver, precision):
    """Tuned (txmax, tymax) for a solver/precision, or None."""
    if not os.path.exists(CACHE):
        return None
    with open(CACHE) as f:
        d = json.load(f)
    v = d.get("%d/%s" % (solver, precision))
    return tuple(v["tiling"]) if v else None


def load_plane_block(precision):
    """Tuned plane block of the 3-D step, or None."""
    if not os.path.exists(CACHE):
        return None
    with open(CACHE) as f:
        d = json.load(f)
    v = d.get("3d/%d/%s" % (capi.EULER, precision))
    return int(v["plane_block"]) if v else None


def _store(key, value, out):
    d = {}
    if os.path.exists(out):
        with open(out) as f:
            d = json.load(f)
    d[key] = value
    with open(out, "w") as f:
        json.dump(d, f, indent=1)


def tune3(precision="fp32", npatch=512, steps=10, verbose=True, out=CACHE):
    """Plane block of the batched 3-D step (workgroups per grid vs recomputed
    planes) on the batched 3-D Euler workload."""
    from . import capi3
    sizes, _, h, qs = workload.batched_euler3_level(npatch)
    cells = int(np.prod(sizes, axis=1).sum())
    dt = 0.5 * h / max(workload.max_wave_speed3(q) for q in qs[:64])
    prec = capi3.FP32 if precision == "fp32" else capi3.FP64
    results = []
    for pb in PLANE_BLOCKS:
        capi3.init3(capi3.make_config3(capi3.EULER, precision=prec))
        capi3.set_patches3(sizes, pb)
        for k, q in enumerate(qs):
            capi3.put_patch3(k, q)
        for _ in range(2):
            capi3.step_patches3(h, h, h, dt)
        capi3.synchronize3()
        t0 = time.perf_counter()
        for _ in range(steps):
            capi3.step_patches3(h, h, h, dt)
        capi3.synchronize3()
        ms = (time.perf_counter() - t0) / steps * 1e3
        capi3.finalize3()
        results.append({"plane_block": pb, "ms": ms, "cell_updates_per_s": cells / ms * 1e3})
        if verbose:
            print("plane block %d  %8.3f ms  %.3e cell-updates/s" % (pb, ms, cells / ms * 1e3),
                  flush=True)
    best = min(results, key=lambda r: r["ms"])
    _store("3d/%d/%s" % (capi.EULER, precision),
           {"plane_block": best["plane_block"], "ms": best["ms"], "npatch": npatch,
            "all": results, "when": time.strftime("%Y-%m-%d %H:%M:%S")}, out)
    return best


def tune(precision="fp32", npatch=4096, steps=10, verbose=True, out=CACHE):
    prec = PRECISIONS[precision]
    boxes, h, mx, my = workload.batched_euler_level(npatch)
    capi.init(capi.make_config(capi.EULER, precision=prec))
    capi.set_level(1, boxes, h, h)
    qs = [workload.patch_initial_state(boxes[k], h, 1000 + k) for k in range(len(boxes))]
    capi.put_level(1, qs, which=0)
    capi.put_level(1, qs, which=1)
    cells = int((mx * my).sum())
    dt = 0.2 * h
    results = []
    for txmax, tymax in CANDIDATES[precision]:
        try:
            capi.set_tiling(txmax, tymax)
        except capi.AmrgpuError:
            continue
        for _ in range(2):
            capi.sweep_level(1, dt, want_cfl=False)
        capi.synchronize()
        capi.event_record(0)
        for _ in range(steps):
            capi.sweep_level(1, dt, want_cfl=False)
        capi.event_record(1)
        ms = capi.event_elapsed(0, 1) / steps
        results.append({"tiling": [txmax, tymax], "ms": ms, "cell_updates_per_s": cells / ms * 1e3})
        if verbose:
            print("tiling %2dx%-2d  %8.3f ms  %.3e cell-updates/s" % (txmax, tymax, ms, cells / ms * 1e3),
                  flush=True)
    capi.finalize()
    best = min(results, key=lambda r: r["ms"])
    _store("%d/%s" % (capi.EULER, precision),
           {"tiling": best["tiling"], "ms": best["ms"], "npatch": npatch, "all": results,
            "when": time.strftime("%Y-%m-%d %H:%M:%S")}, out)
    return best


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["euler2d", "euler3d"], default="euler2d")
    ap.add_argument("--precision", choices=["mixed", "fp32", "fp64"], default="mixed")
    ap.add_argument("--npatch", type=int, default=0)
    ap.add_argument("--out", default=CACHE, help="tuning file to update (default: the package's)")
    a = ap.parse_args()
    if a.workload == "euler3d":
        p3 = "fp64" if a.precision == "fp64" else "fp32"
        print(json.dumps(tune3(p3, a.npatch or 512, out=a.out)))
    else:
        print(json.dumps(tune(a.precision, a.npatch or 4096, out=a.out)))
