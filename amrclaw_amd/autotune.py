"""Autotuner for the sweep's tile shape.

Times the level sweep of a representative level for each candidate pair of
tile caps (txmax, tymax) and records the fastest in tuning.json, keyed by
solver and precision.  The tile caps fix the LDS footprint of a workgroup
and with it how many workgroups (waves) share a CU, so this also picks the
waves per CU.

    python -m amrclaw_amd.autotune [--precision fp32|fp64] [--npatch N]
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


def load(solver, precision):
    """Tuned (txmax, tymax) for a solver/precision, or None."""
    if not os.path.exists(CACHE):
        return None
    with open(CACHE) as f:
        d = json.load(f)
    v = d.get("%d/%s" % (solver, precision))
    return tuple(v["tiling"]) if v else None


def tune(precision="fp32", npatch=4096, steps=10, verbose=True):
    prec = capi.FP32 if precision == "fp32" else capi.FP64
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
    d = {}
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            d = json.load(f)
    d["%d/%s" % (capi.EULER, precision)] = {"tiling": best["tiling"], "ms": best["ms"],
                                            "npatch": npatch, "all": results,
                                            "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    with open(CACHE, "w") as f:
        json.dump(d, f, indent=1)
    return best


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--npatch", type=int, default=4096)
    a = ap.parse_args()
    print(json.dumps(tune(a.precision, a.npatch)))
