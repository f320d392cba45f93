"""Autotuner for the sweep's tile shape and waves per CU.

Times the level sweep of a representative level for each candidate pair of
tile caps (txmax, tymax) and, on the mixed path, each occupancy bound (the
workgroups per CU, i.e. waves per SIMD, its register allocation is bounded
for: 5, 6 or 7), and records the fastest in tuning.json, keyed by solver and
precision.  The tile caps fix the LDS footprint of a workgroup, which also
caps the workgroups per CU; on the FP32 and FP64 paths that is the only
occupancy knob.

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
# occupancy bounds per precision (0: the library's fixed bound)
OCCUPANCY = {"fp32": [0], "fp64": [0], "mixed": [5, 6, 7]}
PRECISIONS = {"fp32": capi.FP32, "fp64": capi.FP64, "mixed": capi.MIXED}


def load(solver, precision):
    """Tuned (txmax, tymax) for a solver/precision, or None."""
    if not os.path.exists(CACHE):
        return None
    with open(CACHE) as f:
        d = json.load(f)
    v = d.get("%d/%s" % (solver, precision))
    return tuple(v["tiling"]) if v else None


def load_occupancy(solver, precision):
    """Tuned occupancy bound (workgroups per CU) for a solver/precision, or None."""
    if not os.path.exists(CACHE):
        return None
    with open(CACHE) as f:
        d = json.load(f)
    v = d.get("%d/%s" % (solver, precision))
    return int(v["wgs"]) if v and v.get("wgs") else None


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
    """Tile caps and occupancy bound of the 2-D level update, timed on the
    step the benchmark measures: advance_level of level 2 of the two-level
    batched Euler hierarchy (ghost fill with time-interpolated coarse values,
    the sweep with its register fluxes, qad + fluxad), level 1 advanced once
    so that level 2 sits inside the coarse step."""
    prec = PRECISIONS[precision]
    H, cells = workload.batched_euler_hierarchy(npatch)
    dt2 = H.levels[1]["possk"]
    cfg = capi.make_config(capi.EULER, order=2, transverse=2, limiter=4, cfl=0.9, cflv1=1.0,
                           gamma=1.4, efix=True, precision=prec)
    results = []
    for wgs in OCCUPANCY[precision]:
        for txmax, tymax in CANDIDATES[precision]:
            capi.init(cfg)
            try:
                capi.set_tiling(txmax, tymax)
                if wgs:
                    capi.set_occupancy(wgs)
            except capi.AmrgpuError:
                capi.finalize()
                continue
            capi.load_hierarchy(H)
            capi.set_time(1, 0.0, 2.0 * dt2)
            capi.advance_level(1)

            def step():
                capi.set_time(2, dt2, dt2)
                capi.advance_level(2, want_cfl=False)

            for _ in range(3):
                step()
            capi.synchronize()
            capi.event_record(0)
            for _ in range(steps):
                step()
            capi.event_record(1)
            ms = capi.event_elapsed(0, 1) / steps
            capi.finalize()
            results.append({"tiling": [txmax, tymax], "wgs": wgs, "ms": ms,
                            "cell_updates_per_s": cells / ms * 1e3})
            if verbose:
                print("tiling %2dx%-2d  wgs %d  %8.3f ms  %.3e cell-updates/s"
                      % (txmax, tymax, wgs, ms, cells / ms * 1e3), flush=True)
    best = min(results, key=lambda r: r["ms"])
    _store("%d/%s" % (capi.EULER, precision),
           {"tiling": best["tiling"], "wgs": best["wgs"], "ms": best["ms"], "npatch": npatch,
            "step": "advance_level(2) of the two-level batched hierarchy",
            "all": results, "when": time.strftime("%Y-%m-%d %H:%M:%S")}, out)
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
