#!/usr/bin/env python3
"""Benchmark: cell-updates/s of the batched 2-D Euler level update on one GPU.

Workload (one "step" = one level update of one level): N = 4096 patches,
mx, my ~ uniform even numbers in [8, 128] (numpy default_rng(20261202)),
2 ghost cells, shelf-packed on the unit square; 2-D Euler (gamma 1.4) Roe
solver with entropy fix, MC limiter, transverse corrections (method 2,2);
four-quadrant initial state with a seeded +-1e-3 perturbation.  Inputs are
resident in HBM before the timed region.

The headline is the FP32-batched path; the FP64 batched path runs the same
steps on the same layout and dt for the speed-up and the L1 error of FP32
against FP64 (the FP64 path is bit-exact with the CPU oracle; see
tests/test_gpu_sweep.py).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--npatch N]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--npatch", type=int, default=4096)
    ap.add_argument("--txmax", type=int, default=0)
    ap.add_argument("--tymax", type=int, default=0)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP64 comparison run")
    ap.add_argument("--cpu-seconds", type=float, default=8.0,
                    help="CPU baseline budget (0 disables it)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB "
                         "(reference kernel for scaling FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy of known size: under rocprofv3 --pmc its FETCH_SIZE and
    WRITE_SIZE per dispatch give the factor from counter units to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "dispatches": 5}))


def setup_level(capi, workload, precision, npatch, txmax, tymax):
    boxes, h, mx, my = workload.batched_euler_level(npatch)
    cfg = capi.make_config(capi.EULER, order=2, transverse=2, limiter=4,
                           cfl=0.9, cflv1=1.0, gamma=1.4, efix=True,
                           precision=precision)
    capi.init(cfg)
    if not (txmax or tymax):
        from amrclaw_amd import autotune
        tuned = autotune.load(capi.EULER, "fp32" if precision == capi.FP32 else "fp64")
        if tuned:
            txmax, tymax = tuned
    if txmax or tymax:
        capi.set_tiling(txmax, tymax)
    capi.set_level(1, boxes, h, h)
    qs = [workload.patch_initial_state(boxes[k], h, 1000 + k) for k in range(len(boxes))]
    # both time levels hold the initial data so every step reads valid ghost
    # cells
    capi.put_level(1, qs, which=0)
    capi.put_level(1, qs, which=1)
    return boxes, h, mx, my


def run(capi, steps, warmup, dt, sync):
    for _ in range(warmup):
        capi.sweep_level(1, dt, want_cfl=False)
    sync()
    t0 = time.perf_counter()
    capi.event_record(0)
    for _ in range(steps):
        capi.sweep_level(1, dt, want_cfl=False)
    capi.event_record(1)
    sync()
    wall = time.perf_counter() - t0
    return wall, capi.event_elapsed(0, 1) / 1e3


def download(capi, boxes):
    shapes = [(int(b[3] - b[2]) + 5, int(b[1] - b[0]) + 5, 4) for b in boxes]
    return [q[2:-2, 2:-2] for q in capi.get_level(1, shapes)]


def cpu_baseline(workload, boxes, h, dt, budget):
    """The reference's stepgrid (built from its own Fortran by oracle/Makefile)
    with its OpenMP patch loop, on the first patches of the same workload;
    falls back to the oracle's C port when the reference build is absent."""
    from oracle import orc
    p = orc.make_params(orc.EULER, order=2, transverse=2, limiter=4)
    nsamp = min(128, len(boxes))
    qs = [workload.patch_initial_state(boxes[k], h, 1000 + k) for k in range(nsamp)]
    cells = sum((q.shape[0] - 4) * (q.shape[1] - 4) for q in qs)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if orc.ref_available():
        kind = "reference"
        fn = lambda: orc.ref_stepgrid_batch(p, qs, h, h, dt)
    else:
        kind, threads = "port", 1
        fn = lambda: [orc.stepgrid(p, q, h, h, dt) for q in qs]
    t0 = time.perf_counter()
    reps = 0
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget:
            break
    return {"value": cells * reps / el, "unit": "cell-updates/s",
            "cores": threads, "kind": kind,
            "sample": "%d patches (%d cells) of the same workload, %d passes of "
                      "stepgrid in %.1f s" % (nsamp, cells, reps, el)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    import torch
    from amrclaw_amd import capi, workload

    def sync():
        capi.synchronize()
        torch.cuda.synchronize()

    prec = capi.FP32 if a.precision == "fp32" else capi.FP64
    boxes, h, mx, my = setup_level(capi, workload, prec, a.npatch, a.txmax, a.tymax)
    cells = int((mx * my).sum())
    # dt from the Courant number of the initial state, cfl_desired = 0.9
    cfl0, _ = capi.sweep_level(1, 0.1 * h)
    dt = 0.1 * h * 0.9 / cfl0
    # restore the initial state (the probe advanced it)
    boxes, h, mx, my = setup_level(capi, workload, prec, a.npatch, a.txmax, a.tymax)

    wall, kern = run(capi, a.steps, a.warmup, dt, sync)
    esize = 4 if prec == capi.FP32 else 8
    bytes_per_cell = 2 * 4 * esize            # (2 nvar + maux) * s
    value = cells * a.steps / wall
    achieved = cells * bytes_per_cell / (kern / a.steps) / 1e9
    result = {
        "metric": "cell-updates/s, FP32-batched vs FP64 per-patch at fixed L1 err; % HBM peak",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == capi.FP32 else "f64", "data": "synthetic",
        "config": {"workload": "batched 2-D Euler level update: %d patches, mx,my "
                               "even U[8,128] seed 20261202, Roe+efix, MC, "
                               "transverse 2" % len(boxes),
                   "patches": len(boxes), "cells_per_step": cells,
                   "nghost": 2, "precision": a.precision},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None,
                     "kernel": "sweep_kernel<EulerRoe>",
                     "kernel_ms": kern / a.steps * 1e3,
                     "bytes_per_cell_update": bytes_per_cell},
    }
    if not a.no_compare and prec == capi.FP32:
        q32 = download(capi, boxes)
        setup_level(capi, workload, capi.FP64, a.npatch, a.txmax, a.tymax)
        wall64, kern64 = run(capi, a.steps, a.warmup, dt, sync)
        q64 = download(capi, boxes)
        num = sum(float(np.abs(x - y).sum()) for x, y in zip(q32, q64))
        den = sum(float(np.abs(y).sum()) for y in q64)
        m32 = sum(float(x[..., 0].sum()) for x in q32)
        m64 = sum(float(y[..., 0].sum()) for y in q64)
        result["fp64_batched"] = {"value": cells * a.steps / wall64,
                                  "kernel_ms": kern64 / a.steps * 1e3}
        result["speedup_vs_fp64_batched"] = wall64 / wall
        result["l1_rel_vs_fp64"] = num / den
        result["mass_rel_diff_vs_fp64"] = abs(m32 - m64) / abs(m64)
        result["steps_compared"] = a.steps + a.warmup
    capi.finalize()
    if a.cpu_seconds > 0:
        result["cpu_baseline"] = cpu_baseline(workload, boxes, h, dt, a.cpu_seconds)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
