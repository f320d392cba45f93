#!/usr/bin/env python3
"""Benchmark: cell-updates/s of the batched 2-D Euler level update on one GPU.

Workload: a two-level hierarchy (amrclaw_amd.workload.batched_euler_hierarchy).
Level 2 holds N = 4096 patches, mx, my uniform even numbers in [8, 128]
(numpy default_rng(20261202)), 2 ghost cells, shelf-packed on the unit
square (18,855,660 interior cells); level 1 covers the domain at ratio 2.
2-D Euler (gamma 1.4), Roe solver with entropy fix, MC limiter, transverse
corrections (method(2) = method(3) = 2), extrapolation boundaries; the
four-quadrant initial state with a seeded +-1e-3 perturbation.

One step = one level update of level 2 (amrgpu_advance_level: ghost fill
with time-interpolated coarse values, the sweep of all 4096 patches and the
fine-side register work), always at the same point of the coarse step
(level-2 time dt2, between the coarse times 0 and 2 dt2).  The register
work differs by precision: the FP64 and FP32 runs do qad and fluxad as the
reference does; the mixed run works in flux form, so its registers take
the fluxes the sweep computed at the perimeter and qad, which turns the
fine side's fluctuations into fluxes, is skipped: its speed-up over FP64
batched is therefore a little larger than a like-for-like one.
Inputs are resident in HBM before the timed region; the CFL readback is
not part of the step.

Runs: the mixed path (the headline: FP32 wave and flux arithmetic on FP64
storage, flux form with an FP64 update, FP64 registers, coarse interpolation
and totals), then FP64 batched, FP64 with one sweep launch per patch (the
plain per-patch GPU path) and FP32 storage on the same layout and dt; the
relative L1 error and level-mass difference against FP64 are reported, and
the composite conservation drift of each precision on a periodic two-level
hierarchy.  The CPU baseline is the reference's own advanc (compiled from
its Fortran by oracle/Makefile) over a sample of the same hierarchy.

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
# HBM bytes per launch of the dominant kernel(s), from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE runs of the default workloads (scaled by the
# plain-copy calibration: FETCH_SIZE x 2 KiB, WRITE_SIZE x 1 KiB); the
# summaries and commands are the files named here.
MEASURED_TRAFFIC = {
    ("euler2d", "mixed", 4096, "batched"): (1200800 * 2048 + 783700 * 1024,
                                            "profiles/level_update_euler_mixed_4096.md"),
    ("euler2d", "fp32", 4096, "batched"): (459126 * 2048 + 450853 * 1024,
                                           "profiles/level_update_euler_fp32_4096.md"),
    ("euler2d", "mixed", 4096, "march"): (397129 * 2048 + 782660 * 1024,
                                          "profiles/level_update_euler_mixed_4096_march.md"),
    ("euler3d", "fp32", 512, "fused"): (69557 * 2048 + 96870 * 1024,
                                        "profiles/step3_euler_fp32_512_fused.md"),
    ("euler3d", "fp32", 512, "phases"): (
        (271400 + 875200 + 1415000 + 330700) * 2048 + (801800 + 1136000 + 1143000 + 98580) * 1024,
        "profiles/step3_euler_fp32_512.md"),
}
# Vector-ALU wave instructions per launch of the same kernels (rocprofv3
# --pmc SQ_INSTS_VALU of the default workloads; the summaries named here).
MEASURED_VALU = {
    ("euler2d", "mixed", 4096, "march"): (423823602,
                                          "profiles/level_update_euler_mixed_4096_march.md"),
    ("euler2d", "mixed", 4096, "batched"): (757577540,
                                            "profiles/level_update_euler_mixed_4096.md"),
    ("euler3d", "fp32", 512, "fused"): (967839559, "profiles/step3_euler_fp32_512_fused.md"),
}
SIMD_COUNT = 1024                 # 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9                  # peak engine clock
VALU_ISSUE_CYCLES = 2             # a SIMD issues one wave64 VALU instruction per 2 cycles
METRIC = "cell-updates/s, FP32-batched vs FP64 per-patch at fixed L1 err; % HBM peak"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--npatch", type=int, default=4096)
    ap.add_argument("--occupancy", type=int, default=0,
                    help="mixed sweep's workgroups per CU (5, 6, 7; default: tuned)")
    ap.add_argument("--txmax", type=int, default=0)
    ap.add_argument("--tymax", type=int, default=0)
    ap.add_argument("--precision", choices=["mixed", "fp32", "fp64"], default="mixed",
                    help="mixed (default): FP32 waves and fluxes on FP64 storage, "
                         "conservative to FP64 rounding; fp32: FP32 storage too")
    ap.add_argument("--per-patch", action="store_true",
                    help="time the one-launch-per-patch sweep instead")
    ap.add_argument("--persistent", action="store_true",
                    help="persistent sweep launch (workgroups take tiles from a counter)")
    ap.add_argument("--tile", action="store_true",
                    help="time the batched tile sweep (one workgroup per LDS tile) instead "
                         "of the default march sweep")
    ap.add_argument("--march", action="store_true",
                    help="march sweep (the default: column strips marching up the rows in "
                         "registers)")
    ap.add_argument("--graphs", action="store_true",
                    help="replay the level steps as hipGraphs in the timed region")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP64 comparison runs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU baseline budget (0 disables it)")
    ap.add_argument("--workload", choices=["euler2d", "euler3d", "amr3d", "C1", "C2", "C3",
                                           "regrid"],
                    default="euler2d",
                    help="euler3d: the batched 3-D Euler step (512 grids); amr3d: coarse "
                         "Berger-Oliger steps of the 3-D radial run configuration (three "
                         "levels, the whole 3-D level update); C1, C2, C3: the 2-D run "
                         "configurations to their end times; regrid: the device regrid "
                         "(filval) of the headline level; default: the headline 2-D "
                         "level update")
    ap.add_argument("--plane-block", type=int, default=0,
                    help="3-D: planes per workgroup of a sweep (0: library default)")
    ap.add_argument("--step3", choices=["fused", "phases"], default="fused",
                    help="3-D FP32 step: fused (one kernel per tile, faces on chip; the "
                         "library's FP32 default) or phases (x, y, z sweep kernels with the "
                         "flux arrays in HBM, then the update); FP64 always runs phases")
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


def install(capi, H, precision, launch, graphs, txmax, tymax, occupancy=0):
    """Load the hierarchy and bring it to the timed state: level 1 advanced
    once from t = 0 (saveqc fills the level-2 registers' coarse ring),
    level 2 half way through that coarse step."""
    cfg = capi.make_config(capi.EULER, order=2, transverse=2, limiter=4, cfl=0.9,
                           cflv1=1.0, gamma=1.4, efix=True, precision=precision)
    capi.init(cfg)
    from amrclaw_amd import autotune
    key = {capi.FP64: "fp64", capi.FP32: "fp32", capi.MIXED: "mixed"}[precision]
    if not (txmax or tymax):
        tuned = autotune.load(capi.EULER, key) or autotune.load(capi.EULER, "fp32")
        if tuned:
            txmax, tymax = tuned
    if txmax or tymax:
        capi.set_tiling(txmax, tymax)
    if launch == "march":
        if occupancy:
            capi.set_occupancy(occupancy)
    else:
        wgs = occupancy or autotune.load_occupancy(capi.EULER, key)
        if wgs and precision == capi.MIXED:
            capi.set_occupancy(wgs)
    if launch != "batched":
        capi.set_launch_mode({"per-patch": capi.LAUNCH_PER_PATCH,
                              "persistent": capi.LAUNCH_PERSISTENT,
                              "march": capi.LAUNCH_MARCH}[launch])
    capi.load_hierarchy(H)
    dt2 = H.levels[1]["possk"]
    capi.set_time(1, 0.0, 2.0 * dt2)
    capi.advance_level(1)
    capi.set_graphs(graphs)
    return dt2


def step(capi, dt2):
    capi.set_time(2, dt2, dt2)
    capi.advance_level(2, want_cfl=False)


def run(capi, dt2, steps, warmup, sync):
    for _ in range(warmup):
        step(capi, dt2)
    sync()
    capi.set_timing(True)
    t0 = time.perf_counter()
    capi.event_record(0)
    for _ in range(steps):
        step(capi, dt2)
    capi.event_record(1)
    sync()
    wall = time.perf_counter() - t0
    out = {"wall": wall, "events_ms": capi.event_elapsed(0, 1)}
    for name, kind in (("sweep", capi.KT_SWEEP), ("fill", capi.KT_FILL),
                       ("registers", capi.KT_REGISTERS)):
        ms, n = capi.kernel_time(kind)
        out[name] = (ms, n)
    capi.set_timing(False)
    return out


def conservation_drift(capi, precision, npatch=256, coarse_steps=4, launch="batched"):
    """Composite conservation on a periodic domain: the two-level batched
    hierarchy with npatch level-2 patches and periodic boundaries, level 1
    first averaged from level 2, then coarse_steps Berger-Oliger steps
    (advance 1, advance 2 twice, update 1 with refluxing).  Returns
    |M - M0| / |M0| of conck(1) (the FP64 total of level 1)."""
    from amrclaw_amd import workload
    H, _ = workload.batched_euler_hierarchy(npatch)
    cfg = capi.make_config(capi.EULER, order=2, transverse=2, limiter=4, cfl=0.9, cflv1=1.0,
                           gamma=1.4, efix=True, mthbc=(2, 2, 2, 2), precision=precision)
    capi.init(cfg)
    try:
        if launch == "march":
            capi.set_launch_mode(capi.LAUNCH_MARCH)
        capi.load_hierarchy(H)
        dt2 = H.levels[1]["possk"]
        capi.set_time(1, 0.0, 2.0 * dt2)
        capi.set_time(2, 0.0, dt2)
        capi.update_level(1)
        m0 = capi.conck(1)
        for _ in range(coarse_steps):
            capi.advance_level(1, want_cfl=False)
            capi.advance_level(2, want_cfl=False)
            capi.advance_level(2, want_cfl=False)
            capi.update_level(1)
        m = capi.conck(1)
    finally:
        capi.finalize()
    return abs(m - m0) / abs(m0)


def interiors(capi, H):
    L = H.levels[1]
    got = capi.read_level(H, 2, 0)
    return [got[o:o + int(np.prod(s))].reshape(s)[2:-2, 2:-2]
            for o, s in zip(L["off"], L["shapes"])]


def _quiet_stdout(fn):
    """Run fn with the process's stdout (fd 1) sent to stderr: the
    reference's Fortran prints progress there, and stdout carries only the
    result line."""
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        return fn()
    finally:
        os.dup2(saved, 1)
        os.close(saved)


def cpu_baseline(H, dt2, budget, nsample=256, parallel=True):
    """The reference's advanc (its own Fortran, OpenMP over grids as
    advanc.f:60-129, or serial with parallel=False) on level 2 of a sample of
    the hierarchy: level 1 whole, the first nsample level-2 patches; same
    state and times as the GPU step.  Falls back to the oracle's C port
    (serial) without the build."""
    from oracle import orc
    from amrclaw_amd.hierarchy import Hier
    p = orc.make_params(orc.EULER, order=2, transverse=2, limiter=4)
    S = Hier(4)
    L1, L2 = H.levels
    S.add_level(L1["boxes"], L1["hx"], L1["hy"], 2.0 * dt2, 0.0, L1["iregsz"], L1["jregsz"], 2, 2)
    S.add_level(L2["boxes"][:nsample], L2["hx"], L2["hy"], dt2, dt2, L2["iregsz"], L2["jregsz"], 2, 2)
    for l in (1, 2):
        for k in range(len(S.levels[l - 1]["boxes"])):
            S.grid(l, k, 0)[:] = H.grid(l, k, 0)
            S.grid(l, k, 1)[:] = H.grid(l, k, 1)
    b = S.levels[1]["boxes"]
    cells = int(((b[:, 1] - b[:, 0] + 1) * (b[:, 3] - b[:, 2] + 1)).sum())
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)) if parallel else 1
    if orc.ref_available():
        kind = "reference"
        A = orc.RefAMR(p, S)
        A.set_parallel(parallel)
        A.advanc(1)                           # level 1: 0 -> 2 dt2, saveqc(2)
        reset = lambda: A.set_time(2, dt2)
    else:
        kind, threads = "port", 1
        A = orc.OracleAMR(p, S)
        A.advanc(1)
        reset = lambda: A.h.time.__setitem__(2, dt2)
    t0 = time.perf_counter()
    reps = 0
    while True:
        reset()
        A.advanc(2)
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget:
            break
    if kind == "reference":
        A.flush()
    return {"value": cells * reps / el, "unit": "cell-updates/s", "cores": threads,
            "kind": kind,
            "sample": "level-2 update (advanc: bound with coarse interpolation, qad, "
                      "stepgrid, fluxad) of the first %d of the 4096 patches (%d cells) "
                      "over the whole level 1, %d passes in %.1f s" % (nsample, cells, reps, el)}


def cpu_baseline3(sizes, qs, dx, dt, budget):
    """The reference's 3-D stepgrid.f (compiled from its Fortran by
    oracle/Makefile, one thread) over the grids of the batch in turn, for
    about `budget` seconds."""
    from oracle import orc
    p = orc.make_params3(orc.EULER, order=2, transverse=22, limiter=4)
    kind = "reference" if orc.ref3_available() else "port"
    fn = orc.ref3_stepgrid3 if kind == "reference" else orc.stepgrid3
    cells = 0
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget:
        k = n % len(qs)
        fn(p, qs[k], dx, dx, dx, dt)
        cells += int(np.prod(sizes[k]))
        n += 1
    wall = time.perf_counter() - t0
    return {"value": cells / wall, "unit": "cell-updates/s", "cores": 1, "kind": kind,
            "sample": "stepgrid.f (3-D) of the batch's grids in turn, %d grid steps "
                      "(%d cells) in %.1f s" % (n, cells, wall)}


def bench3(a):
    """Batched 3-D Euler step: 512 grids, sizes even U[8,32]^3 (seed
    20261203) on a 256^3 index space of the unit cube, the radial bump
    (rho = p = 2 for r < 0.2) with a seeded +-1e-3 perturbation; Roe solver
    with entropy fix, MC limiter, method(3) = 22 (double transverse).  One
    step = the x-, y-, z-sweeps and the update of every grid (stepgrid.f
    for each; ghost cells are those of the initial state)."""
    import torch
    from amrclaw_amd import capi3, workload

    sizes, corners, h, qs = workload.batched_euler3_level(512)
    cells = int(np.prod(sizes, axis=1).sum())
    smax = max(workload.max_wave_speed3(q) for q in qs)
    dt = 0.5 * h / smax
    # the 3-D step has FP32 and FP64 paths (the mixed path is 2-D)
    p3 = "fp64" if a.precision == "fp64" else "fp32"
    prec = capi3.FP32 if p3 == "fp32" else capi3.FP64
    esize = 4 if prec == capi3.FP32 else 8

    from amrclaw_amd import autotune
    plane_block = a.plane_block or autotune.load_plane_block(p3) or 0

    fused = prec == capi3.FP32 and a.step3 == "fused"

    def install(precision):
        capi3.init3(capi3.make_config3(capi3.EULER, precision=precision))
        capi3.set_patches3(sizes, plane_block)
        if precision == capi3.FP32:
            capi3.set_step_mode3(capi3.STEP_FUSED if fused else capi3.STEP_PHASES)
        for k, q in enumerate(qs):
            capi3.put_patch3(k, q)

    def run(steps, warmup):
        for _ in range(warmup):
            capi3.step_patches3(h, h, h, dt)
        capi3.synchronize3()
        torch.cuda.synchronize()
        capi3.set_timing3(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            capi3.step_patches3(h, h, h, dt)
        capi3.synchronize3()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kt = [capi3.kernel_time3(k) for k in (capi3.KT_X, capi3.KT_Y, capi3.KT_Z,
                                               capi3.KT_UPDATE)]
        capi3.set_timing3(False)
        return wall, kt

    def interiors():
        return [capi3.get_patch3(k, q.shape)[2:-2, 2:-2, 2:-2] for k, q in enumerate(qs)]

    print("bench3: timed run", p3, file=sys.stderr, flush=True)
    install(prec)
    wall, kt = run(a.steps, a.warmup)
    q_main = interiors() if not a.no_compare else None
    capi3.finalize3()
    # the fused step is one launch, timed in the x-sweep's slot
    names = ("fused_step",) if fused else ("x_sweep", "y_sweep", "z_sweep", "update")
    kms = {nm: t[0] / max(t[1], 1) for nm, t in zip(names, kt)}
    result = {
        "metric": METRIC,
        "value": cells * a.steps / wall, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == capi3.FP32 else "f64", "data": "synthetic",
        "config": {"workload": "batched 3-D Euler step: 512 grids, mx,my,mz even U[8,32] "
                               "seed 20261203, radial bump, Roe+efix, MC, method(3)=22",
                   "patches": len(qs), "cells_per_step": cells, "nghost": 2,
                   "precision": p3, "launch": "fused tiles" if fused else "batched",
                   "step3": "fused" if fused else "phases",
                   "plane_block": None if fused else (plane_block or 4)},
        "kernels_ms_per_step": kms,
    }
    step_kern = sum(kms.values())
    dom = max(names[:3], key=lambda nm: kms[nm])
    bpc = 2 * 5 * esize
    achieved = cells * bpc / (step_kern / 1e3) / 1e9
    tkey = ("euler3d", p3, len(qs), "fused" if fused else "phases")
    tr = MEASURED_TRAFFIC.get(tkey) if not a.plane_block else None
    vr = MEASURED_VALU.get(tkey) if not a.plane_block else None
    result["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                          "traffic": tr[0] if tr else None,
                          "traffic_source": tr[1] if tr else None,
                          "kernel": "sweep3_tile_kernel" if fused else
                                    "sweep3_phase_kernel (x+y+z) + update3_kernel",
                          "kernel_ms": step_kern, "dominant": dom,
                          "bytes_per_cell_update": bpc,
                          "algorithmic_bytes_per_launch": cells * bpc}
    if vr:
        # the fused step is bound by its vector instructions, not by HBM
        result["roofline"]["valu"] = {
            "wave_instructions_per_launch": vr[0],
            "lane_instructions_per_cell": vr[0] * 64 / cells,
            "issue_frac": vr[0] * VALU_ISSUE_CYCLES / (SIMD_COUNT * step_kern * 1e-3 * CLOCK_HZ),
            "source": vr[1]}
    if not a.no_compare and prec == capi3.FP32:
        print("bench3: FP64 batched", file=sys.stderr, flush=True)
        install(capi3.FP64)
        wall64, kt64 = run(a.steps, a.warmup)
        q64 = interiors()
        capi3.finalize3()
        num = sum(float(np.abs(x - y).sum()) for x, y in zip(q_main, q64))
        den = sum(float(np.abs(y).sum()) for y in q64)
        m32 = sum(float(x[..., 0].sum()) for x in q_main)
        m64 = sum(float(y[..., 0].sum()) for y in q64)
        result["fp64_batched"] = {"value": cells * a.steps / wall64,
                                  "ms_per_step": wall64 / a.steps * 1e3}
        result["speedup_vs_fp64_batched"] = wall64 / wall
        result["l1_rel_vs_fp64"] = num / den
        result["mass_rel_diff_vs_fp64"] = abs(m32 - m64) / abs(m64)
        result["steps_compared"] = a.steps + a.warmup
    if a.cpu_seconds > 0:
        print("bench3: CPU baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = _quiet_stdout(
            lambda: cpu_baseline3([tuple(s) for s in sizes], qs, h, dt, a.cpu_seconds))
    print(json.dumps(result))


def bench_amr3(a):
    """3-D AMR level update on the 3-D Euler run configuration
    (amrclaw_amd.workload.radial_euler3_hierarchy: rho = p = 2 inside
    r < 0.2 about the centre of the unit cube, 1 outside, extrapolation
    BCs, 64^3 level 1 in eight 32^3 grids, two finer levels at ratio 2).
    The timed run goes from t = 0 to 0.2 in 26 coarse Berger-Oliger steps
    (advanc of level 1, 2 of level 2, 4 of level 3, each: ghost fill,
    saveqc, qad, the three sweeps and the update, fluxsv, fluxad; and the
    updates); --warmup coarse steps run first on a separate copy.  Reports
    the FP32 path's relative L1 error against FP64 at t = 0.2 and the
    level-1 mass drift (conck) of both."""
    import torch
    from amrclaw_amd import capi3, workload

    H = workload.radial_euler3_hierarchy()
    nlev = H.nlev
    steps = workload.C5_STEPS
    ncell = [int(sum((b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1)
                     for b in L["boxes"])) for L in H.levels]
    cells = sum(c * 2 ** l for l, c in enumerate(ncell))
    p3 = "fp64" if a.precision == "fp64" else "fp32"
    prec = capi3.FP32 if p3 == "fp32" else capi3.FP64
    esize = 4 if prec == capi3.FP32 else 8
    cfl_seen = [0.0]

    def coarse(level=1):
        dtn, cfl = capi3.advance_level3(level)
        cfl_seen[0] = max(cfl_seen[0], cfl)
        if level < nlev:
            for _ in range(2):
                coarse(level + 1)
            capi3.update_level3(level)

    def load(precision):
        capi3.init3(capi3.make_config3(capi3.EULER, precision=precision))
        if a.plane_block:
            capi3.set_plane_block3(a.plane_block)
        capi3.load_hierarchy3(H)
        for l in range(nlev - 1, 0, -1):      # setgrd: coarse data averaged from the finer
            capi3.update_level3(l)

    def run(precision, timed):
        if timed and a.warmup:
            load(precision)
            for _ in range(a.warmup):
                coarse()
            capi3.finalize3()
        load(precision)
        m0 = capi3.conck_level3(1)
        capi3.synchronize3()
        torch.cuda.synchronize()
        capi3.set_timing3(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            coarse()
        capi3.synchronize3()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kt = [capi3.kernel_time3(k) for k in (capi3.KT_X, capi3.KT_Y, capi3.KT_Z,
                                               capi3.KT_UPDATE)]
        capi3.set_timing3(False)
        m1 = capi3.conck_level3(1)
        levels = [capi3.read_level3(H, l, 0) for l in range(1, nlev + 1)]
        capi3.finalize3()
        return wall, kt, (m1 - m0) / m0, levels

    print("bench_amr3: timed run", p3, file=sys.stderr, flush=True)
    wall, kt, drift, lv = run(prec, True)
    names = ("x_sweep", "y_sweep", "z_sweep", "update")
    kms = {nm: t[0] / steps for nm, t in zip(names, kt)}
    result = {
        "metric": METRIC,
        "value": cells * steps / wall, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": a.warmup, "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == capi3.FP32 else "f64", "data": "synthetic",
        "config": {"workload": "3-D AMR level update, 3-D Euler run configuration: radial "
                               "bump rho=p=2 for r<0.2 about the centre of [0,1]^3, "
                               "extrapolation BCs, 64^3 level 1 (8 grids of 32^3), levels 2-3 "
                               "at ratio 2, Roe+efix, MC, method(3)=22, t = 0 to 0.2 in 26 "
                               "coarse Berger-Oliger steps; one step = one coarse step",
                   "grids_per_level": [len(L["boxes"]) for L in H.levels],
                   "cells_per_level": ncell, "cell_updates_per_step": cells,
                   "precision": p3, "max_cfl": cfl_seen[0]},
        "kernels_ms_per_step": kms,
        "level1_mass_drift": drift,
    }
    step_kern = sum(kms.values())
    bpc = 2 * 5 * esize
    achieved = cells * bpc / (step_kern / 1e3) / 1e9
    result["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                          "kernel": "sweep3_phase_kernel (x+y+z) + update3_kernel over the "
                                    "seven level steps",
                          "kernel_ms": step_kern, "bytes_per_cell_update": bpc,
                          "algorithmic_bytes_per_step": cells * bpc,
                          "non_sweep_ms": wall / steps * 1e3 - step_kern}
    if not a.no_compare and prec == capi3.FP32:
        print("bench_amr3: FP64 run", file=sys.stderr, flush=True)
        wall64, _, drift64, lv64 = run(capi3.FP64, False)
        rel = [float(np.abs(x - y).sum() / np.abs(y).sum()) for x, y in zip(lv, lv64)]
        result["fp64"] = {"value": cells * steps / wall64, "ms_per_step": wall64 / steps * 1e3,
                          "level1_mass_drift": drift64}
        result["speedup_vs_fp64"] = wall64 / wall
        result["l1_rel_vs_fp64_per_level"] = rel
        # the outflow through the extrapolation boundaries is the same in
        # both; their difference is the FP32 path's conservation error
        result["mass_drift_minus_fp64"] = drift - drift64
    if a.cpu_seconds > 0:
        print("bench_amr3: CPU baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = _quiet_stdout(lambda: cpu_baseline_amr3(H, cells, a.cpu_seconds))
    print(json.dumps(result))


def cpu_baseline_amr3(H, cells, budget):
    """The reference's own 3-D level update (bound, saveqc, qad, stepgrid,
    fluxsv, fluxad, upbnd, update of src/3d, compiled from its Fortran by
    oracle/Makefile; one thread) over whole coarse Berger-Oliger steps of
    the same hierarchy, for about `budget` seconds (at least one step)."""
    from oracle import orc
    p = orc.make_params3(orc.EULER)
    kind = "reference" if orc.ref3_available() else "port"
    impl = orc.RefAMR3(p, H.copy()) if kind == "reference" else orc.OracleAMR3(p, H.copy())

    def coarse(level=1):
        impl.advanc(level)
        if level < H.nlev:
            for _ in range(2):
                coarse(level + 1)
            impl.update(level)

    n = 0
    t0 = time.perf_counter()
    while n == 0 or time.perf_counter() - t0 < budget:
        coarse()
        n += 1
    wall = time.perf_counter() - t0
    return {"value": cells * n / wall, "unit": "cell-updates/s", "cores": 1, "kind": kind,
            "sample": "%d coarse Berger-Oliger steps of the same hierarchy (%d cell "
                      "updates each) in %.1f s" % (n, cells, wall)}


def regrid_layout(H, shift=2):
    """A new level-2 layout for the regrid workload: every level-2 box
    shifted by `shift` fine cells in x, kept where it still nests in one
    level-1 grid (its coarse patch grown by one coarse cell inside that
    grid), so the new grids take part of their values from the old grids
    and the rest from the coarse level."""
    C, F = H.levels[0], H.levels[1]
    cb = np.asarray(C["boxes"])
    out = []
    for b in np.asarray(F["boxes"]):
        n = (b[0] + shift, b[1] + shift, b[2], b[3])
        if n[1] >= F["iregsz"]:
            continue
        ic0, ic1, jc0, jc1 = n[0] // 2 - 1, (n[1] + 1) // 2, n[2] // 2 - 1, (n[3] + 1) // 2
        inside = ((cb[:, 0] <= max(ic0, 0)) & (min(ic1, C["iregsz"] - 1) <= cb[:, 1]) &
                  (cb[:, 2] <= max(jc0, 0)) & (min(jc1, C["jregsz"] - 1) <= cb[:, 3]))
        if inside.any():
            out.append(n)
    return np.asarray(out, dtype=np.int32)


def bench_regrid(a):
    """The regrid front end at the headline's patch count: level 2 of the
    batched hierarchy regridded on the device (amrgpu_regrid_level: filval
    for every new grid, old grids copied, coarse interpolation elsewhere,
    planning included) back and forth between its layout and a shifted
    one; the neighbour lists (makeBndryList) of the new layout through the
    library's bucket index against the oracle's all-pairs restatement; CPU
    baseline: the reference's own filval.f90 on a sample of the new grids."""
    import torch
    from amrclaw_amd import capi, workload
    print("bench: building the regrid workload", file=sys.stderr, flush=True)
    H, _ = workload.batched_euler_hierarchy(a.npatch)
    old = np.asarray(H.levels[1]["boxes"], dtype=np.int32)
    new = regrid_layout(H)
    cells_new = int(((new[:, 1] - new[:, 0] + 1) * (new[:, 3] - new[:, 2] + 1)).sum())
    cells_old = int(((old[:, 1] - old[:, 0] + 1) * (old[:, 3] - old[:, 2] + 1)).sum())
    cfg = capi.make_config(capi.EULER, order=2, transverse=2, limiter=4, cfl=0.9, cflv1=1.0,
                           gamma=1.4, efix=True, precision=capi.FP64)
    capi.init(cfg)
    capi.load_hierarchy(H)
    t2 = H.levels[1]["time"]

    def sync():
        capi.synchronize()
        torch.cuda.synchronize()

    layouts = [new, old]
    for w in range(max(a.warmup, 1) * 2):
        capi.regrid_level(2, layouts[w % 2], t2)
    sync()
    steps = a.steps * 2
    t0 = time.perf_counter()
    for s in range(steps):
        capi.regrid_level(2, layouts[s % 2], t2)
    sync()
    wall = time.perf_counter() - t0
    capi.finalize()
    cells_avg = (cells_new + cells_old) / 2
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        capi.bndry_lists(new, H.levels[1]["iregsz"], H.levels[1]["jregsz"])
    t_lists = (time.perf_counter() - t0) / reps
    result = {"metric": "new-grid cells filled per second (device regrid: filval)",
              "value": cells_avg * steps / wall, "unit": "cell-fills/s", "n_gpus": 1,
              "steps": steps, "warmup": a.warmup, "ms_per_step": wall / steps * 1e3,
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
              "data": "synthetic",
              "config": {"workload": "regrid of the batched 2-D Euler level 2 (%d grids) "
                                     "to a layout shifted by 2 cells (%d grids) and back" %
                                     (len(old), len(new)),
                         "cells_new": cells_new, "cells_old": cells_old},
              "neighbour_lists_ms": {"bucket_index": t_lists * 1e3}}
    try:
        from oracle import orc
        p = orc.make_params(orc.EULER)
        h1 = orc.hier_struct(H)
        L = orc.lib()
        import ctypes as C
        L.orc_bndry_list.argtypes = [C.POINTER(orc.OrcParams), C.POINTER(orc.OrcHier), C.c_int,
                                     C.c_int, C.POINTER(C.c_int)]
        out = (C.c_int * len(old))()
        t0 = time.perf_counter()
        for k in range(len(old)):
            L.orc_bndry_list(C.byref(p), C.byref(h1), 2, k, out)
        result["neighbour_lists_ms"]["all_pairs_oracle"] = (time.perf_counter() - t0) * 1e3
        if a.cpu_seconds > 0 and orc.ref_available():
            print("bench: CPU baseline (reference filval)", file=sys.stderr, flush=True)

            def ref_sample():
                R = orc.ref_load_hier(p, H)
                R.ref_filval.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_double, C.c_int,
                                         C.c_void_p]
                done, ncell = 0, 0
                t0 = time.perf_counter()
                while time.perf_counter() - t0 < a.cpu_seconds and done < len(new):
                    b = new[done]
                    val = np.zeros((b[3] - b[2] + 5, b[1] - b[0] + 5, 4))
                    R.ref_filval(2, (C.c_int * 4)(*[int(x) for x in b]), t2, 4,
                                 val.ctypes.data_as(C.c_void_p))
                    done += 1
                    ncell += (b[1] - b[0] + 1) * (b[3] - b[2] + 1)
                el = time.perf_counter() - t0
                return {"value": ncell / el, "unit": "cell-fills/s", "cores": 1,
                        "kind": "reference",
                        "sample": "the reference's filval.f90 (intcopy, bc2amr, min-mod "
                                  "interpolation) on the first %d of the %d new grids "
                                  "(%d cells) in %.1f s" % (done, len(new), ncell, el)}
            result["cpu_baseline"] = _quiet_stdout(ref_sample)
    except OSError:
        pass
    print(json.dumps(result))


def bench_config(a):
    """A 2-D run configuration (amrclaw_amd.workload.config_hierarchy: C1
    advection, periodic, 50^2 + two levels at ratio 2; C2 acoustics, 50^2 +
    ratios 2, 4; C3 Euler quadrants, 80^2 in four grids + ratios 2, 4) run
    from t = 0 to its end time (1.0, 0.6, 0.8) in coarse Berger-Oliger steps
    of fixed size on static layouts.  FP64 (the configurations' precision)
    unless --precision says otherwise.  One step = one coarse step."""
    import torch
    from amrclaw_amd import capi, workload

    name = a.workload
    c = workload.CONFIGS[name]
    H = workload.config_hierarchy(name)
    nlev = H.nlev
    dt1 = H.levels[0]["possk"]
    steps = int(round(c["t_end"] / dt1))
    ncell = [int(((L["boxes"][:, 1] - L["boxes"][:, 0] + 1) *
                  (L["boxes"][:, 3] - L["boxes"][:, 2] + 1)).sum()) for L in H.levels]
    sub = [1]
    for L in H.levels[:-1]:
        sub.append(sub[-1] * L["ratx"])
    cells = sum(n * r for n, r in zip(ncell, sub))
    pname = a.precision if a.precision != "mixed" or "--precision" in sys.argv else "fp64"
    prec = {"mixed": capi.MIXED, "fp32": capi.FP32, "fp64": capi.FP64}[pname]

    def coarse(level=1):
        capi.advance_level(level)
        if level < nlev:
            for _ in range(H.levels[level - 1]["ratx"]):
                coarse(level + 1)
            capi.update_level(level)

    def load(precision):
        capi.init(capi.make_config(c["solver"], mthbc=c["mthbc"], domain=c["domain"],
                                   precision=precision))
        if a.graphs:
            capi.set_graphs(True)
        capi.load_hierarchy(H)

    if a.warmup:
        load(prec)
        for _ in range(a.warmup):
            coarse()
        capi.finalize()
    load(prec)
    capi.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        coarse()
    capi.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    capi.finalize()
    result = {
        "metric": METRIC, "value": cells * steps / wall, "unit": "cell-updates/s",
        "n_gpus": 1, "steps": steps, "warmup": a.warmup, "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "mixed": "f32"}[pname], "data": "synthetic",
        "config": {"workload": "2-D run configuration %s to t = %g on static layouts "
                               "(levels %s, ratios %s)" % (name, c["t_end"],
                                                          [len(L["boxes"]) for L in H.levels],
                                                          list(c["ratios"])),
                   "cells_per_level": ncell, "cell_updates_per_step": cells,
                   "precision": pname, "graphs": bool(a.graphs)},
    }
    if a.cpu_seconds > 0:
        print("bench_config: CPU baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = _quiet_stdout(lambda: cpu_baseline_config(name, H, cells,
                                                                           a.cpu_seconds))
    print(json.dumps(result))


def cpu_baseline_config(name, H, cells, budget):
    """The reference's own advanc/update (src/2d Fortran, OpenMP over grids)
    over coarse Berger-Oliger steps of the same configuration, about
    `budget` seconds."""
    from oracle import orc
    from amrclaw_amd import workload
    c = workload.CONFIGS[name]
    p = orc.make_params(c["solver"])
    kind = "reference" if orc.ref_available() else "port"
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if kind == "reference":
        impl = orc.RefAMR(p, H.copy())
        impl.set_parallel(True)
    else:
        impl, threads = orc.OracleAMR(p, H.copy()), 1

    def coarse(level=1):
        impl.advanc(level)
        if level < H.nlev:
            for _ in range(H.levels[level - 1]["ratx"]):
                coarse(level + 1)
            impl.update(level)

    n = 0
    t0 = time.perf_counter()
    while n == 0 or time.perf_counter() - t0 < budget:
        coarse()
        n += 1
    wall = time.perf_counter() - t0
    if kind == "reference":
        impl.flush()
    return {"value": cells * n / wall, "unit": "cell-updates/s", "cores": threads, "kind": kind,
            "sample": "%d coarse steps of %s (%d cell updates each) in %.1f s" % (n, name, cells, wall)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.workload == "euler3d":
        bench3(a)
        return
    if a.workload == "amr3d":
        bench_amr3(a)
        return
    if a.workload in ("C1", "C2", "C3"):
        bench_config(a)
        return
    if a.workload == "regrid":
        bench_regrid(a)
        return
    import torch
    from amrclaw_amd import capi, workload

    def sync():
        capi.synchronize()
        torch.cuda.synchronize()

    print("bench: building the workload", file=sys.stderr, flush=True)
    H, cells = workload.batched_euler_hierarchy(a.npatch)
    npatch = len(H.levels[1]["boxes"])
    PREC = {"mixed": capi.MIXED, "fp32": capi.FP32, "fp64": capi.FP64}
    DESC = {"mixed": "FP32 wave and flux arithmetic on FP64 storage (flux form, FP64 update, "
                     "registers, coarse interpolation and totals)",
            "fp32": "FP32 storage, wave and flux arithmetic; FP64 registers, coarse "
                    "interpolation and totals",
            "fp64": "FP64 throughout (the reference's arithmetic, bit for bit)"}
    prec = PREC[a.precision]
    esize = 4 if prec == capi.FP32 else 8     # storage bytes per value
    bytes_per_cell = 2 * 4 * esize            # (2 nvar + maux) * s

    print("bench: timed run", a.precision, file=sys.stderr, flush=True)
    launch = ("per-patch" if a.per_patch else "persistent" if a.persistent else
              "batched" if a.tile else "march")
    dt2 = install(capi, H, prec, launch, a.graphs, a.txmax, a.tymax, a.occupancy)
    r = run(capi, dt2, a.steps, a.warmup, sync)
    q_main = interiors(capi, H) if not a.no_compare else None
    capi.finalize()
    value = cells * a.steps / r["wall"]
    sweep_ms, sweep_n = r["sweep"]
    result = {
        "metric": METRIC,
        "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": r["wall"] / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == capi.FP64 else "f32", "data": "synthetic",
        "config": {"workload": "batched 2-D Euler level update (level 2 of 2): %d patches, "
                               "mx,my even U[8,128] seed 20261202, Roe+efix, MC, "
                               "transverse 2, extrapolation BCs" % npatch,
                   "patches": npatch, "cells_per_step": cells, "nghost": 2,
                   "precision": a.precision, "precision_detail": DESC[a.precision],
                   "launch": launch,
                   "graphs": bool(a.graphs)},
        "step_ms_events": r["events_ms"] / a.steps,
        "kernels_ms_per_step": {k: (r[k][0] / a.steps if r[k][1] else None)
                                for k in ("sweep", "fill", "registers")},
    }
    if sweep_n:
        tuned_default = not (a.txmax or a.tymax or a.occupancy) and launch in ("march", "batched")
        key = ("euler2d", a.precision, npatch, launch)
        tr = MEASURED_TRAFFIC.get(key) if tuned_default else None
        vr = MEASURED_VALU.get(key) if tuned_default else None
        kern_ms = sweep_ms / sweep_n        # one timed record per step's sweep
        achieved = cells * bytes_per_cell / (kern_ms / 1e3) / 1e9
        result["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                              "traffic": tr[0] if tr else None,
                              "traffic_source": tr[1] if tr else None,
                              "kernel": ("sweep_march_kernel<EulerRoe>" if launch == "march"
                                         else "sweep_kernel<EulerRoe>"),
                              "kernel_ms": kern_ms,
                              "bytes_per_cell_update": bytes_per_cell,
                              "algorithmic_bytes_per_launch": cells * bytes_per_cell}
        if vr:
            # the kernel issues VALU work, it does not saturate HBM: the
            # vector instructions it executes per launch (SQ_INSTS_VALU, one
            # count per wave instruction) per cell update, and the issue rate
            # they imply at this run's kernel time
            result["roofline"]["valu"] = {
                "wave_instructions_per_launch": vr[0],
                "lane_instructions_per_cell": vr[0] * 64 / cells,
                "issue_frac": vr[0] * VALU_ISSUE_CYCLES / (SIMD_COUNT * kern_ms * 1e-3 * CLOCK_HZ),
                "source": vr[1]}
    if not a.no_compare and prec != capi.FP64:
        h2 = H.levels[1]["hx"]

        def mass(qs):
            return sum(float(x[..., 0].sum()) for x in qs) * h2 * h2

        runs = {}
        for name, p_, per_patch in (("fp64_batched", capi.FP64, "batched"),
                                    ("fp64_per_patch", capi.FP64, "per-patch"),
                                    ("fp32" if prec == capi.MIXED else "mixed",
                                     capi.FP32 if prec == capi.MIXED else capi.MIXED, launch)):
            print("bench:", name, file=sys.stderr, flush=True)
            install(capi, H, p_, per_patch, False, 0, 0)
            rr = run(capi, dt2, a.steps, a.warmup, sync)
            runs[name] = (rr, interiors(capi, H))
            capi.finalize()
        r64, q64 = runs["fp64_batched"]
        rpp, qpp = runs["fp64_per_patch"]
        den = sum(float(np.abs(y).sum()) for y in q64)
        m64 = mass(q64)

        def l1(qs):
            return sum(float(np.abs(x - y).sum()) for x, y in zip(qs, q64)) / den

        result["fp64_batched"] = {"value": cells * a.steps / r64["wall"],
                                  "ms_per_step": r64["wall"] / a.steps * 1e3,
                                  "sweep_ms": r64["sweep"][0] / max(r64["sweep"][1], 1)}
        result["fp64_per_patch"] = {"value": cells * a.steps / rpp["wall"],
                                    "ms_per_step": rpp["wall"] / a.steps * 1e3,
                                    "sweep_launches_per_step": npatch,
                                    "identical_to_fp64_batched":
                                        all(np.array_equal(x, y) for x, y in zip(q64, qpp))}
        other = "fp32" if prec == capi.MIXED else "mixed"
        ro, qo = runs[other]
        result[other] = {"value": cells * a.steps / ro["wall"],
                         "ms_per_step": ro["wall"] / a.steps * 1e3,
                         "sweep_ms": ro["sweep"][0] / max(ro["sweep"][1], 1),
                         "precision_detail": DESC[other],
                         "speedup_vs_fp64_per_patch": rpp["wall"] / ro["wall"],
                         "l1_rel_vs_fp64": l1(qo),
                         "mass_rel_diff_vs_fp64": abs(mass(qo) - m64) / abs(m64)}
        result["speedup_vs_fp64_per_patch"] = rpp["wall"] / r["wall"]
        result["speedup_vs_fp64_batched"] = r64["wall"] / r["wall"]
        result["l1_rel_vs_fp64"] = l1(q_main)
        result["mass_rel_diff_vs_fp64"] = abs(mass(q_main) - m64) / abs(m64)
        # FP32 storage: rounding the FP64 initial state alone moves the
        # level's mass by this much
        r64_0 = r32_0 = 0.0
        for k in range(npatch):
            rho = H.grid(2, k, 0)[2:-2, 2:-2, 0]
            r64_0 += float(rho.sum())
            r32_0 += float(rho.astype(np.float32).astype(np.float64).sum())
        result["fp32_storage_initial_mass_rel_diff"] = abs(r32_0 - r64_0) / abs(r64_0)
        result["steps_compared"] = a.steps + a.warmup
        print("bench: periodic conservation", file=sys.stderr, flush=True)
        result["conservation_drift"] = {
            "setup": "periodic two-level hierarchy, 256 level-2 patches, 4 coarse steps "
                     "(each: level 1, two level-2 steps, update with refluxing); "
                     "|conck(1) - conck0(1)| / conck0(1)",
            "launch": launch,
            "mixed": conservation_drift(capi, capi.MIXED, launch=launch),
            "fp32": conservation_drift(capi, capi.FP32, launch=launch),
            "fp64": conservation_drift(capi, capi.FP64, launch=launch)}
    if a.cpu_seconds > 0:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = _quiet_stdout(lambda: cpu_baseline(H, dt2, a.cpu_seconds))
        # the same sample on one core (the reference's OpenMP loops serial)
        result["cpu_baseline_1thread"] = _quiet_stdout(
            lambda: cpu_baseline(H, dt2, a.cpu_seconds / 2, nsample=64, parallel=False))
    print(json.dumps(result))


if __name__ == "__main__":
    main()
