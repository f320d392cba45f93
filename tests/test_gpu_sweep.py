"""GPU parity of the fused sweep (step2 + flux2 + limiter + update + CFL)
against the oracle, through the C ABI.

FP64: bit-exact (the kernel keeps the reference's evaluation order and is
built without FMA contraction).  FP32 bulk: relative L1 <= 1e-5 of the FP64
oracle for one step, the tolerance the batched-FP32 target states.
"""
import numpy as np
import pytest

from amrclaw_amd import capi
from oracle import orc
from tests.common import both_params, random_patch, rel_l1

pytestmark = pytest.mark.gpu

CASES = [(orc.ADVECTION, 2, 2, 4), (orc.ADVECTION, 2, 1, 3), (orc.ADVECTION, 1, 0, 0),
         (orc.ACOUSTICS, 2, 2, 4), (orc.ACOUSTICS, 2, 2, 2), (orc.ACOUSTICS, 2, 0, 1),
         (orc.EULER, 2, 2, 4), (orc.EULER, 2, 2, 1), (orc.EULER, 2, 1, 5),
         (orc.EULER, 1, 2, 0), (orc.EULER, 2, 2, 3)]
SIZES = [(7, 5), (28, 12), (61, 37), (128, 9), (9, 128), (100, 100)]


def _interior_fluxes(f, mx, my):
    fm, fp, gm, gp = f
    # x-edges 1..mx+1 of interior rows, y-edges 1..my+1 of interior columns
    return [fm[2:my + 2, 2:mx + 3], fp[2:my + 2, 2:mx + 3],
            gm[2:my + 3, 2:mx + 2], gp[2:my + 3, 2:mx + 2]]


@pytest.mark.parametrize("case", CASES)
def test_stepgrid_fp64_bit_exact(gpu, case):
    solver, order, trans, lim = case
    p, c = both_params(solver, order=order, transverse=trans, limiter=lim)
    capi.init(c)
    rng = np.random.default_rng(hash(case) % 2**32)
    for mx, my in SIZES:
        print("stepgrid fp64", case, (mx, my), flush=True)
        q = random_patch(rng, solver, mx, my)
        dx, dy, dt = 1.0 / mx, 1.0 / my, 0.2 / max(mx, my)
        qo, fo, co, do = orc.stepgrid(p, q, dx, dy, dt)
        qg, fg, cg, dg = capi.stepgrid(q, dx, dy, dt, want_fluxes=True)
        assert np.array_equal(qg[2:-2, 2:-2], qo[2:-2, 2:-2]), (mx, my)
        for a, b in zip(_interior_fluxes(fg, mx, my), _interior_fluxes(fo, mx, my)):
            assert np.array_equal(a, b), (mx, my)
        assert cg == co and dg == do
    capi.finalize()


@pytest.mark.parametrize("case", [c for c in CASES if c[0] != orc.ADVECTION])
def test_stepgrid_fp32_within_tolerance(gpu, case):
    solver, order, trans, lim = case
    p, c = both_params(solver, order=order, transverse=trans, limiter=lim,
                       precision=capi.FP32)
    capi.init(c)
    rng = np.random.default_rng(5 + hash(case) % 2**31)
    for mx, my in SIZES[:4]:
        print("stepgrid fp32", case, (mx, my), flush=True)
        q = random_patch(rng, solver, mx, my)
        dx, dy, dt = 1.0 / mx, 1.0 / my, 0.2 / max(mx, my)
        qo, _, co, _ = orc.stepgrid(p, q, dx, dy, dt)
        qg, _, cg, _ = capi.stepgrid(q, dx, dy, dt)
        assert rel_l1(qg[2:-2, 2:-2], qo[2:-2, 2:-2]) <= 1e-5
        assert abs(cg - co) <= 1e-5 * co
    capi.finalize()


@pytest.mark.parametrize("precision", [capi.FP64, capi.FP32])
@pytest.mark.parametrize("tiling", [(28, 12), (12, 12), (60, 12)])
def test_level_sweep_many_patches(gpu, precision, tiling):
    """One launch over a level of uneven patches equals per-patch stepgrid."""
    solver = orc.EULER
    p, c = both_params(solver, precision=precision)
    capi.init(c)
    capi.set_tiling(*tiling)
    rng = np.random.default_rng(20261202)
    n = 40
    mxs = rng.integers(4, 65, size=n) * 2
    mys = rng.integers(4, 65, size=n) * 2
    boxes, x0 = [], 0
    for mx, my in zip(mxs, mys):
        boxes.append((x0, x0 + mx - 1, 0, my - 1))
        x0 += mx
    hx = hy = 1.0 / 512
    print("level sweep", precision, tiling, flush=True)
    capi.set_level(1, boxes, hx, hy)
    qs = []
    for k, (mx, my) in enumerate(zip(mxs, mys)):
        q = random_patch(rng, solver, int(mx), int(my))
        qs.append(q)
        capi.put_grid(1, k, q)
    dt = 0.2 * hx
    cfl, dtnew = capi.sweep_level(1, dt)
    cmax, dmin = 0.0, float(np.float32(10e32))
    for k, (mx, my) in enumerate(zip(mxs, mys)):
        qo, _, co, do = orc.stepgrid(p, qs[k], hx, hy, dt)
        qg = capi.get_grid(1, k, qs[k].shape)
        if precision == capi.FP64:
            assert np.array_equal(qg[2:-2, 2:-2], qo[2:-2, 2:-2]), k
        else:
            assert rel_l1(qg[2:-2, 2:-2], qo[2:-2, 2:-2]) <= 1e-5, k
        # the old-time array keeps the input including ghosts
        qold = capi.get_grid(1, k, qs[k].shape, which=1)
        if precision == capi.FP64:
            assert np.array_equal(qold, qs[k])
        cmax = max(cmax, co)
        dmin = min(dmin, do)
    if precision == capi.FP64:
        assert cfl == cmax and dtnew == dmin
    else:
        assert abs(cfl - cmax) <= 1e-5 * cmax
    capi.finalize()


@pytest.mark.parametrize("precision", [capi.FP64, capi.FP32])
def test_bulk_level_transfer_round_trip(gpu, precision):
    """put_level/get_level (alloc + per-grid offsets) equal per-grid puts."""
    _, c = both_params(orc.EULER, precision=precision)
    capi.init(c)
    rng = np.random.default_rng(9)
    boxes = [(0, 9, 0, 5), (10, 41, 0, 17), (42, 49, 3, 30)]
    capi.set_level(2, boxes, 0.01, 0.01)
    qs = [random_patch(rng, orc.EULER, b[1] - b[0] + 1, b[3] - b[2] + 1) for b in boxes]
    capi.put_level(2, qs, which=0)
    back = capi.get_level(2, [q.shape for q in qs], which=0)
    for k, q in enumerate(qs):
        one = capi.get_grid(2, k, q.shape)
        ref = q if precision == capi.FP64 else q.astype(np.float32).astype(np.float64)
        assert np.array_equal(back[k], ref) and np.array_equal(one, ref)
    capi.finalize()


@pytest.mark.parametrize("case", [c for c in CASES if c[1] == 2][:6])
def test_stepgrid_mixed_within_tolerance(gpu, case):
    """Mixed path (FP32 waves on FP64 storage, flux form, FP64 update): one
    step within 1e-5 relative L1 of the FP64 oracle; ghost cells untouched."""
    solver, order, trans, lim = case
    p, c = both_params(solver, order=order, transverse=trans, limiter=lim,
                       precision=capi.MIXED)
    capi.init(c)
    rng = np.random.default_rng(17 + hash(case) % 2**31)
    for mx, my in SIZES[:4]:
        print("stepgrid mixed", case, (mx, my), flush=True)
        q = random_patch(rng, solver, mx, my)
        dx, dy, dt = 1.0 / mx, 1.0 / my, 0.2 / max(mx, my)
        qo, _, co, _ = orc.stepgrid(p, q, dx, dy, dt)
        qg, _, cg, _ = capi.stepgrid(q, dx, dy, dt)
        assert rel_l1(qg[2:-2, 2:-2], qo[2:-2, 2:-2]) <= 1e-5
        assert np.array_equal(qg[:2], q[:2]) and np.array_equal(qg[:, -2:], q[:, -2:])
        assert abs(cg - co) <= 1e-5 * co
    capi.finalize()


def test_stepgrid_mixed_conserves_to_fp64_rounding(gpu):
    """Periodic-like check of the flux form: with all ghost cells equal to a
    constant far state, sum over the interior changes only by the boundary
    fluxes; the mixed path's update telescopes, so the FP64 interior total
    equals q_total - dt/dx * (F_right - F_left) summed over rows to FP64
    rounding, where F are the boundary fluxes of the uniform ghost state
    (both ghost neighbours are the same state on every row)."""
    solver = orc.EULER
    p, c = both_params(solver, precision=capi.MIXED)
    capi.init(c)
    rng = np.random.default_rng(5)
    mx, my = 40, 30
    q = random_patch(rng, solver, mx, my)
    far = q[0, 0].copy()
    ring = np.ones(q.shape[:2], dtype=bool)
    ring[3:-3, 3:-3] = False   # interior cells next to the ghosts also uniform
    q[ring] = far
    dx = dy = 1.0 / 64
    dt = 0.2 * dx
    qg, _, _, _ = capi.stepgrid(q, dx, dy, dt)
    capi.finalize()
    m0 = float(q[2:-2, 2:-2, 0].sum())
    m1 = float(qg[2:-2, 2:-2, 0].sum())
    # the boundary fluxes of a uniform state cancel between opposite sides
    assert abs(m1 - m0) <= 1e-13 * abs(m0), (m1 - m0) / m0


@pytest.mark.parametrize("precision", [capi.FP32, capi.MIXED])
def test_fp32_limiter_tiny_waves(gpu, precision):
    """Tails below 1e-19 next to exact zeros (as an advected step profile
    leaves on a fine level): in FP32 |wave|^2 is subnormal there and the
    neighbour dot product is zero; the limiter ratio must stay 0, not
    0 * (1 / subnormal) = NaN.  The step matches the FP64 oracle."""
    p, c = both_params(orc.ADVECTION, precision=precision)
    capi.init(c)
    mx, my = 40, 24
    q = np.zeros((my + 4, mx + 4, 1))
    q[6:14, 8:20, 0] = 1.0
    tails = [4.2e-21, 3.5e-28, 1e-20, 6e-38, 2e-25]
    for k, v in enumerate(tails):
        q[4 + 3 * k, 26 + k, 0] = v
        q[20, 5 + 4 * k, 0] = v
    if precision == capi.FP32:
        q = q.astype(np.float32).astype(np.float64)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.4 * dy
    print("tiny waves", precision, flush=True)
    qo, _, co, _ = orc.stepgrid(p, q, dx, dy, dt)
    qg, _, cg, _ = capi.stepgrid(q, dx, dy, dt)
    capi.finalize()
    assert np.isfinite(qg).all()
    assert np.abs(qg - qo)[2:-2, 2:-2].max() <= 1e-6
    assert rel_l1(qg[2:-2, 2:-2], qo[2:-2, 2:-2]) <= 1e-5


def test_mixed_occupancy_bounds_agree(gpu):
    """The mixed sweep's occupancy variants (register bounds of 5, 6 and 7
    workgroups per CU, the autotuner's knob) run the same arithmetic: each
    stays within 1e-5 of the FP64 oracle and within FP32 rounding of the
    others; out-of-range bounds are refused."""
    solver = orc.EULER
    p, c = both_params(solver, precision=capi.MIXED)
    capi.init(c)
    try:
        rng = np.random.default_rng(77)
        q = random_patch(rng, solver, 61, 37)
        dx, dy, dt = 1.0 / 61, 1.0 / 37, 0.2 / 61
        qo, _, co, _ = orc.stepgrid(p, q, dx, dy, dt)
        got = {}
        for wgs in (5, 6, 7, 0):
            print("mixed occupancy", wgs, flush=True)
            capi.set_occupancy(wgs)
            got[wgs], _, cg, _ = capi.stepgrid(q, dx, dy, dt)
            assert rel_l1(got[wgs][2:-2, 2:-2], qo[2:-2, 2:-2]) <= 1e-5, wgs
            assert abs(cg - co) <= 1e-5 * co
        for wgs in (5, 6):
            assert rel_l1(got[wgs], got[7]) <= 1e-6, wgs
        assert np.array_equal(got[0], got[7])
        for bad in (4, 8, -1):
            with pytest.raises(capi.AmrgpuError):
                capi.set_occupancy(bad)
    finally:
        capi.finalize()
