"""GPU parity of the march sweep (launch mode AMRGPU_LAUNCH_MARCH: each grid
cut into column strips that march up the rows with every quantity in
registers, sweep_march.hpp) against the oracle's stepgrid
(src/2d/stepgrid.f:24-234 with step2.f90 / flux2.f): one level of grids
whose widths exercise every strip cut (one to three strips, 16-, 32- and
64-lane segments, strips of one column) and whose heights run from 1 to 70.
FP64 bit for bit (updated interior, ghost ring of the new array, old array,
cflgrid, dtnew) for every solver and several methods; FP32 arithmetic
(mixed and FP32 storage) within 1e-5 of the oracle and within 1e-6 of the
tile sweep in the same precision (the same arithmetic, contracted
differently).  The whole level update with the march sweep is covered by
tests/test_gpu_amr.py (launch variants) and tests/test_gpu_c4.py."""
import numpy as np
import pytest

from amrclaw_amd import capi
from oracle import orc
from tests.common import SOLVERS, both_params, random_patch, rel_l1

pytestmark = pytest.mark.gpu

WIDTHS = [1, 2, 5, 8, 12, 13, 28, 29, 30, 56, 60, 61, 62, 88, 90, 127, 130]


def _level(rng, solver):
    heights = [1, 2, 3] + list(rng.integers(4, 71, size=len(WIDTHS) - 3))
    boxes, qs, x0 = [], [], 0
    for mx, my in zip(WIDTHS, heights):
        boxes.append((x0, x0 + mx - 1, 0, int(my) - 1))
        qs.append(random_patch(rng, solver, mx, int(my)))
        x0 += mx
    return boxes, qs


def _sweep(cfg, boxes, qs, mode, dt, h):
    capi.init(cfg)
    try:
        capi.set_launch_mode(mode)
        capi.set_level(1, boxes, h, h)
        for k, q in enumerate(qs):
            capi.put_grid(1, k, q)
        cfl, dtn = capi.sweep_level(1, dt)
        new = [capi.get_grid(1, k, q.shape) for k, q in enumerate(qs)]
        old = [capi.get_grid(1, k, q.shape, which=1) for k, q in enumerate(qs)]
    finally:
        capi.finalize()
    return new, old, cfl, dtn


METHODS = [dict(), dict(transverse=1), dict(order=1), dict(limiter=1), dict(limiter=2),
           dict(limiter=3), dict(limiter=5), dict(limiter=0), dict(transverse=0)]


@pytest.mark.parametrize("solver", ["advection", "acoustics", "euler"])
@pytest.mark.parametrize("meth", range(len(METHODS)))
def test_march_fp64_bit_exact(gpu, solver, meth):
    sid = SOLVERS[solver]
    kw = METHODS[meth]
    p, cfg = both_params(sid, precision=capi.FP64, **kw)
    rng = np.random.default_rng(500 + 10 * sid + meth)
    boxes, qs = _level(rng, sid)
    h = 1.0 / 256
    dt = (0.1 if sid == orc.EULER else 0.3) * h
    print("march fp64", solver, kw, flush=True)
    new, old, cfl, dtn = _sweep(cfg, boxes, qs, capi.LAUNCH_MARCH, dt, h)
    cmax, dmin = 0.0, float(np.float32(10e32))
    for k, q in enumerate(qs):
        qo, _, co, do = orc.stepgrid(p, q, h, h, dt)
        assert np.array_equal(new[k][2:-2, 2:-2], qo[2:-2, 2:-2]), (k, WIDTHS[k])
        ring = np.ones(q.shape[:2], bool)
        ring[2:-2, 2:-2] = False
        assert np.array_equal(new[k][ring], q[ring]), k        # ghosts keep time t
        assert np.array_equal(old[k], q), k
        cmax, dmin = max(cmax, co), min(dmin, do)
    assert cfl == cmax and dtn == dmin


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP32])
@pytest.mark.parametrize("solver", ["acoustics", "euler"])
def test_march_fp32_arithmetic(gpu, solver, precision):
    sid = SOLVERS[solver]
    p, cfg = both_params(sid, precision=precision)
    rng = np.random.default_rng(900 + sid + 10 * precision)
    boxes, qs = _level(rng, sid)
    h = 1.0 / 256
    dt = (0.1 if sid == orc.EULER else 0.3) * h
    print("march fp32 arithmetic", solver, precision, flush=True)
    new, _, cfl, _ = _sweep(cfg, boxes, qs, capi.LAUNCH_MARCH, dt, h)
    tile, _, cfl_t, _ = _sweep(cfg, boxes, qs, capi.LAUNCH_BATCHED, dt, h)
    for k, q in enumerate(qs):
        qo, _, co, _ = orc.stepgrid(p, q, h, h, dt)
        assert rel_l1(new[k][2:-2, 2:-2], qo[2:-2, 2:-2]) <= 1e-5, k
        assert rel_l1(new[k][2:-2, 2:-2], tile[k][2:-2, 2:-2]) <= 1e-6, k
    assert abs(cfl - cfl_t) <= 1e-6 * cfl_t
