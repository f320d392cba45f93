"""Edge cases of the device paths against the oracle: the smallest and most
ragged grids (one cell wide, two cells high, far wider than a tile), a
level without grids (the batch entry and a hierarchy whose finer level is
empty), in 2-D and 3-D.  FP64 throughout, so every comparison is bit for
bit."""
import numpy as np
import pytest

from amrclaw_amd import capi, capi3
from amrclaw_amd.hierarchy import Hier3
from oracle import orc
from tests.common import (both_params, berger_oliger, nested_hierarchy, random_patch,
                          smooth_euler)
from tests.common3 import berger_oliger3, nested_hierarchy3, smooth_euler3

pytestmark = pytest.mark.gpu

RAGGED = [(1, 1), (1, 6), (6, 1), (2, 3), (3, 2), (255, 2), (2, 255), (62, 1)]


@pytest.mark.parametrize("solver", [orc.ADVECTION, orc.ACOUSTICS, orc.EULER])
def test_stepgrid_ragged_fp64_bit_exact(gpu, solver):
    p, c = both_params(solver)
    capi.init(c)
    try:
        rng = np.random.default_rng(31 + int(solver))
        for mx, my in RAGGED:
            print("ragged stepgrid", solver, (mx, my), flush=True)
            q = random_patch(rng, solver, mx, my)
            dx, dy, dt = 1.0 / mx, 1.0 / my, 0.2 / max(mx, my)
            qo, _, co, do = orc.stepgrid(p, q, dx, dy, dt)
            qg, _, cg, dg = capi.stepgrid(q, dx, dy, dt)
            assert np.array_equal(qg[2:-2, 2:-2], qo[2:-2, 2:-2]), (mx, my)
            assert cg == co and dg == do, (mx, my)
    finally:
        capi.finalize()


def test_sweep_level_without_grids(gpu):
    """A level with no grids sweeps to nothing: Courant number 0 and no
    time-step limit (stepgrid.f's dtnew of an empty level, advanc.f:124)."""
    p, c = both_params(orc.EULER)
    capi.init(c)
    try:
        print("empty level sweep", flush=True)
        capi.set_level(1, np.zeros((0, 4), np.int32), 1.0 / 64, 1.0 / 64)
        cfl, dtn = capi.sweep_level(1, 0.001)
        assert cfl == 0.0 and dtn >= 1e30
    finally:
        capi.finalize()


def test_level_update_empty_finer_level(gpu):
    """A hierarchy whose level 2 has no grids advances level 1 exactly as a
    one-level hierarchy does (no saveqc, registers or averaging to do)."""
    print("empty finer level 2-D", flush=True)
    rng = np.random.default_rng(41)
    H1 = nested_hierarchy(rng, smooth_euler, 4, nlev=1, dt1=0.004)
    He = H1.copy()
    L1 = H1.levels[0]
    He.add_level(np.zeros((0, 4), np.int32), L1["hx"] / 2, L1["hy"] / 2, L1["possk"] / 2, 0.0,
                 L1["iregsz"] * 2, L1["jregsz"] * 2)
    p, cfg = both_params(orc.EULER)
    O = orc.OracleAMR(p, H1.copy())
    capi.init(cfg)
    try:
        capi.load_hierarchy(He)

        class D:
            advanc = staticmethod(capi.advance_level)
            update = staticmethod(capi.update_level)

        for _ in range(2):
            berger_oliger(O, 1, 1, 2)
            berger_oliger(D, 1, 2, 2)
        assert np.array_equal(capi.read_level(He, 1, 0), O.H.levels[0]["qnew"])
    finally:
        capi.finalize()


def test_level_update3_empty_finer_level(gpu):
    print("empty finer level 3-D", flush=True)
    rng = np.random.default_rng(42)
    H1 = nested_hierarchy3(rng, smooth_euler3, 5, nlev=1, dt1=0.004)
    He = H1.copy()
    L1 = H1.levels[0]
    He.add_level(np.zeros((0, 6), np.int32), (L1["hx"] / 2, L1["hy"] / 2, L1["hz"] / 2),
                 L1["possk"] / 2, 0.0, (L1["iregsz"] * 2, L1["jregsz"] * 2, L1["kregsz"] * 2))
    assert isinstance(He, Hier3) and He.nlev == 2
    p = orc.make_params3(orc.EULER)
    O = orc.OracleAMR3(p, H1.copy())
    capi3.init3(capi3.make_config3(orc.EULER))
    try:
        D = capi3.DeviceAMR3(He)
        for _ in range(2):
            berger_oliger3(O, O.H, 1)
            berger_oliger3(D, He, 1)
        assert np.array_equal(capi3.read_level3(He, 1, 0), O.H.levels[0]["qnew"])
        assert capi3.conck_level3(1) == O.conck(1)
    finally:
        capi3.finalize3()
