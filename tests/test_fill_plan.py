"""CPU checks of the ghost-fill planner (libamrgpu's host side): the planned
program, executed with numpy (tests/fill_exec.py), equals the oracle's
bound on random hierarchies.  Needs no device."""
import numpy as np
import pytest

from amrclaw_amd import capi
from oracle import orc
from tests.common import random_hierarchy
from tests.fill_exec import execute

CASES = [((1, 1, 1, 1), False, False), ((3, 3, 3, 3), False, False),
         ((1, 3, 3, 1), False, False), ((2, 2, 2, 2), True, True),
         ((2, 2, 1, 3), True, False), ((3, 1, 2, 2), False, True)]


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("phase", [(0, 1), (1, 1), (2, 1)])
def test_planned_fill_equals_oracle_bound(seed, case, phase):
    bc, xper, yper = case
    rng = np.random.default_rng(1000 + seed)
    H = random_hierarchy(rng, nvar=3, mthbc=bc, xper=xper, yper=yper, fine_phase=phase)
    p = orc.make_params(orc.ACOUSTICS)
    for l in range(1, H.nlev + 1):
        t = H.levels[l - 1]["time"]
        Ho = H.copy()
        for k in range(len(H.levels[l - 1]["boxes"])):
            orc.bound(p, Ho, l, k, t)
        Hp = H.copy()
        ops, bufs, masks = capi.plan_fill(Hp, l)
        execute(Hp, l, t, ops, bufs, masks, bc)
        for k in range(len(H.levels[l - 1]["boxes"])):
            assert np.array_equal(Hp.grid(l, k), Ho.grid(l, k)), (l, k)


def test_plan_rejects_unset_level1_cells():
    """A level-1 layout with a hole cannot be filled (filpatch.f90:168)."""
    from tests.common import Hier
    H = Hier(1)
    H.add_level([(0, 4, 0, 9), (6, 9, 0, 9)], 0.1, 0.1, 0.1, 0.0, 10, 10)
    with pytest.raises(capi.AmrgpuError, match="level 1"):
        capi.plan_fill(H, 1)
