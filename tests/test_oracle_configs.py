"""CPU pin of the oracle's level update on the three 2-D run configurations
(advection, periodic, ratios 2, 2; acoustics, ratios 2, 4; Euler quadrants
over four level-1 grids, ratios 2, 4) against the reference's own Fortran
through Berger-Oliger steps: solution, old-time arrays, registers and the
dt / CFL sequence bit for bit."""
import numpy as np
import pytest

from oracle import orc
from tests.configs import CONFIGS, berger_oliger, config_hierarchy

needs_ref = pytest.mark.skipif(not orc.ref_available(), reason="reference build absent")


def test_config_hierarchies_are_nested():
    """Every level-(l+1) grid, coarsened and grown by one coarse cell, lies
    inside one level-l grid; grids of a level are disjoint."""
    for name in CONFIGS:
        H = config_hierarchy(name)
        for l in range(2, H.nlev + 1):
            r = H.levels[l - 2]["ratx"]
            cover = np.zeros((H.levels[l - 1]["jregsz"], H.levels[l - 1]["iregsz"]), dtype=int)
            for b in H.levels[l - 1]["boxes"]:
                cover[b[2]:b[3] + 1, b[0]:b[1] + 1] += 1
                c = (b[0] // r - 1, (b[1] + 1) // r, b[2] // r - 1, (b[3] + 1) // r)
                assert b[0] % r == 0 and (b[1] + 1) % r == 0
                assert any(p[0] <= c[0] and c[1] <= p[1] and p[2] <= c[2] and c[3] <= p[3]
                           for p in H.levels[l - 2]["boxes"]), (name, l, b)
            assert cover.max() == 1, (name, l)


@needs_ref
@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_config_level_update_matches_reference(name):
    c = CONFIGS[name]
    H = config_hierarchy(name)
    p = orc.make_params(c["solver"])
    O = orc.OracleAMR(p, H.copy())
    R = orc.RefAMR(p, H)
    lo, lr = [], []
    for step in range(2):
        berger_oliger(O, O.H, 1, lo)
        berger_oliger(R, H, 1, lr)
        assert lo == lr, "dtlevnew / cfl sequence"
        for l in range(1, H.nlev + 1):
            for k in range(len(H.levels[l - 1]["boxes"])):
                assert np.array_equal(O.H.grid(l, k, 0), R.grid(l, k, 0)), ("new", step, l, k)
                if l < H.nlev:
                    assert np.array_equal(O.H.grid(l, k, 1), R.grid(l, k, 1)), ("old", step, l, k)
                if l > 1:
                    assert np.array_equal(O.reg(l, k), R.reg(l, k)), ("reg", step, l, k)


@pytest.mark.parametrize("name", ["C1"])
def test_config_periodic_mass_conserved(name):
    """C1 is periodic: with coarse data averaged from the finer levels,
    refluxing keeps the level-1 mass (conck.f) to rounding."""
    H = config_hierarchy(name)
    O = orc.OracleAMR(orc.make_params(CONFIGS[name]["solver"]), H)
    for l in range(H.nlev - 1, 0, -1):
        O.update(l)
    m0 = O.conck(1)
    for _ in range(3):
        berger_oliger(O, H, 1)
        assert abs(O.conck(1) - m0) <= 1e-13 * abs(m0)


needs_ref3 = pytest.mark.skipif(not orc.ref3_available(), reason="3-D reference build absent")


@needs_ref3
def test_config_c5_level_update_matches_reference():
    """C5 (3-D Euler, radial bump rho = p = 2 for r < 0.2 about the centre
    of the unit cube, extrapolation BCs, 64^3 level 1 in eight 32^3 grids,
    two levels at ratio 2): the first level steps of a coarse step (level 1,
    level 2, two of level 3, the level-2 update) of the oracle bit for bit
    with the reference's src/3d Fortran."""
    from tests.common3 import berger_oliger3
    from tests.configs3 import config_hierarchy3
    H = config_hierarchy3()
    assert [len(L["boxes"]) for L in H.levels] == [8, 76, 263]
    p = orc.make_params3(orc.EULER)
    O = orc.OracleAMR3(p, H.copy())
    R = orc.RefAMR3(p, H)
    lo, lr = [], []
    for impl, log in ((O, lo), (R, lr)):
        log.append(impl.advanc(1))
        berger_oliger3(impl, H, 2, log)
    assert lo == lr
    for l in range(1, H.nlev + 1):
        for k in range(len(H.levels[l - 1]["boxes"])):
            assert np.array_equal(O.H.grid(l, k, 0), R.grid(l, k, 0)), (l, k)
            if l > 1:
                assert np.array_equal(O.reg(l, k), R.reg(l, k)), ("reg", l, k)
