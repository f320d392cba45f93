"""CPU pin of the oracle's coarse-fine flux registers and level update
(prepf/prepc/setuse, saveqc/cstore, qad, fluxsv, fluxad, upbnd, update)
against the reference's own Fortran, through Berger-Oliger steps on properly
nested hierarchies; plus the conservation invariant of refluxing."""
import numpy as np
import pytest

from oracle import orc
from tests.common import (acoustic_states, berger_oliger, nested_hierarchy,
                          smooth_euler)

needs_ref = pytest.mark.skipif(not orc.ref_available(), reason="reference build absent")

PHYS = {"acoustics": (orc.ACOUSTICS, 3, acoustic_states),
        "euler": (orc.EULER, 4, smooth_euler)}
BCS = [((1, 1, 1, 1), False, False), ((3, 3, 3, 3), False, False),
       ((2, 2, 2, 2), True, True), ((2, 2, 3, 1), True, False)]


@needs_ref
@pytest.mark.parametrize("phys", ["acoustics", "euler"])
@pytest.mark.parametrize("case", BCS)
@pytest.mark.parametrize("seed", range(3))
def test_level_update_matches_reference(phys, case, seed):
    solver, nvar, states = PHYS[phys]
    bc, xper, yper = case
    rng = np.random.default_rng(300 + seed)
    H = nested_hierarchy(rng, states, nvar, mthbc=bc, xper=xper, yper=yper,
                         dt1=0.004 if phys == "acoustics" else 0.006)
    p = orc.make_params(solver)
    O = orc.OracleAMR(p, H.copy())
    R = orc.RefAMR(p, H)
    for l in range(1, H.nlev):
        for k in range(len(H.levels[l - 1]["boxes"])):
            assert np.array_equal(O.listbc(l, k), R.listbc(l, k)), ("listbc", l, k)
    lo, lr = [], []
    for step in range(2):
        berger_oliger(O, 1, H.nlev, 2, lo)
        berger_oliger(R, 1, H.nlev, 2, lr)
        assert lo == lr, "dtlevnew / cfl sequence"
        for l in range(1, H.nlev + 1):
            for k in range(len(H.levels[l - 1]["boxes"])):
                assert np.array_equal(O.H.grid(l, k, 0), R.grid(l, k, 0)), ("new", step, l, k)
                if l < H.nlev:
                    assert np.array_equal(O.H.grid(l, k, 1), R.grid(l, k, 1)), ("old", step, l, k)
                if l > 1:
                    assert np.array_equal(O.reg(l, k), R.reg(l, k)), ("reg", step, l, k)


@pytest.mark.parametrize("phys", ["acoustics", "euler"])
def test_periodic_composite_mass_is_conserved(phys):
    """With periodic boundaries, refluxing and averaging keep the level-1
    mass (conck.f) constant to rounding over coarse steps."""
    solver, nvar, states = PHYS[phys]
    rng = np.random.default_rng(42)
    H = nested_hierarchy(rng, states, nvar, mthbc=(2, 2, 2, 2), xper=True, yper=True,
                         dt1=0.004 if phys == "acoustics" else 0.006)
    p = orc.make_params(solver)
    O = orc.OracleAMR(p, H)
    # start from averaged coarse data, as the reference does after setgrd
    # (setgrd.f:127 calls update level by level)
    for l in range(H.nlev - 1, 0, -1):
        O.update(l)
    m0 = O.conck(1)
    for _ in range(3):
        berger_oliger(O, 1, H.nlev, 2)
        m = O.conck(1)
        assert abs(m - m0) <= 1e-13 * abs(m0) + 1e-14, (m, m0)
