"""CPU pin of the oracle's 3-D level update around the step against the
reference's own src/3d Fortran: ghost filling (bound.f, filpatch.f,
intfil.f, prefilp.f, trimbd.f, bc3amr.f90), the coarse-fine registers
(prepf, prepc, setuse, saveqc, cstore, qad, fluxsv, fluxad) and
refluxing/averaging (upbnd, update) through Berger-Oliger steps; plus closed
forms: interpolation of linear data, conservation with periodic domains."""
import numpy as np
import pytest

from amrclaw_amd.hierarchy import Hier3
from oracle import orc
from tests.common3 import (advection3_states, berger_oliger3, nested_hierarchy3,
                           random_hierarchy3, smooth_euler3)

needs_ref3 = pytest.mark.skipif(not orc.ref3_available(), reason="3-D reference build absent")

BCS = [((1,) * 6, (False,) * 3), ((3,) * 6, (False,) * 3), ((1, 3, 3, 1, 1, 3), (False,) * 3),
       ((2,) * 6, (True,) * 3)]


@needs_ref3
@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("case", BCS)
@pytest.mark.parametrize("phase", [1, 0, 2])
def test_bound3_matches_reference(seed, case, phase):
    bc, per = case
    rng = np.random.default_rng(50 + seed)
    H = random_hierarchy3(rng, mthbc=bc, per=per, fine_phase=phase)
    p = orc.make_params3(orc.EULER)
    for l in range(1, H.nlev + 1):
        t = H.levels[l - 1]["time"]
        for k in range(len(H.levels[l - 1]["boxes"])):
            Ho = H.copy()
            orc.bound3(p, Ho, l, k, t)
            r = orc.ref3_bound(p, H, l, k, t)
            assert np.array_equal(Ho.grid(l, k), r), (l, k)


@needs_ref3
@pytest.mark.parametrize("seed", range(2))
def test_bound3_ratio_mix_matches_reference(seed):
    """Different ratios per direction and per level (intratx/y/z)."""
    rng = np.random.default_rng(90 + seed)
    H = random_hierarchy3(rng, ratios=((2, 4, 2), (4, 2, 2)), n=(8, 6, 6))
    p = orc.make_params3(orc.EULER)
    for l in range(1, H.nlev + 1):
        t = H.levels[l - 1]["time"]
        for k in range(len(H.levels[l - 1]["boxes"])):
            Ho = H.copy()
            orc.bound3(p, Ho, l, k, t)
            assert np.array_equal(Ho.grid(l, k), orc.ref3_bound(p, H, l, k, t)), (l, k)


def test_bound3_interpolates_linear_data():
    """Coarse-fine interpolation (MC-limited slopes in x, y, z, linear time
    blend) is exact for data linear in x, y, z and t away from the edges."""
    rng = np.random.default_rng(3)
    coef = rng.normal(size=(5, 5))
    H = Hier3(5)
    n = 8
    H.add_level([(0, 7, 0, 7, 0, 7)], (1 / n,) * 3, 0.1, 1.0, (n,) * 3)
    H.add_level([(4, 11, 4, 9, 6, 11)], (0.5 / n,) * 3, 0.05, 0.95, (2 * n,) * 3)

    def f(x, y, z, t):
        return np.stack([c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * t for c in coef], -1)

    for l, L in enumerate(H.levels, start=1):
        b = L["boxes"][0]
        k = np.arange(b[4] - 2, b[5] + 3)
        j = np.arange(b[2] - 2, b[3] + 3)
        i = np.arange(b[0] - 2, b[1] + 3)
        Z, Y, X = np.meshgrid((k + 0.5) * L["hz"], (j + 0.5) * L["hy"], (i + 0.5) * L["hx"],
                              indexing="ij")
        H.grid(l, 0, 0)[:] = f(X, Y, Z, L["time"])
        H.grid(l, 0, 1)[:] = f(X, Y, Z, L["time"] - L["possk"] * (2 if l == 1 else 1))
    # level-1 old time: T - possk(1) = 0.9; level 2 at 0.95 blends 1/2, 1/2
    H.levels[0]["qold"][:] = 0
    b = H.levels[0]["boxes"][0]
    k = np.arange(b[4] - 2, b[5] + 3)
    Z, Y, X = np.meshgrid((k + 0.5) / n, (k + 0.5) / n, (k + 0.5) / n, indexing="ij")
    H.grid(1, 0, 1)[:] = f(X, Y, Z, 0.9)
    p = orc.make_params3(orc.EULER)
    want = H.grid(2, 0).copy()
    orc.bound3(p, H, 2, 0, 0.95)
    got = H.grid(2, 0)
    assert np.allclose(got, want, rtol=0, atol=1e-12)


@needs_ref3
@pytest.mark.parametrize("solver", ["euler", "advection"])
@pytest.mark.parametrize("case", BCS[:3] + [BCS[3]])
@pytest.mark.parametrize("seed", range(2))
def test_level_update3_matches_reference(solver, case, seed):
    bc, per = case
    if solver == "advection" and 3 in bc:
        pytest.skip("wall boundaries negate a velocity component advection does not have")
    rng = np.random.default_rng(400 + seed)
    if solver == "euler":
        H = nested_hierarchy3(rng, smooth_euler3, 5, mthbc=bc, per=per, dt1=0.004)
        p = orc.make_params3(orc.EULER)
    else:
        H = nested_hierarchy3(rng, advection3_states, 1, mthbc=bc, per=per, dt1=0.01)
        p = orc.make_params3(orc.ADVECTION)
    assert all(len(L["boxes"]) for L in H.levels)
    O = orc.OracleAMR3(p, H.copy())
    R = orc.RefAMR3(p, H)
    for l in range(1, H.nlev):
        for k in range(len(H.levels[l - 1]["boxes"])):
            assert np.array_equal(O.listbc(l, k), R.listbc(l, k)), ("listbc", l, k)
    lo, lr = [], []
    for step in range(2):
        berger_oliger3(O, O.H, 1, lo)
        berger_oliger3(R, H, 1, lr)
        assert lo == lr, "dtlevnew / cfl sequence"
        for l in range(1, H.nlev + 1):
            for k in range(len(H.levels[l - 1]["boxes"])):
                assert np.array_equal(O.H.grid(l, k, 0), R.grid(l, k, 0)), ("new", step, l, k)
                if l < H.nlev:
                    assert np.array_equal(O.H.grid(l, k, 1), R.grid(l, k, 1)), ("old", step, l, k)
                if l > 1:
                    assert np.array_equal(O.reg(l, k), R.reg(l, k)), ("reg", step, l, k)


@needs_ref3
def test_level_update3_ratio_mix_matches_reference():
    rng = np.random.default_rng(77)
    H = nested_hierarchy3(rng, smooth_euler3, 5, n=(8, 8, 6), ratios=((2, 2, 2), (4, 4, 4)),
                          dt1=0.004, hi=8)
    p = orc.make_params3(orc.EULER)
    O = orc.OracleAMR3(p, H.copy())
    R = orc.RefAMR3(p, H)
    lo, lr = [], []
    berger_oliger3(O, O.H, 1, lo)
    berger_oliger3(R, H, 1, lr)
    assert lo == lr
    for l in range(1, H.nlev + 1):
        for k in range(len(H.levels[l - 1]["boxes"])):
            assert np.array_equal(O.H.grid(l, k, 0), R.grid(l, k, 0)), (l, k)


@pytest.mark.parametrize("solver", ["euler", "advection"])
def test_periodic_composite_mass_is_conserved3(solver):
    """Fully periodic: refluxing and averaging keep the level-1 mass
    (conck.f) constant to rounding over coarse steps."""
    rng = np.random.default_rng(42)
    per = (True,) * 3
    if solver == "euler":
        H = nested_hierarchy3(rng, smooth_euler3, 5, mthbc=(2,) * 6, per=per, dt1=0.004)
        p = orc.make_params3(orc.EULER)
    else:
        H = nested_hierarchy3(rng, advection3_states, 1, mthbc=(2,) * 6, per=per, dt1=0.01)
        p = orc.make_params3(orc.ADVECTION)
    O = orc.OracleAMR3(p, H)
    for l in range(H.nlev - 1, 0, -1):
        O.update(l)
    m0 = O.conck(1)
    for _ in range(3):
        berger_oliger3(O, H, 1)
        assert abs(O.conck(1) - m0) <= 1e-13 * max(abs(m0), 1.0), (O.conck(1), m0)


def test_partial_periodicity_is_refused():
    rng = np.random.default_rng(1)
    H = nested_hierarchy3(rng, smooth_euler3, 5, mthbc=(2, 2, 1, 1, 1, 1),
                          per=(True, False, False), dt1=0.004)
    with pytest.raises(RuntimeError):
        orc.OracleAMR3(orc.make_params3(orc.EULER), H)


@pytest.mark.parametrize("offset, blends", [(0.002, False), (0.05, True)])
def test_bound3_time_tolerance(offset, blends):
    """intfil.f (3-D) treats a source time within possk/10 of its new time
    as the new time (no blend); the 2-D code's tolerance is 1e-4.  With the
    finest level offset by 0.002 possk(2) from level 2, its ghost cells take
    level 2's new values only: scrambling level 2's old values changes
    nothing.  At 0.05 possk(2) they are blended."""
    rng = np.random.default_rng(11)
    H = random_hierarchy3(rng)
    L2 = H.levels[1]
    H.levels[2]["time"] = L2["time"] - offset * L2["possk"]
    p = orc.make_params3(orc.EULER)
    t = H.levels[2]["time"]
    A = H.copy()
    B = H.copy()
    B.levels[1]["qold"][:] = rng.normal(size=B.levels[1]["qold"].shape)
    changed = False
    for k in range(len(H.levels[2]["boxes"])):
        orc.bound3(p, A, 3, k, t)
        orc.bound3(p, B, 3, k, t)
        changed |= not np.array_equal(A.grid(3, k), B.grid(3, k))
    assert changed == blends
    if orc.ref3_available():
        for k in range(len(H.levels[2]["boxes"])):
            assert np.array_equal(A.grid(3, k), orc.ref3_bound(p, H, 3, k, t)), k


@needs_ref3
@pytest.mark.parametrize("case", [BCS[0], BCS[1], BCS[3]])
@pytest.mark.parametrize("seed", range(2))
def test_flagging3_matches_reference(case, seed):
    """spest (bound + flag2refine + setflags) and errest (copysol, bound with
    2 nghost at t - possk, coarsen, the two stepgrids, errf1) of every level
    after a coarse step, bit for bit with the reference's 3-D Fortran."""
    bc, per = case
    rng = np.random.default_rng(600 + seed)
    H = nested_hierarchy3(rng, smooth_euler3, 5, n=(12, 12, 8), mthbc=bc, per=per, dt1=0.004)
    for l, L in enumerate(H.levels, start=1):          # Richardson needs even grids
        assert all((b[1] - b[0] + 1) % 2 == 0 and (b[3] - b[2] + 1) % 2 == 0 and
                   (b[5] - b[4] + 1) % 2 == 0 for b in L["boxes"])
    p = orc.make_params3(orc.EULER)
    O = orc.OracleAMR3(p, H.copy())
    R = orc.RefAMR3(p, H)
    berger_oliger3(O, O.H, 1)
    berger_oliger3(R, H, 1)
    # regrid flags the levels from fine to coarse (regrid.f); errf1 leaves
    # its flags in component 1 of the flagged level's store2, which a later
    # flagging of a finer level would read as old values
    for level in range(H.nlev - 1, 0, -1):
        tolsp, tol = 0.05, 1e-4
        f, rf, n = orc.flag_level3(O, level, tolsp, tol=tol, richardson=True)
        dom, rref = orc.ref3_flag_level(R, level, tolsp, tol=tol, richardson=True)
        assert np.array_equal(orc.domain_flags3(O.H, level, f), dom), level
        assert n > 0
        for k, (a, b) in enumerate(zip(rf, rref)):
            assert np.array_equal(a, b), (level, k)
        assert sum(int((r == 2.0).sum()) for r in rf) > 0


def test_errest3_uniform_state_flags_nothing():
    """A uniform state has no Richardson error: no flags at any positive
    tolerance, every interior 2x2x2 block flagged at tol = 0."""
    rng = np.random.default_rng(5)
    H = nested_hierarchy3(rng, smooth_euler3, 5, n=(12, 12, 8), dt1=0.004)
    for L in H.levels:
        if len(L["boxes"]):
            for which in ("qnew", "qold"):
                L[which].reshape(-1, 5)[:] = (1.1, 0.2, -0.1, 0.05, 2.9)
    O = orc.OracleAMR3(orc.make_params3(orc.EULER), H)
    for L in H.levels:
        L["time"] = L["possk"]
    O.h = orc.hier3_struct(H)
    f, rf, _ = orc.flag_level3(O, 1, 0.05, tol=1e-12, richardson=True)
    assert all(int((r != 0).sum()) == 0 for r in rf)
    assert all(int((x != 0).sum()) == 0 for x in f)
    f, rf, _ = orc.flag_level3(O, 1, 0.05, tol=0.0, richardson=True)
    for b, r in zip(H.levels[0]["boxes"], rf):
        assert int((r == 2.0).sum()) == (b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1)


def test_conck3_closed_form():
    """conck.f (3-D): integer densities make the sum exact in any order, so
    the oracle's total is sum * hx * hy * hz exactly; poisoned ghost cells
    and random other components catch a wrong index or component."""
    rng = np.random.default_rng(4)
    H = nested_hierarchy3(rng, smooth_euler3, 5, nlev=2, dt1=0.004)
    total = 0
    for k in range(len(H.levels[0]["boxes"])):
        g = H.grid(1, k, 0)
        g[:] = rng.normal(size=g.shape)
        g[..., 0] = 1e30
        dens = rng.integers(1, 1000, size=g[2:-2, 2:-2, 2:-2, 0].shape)
        g[2:-2, 2:-2, 2:-2, 0] = dens
        total += int(dens.sum())
    O = orc.OracleAMR3(orc.make_params3(orc.EULER), H)
    L = H.levels[0]
    assert O.conck(1) == float(total) * L["hx"] * L["hy"] * L["hz"]
    assert orc.mass3([H.grid(1, k, 0) for k in range(len(L["boxes"]))],
                     (L["hx"], L["hy"], L["hz"])) == O.conck(1)
