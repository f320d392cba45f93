"""GPU parity of the 3-D level update around the step (include/amrgpu3.h,
hierarchy section) against the oracle, itself pinned bit for bit to the
reference's src/3d Fortran in test_oracle_amr3.py.

FP64: the ghost fill (bound.f and its recursion, periodic wrap, bc3amr),
and whole Berger-Oliger steps (saveqc, qad, stepgrid, fluxsv, fluxad,
upbnd, update) bit for bit: new and old solutions, registers, dt / CFL
sequence.  FP32 storage: relative L1 <= 1e-5 of the FP64 oracle.
"""
import numpy as np
import pytest

from amrclaw_amd import capi3
from oracle import orc
from tests.common import rel_l1
from tests.common3 import (advection3_states, berger_oliger3, nested_hierarchy3,
                           random_hierarchy3, smooth_euler3)

pytestmark = pytest.mark.gpu

BCS = [((1,) * 6, (False,) * 3), ((3,) * 6, (False,) * 3), ((1, 3, 3, 1, 1, 3), (False,) * 3),
       ((2,) * 6, (True,) * 3)]


def _init(solver, precision=capi3.FP64):
    capi3.init3(capi3.make_config3(solver, precision=precision))


@pytest.mark.parametrize("case", BCS)
@pytest.mark.parametrize("seed", range(2))
@pytest.mark.parametrize("phase", [1, 0])
def test_bound3_fp64_bit_exact(gpu, case, seed, phase):
    bc, per = case
    print("bound3", case, seed, phase, flush=True)
    rng = np.random.default_rng(50 + seed)
    H = random_hierarchy3(rng, mthbc=bc, per=per, fine_phase=phase)
    p = orc.make_params3(orc.EULER)
    _init(orc.EULER)
    try:
        capi3.load_hierarchy3(H)
        for l in range(1, H.nlev + 1):
            t = H.levels[l - 1]["time"]
            capi3.bound_level3(l, t)
            got = capi3.read_level3(H, l, 0)
            Ho = H.copy()
            for k in range(len(H.levels[l - 1]["boxes"])):
                orc.bound3(p, Ho, l, k, t)
            assert np.array_equal(got, Ho.levels[l - 1]["qnew"]), l
    finally:
        capi3.finalize3()


def test_bound3_ratio_mix_fp64_bit_exact(gpu):
    print("bound3 ratios", flush=True)
    rng = np.random.default_rng(91)
    H = random_hierarchy3(rng, ratios=((2, 4, 2), (4, 2, 2)), n=(8, 6, 6))
    p = orc.make_params3(orc.EULER)
    _init(orc.EULER)
    try:
        capi3.load_hierarchy3(H)
        for l in range(1, H.nlev + 1):
            t = H.levels[l - 1]["time"]
            capi3.bound_level3(l, t)
            Ho = H.copy()
            for k in range(len(H.levels[l - 1]["boxes"])):
                orc.bound3(p, Ho, l, k, t)
            assert np.array_equal(capi3.read_level3(H, l, 0), Ho.levels[l - 1]["qnew"]), l
    finally:
        capi3.finalize3()


def _hier(solver, case, seed, **kw):
    bc, per = case
    rng = np.random.default_rng(400 + seed)
    if solver == orc.EULER:
        return nested_hierarchy3(rng, smooth_euler3, 5, mthbc=bc, per=per, dt1=0.004, **kw)
    return nested_hierarchy3(rng, advection3_states, 1, mthbc=bc, per=per, dt1=0.01, **kw)


@pytest.mark.parametrize("solver", [orc.EULER, orc.ADVECTION])
@pytest.mark.parametrize("case", BCS)
@pytest.mark.parametrize("seed", range(2))
def test_level_update3_fp64_bit_exact(gpu, solver, case, seed):
    if solver == orc.ADVECTION and 3 in case[0]:
        pytest.skip("wall boundaries negate a velocity component advection does not have")
    print("amr3 fp64", solver, case, seed, flush=True)
    H = _hier(solver, case, seed)
    p = orc.make_params3(solver)
    O = orc.OracleAMR3(p, H.copy())
    _init(solver)
    try:
        D = capi3.DeviceAMR3(H)
        lo, ld = [], []
        for step in range(2):
            berger_oliger3(O, O.H, 1, lo)
            berger_oliger3(D, H, 1, ld)
            assert lo == ld, ("dtlevnew / cfl sequence", step)
        for l in range(1, H.nlev + 1):
            assert np.array_equal(capi3.read_level3(H, l, 0), O.H.levels[l - 1]["qnew"]), ("new", l)
            if l < H.nlev:
                assert np.array_equal(capi3.read_level3(H, l, 1), O.H.levels[l - 1]["qold"]), ("old", l)
            if l > 1:
                for k in range(len(H.levels[l - 1]["boxes"])):
                    sv, qc = capi3.get_registers3(l, k, H.nvar)
                    ref = O.reg(l, k).reshape(2, -1, H.nvar)
                    assert np.array_equal(qc, ref[1]), ("qc1d", l, k)
                    assert np.array_equal(sv, ref[0]), ("svdflx", l, k)
        assert capi3.conck_level3(1) == O.conck(1)   # sequential sum, bit for bit
    finally:
        capi3.finalize3()


def test_level_update3_ratio_mix_fp64_bit_exact(gpu):
    print("amr3 ratios", flush=True)
    rng = np.random.default_rng(77)
    H = nested_hierarchy3(rng, smooth_euler3, 5, n=(8, 8, 6), ratios=((2, 2, 2), (4, 4, 4)),
                          dt1=0.004, hi=8)
    p = orc.make_params3(orc.EULER)
    O = orc.OracleAMR3(p, H.copy())
    _init(orc.EULER)
    try:
        D = capi3.DeviceAMR3(H)
        lo, ld = [], []
        berger_oliger3(O, O.H, 1, lo)
        berger_oliger3(D, H, 1, ld)
        assert lo == ld
        for l in range(1, H.nlev + 1):
            assert np.array_equal(capi3.read_level3(H, l, 0), O.H.levels[l - 1]["qnew"]), l
    finally:
        capi3.finalize3()


@pytest.mark.parametrize("case", BCS[:1] + BCS[3:])
def test_level_update3_fp32(gpu, case):
    print("amr3 fp32", case, flush=True)
    H = _hier(orc.EULER, case, 0)
    for L in H.levels:
        L["qnew"][:] = L["qnew"].astype(np.float32)
        L["qold"][:] = L["qold"].astype(np.float32)
    p = orc.make_params3(orc.EULER)
    O = orc.OracleAMR3(p, H.copy())
    _init(orc.EULER, capi3.FP32)
    try:
        D = capi3.DeviceAMR3(H)
        lo, ld = [], []
        for _ in range(2):
            berger_oliger3(O, O.H, 1, lo)
            berger_oliger3(D, H, 1, ld)
        for l in range(1, H.nlev + 1):
            got = capi3.read_level3(H, l, 0)
            want = O.H.levels[l - 1]["qnew"]
            assert rel_l1(got, want) <= 1e-5, (l, rel_l1(got, want))
        for (la, (_, ca)), (lb, (_, cb)) in zip(lo, ld):
            assert la == lb and abs(ca - cb) <= 1e-5 * max(ca, 1e-30)
    finally:
        capi3.finalize3()


@pytest.mark.parametrize("solver", [orc.EULER, orc.ADVECTION])
def test_periodic_mass_conserved3_on_device(gpu, solver):
    """Fully periodic: refluxing and averaging on the device keep the
    level-1 mass constant to rounding."""
    print("amr3 conservation", solver, flush=True)
    H = _hier(solver, BCS[3], 5)
    _init(solver)
    try:
        D = capi3.DeviceAMR3(H)
        for l in range(H.nlev - 1, 0, -1):
            D.update(l)
        m0 = D.conck(1)
        for _ in range(3):
            berger_oliger3(D, H, 1)
            m = D.conck(1)
            assert abs(m - m0) <= 1e-13 * max(abs(m0), 1.0), (m, m0)
    finally:
        capi3.finalize3()


@pytest.mark.parametrize("case", [BCS[0], BCS[1], BCS[3]])
@pytest.mark.parametrize("seed", range(2))
def test_flag_level3_fp64_bit_exact(gpu, case, seed):
    """amrgpu3_flag_level after a coarse step: spest's flags and errest /
    errf1's Richardson flags of every level below the finest bit for bit
    with the oracle (flagged fine to coarse, as regrid.f does, since the
    reference's errf1 overwrites component 1 of store2)."""
    bc, per = case
    print("flag3 fp64", case, seed, flush=True)
    rng = np.random.default_rng(600 + seed)
    H = nested_hierarchy3(rng, smooth_euler3, 5, n=(12, 12, 8), mthbc=bc, per=per, dt1=0.004)
    p = orc.make_params3(orc.EULER)
    O = orc.OracleAMR3(p, H.copy())
    _init(orc.EULER)
    try:
        D = capi3.DeviceAMR3(H)
        berger_oliger3(O, O.H, 1)
        berger_oliger3(D, H, 1)
        for level in range(H.nlev - 1, 0, -1):
            tolsp, tol = 0.05, 1e-4
            f, rf, n = orc.flag_level3(O, level, tolsp, tol=tol, richardson=True)
            nd, rd = D.flag(level, tolsp, richardson=True, tol=tol)
            assert nd == n and n > 0, (level, nd, n)
            assert rd == sum(int((r == 2.0).sum()) for r in rf) and rd > 0, level
            for k, shape in enumerate(H.levels[level - 1]["shapes"]):
                gf, gr = capi3.get_level_flags3(level, k, shape[:3])
                assert np.array_equal(gf, f[k]), ("spest", level, k)
                assert np.array_equal(gr, rf[k]), ("errf1", level, k)
        # the finest level: spest only (no old time for errest)
        f, _, n = orc.flag_level3(O, H.nlev, 0.05)
        nd, rd = D.flag(H.nlev, 0.05)
        assert nd == n and rd == 0
        for k, shape in enumerate(H.levels[H.nlev - 1]["shapes"]):
            assert np.array_equal(capi3.get_level_flags3(H.nlev, k, shape[:3])[0], f[k]), k
        # flagging leaves the solutions as they were (bound fills ghost cells only)
        for l in range(1, H.nlev):
            assert np.array_equal(capi3.read_level3(H, l, 1), O.H.levels[l - 1]["qold"]), l
    finally:
        capi3.finalize3()


@pytest.mark.parametrize("precision", [capi3.FP64, capi3.FP32])
def test_flag_level3_uniform_state(gpu, precision):
    """A uniform state has no Richardson error in either precision: nothing
    flagged at tol 1e-12, every interior cell at tol 0."""
    print("flag3 uniform", precision, flush=True)
    rng = np.random.default_rng(5)
    H = nested_hierarchy3(rng, smooth_euler3, 5, n=(12, 12, 8), dt1=0.004)
    for L in H.levels:
        if len(L["boxes"]):
            for which in ("qnew", "qold"):
                L[which].reshape(-1, 5)[:] = (1.1, 0.2, -0.1, 0.05, 2.9)
        L["time"] = L["possk"]
    _init(orc.EULER, precision)
    try:
        D = capi3.DeviceAMR3(H)
        nd, rd = D.flag(1, 0.05, richardson=True, tol=1e-12)
        assert nd == 0 and rd == 0
        nd, rd = D.flag(1, 0.05, richardson=True, tol=0.0)
        boxes = H.levels[0]["boxes"]
        assert rd == sum((b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1) for b in boxes)
        for k, shape in enumerate(H.levels[0]["shapes"]):
            _, r = capi3.get_level_flags3(1, k, shape[:3])
            assert (r[2:-2, 2:-2, 2:-2] == 2.0).all() and r.sum() == 2.0 * r[2:-2, 2:-2, 2:-2].size
    finally:
        capi3.finalize3()
