"""GPU parity of the cell flagging (amrgpu_flag_level) against the oracle's
flagregions2 + flag2refine2 (pinned against the reference's Fortran) on
random hierarchies: ghost fill at the level's time, regions, undivided
differences.  Flags are compared exactly (FP64 values -1 / 0 / 2)."""
import numpy as np
import pytest

from amrclaw_amd import capi
from oracle import orc
from tests.common import random_hierarchy
from tests.test_oracle_flag import random_regions

pytestmark = pytest.mark.gpu


def _oracle_flags(H, Ho, l, which, mbuff, tol, regs):
    L = H.levels[l - 1]
    out = []
    for k, b in enumerate(L["boxes"]):
        mx, my = int(b[1] - b[0] + 1), int(b[3] - b[2] + 1)
        xlo = H.domain[0] + int(b[0]) * L["hx"]
        ylo = H.domain[2] + int(b[2]) * L["hy"]
        out.append(orc.flag_grid(Ho.grid(l, k, which), mx, my, mbuff, xlo, ylo, L["hx"], L["hy"],
                                 L["time"], l, tol, regs))
    return out


def _shapes(H, l):
    return [(int(b[1] - b[0] + 1), int(b[3] - b[2] + 1)) for b in H.levels[l - 1]["boxes"]]


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (2, 2, 2, 2), (3, 1, 2, 2)])
def test_flag_level_after_ghost_fill(gpu, seed, bc):
    print("flag new", seed, bc, flush=True)
    rng = np.random.default_rng(4000 + seed)
    xper, yper = bc[0] == 2, bc[2] == 2
    H = random_hierarchy(rng, nvar=3, mthbc=bc, xper=xper, yper=yper, fine_phase=(1, 1))
    p = orc.make_params(orc.ACOUSTICS)
    capi.init(capi.make_config(capi.ACOUSTICS, mthbc=bc))
    try:
        capi.load_hierarchy(H)
        for l in range(1, H.nlev + 1):
            t = H.levels[l - 1]["time"]
            Ho = H.copy()
            for k in range(len(H.levels[l - 1]["boxes"])):
                orc.bound(p, Ho, l, k, t)
            regs = random_regions(rng, int(rng.integers(0, 4)), t)
            tol, mbuff = 2.5, int(rng.choice([2, 3, 4]))
            want = _oracle_flags(H, Ho, l, 0, mbuff, tol, regs)
            got, n = capi.flag_level(l, tol, t - 7.0, regs, mbuff, _shapes(H, l))
            for k, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), (l, k)
            assert n == sum(int((w == orc.DOFLAG).sum()) for w in want)
    finally:
        capi.finalize()


@pytest.mark.parametrize("precision", [capi.FP64, capi.FP32])
def test_flag_level_initial_time_uses_old(gpu, precision):
    """start_time + possk == time: the old values are flagged, no ghost fill
    (flagger.f:98-117)."""
    print("flag old", precision, flush=True)
    rng = np.random.default_rng(17)
    H = random_hierarchy(rng, nvar=3)
    for L in H.levels:
        L["qnew"][:] = L["qnew"].astype(np.float32)
        L["qold"][:] = L["qold"].astype(np.float32)
    capi.init(capi.make_config(capi.ACOUSTICS, precision=precision))
    try:
        capi.load_hierarchy(H)
        for l in range(1, H.nlev + 1):
            L = H.levels[l - 1]
            want = _oracle_flags(H, H, l, 1, 3, 1.5, [])
            got, n = capi.flag_level(l, 1.5, L["time"] - L["possk"], [], 3, _shapes(H, l))
            for k, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), (l, k)
    finally:
        capi.finalize()


def test_flag_level_rejects_inverted_region(gpu):
    rng = np.random.default_rng(3)
    H = random_hierarchy(rng, nvar=3)
    capi.init(capi.make_config(capi.ACOUSTICS))
    try:
        capi.load_hierarchy(H)
        t = H.levels[0]["time"]
        bad = [(3, 1, t - 1, t + 1, -1.0, 2.0, -1.0, 2.0)]
        with pytest.raises(capi.AmrgpuError, match="minlevel > maxlevel"):
            capi.flag_level(1, 1.0, t - 5.0, bad, 2)
    finally:
        capi.finalize()


@pytest.mark.parametrize("ibuff", [1, 3])
@pytest.mark.parametrize("case", [((1, 1, 1, 1), False, False), ((3, 3, 3, 3), False, False),
                                  ((2, 2, 2, 2), True, True)])
def test_flag_level_richardson_fp64(gpu, case, ibuff):
    """flagger with regions, gradient and Richardson estimation after one
    Berger-Oliger step, device against the oracle (pinned against the
    reference's flagger), every level below mxnest; exact flags."""
    from tests.common import berger_oliger, both_params, nested_hierarchy, smooth_euler
    from tests.test_gpu_amr import DeviceAMR
    print("flag richardson", case, ibuff, flush=True)
    bc, xper, yper = case
    rng = np.random.default_rng(5100 + ibuff)
    H = nested_hierarchy(rng, smooth_euler, 4, mthbc=bc, xper=xper, yper=yper, dt1=0.006)
    p, cfg = both_params(orc.EULER)
    for k in range(4):
        cfg.mthbc[k] = bc[k]
    O = orc.OracleAMR(p, H.copy())
    capi.init(cfg)
    try:
        D = DeviceAMR(H)
        berger_oliger(O, 1, H.nlev, 2)
        berger_oliger(D, 1, H.nlev, 2)
        mb = max(2, ibuff + 1)
        for l in range(1, H.nlev):
            regs = random_regions(rng, 2, O.h.time[l])
            want = O.flag_level(l, 0.4, -1.0, regs, ibuff=ibuff, gradient=True, richardson=True,
                                tol=1e-5)
            got, n = capi.flag_level(l, 0.4, -1.0, regs, mb, _shapes(H, l), gradient=True,
                                     richardson=True, tol=1e-5)
            for k, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), (l, k)
            assert n == sum(int((w == orc.DOFLAG).sum()) for w in want)
            # Richardson alone, at the initial-time branch (old values)
            t0 = O.h.time[l] - H.levels[l - 1]["possk"]
            want = O.flag_level(l, 0.4, t0, (), ibuff=ibuff, gradient=False, richardson=True,
                                tol=2e-6)
            got, _ = capi.flag_level(l, 0.4, t0, (), mb, _shapes(H, l), gradient=False,
                                     richardson=True, tol=2e-6)
            for k, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), ("old", l, k)
    finally:
        capi.finalize()
