"""GPU parity of the ghost fill (amrgpu_bound_level) against the oracle's
bound on random hierarchies: FP64 bit-exact; FP32 storage equal to the FP64
result rounded to FP32 (the fill arithmetic runs in FP64)."""
import numpy as np
import pytest

from amrclaw_amd import capi
from oracle import orc
from tests.common import random_hierarchy

pytestmark = pytest.mark.gpu

CASES = [((1, 1, 1, 1), False, False), ((3, 3, 3, 3), False, False),
         ((1, 3, 3, 1), False, False), ((2, 2, 2, 2), True, True),
         ((2, 2, 1, 3), True, False), ((3, 1, 2, 2), False, True)]


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("phase", [(0, 1), (1, 1), (2, 1)])
def test_bound_level_fp64_bit_exact(gpu, seed, case, phase):
    bc, xper, yper = case
    print("bound fp64", seed, case, phase, flush=True)
    rng = np.random.default_rng(2000 + seed)
    H = random_hierarchy(rng, nvar=3, mthbc=bc, xper=xper, yper=yper, fine_phase=phase)
    p = orc.make_params(orc.ACOUSTICS)
    capi.init(capi.make_config(capi.ACOUSTICS, mthbc=bc))
    capi.load_hierarchy(H)
    for l in range(1, H.nlev + 1):
        t = H.levels[l - 1]["time"]
        Ho = H.copy()
        for k in range(len(H.levels[l - 1]["boxes"])):
            orc.bound(p, Ho, l, k, t)
        capi.bound_level(l, t)
        got = capi.read_level(H, l)
        assert np.array_equal(got, Ho.levels[l - 1]["qnew"]), l
    capi.finalize()


@pytest.mark.parametrize("case", CASES[:4])
def test_bound_level_fp32_storage(gpu, case):
    bc, xper, yper = case
    print("bound fp32", case, flush=True)
    rng = np.random.default_rng(77)
    H = random_hierarchy(rng, nvar=3, mthbc=bc, xper=xper, yper=yper, fine_phase=(1, 1))
    # values representable in FP32 so both paths start from the same data
    for L in H.levels:
        L["qnew"][:] = L["qnew"].astype(np.float32)
        L["qold"][:] = L["qold"].astype(np.float32)
    p = orc.make_params(orc.ACOUSTICS)
    capi.init(capi.make_config(capi.ACOUSTICS, mthbc=bc, precision=capi.FP32))
    capi.load_hierarchy(H)
    for l in range(1, H.nlev + 1):
        t = H.levels[l - 1]["time"]
        Ho = H.copy()
        for k in range(len(H.levels[l - 1]["boxes"])):
            orc.bound(p, Ho, l, k, t)
        capi.bound_level(l, t)
        got = capi.read_level(H, l)
        want = Ho.levels[l - 1]["qnew"].astype(np.float32).astype(np.float64)
        assert np.array_equal(got, want), l
    capi.finalize()


def test_bound_level_rejects_unavailable_time(gpu):
    rng = np.random.default_rng(5)
    H = random_hierarchy(rng, nvar=3)
    capi.init(capi.make_config(capi.ACOUSTICS))
    capi.load_hierarchy(H)
    with pytest.raises(capi.AmrgpuError, match="not available"):
        capi.bound_level(H.nlev, 5.0)
    capi.finalize()
