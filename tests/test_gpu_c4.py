"""GPU parity on the full batched 2-D Euler workload: 4096 level-2 patches
with mx, my even U[8, 128] (seed 20261202, 18,855,660 interior cells) over
a level-1 cover of 1600^2 cells at ratio 2, quadrant state with a seeded
1e-3 perturbation (amrclaw_amd.workload.batched_euler_hierarchy).  The step
checked is one coarse Berger-Oliger step as tick.f sequences it
(src/2d/tick.f:240,295; advanc.f:10-142; update.f:16): advanc(1) from t = 0
(saveqc fills the registers' coarse ring, fluxsv the coarse fluxes),
advanc(2) twice (the second from half way through the coarse step, the
step the benchmark times: its ghost fill blends the two coarse times), then
update(1) (upbnd refluxing and averaging).  The fine level must catch up
before update: the registers then hold the whole coarse-fine flux
difference, which the fluctuation form (FP64, the reference's) and the
flux form (mixed) compute as the same quantity.

FP64, batched, one launch per patch and the march sweep: the whole of both levels, the dt /
CFL results and the registers bit for bit against the oracle.  Mixed (FP32
waves on FP64 storage, the benchmark's headline path) and FP32 storage:
relative L1 <= 1e-5 against the oracle on both levels, level-2 mass within
1e-10 of FP64's."""
import numpy as np
import pytest

from amrclaw_amd import autotune, capi, workload
from oracle import orc
from tests.common import both_params, rel_l1

pytestmark = pytest.mark.gpu

NPATCH = 4096


@pytest.fixture(scope="module")
def c4(gpu):
    print("c4: building the 4096-patch hierarchy and its oracle step", flush=True)
    H, ncell = workload.batched_euler_hierarchy(NPATCH)
    assert ncell == 18855660
    dt2 = H.levels[1]["possk"]
    H.levels[0].update(time=0.0, possk=2.0 * dt2)
    H.levels[1].update(time=0.0, possk=dt2)
    p, _ = both_params(orc.EULER, precision=capi.FP64)
    O = orc.OracleAMR(p, H.copy())
    log = [O.advanc(1), O.advanc(2), O.advanc(2)]
    O.update(1)
    print("c4: oracle done", log, flush=True)
    return H, O, log


def _device_run(H, precision, launch=None):
    _, cfg = both_params(orc.EULER, precision=precision)
    capi.init(cfg)
    key = {capi.FP64: "fp64", capi.FP32: "fp32", capi.MIXED: "mixed"}[precision]
    tuned = autotune.load(capi.EULER, key) or autotune.load(capi.EULER, "fp32")
    if tuned:
        capi.set_tiling(*tuned)
    wgs = autotune.load_occupancy(capi.EULER, key)
    if wgs and precision == capi.MIXED:
        capi.set_occupancy(wgs)
    if launch is not None:
        capi.set_launch_mode(launch)
    capi.load_hierarchy(H)
    log = [capi.advance_level(1), capi.advance_level(2), capi.advance_level(2)]
    capi.update_level(1)
    return log


def _mass(H, q, l):
    L = H.levels[l - 1]
    tot = 0.0
    for o, s in zip(L["off"], L["shapes"]):
        tot += q[o:o + int(np.prod(s))].reshape(s)[2:-2, 2:-2, 0].sum()
    return tot * L["hx"] * L["hy"]


@pytest.mark.parametrize("launch", [None, capi.LAUNCH_PER_PATCH, capi.LAUNCH_MARCH])
def test_c4_fp64_bit_exact(c4, launch):
    H, O, olog = c4
    print("c4 fp64", launch, flush=True)
    try:
        dlog = _device_run(H, capi.FP64, launch)
        assert dlog == olog, (dlog, olog)
        for l in (1, 2):
            assert np.array_equal(capi.read_level(H, l, 0), O.H.levels[l - 1]["qnew"]), l
        if launch is None:
            for k in range(0, NPATCH, 97):
                sv, qc = capi.get_registers(2, k, 4, 4096)
                ref = O.reg(2, k).reshape(2, -1, 4)
                assert np.array_equal(sv, ref[0]) and np.array_equal(qc, ref[1]), k
    finally:
        capi.finalize()


@pytest.mark.parametrize("launch", [None, capi.LAUNCH_MARCH])
@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP32])
def test_c4_fp32_within_tolerance(c4, precision, launch):
    H, O, olog = c4
    print("c4 fp32-arithmetic", precision, launch, flush=True)
    try:
        dlog = _device_run(H, precision, launch)
        for l in (1, 2):
            got = capi.read_level(H, l, 0)
            want = O.H.levels[l - 1]["qnew"]
            err = rel_l1(got, want)
            print("  level", l, "rel L1", err, flush=True)
            assert err <= 1e-5, (l, err)
            if l == 2:
                m, m64 = _mass(H, got, 2), _mass(H, want, 2)
                assert abs(m - m64) <= 1e-10 * abs(m64), (m, m64)
        for (dta, ca), (dtb, cb) in zip(dlog, olog):
            assert abs(ca - cb) <= 1e-5 * cb and abs(dta - dtb) <= 1e-5 * dtb
    finally:
        capi.finalize()
