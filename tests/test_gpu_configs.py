"""GPU parity of the level update on the three 2-D run configurations
(advection, periodic, ratios 2, 2; acoustics on [-1, 1]^2, ratios 2, 4;
Euler quadrants over four level-1 grids, ratios 2, 4) against the oracle,
itself pinned to the reference's Fortran on the same hierarchies in
test_oracle_configs.py.  FP64 bit for bit (solution, old-time arrays,
registers, dt / CFL sequence); FP32 storage and the mixed path within the
relative L1 tolerance 1e-5; the periodic advection run conserves mass to
FP64 rounding on the mixed path."""
import numpy as np
import pytest

from amrclaw_amd import capi
from oracle import orc
from tests.common import rel_l1
from tests.configs import CONFIGS, berger_oliger, config_hierarchy
from tests.test_gpu_amr import DeviceAMR

pytestmark = pytest.mark.gpu


def _setup(name, precision):
    c = CONFIGS[name]
    H = config_hierarchy(name)
    if precision == capi.FP32:
        for L in H.levels:
            L["qnew"][:] = L["qnew"].astype(np.float32)
            L["qold"][:] = L["qold"].astype(np.float32)
    p = orc.make_params(c["solver"])
    cfg = capi.make_config(c["solver"], mthbc=c["mthbc"], domain=c["domain"], precision=precision)
    capi.init(cfg)
    return H, p


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_config_fp64_bit_exact(gpu, name):
    print("config fp64", name, flush=True)
    H, p = _setup(name, capi.FP64)
    try:
        O = orc.OracleAMR(p, H.copy())
        D = DeviceAMR(H)
        lo, ld = [], []
        for step in range(2):
            berger_oliger(O, O.H, 1, lo)
            berger_oliger(D, H, 1, ld)
            assert lo == ld, ("dtlevnew / cfl sequence", step)
        for l in range(1, H.nlev + 1):
            assert np.array_equal(capi.read_level(H, l, 0), O.H.levels[l - 1]["qnew"]), ("new", l)
            if l < H.nlev:
                assert np.array_equal(capi.read_level(H, l, 1), O.H.levels[l - 1]["qold"]), ("old", l)
            if l > 1:
                for k in range(len(H.levels[l - 1]["boxes"])):
                    sv, qc = capi.get_registers(l, k, H.nvar, 8192)
                    ref = O.reg(l, k).reshape(2, -1, H.nvar)
                    assert np.array_equal(qc, ref[1]), ("qc1d", l, k)
                    assert np.array_equal(sv, ref[0]), ("svdflx", l, k)
        assert capi.conck(1) == O.conck(1)   # sequential sum, bit for bit
    finally:
        capi.finalize()


@pytest.mark.parametrize("precision", [capi.FP32, capi.MIXED])
@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_config_reduced_precision(gpu, name, precision):
    print("config reduced", name, precision, flush=True)
    H, p = _setup(name, precision)
    try:
        O = orc.OracleAMR(p, H.copy())
        D = DeviceAMR(H)
        lo, ld = [], []
        for _ in range(2):
            berger_oliger(O, O.H, 1, lo)
            berger_oliger(D, H, 1, ld)
        for l in range(1, H.nlev + 1):
            got = capi.read_level(H, l, 0)
            want = O.H.levels[l - 1]["qnew"]
            assert rel_l1(got, want) <= 1e-5, (l, rel_l1(got, want))
        for (la, (_, ca)), (lb, (_, cb)) in zip(lo, ld):
            assert la == lb and abs(ca - cb) <= 1e-5 * max(ca, 1e-30)
    finally:
        capi.finalize()


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_config_periodic_mass_conserved(gpu, precision):
    print("config conservation", precision, flush=True)
    H, _ = _setup("C1", precision)
    try:
        D = DeviceAMR(H)
        for l in range(H.nlev - 1, 0, -1):
            D.update(l)
        m0 = D.conck(1)
        for _ in range(3):
            berger_oliger(D, H, 1)
            m = D.conck(1)
            assert abs(m - m0) <= 1e-13 * abs(m0), (m, m0)
    finally:
        capi.finalize()


@pytest.mark.parametrize("precision", [capi.FP64, capi.FP32])
def test_config_c5_3d(gpu, precision):
    """C5 (3-D Euler radial bump, 64^3 level 1, two levels at ratio 2): one
    coarse Berger-Oliger step on the device against the oracle; FP64 bit for
    bit (solution, registers, dt / CFL), FP32 within 1e-5 relative L1."""
    from amrclaw_amd import capi3
    from tests.common3 import berger_oliger3
    from tests.configs3 import config_hierarchy3
    print("config C5", precision, flush=True)
    H = config_hierarchy3()
    if precision == capi.FP32:
        for L in H.levels:
            L["qnew"][:] = L["qnew"].astype(np.float32)
            L["qold"][:] = L["qold"].astype(np.float32)
    p = orc.make_params3(orc.EULER)
    O = orc.OracleAMR3(p, H.copy())
    capi3.init3(capi3.make_config3(orc.EULER, precision=precision))
    try:
        D = capi3.DeviceAMR3(H)
        lo, ld = [], []
        berger_oliger3(D, H, 1, ld)
        berger_oliger3(O, O.H, 1, lo)
        for l in range(1, H.nlev + 1):
            got = capi3.read_level3(H, l, 0)
            want = O.H.levels[l - 1]["qnew"]
            if precision == capi.FP64:
                assert np.array_equal(got, want), l
            else:
                assert rel_l1(got, want) <= 1e-5, (l, rel_l1(got, want))
        if precision == capi.FP64:
            assert lo == ld
            for l in range(2, H.nlev + 1):
                for k in range(len(H.levels[l - 1]["boxes"])):
                    sv, qc = capi3.get_registers3(l, k, H.nvar)
                    ref = O.reg(l, k).reshape(2, -1, H.nvar)
                    assert np.array_equal(sv, ref[0]) and np.array_equal(qc, ref[1]), (l, k)
    finally:
        capi3.finalize3()
