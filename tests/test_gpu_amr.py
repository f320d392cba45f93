"""GPU parity of the whole level update (amrgpu_advance_level /
amrgpu_update_level / amrgpu_conck) against the oracle's advanc / update /
conck through Berger-Oliger steps on properly nested three-level
hierarchies: ghost fill, saveqc, sweep with fluxsv / fluxad, qad, upbnd and
averaging.  FP64: solution, old-time arrays, registers and the dt / CFL
sequence bit-exact.  FP32 storage (FP64 registers): within the relative L1
tolerance 1e-5 of the FP64 oracle."""
import numpy as np
import pytest

from amrclaw_amd import capi
from oracle import orc
from tests.common import (acoustic_states, berger_oliger, both_params, nested_hierarchy,
                          rel_l1, smooth_euler)

pytestmark = pytest.mark.gpu

PHYS = {"acoustics": (orc.ACOUSTICS, 3, acoustic_states, 0.004),
        "euler": (orc.EULER, 4, smooth_euler, 0.006)}
BCS = [((1, 1, 1, 1), False, False), ((3, 3, 3, 3), False, False),
       ((2, 2, 2, 2), True, True), ((2, 2, 3, 1), True, False)]


class DeviceAMR:
    """The library's level update over a hierarchy loaded on the device."""

    def __init__(self, H):
        self.H = H
        capi.load_hierarchy(H)

    def advanc(self, level):
        return capi.advance_level(level)

    def update(self, level):
        capi.update_level(level)

    def conck(self, level):
        return capi.conck(level)


def _run(phys, case, seed, precision, tiling=None, steps=2, mode=None, graphs=False):
    solver, nvar, states, dt1 = PHYS[phys]
    bc, xper, yper = case
    rng = np.random.default_rng(300 + seed)
    H = nested_hierarchy(rng, states, nvar, mthbc=bc, xper=xper, yper=yper, dt1=dt1)
    if precision == capi.FP32:
        for L in H.levels:
            L["qnew"][:] = L["qnew"].astype(np.float32)
            L["qold"][:] = L["qold"].astype(np.float32)
    p, cfg = both_params(solver, precision=precision)
    for k in range(4):
        cfg.mthbc[k] = bc[k]
    O = orc.OracleAMR(p, H.copy())
    capi.init(cfg)
    if tiling:
        capi.set_tiling(*tiling)
    if mode is not None:
        capi.set_launch_mode(mode)
    if graphs:
        capi.set_graphs(True)
    D = DeviceAMR(H)
    lo, ld = [], []
    for _ in range(steps):
        berger_oliger(O, 1, H.nlev, 2, lo)
        berger_oliger(D, 1, H.nlev, 2, ld)
    return H, O, lo, ld


@pytest.mark.parametrize("phys", ["acoustics", "euler"])
@pytest.mark.parametrize("case", BCS)
@pytest.mark.parametrize("seed", range(2))
def test_level_update_fp64_bit_exact(gpu, phys, case, seed):
    print("amr fp64", phys, case, seed, flush=True)
    H, O, lo, ld = _run(phys, case, seed, capi.FP64)
    try:
        assert lo == ld, "dtlevnew / cfl sequence"
        for l in range(1, H.nlev + 1):
            got = capi.read_level(H, l, 0)
            assert np.array_equal(got, O.H.levels[l - 1]["qnew"]), ("new", l)
            if l > 1:
                for k in range(len(H.levels[l - 1]["boxes"])):
                    sv, qc = capi.get_registers(l, k, H.nvar, 4096)
                    ref = O.reg(l, k).reshape(2, -1, H.nvar)
                    assert np.array_equal(qc, ref[1]), ("qc1d", l, k)
                    assert np.array_equal(sv, ref[0]), ("svdflx", l, k)
        # the time-t arrays (ghost cells included): every level below the
        # finest keeps them for time interpolation
        for l in range(1, H.nlev):
            assert np.array_equal(capi.read_level(H, l, 1), O.H.levels[l - 1]["qold"]), ("old", l)
        assert capi.conck(1) == O.conck(1)   # the reference's sequential sum, bit for bit
    finally:
        capi.finalize()


@pytest.mark.parametrize("tiling", [(8, 8), (12, 4), (4, 12)])
def test_level_update_fp64_multi_tile(gpu, tiling):
    """Grids split into many tiles: register entries and perimeter fluxes
    cross tile boundaries."""
    print("amr fp64 tiles", tiling, flush=True)
    H, O, lo, ld = _run("euler", BCS[2], 5, capi.FP64, tiling=tiling)
    try:
        assert lo == ld
        for l in range(1, H.nlev + 1):
            assert np.array_equal(capi.read_level(H, l, 0), O.H.levels[l - 1]["qnew"]), l
    finally:
        capi.finalize()


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("mode", [capi.LAUNCH_BATCHED, capi.LAUNCH_PER_PATCH,
                                  capi.LAUNCH_PERSISTENT, capi.LAUNCH_MARCH])
def test_level_update_launch_variants(gpu, mode, graphs):
    """One launch per patch, the persistent and the march sweep, and hipGraph
    replay of the level steps give the oracle's results bit for bit (graphs: every replay after the first of a
    key is checked through the final state after three coarse steps, with
    the program's own synchronising work between the replays)."""
    print("amr variants", mode, graphs, flush=True)
    H, O, lo, ld = _run("euler", BCS[3], 7, capi.FP64, steps=3, mode=mode, graphs=graphs)
    try:
        assert lo == ld
        for l in range(1, H.nlev + 1):
            assert np.array_equal(capi.read_level(H, l, 0), O.H.levels[l - 1]["qnew"]), l
    finally:
        capi.finalize()


@pytest.mark.parametrize("precision", [capi.FP32, capi.MIXED])
@pytest.mark.parametrize("phys", ["acoustics", "euler"])
@pytest.mark.parametrize("case", BCS[:3])
def test_level_update_fp32(gpu, phys, case, precision):
    print("amr fp32", phys, case, precision, flush=True)
    H, O, lo, ld = _run(phys, case, 1, precision)
    try:
        for l in range(1, H.nlev + 1):
            got = capi.read_level(H, l, 0)
            want = O.H.levels[l - 1]["qnew"]
            assert rel_l1(got, want) <= 1e-5, (l, rel_l1(got, want))
        for (la, (dta, ca)), (lb, (dtb, cb)) in zip(lo, ld):
            assert la == lb
            assert abs(ca - cb) <= 1e-5 * max(ca, 1e-30)
    finally:
        capi.finalize()


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_periodic_mass_conserved_on_device(gpu, precision):
    """Refluxing keeps the level-1 mass constant to rounding (periodic
    domain, coarse data first averaged from the finer levels).  The mixed
    path (FP32 waves, FP64 storage, flux form, no qad) conserves to FP64
    rounding as well."""
    print("amr conservation", precision, flush=True)
    solver, nvar, states, dt1 = PHYS["euler"]
    rng = np.random.default_rng(42)
    H = nested_hierarchy(rng, states, nvar, mthbc=(2, 2, 2, 2), xper=True, yper=True, dt1=dt1)
    p, cfg = both_params(solver, precision=precision)
    for k in range(4):
        cfg.mthbc[k] = 2
    capi.init(cfg)
    try:
        D = DeviceAMR(H)
        for l in range(H.nlev - 1, 0, -1):
            D.update(l)
        m0 = D.conck(1)
        for _ in range(3):
            berger_oliger(D, 1, H.nlev, 2)
            m = D.conck(1)
            assert abs(m - m0) <= 1e-13 * abs(m0) + 1e-14, (m, m0)
    finally:
        capi.finalize()
