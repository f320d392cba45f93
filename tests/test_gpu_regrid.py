"""GPU parity across a regrid: the device keeps the coarse levels' state
while the finer levels get new layouts mid-run, as the reference's regrid
does between coarse steps (src/2d/regrid.f:72-118, then tick.f:240 calls
advanc on the new layout).  The library must re-plan its ghost-fill
programs, registers and tile tables, and drop every captured level step
(hipGraph) that names the old tables.  Checked bit for bit in FP64 against
the oracle, which starts from the same state with the new layout, with
graph replay off and on, and with the geometry re-installed in between
(amrgpu_set_geometry, which changes nothing when the geometry is the
same)."""
import numpy as np
import pytest

from amrclaw_amd import capi
from amrclaw_amd.hierarchy import Hier
from oracle import orc
from tests.common import (berger_oliger, both_params, disjoint_boxes, nested_boxes,
                          nested_hierarchy, smooth_euler)

pytestmark = pytest.mark.gpu


class DeviceAMR:
    def advanc(self, level):
        return capi.advance_level(level)

    def update(self, level):
        capi.update_level(level)


def _nested_boxes(rng, coarse_boxes, isz, jsz, ratio, n):
    """Disjoint fine boxes properly nested in the coarse boxes (one coarse
    cell of margin away from every coarse edge inside the domain)."""
    within = []
    for (a, b, c, d) in coarse_boxes:
        within.append((0 if a == 0 else (a + 1) * ratio,
                       isz * ratio - 1 if b == isz - 1 else b * ratio - 1,
                       0 if c == 0 else (c + 1) * ratio,
                       jsz * ratio - 1 if d == jsz - 1 else d * ratio - 1))
    return disjoint_boxes(rng, isz * ratio, jsz * ratio, n, lo=4, hi=max(8, isz * ratio // 3),
                          align=ratio, within=within, tries=4000)


def _regridded(H, rng, first_new):
    """A copy of H whose levels below first_new keep their arrays and whose
    levels first_new.. get new boxes and new (smooth) data, old = new, at
    the time of the level below."""
    H2 = Hier(H.nvar, domain=H.domain, mthbc=H.mthbc, xper=H.xper, yper=H.yper)
    for l, L in enumerate(H.levels, start=1):
        if l < first_new:
            H2.add_level(L["boxes"], L["hx"], L["hy"], L["possk"], L["time"], L["iregsz"],
                         L["jregsz"], L["ratx"], L["raty"])
            H2.levels[-1]["qnew"][:] = L["qnew"]
            H2.levels[-1]["qold"][:] = L["qold"]
        else:
            P = H2.levels[-1]
            boxes = _nested_boxes(rng, [tuple(b) for b in P["boxes"]], P["iregsz"],
                                  P["jregsz"], P["ratx"], 4)
            assert boxes, "no nested boxes drawn"
            H2.add_level(boxes, L["hx"], L["hy"], L["possk"], P["time"], L["iregsz"],
                         L["jregsz"], L["ratx"], L["raty"])
            for k in range(len(boxes)):
                g = H2.grid(l, k, 0)
                g[:] = smooth_euler(rng, g.shape[:2])
                H2.grid(l, k, 1)[:] = g
    return H2


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("first_new", [2, 3])
@pytest.mark.parametrize("regeom", [False, True])
def test_regrid_cycle_fp64_bit_exact(gpu, graphs, first_new, regeom):
    print("regrid cycle", graphs, first_new, regeom, flush=True)
    rng = np.random.default_rng(900 + 2 * first_new + regeom)
    H = nested_hierarchy(rng, smooth_euler, 4, mthbc=(1, 1, 1, 1), dt1=0.006)
    p, cfg = both_params(orc.EULER, precision=capi.FP64)
    O = orc.OracleAMR(p, H.copy())
    capi.init(cfg)
    try:
        capi.set_graphs(graphs)
        capi.load_hierarchy(H)
        D = DeviceAMR()
        lo, ld = [], []
        for _ in range(2):          # two coarse steps: graphs are captured and replayed
            berger_oliger(O, 1, H.nlev, 2, lo)
            berger_oliger(D, 1, H.nlev, 2, ld)
        assert lo == ld
        for l in range(1, H.nlev + 1):
            assert np.array_equal(capi.read_level(H, l, 0), O.H.levels[l - 1]["qnew"]), l
        # regrid: new layouts from level first_new up, coarse levels untouched
        H2 = _regridded(O.H, rng, first_new)
        O2 = orc.OracleAMR(p, H2.copy())
        if regeom:
            lv = H2.levels
            capi.set_geometry([L["iregsz"] for L in lv], [L["jregsz"] for L in lv],
                              [L["ratx"] for L in lv], [L["raty"] for L in lv],
                              [L["hx"] for L in lv], [L["hy"] for L in lv])
        for l in range(first_new, H2.nlev + 1):
            L = H2.levels[l - 1]
            capi.set_level(l, L["boxes"], L["hx"], L["hy"])
            capi.put_level(l, [H2.grid(l, k, 0) for k in range(len(L["boxes"]))], 0)
            capi.put_level(l, [H2.grid(l, k, 1) for k in range(len(L["boxes"]))], 1)
            capi.set_time(l, L["time"], L["possk"])
        lo, ld = [], []
        for _ in range(2):
            berger_oliger(O2, 1, H2.nlev, 2, lo)
            berger_oliger(D, 1, H2.nlev, 2, ld)
        assert lo == ld, "dtlevnew / cfl after the regrid"
        for l in range(1, H2.nlev + 1):
            assert np.array_equal(capi.read_level(H2, l, 0), O2.H.levels[l - 1]["qnew"]), ("new", l)
        for l in range(1, H2.nlev):
            assert np.array_equal(capi.read_level(H2, l, 1), O2.H.levels[l - 1]["qold"]), ("old", l)
        # registers of the regridded levels; the levels below first_new keep
        # their arenas across the regrid as the reference keeps their
        # node(ffluxptr) blocks (only slots along the physical boundary,
        # which no coarse cell reads, then differ from the fresh oracle)
        for l in range(max(2, first_new), H2.nlev + 1):
            for k in range(len(H2.levels[l - 1]["boxes"])):
                sv, qc = capi.get_registers(l, k, H2.nvar, 4096)
                ref = O2.reg(l, k).reshape(2, -1, H2.nvar)
                assert np.array_equal(sv, ref[0]) and np.array_equal(qc, ref[1]), ("regs", l, k)
    finally:
        capi.finalize()


def test_refused_tiling_leaves_level_intact(gpu):
    """Tile caps whose tiles need more LDS than a CU holds are refused before
    any level is re-planned: the level keeps advancing exactly as before."""
    print("refused tiling", flush=True)
    rng = np.random.default_rng(77)
    # one 64x64 level-1 grid: caps 61x61 tile it 32x32, which needs 177 KB of
    # LDS in FP64 (nvar 4)
    H = nested_hierarchy(rng, smooth_euler, 4, nx=64, ny=64, mthbc=(1, 1, 1, 1), dt1=0.004,
                         nlev=2, level1_split=(1, 1))
    p, cfg = both_params(orc.EULER, precision=capi.FP64)
    O = orc.OracleAMR(p, H.copy())
    capi.init(cfg)
    try:
        capi.load_hierarchy(H)
        with pytest.raises(RuntimeError):
            capi.set_tiling(61, 61)
        D = DeviceAMR()
        lo, ld = [], []
        berger_oliger(O, 1, H.nlev, 2, lo)
        berger_oliger(D, 1, H.nlev, 2, ld)
        assert lo == ld
        for l in range(1, H.nlev + 1):
            assert np.array_equal(capi.read_level(H, l, 0), O.H.levels[l - 1]["qnew"]), l
    finally:
        capi.finalize()


def _filval_level(p, Hc, level, boxes, t):
    """The oracle's filval for new boxes of `level` of Hc (level-1 already
    in its new layout, `level` still the old one): the new grids' arrays."""
    out = []
    for b in boxes:
        val = np.zeros((b[3] - b[2] + 5, b[1] - b[0] + 5, Hc.nvar))
        out.append(orc.filval(p, Hc, level, b, t, val))
    return out


@pytest.mark.parametrize("bc", [((1, 1, 1, 1), False, False), ((2, 2, 2, 2), True, True),
                                ((2, 2, 3, 3), True, False)])
@pytest.mark.parametrize("first_new", [2, 3])
def test_device_regrid_matches_filval(gpu, bc, first_new):
    """amrgpu_regrid_level replaces a level's grids and fills them on the
    device as gfixup/filval do (old grids copied over the new grid and its
    ghost ring, min-mod interpolation from the coarse level elsewhere):
    levels first_new.. regridded coarse to fine after a coarse step, the
    new grids' interiors and every ghost cell an old grid covered bit-exact
    against the oracle's filval (pinned to the reference's filval.f90), and
    two more coarse steps on the new layout bit-exact against the oracle."""
    print("device regrid", bc, first_new, flush=True)
    mthbc, xp, yp = bc
    rng = np.random.default_rng(700 + first_new + 10 * xp + 20 * yp)
    H = nested_hierarchy(rng, smooth_euler, 4, mthbc=mthbc, xper=xp, yper=yp, dt1=0.006)
    p, _ = both_params(orc.EULER, precision=capi.FP64)
    cfg = capi.make_config(capi.EULER, precision=capi.FP64, mthbc=mthbc)
    O = orc.OracleAMR(p, H.copy())
    capi.init(cfg)
    try:
        capi.load_hierarchy(H)
        D = DeviceAMR()
        lo, ld = [], []
        berger_oliger(O, 1, H.nlev, 2, lo)
        berger_oliger(D, 1, H.nlev, 2, ld)
        assert lo == ld
        Hc = O.H.copy()
        needs = float(np.float32(10.e33))
        for l in range(first_new, H.nlev + 1):
            t = Hc.levels[l - 2]["time"]
            boxes = []
            for _ in range(20):                # small periodic levels leave little room
                boxes = boxes or nested_boxes(rng, Hc, l, 4)
            assert boxes, "no nested boxes drawn"
            vals = _filval_level(p, Hc, l, boxes, t)
            capi.regrid_level(l, boxes, t)
            # Hc's level l becomes the new layout with the filval values
            L = Hc.levels[l - 1]
            H2 = Hier(Hc.nvar, domain=Hc.domain, mthbc=Hc.mthbc, xper=Hc.xper, yper=Hc.yper)
            for m, M in enumerate(Hc.levels, start=1):
                if m == l:
                    H2.add_level(boxes, M["hx"], M["hy"], M["possk"], t, M["iregsz"],
                                 M["jregsz"], M["ratx"], M["raty"])
                    for k, v in enumerate(vals):
                        H2.grid(l, k, 0)[:] = v
                        H2.grid(l, k, 1)[:] = v
                else:
                    H2.add_level(M["boxes"], M["hx"], M["hy"], M["possk"], M["time"],
                                 M["iregsz"], M["jregsz"], M["ratx"], M["raty"])
                    H2.levels[-1]["qnew"][:] = M["qnew"]
                    H2.levels[-1]["qold"][:] = M["qold"]
            got = capi.read_level(H2, l, 0)
            for k, v in enumerate(vals):
                shape = v.shape
                g = got[H2.levels[l - 1]["off"][k]:H2.levels[l - 1]["off"][k] + v.size].reshape(shape)
                assert np.array_equal(g[2:-2, 2:-2], v[2:-2, 2:-2]), ("interior", l, k)
                covered = v[..., 0] != needs
                assert np.array_equal(g[covered], v[covered]), ("ghosts", l, k)
            Hc = H2
        O2 = orc.OracleAMR(p, Hc.copy())
        lo, ld = [], []
        for _ in range(2):
            berger_oliger(O2, 1, Hc.nlev, 2, lo)
            berger_oliger(D, 1, Hc.nlev, 2, ld)
        assert lo == ld, "dtlevnew / cfl after the device regrid"
        for l in range(1, Hc.nlev + 1):
            assert np.array_equal(capi.read_level(Hc, l, 0), O2.H.levels[l - 1]["qnew"]), l
    finally:
        capi.finalize()


def test_device_regrid_refusals(gpu):
    """Level 1, a level above mxnest and a coarse level without grids are
    refused before anything changes."""
    print("device regrid refusals", flush=True)
    rng = np.random.default_rng(3)
    H = nested_hierarchy(rng, smooth_euler, 4, mthbc=(1, 1, 1, 1), dt1=0.006, nlev=2)
    p, cfg = both_params(orc.EULER, precision=capi.FP64)
    capi.init(cfg)
    try:
        capi.load_hierarchy(H)
        with pytest.raises(capi.AmrgpuError):
            capi.regrid_level(1, [(0, 3, 0, 3)], 0.0)
        with pytest.raises(capi.AmrgpuError):
            capi.regrid_level(5, [(0, 3, 0, 3)], 0.0)
        before = capi.read_level(H, 2, 0)
        assert np.array_equal(before, H.levels[1]["qnew"])
    finally:
        capi.finalize()
