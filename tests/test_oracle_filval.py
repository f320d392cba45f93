"""CPU checks pinning the oracle's regrid front end against the reference's
own Fortran and against closed forms: filval (filval.f90:3-187 with
intcopy.f, preintcopy.f and bc2amr, gfixup.f's new-grid fill) and the
neighbour lists of makeBndryList (init_bndryList.f:50-107), which the
library's planner builds with a bucket index instead of the reference's
all-pairs scan."""
import numpy as np
import pytest

from amrclaw_amd import capi
from oracle import orc
from tests.common import acoustic_states, nested_boxes, nested_hierarchy

needs_ref = pytest.mark.skipif(not orc.ref_available(), reason="reference build absent")

# (mthbc, xper, yper)
BCS = [((1, 1, 1, 1), False, False), ((3, 3, 3, 3), False, False), ((1, 3, 3, 1), False, False),
       ((2, 2, 2, 2), True, True), ((2, 2, 1, 3), True, False), ((1, 3, 2, 2), False, True)]


def regrid_case(seed, bc, ratio=2, level=3):
    rng = np.random.default_rng(seed)
    mthbc, xp, yp = bc
    H = nested_hierarchy(rng, acoustic_states, 3, nx=24, ny=20, nlev=3, ratio=ratio,
                         mthbc=mthbc, xper=xp, yper=yp, nfine=5)
    boxes = nested_boxes(rng, H, level, 6)
    return rng, H, boxes


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@needs_ref
@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("level,ratio", [(2, 2), (3, 2), (2, 4)])
def test_filval_matches_reference(seed, bc, level, ratio):
    rng, H, boxes = regrid_case(seed, bc, ratio, level)
    assert boxes, "no nested box drawn"
    p = orc.make_params(orc.ACOUSTICS)
    t = H.levels[level - 1]["time"]
    for b in boxes:
        shape = (b[3] - b[2] + 5, b[1] - b[0] + 5, H.nvar)
        init = rng.normal(size=shape)                 # what the heap held
        got = orc.filval(p, H, level, b, t, init.copy())
        want = orc.ref_filval(p, H, level, b, t, init.copy())
        assert np.array_equal(bits(got), bits(want)), b


def test_filval_closed_forms():
    """Old grids covering the new grid and its ghost ring are copied as
    they are; where no old grid reaches, a coarse level linear in x and y
    is interpolated exactly at the fine cell centres (min-mod slopes of a
    linear function are its gradient), a constant coarse level gives the
    constant, and component 1 of ghost cells no old grid covers holds
    NEEDS_TO_BE_SET."""
    rng = np.random.default_rng(5)
    H = nested_hierarchy(rng, acoustic_states, 3, nx=16, ny=16, nlev=2, ratio=2, nfine=3)
    p = orc.make_params(orc.ACOUSTICS)
    C = H.levels[0]
    hxc, hyc = C["hx"], C["hy"]
    for k, b in enumerate(C["boxes"]):
        g = H.grid(1, k)
        jj, ii = np.meshgrid(np.arange(b[2] - 2, b[3] + 3), np.arange(b[0] - 2, b[1] + 3),
                             indexing="ij")
        xc, yc = (ii + 0.5) * hxc, (jj + 0.5) * hyc
        for m in range(3):
            g[..., m] = (m + 1) * 0.25 + 2.0 * (m - 1) * xc - 3.0 * yc
    # new box away from the old fine grids: no copy, pure interpolation
    F = H.levels[1]
    F["boxes"] = np.zeros((0, 4), dtype=np.int32)
    F["shapes"], F["off"] = [], np.zeros(0, dtype=np.int64)
    b = (8, 19, 6, 15)
    val = orc.filval(p, H, 2, b, 0.0, np.zeros((b[3] - b[2] + 5, b[1] - b[0] + 5, 3)))
    hx, hy = F["hx"], F["hy"]
    jj, ii = np.meshgrid(np.arange(b[2], b[3] + 1), np.arange(b[0], b[1] + 1), indexing="ij")
    x, y = (ii + 0.5) * hx, (jj + 0.5) * hy
    for m in range(3):
        exact = (m + 1) * 0.25 + 2.0 * (m - 1) * x - 3.0 * y
        assert np.allclose(val[2:-2, 2:-2, m], exact, rtol=0, atol=1e-13), m
    needs = float(np.float32(10.e33))
    assert np.all(val[:2, :, 0] == needs) and np.all(val[:, :2, 0] == needs)
    # constant coarse data: constant interior
    for k in range(len(C["boxes"])):
        H.grid(1, k)[...] = 1.5
    val = orc.filval(p, H, 2, b, 0.0, np.zeros((b[3] - b[2] + 5, b[1] - b[0] + 5, 3)))
    assert np.all(val[2:-2, 2:-2, :] == 1.5)


def test_filval_copies_old_grids():
    """A new grid equal to an old one takes the old grid's interior as it
    is, and its ghost cells from the old grids that cover them."""
    rng, H, boxes = regrid_case(11, BCS[0], 2, 3)
    p = orc.make_params(orc.ACOUSTICS)
    F = H.levels[2]
    for k, b in enumerate(F["boxes"]):
        val = orc.filval(p, H, 3, tuple(b), 0.0, np.zeros(H.grid(3, k).shape))
        assert np.array_equal(val[2:-2, 2:-2], H.grid(3, k)[2:-2, 2:-2]), k


@needs_ref
@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("bc", BCS[:1] + BCS[3:4])
def test_bndry_lists_match_reference(seed, bc):
    """makeBndryList's lists three ways: the reference's, the oracle's
    all-pairs restatement and the library's bucket index
    (amrgpu_bndry_lists, host only), same grids in the same order."""
    rng, H, _ = regrid_case(seed, bc, 2, 3)
    p = orc.make_params(orc.ACOUSTICS)
    for level in (1, 2, 3):
        want = orc.ref_bndry_lists(p, H, level)
        n = len(H.levels[level - 1]["boxes"])
        assert [orc.bndry_list(p, H, level, k) for k in range(n)] == want
        L = H.levels[level - 1]
        assert capi.bndry_lists(L["boxes"], L["iregsz"], L["jregsz"]) == want


def test_bndry_lists_many_grids():
    """At thousands of grids the bucket index gives the all-pairs lists."""
    from amrclaw_amd import workload
    H, _ = workload.batched_euler_hierarchy(1024)
    L = H.levels[1]
    boxes = np.asarray(L["boxes"])
    got = capi.bndry_lists(boxes, L["iregsz"], L["jregsz"])
    lo = boxes[:, [0, 2]] - 2
    hi = boxes[:, [1, 3]] + 2
    for k in range(len(boxes)):
        meet = np.nonzero((np.maximum(boxes[:, 0], lo[k, 0]) <= np.minimum(boxes[:, 1], hi[k, 0])) &
                          (np.maximum(boxes[:, 2], lo[k, 1]) <= np.minimum(boxes[:, 3], hi[k, 1])))[0]
        meet = [int(s) for s in meet if s != k][::-1]
        assert got[k] == meet, k
