"""CPU checks pinning the oracle's ghost filling (bound/filrecur/intfil/
trimbd/prefilrecur/bc2amr) against the reference's own Fortran and against
closed forms."""
import numpy as np
import pytest

from oracle import orc
from tests.common import Hier, random_hierarchy

needs_ref = pytest.mark.skipif(not orc.ref_available(), reason="reference build absent")

BCS = [(1, 1, 1, 1), (3, 3, 3, 3), (1, 3, 3, 1)]


@needs_ref
@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("phase", [(1, 1), (0, 1), (2, 1)])
def test_bound_matches_reference(seed, bc, phase):
    rng = np.random.default_rng(seed)
    H = random_hierarchy(rng, nvar=3, mthbc=bc, fine_phase=phase)
    p = orc.make_params(orc.ACOUSTICS)
    for l in range(1, H.nlev + 1):
        t = H.levels[l - 1]["time"]
        for k in range(len(H.levels[l - 1]["boxes"])):
            Ho = H.copy()
            orc.bound(p, Ho, l, k, t)
            r = orc.ref_bound(p, H, l, k, t)
            assert np.array_equal(Ho.grid(l, k), r), (l, k)


@needs_ref
@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("per", [(True, True), (True, False), (False, True)])
def test_bound_periodic_matches_reference(seed, per):
    rng = np.random.default_rng(100 + seed)
    bc = tuple(2 if per[0] else 1 for _ in range(2)) + tuple(2 if per[1] else 3 for _ in range(2))
    H = random_hierarchy(rng, nvar=3, mthbc=bc, xper=per[0], yper=per[1], fine_phase=(1, 1))
    p = orc.make_params(orc.ACOUSTICS)
    for l in range(1, H.nlev + 1):
        t = H.levels[l - 1]["time"]
        for k in range(len(H.levels[l - 1]["boxes"])):
            Ho = H.copy()
            orc.bound(p, Ho, l, k, t)
            r = orc.ref_bound(p, H, l, k, t)
            assert np.array_equal(Ho.grid(l, k), r), (l, k)


def _linear_hierarchy(ratio=2, phase=1):
    """Two levels holding the same linear function a + b x + c y + d t."""
    rng = np.random.default_rng(0)
    H = Hier(2, mthbc=(1, 1, 1, 1))
    nx = ny = 16
    H.add_level([(0, 7, 0, 15), (8, 15, 0, 15)], 1 / nx, 1 / ny, 0.1, 1.0, nx, ny, ratio, ratio)
    fine = [(8, 17, 8, 19)]
    H.add_level(fine, 1 / nx / ratio, 1 / ny / ratio, 0.1 / ratio, 1.0 - 0.1 + phase * 0.1 / ratio,
                nx * ratio, ny * ratio, ratio, ratio)
    coef = rng.normal(size=(2, 4))

    def f(x, y, t):
        return np.stack([c[0] + c[1] * x + c[2] * y + c[3] * t for c in coef], -1)

    for l, L in enumerate(H.levels, start=1):
        for k, b in enumerate(L["boxes"]):
            i = np.arange(b[0] - 2, b[1] + 3)
            j = np.arange(b[2] - 2, b[3] + 3)
            X, Y = np.meshgrid((i + 0.5) * L["hx"], (j + 0.5) * L["hy"])
            H.grid(l, k, 0)[:] = f(X, Y, L["time"] if l == 1 else L["time"])
            H.grid(l, k, 1)[:] = f(X, Y, L["time"] - L["possk"])
    return H, f


@pytest.mark.parametrize("phase", [0, 1])
def test_bound_reproduces_linear_data(phase):
    """Coarse-fine interpolation (limited linear slopes, linear time blend)
    is exact for data linear in x, y and t away from limiter clipping."""
    H, f = _linear_hierarchy(phase=phase)
    p = orc.make_params(orc.ACOUSTICS)
    p.meqn = 2
    L = H.levels[1]
    t = L["time"]
    orc.bound(p, H, 2, 0, t)
    b = L["boxes"][0]
    i = np.arange(b[0] - 2, b[1] + 3)
    j = np.arange(b[2] - 2, b[3] + 3)
    X, Y = np.meshgrid((i + 0.5) * L["hx"], (j + 0.5) * L["hy"])
    assert np.allclose(H.grid(2, 0), f(X, Y, t), rtol=1e-12, atol=1e-12)


def test_same_level_copy_is_exact():
    """Ghost cells covered by a neighbour on the same level are copies."""
    rng = np.random.default_rng(3)
    H = Hier(1)
    H.add_level([(0, 9, 0, 9), (10, 19, 0, 9)], 0.05, 0.1, 0.1, 0.5, 20, 10)
    H.levels[0]["qnew"][:] = rng.normal(size=H.levels[0]["qnew"].shape)
    p = orc.make_params(orc.ADVECTION)
    orc.bound(p, H, 1, 0, 0.5)
    left, right = H.grid(1, 0), H.grid(1, 1)
    assert np.array_equal(left[2:12, 12:14], right[2:12, 2:4])
    # extrapolation at the physical boundary: ghost = first interior cell
    assert np.array_equal(left[2:12, 0], left[2:12, 2]) and np.array_equal(left[2:12, 1], left[2:12, 2])


def test_conck_closed_form():
    """conck.f:36-63: with integer densities the sum is exact in any order,
    so the oracle's total must be sum * hx * hy exactly; ghost cells hold
    1e30 and the other components random values, so a total that reads a
    ghost cell or another component fails."""
    from tests.common import nested_hierarchy, smooth_euler
    rng = np.random.default_rng(3)
    H = nested_hierarchy(rng, smooth_euler, 4, nlev=2, dt1=0.004)
    total = 0
    for k in range(len(H.levels[0]["boxes"])):
        g = H.grid(1, k, 0)
        g[:] = rng.normal(size=g.shape)
        g[..., 0] = 1e30
        dens = rng.integers(1, 1000, size=g[2:-2, 2:-2, 0].shape)
        g[2:-2, 2:-2, 0] = dens
        total += int(dens.sum())
    O = orc.OracleAMR(orc.make_params(orc.EULER), H)
    L = H.levels[0]
    assert O.conck(1) == float(total) * L["hx"] * L["hy"]
