"""CPU checks pinning the oracle's capacity and aux paths (SURVEY's
capacity / aux row): the sweep with edge velocities in aux and a capacity
function (step2.f90:91-95 dtdx1d, flux2.f:127-237, stepgrid.f:166-172)
against the reference's own Fortran, the variable-coefficient advection
solver against closed forms, and capacity-weighted conservation."""
import numpy as np
import pytest

from oracle import orc

needs_ref = pytest.mark.skipif(not orc.ref_available(), reason="reference build absent")


def vc_aux(rng, mx, my, capa=True, uniform=None):
    """aux(1) x-velocity at each cell's left edge, aux(2) y-velocity at its
    bottom edge, aux(3) the capacity (cell area over dx dy)."""
    shape = (my + 4, mx + 4)
    if uniform is not None:
        a1, a2 = np.full(shape, uniform[0]), np.full(shape, uniform[1])
    else:
        a1, a2 = rng.normal(size=shape), rng.normal(size=shape)
    a3 = 0.5 + 1.5 * rng.random(shape) if capa else np.ones(shape)
    return np.ascontiguousarray(np.stack([a1, a2, a3], -1))


@needs_ref
@pytest.mark.parametrize("order,trans", [(2, 2), (2, 1), (2, 0), (1, 2), (1, 0)])
@pytest.mark.parametrize("lim", [0, 1, 3, 4])
@pytest.mark.parametrize("mcapa", [0, 3])
def test_vcadv_capacity_step_matches_reference(order, trans, lim, mcapa):
    rng = np.random.default_rng(50 * lim + 7 * order + trans + mcapa)
    p = orc.make_params(orc.VCADV, order=order, transverse=trans, limiter=lim, maux=3,
                        mcapa=mcapa)
    mx, my = 11, 7
    q = rng.normal(size=(my + 4, mx + 4, 1))
    aux = vc_aux(rng, mx, my)
    dx, dy, dt = 0.1, 0.13, 0.01
    f1, c1 = orc.step2(p, q, dx, dy, dt, aux)
    f2, c2 = orc.ref_step2(p, q, dx, dy, dt, aux)
    assert c1 == c2
    for a, b in zip(f1, f2):
        assert np.array_equal(a, b)
    r1 = orc.stepgrid(p, q, dx, dy, dt, aux)
    r2 = orc.ref_stepgrid(p, q, dx, dy, dt, aux)
    assert np.array_equal(r1[0], r2[0])
    assert r1[2] == r2[2] and r1[3] == r2[3]


@needs_ref
@pytest.mark.parametrize("solver", [orc.ADVECTION, orc.ACOUSTICS, orc.EULER])
def test_capacity_with_builtin_solvers_matches_reference(solver):
    """A capacity function with the solvers that take no aux: only dtdx1d
    and the update change."""
    from tests.common import random_patch
    rng = np.random.default_rng(solver)
    p = orc.make_params(solver, maux=1, mcapa=1)
    q = random_patch(rng, solver, 9, 6)
    aux = np.ascontiguousarray(0.5 + 1.5 * rng.random((10, 13, 1)))
    r1 = orc.stepgrid(p, q, 0.1, 0.13, 0.01, aux)
    r2 = orc.ref_stepgrid(p, q, 0.1, 0.13, 0.01, aux)
    assert np.array_equal(r1[0], r2[0]) and r1[2] == r2[2]


@pytest.mark.parametrize("order,trans,lim", [(2, 2, 3), (2, 1, 4), (1, 0, 0)])
def test_vcadv_uniform_velocity_is_constant_advection(order, trans, lim):
    """Uniform edge velocities (u, v) and unit capacity: the
    variable-coefficient solver is the constant-coefficient one, bit for
    bit (same fluctuations, same transverse splits)."""
    rng = np.random.default_rng(3)
    mx, my = 10, 8
    u, v = 0.7, -0.4
    q = rng.normal(size=(my + 4, mx + 4, 1))
    pv = orc.make_params(orc.VCADV, order=order, transverse=trans, limiter=lim, maux=3)
    pa = orc.make_params(orc.ADVECTION, order=order, transverse=trans, limiter=lim, ubar=u,
                         vbar=v)
    aux = vc_aux(rng, mx, my, capa=False, uniform=(u, v))
    rv = orc.stepgrid(pv, q, 0.1, 0.1, 0.05, aux)
    ra = orc.stepgrid(pa, q, 0.1, 0.1, 0.05)
    assert np.array_equal(rv[0], ra[0]) and rv[2] == ra[2]


def test_vcadv_velocity_of_the_crossed_edge():
    """One nonzero interface jump: the normal fluctuation moves at the
    velocity stored with the cell to the right of the interface (its left
    edge), and its transverse parts at the velocities of the edges above
    and below the cell it entered; first order, so only those terms act."""
    p = orc.make_params(orc.VCADV, order=1, transverse=1, limiter=0, maux=3)
    mx, my = 6, 5
    q = np.zeros((my + 4, mx + 4, 1))
    q[:, 5:, 0] = 1.0                          # jump across x-edge between cells 2 and 3
    aux = np.zeros((my + 4, mx + 4, 3))
    aux[..., 2] = 1.0
    aux[:, 5, 0] = 0.8                         # x-velocity at the left edge of cell i = 4 (1-based 3)
    aux[:, 4, 1] = 0.3                         # y-velocities of the cells left of the edge
    f, _ = orc.step2(p, q, 1.0, 1.0, 0.5, aux)
    fm, fp, gm, gp = f
    # A+dq = 0.8 * 1 goes right: fp at that edge is -apdq
    assert np.allclose(fp[2:-2, 5, 0], -0.8)
    assert np.allclose(fm[2:-2, 5, 0], 0.0)


def test_capacity_conservation():
    """Edge velocities differenced from a stream function that vanishes on
    the boundary of the unit square (discretely divergence-free, no flow
    through the boundary edges) and a random capacity: the capacity-weighted
    total sum(kappa q) is conserved to rounding (the transverse and
    second-order terms move amounts between neighbouring cells)."""
    rng = np.random.default_rng(9)
    mx, my = 12, 10
    dx, dy = 1.0 / mx, 1.0 / my
    p = orc.make_params(orc.VCADV, maux=3, mcapa=3, limiter=3)
    q = rng.random(size=(my + 4, mx + 4, 1))
    psi = lambda x, y: np.sin(np.pi * x) * np.sin(np.pi * y)
    jj, ii = np.meshgrid(np.arange(-1, my + 3), np.arange(-1, mx + 3), indexing="ij")
    xc, yc = (ii - 1) * dx, (jj - 1) * dy             # lower-left corner of cell (i, j)
    aux = np.zeros((my + 4, mx + 4, 3))
    aux[..., 0] = (psi(xc, yc + dy) - psi(xc, yc)) / dy
    aux[..., 1] = -(psi(xc + dx, yc) - psi(xc, yc)) / dx
    aux[..., 2] = 0.5 + rng.random((my + 4, mx + 4))
    qn, _, cfl, _ = orc.stepgrid(p, q, dx, dy, 0.01, aux)
    assert 0 < cfl < 1
    k = aux[2:-2, 2:-2, 2]
    before = float((k * q[2:-2, 2:-2, 0]).sum())
    after = float((k * qn[2:-2, 2:-2, 0]).sum())
    assert abs(after - before) <= 1e-13 * abs(before), (before, after)


CAPA_BCS = [((1, 1, 1, 1), False, False), ((2, 2, 2, 2), True, True), ((1, 1, 2, 2), False, True)]


@needs_ref
@pytest.mark.parametrize("case", CAPA_BCS)
@pytest.mark.parametrize("seed", range(3))
def test_capacity_level_update_matches_reference(case, seed):
    """The whole level update with aux and capacity (advanc with qad's
    xleft / yleft face values and the coarse aux ring of saveqc/cstore,
    stepgrid, fluxsv, fluxad, upbnd's capacity scaling, update's
    capacity-weighted averaging) against the reference through two
    Berger-Oliger steps, bit for bit: solutions at both times, registers
    and the dt / CFL sequence."""
    from tests.common import berger_oliger, nested_hierarchy, positive_states, vc_annulus_aux
    bc, xper, yper = case
    rng = np.random.default_rng(500 + seed)
    H = nested_hierarchy(rng, positive_states, 1, mthbc=bc, xper=xper, yper=yper, dt1=0.008,
                         naux=3)
    vc_annulus_aux(H)
    p = orc.make_params(orc.VCADV, limiter=3, maux=3, mcapa=3,
                        auxtype=(orc.AUX_XLEFT, orc.AUX_YLEFT, orc.AUX_CAPACITY))
    O = orc.OracleAMR(p, H.copy())
    R = orc.RefAMR(p, H)
    lo, lr = [], []
    for step in range(2):
        berger_oliger(O, 1, H.nlev, 2, lo)
        berger_oliger(R, 1, H.nlev, 2, lr)
        assert lo == lr, "dtlevnew / cfl sequence"
        for l in range(1, H.nlev + 1):
            for k in range(len(H.levels[l - 1]["boxes"])):
                assert np.array_equal(O.H.grid(l, k, 0), R.grid(l, k, 0)), ("new", step, l, k)
                if l < H.nlev:
                    assert np.array_equal(O.H.grid(l, k, 1), R.grid(l, k, 1)), ("old", step, l, k)
                if l > 1:
                    assert np.array_equal(O.reg(l, k), R.reg(l, k)), ("reg", step, l, k)


def test_capacity_composite_mass_is_conserved():
    """Periodic domain, divergence-free edge velocities, a capacity whose
    coarse values are the means of the fine ones: refluxing with upbnd's
    capacity scaling and the capacity-weighted averaging keep the level-1
    total sum(kappa q) hx hy (conck.f with capacity) constant to rounding."""
    from tests.common import berger_oliger, nested_hierarchy, positive_states, vc_annulus_aux
    rng = np.random.default_rng(77)
    H = nested_hierarchy(rng, positive_states, 1, mthbc=(2, 2, 2, 2), xper=True, yper=True,
                         dt1=0.008, naux=3)
    vc_annulus_aux(H)
    p = orc.make_params(orc.VCADV, limiter=3, maux=3, mcapa=3,
                        auxtype=(orc.AUX_XLEFT, orc.AUX_YLEFT, orc.AUX_CAPACITY))
    O = orc.OracleAMR(p, H)
    for l in range(H.nlev - 1, 0, -1):
        O.update(l)
    m0 = O.conck(1)
    for _ in range(3):
        berger_oliger(O, 1, H.nlev, 2)
        m = O.conck(1)
        assert abs(m - m0) <= 1e-13 * abs(m0), (m, m0)


def test_wall_needs_the_velocity_component():
    """bc2amr's solid wall negates q(2) on the x-sides and q(3) on the
    y-sides: the oracle refuses it for the one-equation solver instead of
    writing past the patch (the reference's loops would)."""
    from tests.common import berger_oliger, nested_hierarchy, positive_states, vc_annulus_aux
    rng = np.random.default_rng(3)
    H = nested_hierarchy(rng, positive_states, 1, mthbc=(3, 3, 3, 3), dt1=0.008, naux=3)
    vc_annulus_aux(H)
    p = orc.make_params(orc.VCADV, limiter=3, maux=3, mcapa=3,
                        auxtype=(orc.AUX_XLEFT, orc.AUX_YLEFT, orc.AUX_CAPACITY))
    O = orc.OracleAMR(p, H)
    with pytest.raises(RuntimeError):
        berger_oliger(O, 1, H.nlev, 2)
