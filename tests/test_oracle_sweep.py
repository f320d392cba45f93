"""CPU checks that pin the oracle's sweep (step2/flux2/limiter/stepgrid) and
Riemann solvers.

* against the reference's own Fortran, compiled unmodified (oracle/_ref),
  when that build is present;
* against golden vectors written by that build (tests/golden/step2_ref.npz);
* against closed forms: Riemann-solver identities, invariance of uniform
  states, reduction to the 1-D scheme, exact translation, conservation.
"""
import os

import numpy as np
import pytest

from oracle import orc
from tests.common import random_patch

GOLD = os.path.join(os.path.dirname(__file__), "golden", "step2_ref.npz")


# --------------------------------------------------------------- reference --
@pytest.mark.skipif(not orc.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("solver", [orc.ADVECTION, orc.ACOUSTICS, orc.EULER])
@pytest.mark.parametrize("order,trans", [(2, 2), (2, 1), (2, 0), (1, 2), (1, 0)])
@pytest.mark.parametrize("lim", [0, 1, 2, 3, 4, 5])
def test_step2_matches_reference_build(solver, order, trans, lim):
    rng = np.random.default_rng(1000 * solver + 10 * lim + order + 3 * trans)
    p = orc.make_params(solver, order=order, transverse=trans, limiter=lim)
    q = random_patch(rng, solver, 9, 6)
    f1, c1 = orc.step2(p, q, 0.1, 0.13, 0.02)
    f2, c2 = orc.ref_step2(p, q, 0.1, 0.13, 0.02)
    assert c1 == c2
    for a, b in zip(f1, f2):
        assert np.array_equal(a, b)
    r1 = orc.stepgrid(p, q, 0.1, 0.13, 0.02)
    r2 = orc.ref_stepgrid(p, q, 0.1, 0.13, 0.02)
    assert np.array_equal(r1[0], r2[0])
    assert r1[2] == r2[2] and r1[3] == r2[3]


# ------------------------------------------------------------------ golden --
def _golden_cases():
    z = np.load(GOLD)
    names = sorted({k.split("/")[0] for k in z.files})
    return z, names


@pytest.mark.parametrize("name", _golden_cases()[1])
def test_step2_matches_golden(name):
    z, _ = _golden_cases()
    solver, order, trans, lim, mx, my, dx, dy, dt = z[name + "/meta"]
    p = orc.make_params(int(solver), order=int(order), transverse=int(trans),
                        limiter=int(lim))
    q = z[name + "/q"]
    f, cfl = orc.step2(p, q, dx, dy, dt)
    for k, a in zip(("fm", "fp", "gm", "gp"), f):
        assert np.array_equal(a, z[name + "/" + k]), k
    assert cfl == z[name + "/cfl"][0]
    qn, _, cflg, dtn = orc.stepgrid(p, q, dx, dy, dt)
    assert np.array_equal(qn, z[name + "/qnew"])
    assert dtn == z[name + "/dtnew"][0]


# ------------------------------------------------------------ closed forms --
def _slice(states):
    """rows = cells 1-mbc .. mx+mbc of one slice"""
    return np.ascontiguousarray(states)


@pytest.mark.parametrize("ixy", [1, 2])
def test_roe_property_euler(ixy):
    """sum_p s_p W_p = f(q_r) - f(q_l) and sum_p W_p = q_r - q_l (Roe)."""
    rng = np.random.default_rng(7 + ixy)
    for efix in (False, True):
        p = orc.make_params(orc.EULER, efix=efix)
        q = random_patch(rng, orc.EULER, 20, 1)[1]      # one row, 24 cells
        wave, s, amdq, apdq = orc.rpn2(p, ixy, q, q)
        for i in range(1, q.shape[0]):
            dq = q[i] - q[i - 1]
            df = orc.flux(p, ixy, q[i]) - orc.flux(p, ixy, q[i - 1])
            assert np.allclose(wave[i].sum(0), dq, rtol=1e-12, atol=1e-12)
            assert np.allclose((s[i][:, None] * wave[i]).sum(0), df, rtol=1e-11, atol=1e-12)
            assert np.allclose(amdq[i] + apdq[i], df, rtol=1e-11, atol=1e-12)


def test_entropy_fix_only_changes_transonic_interfaces():
    """Without a sonic point the Harten-Hyman split equals the Roe split."""
    rng = np.random.default_rng(3)
    p0 = orc.make_params(orc.EULER, efix=False)
    p1 = orc.make_params(orc.EULER, efix=True)
    q = random_patch(rng, orc.EULER, 30, 1)[1]
    w0, s0, am0, ap0 = orc.rpn2(p0, 1, q, q)
    w1, s1, am1, ap1 = orc.rpn2(p1, 1, q, q)
    gamma = 1.4
    for i in range(1, q.shape[0]):
        l, r = q[i - 1], q[i]
        cl = np.sqrt(gamma * (gamma - 1) * (l[3] - 0.5 * (l[1] ** 2 + l[2] ** 2) / l[0]) / l[0])
        cr = np.sqrt(gamma * (gamma - 1) * (r[3] - 0.5 * (r[1] ** 2 + r[2] ** 2) / r[0]) / r[0])
        ul, ur = l[1] / l[0], r[1] / r[0]
        sonic = (ul - cl) * (s0[i][0]) < 0 or (ul - cl < 0 < ur - cr) or (ul + cl < 0 < ur + cr)
        if not sonic and (ul - cl) * (ur - cr) > 0 and (ul + cl) * (ur + cr) > 0:
            assert np.allclose(am0[i], am1[i], rtol=1e-10, atol=1e-12)
        assert np.allclose(am1[i] + ap1[i], am0[i] + ap0[i], rtol=1e-12, atol=1e-12)


def test_euler_single_shock_is_one_wave():
    """Two states joined by a 3-shock (Rankine-Hugoniot) give one wave."""
    gamma = 1.4
    rho_l, u_l, p_l = 1.0, 0.0, 1.0
    M = 2.0                                     # shock Mach number (right-going)
    c_l = np.sqrt(gamma * p_l / rho_l)
    # right-going shock into state l from the left: post-shock state on left
    rho_r = rho_l
    p_post = p_l * (2 * gamma * M * M - (gamma - 1)) / (gamma + 1)
    rho_post = rho_l * (gamma + 1) * M * M / ((gamma - 1) * M * M + 2)
    s_shock = M * c_l
    u_post = s_shock * (1 - rho_l / rho_post)
    def cons(r, u, p):
        return np.array([r, r * u, 0.0, p / (gamma - 1) + 0.5 * r * u * u])
    ql = cons(rho_post, u_post, p_post)
    qr = cons(rho_r, u_l, p_l)
    p = orc.make_params(orc.EULER, efix=False)
    q = np.stack([ql, ql, qr, qr, qr])
    wave, s, amdq, apdq = orc.rpn2(p, 1, q, q)
    w = wave[2]                                  # interface between ql and qr
    norms = np.linalg.norm(w, axis=1)
    assert norms[3] > 0.1 and norms[:3].max() < 1e-9 * norms[3]
    assert abs(s[2][3] - s_shock) < 1e-9
    assert np.allclose(apdq[2], s_shock * (qr - ql), rtol=1e-9)


@pytest.mark.parametrize("ixy", [1, 2])
def test_acoustics_eigen_identities(ixy):
    rng = np.random.default_rng(11)
    p = orc.make_params(orc.ACOUSTICS, rho=1.0, bulk=4.0)
    q = rng.normal(size=(12, 3))
    wave, s, amdq, apdq = orc.rpn2(p, ixy, q, q)
    mu = 1 if ixy == 1 else 2
    for i in range(1, 12):
        dq = q[i] - q[i - 1]
        # A dq with A = [[0,K,0],[1/rho,0,0],[0,0,0]] in the normal direction
        adq = np.zeros(3)
        adq[0] = 4.0 * dq[mu]
        adq[mu] = dq[0] / 1.0
        w_sum = wave[i].sum(0)
        assert np.allclose(w_sum[[0, mu]], dq[[0, mu]], atol=1e-13)
        assert np.allclose(amdq[i] + apdq[i], adq, atol=1e-12)
        assert np.allclose(s[i], [-2.0, 2.0])


def test_transverse_splitting_sums_to_B_asdq():
    """rpt2: B^- A*dq + B^+ A*dq = B A*dq for acoustics (linear)."""
    rng = np.random.default_rng(5)
    p = orc.make_params(orc.ACOUSTICS, rho=1.0, bulk=4.0)
    q = rng.normal(size=(10, 3))
    asdq = rng.normal(size=(10, 3))
    bm, bp = orc.rpt2(p, 1, 1, q, q, asdq)
    b = np.zeros_like(asdq)        # transverse (y) Jacobian applied to asdq
    b[:, 0] = 4.0 * asdq[:, 2]
    b[:, 2] = asdq[:, 0]
    assert np.allclose((bm + bp)[1:], b[1:], atol=1e-12)


@pytest.mark.parametrize("solver", [orc.ADVECTION, orc.ACOUSTICS, orc.EULER])
def test_uniform_state_is_invariant(solver):
    rng = np.random.default_rng(2)
    p = orc.make_params(solver)
    q0 = random_patch(rng, solver, 1, 1)[0, 0]
    q = np.broadcast_to(q0, (10, 12, q0.size)).copy()
    qn, f, cfl, dtn = orc.stepgrid(p, q, 0.1, 0.1, 0.01)
    assert np.allclose(qn, q, rtol=0, atol=1e-14 * np.abs(q).max())


def test_advection_exact_translation_at_cfl_one():
    """Upwind (Godunov) with u dt/dx = 1 shifts the data one cell (exact up
    to the rounding of q_i - (q_i - q_{i-1}))."""
    rng = np.random.default_rng(4)
    p = orc.make_params(orc.ADVECTION, order=1, transverse=0, ubar=1.0, vbar=0.0)
    q = rng.normal(size=(9, 12, 1))
    qn, f, cfl, dtn = orc.stepgrid(p, q, 0.1, 0.1, 0.1)
    assert np.allclose(qn[2:-2, 2:-2], q[2:-2, 1:-3], rtol=0, atol=1e-15 * np.abs(q).max())
    assert cfl == 1.0


@pytest.mark.parametrize("lim", [1, 2, 3, 4])
def test_slab_reduces_to_1d_scheme(lim):
    """q independent of y: transverse terms vanish and every row performs the
    1-D high-resolution update with the limiter phi(theta)."""
    rng = np.random.default_rng(6)
    u = 0.7
    p = orc.make_params(orc.ADVECTION, ubar=u, vbar=0.4, limiter=lim)
    row = rng.normal(size=14)
    q = np.broadcast_to(row[None, :, None], (8, 14, 1)).copy()
    dx, dt = 0.1, 0.05
    qn, f, cfl, dtn = orc.stepgrid(p, q, dx, dx, dt)
    nu = u * dt / dx

    def phi(t):
        return {1: max(0, min(1, t)), 2: max(0, min(1, 2 * t), min(2, t)),
                3: (t + abs(t)) / (1 + abs(t)),
                4: max(0, min((1 + t) / 2, 2, 2 * t))}[lim]

    dq = np.diff(row)                         # wave at interface i+1/2
    lim_w = np.zeros_like(dq)
    for k in range(1, len(dq)):
        lim_w[k] = dq[k] * phi(dq[k - 1] / dq[k]) if dq[k] != 0 else 0.0
    F = u * row[:-1] + 0.5 * u * (1 - nu) * lim_w  # flux at interface k+1/2
    expect = row.copy()
    for i in range(2, 12):
        expect[i] = row[i] - nu * (F[i] - F[i - 1]) / u
    assert np.allclose(qn[2:-2, 2:-2, 0], expect[None, 2:12], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("solver", [orc.ADVECTION, orc.ACOUSTICS, orc.EULER])
def test_periodic_conservation(solver):
    """Periodic ghost cells: the boundary fluxes of the patch cancel in
    flux form, so the interior sum is conserved to rounding (conck.f)."""
    rng = np.random.default_rng(8)
    p = orc.make_params(solver)
    inner = random_patch(rng, solver, 10, 10, mbc=0)
    q = np.pad(inner, ((2, 2), (2, 2), (0, 0)), mode="wrap")
    qn, f, cfl, dtn = orc.stepgrid(p, q, 0.1, 0.1, 0.01)
    d = (qn[2:-2, 2:-2] - q[2:-2, 2:-2]).sum((0, 1))
    assert np.all(np.abs(d) <= 1e-13 * np.abs(q).sum((0, 1)) + 1e-15)
