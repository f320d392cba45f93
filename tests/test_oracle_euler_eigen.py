"""Closed-form pins of the oracle's Euler Riemann solvers (2-D rpn2/rpt2,
3-D rpn3/rpt3/rptt3), which the reference links from clawpack/riemann and
therefore does not hold (the reference build forwards its solver calls to
these restatements, so only mathematics can pin them).

Everything the checks compare against is computed here independently of
the oracle's formulas:

* the Roe average (velocity and enthalpy weighted by sqrt(rho)) from the
  conserved states;
* the Roe matrix as the exact flux Jacobian at a state with that velocity
  and enthalpy (the Euler Jacobian depends on u and H only), differentiated
  by complex step from a plain flux function;
* its spectral projectors P_k = prod_{j!=k} (A - l_j)/(l_k - l_j) for the
  three distinct eigenvalues u-a, u, u+a, so that A-dq = sum min(l_k,0) P_k dq
  needs no choice of eigenvector basis.

The Harten-Hyman entropy fix (LeVeque 2002 sec. 15.3.5) is pinned by what
defines it: off-sonic it is the Roe split; on a transonic rarefaction the
wave is split into a part moving at the left characteristic speed and a
part at the right one whose weights add to one; the fluctuation is
continuous where a characteristic speed crosses zero; and the whole split
is symmetric under reflection x -> -x (which maps the 1-wave fix onto the
4-wave fix).  Call sites these solvers serve: src/2d/flux2.f:125,202-233,
src/2d/qad.f:129, src/3d/flux3.f:220-224,302-321,411-422.
"""
import numpy as np
import pytest

from oracle import orc
from tests.common import euler_states

GAMMA = 1.4


# ------------------------------------------------------------ independent math
def euler_flux(q, k, gamma=GAMMA):
    """Flux of the Euler equations along velocity component k (0-based) for
    q = (rho, rho u_1 .. rho u_d, E); works on complex arrays."""
    nd = q.shape[0] - 2
    rho = q[0]
    vel = q[1:1 + nd] / rho
    p = (gamma - 1.0) * (q[-1] - 0.5 * rho * (vel * vel).sum())
    f = q * vel[k]
    f[1 + k] = f[1 + k] + p
    f[-1] = f[-1] + p * vel[k]
    return f


def primitive(q, gamma=GAMMA):
    nd = q.shape[0] - 2
    rho = q[0]
    vel = q[1:1 + nd] / rho
    p = (gamma - 1.0) * (q[-1] - 0.5 * rho * (vel * vel).sum())
    return rho, vel, p


def roe_state(L, R, gamma=GAMMA):
    """Roe-averaged velocity vector, enthalpy and sound speed."""
    rl, vl, pl = primitive(L, gamma)
    rr, vr, pr = primitive(R, gamma)
    wl, wr = np.sqrt(rl), np.sqrt(rr)
    hl, hr = (L[-1] + pl) / rl, (R[-1] + pr) / rr
    vel = (wl * vl + wr * vr) / (wl + wr)
    h = (wl * hl + wr * hr) / (wl + wr)
    a = np.sqrt((gamma - 1.0) * (h - 0.5 * (vel * vel).sum()))
    return vel, h, a


def roe_matrix(vel, h, k, gamma=GAMMA):
    """Jacobian of the flux along k at (rho = 1, velocity vel, enthalpy h)."""
    nd = vel.size
    e = (h + 0.5 * (gamma - 1.0) * (vel * vel).sum()) / gamma
    q0 = np.concatenate([[1.0], vel, [e]]).astype(complex)
    n = nd + 2
    J = np.zeros((n, n))
    hstep = 1e-30
    for j in range(n):
        qc = q0.copy()
        qc[j] += 1j * hstep
        J[:, j] = euler_flux(qc, k, gamma).imag / hstep
    return J


def projectors(A, lams):
    """Spectral projectors of a diagonalizable A with distinct eigenvalues lams."""
    n = A.shape[0]
    I = np.eye(n)
    Ps = []
    for i, li in enumerate(lams):
        P = I.copy()
        for j, lj in enumerate(lams):
            if j != i:
                P = P @ (A - lj * I) / (li - lj)
        Ps.append(P)
    return Ps


def split(A, lams, x):
    """(A- x, A+ x) from the spectral decomposition."""
    Ps = projectors(A, lams)
    am = sum(min(l, 0.0) * (P @ x) for l, P in zip(lams, Ps))
    ap = sum(max(l, 0.0) * (P @ x) for l, P in zip(lams, Ps))
    return am, ap


def close(x, y, scale, tol=1e-10):
    return np.abs(np.asarray(x) - np.asarray(y)).max() <= tol * scale


def pairs_2d(L, R):
    """Rows L0, R0, L1, R1, ...: interface row 2k+1 joins (L_k, R_k)."""
    q = np.empty((2 * len(L), L.shape[1]))
    q[0::2], q[1::2] = L, R
    return q


def prim_states(rho, vel, p, gamma=GAMMA):
    vel = np.atleast_2d(vel)
    rho = np.asarray(rho, dtype=float)
    p = np.asarray(p, dtype=float)
    e = p / (gamma - 1.0) + 0.5 * rho * (vel * vel).sum(-1)
    return np.column_stack([rho, rho[:, None] * vel, e])


def random_pairs(rng, n, nd):
    """Random physical pairs, a fraction of them transonic: velocities of
    the order of the sound speed."""
    rho = 0.3 + 1.5 * rng.random((2, n))
    p = 0.2 + 1.5 * rng.random((2, n))
    vel = 1.3 * rng.normal(size=(2, n, nd))
    return prim_states(rho[0], vel[0], p[0]), prim_states(rho[1], vel[1], p[1])


def normal_index_2d(ixy):
    return (1, 2) if ixy == 1 else (2, 1)   # (mu, mv), 0-based components


TRIPLE3 = {1: (1, 2, 3), 2: (2, 3, 1), 3: (3, 1, 2)}  # (mu, mv, mw), 0-based


# --------------------------------------------------------------------- 2-D
@pytest.mark.parametrize("ixy", [1, 2])
@pytest.mark.parametrize("efix", [False, True])
def test_rpn2_waves_are_roe_eigenvectors(ixy, efix):
    """rpn2: each wave lies in the eigenspace of the Roe matrix with its
    speed; the shear and entropy waves have the textbook eigenvector forms
    (0, 0, 1, v) and (1, u, v, |u|^2/2); without the fix A-dq, A+dq are the
    spectral split of dq."""
    rng = np.random.default_rng(31 + ixy + 2 * efix)
    mu, mv = normal_index_2d(ixy)
    p = orc.make_params(orc.EULER, efix=efix)
    L, R = random_pairs(rng, 60, 2)
    q = pairs_2d(L, R)
    wave, s, amdq, apdq = orc.rpn2(p, ixy, q, q)
    for k in range(len(L)):
        i = 2 * k + 1
        dq = R[k] - L[k]
        vel, h, a = roe_state(L[k], R[k])
        un = vel[mu - 1]
        A = roe_matrix(vel, h, mu - 1)
        lams = (un - a, un, un + a)
        ev = np.sort(np.linalg.eigvals(A).real)
        assert close(ev, [un - a, un, un, un + a], 1 + abs(un) + a)
        P1, P2, P3 = projectors(A, lams)
        sc = np.abs(dq).max() * (1 + abs(un) + a + np.abs(vel).max() ** 2 + h)
        assert close(s[i], [un - a, un, un, un + a], 1 + abs(un) + a)
        W = wave[i]
        assert close(W[0], P1 @ dq, sc)
        assert close(W[3], P3 @ dq, sc)
        assert close(W[1] + W[2], P2 @ dq, sc)
        for w, lam in zip(W, (un - a, un, un, un + a)):
            assert close(A @ w, lam * w, sc * (1 + abs(lam)))
        # shear wave: no density, no normal momentum, energy = v * (transverse momentum)
        assert W[1][0] == 0.0 and W[1][mu] == 0.0
        assert close(W[1][3], vel[mv - 1] * W[1][mv], sc)
        # entropy wave: multiple of (1, u, v, |u|^2 / 2)
        ent = np.zeros(4)
        ent[0], ent[mu], ent[mv], ent[3] = 1.0, un, vel[mv - 1], 0.5 * (vel * vel).sum()
        assert close(W[2], W[2][0] * ent, sc)
        df = euler_flux(R[k], mu - 1) - euler_flux(L[k], mu - 1)
        assert close(amdq[i] + apdq[i], df, sc)
        if not efix:
            am, ap = split(A, lams, dq)
            assert close(amdq[i], am, sc)
            assert close(apdq[i], ap, sc)


@pytest.mark.parametrize("ixy", [1, 2])
def test_rpt2_is_transverse_roe_split(ixy):
    """rpt2: B-asdq and B+asdq are the spectral split of asdq by the Roe
    matrix of the transverse direction at the interface's Roe average; with
    asdq = dq they sum to the transverse flux jump (Roe property of B)."""
    rng = np.random.default_rng(41 + ixy)
    mu, mv = normal_index_2d(ixy)
    p = orc.make_params(orc.EULER)
    L, R = random_pairs(rng, 40, 2)
    q = pairs_2d(L, R)
    asdq = rng.normal(size=q.shape)
    for imp in (1, 2):
        bm, bp = orc.rpt2(p, ixy, imp, q, q, asdq)
        for k in range(len(L)):
            i = 2 * k + 1
            vel, h, a = roe_state(L[k], R[k])
            vt = vel[mv - 1]
            B = roe_matrix(vel, h, mv - 1)
            m, pl = split(B, (vt - a, vt, vt + a), asdq[i])
            sc = np.abs(asdq[i]).max() * (1 + abs(vt) + a + np.abs(vel).max() ** 2 + h)
            assert close(bm[i], m, sc) and close(bp[i], pl, sc)
    jump = np.zeros_like(q)
    jump[1::2] = R - L
    bm, bp = orc.rpt2(p, ixy, 1, q, q, jump)
    for k in range(len(L)):
        i = 2 * k + 1
        dg = euler_flux(R[k], mv - 1) - euler_flux(L[k], mv - 1)
        assert close(bm[i] + bp[i], dg, 1 + np.abs(dg).max(), 1e-12)


def _char_speeds(L, R, W1, W4, mu):
    """Independent characteristic speeds the Harten-Hyman fix compares:
    u-c left of / right of the 1-wave, u+c left of / right of the 4-wave."""
    def spd(q, sgn):
        rho, vel, p = primitive(q)
        with np.errstate(invalid="ignore"):    # a non-physical middle state: NaN
            return vel[mu - 1] + sgn * np.sqrt(GAMMA * p / rho)
    return spd(L, -1), spd(L + W1, -1), spd(R - W4, +1), spd(R, +1)


@pytest.mark.parametrize("ixy", [1, 2])
def test_harten_hyman_split_and_roe_off_sonic(ixy):
    """With the fix: off-sonic interfaces give the Roe split exactly as
    computed here; at a transonic 1-rarefaction with every other wave
    right-going, A-dq = c W1 with the wave split into a part moving at the
    left speed s0 and one at the right speed s1 whose weights c/s0 and
    (l1 - c)/s1 add to one."""
    rng = np.random.default_rng(51 + ixy)
    mu, mv = normal_index_2d(ixy)
    p = orc.make_params(orc.EULER, efix=True)
    L, R = random_pairs(rng, 400, 2)
    q = pairs_2d(L, R)
    wave, s, amdq, apdq = orc.rpn2(p, ixy, q, q)
    n_off = n_tr = 0
    for k in range(len(L)):
        i = 2 * k + 1
        W = wave[i]
        s0, s1, s2, s3 = _char_speeds(L[k], R[k], W[0], W[3], mu)
        vel, h, a = roe_state(L[k], R[k])
        un = vel[mu - 1]
        lam1, lam4 = un - a, un + a
        dq = R[k] - L[k]
        sc = np.abs(dq).max() * (1 + abs(un) + a + np.abs(vel).max() ** 2 + h)
        tr1 = s0 < 0 < s1
        tr4 = s2 < 0 < s3
        if not tr1 and not tr4:
            # no sonic point in either acoustic family
            A = roe_matrix(vel, h, mu - 1)
            am, _ = split(A, (lam1, un, lam4), dq)
            if (s0 >= 0) == (lam1 >= 0) and (s3 > 0) == (lam4 > 0):
                assert close(amdq[i], am, sc), k
                n_off += 1
        elif tr1 and un > 0:
            c = amdq[i] @ W[0] / (W[0] @ W[0])
            assert close(amdq[i], c * W[0], sc), k
            assert abs(c / s0 + (lam1 - c) / s1 - 1.0) < 1e-10 * (1 + abs(1 / s0) + abs(1 / s1))
            if s0 <= lam1 <= s1:         # Roe speed between the two: weights in [0, 1]
                assert s0 <= c <= 0.0
            n_tr += 1
    assert n_off > 100 and n_tr > 5, (n_off, n_tr)


def _fix_direction_conflict(s0, s1, s2, s3, u):
    """The split evaluates the 4-wave fix only when the contact (speed u)
    goes left (clawpack's rpn2 efix order: 1-wave, then 2/3 waves, and the
    4-wave only if those are left-going), so the two branches are mirror
    images except where a transonic acoustic wave and the contact disagree
    in direction: a transonic 4-wave with u >= 0 or a transonic 1-wave with
    u <= 0, which needs jumps too strong for a physical Riemann fan."""
    return (s2 < 0 < s3 and u >= 0) or (s0 < 0 < s1 and u <= 0)


def _mirror_2d(q, mu):
    m = q.copy()
    m[..., mu] = -m[..., mu]
    return m


@pytest.mark.parametrize("ixy", [1, 2])
def test_harten_hyman_reflection_symmetry(ixy):
    """Reflecting x -> -x (swap the states, negate the normal momentum)
    must give A-dq' = M A+dq and A+dq' = M A-dq, fix included: the 4-wave
    branch of the fix is the mirror image of the 1-wave branch."""
    rng = np.random.default_rng(61 + ixy)
    mu, _ = normal_index_2d(ixy)
    p = orc.make_params(orc.EULER, efix=True)
    L, R = random_pairs(rng, 400, 2)
    q = pairs_2d(L, R)
    qm = pairs_2d(_mirror_2d(R, mu), _mirror_2d(L, mu))
    wave, s, amdq, apdq = orc.rpn2(p, ixy, q, q)
    _, _, amm, apm = orc.rpn2(p, ixy, qm, qm)
    n_tr = n_skip = 0
    for k in range(len(L)):
        i = 2 * k + 1
        s0, s1, s2, s3 = _char_speeds(L[k], R[k], wave[i][0], wave[i][3], mu)
        if _fix_direction_conflict(s0, s1, s2, s3, s[i][1]):
            n_skip += 1
            continue
        sc = np.abs(R[k] - L[k]).max() * (1 + np.abs(q).max() ** 2)
        assert close(amm[i], _mirror_2d(apdq[i], mu), sc, 1e-11), k
        assert close(apm[i], _mirror_2d(amdq[i], mu), sc, 1e-11), k
        n_tr += (s0 < 0 < s1) or (s2 < 0 < s3)
    assert n_tr > 10 and n_skip < len(L) // 20, (n_tr, n_skip)


def _rarefaction_family(theta, gamma=GAMMA, vt=(0.3, -0.2)):
    """Left state (rho 1, p 1) and the right state on its exact 1-rarefaction
    at rho 0.55 (isentropic, u + 2c/(gamma-1) constant), both moving with
    normal velocity shifted by theta; transverse velocities vt."""
    rl, pl = 1.0, 1.0
    cl = np.sqrt(gamma * pl / rl)
    rr = 0.55
    pr = pl * (rr / rl) ** gamma
    cr = np.sqrt(gamma * pr / rr)
    ur = 2.0 * (cl - cr) / (gamma - 1.0)
    return (np.array([rl, theta, vt[0], pl]), np.array([rr, theta + ur, vt[1], pr]))


def _hh_amdq(p, ixy, theta):
    mu, mv = normal_index_2d(ixy)
    pl, pr = _rarefaction_family(theta)

    def cons(w):
        vel = np.zeros(2)
        vel[mu - 1], vel[mv - 1] = w[1], w[2]
        return prim_states([w[0]], vel[None], [w[3]])[0]
    L, R = cons(pl), cons(pr)
    q = pairs_2d(L[None], R[None])
    wave, s, amdq, apdq = orc.rpn2(p, ixy, q, q)
    s0, s1, _, _ = _char_speeds(L, R, wave[1][0], wave[1][3], mu)
    return amdq[1], s0, s1


@pytest.mark.parametrize("ixy", [1, 2])
def test_harten_hyman_continuous_at_sonic_points(ixy):
    """Along a Galilean family of exact 1-rarefactions, A-dq has no jump
    where s0 = u_l - c_l crosses 0 (fix starts) nor where s1 = u - c right
    of the 1-wave crosses 0 (fix ends): a wrong weight in the transonic
    formula would jump by O(|W1|) there."""
    p = orc.make_params(orc.EULER, efix=True)

    def root(which, lo, hi):
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            v = _hh_amdq(p, ixy, mid)[which]
            if v < 0:
                lo = mid
            else:
                hi = mid
        return 0.5 * (lo + hi)

    th0 = root(1, 0.0, 3.0)      # s0 crosses 0
    th1 = root(2, -2.0, th0)     # s1 crosses 0 (below th0: s1 > s0)
    assert th1 < th0
    # transonic in between
    _, s0, s1 = _hh_amdq(p, ixy, 0.5 * (th0 + th1))
    assert s0 < 0 < s1
    eps = 1e-9
    for th in (th0, th1):
        a, _, _ = _hh_amdq(p, ixy, th - eps)
        b, _, _ = _hh_amdq(p, ixy, th + eps)
        assert np.abs(a - b).max() < 1e-6, (th, a, b)
    # and the fix changes the answer inside the transonic range
    p0 = orc.make_params(orc.EULER, efix=False)
    a, _, _ = _hh_amdq(p, ixy, 0.5 * (th0 + th1))
    b, _, _ = _hh_amdq(p0, ixy, 0.5 * (th0 + th1))
    assert np.abs(a - b).max() > 1e-3


# --------------------------------------------------------------------- 3-D
@pytest.mark.parametrize("ixyz", [1, 2, 3])
@pytest.mark.parametrize("efix", [False, True])
def test_rpn3_waves_are_roe_eigenvectors(ixyz, efix):
    """rpn3 (3-wave split): acoustic waves are the spectral parts of dq for
    u -+ a, the middle wave is its part in the triple eigenvalue u (entropy
    and both shear jumps); without the fix A-dq, A+dq are the spectral split."""
    rng = np.random.default_rng(71 + ixyz + 3 * efix)
    mu, _, _ = TRIPLE3[ixyz]
    p = orc.make_params3(orc.EULER, efix=efix)
    L, R = random_pairs(rng, 50, 3)
    q = pairs_2d(L, R)
    wave, s, amdq, apdq = orc.rpn3(p, ixyz, q, q)
    for k in range(len(L)):
        i = 2 * k + 1
        dq = R[k] - L[k]
        vel, h, a = roe_state(L[k], R[k])
        un = vel[mu - 1]
        A = roe_matrix(vel, h, mu - 1)
        lams = (un - a, un, un + a)
        ev = np.sort(np.linalg.eigvals(A).real)
        assert close(ev, [un - a, un, un, un, un + a], 1 + abs(un) + a)
        Ps = projectors(A, lams)
        sc = np.abs(dq).max() * (1 + abs(un) + a + np.abs(vel).max() ** 2 + h)
        assert close(s[i], lams, 1 + abs(un) + a)
        for w, P in zip(wave[i], Ps):
            assert close(w, P @ dq, sc)
        df = euler_flux(R[k], mu - 1) - euler_flux(L[k], mu - 1)
        assert close(amdq[i] + apdq[i], df, sc)
        if not efix:
            am, ap = split(A, lams, dq)
            assert close(amdq[i], am, sc) and close(apdq[i], ap, sc)


@pytest.mark.parametrize("ixyz", [1, 2, 3])
@pytest.mark.parametrize("icoor", [2, 3])
@pytest.mark.parametrize("double", [False, True])
def test_split3_is_transverse_roe_split(ixyz, icoor, double):
    """rpt3 / rptt3: the spectral split of a fluctuation by the Roe matrix
    of direction mv (icoor 2) or mw (icoor 3) at the interface's Roe average."""
    rng = np.random.default_rng(81 + ixyz + 3 * icoor + 7 * double)
    mu, mv, mw = TRIPLE3[ixyz]
    md = mv if icoor == 2 else mw
    p = orc.make_params3(orc.EULER)
    L, R = random_pairs(rng, 30, 3)
    q = pairs_2d(L, R)
    asdq = rng.normal(size=q.shape)
    bm, bp = orc.rpt3(p, ixyz, icoor, q, q, asdq, double=double)
    for k in range(len(L)):
        i = 2 * k + 1
        vel, h, a = roe_state(L[k], R[k])
        vt = vel[md - 1]
        B = roe_matrix(vel, h, md - 1)
        m, pl = split(B, (vt - a, vt, vt + a), asdq[i])
        sc = np.abs(asdq[i]).max() * (1 + abs(vt) + a + np.abs(vel).max() ** 2 + h)
        assert close(bm[i], m, sc) and close(bp[i], pl, sc)


@pytest.mark.parametrize("ixyz", [1, 2, 3])
def test_rpn3_harten_hyman(ixyz):
    """3-D fix: the transonic 1-wave two-speed split and reflection symmetry
    (which maps the 1-wave fix onto the 3-wave fix)."""
    rng = np.random.default_rng(91 + ixyz)
    mu, _, _ = TRIPLE3[ixyz]
    p = orc.make_params3(orc.EULER, efix=True)
    L, R = random_pairs(rng, 400, 3)
    q = pairs_2d(L, R)
    Lm, Rm = _mirror_2d(R, mu), _mirror_2d(L, mu)
    qm = pairs_2d(Lm, Rm)
    wave, s, amdq, apdq = orc.rpn3(p, ixyz, q, q)
    _, _, amm, apm = orc.rpn3(p, ixyz, qm, qm)
    n_tr = n_skip = 0
    for k in range(len(L)):
        i = 2 * k + 1
        W = wave[i]
        s0, s1, s2, s3 = _char_speeds(L[k], R[k], W[0], W[2], mu)
        if _fix_direction_conflict(s0, s1, s2, s3, s[i][1]):
            n_skip += 1
            continue
        sc = np.abs(R[k] - L[k]).max() * (1 + np.abs(q).max() ** 2)
        assert close(amm[i], _mirror_2d(apdq[i], mu), sc, 1e-11), k
        assert close(apm[i], _mirror_2d(amdq[i], mu), sc, 1e-11), k
        vel, h, a = roe_state(L[k], R[k])
        un = vel[mu - 1]
        if s0 < 0 < s1 and un > 0:
            c = amdq[i] @ W[0] / (W[0] @ W[0])
            assert close(amdq[i], c * W[0], sc), k
            lam1 = un - a
            assert abs(c / s0 + (lam1 - c) / s1 - 1.0) < 1e-10 * (1 + abs(1 / s0) + abs(1 / s1))
            n_tr += 1
    assert n_tr > 5 and n_skip < len(L) // 20, (n_tr, n_skip)


# --------------------------------------------------------- shallow water
# The physics of the solver plug-in parity case (not built into the library):
# q = (h, hu, hv), f = (hu, hu^2 + g h^2 / 2, huv).  Its Roe matrix is the
# flux Jacobian at the mean depth and the sqrt(h)-weighted velocities.
GRAV = 1.7


def sw_flux(q, k, g=GRAV):
    h = q[0]
    vel = q[1:3] / h
    f = q * vel[k]
    f[1 + k] = f[1 + k] + 0.5 * g * h * h
    return f


def sw_roe_matrix(L, R, k, g=GRAV):
    wl, wr = np.sqrt(L[0]), np.sqrt(R[0])
    vel = (L[1:3] / wl + R[1:3] / wr) / (wl + wr)
    hbar = 0.5 * (L[0] + R[0])
    q0 = np.array([hbar, hbar * vel[0], hbar * vel[1]], dtype=complex)
    J = np.zeros((3, 3))
    for j in range(3):
        qc = q0.copy()
        qc[j] += 1e-30j
        J[:, j] = sw_flux(qc, k, g).imag / 1e-30
    return J, vel, np.sqrt(g * hbar)


def sw_pairs(rng, n):
    h = 0.3 + 1.5 * rng.random((2, n))
    vel = 1.2 * rng.normal(size=(2, n, 2))
    return [np.column_stack([h[s], h[s][:, None] * vel[s]]) for s in (0, 1)]


@pytest.mark.parametrize("ixy", [1, 2])
def test_shallow_water_roe_split(ixy):
    """rpn2 / rpt2 of the shallow-water restatement: waves are the spectral
    parts of the jump for u - c, u, u + c of the Roe matrix, the Roe
    property holds, A-dq / A+dq and B-asdq / B+asdq are spectral splits."""
    rng = np.random.default_rng(101 + ixy)
    mu, mv = normal_index_2d(ixy)
    p = orc.make_params(orc.SHALLOW, grav=GRAV)
    L, R = sw_pairs(rng, 60)
    q = pairs_2d(L, R)
    wave, s, amdq, apdq = orc.rpn2(p, ixy, q, q)
    asdq = rng.normal(size=q.shape)
    bm, bp = orc.rpt2(p, ixy, 1, q, q, asdq)
    for k in range(len(L)):
        i = 2 * k + 1
        dq = R[k] - L[k]
        A, vel, c = sw_roe_matrix(L[k], R[k], mu - 1)
        un, vt = vel[mu - 1], vel[mv - 1]
        lams = (un - c, un, un + c)
        assert close(np.sort(np.linalg.eigvals(A).real), lams, 1 + abs(un) + c)
        sc = np.abs(dq).max() * (1 + abs(un) + c) ** 2
        assert close(s[i], lams, 1 + abs(un) + c)
        for w, P in zip(wave[i], projectors(A, lams)):
            assert close(w, P @ dq, sc)
        df = sw_flux(R[k], mu - 1) - sw_flux(L[k], mu - 1)
        assert close(amdq[i] + apdq[i], df, sc)
        am, ap = split(A, lams, dq)
        assert close(amdq[i], am, sc) and close(apdq[i], ap, sc)
        B, _, _ = sw_roe_matrix(L[k], R[k], mv - 1)
        m, pl = split(B, (vt - c, vt, vt + c), asdq[i])
        sc2 = np.abs(asdq[i]).max() * (1 + abs(vt) + c) ** 2
        assert close(bm[i], m, sc2) and close(bp[i], pl, sc2)
