"""Shared test helpers: seeded inputs in the reference layout."""
import numpy as np

from oracle import orc
from amrclaw_amd import capi

SOLVERS = {"advection": orc.ADVECTION, "acoustics": orc.ACOUSTICS,
           "euler": orc.EULER}


def euler_states(rng, shape, gamma=1.4):
    rho = 0.5 + rng.random(shape)
    u = rng.normal(size=shape)
    v = rng.normal(size=shape)
    p = 0.3 + rng.random(shape)
    return np.stack([rho, rho * u, rho * v,
                     p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)], -1)


def random_patch(rng, solver, mx, my, mbc=2):
    shape = (my + 2 * mbc, mx + 2 * mbc)
    if solver == orc.EULER:
        return euler_states(rng, shape)
    nv = 1 if solver == orc.ADVECTION else 3
    return rng.normal(size=shape + (nv,))


def quadrant_euler(x, y, gamma=1.4, xs=0.8, ys=0.8):
    """Four-quadrant 2-D Riemann problem (examples/euler_2d_quadrants/qinit.f
    :33-69 with the rounded states of the batched Euler configuration)."""
    x = np.asarray(x)
    y = np.asarray(y)
    st = {1: (1.5, 0.0, 0.0, 1.5), 2: (0.5323, 1.206, 0.0, 0.3),
          3: (0.138, 1.206, 1.206, 0.029), 4: (0.5323, 0.0, 1.206, 0.3)}
    iq = np.where(x >= xs, np.where(y >= ys, 1, 4), np.where(y >= ys, 2, 3))
    q = np.zeros(x.shape + (4,))
    for k, (r, u, v, p) in st.items():
        m = iq == k
        q[m, 0] = r
        q[m, 1] = r * u
        q[m, 2] = r * v
        q[m, 3] = p / (gamma - 1.0) + 0.5 * r * (u * u + v * v)
    return q


def both_params(solver, **kw):
    """Matching oracle params and library config."""
    ckw = dict(kw)
    precision = ckw.pop("precision", capi.FP64)
    p = orc.make_params(solver, **ckw)
    c = capi.make_config(solver, precision=precision, **ckw)
    return p, c


def rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))
