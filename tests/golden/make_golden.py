"""Write golden vectors of the sweep from the reference's own Fortran.

Runs the reference's step2.f90/flux2.f/inlinelimiter.f and stepgrid.f,
compiled unmodified by `make -C oracle ref` into oracle/_ref/libamrref.so,
on small seeded patches, and stores inputs and outputs in step2_ref.npz.
The physics plugin (rpn2/rpt2) is the oracle's restated solver, so these
vectors pin the reference's sweep, limiter and update code paths.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import orc  # noqa: E402

CASES = [
    # name, solver, order, transverse, limiter, (mx, my)
    ("adv_mc", orc.ADVECTION, 2, 2, 4, (7, 5)),
    ("adv_vanleer_t1", orc.ADVECTION, 2, 1, 3, (6, 6)),
    ("adv_godunov", orc.ADVECTION, 1, 0, 0, (5, 4)),
    ("ac_mc", orc.ACOUSTICS, 2, 2, 4, (6, 7)),
    ("ac_superbee", orc.ACOUSTICS, 2, 2, 2, (5, 5)),
    ("ac_minmod_t0", orc.ACOUSTICS, 2, 0, 1, (4, 6)),
    ("eu_mc", orc.EULER, 2, 2, 4, (6, 5)),
    ("eu_minmod", orc.EULER, 2, 2, 1, (5, 6)),
    ("eu_bw_t1", orc.EULER, 2, 1, 5, (4, 4)),
    ("eu_order1", orc.EULER, 1, 2, 0, (5, 5)),
]


def states(rng, solver, mx, my):
    shape = (my + 4, mx + 4)
    if solver == orc.EULER:
        rho = 0.5 + rng.random(shape)
        u, v = rng.normal(size=shape), rng.normal(size=shape)
        p = 0.3 + rng.random(shape)
        return np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)], -1)
    nv = 1 if solver == orc.ADVECTION else 3
    return rng.normal(size=shape + (nv,))


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for name, solver, order, trans, lim, (mx, my) in CASES:
        p = orc.make_params(solver, order=order, transverse=trans, limiter=lim)
        q = states(rng, solver, mx, my)
        dx, dy, dt = 0.1, 0.125, 0.02
        f, cfl = orc.ref_step2(p, q, dx, dy, dt)
        qn, _, cflg, dtn = orc.ref_stepgrid(p, q, dx, dy, dt)
        out[name + "/q"] = q
        out[name + "/meta"] = np.array([solver, order, trans, lim, mx, my, dx, dy, dt])
        for k, a in zip(("fm", "fp", "gm", "gp"), f):
            out[name + "/" + k] = a
        out[name + "/cfl"] = np.array([cfl])
        out[name + "/qnew"] = qn
        out[name + "/dtnew"] = np.array([dtn])
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "step2_ref.npz"), **out)


if __name__ == "__main__":
    main()
