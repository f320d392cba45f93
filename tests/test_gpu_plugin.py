"""GPU parity of a Riemann solver defined outside the library: the 2-D
shallow-water Roe solver of tests/plugin/shallow_water.hip, registered at
run time through the plug-in contract (amrgpu_register_solver), drives the
library's sweep, ghost fill, registers (qad calls its rpn2, qad.f:129) and
level update unchanged.  FP64 bit for bit against the oracle's restatement
of the same solver (oracle/orc_riemann.c rpn2_shallow / rpt2_shallow, pinned
by tests/test_oracle_euler_eigen.py); FP32 and mixed within relative L1
1e-5.  The reference's counterpart is linking another rpn2/rpt2
(src/2d/step2.f90:7, flux2.f:25-28)."""
import os

import numpy as np
import pytest

from amrclaw_amd import capi
from oracle import orc
from tests.common import berger_oliger, nested_hierarchy, rel_l1

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUGIN = os.path.join(ROOT, "tests", "plugin", "libshallow_water.so")
GRAV = 1.7
_reg = {}


def solver_id():
    if "id" not in _reg:
        ops, keep = capi.load_solver_plugin(PLUGIN, "amrgpu_plugin_shallow_water")
        _reg["keep"] = (ops, keep)
        _reg["id"] = capi.register_solver(ops)
    capi.set_solver_params([GRAV])
    return _reg["id"]


def sw_states(rng, shape):
    """Smooth positive depths with velocities of the order of the celerity."""
    ny, nx = shape
    y, x = np.meshgrid(np.linspace(0, 1, ny), np.linspace(0, 1, nx), indexing="ij")
    ph = rng.random(3) * 6.28
    h = 1.0 + 0.3 * np.sin(3 * x + ph[0]) * np.cos(2 * y + ph[1]) + 0.05 * rng.random(shape)
    u = 0.5 * np.sin(2 * y + ph[2]) + 0.1 * rng.normal(size=shape)
    v = 0.5 * np.cos(2 * x) + 0.1 * rng.normal(size=shape)
    return np.stack([h, h * u, h * v], -1)


def configs(precision, mthbc=(1, 1, 1, 1), **kw):
    """Oracle params and library config of the plug-in solver; the boundary
    conditions go to the library only (the oracle reads them from the
    hierarchy)."""
    p = orc.make_params(orc.SHALLOW, grav=GRAV, **kw)
    c = capi.make_config(solver_id(), nvar=3, mwaves=3, precision=precision, mthbc=mthbc, **kw)
    return p, c


@pytest.mark.parametrize("order,trans,lim", [(2, 2, 4), (2, 1, 1), (2, 0, 3), (1, 2, 0),
                                             (2, 2, 2), (2, 2, 5)])
def test_plugin_stepgrid_fp64_bit_exact(gpu, order, trans, lim):
    print("plugin stepgrid", order, trans, lim, flush=True)
    rng = np.random.default_rng(10 * order + trans + 100 * lim)
    p, c = configs(capi.FP64, order=order, transverse=trans, limiter=lim)
    capi.init(c)
    try:
        for mx, my in ((23, 17), (40, 9), (5, 61)):
            q = sw_states(rng, (my + 4, mx + 4))
            dx, dy = 1.0 / 40, 1.0 / 35
            dt = 0.2 * dx
            qg, fg, cg, dg = capi.stepgrid(q, dx, dy, dt, want_fluxes=True)
            qo, fo, co, do = orc.stepgrid(p, q, dx, dy, dt)
            assert np.array_equal(qg[2:-2, 2:-2], qo[2:-2, 2:-2]), (mx, my)
            assert cg == co and dg == do
    finally:
        capi.finalize()


def _amr(precision, seed):
    rng = np.random.default_rng(seed)
    H = nested_hierarchy(rng, sw_states, 3, mthbc=(1, 1, 3, 3), dt1=0.006)
    p, c = configs(precision, mthbc=(1, 1, 3, 3))
    O = orc.OracleAMR(p, H.copy())
    capi.init(c)
    capi.load_hierarchy(H)

    class D:
        advanc = staticmethod(capi.advance_level)
        update = staticmethod(capi.update_level)
    lo, ld = [], []
    for _ in range(2):
        berger_oliger(O, 1, H.nlev, 2, lo)
        berger_oliger(D, 1, H.nlev, 2, ld)
    return H, O, lo, ld


@pytest.mark.parametrize("seed", [0, 1])
def test_plugin_level_update_fp64_bit_exact(gpu, seed):
    print("plugin amr fp64", seed, flush=True)
    try:
        H, O, lo, ld = _amr(capi.FP64, 1200 + seed)
        assert lo == ld
        for l in range(1, H.nlev + 1):
            assert np.array_equal(capi.read_level(H, l, 0), O.H.levels[l - 1]["qnew"]), l
            if l > 1:
                for k in range(len(H.levels[l - 1]["boxes"])):
                    sv, qc = capi.get_registers(l, k, 3, 4096)
                    ref = O.reg(l, k).reshape(2, -1, 3)
                    assert np.array_equal(sv, ref[0]) and np.array_equal(qc, ref[1]), (l, k)
    finally:
        capi.finalize()


@pytest.mark.parametrize("precision", [capi.FP32, capi.MIXED])
def test_plugin_level_update_fp32(gpu, precision):
    print("plugin amr", precision, flush=True)
    try:
        H, O, lo, ld = _amr(precision, 1300)
        for l in range(1, H.nlev + 1):
            err = rel_l1(capi.read_level(H, l, 0), O.H.levels[l - 1]["qnew"])
            assert err <= 1e-5, (l, err)
    finally:
        capi.finalize()
