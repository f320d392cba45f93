"""GPU parity of the batched 3-D step (step3 + flux3 + stepgrid) against the
oracle, through the C ABI of include/amrgpu3.h.

FP64: bit-exact (the kernels keep the reference's evaluation order, apply
every face's transverse increments in the reference's line order, and are
built without FMA contraction); the oracle itself is pinned bit for bit to
the reference's src/3d build in test_oracle_sweep3.py.  FP32 bulk: relative
L1 <= 1e-5 of the FP64 oracle for one step.
"""
import numpy as np
import pytest

from amrclaw_amd import capi3
from oracle import orc
from tests.common import rel_l1
from tests.test_oracle_sweep3 import patch

pytestmark = pytest.mark.gpu

CASES = [(orc.EULER, 2, 22, 4), (orc.EULER, 2, 21, 1), (orc.EULER, 2, 11, 3),
         (orc.EULER, 2, 10, 2), (orc.EULER, 1, 0, 0), (orc.EULER, 2, 22, 5),
         (orc.EULER, 2, 20, 4), (orc.EULER, 2, -1, 4), (orc.ADVECTION, 2, 22, 4),
         (orc.ADVECTION, 2, 11, 1), (orc.ADVECTION, 1, 10, 0), (orc.ADVECTION, 2, 20, 3)]
SIZES = [(3, 4, 5), (1, 2, 3), (8, 8, 8), (13, 6, 9), (32, 17, 11)]


def params(case, precision=capi3.FP64, efix=True):
    solver, order, trans, lim = case
    p = orc.make_params3(solver, order=order, transverse=trans, limiter=lim, efix=efix)
    c = capi3.make_config3(solver, order=order, transverse=trans, limiter=lim, efix=efix,
                           precision=precision)
    return p, c


def interior_faces(f, mx, my, mz):
    """The faces the step writes: x-faces 1..mx+1 of interior rows, etc.
    (arrays (mz+4, my+4, mx+4, nvar); cell x sits at index x + 1)."""
    fm, fp, gm, gp, hm, hp = f
    return [fm[2:mz + 2, 2:my + 2, 2:mx + 3], fp[2:mz + 2, 2:my + 2, 2:mx + 3],
            gm[2:mz + 2, 2:my + 3, 2:mx + 2], gp[2:mz + 2, 2:my + 3, 2:mx + 2],
            hm[2:mz + 3, 2:my + 2, 2:mx + 2], hp[2:mz + 3, 2:my + 2, 2:mx + 2]]


def mesh(mx, my, mz):
    dx, dy, dz = 1.0 / (mx + 3), 1.0 / (my + 5), 1.0 / (mz + 2)
    return dx, dy, dz, 0.08 * min(dx, dy, dz)


@pytest.mark.parametrize("case", CASES)
def test_stepgrid3_fp64_bit_exact(gpu, case):
    p, c = params(case)
    capi3.init3(c)
    rng = np.random.default_rng(31 + CASES.index(case))
    for mx, my, mz in SIZES:
        print("stepgrid3 fp64", case, (mx, my, mz), flush=True)
        q = patch(rng, case[0], mx, my, mz)
        dx, dy, dz, dt = mesh(mx, my, mz)
        qo, fo, co, do = orc.stepgrid3(p, q, dx, dy, dz, dt)
        qg, fg, cg, dg = capi3.stepgrid3(q, dx, dy, dz, dt, want_fluxes=True)
        assert np.array_equal(qg, qo), (mx, my, mz)
        for k, (a, b) in enumerate(zip(interior_faces(fg, mx, my, mz),
                                       interior_faces(fo, mx, my, mz))):
            assert np.array_equal(a, b), (mx, my, mz, k)
        assert cg == co and dg == do
    capi3.finalize3()


def test_stepgrid3_no_entropy_fix(gpu):
    case = (orc.EULER, 2, 22, 4)
    p, c = params(case, efix=False)
    capi3.init3(c)
    rng = np.random.default_rng(77)
    q = patch(rng, orc.EULER, 9, 7, 6)
    dx, dy, dz, dt = mesh(9, 7, 6)
    qo, fo, co, do = orc.stepgrid3(p, q, dx, dy, dz, dt)
    qg, fg, cg, dg = capi3.stepgrid3(q, dx, dy, dz, dt)
    assert np.array_equal(qg, qo) and cg == co
    capi3.finalize3()


@pytest.mark.parametrize("plane_block", [1, 3, 0])
def test_batch_fp64_matches_per_grid_oracle(gpu, plane_block):
    """Grids of uneven sizes in one batch (one launch per sweep), three steps
    without refreshing ghost cells, each grid bit-exact with the oracle."""
    case = (orc.EULER, 2, 22, 4)
    p, c = params(case)
    capi3.init3(c)
    rng = np.random.default_rng(100 + plane_block)
    sizes = [(int(a), int(b), int(d)) for a, b, d in rng.integers(1, 21, size=(9, 3))]
    sizes[0] = (32, 32, 32)
    qs = [patch(rng, orc.EULER, *s) for s in sizes]
    capi3.set_patches3(sizes, plane_block)
    for k, q in enumerate(qs):
        capi3.put_patch3(k, q)
    dx, dy, dz, dt = 1.0 / 40, 1.0 / 36, 1.0 / 44, 0.06 / 44
    ref = [q.copy() for q in qs]
    for step in range(3):
        print("batch3 fp64 plane_block", plane_block, "step", step, flush=True)
        cfl = capi3.step_patches3(dx, dy, dz, dt, npatch=len(sizes))
        for k, s in enumerate(sizes):
            ref[k], f, co, _ = orc.stepgrid3(p, ref[k], dx, dy, dz, dt)
            assert cfl[k] == co, (step, k)
            if step == 2:
                fg = capi3.get_fluxes3(k, ref[k].shape)
                for a, b in zip(interior_faces(fg, *s), interior_faces(f, *s)):
                    assert np.array_equal(a, b), (k, s)
    for k, s in enumerate(sizes):
        assert np.array_equal(capi3.get_patch3(k, ref[k].shape), ref[k]), (k, s)
    capi3.finalize3()


@pytest.mark.parametrize("case", [c for c in CASES if c[0] == orc.EULER][:5])
def test_stepgrid3_fp32_within_tolerance(gpu, case):
    p, c = params(case, precision=capi3.FP32)
    capi3.init3(c)
    rng = np.random.default_rng(900 + CASES.index(case))
    for mx, my, mz in SIZES:
        print("stepgrid3 fp32", case, (mx, my, mz), flush=True)
        q = patch(rng, case[0], mx, my, mz)
        dx, dy, dz, dt = mesh(mx, my, mz)
        qo, fo, co, _ = orc.stepgrid3(p, q, dx, dy, dz, dt)
        qg, fg, cg, _ = capi3.stepgrid3(q, dx, dy, dz, dt, want_fluxes=True)
        s = (slice(2, -2),) * 3
        assert rel_l1(qg[s], qo[s]) <= 1e-5, (mx, my, mz)
        assert cg == pytest.approx(co, rel=1e-5)
        # fp and gp are rebuilt from the merged form (fm + (faddp - faddm))
        for k, (a, b) in enumerate(zip(interior_faces(fg, mx, my, mz),
                                       interior_faces(fo, mx, my, mz))):
            assert np.abs(a - b).max() <= 1e-5 * max(np.abs(b).max(), 1.0), (mx, my, mz, k)
    capi3.finalize3()


def test_uniform_state_invariant_fp32(gpu):
    p, c = params((orc.EULER, 2, 22, 4), precision=capi3.FP32)
    capi3.init3(c)
    q = np.empty((10, 9, 12, 5))
    q[:] = (1.2, 0.3, -0.2, 0.1, 2.7)
    qg, _, cg, _ = capi3.stepgrid3(q, 0.1, 0.1, 0.1, 0.01)
    assert np.allclose(qg, q, rtol=2e-6, atol=0)
    assert cg > 0
    capi3.finalize3()


@pytest.mark.parametrize("precision", [capi3.FP64, capi3.FP32])
def test_flag_patches3_matches_oracle(gpu, precision):
    """flag2refine (3-D) on every grid of a batch: FP64 flags equal the
    oracle's (pinned to the reference's flag2refine.f); with FP32 storage the
    flags are those of the FP32-rounded state.  Flagged count and conck
    (the reference's sequential sum, bit for bit) too."""
    c = capi3.make_config3(orc.EULER, precision=precision)
    capi3.init3(c)
    rng = np.random.default_rng(55)
    sizes = [(int(a), int(b), int(d)) for a, b, d in rng.integers(1, 17, size=(7, 3))]
    sizes[0] = (32, 20, 9)
    qs = [patch(rng, orc.EULER, *s) for s in sizes]
    if precision == capi3.FP32:
        qs = [q.astype(np.float32).astype(np.float64) for q in qs]
    capi3.set_patches3(sizes)
    for k, q in enumerate(qs):
        capi3.put_patch3(k, q)
    tol = 1.3
    print("flag3", precision, flush=True)
    n = capi3.flag_patches3(tol)
    total = 0
    for k, q in enumerate(qs):
        f, nk = orc.flag2refine3(q, tol)
        total += nk
        assert np.array_equal(capi3.get_flags3(k, f.shape), f), k
    assert n == total and 0 < total
    h = (1.0 / 32, 1.0 / 20, 1.0 / 9)
    m = capi3.conck3(h)
    assert m == orc.mass3(qs, h)   # the sequential sum of the same doubles, bit for bit
    capi3.finalize3()


FUSED_SIZES = [(32, 32, 32), (1, 2, 3), (10, 10, 10), (11, 21, 9), (20, 8, 26), (3, 30, 4),
               (13, 6, 9), (16, 12, 22), (15, 15, 5), (14, 9, 15)]


@pytest.mark.parametrize("case", CASES)
def test_fused_batch_fp32_within_tolerance(gpu, case):
    """The fused FP32 step (amrgpu3_set_step_mode(1): one workgroup per tile
    of at most 13 cells per direction, every face on chip, each face value added to the
    two cells it separates) on a batch of uneven grids, including grids
    thinner than a tile and grids cut into several tiles per direction:
    each grid within relative L1 1e-5 of the FP64 oracle for each of two
    steps (ghost cells kept), and its Courant numbers those of the phase
    kernels (the same FP32 wave speeds over the same interfaces)."""
    p, c = params(case, precision=capi3.FP32)
    capi3.init3(c)
    rng = np.random.default_rng(700 + CASES.index(case))
    qs = [patch(rng, case[0], *s) for s in FUSED_SIZES]
    dx, dy, dz, dt = 1.0 / 40, 1.0 / 36, 1.0 / 44, 0.06 / 44
    cfls = {}
    for mode in (capi3.STEP_PHASES, capi3.STEP_FUSED):
        capi3.set_patches3(FUSED_SIZES)
        capi3.set_step_mode3(mode)
        for k, q in enumerate(qs):
            capi3.put_patch3(k, q)
        ref = [q.copy() for q in qs]
        cfls[mode] = []
        for step in range(2):
            print("fused3", case, "mode", mode, "step", step, flush=True)
            cfls[mode].append(capi3.step_patches3(dx, dy, dz, dt, npatch=len(qs)))
            for k, s in enumerate(FUSED_SIZES):
                ref[k], _, co, _ = orc.stepgrid3(p, ref[k], dx, dy, dz, dt)
                got = capi3.get_patch3(k, ref[k].shape)
                inner = (slice(2, -2),) * 3
                assert rel_l1(got[inner], ref[k][inner]) <= 1e-5, (mode, step, k, s)
                # ghost cells are the ones put_patch wrote (stored in FP32)
                ghost = np.ones(got.shape[:3], dtype=bool)
                ghost[inner] = False
                want = qs[k][ghost].astype(np.float32).astype(np.float64)
                assert np.array_equal(got[ghost], want), (mode, step, k)
                # after the first step the FP32 states differ at rounding level
                if step == 0:
                    assert cfls[mode][0][k] == pytest.approx(co, rel=1e-5), (mode, k)
    # (the two kernels' FP32 code may contract multiply-adds differently)
    assert np.allclose(cfls[0][0], cfls[1][0], rtol=1e-6, atol=0)
    capi3.finalize3()


def test_fused_step_keeps_no_flux_arrays(gpu):
    p, c = params((orc.EULER, 2, 22, 4), precision=capi3.FP32)
    capi3.init3(c)
    rng = np.random.default_rng(5)
    q = patch(rng, orc.EULER, 6, 5, 4)
    capi3.set_patches3([(6, 5, 4)])
    capi3.put_patch3(0, q)
    with pytest.raises(capi3.AmrgpuError, match="no step has run"):
        capi3.get_fluxes3(0, q.shape)
    capi3.step_patches3(0.1, 0.1, 0.1, 0.002)
    with pytest.raises(capi3.AmrgpuError, match="fused"):
        capi3.get_fluxes3(0, q.shape)
    capi3.set_step_mode3(capi3.STEP_PHASES)
    capi3.step_patches3(0.1, 0.1, 0.1, 0.002)
    capi3.get_fluxes3(0, q.shape)
    capi3.finalize3()


def test_fused_step_is_deterministic(gpu):
    """No atomics: two fused steps from the same state give the same bits."""
    p, c = params((orc.EULER, 2, 22, 4), precision=capi3.FP32)
    capi3.init3(c)
    rng = np.random.default_rng(8)
    qs = [patch(rng, orc.EULER, *s) for s in FUSED_SIZES]
    out = []
    for _ in range(2):
        capi3.set_patches3(FUSED_SIZES)
        for k, q in enumerate(qs):
            capi3.put_patch3(k, q)
        capi3.step_patches3(1.0 / 40, 1.0 / 36, 1.0 / 44, 0.06 / 44)
        out.append([capi3.get_patch3(k, q.shape) for k, q in enumerate(qs)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    capi3.finalize3()
