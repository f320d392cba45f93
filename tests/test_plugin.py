"""CPU checks of the Riemann-solver plug-in contract (include/amrgpu.h
amrgpu_register_solver, include/amrgpu_solver.hpp): a solver built outside
the library (tests/plugin/shallow_water.hip) exports a launcher table of the
right shape, the library registers it and refuses tables built against
other headers, bad ids and bad constants, all before any device call."""
import ctypes as C
import os
import subprocess
import tempfile

import pytest

from amrclaw_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUGIN = os.path.join(ROOT, "tests", "plugin", "libshallow_water.so")
ENTRY = "amrgpu_plugin_shallow_water"

needs_plugin = pytest.mark.skipif(not os.path.exists(PLUGIN), reason="plug-in not built")


@needs_plugin
def test_plugin_table():
    ops, _ = capi.load_solver_plugin(PLUGIN, ENTRY)
    assert ops.abi == capi.SOLVER_ABI
    assert (ops.nvar, ops.mwaves) == (3, 3)
    assert ops.name == b"shallow water (Roe)"
    assert all(getattr(ops, f) for f in ("sweep_f64", "sweep_f32", "sweep_mixed", "qad"))


def test_ops_struct_layout():
    """sizeof(amrgpu_solver_ops) and its offsets as the C compiler lays them
    out equal the ctypes mirror's."""
    hdr = os.path.join(ROOT, "include", "amrgpu.h")
    src = ('#include "%s"\n#include <stdio.h>\n#include <stddef.h>\nint main(void){'
           'printf("%%zu %%zu %%zu\\n", sizeof(amrgpu_solver_ops), offsetof(amrgpu_solver_ops,'
           ' name), offsetof(amrgpu_solver_ops, qad));return 0;}\n' % hdr)
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", c, "-o", exe], check=True)
        out = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                              check=True).stdout.split()]
    assert out == [C.sizeof(capi.AmrgpuSolverOps), capi.AmrgpuSolverOps.name.offset,
                   capi.AmrgpuSolverOps.qad.offset]


@needs_plugin
def test_register_and_refusals():
    ops, keep = capi.load_solver_plugin(PLUGIN, ENTRY)
    sid = capi.register_solver(ops)
    assert sid >= capi.SOLVER_PLUGIN0
    bad = capi.AmrgpuSolverOps.from_buffer_copy(ops)
    bad.physics_bytes += 8                     # built against other headers
    with pytest.raises(capi.AmrgpuError, match="other headers"):
        capi.register_solver(bad)
    bad = capi.AmrgpuSolverOps.from_buffer_copy(ops)
    bad.abi = 99
    with pytest.raises(capi.AmrgpuError, match="ABI"):
        capi.register_solver(bad)
    bad = capi.AmrgpuSolverOps.from_buffer_copy(ops)
    bad.qad = None
    with pytest.raises(capi.AmrgpuError, match="NULL"):
        capi.register_solver(bad)
    with pytest.raises(capi.AmrgpuError):
        capi.set_solver_params([1.0] * 17)
    # a config naming the plug-in with the wrong nvar, or an unknown id, fails
    cfg = capi.make_config(sid, nvar=4, mwaves=3)
    assert capi.lib().amrgpu_init(C.byref(cfg)) != 0
    cfg = capi.make_config(sid + 50, nvar=3, mwaves=3)
    assert capi.lib().amrgpu_init(C.byref(cfg)) != 0
