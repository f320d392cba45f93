"""CPU checks of the C ABI: the built library exports every entry point that
include/amrgpu.h declares, the ctypes config struct matches the C struct,
and argument validation fails cleanly before any device call."""
import ctypes as C
import os
import re

import pytest

from amrclaw_amd import capi

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "include", "amrgpu.h")


def declared_functions():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"^int\s+(amrgpu_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "amrgpu_init" in names and "amrgpu_sweep_level" in names


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_config_struct_layout():
    """sizeof(amrgpu_config) as laid out by the C compiler equals ctypes'."""
    import subprocess
    import tempfile
    src = '#include "%s"\n#include <stdio.h>\n#include <stddef.h>\n' \
          'int main(void){printf("%%zu %%zu %%zu %%zu\\n", sizeof(amrgpu_config),' \
          ' offsetof(amrgpu_config, gamma), offsetof(amrgpu_config, precision),' \
          ' offsetof(amrgpu_config, auxtype));return 0;}\n' % HDR
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == C.sizeof(capi.AmrgpuConfig)
    assert int(out[1]) == capi.AmrgpuConfig.gamma.offset
    assert int(out[2]) == capi.AmrgpuConfig.precision.offset
    assert int(out[3]) == capi.AmrgpuConfig.auxtype.offset


@pytest.mark.parametrize("field,value", [("solver", 9), ("nvar", 7), ("nghost", 3),
                                         ("mcapa", 1), ("precision", 5)])
def test_init_rejects_bad_config(field, value):
    cfg = capi.make_config(capi.EULER)
    setattr(cfg, field, value)
    rc = capi.lib().amrgpu_init(C.byref(cfg))
    assert rc != 0
    buf = C.create_string_buffer(256)
    capi.lib().amrgpu_last_error(buf, 256)
    assert buf.value


@pytest.mark.parametrize("kw,msg", [
    (dict(solver=capi.VCADV, naux=1), "aux(1:2)"),
    (dict(solver=capi.VCADV, naux=3, mcapa=3, precision=capi.MIXED), "mixed"),
    (dict(solver=capi.VCADV, naux=3, mcapa=4), "mcapa"),
    (dict(solver=capi.VCADV, naux=3, auxtype=(1, 2, 7)), "auxtype"),
    (dict(solver=capi.ADVECTION, naux=9), "naux"),
    (dict(solver=capi.ADVECTION, mthbc=(3, 3, 1, 1)), "wall"),
    (dict(solver=capi.VCADV, naux=2, mthbc=(1, 1, 3, 3)), "wall")])
def test_init_rejects_bad_aux_config(kw, msg):
    """Aux and capacity settings the library does not run are refused at
    init with a message (the mixed precision has no aux sweep)."""
    solver = kw.pop("solver")
    cfg = capi.make_config(solver, **kw)
    rc = capi.lib().amrgpu_init(C.byref(cfg))
    assert rc != 0
    buf = C.create_string_buffer(256)
    capi.lib().amrgpu_last_error(buf, 256)
    assert msg in buf.value.decode().lower()


def test_calls_before_init_fail_cleanly():
    capi.lib().amrgpu_finalize()
    assert capi.lib().amrgpu_sweep_level(1, 0.1, None, None) != 0
    assert capi.lib().amrgpu_synchronize() != 0


# ---- 3-D step: include/amrgpu3.h ----
HDR3 = os.path.join(os.path.dirname(HDR), "amrgpu3.h")


def declared_functions3():
    txt = open(HDR3).read()
    return sorted(set(re.findall(r"^int\s+(amrgpu3_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_3d_symbol():
    from amrclaw_amd import capi3
    names = declared_functions3()
    assert "amrgpu3_step_patches" in names and "amrgpu3_stepgrid" in names
    L = capi3.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_config3_struct_layout():
    import subprocess
    import tempfile
    from amrclaw_amd import capi3
    src = '#include "%s"\n#include <stdio.h>\n#include <stddef.h>\n' \
          'int main(void){printf("%%zu %%zu %%zu %%zu\\n", sizeof(amrgpu3_config),' \
          ' offsetof(amrgpu3_config, wbar), offsetof(amrgpu3_config, gamma),' \
          ' offsetof(amrgpu3_config, precision));return 0;}\n' % HDR3
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", c, "-o", exe], check=True)
        out = [int(v) for v in
               subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    K = capi3.Amrgpu3Config
    assert out == [C.sizeof(K), K.wbar.offset, K.gamma.offset, K.precision.offset]


@pytest.mark.parametrize("field,value", [("solver", 2), ("nvar", 4), ("mwaves", 4), ("nghost", 3),
                                         ("naux", 1), ("mcapa", 1), ("precision", 5)])
def test_init3_rejects_bad_config(field, value):
    from amrclaw_amd import capi3
    cfg = capi3.make_config3(capi.EULER)
    setattr(cfg, field, value)
    assert capi3.lib().amrgpu3_init(C.byref(cfg)) != 0
    buf = C.create_string_buffer(256)
    capi3.lib().amrgpu3_last_error(buf, 256)
    assert buf.value


@pytest.mark.parametrize("trans", [12, 33, 3 * 10 + 1])
def test_init3_rejects_bad_transverse_option(trans):
    """method(3) = 12 asks flux3 for the cqxx splits only m3 = 2 computes;
    m3 or m4 above 2 has no meaning in flux3.f:205-211."""
    from amrclaw_amd import capi3
    cfg = capi3.make_config3(3, transverse=trans)
    assert capi3.lib().amrgpu3_init(C.byref(cfg)) != 0


def test_3d_calls_before_init_fail_cleanly():
    from amrclaw_amd import capi3
    capi3.lib().amrgpu3_finalize()
    assert capi3.lib().amrgpu3_step_patches(0.1, 0.1, 0.1, 0.01, None) != 0
    assert capi3.lib().amrgpu3_synchronize() != 0
