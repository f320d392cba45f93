"""The Fortran side of the drop-in boundary compiles against the
reference's amr_module, defines advanc/update with the reference's names,
and every amrgpu_* routine it binds is exported by libamrgpu.so."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "amrclaw_amd", "fortran", "amrgpu_shim.f90")
# the modules of the reference's whole 2-D program (make -C oracle driver):
# the shim uses amr_module and gauges_module
MODS = os.path.join(ROOT, "oracle", "_ref", "drv")
LIB = os.path.join(ROOT, "amrclaw_amd", "libamrgpu.so")
FC = shutil.which("amdflang") or "/opt/rocm/llvm/bin/amdflang"


@pytest.mark.skipif(not os.path.exists(FC) or not os.path.exists(os.path.join(MODS, "amr_module.mod")),
                    reason="amdflang or the reference module build absent")
def test_shim_compiles_and_binds_exported_symbols(tmp_path):
    obj = tmp_path / "shim.o"
    r = subprocess.run([FC, "-c", "-I" + MODS, SHIM, "-o", str(obj), "-module-dir", str(tmp_path)],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    syms = subprocess.run(["nm", str(obj)], capture_output=True, text=True).stdout.split("\n")
    defined = {l.split()[-1] for l in syms if " T " in l}
    undefined = {l.split()[-1] for l in syms if " U " in l}
    assert "advanc_" in defined and "update_" in defined
    # C symbols of the library (the application's Fortran hook ends in "_")
    bound = {s for s in undefined if s.startswith("amrgpu_") and not s.endswith("_")}
    assert {"amrgpu_advance_level", "amrgpu_update_level", "amrgpu_put_level",
            "amrgpu_get_level", "amrgpu_set_level", "amrgpu_init", "amrgpu_get_registers",
            "amrgpu_set_gauges", "amrgpu_set_gauge_sources", "amrgpu_get_gauge_records"} <= bound
    # the application supplies its physics next to setprob
    assert "amrgpu_app_physics_" in undefined
    exported = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                              text=True).stdout
    for s in bound:
        assert (" " + s + "\n") in exported, s


SHIM3 = os.path.join(ROOT, "amrclaw_amd", "fortran", "amrgpu3_shim.f90")
MODS3 = os.path.join(ROOT, "oracle", "_ref", "obj3")


@pytest.mark.skipif(not os.path.exists(FC) or not os.path.exists(os.path.join(MODS3, "amr_module.mod")),
                    reason="amdflang or the reference 3-D module build absent")
def test_3d_shim_compiles_and_binds_exported_symbols(tmp_path):
    """The 3-D shim defines stepgrid with src/3d/stepgrid.f's name against the
    reference's 3-D amr_module and binds only exported amrgpu3_* symbols."""
    obj = tmp_path / "shim3.o"
    r = subprocess.run([FC, "-c", "-I" + MODS3, SHIM3, "-o", str(obj), "-module-dir", str(tmp_path)],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    syms = subprocess.run(["nm", str(obj)], capture_output=True, text=True).stdout.split("\n")
    defined = {l.split()[-1] for l in syms if " T " in l}
    undefined = {l.split()[-1] for l in syms if " U " in l}
    assert "stepgrid_" in defined
    bound = {s for s in undefined if s.startswith("amrgpu3_")}
    assert {"amrgpu3_stepgrid", "amrgpu3_last_error"} <= bound
    exported = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                              text=True).stdout
    for s in bound:
        assert (" " + s + "\n") in exported, s


SHIM3L = os.path.join(ROOT, "amrclaw_amd", "fortran", "amrgpu3_level_shim.f90")


@pytest.mark.skipif(not os.path.exists(FC) or not os.path.exists(os.path.join(MODS3, "amr_module.mod")),
                    reason="amdflang or the reference 3-D module build absent")
def test_3d_level_shim_compiles_and_binds_exported_symbols(tmp_path):
    """The 3-D level shim defines advanc and update with src/3d's names
    against the reference's 3-D amr_module and binds only exported
    amrgpu3_* symbols."""
    objs = []
    for src in (SHIM3, SHIM3L):
        obj = tmp_path / (os.path.basename(src) + ".o")
        r = subprocess.run([FC, "-c", "-I" + MODS3, src, "-o", str(obj), "-module-dir",
                            str(tmp_path), "-I" + str(tmp_path)],
                           capture_output=True, text=True, cwd=tmp_path)
        assert r.returncode == 0, r.stderr
        objs.append(obj)
    syms = subprocess.run(["nm", str(objs[1])], capture_output=True, text=True).stdout.split("\n")
    defined = {l.split()[-1] for l in syms if " T " in l}
    undefined = {l.split()[-1] for l in syms if " U " in l}
    assert {"advanc_", "update_", "amrgpu3_upload_hierarchy_",
            "amrgpu3_download_hierarchy_", "amrgpu3_spest_errest_"} <= defined
    bound = {s for s in undefined if s.startswith("amrgpu3_") and not s.endswith("_")}
    assert {"amrgpu3_advance_level", "amrgpu3_update_level", "amrgpu3_set_geometry",
            "amrgpu3_set_level", "amrgpu3_put_level", "amrgpu3_get_level",
            "amrgpu3_flag_level", "amrgpu3_get_level_flags"} <= bound
    # the flagging replacement hands its flags to the reference's own routines
    assert {"setflags_", "igetsp_"} <= undefined
    exported = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                              text=True).stdout
    for s in bound:
        assert (" " + s + "\n") in exported, s
