"""Build libamrgpu.so in-tree with hipcc for gfx950.

    python -m amrclaw_amd.build

Every HIP translation unit is built with packed FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops); the FP64 sweep is built
with -ffp-contract=off so that no multiply-add is fused and its results match
the reference's FP64 arithmetic bit for bit.  The FP32 and mixed sweeps keep
IEEE semantics too (no -ffinite-math-only / -ffast-math): the Roe solver's
entropy fix takes its branches on comparisons that see NaN sound speeds of
non-physical states, as the reference's does.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libamrgpu.so")
ARCH = os.environ.get("AMRGPU_ARCH", "gfx950")

COMMON = ["-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + os.path.join(HERE, "..", "include")]
UNITS = [
    ("sweep_f64.hip", ["-O3", "-ffp-contract=off"]),
    ("sweep_f32.hip", ["-O3"]),
    ("util.hip", ["-O2"]),
    ("noalloc.hip", ["-O2"]),
    ("fill.hip", ["-O3", "-ffp-contract=off"]),
    ("regs.hip", ["-O3", "-ffp-contract=off"]),
    ("flag.hip", ["-O3", "-ffp-contract=off"]),
    ("gauges.hip", ["-O3", "-ffp-contract=off"]),
    ("fill_plan.cpp", ["-O2", "-x", "hip"]),
    ("regs_plan.cpp", ["-O2", "-x", "hip"]),
    ("capi.cpp", ["-O2", "-x", "hip"]),
    ("sweep3_f64.hip", ["-O3", "-ffp-contract=off"]),
    ("sweep3_f32.hip", ["-O3"]),
    ("flag3.hip", ["-O3", "-ffp-contract=off"]),
    ("capi3.cpp", ["-O2", "-x", "hip"]),
    ("amr3.hip", ["-O3", "-ffp-contract=off"]),
    ("amr3_plan.cpp", ["-O2", "-x", "hip"]),
    ("capi3_amr.cpp", ["-O2", "-x", "hip"]),
]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


def _compile(unit, flags):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(OBJ, unit.rsplit(".", 1)[0] + ".o")
    cmd = [_hipcc()] + COMMON + flags + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(UNITS))) as ex:
        objs = list(ex.map(lambda u: _compile(*u), UNITS))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_solver_plugin(src, out, verbose=False):
    """Build a plug-in Riemann solver (a HIP translation unit written against
    include/amrgpu_solver.hpp) into its own shared library, FP64-exact
    (-ffp-contract=off).  The library is registered at run time through
    amrgpu_register_solver; it does not link libamrgpu."""
    os.makedirs(OBJ, exist_ok=True)
    tmp = out + ".tmp"
    cmd = [_hipcc()] + COMMON + ["-O3", "-ffp-contract=off", "-I" + CSRC, "-shared", src,
                                 "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("plug-in build failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
