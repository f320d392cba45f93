"""Runs of the reference's own 2-D driver (amr2, tick, regrid, flagging,
valout, gauges: src/2d compiled unmodified by `make -C oracle driver`) on
the run configurations, written with amrclaw_amd.rundata:

  amr2_ref   the reference's advanc.f / update.f on the CPU;
  amr2_gpu   the same program with advanc / update replaced by the
             library's Fortran shim (amrclaw_amd/fortran/amrgpu_shim.f90)
             over libamrgpu.so.

setprob.data (oracle/driver/app.f90): problem (1 square, 2 disk, 3
acoustic pulse, 4 Euler quadrants), solver (1 advection, 2 acoustics,
3 Euler), ubar, vbar, rho, K, gamma, efix, precision of a device run."""
import os
import subprocess
import numpy as np

from amrclaw_amd.rundata import RunData

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = {"ref": os.path.join(ROOT, "oracle", "_ref", "amr2_ref"),
       "gpu": os.path.join(ROOT, "oracle", "_ref", "amr2_gpu"),
       "annulus_ref": os.path.join(ROOT, "oracle", "_ref", "annulus_ref"),
       "annulus_gpu": os.path.join(ROOT, "oracle", "_ref", "annulus_gpu")}
PERIODIC = dict(bc_lower=["periodic", "periodic"], bc_upper=["periodic", "periodic"])
EXTRAP = dict(bc_lower=["extrap", "extrap"], bc_upper=["extrap", "extrap"])


def square(**kw):
    """tests/advection_2d_square/setrun.py of the reference, verbatim in
    its parameters (gauges 1 and 2, regions, van Leer, t = 0.4)."""
    p = dict(num_cells=[50, 50], num_eqn=1, num_waves=1, tfinal=0.4, num_output_times=1,
             output_t0=True, dt_initial=0.016, cfl_desired=0.9, cfl_max=1.0,
             limiter=["vanleer"], amr_levels_max=3, refinement_ratios_x=[2, 2],
             refinement_ratios_y=[2, 2], refinement_ratios_t=[2, 2], flag2refine_tol=0.05,
             regrid_interval=2, regrid_buffer_width=3, clustering_cutoff=0.7,
             regions=[[1, 1, 0, 1e10, 0., 1., 0., 1.], [1, 3, 0, 1e10, 0., 1., 0., 0.7]],
             gauges=[[1, 0.65, 0.4, 0., 10.], [2, 0.2, 0.8, 0., 10.]],
             probdata=[1, 1, 0.5, 1.0, 1.0, 4.0, 1.4, 1, 0], **PERIODIC)
    p.update(kw)
    return RunData(**p)


def annulus(**kw):
    """tests/advection_2d_annulus/setrun.py of the reference in its
    parameters: variable-coefficient advection on the polar grid
    0.2 <= r <= 1, 0 <= theta <= 2 pi (20 x 120, extrapolation in r,
    periodic in theta), edge velocities and the capacity in aux
    (xleft, yleft, capacity; capa_index 3), van Leer, three levels with
    ratios 2, 2, regions, gauges 1 and 2, t = 0.5; setprob.data holds the
    two Gaussians of qinit.f."""
    p = dict(lower=[0.2, 0.0], upper=[1.0, 2.0 * np.pi], num_cells=[20, 120], num_eqn=1,
             num_waves=1, num_aux=3, capa_index=3, tfinal=0.5, num_output_times=1,
             output_t0=True, dt_initial=0.005, cfl_desired=0.9, cfl_max=1.0, steps_max=1000,
             limiter=["vanleer"], amr_levels_max=3, refinement_ratios_x=[2, 2],
             refinement_ratios_y=[2, 2], refinement_ratios_t=[2, 2],
             aux_type=["xleft", "yleft", "capacity"], flag2refine_tol=0.02, regrid_interval=2,
             regrid_buffer_width=3, clustering_cutoff=0.7,
             regions=[[1, 2, 0., 10., 0.2, 1., 0., 2. * np.pi],
                      [3, 3, 0., 10., 0.5, 1., 0., 0.5 * np.pi]],
             gauges=[[1, 0.5, 0.5 * np.pi, 0., 10.], [2, 0.5, 1.5 * np.pi, 0., 10.]],
             probdata=[1.0, 40.0, -0.5, 0.0, -1.0, 40.0, 0.5, 0.0],
             bc_lower=["extrap", "periodic"], bc_upper=["extrap", "periodic"])
    p.update(kw)
    return RunData(**p)


def config(name, precision=0, **kw):
    """The 2-D run configurations: C1 advection of a disk (periodic 50^2,
    ratios 2, 2, MC, t = 1), C2 acoustic pulse (50^2 on [-1, 1]^2, ratios
    2, 4, t = 0.6), C3 Euler quadrants (80^2, ratios 2, 4, t = 0.8); flag
    tolerances 0.05 / 0.02 / 0.1, regrid every 2 steps, buffer 3, CFL 0.9;
    two gauges each (C1's second with a time window)."""
    common = dict(output_t0=True, num_output_times=2, cfl_desired=0.9, cfl_max=1.0,
                  limiter=["mc"], amr_levels_max=3, regrid_interval=2,
                  regrid_buffer_width=3, clustering_cutoff=0.7, output_format="binary",
                  tprint=True)
    if name == "C1":
        p = dict(num_cells=[50, 50], num_eqn=1, num_waves=1, tfinal=1.0, dt_initial=0.016,
                 refinement_ratios_x=[2, 2], refinement_ratios_y=[2, 2],
                 refinement_ratios_t=[2, 2], flag2refine_tol=0.05,
                 gauges=[[1, 0.3, 0.6, 0.0, 10.0], [2, 0.5, 0.5, 0.2, 0.8]],
                 probdata=[2, 1, 0.5, 1.0, 1.0, 4.0, 1.4, 1, precision], **PERIODIC)
    elif name == "C2":
        p = dict(num_cells=[50, 50], lower=[-1.0, -1.0], upper=[1.0, 1.0], num_eqn=3,
                 num_waves=2, tfinal=0.6, dt_initial=0.01, refinement_ratios_x=[2, 4],
                 refinement_ratios_y=[2, 4], refinement_ratios_t=[2, 4], flag2refine_tol=0.02,
                 gauges=[[1, 0.2, 0.1, 0.0, 10.0], [2, -0.5, 0.5, 0.0, 10.0]],
                 probdata=[3, 2, 0.5, 1.0, 1.0, 4.0, 1.4, 1, precision], **EXTRAP)
    elif name == "C3":
        p = dict(num_cells=[80, 80], num_eqn=4, num_waves=4, tfinal=0.8, dt_initial=0.002,
                 refinement_ratios_x=[2, 4], refinement_ratios_y=[2, 4],
                 refinement_ratios_t=[2, 4], flag2refine_tol=0.1,
                 gauges=[[1, 0.5, 0.5, 0.0, 10.0], [2, 0.85, 0.85, 0.0, 10.0]],
                 probdata=[4, 3, 0.5, 1.0, 1.0, 4.0, 1.4, 1, precision], **EXTRAP)
    else:
        raise ValueError(name)
    common.update(p)
    common.update(kw)
    return RunData(**common)


def run(which, rd, outdir, timeout=600, threads=8, extra_env=None):
    """Write the data files into outdir and run one driver there; returns
    the completed process (stdout and stderr in outdir/run.log)."""
    os.makedirs(outdir, exist_ok=True)
    rd.write(outdir)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), **(extra_env or {}))
    with open(os.path.join(outdir, "run.log"), "w") as log:
        r = subprocess.run([EXE[which]], cwd=outdir, stdout=log, stderr=subprocess.STDOUT,
                           env=env, timeout=timeout)
    return r


def outputs(outdir):
    """The files a run writes that depend on the solution: frames (fort.q /
    fort.b / fort.t), gauges and the conservation lines of fort.amr."""
    out = {}
    for f in sorted(os.listdir(outdir)):
        if f.startswith(("fort.q", "fort.b", "fort.t", "gauge")):
            with open(os.path.join(outdir, f), "rb") as fh:
                out[f] = fh.read()
    with open(os.path.join(outdir, "fort.amr")) as fh:
        out["fort.amr:steps"] = "".join(l for l in fh if "AMRCLAW" in l or "mass" in l)
    return out


LFDIM = 5000                      # amr_module.f90:235, lfree(lfdim, 2)


def RESTART_FILES(chk):
    """What a restart from checkpoint `chk` reads besides the data files:
    the checkpoint and its time stamp (restrt.f:39-60), the timing table
    amr2.f90 appends to, and the gauge files gauges_module.f90:198 opens
    with status 'old' to append to."""
    return [chk, chk.replace("chk", "tck"), "timing.csv", "gauge00001.txt", "gauge00002.txt"]


def checkpoint_records(path):
    """The records of a checkpoint file (check.f:66-82, unformatted
    sequential: each record between two 4-byte lengths) with the wall and
    CPU timers of the last record cut off: lenmax/lendim/memsize/maxgr,
    alloc (solutions, registers, boundary lists), hxposs..icheck, the free
    list, the node tables and run state, and the counters avenumgrids,
    iregridcount, evol, rvol, rvoll, lentot, tmass0, cflmax (236 bytes;
    tvoll and the time* timers follow).

    alloc's free blocks (the free list lfree(1:lenf) of the next record,
    igetsp.f:17-38) are zeroed: they hold whatever the last released block
    held, such as saveqc.f:60-104's temporary of coarse values, which a run
    that builds those values elsewhere never writes; restrt.f reads them
    back but nothing reads a word before igetsp hands it out again."""
    import struct
    b = open(path, "rb").read()
    out, i = [], 0
    while i < len(b):
        n = struct.unpack("<i", b[i:i + 4])[0]
        out.append(b[i + 4:i + 4 + n])
        i += 8 + n
    lfree = np.frombuffer(out[3][:8 * LFDIM], "<i4").reshape(2, LFDIM)
    lenf = struct.unpack("<i", out[3][8 * LFDIM:8 * LFDIM + 4])[0]
    alloc = np.frombuffer(out[1], "<i8").copy()
    for k in range(lenf):
        start, n = int(lfree[0, k]), int(lfree[1, k])
        if n > 0:
            alloc[max(start - 1, 0):start - 1 + n] = 0
    out[1] = alloc.tobytes()
    out[-1] = out[-1][:236]
    return out


def first_difference(a, b, context=3):
    """Where two outputs first differ: the lines around it for text, the
    byte offset for binary."""
    if isinstance(a, str):
        la, lb = a.splitlines(), b.splitlines()
        for i, (x, y) in enumerate(zip(la, lb)):
            if x != y:
                lo = max(0, i - context)
                return "line %d:\n ref: %s\n dev: %s" % (
                    i, "\n      ".join(la[lo:i + context]), "\n      ".join(lb[lo:i + context]))
        return "lengths %d / %d lines" % (len(la), len(lb))
    n = min(len(a), len(b))
    for i in range(n):
        if a[i] != b[i]:
            return "byte %d of %d / %d" % (i, len(a), len(b))
    return "lengths %d / %d bytes" % (len(a), len(b))

