"""Run-parameter files for the reference's 2-D driver (amr2.f90).

The reference's Python front end (setrun.py) fills a clawpack
``ClawRunData`` and writes ``claw.data``, ``amr.data``, ``regions.data``,
``gauges.data`` and the application's ``setprob.data``; the Fortran driver
reads them back in a fixed order (src/2d/amr2.f90:152-392,
regions_module.f90:50-63, gauges_module.f90:119-170).  clawpack's writer is
not part of the reference tree, so this module writes the same files from a
plain description of a run, with the field names of ``ClawRunData``
(``num_cells``, ``limiter``, ``bc_lower``, ``amr_levels_max``,
``regrid_interval``, ...), for driving the kept driver with the device
library linked in.

Every data file opens with comment lines starting with '#' and one blank
line: opendatafile.f reads lines until the first one not starting with '#'
and discards it.  Values are written so that list-directed Fortran reads
them back exactly (floats in Python's shortest round-trip form, logicals as
T / F).
"""
import os

LIMITERS = {"none": 0, "minmod": 1, "superbee": 2, "vanleer": 3, "mc": 4,
            "beam-warming": 5}
BCS = {"user": 0, "extrap": 1, "periodic": 2, "wall": 3}
TRANSVERSE = {"none": 0, "increment": 1, "all": 2}
OUTPUT_FORMATS = {"ascii": 1, "binary": 3}


def _fmt(v):
    if isinstance(v, bool):
        return "T" if v else "F"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        return repr(v)
    return str(v)


def _line(*vals):
    return " ".join(_fmt(v) for v in vals)


class RunData:
    """Parameters of one run.  Defaults follow clawpack's data classes;
    ``probdata`` is the ordered list of setprob.data values the
    application's setprob reads."""

    def __init__(self, **kw):
        # claw.data (amr2.f90:152-327)
        self.lower = [0.0, 0.0]
        self.upper = [1.0, 1.0]
        self.num_cells = [50, 50]
        self.num_eqn = 1
        self.num_waves = 1
        self.num_aux = 0
        self.t0 = 0.0
        self.output_style = 1
        self.num_output_times = 1
        self.tfinal = 1.0
        self.output_t0 = True
        self.output_times = []
        self.output_step_interval = 1
        self.total_steps = 1
        self.output_format = "ascii"
        self.output_q_components = None          # all
        self.output_aux_components = None        # none
        self.output_aux_onlyonce = True
        self.dt_initial = 1e-5
        self.dt_max = 1e99
        self.cfl_max = 1.0
        self.cfl_desired = 0.9
        self.steps_max = 100000
        self.dt_variable = True
        self.order = 2
        self.transverse_waves = "all"
        self.dimensional_split = 0
        self.verbosity = 0
        self.source_split = 0
        self.capa_index = 0
        self.use_fwaves = False
        self.limiter = ["mc"]
        self.num_ghost = 2
        self.bc_lower = ["extrap", "extrap"]
        self.bc_upper = ["extrap", "extrap"]
        self.restart = False
        self.restart_file = "fort.chk00000"
        self.checkpt_style = 0
        self.checkpt_times = []
        self.checkpt_interval = 1
        # amr.data (amr2.f90:331-380)
        self.amr_levels_max = 1
        self.refinement_ratios_x = []
        self.refinement_ratios_y = []
        self.refinement_ratios_t = []
        self.aux_type = []
        self.flag_richardson = False
        self.flag_richardson_tol = 0.1
        self.flag2refine = True
        self.flag2refine_tol = 0.05
        self.regrid_interval = 2
        self.regrid_buffer_width = 3
        self.clustering_cutoff = 0.7
        self.verbosity_regrid = 0
        self.dprint = self.eprint = self.edebug = self.gprint = self.nprint = False
        self.pprint = self.rprint = self.sprint = self.tprint = self.uprint = False
        # regions.data, gauges.data, setprob.data
        self.regions = []          # [min_level, max_level, t1, t2, x1, x2, y1, y2]
        self.gauges = []           # [gauge_no, x, y, t1, t2]
        self.gauge_display_format = "e15.7"
        self.gauge_min_time_increment = 0.0
        self.probdata = []
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError("unknown run parameter %r" % k)
            setattr(self, k, v)

    # ------------------------------------------------------------------
    def claw_lines(self):
        L = [_line(2)]                                   # ndim, read as (i1)
        L.append(_line(float(self.lower[0]), float(self.lower[1])))
        L.append(_line(float(self.upper[0]), float(self.upper[1])))
        L.append(_line(int(self.num_cells[0]), int(self.num_cells[1])))
        L.append(_line(int(self.num_eqn)))
        L.append(_line(int(self.num_waves)))
        L.append(_line(int(self.num_aux)))
        L.append(_line(float(self.t0)))
        L.append(_line(int(self.output_style)))
        if self.output_style == 1:
            L.append(_line(int(self.num_output_times)))
            L.append(_line(float(self.tfinal)))
            L.append(_line(bool(self.output_t0)))
        elif self.output_style == 2:
            L.append(_line(len(self.output_times)))
            L.append(_line(*[float(t) for t in self.output_times]))
        elif self.output_style == 3:
            L.append(_line(int(self.output_step_interval)))
            L.append(_line(int(self.total_steps)))
            L.append(_line(bool(self.output_t0)))
        else:
            raise ValueError("output_style must be 1, 2 or 3")
        L.append(_line(OUTPUT_FORMATS[self.output_format]))
        qc = self.output_q_components or [1] * self.num_eqn
        L.append(_line(*[int(bool(c)) for c in qc]))
        if self.num_aux > 0:
            ac = self.output_aux_components or [0] * self.num_aux
            L.append(_line(*[int(bool(c)) for c in ac]))
            L.append(_line(bool(self.output_aux_onlyonce)))
        L.append(_line(float(self.dt_initial)))
        L.append(_line(float(self.dt_max)))
        L.append(_line(float(self.cfl_max)))
        L.append(_line(float(self.cfl_desired)))
        L.append(_line(int(self.steps_max)))
        L.append(_line(bool(self.dt_variable)))
        L.append(_line(int(self.order)))
        tw = self.transverse_waves
        L.append(_line(TRANSVERSE[tw] if isinstance(tw, str) else int(tw)))
        L.append(_line(int(self.dimensional_split)))
        L.append(_line(int(self.verbosity)))
        L.append(_line(int(self.source_split)))
        L.append(_line(int(self.capa_index)))
        L.append(_line(bool(self.use_fwaves)))
        lim = [LIMITERS[x] if isinstance(x, str) else int(x) for x in self.limiter]
        if len(lim) == 1 and self.num_waves > 1:
            lim = lim * self.num_waves
        L.append(_line(*lim))
        L.append(_line(int(self.num_ghost)))
        bc = [BCS[b] if isinstance(b, str) else int(b) for b in self.bc_lower + self.bc_upper]
        L.append(_line(bc[0], bc[1]))                   # mthbc(1), mthbc(3)
        L.append(_line(bc[2], bc[3]))                   # mthbc(2), mthbc(4)
        L.append(_line(bool(self.restart)))
        L.append(_line(self.restart_file))
        L.append(_line(int(self.checkpt_style)))
        if abs(self.checkpt_style) == 2:
            L.append(_line(len(self.checkpt_times)))
            L.append(_line(*[float(t) for t in self.checkpt_times]))
        elif abs(self.checkpt_style) == 3:
            L.append(_line(int(self.checkpt_interval)))
        return L

    def amr_lines(self):
        n = int(self.amr_levels_max)
        m = max(1, n - 1)

        def ratios(r):
            r = list(r) + [1] * m
            return [int(x) for x in r[:m]]
        L = [_line(n), _line(*ratios(self.refinement_ratios_x)),
             _line(*ratios(self.refinement_ratios_y)),
             _line(*ratios(self.refinement_ratios_t)), ""]
        if self.num_aux > 0:
            L.append(_line(*(self.aux_type or ["center"] * self.num_aux)))
        L.append("")
        L += [_line(bool(self.flag_richardson)), _line(float(self.flag_richardson_tol)),
              _line(bool(self.flag2refine)), _line(float(self.flag2refine_tol)),
              _line(int(self.regrid_interval)), _line(int(self.regrid_buffer_width)),
              _line(float(self.clustering_cutoff)), _line(int(self.verbosity_regrid))]
        for f in ("dprint", "eprint", "edebug", "gprint", "nprint", "pprint", "rprint",
                  "sprint", "tprint", "uprint"):
            L.append(_line(bool(getattr(self, f))))
        return L

    def regions_lines(self):
        L = ["%-2d" % len(self.regions)]                # read as (i2)
        for r in self.regions:
            L.append(_line(int(r[0]), int(r[1]), *[float(x) for x in r[2:8]]))
        return L

    def gauges_lines(self):
        n = len(self.gauges)
        L = [_line(n)]
        for g in self.gauges:
            L.append(_line(int(g[0]), *[float(x) for x in g[1:5]]))
        L += ["", "", _line(*([1] * n)), "", "", _line(*([self.gauge_display_format] * n)),
              "", "", _line(*([float(self.gauge_min_time_increment)] * n)), "", ""]
        for _ in range(n):
            L.append(_line(*([True] * self.num_eqn)))
        if self.num_aux > 0:
            L += ["", ""]
            for _ in range(n):
                L.append(_line(*([False] * self.num_aux)))
        return L

    def write(self, outdir="."):
        """Write claw.data, amr.data, regions.data, gauges.data and
        setprob.data into outdir."""
        files = {"claw.data": self.claw_lines(), "amr.data": self.amr_lines(),
                 "regions.data": self.regions_lines(), "gauges.data": self.gauges_lines(),
                 "setprob.data": [_line(v) for v in self.probdata]}
        for name, lines in files.items():
            with open(os.path.join(outdir, name), "w") as f:
                f.write("# %s written by amrclaw_amd.rundata\n\n" % name)
                f.write("\n".join(lines) + "\n")
        return sorted(files)
