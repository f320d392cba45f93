"""ctypes binding of libamrgpu.so (include/amrgpu.h).

This is the Python mirror of the binding a Fortran driver makes through
ISO_C_BINDING: the same entry points, the same argument meaning.  Loading
fails loudly if the HIP library has not been built; there is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libamrgpu.so")

ADVECTION, ACOUSTICS, EULER, VCADV = 1, 2, 3, 4
AUX_CENTER, AUX_XLEFT, AUX_YLEFT, AUX_CAPACITY = 0, 1, 2, 3
FP64, FP32, MIXED = 0, 1, 2


class AmrgpuConfig(C.Structure):
    _fields_ = [
        ("solver", C.c_int), ("nvar", C.c_int), ("mwaves", C.c_int),
        ("naux", C.c_int), ("nghost", C.c_int),
        ("method", C.c_int * 7), ("mthlim", C.c_int * 8),
        ("mcapa", C.c_int), ("use_fwaves", C.c_int),
        ("cfl", C.c_double), ("cflv1", C.c_double),
        ("mthbc", C.c_int * 4),
        ("xlower", C.c_double), ("xupper", C.c_double),
        ("ylower", C.c_double), ("yupper", C.c_double),
        ("ubar", C.c_double), ("vbar", C.c_double),
        ("rho", C.c_double), ("bulk", C.c_double),
        ("gamma", C.c_double), ("efix", C.c_int),
        ("precision", C.c_int), ("auxtype", C.c_int * 8),
    ]


class FillOp(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("type", "dst", "src", "src_level", "di0", "dj0", "ni",
                                       "nj", "si0", "sj0", "mask_off", "ilo", "jlo", "iplo",
                                       "jplo", "ratx", "raty", "level")] + \
               [("bc", C.c_int * 4), ("stage", C.c_int), ("flags", C.c_int), ("pad", C.c_int)]


class FillBuf(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("kind", "level", "patch", "nrow", "ncol", "plane", "pad")] + \
               [("off", C.c_int64)]


class AmrgpuSolverOps(C.Structure):
    """amrgpu_solver_ops: the launcher table of a plug-in Riemann solver."""
    _fields_ = [("abi", C.c_int), ("nvar", C.c_int), ("mwaves", C.c_int),
                ("sweep_args_bytes", C.c_int), ("qad_args_bytes", C.c_int),
                ("physics_bytes", C.c_int), ("name", C.c_char_p),
                ("sweep_f64", C.c_void_p), ("sweep_f32", C.c_void_p),
                ("sweep_mixed", C.c_void_p), ("qad", C.c_void_p)]


SOLVER_ABI = 1
SOLVER_PLUGIN0 = 100


class AmrgpuError(RuntimeError):
    pass


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    """Load libamrgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise AmrgpuError("libamrgpu.so is not built; run "
                              "`python -m amrclaw_amd.build`")
        L = C.CDLL(LIBPATH)
        dpo = C.POINTER(C.c_double)
        L.amrgpu_init.argtypes = [C.POINTER(AmrgpuConfig)]
        L.amrgpu_register_solver.argtypes = [C.POINTER(AmrgpuSolverOps), C.POINTER(C.c_int)]
        L.amrgpu_set_solver_params.argtypes = [C.c_int, C.c_void_p]
        L.amrgpu_finalize.argtypes = []
        L.amrgpu_last_error.argtypes = [C.c_char_p, C.c_int]
        L.amrgpu_set_tiling.argtypes = [C.c_int, C.c_int]
        L.amrgpu_set_occupancy.argtypes = [C.c_int]
        L.amrgpu_stepgrid.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double,
                                      C.c_double, C.c_double, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p,
                                      dpo, dpo]
        L.amrgpu_stepgrid_noalloc.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                              C.c_double, C.c_double, C.c_double, dpo, dpo]
        L.amrgpu_set_level.argtypes = [C.c_int, C.c_int, _ip, _ip, _ip, _ip,
                                       C.c_double, C.c_double]
        L.amrgpu_put_grid.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
        L.amrgpu_get_grid.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
        _lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.amrgpu_put_level.argtypes = [C.c_int, C.c_int, _dp, _lp]
        L.amrgpu_put_level_aux.argtypes = [C.c_int, _dp, _lp]
        L.amrgpu_get_level.argtypes = [C.c_int, C.c_int, _dp, _lp]
        L.amrgpu_sweep_level.argtypes = [C.c_int, C.c_double, C.c_void_p,
                                         C.c_void_p]
        L.amrgpu_set_geometry.argtypes = [C.c_int, _ip, _ip, _ip, _ip, _dp, _dp]
        L.amrgpu_set_time.argtypes = [C.c_int, C.c_double, C.c_double]
        L.amrgpu_bound_level.argtypes = [C.c_int, C.c_double]
        L.amrgpu_advance_level.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.amrgpu_update_level.argtypes = [C.c_int]
        L.amrgpu_conck.argtypes = [C.c_int, C.POINTER(C.c_double)]
        L.amrgpu_get_registers.argtypes = [C.c_int, C.c_int, _dp, C.POINTER(C.c_int)]
        L.amrgpu_set_corners.argtypes = [C.c_int, _dp, _dp]
        L.amrgpu_set_launch_mode.argtypes = [C.c_int]
        L.amrgpu_set_graphs.argtypes = [C.c_int]
        L.amrgpu_set_timing.argtypes = [C.c_int]
        L.amrgpu_kernel_time.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.amrgpu_flag_level.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int,
                                        C.c_double, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                        C.c_void_p, C.POINTER(C.c_int64)]
        L.amrgpu_event_record.argtypes = [C.c_int]
        L.amrgpu_event_elapsed.argtypes = [C.c_int, C.c_int,
                                           C.POINTER(C.c_float)]
        L.amrgpu_synchronize.argtypes = []
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        buf = C.create_string_buffer(512)
        lib().amrgpu_last_error(buf, 512)
        raise AmrgpuError("amrgpu error %d: %s" % (rc, buf.value.decode()))


def make_config(solver, *, order=2, transverse=2, limiter=4, cfl=0.9,
                cflv1=1.0, mthbc=(1, 1, 1, 1), domain=(0.0, 1.0, 0.0, 1.0),
                ubar=0.5, vbar=1.0, rho=1.0, bulk=4.0, gamma=1.4, efix=True,
                precision=FP64, nvar=None, mwaves=None, naux=0, mcapa=0, auxtype=()):
    """Configuration of one run (claw.data / amr.data fields, setprob
    constants).  limiter: one mthlim code for every wave, or a list.
    A plug-in solver (id from register_solver) needs nvar and mwaves.
    naux / mcapa / auxtype: num_aux, capa_index (1-based, 0 for none) and
    the AUX_* code of each aux component (amr.data aux_type)."""
    c = AmrgpuConfig()
    c.solver = solver
    if solver in (ADVECTION, ACOUSTICS, EULER, VCADV):
        c.nvar, c.mwaves = {ADVECTION: (1, 1), ACOUSTICS: (3, 2), EULER: (4, 4),
                            VCADV: (1, 1)}[solver]
    else:
        c.nvar, c.mwaves = nvar, mwaves
    c.naux, c.nghost = naux, 2
    for k in range(8):
        c.auxtype[k] = auxtype[k] if k < len(auxtype) else AUX_CENTER
    meth = [0, order, transverse, 0, 0, 0, 0]
    for k in range(7):
        c.method[k] = meth[k]
    lim = limiter if isinstance(limiter, (list, tuple)) else [limiter] * c.mwaves
    for k in range(8):
        c.mthlim[k] = lim[k] if k < len(lim) else 0
    c.mcapa, c.use_fwaves = mcapa, 0
    c.cfl, c.cflv1 = cfl, cflv1
    for k in range(4):
        c.mthbc[k] = mthbc[k]
    c.xlower, c.xupper, c.ylower, c.yupper = domain
    c.ubar, c.vbar, c.rho, c.bulk = ubar, vbar, rho, bulk
    c.gamma, c.efix = gamma, int(bool(efix))
    c.precision = precision
    return c


def load_solver_plugin(path, entry):
    """Load a plug-in solver library (built against include/amrgpu_solver.hpp)
    and fill its launcher table through its exported `entry` function.
    Returns (table, the loaded library); keep both alive while the solver
    is registered."""
    P = C.CDLL(path)
    fn = getattr(P, entry)
    fn.argtypes = [C.POINTER(AmrgpuSolverOps)]
    fn.restype = C.c_int
    ops = AmrgpuSolverOps()
    if fn(C.byref(ops)) != 0:
        raise AmrgpuError("%s(%s) failed" % (entry, path))
    return ops, P


def register_solver(ops):
    """amrgpu_register_solver: returns the solver id for make_config."""
    sid = C.c_int()
    _check(lib().amrgpu_register_solver(C.byref(ops), C.byref(sid)))
    return sid.value


def set_solver_params(params):
    """amrgpu_set_solver_params: the plug-in solver's constants (/cparam/)."""
    a = np.ascontiguousarray(params, dtype=np.float64)
    _check(lib().amrgpu_set_solver_params(len(a), a.ctypes.data_as(C.c_void_p) if len(a) else None))


def init(cfg):
    _check(lib().amrgpu_init(C.byref(cfg)))


def finalize():
    _check(lib().amrgpu_finalize())


def set_tiling(txmax, tymax):
    _check(lib().amrgpu_set_tiling(txmax, tymax))


def set_occupancy(wgs):
    """Workgroups per CU the mixed sweep is register-bounded for (5, 6, 7; 0 default)."""
    _check(lib().amrgpu_set_occupancy(wgs))


def stepgrid(q, dx, dy, dt, want_fluxes=False):
    """amrgpu_stepgrid on one patch.  q: (my+4, mx+4, nvar) array (the
    reference's q(nvar, mitot, mjtot) in Fortran order).  Returns
    (qnew, [fm, fp, gm, gp] or None, cflgrid, dtnew)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    my, mx = q.shape[0] - 4, q.shape[1] - 4
    out = np.empty_like(q)
    fl = [np.zeros_like(q) for _ in range(4)] if want_fluxes else None
    ptrs = [f.ctypes.data_as(C.c_void_p) for f in fl] if fl else [None] * 4
    cfl, dtn = C.c_double(), C.c_double()
    _check(lib().amrgpu_stepgrid(mx, my, q, out, dx, dy, dt, *ptrs,
                                 C.byref(cfl), C.byref(dtn)))
    return out, fl, cfl.value, dtn.value


def stepgrid_noalloc(mx, my, q_in_dev, q_out_dev, dx, dy, dt):
    """amrgpu_stepgrid_noalloc: q_in_dev / q_out_dev are device addresses
    of (my+4, mx+4, nvar) FP64 arrays (the reference's layout).  Returns
    (cflgrid, dtnew)."""
    cfl, dtn = C.c_double(), C.c_double()
    _check(lib().amrgpu_stepgrid_noalloc(mx, my, C.c_void_p(q_in_dev), C.c_void_p(q_out_dev),
                                         dx, dy, dt, C.byref(cfl), C.byref(dtn)))
    return cfl.value, dtn.value


def set_level(level, boxes, hx, hy):
    """boxes: (n, 4) int array of (ilo, ihi, jlo, jhi) per grid."""
    b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 4))
    cols = [np.ascontiguousarray(b[:, k]) for k in range(4)]
    _check(lib().amrgpu_set_level(level, b.shape[0], *cols, hx, hy))


def put_grid(level, grid, q, which=0):
    _check(lib().amrgpu_put_grid(level, grid, which,
                                 np.ascontiguousarray(q, dtype=np.float64)))


def get_grid(level, grid, shape, which=0):
    out = np.empty(shape, dtype=np.float64)
    _check(lib().amrgpu_get_grid(level, grid, which, out))
    return out


def put_level(level, qs, which=0):
    """Upload all grids of a level: qs is a list of (my+4, mx+4, nvar)
    arrays in set_level order (packed like the reference's alloc)."""
    sizes = np.array([q.size for q in qs], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    packed = np.concatenate([np.ascontiguousarray(q, dtype=np.float64).ravel() for q in qs])
    _check(lib().amrgpu_put_level(level, which, packed, offs))


def put_level_aux(level, aux, offsets):
    """amrgpu_put_level_aux: every grid's aux array (naux values per cell,
    component fastest, ghost cells included) at element offsets of aux."""
    _check(lib().amrgpu_put_level_aux(level, np.ascontiguousarray(aux, dtype=np.float64),
                                      np.ascontiguousarray(offsets, dtype=np.int64)))


def get_level(level, shapes, which=0):
    """Download all grids of a level as a list of arrays of the given shapes."""
    sizes = np.array([int(np.prod(s)) for s in shapes], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    packed = np.empty(int(sizes.sum()), dtype=np.float64)
    _check(lib().amrgpu_get_level(level, which, packed, offs))
    return [packed[o:o + n].reshape(s) for o, n, s in zip(offs, sizes, shapes)]


def sweep_level(level, dt, want_cfl=True):
    if want_cfl:
        cfl, dtn = C.c_double(), C.c_double()
        _check(lib().amrgpu_sweep_level(level, dt, C.byref(cfl), C.byref(dtn)))
        return cfl.value, dtn.value
    _check(lib().amrgpu_sweep_level(level, dt, None, None))
    return None


def set_geometry(iregsz, jregsz, intratx, intraty, hxposs, hyposs):
    """Per-level geometry for levels 1..len(iregsz) (mxnest = len)."""
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    _check(lib().amrgpu_set_geometry(len(iregsz), i32(iregsz), i32(jregsz), i32(intratx),
                                     i32(intraty), f64(hxposs), f64(hyposs)))


def set_time(level, time, possk):
    _check(lib().amrgpu_set_time(level, time, possk))


def bound_level(level, time):
    _check(lib().amrgpu_bound_level(level, time))


def advance_level(level, want_cfl=True):
    """advanc for one level; returns (dtlevnew, cfl_level), or None without
    waiting for the device when want_cfl is False."""
    if not want_cfl:
        _check(lib().amrgpu_advance_level(level, None, None))
        return None
    dt, cfl = C.c_double(), C.c_double()
    _check(lib().amrgpu_advance_level(level, C.byref(dt), C.byref(cfl)))
    return dt.value, cfl.value


LAUNCH_BATCHED, LAUNCH_PER_PATCH, LAUNCH_PERSISTENT, LAUNCH_MARCH = 0, 1, 2, 3
KT_SWEEP, KT_FILL, KT_REGISTERS, KT_UPDATE, KT_FLAG = range(5)


def set_launch_mode(mode):
    _check(lib().amrgpu_set_launch_mode(mode))


def set_graphs(on):
    _check(lib().amrgpu_set_graphs(int(bool(on))))


def set_timing(on):
    _check(lib().amrgpu_set_timing(int(bool(on))))


def kernel_time(kind):
    """(total ms, launches) of one kernel class since set_timing(True)."""
    ms, n = C.c_double(), C.c_int64()
    _check(lib().amrgpu_kernel_time(kind, C.byref(ms), C.byref(n)))
    return ms.value, n.value


def update_level(level):
    _check(lib().amrgpu_update_level(level))


def conck(level):
    t = C.c_double()
    _check(lib().amrgpu_conck(level, C.byref(t)))
    return t.value


def get_registers(level, grid, nvar, maxlen):
    """(svdflx, qc1d) of one fine grid, each (lenbc, nvar)."""
    buf = np.zeros(2 * nvar * maxlen)
    n = C.c_int()
    _check(lib().amrgpu_get_registers(level, grid, buf, C.byref(n)))
    lb = n.value
    return buf[:nvar * lb].reshape(lb, nvar), buf[nvar * lb:2 * nvar * lb].reshape(lb, nvar)


def set_corners(level, xlo, ylo):
    _check(lib().amrgpu_set_corners(level, np.ascontiguousarray(xlo, dtype=np.float64),
                                    np.ascontiguousarray(ylo, dtype=np.float64)))


def flag_level(level, tolsp, start_time, regions=(), mbuff=4, shapes=None, gradient=True,
               richardson=False, tol=0.0, iorder=2):
    """Flag the cells of every grid of a level.  With shapes (list of (mx,
    my)) the flags are returned as a list of (my+2 mbuff, mx+2 mbuff)
    arrays; returns (flags or None, number of DOFLAG cells)."""
    reg = np.ascontiguousarray(np.asarray(regions, dtype=np.float64).reshape(-1))
    n = C.c_int64()
    rp = reg.ctypes.data_as(C.c_void_p) if reg.size else None
    head = (level, int(bool(gradient)), tolsp, int(bool(richardson)), tol, iorder, start_time,
            len(regions), rp, mbuff)
    if shapes is None:
        _check(lib().amrgpu_flag_level(*head, None, None, C.byref(n)))
        return None, n.value
    sizes = [(mx + 2 * mbuff) * (my + 2 * mbuff) for mx, my in shapes]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    out = np.zeros(int(sum(sizes)))
    _check(lib().amrgpu_flag_level(*head, out.ctypes.data_as(C.c_void_p),
                                   offs.ctypes.data_as(C.c_void_p), C.byref(n)))
    flags = [out[o:o + sz].reshape(my + 2 * mbuff, mx + 2 * mbuff)
             for o, sz, (mx, my) in zip(offs, sizes, shapes)]
    return flags, n.value


def load_hierarchy(H):
    """Install a tests.common.Hier-like hierarchy: geometry, layouts, both
    time levels of every grid and the level times."""
    lv = H.levels
    set_geometry([L["iregsz"] for L in lv], [L["jregsz"] for L in lv],
                 [L["ratx"] for L in lv], [L["raty"] for L in lv],
                 [L["hx"] for L in lv], [L["hy"] for L in lv])
    for l, L in enumerate(lv, start=1):
        set_level(l, L["boxes"], L["hx"], L["hy"])
        if len(L["boxes"]):
            _check(lib().amrgpu_put_level(l, 0, L["qnew"], L["off"]))
            _check(lib().amrgpu_put_level(l, 1, L["qold"], L["off"]))
            if getattr(H, "naux", 0) > 0:
                put_level_aux(l, L["aux"], L["auxoff"])
        set_time(l, L["time"], L["possk"])


def read_level(H, l, which=0):
    """Download level l into a packed array laid out like H.levels[l-1]."""
    L = H.levels[l - 1]
    out = np.zeros_like(L["qnew"])
    _check(lib().amrgpu_get_level(l, which, out, L["off"]))
    return out


def plan_fill(H, level, domain=None, mthbc=None):
    """Plan the ghost fill of `level` of a Hier-like hierarchy on the host
    (no device).  Returns (ops, bufs, masks) as ctypes arrays / bytes."""
    lv = H.levels
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    boxes = i32(np.concatenate([np.asarray(L["boxes"]).ravel() for L in lv]))
    nops, nbufs = C.c_int(), C.c_int()
    nmask, nscr = C.c_int64(), C.c_int64()
    L = lib()
    L.amrgpu_plan_fill.argtypes = [C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _ip,
                                   C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    _check(L.amrgpu_plan_fill(len(lv), i32([len(x["boxes"]) for x in lv]), boxes,
                              i32([x["iregsz"] for x in lv]), i32([x["jregsz"] for x in lv]),
                              i32([x["ratx"] for x in lv]), i32([x["raty"] for x in lv]),
                              f64([x["hx"] for x in lv]), f64([x["hy"] for x in lv]),
                              f64(domain if domain is not None else H.domain),
                              i32(mthbc if mthbc is not None else H.mthbc), H.nvar, level,
                              C.byref(nops), C.byref(nbufs), C.byref(nmask), C.byref(nscr)))
    ops = (FillOp * max(nops.value, 1))()
    bufs = (FillBuf * max(nbufs.value, 1))()
    masks = (C.c_ubyte * max(nmask.value, 1))()
    L.amrgpu_plan_fill_get.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    _check(L.amrgpu_plan_fill_get(C.cast(ops, C.c_void_p), C.cast(bufs, C.c_void_p),
                                  C.cast(masks, C.c_void_p)))
    return list(ops)[:nops.value], list(bufs)[:nbufs.value], bytes(masks)[:nmask.value]


def regrid_level(level, boxes, time):
    """amrgpu_regrid_level: replace the grids of `level` by boxes ((n, 4)
    ilo, ihi, jlo, jhi) filled on the device as filval fills new grids."""
    b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 4))
    cols = [np.ascontiguousarray(b[:, k]) for k in range(4)]
    L = lib()
    L.amrgpu_regrid_level.argtypes = [C.c_int, C.c_int, _ip, _ip, _ip, _ip, C.c_double]
    _check(L.amrgpu_regrid_level(level, len(b), *cols, time))


def bndry_lists(boxes, iregsz, jregsz):
    """amrgpu_bndry_lists (host only): every grid's neighbour list as
    makeBndryList builds it, as a list of lists of 0-based grid indices."""
    b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 4))
    n = len(b)
    counts = np.zeros(max(n, 1), dtype=np.int32)
    total = C.c_int64()
    L = lib()
    L.amrgpu_bndry_lists.argtypes = [C.c_int, _ip, C.c_int, C.c_int, _ip, C.c_void_p, C.c_int64,
                                     C.POINTER(C.c_int64)]
    L.amrgpu_bndry_lists(n, b, iregsz, jregsz, counts, None, 0, C.byref(total))
    items = np.zeros(max(total.value, 1), dtype=np.int32)
    _check(L.amrgpu_bndry_lists(n, b, iregsz, jregsz, counts, items.ctypes.data_as(C.c_void_p),
                                total.value, C.byref(total)))
    out, pos = [], 0
    for k in range(n):
        out.append([int(x) for x in items[pos:pos + counts[k]]])
        pos += counts[k]
    return out


def plan_listbc(H, level, grid):
    """Host-planned listbc entries of one coarse grid (no device)."""
    lv = H.levels
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    boxes = i32(np.concatenate([np.asarray(L["boxes"]).ravel() for L in lv]))
    L = lib()
    L.amrgpu_plan_listbc.argtypes = [C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip, C.c_int, C.c_int,
                                     C.c_int, _ip, C.c_int, C.POINTER(C.c_int)]
    out = np.zeros(5 * 4096, dtype=np.int32)
    n = C.c_int()
    _check(L.amrgpu_plan_listbc(len(lv), i32([len(x["boxes"]) for x in lv]), boxes,
                                i32([x["iregsz"] for x in lv]), i32([x["jregsz"] for x in lv]),
                                i32([x["ratx"] for x in lv]), i32([x["raty"] for x in lv]),
                                i32(H.mthbc), H.nvar, level, grid, out, 4096, C.byref(n)))
    return out[:5 * n.value].reshape(-1, 5)


def event_record(slot):
    _check(lib().amrgpu_event_record(slot))


def event_elapsed(a, b):
    ms = C.c_float()
    _check(lib().amrgpu_event_elapsed(a, b, C.byref(ms)))
    return ms.value


def synchronize():
    _check(lib().amrgpu_synchronize())
