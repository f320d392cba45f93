"""ctypes front end of the C oracle (liborc.so) and of the reference build
(_ref/libamrref.so).  TEST INFRASTRUCTURE ONLY.

Arrays follow the reference layout: q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc) in
Fortran order, i.e. numpy arrays of shape (my+2mbc, mx+2mbc, meqn) in C order.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ADVECTION, ACOUSTICS, EULER, SHALLOW, VCADV = 1, 2, 3, 4, 5


class OrcParams(C.Structure):
    _fields_ = [
        ("solver", C.c_int), ("meqn", C.c_int), ("mwaves", C.c_int),
        ("maux", C.c_int), ("mbc", C.c_int),
        ("method", C.c_int * 8), ("mthlim", C.c_int * 10),
        ("mcapa", C.c_int), ("use_fwaves", C.c_int),
        ("cfl", C.c_double), ("cflv1", C.c_double),
        ("ubar", C.c_double), ("vbar", C.c_double),
        ("rho", C.c_double), ("bulk", C.c_double), ("cc", C.c_double),
        ("zz", C.c_double), ("gamma", C.c_double), ("efix", C.c_int),
        ("wbar", C.c_double), ("grav", C.c_double),
        ("auxtype", C.c_int * 8),
    ]


AUX_CENTER, AUX_XLEFT, AUX_YLEFT, AUX_CAPACITY = 0, 1, 2, 3


def make_params(solver, *, order=2, transverse=2, limiter=4, mwaves=None,
                maux=0, mcapa=0, cfl=0.9, cflv1=1.0, ubar=0.5, vbar=1.0,
                rho=1.0, bulk=4.0, gamma=1.4, efix=True, mbc=2, grav=1.0, auxtype=()):
    """Build the params of one configuration.  limiter: mthlim code applied
    to every wave (0 none, 1 minmod, 2 superbee, 3 van Leer, 4 MC, 5 BW)."""
    p = OrcParams()
    p.solver = solver
    p.meqn, nw = {ADVECTION: (1, 1), ACOUSTICS: (3, 2), EULER: (4, 4), SHALLOW: (3, 3),
                  VCADV: (1, 1)}[solver]
    p.mwaves = nw if mwaves is None else mwaves
    p.maux, p.mbc = maux, mbc
    for k in range(8):
        p.method[k] = 0
    p.method[2], p.method[3] = order, transverse
    lim = limiter if isinstance(limiter, (list, tuple)) else [limiter] * p.mwaves
    for k in range(10):
        p.mthlim[k] = lim[k] if k < len(lim) else 0
    p.mcapa, p.use_fwaves = mcapa, 0
    p.cfl, p.cflv1 = cfl, cflv1
    p.ubar, p.vbar = ubar, vbar
    p.rho, p.bulk = rho, bulk
    p.cc = float(np.sqrt(bulk / rho))
    p.zz = rho * p.cc
    p.gamma, p.efix = gamma, int(bool(efix))
    p.grav = grav
    for k in range(8):
        p.auxtype[k] = auxtype[k] if k < len(auxtype) else AUX_CENTER
    return p


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = C.CDLL(path)
        P = C.POINTER(OrcParams)
        L.orc_step2.argtypes = [P, C.c_int, C.c_int, _dp, _dp, C.c_double,
                                C.c_double, C.c_double, C.POINTER(C.c_double),
                                _dp, _dp, _dp, _dp]
        L.orc_stepgrid.argtypes = [P, C.c_int, C.c_int, _dp, _dp, C.c_double,
                                   C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                                   C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
        L.orc_rpn2.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                               _dp, _dp]
        L.orc_rpt2.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                               _dp, _dp, _dp, _dp]
        L.orc_flux.argtypes = [P, C.c_int, _dp, _dp]
        L.orc_set_plugin.argtypes = [P]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libamrref.so"))


def ref():
    """The reference's own routines (built by `make -C oracle ref`)."""
    global _ref
    if _ref is None:
        lib()  # libamrref links liborc; load it first
        R = C.CDLL(os.path.join(HERE, "_ref", "libamrref.so"))
        R.ref_set_method.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.c_int, C.c_int, C.c_int, C.c_double,
                                     C.c_double, C.c_int]
        R.ref_step2.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                _dp, _dp, C.c_double, C.c_double, C.c_double,
                                C.POINTER(C.c_double), _dp, _dp, _dp, _dp]
        R.ref_stepgrid.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   _dp, _dp, C.c_double, C.c_double,
                                   C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                                   C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
        R.ref_stepgrid_batch.argtypes = [
            C.c_int, C.c_int, C.c_int,
            np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"),
            np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"),
            np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
            _dp, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
        _ref = R
    return _ref


def _aux_or_dummy(p, aux, mx, my):
    if aux is None:
        return np.zeros((my + 2 * p.mbc, mx + 2 * p.mbc, max(p.maux, 1)))
    return np.ascontiguousarray(aux, dtype=np.float64)


def step2(p, q, dx, dy, dt, aux=None):
    """orc_step2 on one patch; q of shape (my+2mbc, mx+2mbc, meqn)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    my, mx = q.shape[0] - 2 * p.mbc, q.shape[1] - 2 * p.mbc
    aux = _aux_or_dummy(p, aux, mx, my)
    f = [np.zeros_like(q) for _ in range(4)]
    cfl = C.c_double()
    lib().orc_step2(C.byref(p), mx, my, q, aux, dx, dy, dt, C.byref(cfl), *f)
    return f, cfl.value


def stepgrid(p, q, dx, dy, dt, aux=None):
    """orc_stepgrid on one patch: returns (qnew, fluxes, cflgrid, dtnew)."""
    qn = np.array(q, dtype=np.float64, order="C", copy=True)
    my, mx = qn.shape[0] - 2 * p.mbc, qn.shape[1] - 2 * p.mbc
    aux = _aux_or_dummy(p, aux, mx, my)
    f = [np.zeros_like(qn) for _ in range(4)]
    cfl, dtn = C.c_double(), C.c_double()
    lib().orc_stepgrid(C.byref(p), mx, my, qn, aux, dx, dy, dt, *f,
                       C.byref(cfl), C.byref(dtn))
    return qn, f, cfl.value, dtn.value


def _ref_setup(p):
    R = ref()
    lib().orc_set_plugin(C.byref(p))
    meth = (C.c_int * 7)(*[p.method[k] for k in range(1, 8)])
    lim = (C.c_int * 10)(*list(p.mthlim))
    R.ref_set_method(meth, lim, p.mwaves, p.mcapa, p.use_fwaves, p.cfl,
                     p.cflv1, p.mbc)
    return R


def ref_step2(p, q, dx, dy, dt, aux=None):
    """The reference's step2.f90 on one patch (same conventions as step2)."""
    R = _ref_setup(p)
    q = np.ascontiguousarray(q, dtype=np.float64)
    my, mx = q.shape[0] - 2 * p.mbc, q.shape[1] - 2 * p.mbc
    aux = _aux_or_dummy(p, aux, mx, my)
    f = [np.zeros_like(q) for _ in range(4)]
    cfl = C.c_double()
    R.ref_step2(p.meqn, p.maux, p.mbc, mx, my, q, aux, dx, dy, dt,
                C.byref(cfl), *f)
    return f, cfl.value


def ref_stepgrid(p, q, dx, dy, dt, aux=None, time=0.0):
    """The reference's stepgrid.f on one patch."""
    R = _ref_setup(p)
    qn = np.array(q, dtype=np.float64, order="C", copy=True)
    my, mx = qn.shape[0] - 2 * p.mbc, qn.shape[1] - 2 * p.mbc
    aux = _aux_or_dummy(p, aux, mx, my)
    f = [np.zeros_like(qn) for _ in range(4)]
    cfl, dtn = C.c_double(), C.c_double()
    R.ref_stepgrid(p.meqn, p.maux, p.mbc, mx, my, qn, aux, dx, dy, dt, time,
                   *f, C.byref(dtn), C.byref(cfl))
    return qn, f, cfl.value, dtn.value


def ref_stepgrid_batch(p, qs, dx, dy, dt):
    """The reference's stepgrid over a list of patches with its OpenMP patch
    loop (advanc.f:106-129); patches are updated in place in a packed copy,
    which is returned with the level's max Courant number."""
    R = _ref_setup(p)
    mys = np.array([q.shape[0] - 2 * p.mbc for q in qs], dtype=np.int32)
    mxs = np.array([q.shape[1] - 2 * p.mbc for q in qs], dtype=np.int32)
    sizes = np.array([q.size for q in qs], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    packed = np.concatenate([np.ascontiguousarray(q).ravel() for q in qs])
    cfl = C.c_double()
    R.ref_stepgrid_batch(len(qs), p.meqn, p.mbc, mxs, mys, offs, packed, dx,
                         dy, dt, C.byref(cfl))
    return packed, offs, cfl.value


def rpn2(p, ixy, ql, qr):
    """Solve the Riemann problems between consecutive slice cells.
    ql, qr: (n, meqn) slices (cell index 1-mbc.. as rows); returns
    wave (n, mwaves, meqn), s (n, mwaves), amdq, apdq (n, meqn)."""
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    n = ql.shape[0]
    mx = n - 2 * p.mbc
    wave = np.zeros((n, p.mwaves, p.meqn))
    s = np.zeros((n, p.mwaves))
    amdq = np.zeros((n, p.meqn))
    apdq = np.zeros((n, p.meqn))
    aux = np.zeros((n, max(p.maux, 1)))
    lib().orc_rpn2(C.byref(p), ixy, mx, p.meqn, p.mwaves, p.maux, p.mbc, mx,
                   ql, qr, aux, aux, wave, s, amdq, apdq)
    return wave, s, amdq, apdq


def rpt2(p, ixy, imp, ql, qr, asdq):
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    asdq = np.ascontiguousarray(asdq, dtype=np.float64)
    n = ql.shape[0]
    mx = n - 2 * p.mbc
    bm = np.zeros((n, p.meqn))
    bp = np.zeros((n, p.meqn))
    aux = np.zeros((n, max(p.maux, 1)))
    lib().orc_rpt2(C.byref(p), ixy, imp, mx, p.meqn, p.mwaves, p.maux, p.mbc,
                   mx, ql, qr, aux, aux, aux, asdq, bm, bp)
    return bm, bp


def flux(p, ixy, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    f = np.zeros(p.meqn)
    lib().orc_flux(C.byref(p), ixy, q, f)
    return f


def _slice3(p, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    n = q.shape[0]
    return q, n, n - 2 * p.mbc


def rpn3(p, ixyz, ql, qr):
    """orc_rpn3 over consecutive slice cells (rows = cells 1-mbc..): returns
    wave (n, mwaves, meqn), s (n, mwaves), amdq, apdq (n, meqn); row i holds
    the interface between rows i-1 and i."""
    ql, n, mx = _slice3(p, ql)
    qr, _, _ = _slice3(p, qr)
    wave = np.zeros((n, p.mwaves, p.meqn))
    s = np.zeros((n, p.mwaves))
    amdq = np.zeros((n, p.meqn))
    apdq = np.zeros((n, p.meqn))
    L = lib()
    L.orc_rpn3.argtypes = [C.POINTER(OrcParams)] + [C.c_int] * 6 + [_dp] * 6
    L.orc_rpn3(C.byref(p), ixyz, mx, p.meqn, p.mwaves, p.mbc, mx, ql, qr, wave, s, amdq, apdq)
    return wave, s, amdq, apdq


def rpt3(p, ixyz, icoor, ql, qr, asdq, double=False):
    """orc_rpt3 (or orc_rptt3 with double=True): split asdq along direction
    icoor (2: the sweep's mv, 3: its mw) into (bm, bp), rows as in rpn3."""
    ql, n, mx = _slice3(p, ql)
    qr, _, _ = _slice3(p, qr)
    asdq = np.ascontiguousarray(asdq, dtype=np.float64)
    bm = np.zeros((n, p.meqn))
    bp = np.zeros((n, p.meqn))
    L = lib()
    if double:
        L.orc_rptt3.argtypes = [C.POINTER(OrcParams)] + [C.c_int] * 9 + [_dp] * 5
        L.orc_rptt3(C.byref(p), ixyz, icoor, 1, 1, mx, p.meqn, p.mwaves, p.mbc, mx, ql, qr,
                    asdq, bm, bp)
    else:
        L.orc_rpt3.argtypes = [C.POINTER(OrcParams)] + [C.c_int] * 8 + [_dp] * 5
        L.orc_rpt3(C.byref(p), ixyz, icoor, 1, mx, p.meqn, p.mwaves, p.mbc, mx, ql, qr, asdq,
                   bm, bp)
    return bm, bp


# ------------------------------------------------------------ hierarchy ---
MAXLV = 12


class OrcHier(C.Structure):
    _fields_ = [
        ("nlev", C.c_int), ("mxnest", C.c_int),
        ("ngrids", C.c_int * MAXLV),
        ("boxes", C.POINTER(C.c_int) * MAXLV),
        ("qnew", C.POINTER(C.c_double) * MAXLV),
        ("qold", C.POINTER(C.c_double) * MAXLV),
        ("qoff", C.POINTER(C.c_int64) * MAXLV),
        ("hx", C.c_double * MAXLV), ("hy", C.c_double * MAXLV),
        ("possk", C.c_double * MAXLV), ("time", C.c_double * MAXLV),
        ("iregsz", C.c_int * MAXLV), ("jregsz", C.c_int * MAXLV),
        ("intratx", C.c_int * MAXLV), ("intraty", C.c_int * MAXLV),
        ("xlower", C.c_double), ("xupper", C.c_double),
        ("ylower", C.c_double), ("yupper", C.c_double),
        ("xperdom", C.c_int), ("yperdom", C.c_int), ("spheredom", C.c_int),
        ("mthbc", C.c_int * 4),
        ("aux", C.POINTER(C.c_double) * MAXLV),
        ("auxoff", C.POINTER(C.c_int64) * MAXLV),
    ]


def hier_struct(H):
    """OrcHier view of a tests.common.Hier (shares its arrays)."""
    h = OrcHier()
    h.nlev = h.mxnest = H.nlev
    for l, L in enumerate(H.levels, start=1):
        h.ngrids[l] = len(L["boxes"])
        h.boxes[l] = L["boxes"].ctypes.data_as(C.POINTER(C.c_int))
        h.qnew[l] = L["qnew"].ctypes.data_as(C.POINTER(C.c_double))
        h.qold[l] = L["qold"].ctypes.data_as(C.POINTER(C.c_double))
        h.qoff[l] = L["off"].ctypes.data_as(C.POINTER(C.c_int64))
        h.hx[l], h.hy[l] = L["hx"], L["hy"]
        h.possk[l], h.time[l] = L["possk"], L["time"]
        h.iregsz[l], h.jregsz[l] = L["iregsz"], L["jregsz"]
        h.intratx[l], h.intraty[l] = L["ratx"], L["raty"]
    h.xlower, h.xupper, h.ylower, h.yupper = H.domain
    h.xperdom, h.yperdom, h.spheredom = int(H.xper), int(H.yper), 0
    for k in range(4):
        h.mthbc[k] = H.mthbc[k]
    if getattr(H, "naux", 0) > 0:
        for l, L in enumerate(H.levels, start=1):
            h.aux[l] = L["aux"].ctypes.data_as(C.POINTER(C.c_double))
            h.auxoff[l] = L["auxoff"].ctypes.data_as(C.POINTER(C.c_int64))
    return h


def bound(p, H, level, grid, time):
    """orc_bound on grid `grid` (0-based) of `level`, in place on H."""
    L = lib()
    L.orc_bound.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcHier), C.c_int,
                            C.c_int, C.c_double, C.c_void_p]
    h = hier_struct(H)
    val = H.grid(level, grid, 0)
    rc = L.orc_bound(C.byref(p), C.byref(h), level, grid, time,
                     val.ctypes.data_as(C.c_void_p))
    if rc:
        raise RuntimeError("orc_bound error %d" % rc)


def ref_load_hier(p, H):
    """Install the hierarchy in the reference's amr_module (harness)."""
    R = _ref_setup(p)
    ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    R.ref_hier_setup.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, ip, ip, dp, dp,
                                 dp, dp, ip, ip, ip, ip, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_int, C.c_int, ip]
    R.ref_hier_xfer.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int]
    lv = H.levels
    f64 = lambda k: np.array([L[k] for L in lv], dtype=np.float64)
    i32 = lambda k: np.array([L[k] for L in lv], dtype=np.int32)
    boxes = np.ascontiguousarray(np.concatenate([L["boxes"].ravel() for L in lv]).astype(np.int32))
    ngr = np.array([len(L["boxes"]) for L in lv], dtype=np.int32)
    R.ref_hier_setup(H.nlev, H.nlev, H.nvar, H.ng, ngr, boxes, f64("hx"), f64("hy"),
                     f64("possk"), f64("time"), i32("iregsz"), i32("jregsz"),
                     i32("ratx"), i32("raty"), *H.domain, int(H.xper), int(H.yper),
                     np.array(H.mthbc, dtype=np.int32))
    for l in range(1, H.nlev + 1):
        for k in range(len(lv[l - 1]["boxes"])):
            for which in (0, 1):
                g = np.ascontiguousarray(H.grid(l, k, which))
                R.ref_hier_xfer(l, k, which, H.nvar, g.ctypes.data_as(C.c_void_p), 1)
    naux = getattr(H, "naux", 0)
    R.ref_hier_set_aux.argtypes = [C.c_int, C.POINTER(C.c_int)]
    R.ref_hier_set_aux(naux, (C.c_int * 8)(*list(p.auxtype)))
    if naux > 0:
        R.ref_hier_put_aux.argtypes = [C.c_int, C.c_int, C.c_void_p]
        for l in range(1, H.nlev + 1):
            for k in range(len(lv[l - 1]["boxes"])):
                a = np.ascontiguousarray(H.aux(l, k))
                R.ref_hier_put_aux(l, k, a.ctypes.data_as(C.c_void_p))
    return R


def ref_get_grid(H, level, grid, which=0):
    R = ref()
    out = np.zeros(H.levels[level - 1]["shapes"][grid])
    R.ref_hier_xfer(level, grid, which, H.nvar, out.ctypes.data_as(C.c_void_p), 0)
    return out


def ref_bound(p, H, level, grid, time):
    """The reference's bound.f90 on a freshly installed copy of H; returns
    the grid's new array after the call."""
    R = ref_load_hier(p, H)
    R.ref_bound.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int]
    R.ref_bound(level, grid, time, H.nvar)
    return ref_get_grid(H, level, grid)


def filval(p, H, level, box, time, val):
    """orc_filval: the new grid with box (ilo, ihi, jlo, jhi) of `level`
    (H's grids of `level` are the old ones), into val (my+4, mx+4, nvar),
    updated in place and returned."""
    L = lib()
    L.orc_filval.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcHier), C.c_int,
                             C.POINTER(C.c_int), C.c_double, C.c_void_p]
    h = hier_struct(H)
    b = (C.c_int * 4)(*[int(x) for x in box])
    assert val.flags.c_contiguous and val.dtype == np.float64
    rc = L.orc_filval(C.byref(p), C.byref(h), level, b, time, val.ctypes.data_as(C.c_void_p))
    if rc:
        raise RuntimeError("orc_filval error %d" % rc)
    return val


def ref_filval(p, H, level, box, time, val):
    """The reference's filval.f90 on a freshly installed copy of H."""
    R = ref_load_hier(p, H)
    R.ref_filval.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_double, C.c_int, C.c_void_p]
    b = (C.c_int * 4)(*[int(x) for x in box])
    assert val.flags.c_contiguous and val.dtype == np.float64
    R.ref_filval(level, b, time, H.nvar, val.ctypes.data_as(C.c_void_p))
    return val


def bndry_list(p, H, level, k):
    """orc_bndry_list: grid k's neighbours (0-based) in makeBndryList's order."""
    L = lib()
    L.orc_bndry_list.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcHier), C.c_int, C.c_int,
                                 C.POINTER(C.c_int)]
    h = hier_struct(H)
    out = (C.c_int * max(len(H.levels[level - 1]["boxes"]), 1))()
    n = L.orc_bndry_list(C.byref(p), C.byref(h), level, k, out)
    return list(out)[:n]


def ref_bndry_lists(p, H, level):
    """Every grid's neighbour list as the reference's makeBndryList built it
    (on a freshly installed copy of H)."""
    R = ref_load_hier(p, H)
    R.ref_bndry_list.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int)]
    R.ref_bndry_list.restype = C.c_int
    n = len(H.levels[level - 1]["boxes"])
    lists = []
    for k in range(n):
        out = (C.c_int * max(n, 1))()
        c = R.ref_bndry_list(level, k, out)
        lists.append(list(out)[:c])
    return lists


# ------------------------------------------------------------ registers ---
class OrcRegs(C.Structure):
    _fields_ = [
        ("nlev", C.c_int),
        ("lenbc", C.POINTER(C.c_int) * MAXLV),
        ("roff", C.POINTER(C.c_int64) * MAXLV),
        ("reg", C.POINTER(C.c_double) * MAXLV),
        ("nreg", C.c_int64 * MAXLV),
        ("lst_start", C.POINTER(C.c_int) * MAXLV),
        ("lst_count", C.POINTER(C.c_int) * MAXLV),
        ("listbc", C.POINTER(C.c_int) * MAXLV),
        ("nlist", C.c_int * MAXLV),
    ]


class OracleAMR:
    """The oracle's level update over a hierarchy (advanc / update / conck),
    working in place on the arrays of a tests.common.Hier."""

    def __init__(self, p, H):
        L = lib()
        P, Hh, Rr = C.POINTER(OrcParams), C.POINTER(OrcHier), C.POINTER(OrcRegs)
        L.orc_prep_regs.argtypes = [P, Hh, Rr]
        L.orc_free_regs.argtypes = [Rr]
        L.orc_advanc.argtypes = [P, Hh, Rr, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_update.argtypes = [P, Hh, Rr, C.c_int]
        L.orc_conck.argtypes = [P, Hh, C.c_int]
        L.orc_conck.restype = C.c_double
        self.p, self.H = p, H
        self.h = hier_struct(H)
        self.r = OrcRegs()
        rc = L.orc_prep_regs(C.byref(p), C.byref(self.h), C.byref(self.r))
        if rc:
            raise RuntimeError("orc_prep_regs error %d" % rc)

    def __del__(self):
        try:
            lib().orc_free_regs(C.byref(self.r))
        except Exception:
            pass

    def advanc(self, level):
        dtn, cfl = C.c_double(), C.c_double()
        rc = lib().orc_advanc(C.byref(self.p), C.byref(self.h), C.byref(self.r), level,
                              C.byref(dtn), C.byref(cfl))
        if rc:
            raise RuntimeError("orc_advanc error %d" % rc)
        self.H.levels[level - 1]["time"] = self.h.time[level]
        return dtn.value, cfl.value

    def update(self, level):
        rc = lib().orc_update(C.byref(self.p), C.byref(self.h), C.byref(self.r), level)
        if rc:
            raise RuntimeError("orc_update error %d" % rc)

    def conck(self, level):
        return lib().orc_conck(C.byref(self.p), C.byref(self.h), level)

    def flag_level(self, level, tolsp, start_time, regions=(), ibuff=3, gradient=True,
                   richardson=False, tol=0.0, iorder=2):
        """flagger.f for every grid of `level`: ghost fill of the new values
        (unless start_time + possk equals the level's time: old values),
        regions, gradient and/or Richardson flags.  Returns one flag array
        (ny + 2 mbuff, nx + 2 mbuff) per grid."""
        L = lib()
        L.orc_flag_level.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcHier), C.c_int, C.c_double,
                                     _dp, C.c_int, C.c_int, C.c_int, _dp,
                                     np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                     C.c_int, C.c_int, C.c_double, C.c_int]
        Lv = self.H.levels[level - 1]
        t = self.h.time[level]
        use_old = not (start_time + Lv["possk"] != t)
        if not use_old:
            L.orc_bound.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcHier), C.c_int,
                                    C.c_int, C.c_double, C.c_void_p]
            for k in range(len(Lv["boxes"])):
                val = self.H.grid(level, k, 0)
                rc = L.orc_bound(C.byref(self.p), C.byref(self.h), level, k, t,
                                 val.ctypes.data_as(C.c_void_p))
                if rc:
                    raise RuntimeError("orc_bound error %d" % rc)
        mbuff = max(self.p.mbc, ibuff + 1)
        shapes = [(int(b[3] - b[2]) + 1 + 2 * mbuff, int(b[1] - b[0]) + 1 + 2 * mbuff)
                  for b in Lv["boxes"]]
        sizes = [a * b for a, b in shapes]
        foff = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        out = np.zeros(max(int(sum(sizes)), 1))
        reg = np.ascontiguousarray(np.asarray(regions, dtype=np.float64).reshape(-1))
        if reg.size == 0:
            reg = np.zeros(8)
        rc = L.orc_flag_level(C.byref(self.p), C.byref(self.h), level, tolsp, reg, len(regions),
                              int(use_old), mbuff, out, foff, int(gradient), int(richardson),
                              tol, iorder)
        if rc:
            raise RuntimeError("orc_flag_level error %d" % rc)
        return [out[o:o + n].reshape(sh) for o, n, sh in zip(foff, sizes, shapes)]

    def reg(self, level, k):
        n = 2 * self.p.meqn * self.r.lenbc[level][k]
        off = self.r.roff[level][k]
        return np.ctypeslib.as_array(self.r.reg[level], shape=(self.r.nreg[level],))[off:off + n].copy()

    def listbc(self, level, k):
        if not self.r.listbc[level]:
            return np.zeros((0, 5), dtype=np.int32)
        a = np.ctypeslib.as_array(self.r.listbc[level], shape=(5 * self.r.nlist[level],))
        s, c = self.r.lst_start[level][k], self.r.lst_count[level][k]
        return a[5 * s:5 * (s + c)].reshape(-1, 5).copy()


class RefAMR:
    """The reference's own routines over a copy of a hierarchy installed in
    amr_module (prepf/prepc, bound, saveqc, qad, stepgrid, fluxsv, fluxad,
    update)."""

    def __init__(self, p, H):
        self.p, self.H = p, H
        R = ref_load_hier(p, H)
        R.ref_prep_regs.argtypes = [C.c_int, C.c_int]
        R.ref_advanc.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        R.ref_update.argtypes = [C.c_int, C.c_int]
        R.ref_get_reg.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p]
        R.ref_get_listbc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_int)]
        R.ref_get_time.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double)]
        R.ref_prep_regs(H.nlev, H.nvar)
        self.R = R

    def advanc(self, level):
        dtn, cfl = C.c_double(), C.c_double()
        self.R.ref_advanc(level, self.H.nvar, C.byref(dtn), C.byref(cfl))
        return dtn.value, cfl.value

    def flag_level(self, level, tolsp, start_time, regions=(), ibuff=3, gradient=True,
                   richardson=False, tol=0.0, iorder=2):
        """The reference's flagger.f on the installed hierarchy."""
        R = self.R
        R.ref_flag_level.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
        Lv = self.H.levels[level - 1]
        mbuff = max(2, ibuff + 1)
        shapes = [(int(b[3] - b[2]) + 1 + 2 * mbuff, int(b[1] - b[0]) + 1 + 2 * mbuff)
                  for b in Lv["boxes"]]
        sizes = [a * b for a, b in shapes]
        out = np.zeros(max(int(sum(sizes)), 1))
        reg = np.ascontiguousarray(np.asarray(regions, dtype=np.float64).reshape(-1))
        if reg.size == 0:
            reg = np.zeros(8)
        R.ref_flag_level(level, self.H.nvar, start_time, tolsp, tol, iorder, ibuff, int(gradient),
                         int(richardson), len(regions), reg, out)
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(int)
        return [out[o:o + n].reshape(sh) for o, n, sh in zip(offs, sizes, shapes)]

    def set_parallel(self, on):
        """OpenMP over the grids of a level, as the reference's advanc."""
        self.R.ref_set_parallel.argtypes = [C.c_int]
        self.R.ref_set_parallel(int(bool(on)))

    def set_time(self, level, t):
        self.R.ref_set_time.argtypes = [C.c_int, C.c_double]
        self.R.ref_set_time(level, t)

    def flush(self):
        self.R.ref_flush()

    def update(self, level):
        self.R.ref_update(level, self.H.nvar)

    def grid(self, level, k, which=0):
        return ref_get_grid(self.H, level, k, which)

    def reg(self, level, k):
        b = self.H.levels[level - 1]["boxes"][k]
        r = self.H.levels[level - 2]["ratx"]
        lenbc = 2 * ((b[1] - b[0] + 1) // r + (b[3] - b[2] + 1) // r)
        out = np.zeros(2 * self.H.nvar * lenbc)
        self.R.ref_get_reg(level, k, self.H.nvar, out.ctypes.data_as(C.c_void_p))
        return out

    def listbc(self, level, k):
        n = C.c_int()
        buf = np.zeros(5 * 100000, dtype=np.int32)
        self.R.ref_get_listbc(level, k, 100000, buf.ctypes.data_as(C.c_void_p), C.byref(n))
        out = buf[:5 * n.value].reshape(-1, 5).copy()
        # kid is a node number; the harness numbers the grids of level l+1
        # consecutively after those of levels 1..l
        first = 1 + sum(len(L["boxes"]) for L in self.H.levels[:level])
        out[:, 3] -= first
        return out

    def time(self, level, k=0):
        t = C.c_double()
        self.R.ref_get_time(level, k, C.byref(t))
        return t.value


# ---------------------------------------------------------------- flagging
UNSET, DONTFLAG, DOFLAG = -1.0, 0.0, 2.0


def flag_grid(q, mx, my, mbuff, xlower, ylower, dx, dy, t, level, tolsp, regions=()):
    """Oracle flags of one grid: UNSET, flagregions2, flag2refine2.  q is the
    grid's AoS array (my+4, mx+4, meqn); regions rows are (min_level,
    max_level, t_low, t_hi, x_low, x_hi, y_low, y_hi).  Returns flags of
    shape (my + 2 mbuff, mx + 2 mbuff)."""
    L = lib()
    L.orc_flagregions2.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_int, C.c_double, _dp, C.c_int, _dp]
    L.orc_flag2refine2.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp]
    q = np.ascontiguousarray(q, dtype=np.float64)
    meqn = q.shape[-1]
    f = np.full((my + 2 * mbuff) * (mx + 2 * mbuff), UNSET)
    reg = np.ascontiguousarray(np.asarray(regions, dtype=np.float64).reshape(-1), dtype=np.float64)
    if reg.size == 0:
        reg = np.zeros(8)
    rc = L.orc_flagregions2(mx, my, mbuff, xlower, ylower, dx, dy, level, t, reg, len(regions), f)
    if rc:
        raise RuntimeError("orc_flagregions2 error %d" % rc)
    L.orc_flag2refine2(mx, my, 2, mbuff, meqn, tolsp, q.reshape(-1), f)
    return f.reshape(my + 2 * mbuff, mx + 2 * mbuff)


def ref_flag_grid(q, mx, my, mbuff, xlower, ylower, dx, dy, t, level, tolsp, regions=()):
    """The reference's flagregions2 + flag2refine2 on the same inputs."""
    R = ref()
    R.ref_flag_grid.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, _dp,
                                C.c_int, _dp, _dp]
    q = np.ascontiguousarray(q, dtype=np.float64)
    meqn = q.shape[-1]
    f = np.zeros((my + 2 * mbuff) * (mx + 2 * mbuff))
    reg = np.ascontiguousarray(np.asarray(regions, dtype=np.float64).reshape(-1), dtype=np.float64)
    if reg.size == 0:
        reg = np.zeros(8)
    R.ref_flag_grid(mx, my, mbuff, meqn, xlower, ylower, dx, dy, t, level, tolsp,
                    q.reshape(-1), len(regions), reg, f)
    return f.reshape(my + 2 * mbuff, mx + 2 * mbuff)


# ------------------------------------------------------------------ 3-D
_ref3 = None


def make_params3(solver, *, order=2, transverse=22, limiter=4, cfl=0.9, cflv1=1.0, ubar=0.5,
                 vbar=1.0, wbar=-0.75, gamma=1.4, efix=True, mbc=2):
    """Params of a 3-D run: advection (1 equation, 1 wave) or Euler (5
    equations, 3 waves); transverse is method(3) = 10 m3 + m4."""
    p = OrcParams()
    p.solver = solver
    p.meqn, p.mwaves = {ADVECTION: (1, 1), EULER: (5, 3)}[solver]
    p.maux, p.mbc = 0, mbc
    for k in range(8):
        p.method[k] = 0
    p.method[2], p.method[3] = order, transverse
    lim = limiter if isinstance(limiter, (list, tuple)) else [limiter] * p.mwaves
    for k in range(10):
        p.mthlim[k] = lim[k] if k < len(lim) else 0
    p.mcapa, p.use_fwaves = 0, 0
    p.cfl, p.cflv1 = cfl, cflv1
    p.ubar, p.vbar, p.wbar = ubar, vbar, wbar
    p.gamma, p.efix = gamma, int(bool(efix))
    return p


def stepgrid3(p, q, dx, dy, dz, dt):
    """orc_stepgrid3 on one patch q (mz+4, my+4, mx+4, meqn): returns
    (qnew, [fm, fp, gm, gp, hm, hp], cflgrid, dtnew)."""
    qn = np.array(q, dtype=np.float64, order="C", copy=True)
    mz, my, mx = (n - 2 * p.mbc for n in qn.shape[:3])
    f = [np.zeros_like(qn) for _ in range(6)]
    cfl, dtn = C.c_double(), C.c_double()
    L = lib()
    L.orc_stepgrid3.argtypes = [C.POINTER(OrcParams), C.c_int, C.c_int, C.c_int, _dp, C.c_double,
                                C.c_double, C.c_double, C.c_double] + [_dp] * 6 + \
                               [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_stepgrid3(C.byref(p), mx, my, mz, qn, dx, dy, dz, dt, *f, C.byref(cfl), C.byref(dtn))
    return qn, f, cfl.value, dtn.value


def ref3_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libamrref3.so"))


def ref3_stepgrid3(p, q, dx, dy, dz, dt):
    """The reference's 3-D stepgrid.f on the same inputs."""
    global _ref3
    if _ref3 is None:
        lib()
        _ref3 = C.CDLL(os.path.join(HERE, "_ref", "libamrref3.so"))
    R = _ref3
    lib().orc_set_plugin(C.byref(p))
    R.ref3_set_method.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int,
                                  C.c_int, C.c_double, C.c_double, C.c_int]
    meth = (C.c_int * 7)(*[p.method[k] for k in range(1, 8)])
    lim = (C.c_int * 10)(*list(p.mthlim))
    R.ref3_set_method(meth, lim, p.mwaves, p.mcapa, p.use_fwaves, p.cfl, p.cflv1, p.mbc)
    qn = np.array(q, dtype=np.float64, order="C", copy=True)
    mz, my, mx = (n - 2 * p.mbc for n in qn.shape[:3])
    f = [np.zeros_like(qn) for _ in range(6)]
    cfl, dtn = C.c_double(), C.c_double()
    R.ref3_stepgrid.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_double,
                                C.c_double, C.c_double, C.c_double] + [_dp] * 6 + \
                               [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    R.ref3_stepgrid(p.meqn, p.mbc, mx, my, mz, qn, dx, dy, dz, dt, *f, C.byref(cfl), C.byref(dtn))
    return qn, f, cfl.value, dtn.value


def flag2refine3(q, tolsp, mbc=2):
    """orc_flag2refine3 on one 3-D grid q (mz+4, my+4, mx+4, meqn): the flag
    array (mz+4, my+4, mx+4) (0 = goodpt / DONTFLAG, 2 = DOFLAG) and the
    number of DOFLAG cells."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    mz, my, mx = (n - 2 * mbc for n in q.shape[:3])
    f = np.zeros(q.shape[:3])
    L = lib()
    L.orc_flag2refine3.restype = C.c_int
    L.orc_flag2refine3.argtypes = [C.c_int] * 5 + [C.c_double, _dp, C.c_double, C.c_double, _dp]
    n = L.orc_flag2refine3(mx, my, mz, mbc, q.shape[3], tolsp, q, 0.0, 2.0, f)
    return f, n


def mass3(qs, h, mbc=2):
    """conck.f (3-D) over a list of grids: running sum of q(1) times hx hy hz."""
    L = lib()
    L.orc_mass3.restype = C.c_double
    L.orc_mass3.argtypes = [C.c_int] * 5 + [_dp, C.c_double]
    acc = 0.0
    for q in qs:
        q = np.ascontiguousarray(q, dtype=np.float64)
        mz, my, mx = (n - 2 * mbc for n in q.shape[:3])
        acc = L.orc_mass3(mx, my, mz, mbc, q.shape[3], q, acc)
    return acc * h[0] * h[1] * h[2]


def ref3_flag2refine3(q, tolsp, mbc=2):
    """The reference's 3-D flag2refine.f on the same grid."""
    global _ref3
    if _ref3 is None:
        lib()
        _ref3 = C.CDLL(os.path.join(HERE, "_ref", "libamrref3.so"))
    q = np.ascontiguousarray(q, dtype=np.float64)
    mz, my, mx = (n - 2 * mbc for n in q.shape[:3])
    f = np.zeros(q.shape[:3])
    _ref3.ref3_flag2refine.argtypes = [C.c_int] * 5 + [_dp, C.c_double, _dp]
    _ref3.ref3_flag2refine(q.shape[3], mbc, mx, my, mz, q, tolsp, f)
    return f


# ------------------------------------------------- 3-D level update ----
class OrcHier3(C.Structure):
    _fields_ = [
        ("nlev", C.c_int), ("mxnest", C.c_int),
        ("ngrids", C.c_int * MAXLV),
        ("boxes", C.POINTER(C.c_int) * MAXLV),
        ("qnew", C.POINTER(C.c_double) * MAXLV),
        ("qold", C.POINTER(C.c_double) * MAXLV),
        ("qoff", C.POINTER(C.c_int64) * MAXLV),
        ("hx", C.c_double * MAXLV), ("hy", C.c_double * MAXLV), ("hz", C.c_double * MAXLV),
        ("possk", C.c_double * MAXLV), ("time", C.c_double * MAXLV),
        ("iregsz", C.c_int * MAXLV), ("jregsz", C.c_int * MAXLV), ("kregsz", C.c_int * MAXLV),
        ("intratx", C.c_int * MAXLV), ("intraty", C.c_int * MAXLV), ("intratz", C.c_int * MAXLV),
        ("xlower", C.c_double), ("xupper", C.c_double),
        ("ylower", C.c_double), ("yupper", C.c_double),
        ("zlower", C.c_double), ("zupper", C.c_double),
        ("xperdom", C.c_int), ("yperdom", C.c_int), ("zperdom", C.c_int),
        ("mthbc", C.c_int * 6),
    ]


class OrcRegs3(C.Structure):
    _fields_ = OrcRegs._fields_


def hier3_struct(H):
    """OrcHier3 view of an amrclaw_amd.hierarchy.Hier3 (shares its arrays)."""
    h = OrcHier3()
    h.nlev = h.mxnest = H.nlev
    for l, L in enumerate(H.levels, start=1):
        h.ngrids[l] = len(L["boxes"])
        h.boxes[l] = L["boxes"].ctypes.data_as(C.POINTER(C.c_int))
        h.qnew[l] = L["qnew"].ctypes.data_as(C.POINTER(C.c_double))
        h.qold[l] = L["qold"].ctypes.data_as(C.POINTER(C.c_double))
        h.qoff[l] = L["off"].ctypes.data_as(C.POINTER(C.c_int64))
        h.hx[l], h.hy[l], h.hz[l] = L["hx"], L["hy"], L["hz"]
        h.possk[l], h.time[l] = L["possk"], L["time"]
        h.iregsz[l], h.jregsz[l], h.kregsz[l] = L["iregsz"], L["jregsz"], L["kregsz"]
        h.intratx[l], h.intraty[l], h.intratz[l] = L["ratx"], L["raty"], L["ratz"]
    (h.xlower, h.xupper, h.ylower, h.yupper, h.zlower, h.zupper) = H.domain
    h.xperdom, h.yperdom, h.zperdom = int(H.xper), int(H.yper), int(H.zper)
    for k in range(6):
        h.mthbc[k] = H.mthbc[k]
    return h


def bound3(p, H, level, grid, time):
    """orc_bound3 on grid `grid` (0-based) of `level`, in place on H."""
    L = lib()
    L.orc_bound3.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcHier3), C.c_int, C.c_int,
                             C.c_double, C.c_void_p]
    h = hier3_struct(H)
    val = H.grid(level, grid, 0)
    rc = L.orc_bound3(C.byref(p), C.byref(h), level, grid, time, val.ctypes.data_as(C.c_void_p))
    if rc:
        raise RuntimeError("orc_bound3 error %d" % rc)


class OracleAMR3:
    """The oracle's 3-D level update (advanc / update / conck) in place on
    the arrays of a Hier3."""

    def __init__(self, p, H):
        L = lib()
        P, Hh, Rr = C.POINTER(OrcParams), C.POINTER(OrcHier3), C.POINTER(OrcRegs3)
        L.orc_prep_regs3.argtypes = [P, Hh, Rr]
        L.orc_free_regs3.argtypes = [Rr]
        L.orc_advanc3.argtypes = [P, Hh, Rr, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_update3.argtypes = [P, Hh, Rr, C.c_int]
        L.orc_conck3.argtypes = [P, Hh, C.c_int]
        L.orc_conck3.restype = C.c_double
        self.p, self.H = p, H
        self.h = hier3_struct(H)
        self.r = OrcRegs3()
        rc = L.orc_prep_regs3(C.byref(p), C.byref(self.h), C.byref(self.r))
        if rc:
            raise RuntimeError("orc_prep_regs3 error %d" % rc)

    def __del__(self):
        try:
            lib().orc_free_regs3(C.byref(self.r))
        except Exception:
            pass

    def advanc(self, level):
        dtn, cfl = C.c_double(), C.c_double()
        rc = lib().orc_advanc3(C.byref(self.p), C.byref(self.h), C.byref(self.r), level,
                               C.byref(dtn), C.byref(cfl))
        if rc:
            raise RuntimeError("orc_advanc3 error %d" % rc)
        self.H.levels[level - 1]["time"] = self.h.time[level]
        return dtn.value, cfl.value

    def update(self, level):
        rc = lib().orc_update3(C.byref(self.p), C.byref(self.h), C.byref(self.r), level)
        if rc:
            raise RuntimeError("orc_update3 error %d" % rc)

    def conck(self, level):
        return lib().orc_conck3(C.byref(self.p), C.byref(self.h), level)

    def reg(self, level, k):
        n = 2 * self.p.meqn * self.r.lenbc[level][k]
        off = self.r.roff[level][k]
        return np.ctypeslib.as_array(self.r.reg[level], shape=(self.r.nreg[level],))[off:off + n].copy()

    def listbc(self, level, k):
        if not self.r.listbc[level]:
            return np.zeros((0, 6), dtype=np.int32)
        a = np.ctypeslib.as_array(self.r.listbc[level], shape=(6 * self.r.nlist[level],))
        s, c = self.r.lst_start[level][k], self.r.lst_count[level][k]
        return a[6 * s:6 * (s + c)].reshape(-1, 6).copy()


def _ref3_lib(p):
    global _ref3
    if _ref3 is None:
        lib()
        _ref3 = C.CDLL(os.path.join(HERE, "_ref", "libamrref3.so"))
    R = _ref3
    lib().orc_set_plugin(C.byref(p))
    R.ref3_set_method.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int,
                                  C.c_int, C.c_double, C.c_double, C.c_int]
    meth = (C.c_int * 7)(*[p.method[k] for k in range(1, 8)])
    lim = (C.c_int * 10)(*list(p.mthlim))
    R.ref3_set_method(meth, lim, p.mwaves, p.mcapa, p.use_fwaves, p.cfl, p.cflv1, p.mbc)
    return R


def ref3_load_hier(p, H):
    """Install a Hier3 in the reference's 3-D amr_module (harness)."""
    R = _ref3_lib(p)
    ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    R.ref3_hier_setup.argtypes = [C.c_int] * 4 + [ip, ip] + [dp] * 5 + [ip] * 6 + \
        [C.c_double] * 6 + [C.c_int] * 3 + [ip]
    R.ref3_hier_xfer.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int]
    lv = H.levels
    f64 = lambda k: np.array([L[k] for L in lv], dtype=np.float64)
    i32 = lambda k: np.array([L[k] for L in lv], dtype=np.int32)
    boxes = np.ascontiguousarray(np.concatenate([L["boxes"].ravel() for L in lv]).astype(np.int32))
    ngr = np.array([len(L["boxes"]) for L in lv], dtype=np.int32)
    R.ref3_hier_setup(H.nlev, H.nlev, H.nvar, H.ng, ngr, boxes, f64("hx"), f64("hy"), f64("hz"),
                      f64("possk"), f64("time"), i32("iregsz"), i32("jregsz"), i32("kregsz"),
                      i32("ratx"), i32("raty"), i32("ratz"), *H.domain, int(H.xper), int(H.yper),
                      int(H.zper), np.array(H.mthbc, dtype=np.int32))
    for l in range(1, H.nlev + 1):
        for k in range(len(lv[l - 1]["boxes"])):
            for which in (0, 1):
                g = np.ascontiguousarray(H.grid(l, k, which))
                R.ref3_hier_xfer(l, k, which, H.nvar, g.ctypes.data_as(C.c_void_p), 1)
    return R


def ref3_get_grid(H, level, grid, which=0):
    out = np.zeros(H.levels[level - 1]["shapes"][grid])
    _ref3.ref3_hier_xfer(level, grid, which, H.nvar, out.ctypes.data_as(C.c_void_p), 0)
    return out


def ref3_bound(p, H, level, grid, time):
    """The reference's 3-D bound.f on a freshly installed copy of H; returns
    the grid's new array after the call."""
    R = ref3_load_hier(p, H)
    R.ref3_bound.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int]
    R.ref3_bound(level, grid, time, H.nvar)
    return ref3_get_grid(H, level, grid)


class RefAMR3:
    """The reference's own 3-D routines over a copy of a Hier3 installed in
    amr_module (prepf/prepc, bound, saveqc, qad, stepgrid, fluxsv, fluxad,
    update)."""

    def __init__(self, p, H):
        self.p, self.H = p, H
        R = ref3_load_hier(p, H)
        R.ref3_prep_regs.argtypes = [C.c_int, C.c_int]
        R.ref3_advanc.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        R.ref3_update.argtypes = [C.c_int, C.c_int]
        R.ref3_get_reg.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p]
        R.ref3_get_listbc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_int)]
        R.ref3_prep_regs(H.nlev, H.nvar)
        self.R = R

    def advanc(self, level):
        dtn, cfl = C.c_double(), C.c_double()
        self.R.ref3_advanc(level, self.H.nvar, C.byref(dtn), C.byref(cfl))
        return dtn.value, cfl.value

    def update(self, level):
        self.R.ref3_update(level, self.H.nvar)

    def grid(self, level, k, which=0):
        return ref3_get_grid(self.H, level, k, which)

    def reg(self, level, k):
        b = self.H.levels[level - 1]["boxes"][k]
        L = self.H.levels[level - 2]
        ik = (b[1] - b[0] + 1) // L["ratx"]
        jk = (b[3] - b[2] + 1) // L["raty"]
        kk = (b[5] - b[4] + 1) // L["ratz"]
        lenbc = 2 * (ik * jk + jk * kk + kk * ik)
        out = np.zeros(2 * self.H.nvar * lenbc)
        self.R.ref3_get_reg(level, k, self.H.nvar, out.ctypes.data_as(C.c_void_p))
        return out

    def listbc(self, level, k):
        n = C.c_int()
        buf = np.zeros(6 * 200000, dtype=np.int32)
        self.R.ref3_get_listbc(level, k, 200000, buf.ctypes.data_as(C.c_void_p), C.byref(n))
        out = buf[:6 * n.value].reshape(-1, 6).copy()
        first = 1 + sum(len(L["boxes"]) for L in self.H.levels[:level])
        out[:, 4] -= first
        return out


def flag_level3(O, level, tolsp, tol=0.0, iorder=2, richardson=False):
    """spest.f (3-D) and, with richardson, errest.f + errf1.f for every grid
    of `level` of an OracleAMR3: returns (flags, rflags, nflag) with one
    (mz+4, my+4, mx+4) array per grid (goodpt 0 / badpt 2); rflags is None
    without richardson."""
    L = lib()
    P, Hh = C.POINTER(OrcParams), C.POINTER(OrcHier3)
    L.orc_spest3.argtypes = [P, Hh, C.c_int, C.c_int, C.c_double, _dp, _dp, C.POINTER(C.c_int)]
    L.orc_errest3.argtypes = [P, Hh, C.c_int, C.c_int, C.c_double, C.c_int, _dp, _dp]
    H = O.H
    flags, rflags, total = [], [], 0
    for k in range(len(H.levels[level - 1]["boxes"])):
        shape = H.levels[level - 1]["shapes"][k]
        big = np.zeros(shape)
        f = np.zeros(shape[:3])
        n = C.c_int()
        rc = L.orc_spest3(C.byref(O.p), C.byref(O.h), level, k, tolsp, big, f, C.byref(n))
        if rc:
            raise RuntimeError("orc_spest3 error %d" % rc)
        flags.append(f)
        total += n.value
        if richardson:
            r = np.zeros(shape[:3])
            rc = L.orc_errest3(C.byref(O.p), C.byref(O.h), level, k, tol, iorder, big, r)
            if rc:
                raise RuntimeError("orc_errest3 error %d" % rc)
            rflags.append(r)
    return flags, (rflags if richardson else None), total


def ref3_flag_level(R, level, tolsp, tol=0.0, iorder=2, richardson=False):
    """The reference's 3-D flagging of `level` on a RefAMR3's hierarchy:
    (dom1flags over the level's index space with a one-cell border,
    per-grid errf1 flags or None)."""
    H = R.H
    Lv = H.levels[level - 1]
    isz, jsz, ksz = Lv["iregsz"], Lv["jregsz"], Lv["kregsz"]
    dom = np.zeros((ksz + 2, jsz + 2, isz + 2), dtype=np.int8)
    sizes = [int(np.prod(s[:3])) for s in Lv["shapes"]]
    rf = np.zeros(max(sum(sizes), 1))
    R.R.ref3_flag_level.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                    np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS"), _dp]
    R.R.ref3_flag_level(level, H.nvar, tolsp, tol, iorder, int(bool(richardson)), dom, rf)
    rflags = None
    if richardson:
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(int)
        rflags = [rf[o:o + n].reshape(s[:3]) for o, n, s in zip(offs, sizes, Lv["shapes"])]
    return dom, rflags


def domain_flags3(H, level, flags):
    """setflags.f (3-D): the grids' interior flags summed into the level's
    dom1flags(0:isize+1, 0:jsize+1, 0:ksize+1)."""
    Lv = H.levels[level - 1]
    dom = np.zeros((Lv["kregsz"] + 2, Lv["jregsz"] + 2, Lv["iregsz"] + 2), dtype=np.int8)
    for b, f in zip(Lv["boxes"], flags):
        dom[b[4] + 1:b[5] + 2, b[2] + 1:b[3] + 2, b[0] + 1:b[1] + 2] += f[2:-2, 2:-2, 2:-2].astype(np.int8)
    return dom
