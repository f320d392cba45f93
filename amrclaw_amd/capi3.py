"""ctypes binding of the 3-D patch step of libamrgpu.so (include/amrgpu3.h).

The Python mirror of the binding a Fortran driver makes through
ISO_C_BINDING for src/3d/stepgrid.f.  Arrays are the reference's
q(nvar, 1-nghost:mx+nghost, 1-nghost:my+nghost, 1-nghost:mz+nghost) in
Fortran order, i.e. numpy arrays of shape (mz+4, my+4, mx+4, nvar).
Loading fails loudly if the HIP library has not been built; there is no CPU
fallback.
"""
import ctypes as C

import numpy as np

from .capi import ADVECTION, EULER, FP32, FP64, AmrgpuError, lib as _lib2

KT_X, KT_Y, KT_Z, KT_UPDATE = 0, 1, 2, 3


class Amrgpu3Config(C.Structure):
    _fields_ = [
        ("solver", C.c_int), ("nvar", C.c_int), ("mwaves", C.c_int),
        ("naux", C.c_int), ("nghost", C.c_int),
        ("method", C.c_int * 7), ("mthlim", C.c_int * 8),
        ("mcapa", C.c_int), ("use_fwaves", C.c_int),
        ("cfl", C.c_double), ("cflv1", C.c_double),
        ("ubar", C.c_double), ("vbar", C.c_double), ("wbar", C.c_double),
        ("gamma", C.c_double), ("efix", C.c_int),
        ("precision", C.c_int),
    ]


_bound = False
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _bound
    L = _lib2()
    if not _bound:
        dpo = C.POINTER(C.c_double)
        L.amrgpu3_init.argtypes = [C.POINTER(Amrgpu3Config)]
        L.amrgpu3_finalize.argtypes = []
        L.amrgpu3_last_error.argtypes = [C.c_char_p, C.c_int]
        L.amrgpu3_set_patches.argtypes = [C.c_int, _ip, _ip, _ip, C.c_int]
        L.amrgpu3_put_patch.argtypes = [C.c_int, _dp]
        L.amrgpu3_get_patch.argtypes = [C.c_int, _dp]
        L.amrgpu3_step_patches.argtypes = [C.c_double] * 4 + [C.c_void_p]
        L.amrgpu3_set_step_mode.argtypes = [C.c_int]
        L.amrgpu3_get_fluxes.argtypes = [C.c_int] + [C.c_void_p] * 6
        L.amrgpu3_stepgrid.argtypes = [C.c_int] * 3 + [_dp, _dp] + [C.c_double] * 4 + \
                                      [C.c_void_p] * 6 + [dpo, dpo]
        L.amrgpu3_set_timing.argtypes = [C.c_int]
        L.amrgpu3_kernel_time.argtypes = [C.c_int, dpo, C.POINTER(C.c_int64)]
        L.amrgpu3_synchronize.argtypes = []
        L.amrgpu3_flag_patches.argtypes = [C.c_double, C.POINTER(C.c_int64)]
        L.amrgpu3_get_flags.argtypes = [C.c_int, _dp]
        L.amrgpu3_conck.argtypes = [C.c_double] * 3 + [dpo]
        L.amrgpu3_set_geometry.argtypes = [C.c_int] + [_ip] * 6 + [_dp] * 4 + [_ip]
        L.amrgpu3_set_level.argtypes = [C.c_int, C.c_int, _ip]
        L.amrgpu3_set_time.argtypes = [C.c_int, C.c_double, C.c_double]
        L.amrgpu3_set_plane_block.argtypes = [C.c_int]
        L.amrgpu3_get_time.argtypes = [C.c_int, dpo]
        _i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.amrgpu3_put_level.argtypes = [C.c_int, C.c_int, _dp, _i64]
        L.amrgpu3_get_level.argtypes = [C.c_int, C.c_int, _dp, _i64]
        L.amrgpu3_bound_level.argtypes = [C.c_int, C.c_double]
        L.amrgpu3_advance_level.argtypes = [C.c_int, dpo, dpo]
        L.amrgpu3_update_level.argtypes = [C.c_int]
        L.amrgpu3_conck_level.argtypes = [C.c_int, dpo]
        L.amrgpu3_get_registers.argtypes = [C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_int)]
        L.amrgpu3_flag_level.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_int,
                                         C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.amrgpu3_get_level_flags.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        _bound = True
    return L


def _check(rc):
    if rc != 0:
        buf = C.create_string_buffer(512)
        lib().amrgpu3_last_error(buf, 512)
        raise AmrgpuError("amrgpu3 error %d: %s" % (rc, buf.value.decode()))


def make_config3(solver, *, order=2, transverse=22, limiter=4, cfl=0.9, cflv1=1.0, ubar=0.5,
                 vbar=1.0, wbar=-0.75, gamma=1.4, efix=True, precision=FP64, use_fwaves=0):
    """3-D run parameters: advection (1 equation, 1 wave) or Euler (5, 3);
    transverse is method(3) = 10 m3 + m4."""
    c = Amrgpu3Config()
    c.solver = solver
    c.nvar, c.mwaves = {ADVECTION: (1, 1), EULER: (5, 3)}[solver]
    c.naux, c.nghost = 0, 2
    meth = [0, order, transverse, 0, 0, 0, 0]
    for k in range(7):
        c.method[k] = meth[k]
    lim = limiter if isinstance(limiter, (list, tuple)) else [limiter] * c.mwaves
    for k in range(8):
        c.mthlim[k] = lim[k] if k < len(lim) else 0
    c.mcapa, c.use_fwaves = 0, use_fwaves
    c.cfl, c.cflv1 = cfl, cflv1
    c.ubar, c.vbar, c.wbar = ubar, vbar, wbar
    c.gamma, c.efix = gamma, int(bool(efix))
    c.precision = precision
    return c


def init3(cfg):
    _check(lib().amrgpu3_init(C.byref(cfg)))


def finalize3():
    _check(lib().amrgpu3_finalize())


def set_patches3(sizes, plane_block=0):
    """sizes: (n, 3) of (mx, my, mz) per grid."""
    s = np.asarray(sizes, dtype=np.int32).reshape(-1, 3)
    cols = [np.ascontiguousarray(s[:, k]) for k in range(3)]
    _check(lib().amrgpu3_set_patches(s.shape[0], *cols, plane_block))


def put_patch3(p, q):
    _check(lib().amrgpu3_put_patch(p, np.ascontiguousarray(q, dtype=np.float64)))


def get_patch3(p, shape):
    out = np.empty(shape, dtype=np.float64)
    _check(lib().amrgpu3_get_patch(p, out))
    return out


def step_patches3(dx, dy, dz, dt, npatch=None):
    """One step of every grid of the batch; returns the per-grid Courant
    numbers when npatch is given."""
    if npatch is None:
        _check(lib().amrgpu3_step_patches(dx, dy, dz, dt, None))
        return None
    cfl = np.zeros(npatch, dtype=np.float64)
    _check(lib().amrgpu3_step_patches(dx, dy, dz, dt, cfl.ctypes.data_as(C.c_void_p)))
    return cfl


STEP_PHASES, STEP_FUSED = 0, 1


def set_step_mode3(mode):
    """amrgpu3_set_step_mode: STEP_PHASES (flux arrays in HBM, FP64 bit-exact)
    or STEP_FUSED (FP32, every face on chip; the FP32 default)."""
    _check(lib().amrgpu3_set_step_mode(int(mode)))


def get_fluxes3(p, shape):
    f = [np.zeros(shape, dtype=np.float64) for _ in range(6)]
    _check(lib().amrgpu3_get_fluxes(p, *[a.ctypes.data_as(C.c_void_p) for a in f]))
    return f


def stepgrid3(q, dx, dy, dz, dt, want_fluxes=False):
    """amrgpu3_stepgrid on one grid q (mz+4, my+4, mx+4, nvar).  Returns
    (qnew, [fm, fp, gm, gp, hm, hp] or None, cflgrid, dtnew)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    mz, my, mx = (n - 4 for n in q.shape[:3])
    out = np.empty_like(q)
    fl = [np.zeros_like(q) for _ in range(6)] if want_fluxes else None
    ptrs = [f.ctypes.data_as(C.c_void_p) for f in fl] if fl else [None] * 6
    cfl, dtn = C.c_double(), C.c_double()
    _check(lib().amrgpu3_stepgrid(mx, my, mz, q, out, dx, dy, dz, dt, *ptrs, C.byref(cfl),
                                  C.byref(dtn)))
    return out, fl, cfl.value, dtn.value


def set_timing3(on):
    _check(lib().amrgpu3_set_timing(int(on)))


def kernel_time3(kind):
    ms, n = C.c_double(), C.c_int64()
    _check(lib().amrgpu3_kernel_time(kind, C.byref(ms), C.byref(n)))
    return ms.value, n.value


def synchronize3():
    _check(lib().amrgpu3_synchronize())


def flag_patches3(tolsp):
    """flag2refine on every grid of the batch; returns the number flagged."""
    n = C.c_int64()
    _check(lib().amrgpu3_flag_patches(tolsp, C.byref(n)))
    return n.value


def get_flags3(p, shape):
    """Flags of grid p, shape (mz+4, my+4, mx+4)."""
    out = np.empty(shape, dtype=np.float64)
    _check(lib().amrgpu3_get_flags(p, out))
    return out


def conck3(h):
    m = C.c_double()
    _check(lib().amrgpu3_conck(h[0], h[1], h[2], C.byref(m)))
    return m.value


__all__ = ["ADVECTION", "EULER", "FP32", "FP64", "KT_X", "KT_Y", "KT_Z", "KT_UPDATE",
           "Amrgpu3Config", "make_config3", "init3", "finalize3", "set_patches3", "put_patch3",
           "get_patch3", "step_patches3", "get_fluxes3", "stepgrid3", "set_timing3",
           "kernel_time3", "synchronize3", "flag_patches3", "get_flags3", "conck3"]


# ------------------------------------------------------------ hierarchy ---
def load_hierarchy3(H):
    """Install an amrclaw_amd.hierarchy.Hier3: geometry (domain, mthbc,
    ratios), the grids of every level, both time levels and the times."""
    lv = H.levels
    i32 = lambda k: np.ascontiguousarray([L[k] for L in lv], dtype=np.int32)
    f64 = lambda k: np.ascontiguousarray([L[k] for L in lv], dtype=np.float64)
    _check(lib().amrgpu3_set_geometry(H.nlev, i32("iregsz"), i32("jregsz"), i32("kregsz"),
                                      i32("ratx"), i32("raty"), i32("ratz"), f64("hx"), f64("hy"),
                                      f64("hz"), np.ascontiguousarray(H.domain, dtype=np.float64),
                                      np.ascontiguousarray(H.mthbc, dtype=np.int32)))
    for l, L in enumerate(lv, start=1):
        b = np.ascontiguousarray(L["boxes"], dtype=np.int32).reshape(-1)
        _check(lib().amrgpu3_set_level(l, len(L["boxes"]), b if b.size else np.zeros(6, np.int32)))
        _check(lib().amrgpu3_set_time(l, L["time"], L["possk"]))
        if len(L["boxes"]):
            _check(lib().amrgpu3_put_level(l, 0, L["qnew"], L["off"]))
            _check(lib().amrgpu3_put_level(l, 1, L["qold"], L["off"]))


def read_level3(H, l, which=0):
    """Download level l (new: which = 0, old: 1) laid out like H.levels[l-1]."""
    L = H.levels[l - 1]
    out = np.zeros_like(L["qnew"])
    _check(lib().amrgpu3_get_level(l, which, out, L["off"]))
    return out


def set_plane_block3(plane_block):
    """Planes per sweep workgroup of the hierarchy's levels (0: per level)."""
    _check(lib().amrgpu3_set_plane_block(plane_block))


def set_time3(level, time, possk):
    _check(lib().amrgpu3_set_time(level, time, possk))


def get_time3(level):
    t = C.c_double()
    _check(lib().amrgpu3_get_time(level, C.byref(t)))
    return t.value


def bound_level3(level, time):
    """bound.f (3-D) for every grid of the level at `time` (ghost cells only)."""
    _check(lib().amrgpu3_bound_level(level, time))


def advance_level3(level):
    """advanc.f (3-D) for one level: returns (dtlevnew, cfl_level)."""
    dtn, cfl = C.c_double(), C.c_double()
    _check(lib().amrgpu3_advance_level(level, C.byref(dtn), C.byref(cfl)))
    return dtn.value, cfl.value


def update_level3(level):
    """update.f with upbnd.f (3-D): reflux and average onto `level`."""
    _check(lib().amrgpu3_update_level(level))


def conck_level3(level):
    m = C.c_double()
    _check(lib().amrgpu3_conck_level(level, C.byref(m)))
    return m.value


def get_registers3(level, grid, nvar):
    """(svdflx, qc1d) of one fine grid, each (lenbc, nvar)."""
    n = C.c_int()
    _check(lib().amrgpu3_get_registers(level, grid, None, C.byref(n)))
    buf = np.zeros(2 * nvar * n.value)
    _check(lib().amrgpu3_get_registers(level, grid, buf.ctypes.data_as(C.c_void_p), C.byref(n)))
    lb = n.value
    return buf[:nvar * lb].reshape(lb, nvar), buf[nvar * lb:].reshape(lb, nvar)


def flag_level3(level, tolsp, *, richardson=False, tol=0.0, iorder=2):
    """spest (and with richardson errest/errf1) on every grid of the level;
    returns (flagged cells, Richardson-flagged cells)."""
    n, r = C.c_int64(), C.c_int64()
    _check(lib().amrgpu3_flag_level(level, tolsp, 1 if richardson else 0, tol, iorder,
                                    C.byref(n), C.byref(r)))
    return n.value, r.value


def get_level_flags3(level, grid, shape):
    """(amrflags, rflags) of one grid, each of shape (mz+4, my+4, mx+4)."""
    f = np.empty(shape, dtype=np.float64)
    r = np.empty(shape, dtype=np.float64)
    _check(lib().amrgpu3_get_level_flags(level, grid, f.ctypes.data_as(C.c_void_p),
                                         r.ctypes.data_as(C.c_void_p)))
    return f, r


class DeviceAMR3:
    """The library's 3-D level update over a Hier3 loaded on the device
    (the advanc / update / conck interface of the reference's driver)."""

    def __init__(self, H):
        self.H = H
        load_hierarchy3(H)

    def advanc(self, level):
        return advance_level3(level)

    def update(self, level):
        update_level3(level)

    def conck(self, level):
        return conck_level3(level)

    def flag(self, level, tolsp, **kw):
        return flag_level3(level, tolsp, **kw)
