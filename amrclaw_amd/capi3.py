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
        L.amrgpu3_get_fluxes.argtypes = [C.c_int] + [C.c_void_p] * 6
        L.amrgpu3_stepgrid.argtypes = [C.c_int] * 3 + [_dp, _dp] + [C.c_double] * 4 + \
                                      [C.c_void_p] * 6 + [dpo, dpo]
        L.amrgpu3_set_timing.argtypes = [C.c_int]
        L.amrgpu3_kernel_time.argtypes = [C.c_int, dpo, C.POINTER(C.c_int64)]
        L.amrgpu3_synchronize.argtypes = []
        L.amrgpu3_flag_patches.argtypes = [C.c_double, C.POINTER(C.c_int64)]
        L.amrgpu3_get_flags.argtypes = [C.c_int, _dp]
        L.amrgpu3_conck.argtypes = [C.c_double] * 3 + [dpo]
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
