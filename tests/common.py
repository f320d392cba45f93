"""Shared test helpers: seeded inputs in the reference layout."""
import numpy as np

from oracle import orc
from amrclaw_amd import capi
from amrclaw_amd.hierarchy import Hier  # noqa: F401  (re-exported for the tests)

SOLVERS = {"advection": orc.ADVECTION, "acoustics": orc.ACOUSTICS,
           "euler": orc.EULER}


def euler_states(rng, shape, gamma=1.4):
    rho = 0.5 + rng.random(shape)
    u = rng.normal(size=shape)
    v = rng.normal(size=shape)
    p = 0.3 + rng.random(shape)
    return np.stack([rho, rho * u, rho * v,
                     p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)], -1)


def random_patch(rng, solver, mx, my, mbc=2):
    shape = (my + 2 * mbc, mx + 2 * mbc)
    if solver == orc.EULER:
        return euler_states(rng, shape)
    nv = 1 if solver == orc.ADVECTION else 3
    return rng.normal(size=shape + (nv,))


def quadrant_euler(x, y, gamma=1.4, xs=0.8, ys=0.8):
    """Four-quadrant 2-D Riemann problem (examples/euler_2d_quadrants/qinit.f
    :33-69 with the rounded states of the batched Euler configuration)."""
    x = np.asarray(x)
    y = np.asarray(y)
    st = {1: (1.5, 0.0, 0.0, 1.5), 2: (0.5323, 1.206, 0.0, 0.3),
          3: (0.138, 1.206, 1.206, 0.029), 4: (0.5323, 0.0, 1.206, 0.3)}
    iq = np.where(x >= xs, np.where(y >= ys, 1, 4), np.where(y >= ys, 2, 3))
    q = np.zeros(x.shape + (4,))
    for k, (r, u, v, p) in st.items():
        m = iq == k
        q[m, 0] = r
        q[m, 1] = r * u
        q[m, 2] = r * v
        q[m, 3] = p / (gamma - 1.0) + 0.5 * r * (u * u + v * v)
    return q


def both_params(solver, **kw):
    """Matching oracle params and library config."""
    ckw = dict(kw)
    precision = ckw.pop("precision", capi.FP64)
    p = orc.make_params(solver, **ckw)
    c = capi.make_config(solver, precision=precision, **ckw)
    return p, c


def rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def disjoint_boxes(rng, isz, jsz, nbox, lo=4, hi=20, align=2, within=None, tries=400):
    """Random disjoint boxes (ilo, ihi, jlo, jhi) inside [0,isz) x [0,jsz)
    (or inside the union `within` of boxes), corners aligned to `align`."""
    out = []
    for _ in range(tries):
        if len(out) >= nbox:
            break
        w = int(rng.integers(lo // align, hi // align + 1)) * align
        h = int(rng.integers(lo // align, hi // align + 1)) * align
        if w > isz or h > jsz:
            continue
        i0 = int(rng.integers(0, (isz - w) // align + 1)) * align
        j0 = int(rng.integers(0, (jsz - h) // align + 1)) * align
        b = (i0, i0 + w - 1, j0, j0 + h - 1)
        if any(not (b[1] < c[0] or c[1] < b[0] or b[3] < c[2] or c[3] < b[2]) for c in out):
            continue
        if within is not None and not any(c[0] <= b[0] and b[1] <= c[1] and c[2] <= b[2]
                                          and b[3] <= c[3] for c in within):
            continue
        out.append(b)
    return out


def random_hierarchy(rng, nvar=3, nx=24, ny=20, nlev=3, ratio=2, mthbc=(1, 1, 1, 1),
                     xper=False, yper=False, fine_phase=(1, 1), level1_split=(2, 2)):
    """Level 1 tiles the domain; finer levels hold random disjoint boxes.
    Times follow Berger-Oliger: every level l < nlev has just been advanced
    to T = 1.0 (new) from T - possk (old); level nlev sits inside the last
    coarse step at fine_phase (k, r): time = T - possk(l-1) + k possk(l)."""
    H = Hier(nvar, mthbc=mthbc, xper=xper, yper=yper)
    hx, hy = 1.0 / nx, 1.0 / ny
    dt1 = 0.08
    T = 1.0
    isz, jsz = nx, ny
    cuts_i = np.linspace(0, nx, level1_split[0] + 1).astype(int)
    cuts_j = np.linspace(0, ny, level1_split[1] + 1).astype(int)
    boxes = [(cuts_i[a], cuts_i[a + 1] - 1, cuts_j[b], cuts_j[b + 1] - 1)
             for b in range(level1_split[1]) for a in range(level1_split[0])]
    possk = dt1
    for l in range(1, nlev + 1):
        if l > 1:
            isz, jsz = isz * ratio, jsz * ratio
            hx, hy = hx / ratio, hy / ratio
            possk = possk / ratio
            boxes = disjoint_boxes(rng, isz, jsz, 6, lo=4, hi=max(6, isz // 3), align=ratio)
        t = T if l < nlev else T - possk * ratio + fine_phase[0] * possk
        H.add_level(boxes, hx, hy, possk, t, isz, jsz, ratio, ratio)
    for l in range(1, nlev + 1):
        L = H.levels[l - 1]
        L["qnew"][:] = rng.normal(size=L["qnew"].shape)
        L["qold"][:] = rng.normal(size=L["qold"].shape)
    return H


def nested_hierarchy(rng, solver_states, nvar, nx=24, ny=20, nlev=3, ratio=2, mthbc=(1, 1, 1, 1),
                     xper=False, yper=False, dt1=None, nfine=5, level1_split=(2, 2), naux=0):
    """Properly nested hierarchy at one time (t = 0, old = new) for time
    stepping: every level-(l+1) grid, coarsened and grown by one coarse
    cell, lies inside one level-l grid or outside the domain (the nesting
    saveqc and the coarse-fine registers assume).  solver_states(rng, shape)
    returns physical states of shape shape + (nvar,)."""
    H = Hier(nvar, mthbc=mthbc, xper=xper, yper=yper, naux=naux)
    hx, hy = 1.0 / nx, 1.0 / ny
    isz, jsz = nx, ny
    cuts_i = np.linspace(0, nx, level1_split[0] + 1).astype(int)
    cuts_j = np.linspace(0, ny, level1_split[1] + 1).astype(int)
    boxes = [(cuts_i[a], cuts_i[a + 1] - 1, cuts_j[b], cuts_j[b + 1] - 1)
             for b in range(level1_split[1]) for a in range(level1_split[0])]
    possk = dt1 if dt1 is not None else 0.2 * min(hx, hy)
    for l in range(1, nlev + 1):
        if l > 1:
            # a fine grid may reach the domain edge where the coarse grid
            # does, unless the edge is periodic and the coarse level does not
            # cover the whole domain (its ring would wrap onto uncovered cells)
            full = l - 1 == 1
            xe, ye = full or not xper, full or not yper
            within = []
            for (a, b, c, d) in boxes:
                within.append((0 if (a == 0 and xe) else (a + 1) * ratio,
                               isz * ratio - 1 if (b == isz - 1 and xe) else b * ratio - 1,
                               0 if (c == 0 and ye) else (c + 1) * ratio,
                               jsz * ratio - 1 if (d == jsz - 1 and ye) else d * ratio - 1))
            isz, jsz = isz * ratio, jsz * ratio
            hx, hy = hx / ratio, hy / ratio
            possk = possk / ratio
            boxes = disjoint_boxes(rng, isz, jsz, nfine, lo=4, hi=max(8, isz // 3),
                                   align=ratio, within=within, tries=4000)
        H.add_level(boxes, hx, hy, possk, 0.0, isz, jsz, ratio, ratio)
    for l in range(1, nlev + 1):
        L = H.levels[l - 1]
        for k in range(len(L["boxes"])):
            g = H.grid(l, k, 0)
            g[:] = solver_states(rng, g.shape[:2])
            H.grid(l, k, 1)[:] = g
    return H


def nested_boxes(rng, H, level, nbox, lo=4, hi=None):
    """Random disjoint boxes for `level` of H, properly nested in H's level
    level-1 the way regrid makes new grids (each, coarsened and grown by
    one coarse cell, inside one level-1 grid or outside the domain), as a
    new layout for a regrid of `level`."""
    C = H.levels[level - 2]
    F = H.levels[level - 1]
    rx, ry = C["ratx"], C["raty"]
    isz, jsz = C["iregsz"], C["jregsz"]
    full = level - 1 == 1
    xe, ye = full or not H.xper, full or not H.yper
    within = []
    for (a, b, c, d) in C["boxes"]:
        within.append((0 if (a == 0 and xe) else (a + 1) * rx,
                       isz * rx - 1 if (b == isz - 1 and xe) else b * rx - 1,
                       0 if (c == 0 and ye) else (c + 1) * ry,
                       jsz * ry - 1 if (d == jsz - 1 and ye) else d * ry - 1))
    return disjoint_boxes(rng, F["iregsz"], F["jregsz"], nbox, lo=lo,
                          hi=hi or max(8, F["iregsz"] // 3), align=max(rx, ry),
                          within=within, tries=4000)


def vc_annulus_aux(H, amp=0.3):
    """Fill the aux arrays of H (naux = 3) as setaux does for
    variable-coefficient advection with capacity on every grid and ghost
    cell: aux(1) the x-velocity at the cell's left edge and aux(2) the
    y-velocity at its bottom edge, differenced from the stream function
    amp sin(pi x) sin(pi y) (discretely divergence-free, zero normal flow
    through the unit square's edges), aux(3) the capacity 1 + 0.3 x + 0.2 y
    at the cell centre (its exact cell average, so a coarse cell's capacity
    is the mean of its children's)."""
    psi = lambda x, y: amp * np.sin(np.pi * x) * np.sin(np.pi * y)
    x0, _, y0, _ = H.domain
    for l, L in enumerate(H.levels, start=1):
        hx, hy = L["hx"], L["hy"]
        for k, b in enumerate(L["boxes"]):
            a = H.aux(l, k)
            jj, ii = np.meshgrid(np.arange(b[2] - 2, b[3] + 3), np.arange(b[0] - 2, b[1] + 3),
                                 indexing="ij")
            xl, yb = x0 + ii * hx, y0 + jj * hy
            a[..., 0] = (psi(xl, yb + hy) - psi(xl, yb)) / hy
            a[..., 1] = -(psi(xl + hx, yb) - psi(xl, yb)) / hx
            a[..., 2] = 1.0 + 0.3 * (xl + 0.5 * hx) + 0.2 * (yb + 0.5 * hy)


def positive_states(rng, shape):
    """A smooth positive scalar field with noise (one component)."""
    ny, nx = shape
    y, x = np.meshgrid(np.linspace(0, 1, ny), np.linspace(0, 1, nx), indexing="ij")
    return (1.0 + 0.5 * np.sin(5 * x + rng.random() * 6) * np.cos(4 * y)
            + 0.05 * rng.random(shape))[..., None]


def smooth_euler(rng, shape):
    """Positive, smoothly varying Euler states (for hierarchies)."""
    ny, nx = shape
    y, x = np.meshgrid(np.linspace(0, 1, ny), np.linspace(0, 1, nx), indexing="ij")
    ph = rng.random(4) * 6.28
    rho = 1.0 + 0.3 * np.sin(3 * x + ph[0]) * np.cos(2 * y + ph[1]) + 0.05 * rng.random(shape)
    u = 0.4 * np.sin(2 * y + ph[2]) + 0.05 * rng.normal(size=shape)
    v = 0.4 * np.cos(2 * x + ph[3]) + 0.05 * rng.normal(size=shape)
    p = 1.0 + 0.2 * rng.random(shape)
    return np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)], -1)


def acoustic_states(rng, shape):
    return rng.normal(size=shape + (3,))


def berger_oliger(impl, level, nlev, ratio, log=None):
    """One step of `level` and the subcycled steps of the finer levels
    (tick.f:230-300 with fixed time steps)."""
    r = impl.advanc(level)
    if log is not None:
        log.append((level, r))
    if level < nlev:
        for _ in range(ratio):
            berger_oliger(impl, level + 1, nlev, ratio, log)
        impl.update(level)
