"""The 2-D run configurations as static hierarchies for parity tests.

C1 advection (u = 0.5, v = 1.0), periodic unit square, level 1 50x50, two
finer levels at ratio 2, q = 1 inside the disk r < 0.2 about (0.5, 0.5);
C2 acoustics (rho = 1, K = 4) on [-1, 1]^2, extrapolation, level 1 50x50,
ratios 2 and 4, p = 1 + cos(pi r / 0.35) for r < 0.35, u = v = 0;
C3 Euler four-quadrant problem on the unit square, extrapolation, level 1
80x80 (four 40x40 grids, the reference's max1d split), ratios 2 and 4.

Finer levels come from flagging the initial data as flag2refine2 does
(largest centred difference above the configuration's tolerance), buffering
by ibuff = 3 cells, and covering the flagged cells with the tiles of a fixed
grid that contain flags (a stand-in for the reference's clustering, which
stays with its host driver), each level nested one coarse cell inside a
single grid of the level below.  Time steps: CFL 0.9 of the initial state
on level 1, possk divided by the ratios below.
"""
import numpy as np

from amrclaw_amd.hierarchy import Hier
from tests.common import quadrant_euler

CONFIGS = {
    "C1": dict(solver=1, nvar=1, nx=50, ratios=(2, 2), domain=(0.0, 1.0, 0.0, 1.0),
               mthbc=(2, 2, 2, 2), per=True, tol=0.05, speed=1.0),
    "C2": dict(solver=2, nvar=3, nx=50, ratios=(2, 4), domain=(-1.0, 1.0, -1.0, 1.0),
               mthbc=(1, 1, 1, 1), per=False, tol=0.02, speed=2.0),
    "C3": dict(solver=3, nvar=4, nx=80, ratios=(2, 4), domain=(0.0, 1.0, 0.0, 1.0),
               mthbc=(1, 1, 1, 1), per=False, tol=0.1, speed=None),
}


def initial_state(name, x, y):
    if name == "C1":
        r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2)
        return np.where(r < 0.2, 1.0, 0.0)[..., None]
    if name == "C2":
        r = np.sqrt(x * x + y * y)
        p = np.where(r < 0.35, 1.0 + np.cos(np.pi * r / 0.35), 1.0)
        return np.stack([p, np.zeros_like(p), np.zeros_like(p)], -1)
    return quadrant_euler(x, y)


def _fill(H, name, l, ng=2):
    L = H.levels[l - 1]
    xlo, _, ylo, _ = H.domain
    for k, b in enumerate(L["boxes"]):
        i = np.arange(b[0] - ng, b[1] + ng + 1)
        j = np.arange(b[2] - ng, b[3] + ng + 1)
        Y, X = np.meshgrid(ylo + (j + 0.5) * L["hy"], xlo + (i + 0.5) * L["hx"], indexing="ij")
        q = initial_state(name, X, Y)
        H.grid(l, k, 0)[:] = q
        H.grid(l, k, 1)[:] = q


def _flags(name, box, h, domain, tol):
    """Flagged cells of one box (flag2refine2's centred differences)."""
    i = np.arange(box[0] - 1, box[1] + 2)
    j = np.arange(box[2] - 1, box[3] + 2)
    Y, X = np.meshgrid(domain[2] + (j + 0.5) * h, domain[0] + (i + 0.5) * h, indexing="ij")
    q = initial_state(name, X, Y)
    dq = np.maximum(np.abs(q[1:-1, 2:] - q[1:-1, :-2]), np.abs(q[2:, 1:-1] - q[:-2, 1:-1]))
    return dq.max(axis=-1) > tol


def config_hierarchy(name, tile=8, ibuff=3):
    c = CONFIGS[name]
    nx = c["nx"]
    xlo, xhi, ylo, yhi = c["domain"]
    H = Hier(c["nvar"], domain=c["domain"], mthbc=c["mthbc"], xper=c["per"], yper=c["per"])
    h = (xhi - xlo) / nx
    half = nx // 2 if nx > 56 else nx
    boxes = [(a, a + half - 1, b, b + half - 1) for b in range(0, nx, half) for a in range(0, nx, half)]
    isz = nx
    levels = [(boxes, h, isz)]
    for r in c["ratios"]:
        boxes, h, isz = levels[-1]
        fine = []
        for b in boxes:
            f = _flags(name, b, h, c["domain"], c["tol"])
            if not f.any():
                continue
            # buffer by ibuff cells (flagger's bufnst)
            fb = np.zeros_like(f)
            jj, ii = np.nonzero(f)
            for dj in range(-ibuff, ibuff + 1):
                for di in range(-ibuff, ibuff + 1):
                    fb[np.clip(jj + dj, 0, f.shape[0] - 1), np.clip(ii + di, 0, f.shape[1] - 1)] = True
            # tiles of the box shrunk by one cell (proper nesting), kept if flagged
            a0, a1, b0, b1 = b[0] + 1, b[1] - 1, b[2] + 1, b[3] - 1
            for tj in range(b0, b1 + 1, tile):
                for ti in range(a0, a1 + 1, tile):
                    hi_i, hi_j = min(ti + tile - 1, a1), min(tj + tile - 1, b1)
                    if hi_i - ti < 1 or hi_j - tj < 1:
                        continue
                    if fb[tj - b[2]:hi_j - b[2] + 1, ti - b[0]:hi_i - b[0] + 1].any():
                        fine.append((ti * r, (hi_i + 1) * r - 1, tj * r, (hi_j + 1) * r - 1))
        levels.append((fine, h / r, isz * r))
    ratios = list(c["ratios"]) + [2]
    if c["speed"] is None:
        q = initial_state(name, np.array([0.1, 0.9, 0.1, 0.9]), np.array([0.1, 0.1, 0.9, 0.9]))
        rho, u, v = q[:, 0], q[:, 1] / q[:, 0], q[:, 2] / q[:, 0]
        p = 0.4 * (q[:, 3] - 0.5 * rho * (u * u + v * v))
        speed = float((np.maximum(np.abs(u), np.abs(v)) + np.sqrt(1.4 * p / rho)).max())
    else:
        speed = c["speed"]
    dt = 0.9 * levels[0][1] / speed
    for l, (boxes, h, isz) in enumerate(levels, start=1):
        H.add_level(boxes, h, h, dt, 0.0, isz, isz, ratios[l - 1], ratios[l - 1])
        dt = dt / ratios[l - 1]
    for l in range(1, H.nlev + 1):
        _fill(H, name, l)
    return H


def berger_oliger(impl, H, level=1, log=None):
    """One step of `level` and the subcycled steps of the finer levels with
    each level's own ratio (tick.f:230-300 with fixed time steps)."""
    r = impl.advanc(level)
    if log is not None:
        log.append((level, r))
    if level < H.nlev:
        for _ in range(H.levels[level - 1]["ratx"]):
            berger_oliger(impl, H, level + 1, log)
        impl.update(level)
