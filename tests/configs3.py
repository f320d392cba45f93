"""The 3-D run configuration as a static hierarchy for parity tests.

C5: 3-D Euler radial bump on [0, 2]^3 (rho = 1 + 10 exp(-20 r^2) with r the
distance from the corner at the origin, u = v = w = 0, E = rho), walls at
the lower sides and extrapolation at the upper sides, level 1 64^3 in eight
32^3 grids (patches of at most 32^3), two finer levels at ratio 2.

Finer levels come from flagging the initial data (largest centred pressure
difference above 0.05, the example's flag2refine criterion on the
pressure), buffering by 3 cells, and covering the flagged cells with the
tiles of a fixed grid (8^3 coarse cells) that contain flags, each level
nested one coarse cell inside a single grid of the level below.  The time
step is CFL 0.5 of the initial state on level 1, divided by 2 per level.
"""
import numpy as np

from amrclaw_amd.hierarchy import Hier3

GAMMA = 1.4


def radial_state(x, y, z):
    r = np.sqrt(x * x + y * y + z * z)
    rho = 1.0 + 10.0 * np.exp(-20.0 * r * r)
    zero = np.zeros_like(rho)
    return np.stack([rho, zero, zero, zero, rho], -1)


def _cells(box, h, lo, ng):
    i = np.arange(box[0] - ng, box[1] + ng + 1)
    j = np.arange(box[2] - ng, box[3] + ng + 1)
    k = np.arange(box[4] - ng, box[5] + ng + 1)
    return np.meshgrid(lo[2] + (k + 0.5) * h[2], lo[1] + (j + 0.5) * h[1], lo[0] + (i + 0.5) * h[0],
                       indexing="ij")


def _flags(box, h, lo, tol):
    Z, Y, X = _cells(box, h, lo, 1)
    q = radial_state(X, Y, Z)
    p = (GAMMA - 1.0) * (q[..., 4] - 0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 3] ** 2) / q[..., 0])
    d = np.maximum.reduce([np.abs(p[1:-1, 1:-1, 2:] - p[1:-1, 1:-1, :-2]),
                           np.abs(p[1:-1, 2:, 1:-1] - p[1:-1, :-2, 1:-1]),
                           np.abs(p[2:, 1:-1, 1:-1] - p[:-2, 1:-1, 1:-1])])
    return d > tol


def config_hierarchy3(n=64, nlev=3, tile=8, ibuff=3, tol=0.05, split=2):
    lo = (0.0, 0.0, 0.0)
    H = Hier3(5, domain=(0.0, 2.0, 0.0, 2.0, 0.0, 2.0), mthbc=(3, 1, 3, 1, 3, 1))
    h = 2.0 / n
    s = n // split
    boxes = [(a, a + s - 1, b, b + s - 1, c, c + s - 1)
             for c in range(0, n, s) for b in range(0, n, s) for a in range(0, n, s)]
    levels = [(boxes, h, n)]
    for _ in range(nlev - 1):
        boxes, h, sz = levels[-1]
        fine = []
        for b in boxes:
            f = _flags(b, (h, h, h), lo, tol)
            if not f.any():
                continue
            fb = np.zeros_like(f)
            kk, jj, ii = np.nonzero(f)
            for dk in range(-ibuff, ibuff + 1):
                for dj in range(-ibuff, ibuff + 1):
                    for di in range(-ibuff, ibuff + 1):
                        fb[np.clip(kk + dk, 0, f.shape[0] - 1), np.clip(jj + dj, 0, f.shape[1] - 1),
                           np.clip(ii + di, 0, f.shape[2] - 1)] = True
            # tiles of the box shrunk by one cell, except at the domain edge
            # (nesting is not needed across a physical boundary)
            a0 = b[0] if b[0] == 0 else b[0] + 1
            a1 = b[1] if b[1] == sz - 1 else b[1] - 1
            b0 = b[2] if b[2] == 0 else b[2] + 1
            b1 = b[3] if b[3] == sz - 1 else b[3] - 1
            c0 = b[4] if b[4] == 0 else b[4] + 1
            c1 = b[5] if b[5] == sz - 1 else b[5] - 1
            for tk in range(c0, c1 + 1, tile):
                for tj in range(b0, b1 + 1, tile):
                    for ti in range(a0, a1 + 1, tile):
                        hi = (min(ti + tile - 1, a1), min(tj + tile - 1, b1), min(tk + tile - 1, c1))
                        if min(hi[0] - ti, hi[1] - tj, hi[2] - tk) < 1:
                            continue
                        if fb[tk - b[4]:hi[2] - b[4] + 1, tj - b[2]:hi[1] - b[2] + 1,
                              ti - b[0]:hi[0] - b[0] + 1].any():
                            fine.append((2 * ti, 2 * hi[0] + 1, 2 * tj, 2 * hi[1] + 1,
                                         2 * tk, 2 * hi[2] + 1))
        levels.append((fine, h / 2, sz * 2))
    # CFL 0.5 of the initial state (zero velocity: the sound speed)
    cmax = float(np.sqrt(GAMMA * 0.4 * 11.0 / 11.0))
    dt = 0.5 * levels[0][1] / cmax
    for l, (boxes, hl, sz) in enumerate(levels, start=1):
        H.add_level(boxes, (hl, hl, hl), dt, 0.0, (sz, sz, sz), (2, 2, 2))
        dt = dt / 2
    for l in range(1, H.nlev + 1):
        L = H.levels[l - 1]
        for k, b in enumerate(L["boxes"]):
            Z, Y, X = _cells(b, (L["hx"],) * 3, lo, 2)
            q = radial_state(X, Y, Z)
            H.grid(l, k, 0)[:] = q
            H.grid(l, k, 1)[:] = q
    return H
