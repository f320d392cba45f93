"""Synthetic workloads for the level update (batched 2-D Euler).

The batched Euler configuration advances one level of N patches whose sizes
mx, my are drawn independently and uniformly from the even numbers in
[8, 128] (fine grids are whole coarse cells at refinement ratio 2), with
numpy default_rng(seed).  Patches are shelf-packed without overlap into a
square index space on the unit square; the state is the four-quadrant
Riemann problem of examples/euler_2d_quadrants/qinit.f (rounded states,
split at (0.8, 0.8)) sampled at cell centres, times a seeded +-1e-3 relative
perturbation.
"""
import numpy as np

from .hierarchy import Hier, Hier3

GAMMA = 1.4
QUADRANTS = {1: (1.5, 0.0, 0.0, 1.5), 2: (0.5323, 1.206, 0.0, 0.3),
             3: (0.138, 1.206, 1.206, 0.029), 4: (0.5323, 0.0, 1.206, 0.3)}


def quadrant_state(x, y, gamma=GAMMA, xs=0.8, ys=0.8):
    """Conserved (rho, rho u, rho v, E) of the four-quadrant problem."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    iq = np.where(x >= xs, np.where(y >= ys, 1, 4), np.where(y >= ys, 2, 3))
    q = np.zeros(x.shape + (4,))
    for k, (r, u, v, p) in QUADRANTS.items():
        m = iq == k
        q[m, 0] = r
        q[m, 1] = r * u
        q[m, 2] = r * v
        q[m, 3] = p / (gamma - 1.0) + 0.5 * r * (u * u + v * v)
    return q


def patch_sizes(n, seed=20261202, lo=8, hi=128):
    rng = np.random.default_rng(seed)
    mx = 2 * rng.integers(lo // 2, hi // 2 + 1, size=n)
    my = 2 * rng.integers(lo // 2, hi // 2 + 1, size=n)
    return mx.astype(np.int64), my.astype(np.int64)


def shelf_pack(mx, my, width):
    """Place boxes left to right in shelves of height max(my of the shelf).
    Returns (n, 4) int array of (ilo, ihi, jlo, jhi) and the used height."""
    boxes = np.zeros((len(mx), 4), dtype=np.int64)
    x = y = shelf_h = 0
    for k in range(len(mx)):
        if x + mx[k] > width:
            y += shelf_h
            x = shelf_h = 0
        boxes[k] = (x, x + mx[k] - 1, y, y + my[k] - 1)
        x += mx[k]
        shelf_h = max(shelf_h, my[k])
    return boxes, y + shelf_h


def batched_euler_level(n=4096, seed=20261202, width=None):
    """Layout of the batched Euler level: boxes (global indices), mesh width
    h (index space width -> unit square), sizes."""
    mx, my = patch_sizes(n, seed)
    if width is None:
        # square-ish packing: area / fill factor of the shelves
        area = float((mx * my).sum())
        width = int(np.ceil(np.sqrt(area / 0.55) / 64.0) * 64)
    boxes, height = shelf_pack(mx, my, width)
    side = max(width, height)
    return boxes, 1.0 / side, mx, my


def patch_initial_state(box, h, seed, nghost=2, perturb=1e-3):
    """Initial q of one patch including ghost cells, shape (my+4, mx+4, 4)."""
    ilo, ihi, jlo, jhi = [int(v) for v in box]
    i = np.arange(ilo - nghost, ihi + nghost + 1)
    j = np.arange(jlo - nghost, jhi + nghost + 1)
    xc = (i + 0.5) * h
    yc = (j + 0.5) * h
    X, Y = np.meshgrid(xc, yc)
    q = quadrant_state(X, Y)
    rng = np.random.default_rng(seed)
    q *= 1.0 + perturb * (2.0 * rng.random(q.shape[:2] + (1,)) - 1.0)
    return q


def max_wave_speed(q, gamma=GAMMA):
    """max(|u| + c, |v| + c) over the cells of an Euler array (..., 4)."""
    rho = q[..., 0]
    u, v = q[..., 1] / rho, q[..., 2] / rho
    p = (gamma - 1.0) * (q[..., 3] - 0.5 * rho * (u * u + v * v))
    c = np.sqrt(gamma * p / rho)
    return float(max((np.abs(u) + c).max(), (np.abs(v) + c).max()))


def batched_euler_hierarchy(n=4096, seed=20261202, coarse_block=256, courant=0.8):
    """Two-level hierarchy for the level-update benchmark: level 2 holds the
    n batched patches (batched_euler_level), level 1 covers the whole domain
    at ratio 2 with grids of coarse_block^2 cells, so every level-2 grid is
    properly nested.  Level 1 has been advanced from t = 0 to dt1 (old and
    new hold the initial state); level 2 sits half way through that coarse
    step (time dt2 = dt1 / 2), so its ghost fill blends both coarse times.
    dt2 gives the level-2 Courant number `courant` on the initial state.
    Returns the Hier (amrclaw_amd.hierarchy) and the interior cell count of
    level 2."""
    from .hierarchy import Hier
    boxes2, h2, mx, my = batched_euler_level(n, seed)
    side = int(round(1.0 / h2))
    csz = side // 2
    h1 = 2.0 * h2
    cuts = list(range(0, csz, coarse_block)) + [csz]
    boxes1 = [(a, b - 1, c, d - 1) for c, d in zip(cuts[:-1], cuts[1:])
              for a, b in zip(cuts[:-1], cuts[1:])]
    H = Hier(4)
    H.add_level(boxes1, h1, h1, 0.0, 0.0, csz, csz, 2, 2)
    H.add_level(boxes2, h2, h2, 0.0, 0.0, side, side, 2, 2)
    smax = 0.0
    for l, (boxes, h, s0) in enumerate(((boxes1, h1, 500000), (boxes2, h2, 1000)), start=1):
        for k, b in enumerate(boxes):
            g = H.grid(l, k, 0)
            g[:] = patch_initial_state(b, h, s0 + k)
            H.grid(l, k, 1)[:] = g
            if l == 2:
                smax = max(smax, max_wave_speed(g))
    dt2 = courant * h2 / smax
    H.levels[0].update(possk=2.0 * dt2, time=2.0 * dt2)
    H.levels[1].update(possk=dt2, time=dt2)
    return H, int((mx * my).sum())


# ---------------------------------------------------------------------------
# 3-D: the radial-bump Euler problem (rho = p = 2 inside r < 0.2 about the
# centre of the unit cube, 1 outside, at rest), batched over grids of uneven
# sizes.

def bump3_state(x, y, z, gamma=GAMMA, r0=0.2):
    """Conserved (rho, rho u, rho v, rho w, E) of the radial bump."""
    r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
    inside = r < r0
    rho = np.where(inside, 2.0, 1.0)
    p = np.where(inside, 2.0, 1.0)
    q = np.zeros(np.shape(x) + (5,))
    q[..., 0] = rho
    q[..., 4] = p / (gamma - 1.0)
    return q


def patch_sizes3(n, seed=20261203, lo=8, hi=32):
    rng = np.random.default_rng(seed)
    s = 2 * rng.integers(lo // 2, hi // 2 + 1, size=(n, 3))
    return s.astype(np.int64)


def batched_euler3_level(n=512, seed=20261203, fine=256, perturb=1e-3):
    """A level of n 3-D grids with sizes even U[8, 32]^3 on a fine index
    space of fine^3 cells over the unit cube (mesh width h = 1/fine).  Each
    grid sits at a seeded random corner and holds the radial bump sampled at
    its cell centres (ghost cells included) times a seeded +-perturb relative
    perturbation.  Returns (sizes (n, 3), corners (n, 3), h, list of q
    arrays (mz+4, my+4, mx+4, 5))."""
    sizes = patch_sizes3(n, seed)
    rng = np.random.default_rng(seed + 1)
    h = 1.0 / fine
    corners = np.array([[int(rng.integers(0, fine - int(v) + 1)) for v in s] for s in sizes])
    qs = []
    for k in range(n):
        mx, my, mz = (int(v) for v in sizes[k])
        i0, j0, k0 = (int(v) for v in corners[k])
        xc = (np.arange(i0 - 2, i0 + mx + 2) + 0.5) * h
        yc = (np.arange(j0 - 2, j0 + my + 2) + 0.5) * h
        zc = (np.arange(k0 - 2, k0 + mz + 2) + 0.5) * h
        Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
        q = bump3_state(X, Y, Z)
        prng = np.random.default_rng(seed + 100 + k)
        q *= 1.0 + perturb * (2.0 * prng.random(q.shape[:3] + (1,)) - 1.0)
        qs.append(q)
    return sizes, corners, h, qs


def max_wave_speed3(q, gamma=GAMMA):
    rho = q[..., 0]
    u, v, w = q[..., 1] / rho, q[..., 2] / rho, q[..., 3] / rho
    p = (gamma - 1.0) * (q[..., 4] - 0.5 * rho * (u * u + v * v + w * w))
    c = np.sqrt(gamma * p / rho)
    return float(max((np.abs(u) + c).max(), (np.abs(v) + c).max(), (np.abs(w) + c).max()))


# ------------------------------------------------------ 2-D run configurations
# C1 advection (u = 0.5, v = 1.0), periodic unit square, level 1 50x50, two
# finer levels at ratio 2, q = 1 inside the disk r < 0.2 about (0.5, 0.5);
# C2 acoustics (rho = 1, K = 4) on [-1, 1]^2, extrapolation, level 1 50x50,
# ratios 2 and 4, p = 1 + cos(pi r / 0.35) for r < 0.35, u = v = 0; C3 Euler
# four-quadrant problem on the unit square, extrapolation, level 1 80x80
# (four 40x40 grids), ratios 2 and 4.  Finer levels come from flagging the
# initial data as flag2refine2 does (largest centred difference above the
# configuration's tolerance), buffering by ibuff = 3 cells, and covering the
# flagged cells with the tiles of a fixed grid that contain flags, each
# level nested one coarse cell inside a single grid of the level below.
# Time steps: CFL 0.9 of the initial state on level 1, possk divided by the
# ratios below; t_end the configurations' end times.

CONFIGS = {
    "C1": dict(solver=1, nvar=1, nx=50, ratios=(2, 2), domain=(0.0, 1.0, 0.0, 1.0),
               mthbc=(2, 2, 2, 2), per=True, tol=0.05, speed=1.0, t_end=1.0),
    "C2": dict(solver=2, nvar=3, nx=50, ratios=(2, 4), domain=(-1.0, 1.0, -1.0, 1.0),
               mthbc=(1, 1, 1, 1), per=False, tol=0.02, speed=2.0, t_end=0.6),
    "C3": dict(solver=3, nvar=4, nx=80, ratios=(2, 4), domain=(0.0, 1.0, 0.0, 1.0),
               mthbc=(1, 1, 1, 1), per=False, tol=0.1, speed=None, t_end=0.8),
}


def config_state(name, x, y):
    if name == "C1":
        r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2)
        return np.where(r < 0.2, 1.0, 0.0)[..., None]
    if name == "C2":
        r = np.sqrt(x * x + y * y)
        p = np.where(r < 0.35, 1.0 + np.cos(np.pi * r / 0.35), 1.0)
        return np.stack([p, np.zeros_like(p), np.zeros_like(p)], -1)
    return quadrant_state(x, y)


def _config_fill(H, name, l, ng=2):
    L = H.levels[l - 1]
    xlo, _, ylo, _ = H.domain
    for k, b in enumerate(L["boxes"]):
        i = np.arange(b[0] - ng, b[1] + ng + 1)
        j = np.arange(b[2] - ng, b[3] + ng + 1)
        Y, X = np.meshgrid(ylo + (j + 0.5) * L["hy"], xlo + (i + 0.5) * L["hx"], indexing="ij")
        q = config_state(name, X, Y)
        H.grid(l, k, 0)[:] = q
        H.grid(l, k, 1)[:] = q


def _config_flags(name, box, h, domain, tol):
    """Flagged cells of one box (flag2refine2's centred differences)."""
    i = np.arange(box[0] - 1, box[1] + 2)
    j = np.arange(box[2] - 1, box[3] + 2)
    Y, X = np.meshgrid(domain[2] + (j + 0.5) * h, domain[0] + (i + 0.5) * h, indexing="ij")
    q = config_state(name, X, Y)
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
            f = _config_flags(name, b, h, c["domain"], c["tol"])
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
        q = config_state(name, np.array([0.1, 0.9, 0.1, 0.9]), np.array([0.1, 0.1, 0.9, 0.9]))
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
        _config_fill(H, name, l)
    return H


# ---------------------------------------------------------------- 3-D AMR --
# The 3-D Euler run configuration as a static three-level hierarchy: the
# radial bump of bump3_state (rho = p = 2 inside r < 0.2 about the centre of
# the unit cube, 1 outside, at rest), extrapolation at every side, level 1
# 64^3 in eight 32^3 grids, two finer levels at ratio 2 covering the cells
# flag2refine's test flags (largest centred difference of any component
# above `tol`), buffered by 3 cells, with tiles of 8^3 coarse cells, each
# nested one coarse cell inside a single grid of the level below except at
# the domain edge.  The coarse step is t_end / nsteps (t_end = 0.2: initial
# Courant number 0.58 on level 1), halved per level.

C5_T_END = 0.2
C5_STEPS = 26


def _cells3(box, h, ng):
    i = np.arange(box[0] - ng, box[1] + ng + 1)
    j = np.arange(box[2] - ng, box[3] + ng + 1)
    k = np.arange(box[4] - ng, box[5] + ng + 1)
    return np.meshgrid((k + 0.5) * h, (j + 0.5) * h, (i + 0.5) * h, indexing="ij")


def _flags3(box, h, tol):
    """flag2refine.f's test on the interior of `box`: any component's
    centred difference in any direction above tol."""
    Z, Y, X = _cells3(box, h, 1)
    q = bump3_state(X, Y, Z)
    d = np.maximum.reduce([np.abs(q[1:-1, 1:-1, 2:] - q[1:-1, 1:-1, :-2]),
                           np.abs(q[1:-1, 2:, 1:-1] - q[1:-1, :-2, 1:-1]),
                           np.abs(q[2:, 1:-1, 1:-1] - q[:-2, 1:-1, 1:-1])])
    return d.max(axis=-1) > tol


def radial_euler3_hierarchy(n=64, nlev=3, tile=8, ibuff=3, tol=0.1, split=2,
                            t_end=C5_T_END, nsteps=C5_STEPS):
    H = Hier3(5, domain=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), mthbc=(1,) * 6)
    h = 1.0 / n
    s = n // split
    boxes = [(a, a + s - 1, b, b + s - 1, c, c + s - 1)
             for c in range(0, n, s) for b in range(0, n, s) for a in range(0, n, s)]
    levels = [(boxes, h, n)]
    for _ in range(nlev - 1):
        boxes, h, sz = levels[-1]
        fine = []
        for b in boxes:
            f = _flags3(b, h, tol)
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
    dt = t_end / nsteps
    for l, (boxes, hl, sz) in enumerate(levels, start=1):
        H.add_level(boxes, (hl, hl, hl), dt, 0.0, (sz, sz, sz), (2, 2, 2))
        dt = dt / 2
    for l in range(1, H.nlev + 1):
        L = H.levels[l - 1]
        for k, b in enumerate(L["boxes"]):
            Z, Y, X = _cells3(b, L["hx"], 2)
            q = bump3_state(X, Y, Z)
            H.grid(l, k, 0)[:] = q
            H.grid(l, k, 1)[:] = q
    return H
