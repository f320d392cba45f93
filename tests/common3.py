"""3-D hierarchy builders shared by the 3-D level-update tests."""
import numpy as np

from amrclaw_amd.hierarchy import Hier3


def euler3_states(rng, shape, gamma=1.4):
    rho = 0.6 + rng.random(shape)
    u, v, w = (0.6 * rng.normal(size=shape) for _ in range(3))
    p = 0.4 + rng.random(shape)
    return np.stack([rho, rho * u, rho * v, rho * w,
                     p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)], -1)


def smooth_euler3(rng, shape, gamma=1.4):
    """Positive, smoothly varying 3-D Euler states (for time stepping)."""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.linspace(0, 1, nz), np.linspace(0, 1, ny), np.linspace(0, 1, nx),
                          indexing="ij")
    ph = rng.random(6) * 6.28
    rho = 1.0 + 0.3 * np.sin(3 * x + ph[0]) * np.cos(2 * y + ph[1]) * np.cos(2 * z + ph[2]) \
        + 0.05 * rng.random(shape)
    u = 0.4 * np.sin(2 * y + ph[3]) + 0.05 * rng.normal(size=shape)
    v = 0.4 * np.cos(2 * z + ph[4]) + 0.05 * rng.normal(size=shape)
    w = 0.4 * np.sin(2 * x + ph[5]) + 0.05 * rng.normal(size=shape)
    p = 1.0 + 0.2 * rng.random(shape)
    return np.stack([rho, rho * u, rho * v, rho * w,
                     p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)], -1)


def advection3_states(rng, shape):
    return rng.normal(size=shape + (1,))


def _overlap(b, c):
    return not (b[1] < c[0] or c[1] < b[0] or b[3] < c[2] or c[3] < b[2] or
                b[5] < c[4] or c[5] < b[4])


def disjoint_boxes3(rng, sz, nbox, lo=4, hi=8, align=(2, 2, 2), within=None, tries=2000):
    """Random disjoint boxes (ilo, ihi, jlo, jhi, klo, khi) inside
    [0, sz[0]) x [0, sz[1]) x [0, sz[2]) (or inside one of the boxes
    `within`), corners and sizes aligned per direction."""
    out = []
    for _ in range(tries):
        if len(out) >= nbox:
            break
        b = []
        ok = True
        for d in range(3):
            a = align[d]
            w = int(rng.integers(max(1, lo // a), hi // a + 1)) * a
            if w > sz[d]:
                ok = False
                break
            s = int(rng.integers(0, (sz[d] - w) // a + 1)) * a
            b += [s, s + w - 1]
        if not ok:
            continue
        b = tuple(b)
        if any(_overlap(b, c) for c in out):
            continue
        if within is not None and not any(all(c[2 * d] <= b[2 * d] and b[2 * d + 1] <= c[2 * d + 1]
                                              for d in range(3)) for c in within):
            continue
        out.append(b)
    return out


def level1_boxes(n, split):
    cuts = [np.linspace(0, n[d], split[d] + 1).astype(int) for d in range(3)]
    return [(cuts[0][a], cuts[0][a + 1] - 1, cuts[1][b], cuts[1][b + 1] - 1,
             cuts[2][c], cuts[2][c + 1] - 1)
            for c in range(split[2]) for b in range(split[1]) for a in range(split[0])]


def random_hierarchy3(rng, nvar=5, n=(12, 10, 8), nlev=3, ratios=((2, 2, 2), (2, 2, 2)),
                      mthbc=(1,) * 6, per=(False, False, False), fine_phase=1,
                      split=(2, 2, 1), nfine=4):
    """Level 1 tiles the domain; finer levels hold random disjoint boxes
    (not nested).  Times follow Berger-Oliger: levels below the finest have
    just been advanced to T = 1 from T - possk; the finest sits inside the
    last coarse step at fine_phase fine steps."""
    H = Hier3(nvar, mthbc=mthbc, xper=per[0], yper=per[1], zper=per[2])
    h = [1.0 / n[0], 1.0 / n[1], 1.0 / n[2]]
    sz = list(n)
    boxes = level1_boxes(n, split)
    possk, T = 0.08, 1.0
    for l in range(1, nlev + 1):
        r = ratios[l - 1] if l - 1 < len(ratios) else (2, 2, 2)
        if l > 1:
            rp = ratios[l - 2]
            sz = [sz[d] * rp[d] for d in range(3)]
            h = [h[d] / rp[d] for d in range(3)]
            possk = possk / max(rp)
            boxes = disjoint_boxes3(rng, sz, nfine, lo=2, hi=max(4, min(sz) // 2), align=rp)
        t = T if l < nlev else T - possk * max(ratios[l - 2]) + fine_phase * possk
        H.add_level(boxes, h, possk, t, sz, r)
    for L in H.levels:
        L["qnew"][:] = rng.normal(size=L["qnew"].shape)
        L["qold"][:] = rng.normal(size=L["qold"].shape)
    return H


def nested_hierarchy3(rng, states, nvar, n=(12, 10, 8), nlev=3, ratios=((2, 2, 2), (2, 2, 2)),
                      mthbc=(1,) * 6, per=(False, False, False), dt1=None, nfine=5,
                      split=(2, 2, 1), hi=10):
    """Properly nested 3-D hierarchy at t = 0 (old = new) for time stepping:
    every level-(l+1) grid, coarsened and grown by one coarse cell, lies
    inside one level-l grid or reaches a non-periodic domain edge where the
    level-l grid does."""
    H = Hier3(nvar, mthbc=mthbc, xper=per[0], yper=per[1], zper=per[2])
    h = [1.0 / n[0], 1.0 / n[1], 1.0 / n[2]]
    sz = list(n)
    boxes = level1_boxes(n, split)
    possk = dt1 if dt1 is not None else 0.15 * min(h)
    for l in range(1, nlev + 1):
        r = ratios[l - 1] if l - 1 < len(ratios) else (2, 2, 2)
        if l > 1:
            rp = ratios[l - 2]
            full = l - 1 == 1
            within = []
            for b in boxes:
                w = []
                for d in range(3):
                    edge_ok = full or not per[d]
                    lo_ = 0 if (b[2 * d] == 0 and edge_ok) else (b[2 * d] + 1) * rp[d]
                    hi_ = sz[d] * rp[d] - 1 if (b[2 * d + 1] == sz[d] - 1 and edge_ok) \
                        else b[2 * d + 1] * rp[d] - 1
                    w += [lo_, hi_]
                within.append(tuple(w))
            sz = [sz[d] * rp[d] for d in range(3)]
            h = [h[d] / rp[d] for d in range(3)]
            possk = possk / max(rp)
            boxes = disjoint_boxes3(rng, sz, nfine, lo=2, hi=hi, align=rp, within=within,
                                    tries=20000)
        H.add_level(boxes, h, possk, 0.0, sz, r)
    for l in range(1, nlev + 1):
        for k in range(len(H.levels[l - 1]["boxes"])):
            g = H.grid(l, k, 0)
            g[:] = states(rng, g.shape[:3])
            H.grid(l, k, 1)[:] = g
    return H


def berger_oliger3(impl, H, level=1, log=None):
    """One step of `level` and the subcycled steps of the finer levels; the
    subcycle count is the largest of the three ratios (the reference's
    kratio, setgrd/tick)."""
    r = impl.advanc(level)
    if log is not None:
        log.append((level, r))
    if level < H.nlev:
        L = H.levels[level - 1]
        for _ in range(max(L["ratx"], L["raty"], L["ratz"])):
            berger_oliger3(impl, H, level + 1, log)
        impl.update(level)
