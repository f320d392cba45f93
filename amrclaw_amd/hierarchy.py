"""A grid hierarchy in the reference's terms, held on the host: per level
the global index boxes (node(ndilo..ndjhi)), mesh widths (hxposs/hyposs),
step (possk), time (rnode(timemult)), index-space size (iregsz/jregsz) and
refinement ratio to the next level (intratx/intraty), plus packed AoS FP64
new/old arrays, one (my+4, mx+4, nvar) block per grid like the reference's
alloc(node(store1/store2)) (amr_module.f90:30-151).  amrclaw_amd.capi
installs one on the device with load_hierarchy and reads levels back with
read_level."""
import numpy as np


class Hier:
    """A grid hierarchy in the reference's terms: per level the global index
    boxes, mesh widths, step, time, index-space size and refinement ratio
    to the next level, plus packed AoS FP64 new/old solution arrays (one
    (my+4, mx+4, nvar) block per grid, like the reference's alloc)."""

    def __init__(self, nvar, domain=(0.0, 1.0, 0.0, 1.0), mthbc=(1, 1, 1, 1),
                 xper=False, yper=False, ng=2, naux=0):
        self.nvar, self.ng = nvar, ng
        self.naux = naux
        self.domain = domain
        self.mthbc = tuple(mthbc)
        self.xper, self.yper = xper, yper
        self.levels = []

    def add_level(self, boxes, hx, hy, possk, time, iregsz, jregsz, ratx=2, raty=2):
        boxes = np.asarray(boxes, dtype=np.int32).reshape(-1, 4)
        shapes = [(int(b[3] - b[2]) + 1 + 2 * self.ng, int(b[1] - b[0]) + 1 + 2 * self.ng, self.nvar)
                  for b in boxes]
        sizes = np.array([int(np.prod(s)) for s in shapes], dtype=np.int64)
        off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        tot = max(int(sizes.sum()), 1)
        self.levels.append(dict(boxes=np.ascontiguousarray(boxes), shapes=shapes, off=off,
                                qnew=np.zeros(tot), qold=np.zeros(tot), hx=hx, hy=hy,
                                possk=possk, time=time, iregsz=iregsz, jregsz=jregsz,
                                ratx=ratx, raty=raty))
        if self.naux > 0:
            # aux arrays (alloc(node(storeaux))): naux values per cell, the
            # same cells as q
            asz = (sizes // self.nvar) * self.naux
            L = self.levels[-1]
            L["auxoff"] = np.concatenate([[0], np.cumsum(asz)[:-1]]).astype(np.int64)
            L["aux"] = np.zeros(max(int(asz.sum()), 1))

    @property
    def nlev(self):
        return len(self.levels)

    def grid(self, l, k, which=0):
        """View of grid k of level l (1-based) as (my+4, mx+4, nvar)."""
        L = self.levels[l - 1]
        a = L["qnew"] if which == 0 else L["qold"]
        n = int(np.prod(L["shapes"][k]))
        return a[L["off"][k]:L["off"][k] + n].reshape(L["shapes"][k])

    def aux(self, l, k):
        """View of grid k's aux array on level l as (my+4, mx+4, naux)."""
        L = self.levels[l - 1]
        sh = L["shapes"][k][:2] + (self.naux,)
        n = int(np.prod(sh))
        return L["aux"][L["auxoff"][k]:L["auxoff"][k] + n].reshape(sh)

    def copy(self):
        import copy
        H = copy.copy(self)
        H.levels = [dict(L, qnew=L["qnew"].copy(), qold=L["qold"].copy()) for L in self.levels]
        return H


class Hier3:
    """A 3-D grid hierarchy in the reference's terms (src/3d/amr_module.f90):
    per level the global index boxes (ilo, ihi, jlo, jhi, klo, khi), mesh
    widths, step, time, index-space size and refinement ratios to the next
    level, plus packed AoS FP64 new/old arrays, one (mz+4, my+4, mx+4, nvar)
    block per grid.  mthbc = (x-lo, x-hi, y-lo, y-hi, z-lo, z-hi); a
    direction is periodic when its flag is set."""

    def __init__(self, nvar, domain=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), mthbc=(1,) * 6,
                 xper=False, yper=False, zper=False, ng=2):
        self.nvar, self.ng = nvar, ng
        self.domain = tuple(domain)
        self.mthbc = tuple(mthbc)
        self.xper, self.yper, self.zper = xper, yper, zper
        self.levels = []

    def add_level(self, boxes, h, possk, time, regsz, ratios=(2, 2, 2)):
        """boxes: (n, 6) global index boxes; h = (hx, hy, hz); regsz = (iregsz,
        jregsz, kregsz); ratios = refinement to the next finer level."""
        boxes = np.asarray(boxes, dtype=np.int32).reshape(-1, 6)
        ng = self.ng
        shapes = [(int(b[5] - b[4]) + 1 + 2 * ng, int(b[3] - b[2]) + 1 + 2 * ng,
                   int(b[1] - b[0]) + 1 + 2 * ng, self.nvar) for b in boxes]
        sizes = np.array([int(np.prod(s)) for s in shapes], dtype=np.int64)
        off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        tot = max(int(sizes.sum()), 1)
        self.levels.append(dict(boxes=np.ascontiguousarray(boxes), shapes=shapes, off=off,
                                qnew=np.zeros(tot), qold=np.zeros(tot), hx=h[0], hy=h[1],
                                hz=h[2], possk=possk, time=time, iregsz=regsz[0],
                                jregsz=regsz[1], kregsz=regsz[2], ratx=ratios[0],
                                raty=ratios[1], ratz=ratios[2]))

    @property
    def nlev(self):
        return len(self.levels)

    def grid(self, l, k, which=0):
        """View of grid k of level l (1-based) as (mz+4, my+4, mx+4, nvar)."""
        L = self.levels[l - 1]
        a = L["qnew"] if which == 0 else L["qold"]
        n = int(np.prod(L["shapes"][k]))
        return a[L["off"][k]:L["off"][k] + n].reshape(L["shapes"][k])

    def copy(self):
        import copy
        H = copy.copy(self)
        H.levels = [dict(L, qnew=L["qnew"].copy(), qold=L["qold"].copy()) for L in self.levels]
        return H
