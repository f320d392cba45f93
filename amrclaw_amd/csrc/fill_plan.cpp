// fill_plan.cpp -- host planner of the ghost-cell fill programs.
//
// Replays bound.f90:51-158 for every grid of a level on index boxes only:
// the four ghost strips, filrecur (filpatch.f90:71-355) with intfil's
// coverage marks (intfil.f90:98-311), trimbd's unset rectangle
// (trimbd.f90:32-97), the coarse patch formula (filpatch.f90:205-210), the
// recursion to coarser levels, prefilrecur's periodic split
// (prefilp.f90:18-214) and bc2amr's side extents (bc2amr.f90:104-257).
// Every array index a generated operation will touch on the device is
// checked here, so a program that would read or write outside a buffer is
// rejected on the host instead of launched.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "fill_plan.hpp"

namespace amrgpu {

void BoxIndex::build(const std::vector<std::array<int, 4>>& boxes, int iregsz, int jregsz) {
    const int n = (int)boxes.size();
    start.clear();
    items.clear();
    nbx = nby = 0;
    if (n == 0 || iregsz < 1 || jregsz < 1) return;
    double area = 0;
    for (const auto& b : boxes) area += double(b[1] - b[0] + 1) * double(b[3] - b[2] + 1);
    bs = std::max(8, (int)std::ceil(std::sqrt(area / n)));
    nbx = (iregsz + bs - 1) / bs;
    nby = (jregsz + bs - 1) / bs;
    std::vector<int> count((size_t)nbx * nby + 1, 0);
    auto range = [&](const std::array<int, 4>& b, int& bx0, int& bx1, int& by0, int& by1) {
        bx0 = std::min(std::max(b[0], 0) / bs, nbx - 1);
        bx1 = std::min(std::max(b[1], 0) / bs, nbx - 1);
        by0 = std::min(std::max(b[2], 0) / bs, nby - 1);
        by1 = std::min(std::max(b[3], 0) / bs, nby - 1);
    };
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            start.assign((size_t)nbx * nby + 1, 0);
            for (size_t b = 0; b + 1 < count.size(); b++) start[b + 1] = start[b] + count[b];
            items.assign(start.back(), 0);
            std::fill(count.begin(), count.end(), 0);
        }
        for (int k = 0; k < n; k++) {
            int bx0, bx1, by0, by1;
            range(boxes[k], bx0, bx1, by0, by1);
            for (int by = by0; by <= by1; by++)
                for (int bx = bx0; bx <= bx1; bx++) {
                    const size_t b = (size_t)by * nbx + bx;
                    if (pass == 0) count[b]++;
                    else items[start[b] + count[b]++] = k;
                }
        }
    }
}

void BoxIndex::query(int ilo, int ihi, int jlo, int jhi, std::vector<int>& out) const {
    out.clear();
    if (nbx == 0 || ihi < ilo || jhi < jlo) return;
    const int bx0 = std::min(std::max(ilo, 0) / bs, nbx - 1), bx1 = std::min(std::max(ihi, 0) / bs, nbx - 1);
    const int by0 = std::min(std::max(jlo, 0) / bs, nby - 1), by1 = std::min(std::max(jhi, 0) / bs, nby - 1);
    for (int by = by0; by <= by1; by++)
        for (int bx = bx0; bx <= bx1; bx++) {
            const size_t b = (size_t)by * nbx + bx;
            out.insert(out.end(), items.begin() + start[b], items.begin() + start[b + 1]);
        }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

namespace {

struct Planner {
    const HierGeom& H;
    FillPlan& P;
    int nv;
    std::string err;
    std::vector<std::vector<int>> patch_buf;   // [level][grid] -> buffer id
    std::vector<BoxIndex> index;               // [level]: the grids a rectangle meets
    std::vector<int> cand;

    Planner(const HierGeom& h, FillPlan& p, int nvar) : H(h), P(p), nv(nvar) {
        patch_buf.resize(H.nlev + 1);
        index.resize(H.nlev + 1);
        for (int l = 1; l <= H.nlev; l++) {
            patch_buf[l].assign(H.lev[l].boxes.size(), -1);
            index[l].build(H.lev[l].boxes, H.lev[l].iregsz, H.lev[l].jregsz);
        }
    }

    // the grids of `level` that can meet [ilo..ihi] x [jlo..jhi], ascending
    // (a superset: callers intersect each box themselves)
    const std::vector<int>& candidates(int level, int ilo, int ihi, int jlo, int jhi) {
        index[level].query(ilo, ihi, jlo, jhi, cand);
        return cand;
    }

    int pbuf(int level, int grid) {
        int& b = patch_buf[level][grid];
        if (b < 0) {
            const auto& bx = H.lev[level].boxes[grid];
            FillBuf fb{};
            fb.kind = 0;
            fb.level = level;
            fb.patch = grid;
            fb.nrow = bx[1] - bx[0] + 1 + 2 * NGHOST;
            fb.ncol = bx[3] - bx[2] + 1 + 2 * NGHOST;
            fb.plane = fb.nrow * fb.ncol;
            b = (int)P.bufs.size();
            P.bufs.push_back(fb);
        }
        return b;
    }

    int scratch(int nrow, int ncol) {
        FillBuf fb{};
        fb.kind = 1;
        fb.level = 0;
        fb.patch = -1;
        fb.nrow = nrow;
        fb.ncol = ncol;
        fb.plane = nrow * ncol;
        fb.off = P.scratch_elems;
        P.scratch_elems += ((int64_t)nv * fb.plane + 31) / 32 * 32;
        P.bufs.push_back(fb);
        return (int)P.bufs.size() - 1;
    }

    bool fail(const char* msg) {
        if (err.empty()) err = msg;
        return false;
    }

    void push(FillOp op) { P.ops.push_back(op); }

    // bc2amr on a whole buffer with the given physical extent
    bool bc_op(int buf, double hx, double hy, double xlo, double xhi, double ylo, double yhi,
               int stage) {
        if (H.xperdom && (H.yperdom || H.spheredom)) return true;   // bc2amr.f90:109
        const FillBuf& fb = P.bufs[buf];
        const int nrow = fb.nrow, ncol = fb.ncol;
        FillOp op{};
        op.type = FOP_BC;
        op.dst = buf;
        op.ni = nrow;
        op.nj = ncol;
        op.stage = stage;
        const double hxmarg = hx * .01, hymarg = hy * .01;
        bool any = false;
        if (xlo < H.xlower - hxmarg) {
            int nxl = (int)((H.xlower + hxmarg - xlo) / hx);
            int mb = H.mthbc[0];
            if (mb != 1 && mb != 2 && mb != 3 && mb != 4) return fail("bc2amr: user/unknown BC on the left");
            // a buffer lying wholly outside the domain on this side has no
            // source column; the reference reads past the array there and the
            // values are overwritten by the grid's own bc2amr, so skip it
            bool fits = (mb == 1 && nxl < nrow) || (mb == 3 && 2 * nxl <= nrow);
            if (nxl > 0 && (mb == 1 || mb == 3) && fits) {
                op.bc[0] = nxl;
                any = true;
            }
        }
        if (xhi > H.xupper + hxmarg) {
            int nxr = (int)((xhi - H.xupper + hxmarg) / hx);
            int ibeg = std::max(nrow - nxr + 1, 1);
            int mb = H.mthbc[1];
            if (mb != 1 && mb != 2 && mb != 3 && mb != 4) return fail("bc2amr: user/unknown BC on the right");
            bool fits = (mb == 1 && ibeg >= 2) || (mb == 3 && 2 * ibeg - 1 - nrow >= 1);
            if (ibeg <= nrow && (mb == 1 || mb == 3) && fits) {
                op.bc[1] = ibeg;
                any = true;
            }
        }
        if (ylo < H.ylower - hymarg) {
            int nyb = (int)((H.ylower + hymarg - ylo) / hy);
            int mb = H.mthbc[2];
            if (mb != 1 && mb != 2 && mb != 3 && mb != 4) return fail("bc2amr: user/unknown BC at the bottom");
            bool fits = (mb == 1 && nyb < ncol) || (mb == 3 && 2 * nyb <= ncol);
            if (nyb > 0 && (mb == 1 || mb == 3) && fits) {
                op.bc[2] = nyb;
                any = true;
            }
        }
        if (yhi > H.yupper + hymarg) {
            int nyt = (int)((yhi - H.yupper + hymarg) / hy);
            int jbeg = std::max(ncol - nyt + 1, 1);
            int mb = H.mthbc[3];
            if (mb != 1 && mb != 2 && mb != 3 && mb != 4) return fail("bc2amr: user/unknown BC at the top");
            bool fits = (mb == 1 && jbeg >= 2) || (mb == 3 && 2 * jbeg - 1 - ncol >= 1);
            if (jbeg <= ncol && (mb == 1 || mb == 3) && fits) {
                op.bc[3] = jbeg;
                any = true;
            }
        }
        if (any) push(op);
        return true;
    }

    // filrecur: fill rectangle [ilo..ihi]x[jlo..jhi] of `level` into buffer
    // `buf` starting at its 1-based cell (nrowst, ncolst).  Returns the stage
    // after which the rectangle (and, with patchOnly, the buffer) is final.
    int filrecur(int level, int buf, int nrowst, int ncolst, int ilo, int ihi, int jlo, int jhi,
                 bool patchOnly, int msrc = -1) {
        if (!err.empty()) return -1;
        const LevelGeom& G = H.lev[level];
        const int ni = ihi - ilo + 1, nj = jhi - jlo + 1;
        const FillBuf fb = P.bufs[buf];
        if (nrowst - 1 < 0 || ncolst - 1 < 0 || nrowst - 1 + ni > fb.nrow || ncolst - 1 + nj > fb.ncol) {
            fail("filrecur: rectangle outside its buffer");
            return -1;
        }
        std::vector<unsigned char> flag((size_t)ni * nj, 0);
        // intfil: same-level grids intersecting the rectangle, in grid order
        const std::vector<int> ks = candidates(level, ilo, ihi, jlo, jhi);
        for (int k : ks) {
            const auto& b = G.boxes[k];
            if (msrc >= 0) {
                // a grid's own strips take same-level values only from its
                // bndList neighbours: the grids meeting it grown by nghost
                // (init_bndryList.f:69-97, intfil.f90:153-169)
                const auto& m = G.boxes[msrc];
                if (k == msrc || std::max(b[0], m[0] - NGHOST) > std::min(b[1], m[1] + NGHOST) ||
                    std::max(b[2], m[2] - NGHOST) > std::min(b[3], m[3] + NGHOST))
                    continue;
            }
            int ixlo = std::max(b[0], ilo), ixhi = std::min(b[1], ihi);
            int jxlo = std::max(b[2], jlo), jxhi = std::min(b[3], jhi);
            if (!(ixlo <= ixhi && jxlo <= jxhi)) continue;
            FillOp op{};
            op.type = FOP_COPY;
            op.dst = buf;
            op.src = k;
            op.src_level = level;
            op.di0 = ixlo - ilo + nrowst - 1;
            op.dj0 = jxlo - jlo + ncolst - 1;
            op.ni = ixhi - ixlo + 1;
            op.nj = jxhi - jxlo + 1;
            op.si0 = ixlo - b[0] + NGHOST;
            op.sj0 = jxlo - b[2] + NGHOST;
            op.stage = 0;
            push(op);
            P.src_levels_mask |= 1u << level;
            for (int j = jxlo; j <= jxhi; j++)
                for (int i = ixlo; i <= ixhi; i++) flag[(size_t)(j - jlo) * ni + (i - ilo)] = 1;
        }
        // cells outside the domain are left to the boundary conditions
        auto mark = [&](int i0, int i1, int j0, int j1) {
            i0 = std::max(i0, ilo); i1 = std::min(i1, ihi);
            j0 = std::max(j0, jlo); j1 = std::min(j1, jhi);
            for (int j = j0; j <= j1; j++)
                for (int i = i0; i <= i1; i++) flag[(size_t)(j - jlo) * ni + (i - ilo)] = 1;
        };
        if (jhi >= G.jregsz) mark(ilo, ihi, std::max(G.jregsz, jlo), jhi);
        if (jlo < 0) mark(ilo, ihi, jlo, std::min(-1, ncolst + jhi - jlo));
        if (ilo < 0) mark(ilo, std::min(-1, nrowst + ihi - ilo), jlo, jhi);
        if (ihi >= G.iregsz) mark(std::max(G.iregsz, ilo), ihi, jlo, jhi);
        // trimbd
        int un[4] = {0, 0, 0, 0};
        bool set = true;
        {
            int imn = ni + 1, imx = 0, jmn = nj + 1, jmx = 0;
            for (int j = 1; j <= nj; j++)
                for (int i = 1; i <= ni; i++)
                    if (!flag[(size_t)(j - 1) * ni + (i - 1)]) {
                        set = false;
                        imn = std::min(imn, i); imx = std::max(imx, i);
                        jmn = std::min(jmn, j); jmx = std::max(jmx, j);
                    }
            un[0] = imn; un[1] = imx; un[2] = jmn; un[3] = jmx;
        }
        int ready = 0;
        if (!set) {
            if (level == 1) {
                fail("filrecur: level 1 cells left unset");
                return -1;
            }
            const LevelGeom& C = H.lev[level - 1];
            un[0] += ilo - 1; un[1] += ilo - 1; un[2] += jlo - 1; un[3] += jlo - 1;
            const int rx = C.ratx, ry = C.raty, ng = NGHOST;
            int iplo = (un[0] - rx + ng * rx) / rx - ng;
            int iphi = (un[1] + rx) / rx;
            int jplo = (un[2] - ry + ng * ry) / ry - ng;
            int jphi = (un[3] + ry) / ry;
            int mic = iphi - iplo + 1, mjc = jphi - jplo + 1;
            if (mic > ihi - ilo + 3 || mjc > jhi - jlo + 3) {
                fail("filrecur: coarse scratch too small");
                return -1;
            }
            int cbuf = scratch(mic, mjc);
            bool sticks = iplo < 0 || jplo < 0 || iphi >= C.iregsz || jphi >= C.jregsz;
            int rc;
            if ((H.xperdom || H.yperdom || H.spheredom) && sticks)
                rc = prefil(level - 1, cbuf, iplo, iphi, jplo, jphi, iplo, jplo);
            else
                rc = filrecur(level - 1, cbuf, 1, 1, iplo, iphi, jplo, jphi, true);
            if (rc < 0) return -1;
            // interpolation reads coarse cells ic-1..ic+1, jc-1..jc+1 of the
            // scratch for every unset fine cell: check them all
            for (int j = 1; j <= nj; j++)
                for (int i = 1; i <= ni; i++) {
                    if (flag[(size_t)(j - 1) * ni + (i - 1)]) continue;
                    int jc = (int)std::floor((double)(j + jlo - 1) / (double)ry) - jplo + 1;
                    int ic = (int)std::floor((double)(i + ilo - 1) / (double)rx) - iplo + 1;
                    if (ic - 1 < 1 || ic + 1 > mic || jc - 1 < 1 || jc + 1 > mjc) {
                        fail("filrecur: interpolation stencil outside the coarse patch");
                        return -1;
                    }
                }
            FillOp op{};
            op.type = FOP_INTERP;
            op.dst = buf;
            op.src = cbuf;
            op.di0 = nrowst - 1;
            op.dj0 = ncolst - 1;
            op.ni = ni;
            op.nj = nj;
            op.mask_off = (int32_t)P.masks.size();
            P.masks.insert(P.masks.end(), flag.begin(), flag.end());
            op.ilo = ilo;
            op.jlo = jlo;
            op.iplo = iplo;
            op.jplo = jplo;
            op.ratx = rx;
            op.raty = ry;
            op.level = level;
            op.stage = rc + 1;
            push(op);
            ready = rc + 1;
        }
        if (patchOnly) {
            const double dxf = G.hx, dyf = G.hy;
            double xlo = H.xlower + ilo * dxf, ylo = H.ylower + jlo * dyf;
            double xhi = H.xlower + (ihi + 1) * dxf, yhi = H.ylower + (jhi + 1) * dyf;
            if (!bc_op(buf, dxf, dyf, xlo, xhi, ylo, yhi, ready + 1)) return -1;
            ready = ready + 1;
        }
        return ready;
    }

    // prefilrecur (periodic split); buffer cell (1,1) is global (iglo, jglo)
    int prefil(int level, int buf, int ilo, int ihi, int jlo, int jhi, int iglo, int jglo) {
        if (!err.empty()) return -1;
        if (H.spheredom) { fail("spherical domains are not supported"); return -1; }
        const LevelGeom& G = H.lev[level];
        int ist[3], iend[3], ishift[3], jst[3], jend[3], jshift[3];
        if (H.xperdom) {
            ist[0] = ilo; ist[1] = 0; ist[2] = G.iregsz;
            iend[0] = -1; iend[1] = G.iregsz - 1; iend[2] = ihi;
            ishift[0] = G.iregsz; ishift[1] = 0; ishift[2] = -G.iregsz;
        } else {
            ist[0] = G.iregsz; ist[1] = ilo; ist[2] = G.iregsz;
            iend[0] = -1; iend[1] = ihi; iend[2] = -1;
            ishift[0] = ishift[1] = ishift[2] = 0;
        }
        if (H.yperdom) {
            jst[0] = jlo; jst[1] = 0; jst[2] = G.jregsz;
            jend[0] = -1; jend[1] = G.jregsz - 1; jend[2] = jhi;
            jshift[0] = G.jregsz; jshift[1] = 0; jshift[2] = -G.jregsz;
        } else {
            jst[0] = G.jregsz; jst[1] = jlo; jst[2] = G.jregsz;
            jend[0] = -1; jend[1] = jhi; jend[2] = -1;
            jshift[0] = jshift[1] = jshift[2] = 0;
        }
        int ready = 0;
        for (int a = 0; a < 3; a++) {
            int i1 = std::max(ilo, ist[a]), i2 = std::min(ihi, iend[a]);
            if (i1 > i2) continue;
            for (int b = 0; b < 3; b++) {
                int j1 = std::max(jlo, jst[b]), j2 = std::min(jhi, jend[b]);
                if (j1 > j2) continue;
                int mi = i2 - i1 + 1, mj = j2 - j1 + 1;
                int tbuf = scratch(mi, mj);
                int rc = filrecur(level, tbuf, 1, 1, i1 + ishift[a], i2 + ishift[a],
                                  j1 + jshift[b], j2 + jshift[b], true);
                if (rc < 0) return -1;
                FillOp op{};
                op.type = FOP_COPYBACK;
                op.dst = buf;
                op.src = tbuf;
                op.di0 = i1 - iglo;
                op.dj0 = j1 - jglo;
                op.ni = mi;
                op.nj = mj;
                op.stage = rc + 1;
                const FillBuf& db = P.bufs[buf];
                if (op.di0 < 0 || op.dj0 < 0 || op.di0 + mi > db.nrow || op.dj0 + mj > db.ncol) {
                    fail("prefilrecur: piece outside its buffer");
                    return -1;
                }
                push(op);
                ready = std::max(ready, rc + 1);
            }
        }
        return ready;
    }

    // icall.f: copy the level's grids (store1) over [ilo..ihi]x[jlo..jhi]
    // into `buf` with (ilo, jlo) at 1-based cell (iputst, jputst)
    void icall(int level, int buf, int ilo, int ihi, int jlo, int jhi, int iputst, int jputst,
               std::vector<unsigned char>& covered) {
        const LevelGeom& G = H.lev[level];
        const FillBuf fb = P.bufs[buf];
        const std::vector<int> ks = candidates(level, ilo, ihi, jlo, jhi);
        for (int k : ks) {
            const auto& b = G.boxes[k];
            int ixlo = std::max(b[0], ilo), ixhi = std::min(b[1], ihi);
            int jxlo = std::max(b[2], jlo), jxhi = std::min(b[3], jhi);
            if (!(ixlo <= ixhi && jxlo <= jxhi)) continue;
            FillOp op{};
            op.type = FOP_COPY;
            op.flags = FOPF_FORCE_NEW;
            op.dst = buf;
            op.src = k;
            op.src_level = level;
            op.di0 = ixlo - ilo + iputst - 1;
            op.dj0 = jxlo - jlo + jputst - 1;
            op.ni = ixhi - ixlo + 1;
            op.nj = jxhi - jxlo + 1;
            op.si0 = ixlo - b[0] + NGHOST;
            op.sj0 = jxlo - b[2] + NGHOST;
            op.stage = 0;
            if (op.di0 < 0 || op.dj0 < 0 || op.di0 + op.ni > fb.nrow || op.dj0 + op.nj > fb.ncol) {
                fail("icall: copy outside its buffer");
                return;
            }
            push(op);
            for (int j = 0; j < op.nj; j++)
                for (int i = 0; i < op.ni; i++)
                    covered[(size_t)(op.dj0 + j) * fb.nrow + (op.di0 + i)] = 1;
        }
    }

    // saveqc.f:13-109 for grid k of fine level `level`: coarsened patch
    // grown by one coarse cell, copied from the coarse grids (preicall.f for
    // periodic domains), bc2amr, then cstore's ring into the register
    bool saveqc(int level, int k) {
        const LevelGeom& F = H.lev[level];
        const LevelGeom& C = H.lev[level - 1];
        const int levc = level - 1;
        const auto& b = F.boxes[k];
        const int rx = C.ratx, ry = C.raty, ng = NGHOST;
        const int iclo = b[0] / rx - 1, jclo = b[2] / ry - 1;
        const int ichi = (b[1] + 1) / rx, jchi = (b[3] + 1) / ry;
        const int nrow = ichi - iclo + 1, ncol = jchi - jclo + 1;
        const int S = scratch(nrow, ncol);
        std::vector<unsigned char> covered((size_t)nrow * ncol, 0);
        bool sticks = iclo < 0 || ichi == C.iregsz || jclo < 0 || jchi == C.jregsz;
        if (sticks && (H.xperdom || H.yperdom || H.spheredom)) {
            int ist[3], iend[3], ishift[3], jst[3], jend[3], jshift[3];
            if (H.xperdom) {
                ist[0] = iclo; ist[1] = 0; ist[2] = C.iregsz;
                iend[0] = -1; iend[1] = C.iregsz - 1; iend[2] = ichi;
                ishift[0] = C.iregsz; ishift[1] = 0; ishift[2] = -C.iregsz;
            } else {
                ist[0] = C.iregsz; ist[1] = iclo; ist[2] = C.iregsz;
                iend[0] = -ng; iend[1] = ichi; iend[2] = -ng;
                ishift[0] = ishift[1] = ishift[2] = 0;
            }
            if (H.yperdom) {
                jst[0] = jclo; jst[1] = 0; jst[2] = C.jregsz;
                jend[0] = -1; jend[1] = C.jregsz - 1; jend[2] = jchi;
                jshift[0] = C.jregsz; jshift[1] = 0; jshift[2] = -C.jregsz;
            } else {
                jst[0] = C.jregsz; jst[1] = jclo; jst[2] = C.jregsz;
                jend[0] = -ng; jend[1] = jchi; jend[2] = -ng;
                jshift[0] = jshift[1] = jshift[2] = 0;
            }
            for (int a = 0; a < 3; a++) {
                int i1 = std::max(iclo, ist[a]), i2 = std::min(ichi, iend[a]);
                if (i1 > i2) continue;
                for (int c = 0; c < 3; c++) {
                    int j1 = std::max(jclo, jst[c]), j2 = std::min(jchi, jend[c]);
                    if (j1 > j2) continue;
                    icall(levc, S, i1 + ishift[a], i2 + ishift[a], j1 + jshift[c], j2 + jshift[c],
                          i1 - iclo + 1, j1 - jclo + 1, covered);
                }
            }
        } else {
            icall(levc, S, iclo, ichi, jclo, jchi, 1, 1, covered);
        }
        if (!err.empty()) return false;
        const double hxc = C.hx, hyc = C.hy;
        const double xl = (b[0] * F.hx + H.xlower) - hxc, yb = (b[2] * F.hy + H.ylower) - hyc;
        const double xr = ((b[1] + 1) * F.hx + H.xlower) + hxc, yt = ((b[3] + 1) * F.hy + H.ylower) + hyc;
        size_t nops0 = P.ops.size();
        if (!bc_op(S, hxc, hyc, xl, xr, yb, yt, 1)) return false;
        // every ring cell (the cells cstore reads) must come from a coarse
        // grid or from the boundary conditions; one left unset means the fine
        // grid is not properly nested in the coarse level
        std::vector<unsigned char> set = covered;
        for (int j = 1; j < ncol - 1; j++)
            for (int i = 1; i < nrow - 1; i++) set[(size_t)j * nrow + i] = 1;
        set[0] = set[nrow - 1] = set[(size_t)(ncol - 1) * nrow] = set[(size_t)ncol * nrow - 1] = 1;
        if (P.ops.size() > nops0) {
            const FillOp& bc = P.ops.back();
            auto mark = [&](int i0, int i1, int j0, int j1) {
                for (int j = j0; j < j1; j++)
                    for (int i = i0; i < i1; i++) set[(size_t)j * nrow + i] = 1;
            };
            if (bc.bc[0]) mark(0, bc.bc[0], 0, ncol);
            if (bc.bc[1]) mark(bc.bc[1] - 1, nrow, 0, ncol);
            if (bc.bc[2]) mark(0, nrow, 0, bc.bc[2]);
            if (bc.bc[3]) mark(0, nrow, bc.bc[3] - 1, ncol);
        }
        for (int j = 0; j < ncol; j++)
            for (int i = 0; i < nrow; i++)
                if (!set[(size_t)j * nrow + i]) return fail("saveqc: fine grid not properly nested in the coarse level");
        FillOp ring{};
        ring.type = FOP_RING;
        ring.dst = S;
        ring.src = S;
        ring.level = level;
        ring.si0 = k;
        ring.ni = nrow;
        ring.nj = ncol;
        ring.stage = 2;
        push(ring);
        return true;
    }

    // an FP64 arena buffer (kind 2) laid out like a patch: nrow x ncol cells
    int arena(int nrow, int ncol, int64_t& arena_elems) {
        FillBuf fb{};
        fb.kind = 2;
        fb.level = 0;
        fb.patch = -1;
        fb.nrow = nrow;
        fb.ncol = ncol;
        fb.plane = nrow * ncol;
        fb.off = arena_elems;
        arena_elems += ((int64_t)nv * fb.plane + 31) / 32 * 32;
        P.bufs.push_back(fb);
        return (int)P.bufs.size() - 1;
    }

    // bound.f90:51-158 for grid `grid` of `level` into its own storage
    bool bound(int level, int grid) { return bound_into(level, grid, pbuf(level, grid), NGHOST); }

    // bound.f90:51-158 with a ghost ring of width ng into buffer `buf`
    // ((mx + 2 ng) x (my + 2 ng), the grid's interior at (ng, ng))
    bool bound_into(int level, int grid, int buf, int ng) {
        const LevelGeom& G = H.lev[level];
        const auto& b = G.boxes[grid];
        const int ilo = b[0], ihi = b[1], jlo = b[2], jhi = b[3];
        const int mitot = ihi - ilo + 1 + 2 * ng, mjtot = jhi - jlo + 1 + 2 * ng;
        const double hx = G.hx, hy = G.hy;
        const double xleft = ilo * hx + H.xlower, xright = (ihi + 1) * hx + H.xlower;
        const double ybot = jlo * hy + H.ylower, ytop = (jhi + 1) * hy + H.ylower;
        int r[4];
        double xl = xleft - ng * hx, xr = xright + ng * hx;
        if (xl < H.xlower && H.xperdom)
            r[0] = prefil(level, buf, ilo - ng, ilo - 1, jlo, jhi, ilo - ng, jlo - ng);
        else
            r[0] = filrecur(level, buf, 1, ng + 1, ilo - ng, ilo - 1, jlo, jhi, false, grid);
        if (xr > H.xupper && H.xperdom)
            r[1] = prefil(level, buf, ihi + 1, ihi + ng, jlo, jhi, ilo - ng, jlo - ng);
        else
            r[1] = filrecur(level, buf, mitot - ng + 1, ng + 1, ihi + 1, ihi + ng, jlo, jhi, false, grid);
        double yb = ybot - ng * hy, yt = ytop + ng * hy;
        bool xsticks = (xl < H.xlower || xr > H.xupper) && H.xperdom;
        if ((yb < H.ylower && (H.yperdom || H.spheredom)) || xsticks)
            r[2] = prefil(level, buf, ilo - ng, ihi + ng, jlo - ng, jlo - 1, ilo - ng, jlo - ng);
        else
            r[2] = filrecur(level, buf, 1, 1, ilo - ng, ihi + ng, jlo - ng, jlo - 1, false, grid);
        if ((yt > H.yupper && (H.yperdom || H.spheredom)) || xsticks)
            r[3] = prefil(level, buf, ilo - ng, ihi + ng, jhi + 1, jhi + ng, ilo - ng, jlo - ng);
        else
            r[3] = filrecur(level, buf, 1, mjtot - ng + 1, ilo - ng, ihi + ng, jhi + 1, jhi + ng, false, grid);
        if (!err.empty()) return false;
        int ready = std::max(std::max(r[0], r[1]), std::max(r[2], r[3]));
        return bc_op(buf, hx, hy, xl, xr, yb, yt, ready + 1);
    }

    // preintcopy.f: icall over [ilo..ihi] x [jlo..jhi] split into the
    // pieces inside the domain and their periodic images
    void preicall(int level, int buf, int ilo, int ihi, int jlo, int jhi,
                  std::vector<unsigned char>& covered) {
        const LevelGeom& G = H.lev[level];
        int ist[3], iend[3], ishift[3], jst[3], jend[3], jshift[3];
        if (H.xperdom) {
            ist[0] = ilo; ist[1] = 0; ist[2] = G.iregsz;
            iend[0] = -1; iend[1] = G.iregsz - 1; iend[2] = ihi;
            ishift[0] = G.iregsz; ishift[1] = 0; ishift[2] = -G.iregsz;
        } else {
            ist[0] = G.iregsz; ist[1] = ilo; ist[2] = G.iregsz;
            iend[0] = -1; iend[1] = ihi; iend[2] = -1;
            ishift[0] = ishift[1] = ishift[2] = 0;
        }
        if (H.yperdom) {
            jst[0] = jlo; jst[1] = 0; jst[2] = G.jregsz;
            jend[0] = -1; jend[1] = G.jregsz - 1; jend[2] = jhi;
            jshift[0] = G.jregsz; jshift[1] = 0; jshift[2] = -G.jregsz;
        } else {
            jst[0] = G.jregsz; jst[1] = jlo; jst[2] = G.jregsz;
            jend[0] = -1; jend[1] = jhi; jend[2] = -1;
            jshift[0] = jshift[1] = jshift[2] = 0;
        }
        for (int a = 0; a < 3; a++) {
            const int i1 = std::max(ilo, ist[a]), i2 = std::min(ihi, iend[a]);
            if (i1 > i2) continue;
            for (int c = 0; c < 3; c++) {
                const int j1 = std::max(jlo, jst[c]), j2 = std::min(jhi, jend[c]);
                if (j1 > j2) continue;
                icall(level, buf, i1 + ishift[a], i2 + ishift[a], j1 + jshift[c], j2 + jshift[c],
                      i1 - ilo + 1, j1 - jlo + 1, covered);
            }
        }
    }

    // filval.f90:3-188 (no aux) for new grid k of `level`: the coarse patch
    // grown by one coarse cell from level-1's grids (intcopy / preintcopy,
    // stage 0) with bc2amr (stage 1); the old grids of the level (slot
    // FILL_OLDSLOT) copied over the grid and its ghost ring (stage 0); the
    // min-mod interpolation into the interior cells no old grid reached
    // (stage 2)
    bool filval(int level, int k) {
        const LevelGeom& F = H.lev[level];
        const LevelGeom& C = H.lev[level - 1];
        const int levc = level - 1, ng = NGHOST;
        const auto& b = F.boxes[k];
        const int rx = C.ratx, ry = C.raty;
        const int nx = b[1] - b[0] + 1, ny = b[3] - b[2] + 1;
        if (nx % rx || ny % ry || b[0] % rx || b[2] % ry)
            return fail("filval: a new grid is not aligned to the coarse cells");
        const int mic = nx / rx + 2, mjc = ny / ry + 2;                     // gfixup.f:83-84
        const int iclo = b[0] / rx - 1, jclo = b[2] / ry - 1;              // filval.f90:59-62
        const int ichi = (b[1] + 1) / rx, jchi = (b[3] + 1) / ry;
        const int S = scratch(mic, mjc);
        std::vector<unsigned char> cov_c((size_t)mic * mjc, 0);
        const bool sxc = iclo < 0 || ichi >= C.iregsz, syc = jclo < 0 || jchi >= C.jregsz;
        if ((H.xperdom && sxc) || (H.yperdom && syc))
            preicall(levc, S, iclo, ichi, jclo, jchi, cov_c);
        else
            icall(levc, S, iclo, ichi, jclo, jchi, 1, 1, cov_c);
        if (!err.empty()) return false;
        const double dxc = F.hx * rx, dyc = F.hy * ry;                     // filval.f90:46-47
        const double xl = (b[0] * F.hx + H.xlower) - dxc, xr = ((b[1] + 1) * F.hx + H.xlower) + dxc;
        const double yb = (b[2] * F.hy + H.ylower) - dyc, yt = ((b[3] + 1) * F.hy + H.ylower) + dyc;
        if (!bc_op(S, dxc, dyc, xl, xr, yb, yt, 1)) return false;
        const int D = pbuf(level, k);
        const int mitot = nx + 2 * ng, mjtot = ny + 2 * ng;
        std::vector<unsigned char> cov((size_t)mitot * mjtot, 0);
        const bool sxf = b[0] < 0 || b[1] >= F.iregsz, syf = b[2] < 0 || b[3] >= F.jregsz;
        if ((H.xperdom && sxf) || (H.yperdom && syf))
            preicall(FILL_OLDSLOT, D, b[0] - ng, b[1] + ng, b[2] - ng, b[3] + ng, cov);
        else
            icall(FILL_OLDSLOT, D, b[0] - ng, b[1] + ng, b[2] - ng, b[3] + ng, 1, 1, cov);
        if (!err.empty()) return false;
        FillOp op{};
        op.type = FOP_FILVAL;
        op.dst = D;
        op.src = S;
        op.ni = mic;
        op.nj = mjc;
        op.ratx = rx;
        op.raty = ry;
        op.level = level;
        op.stage = 2;
        if (P.masks.size() + (size_t)nx * ny > (size_t)INT32_MAX) return fail("filval: masks too large");
        op.mask_off = (int32_t)P.masks.size();
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) P.masks.push_back(cov[(size_t)(j + ng) * mitot + (i + ng)]);
        push(op);
        return true;
    }

    void index_slot(int slot) {
        if ((int)index.size() <= slot) index.resize(slot + 1);
        index[slot].build(H.lev[slot].boxes, H.lev[slot].iregsz, H.lev[slot].jregsz);
    }
};

}  // namespace

static void finish_plan(FillPlan& P) {
    std::stable_sort(P.ops.begin(), P.ops.end(),
                     [](const FillOp& a, const FillOp& b) { return a.stage < b.stage; });
    P.nstages = P.ops.empty() ? 0 : P.ops.back().stage + 1;
    P.stage_start.assign(P.nstages + 1, 0);
    for (const auto& op : P.ops) P.stage_start[op.stage + 1]++;
    for (int s = 0; s < P.nstages; s++) P.stage_start[s + 1] += P.stage_start[s];
}

bool plan_filval(const HierGeom& H, int level, int nvar, FillPlan& P, std::string& err) {
    P = FillPlan();
    if (level < 2 || level > H.nlev || level >= FILL_OLDSLOT) {
        err = "filval: level out of range";
        return false;
    }
    Planner pl(H, P, nvar);
    pl.index_slot(FILL_OLDSLOT);
    for (int k = 0; k < (int)H.lev[level].boxes.size(); k++)
        if (!pl.filval(level, k)) {
            err = pl.err.empty() ? "filval planning failed" : pl.err;
            return false;
        }
    finish_plan(P);
    return true;
}

bool plan_saveqc(const HierGeom& H, int level, int nvar, FillPlan& P, std::string& err) {
    P = FillPlan();
    Planner pl(H, P, nvar);
    for (int k = 0; k < (int)H.lev[level].boxes.size(); k++)
        if (!pl.saveqc(level, k)) {
            err = pl.err.empty() ? "saveqc planning failed" : pl.err;
            return false;
        }
    finish_plan(P);
    return true;
}

bool plan_richardson(const HierGeom& H, int level, int nvar, RichPlan& R, std::string& err) {
    R = RichPlan();
    FillPlan& P = R.fill;
    Planner pl(H, P, nvar);
    const LevelGeom& G = H.lev[level];
    const int ng = NGHOST, ng2 = 2 * NGHOST;
    std::vector<int> dub(G.boxes.size());
    for (int k = 0; k < (int)G.boxes.size(); k++) {
        const auto& b = G.boxes[k];
        const int mx = b[1] - b[0] + 1, my = b[3] - b[2] + 1;
        if (mx % 2 || my % 2) {
            err = "Richardson estimation needs grids of even size (errest.f:27-28)";
            return false;
        }
        // copysol.f: the old interior into the centre of the wider array
        const int d = pl.scratch(mx + 2 * ng2, my + 2 * ng2);
        dub[k] = d;
        FillOp op{};
        op.type = FOP_COPY;
        op.flags = FOPF_FORCE_OLD;
        op.dst = d;
        op.src = k;
        op.src_level = level;
        op.di0 = ng2;
        op.dj0 = ng2;
        op.ni = mx;
        op.nj = my;
        op.si0 = ng;
        op.sj0 = ng;
        op.stage = 0;
        pl.push(op);
        if (!pl.bound_into(level, k, d, ng2)) {
            err = pl.err.empty() ? "Richardson ghost-fill planning failed" : pl.err;
            return false;
        }
    }
    int last = 0;
    for (const auto& op : P.ops) last = std::max(last, op.stage);
    for (int k = 0; k < (int)G.boxes.size(); k++) {
        const auto& b = G.boxes[k];
        const int mx = b[1] - b[0] + 1, my = b[3] - b[2] + 1;
        // coarsen.f into the coarse array (mx/2 + 2 nghost) x (my/2 + 2 nghost)
        const int c = pl.arena(mx / 2 + 2 * ng, my / 2 + 2 * ng, R.arena_elems);
        FillOp op{};
        op.type = FOP_COARSEN;
        op.dst = c;
        op.src = dub[k];
        op.ni = mx / 2 + 2 * ng;
        op.nj = my / 2 + 2 * ng;
        op.stage = last + 1;
        pl.push(op);
        // flagger's locbig: the grid's values with their ghost cells
        const int f = pl.arena(mx + 2 * ng, my + 2 * ng, R.arena_elems);
        FillOp cp{};
        cp.type = FOP_COPY;
        cp.flags = FOPF_BIG;
        cp.dst = f;
        cp.src = k;
        cp.src_level = level;
        cp.ni = mx + 2 * ng;
        cp.nj = my + 2 * ng;
        cp.stage = 0;
        pl.push(cp);
        PatchDesc pc{}, pf{};
        pc.off = P.bufs[c].off;
        pc.mx = mx / 2;
        pc.my = my / 2;
        pc.mitot = pc.mx + 2 * ng;
        pc.plane = pc.mitot * (pc.my + 2 * ng);
        pc.ilo = b[0] / 2; pc.ihi = pc.ilo + pc.mx - 1;
        pc.jlo = b[2] / 2; pc.jhi = pc.jlo + pc.my - 1;
        pf.off = P.bufs[f].off;
        pf.mx = mx;
        pf.my = my;
        pf.mitot = mx + 2 * ng;
        pf.plane = pf.mitot * (my + 2 * ng);
        pf.ilo = b[0]; pf.ihi = b[1];
        pf.jlo = b[2]; pf.jhi = b[3];
        R.crs.push_back(pc);
        R.fine.push_back(pf);
    }
    finish_plan(P);
    return true;
}

bool plan_fill_level(const HierGeom& H, int level, int nvar, FillPlan& P, std::string& err) {
    P = FillPlan();
    Planner pl(H, P, nvar);
    for (int k = 0; k < (int)H.lev[level].boxes.size(); k++)
        if (!pl.bound(level, k)) {
            err = pl.err.empty() ? "ghost-fill planning failed" : pl.err;
            return false;
        }
    finish_plan(P);
    return true;
}

}  // namespace amrgpu
