// amr3_plan.cpp -- host planner of the 3-D ghost-fill programs and of the
// coarse-fine registers.
//
// Replays bound.f (3-D) for every grid of a level on index boxes only: the
// six ghost slabs (x: all j, interior k; y: interior i and k; z: all i, j),
// filrecur (filpatch.f) with intfil's coverage marks (intfil.f), trimbd's
// unset box (trimbd.f), the coarse patch formula (filpatch.f:128-133), the
// recursion to coarser levels, prefilrecur's periodic split (prefilp.f) and
// bc3amr's side extents (bc3amr.f90).  Every array index a generated
// operation touches on the device is checked here, so a program that would
// read or write outside a buffer is refused on the host instead of launched.
// The registers follow prepf.f, prepc.f and setuse.f (3-D); upbnd's
// once-per-side rule (upbnd.f) and fluxsv's last-write-wins order are
// resolved here into per-cell entry lists and write flags.
#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "amr3_plan.hpp"

namespace amrgpu {

namespace {

constexpr int NG3 = 2;

struct Planner3 {
    const Hier3Geom& H;
    Fill3Plan& P;
    int nv;
    std::string err;
    std::vector<std::vector<int>> patch_buf;

    Planner3(const Hier3Geom& h, Fill3Plan& p, int nvar) : H(h), P(p), nv(nvar) {
        patch_buf.resize(H.nlev + 1);
        for (int l = 1; l <= H.nlev; l++) patch_buf[l].assign(H.lev[l].boxes.size(), -1);
    }

    bool fail(const char* msg) {
        if (err.empty()) err = msg;
        return false;
    }

    int pbuf(int level, int grid) {
        int& b = patch_buf[level][grid];
        if (b < 0) {
            const auto& bx = H.lev[level].boxes[grid];
            Fill3Buf fb{};
            fb.kind = 0;
            fb.level = level;
            fb.patch = grid;
            fb.nrow = bx[1] - bx[0] + 1 + 2 * NG3;
            fb.ncol = bx[3] - bx[2] + 1 + 2 * NG3;
            fb.nfil = bx[5] - bx[4] + 1 + 2 * NG3;
            fb.plane = (int64_t)fb.nrow * fb.ncol * fb.nfil;
            b = (int)P.bufs.size();
            P.bufs.push_back(fb);
        }
        return b;
    }

    int scratch(int nrow, int ncol, int nfil) {
        Fill3Buf fb{};
        fb.kind = 1;
        fb.patch = -1;
        fb.nrow = nrow;
        fb.ncol = ncol;
        fb.nfil = nfil;
        fb.plane = (int64_t)nrow * ncol * nfil;
        fb.off = P.scratch_elems;
        P.scratch_elems += ((int64_t)nv * fb.plane + 31) / 32 * 32;
        P.bufs.push_back(fb);
        return (int)P.bufs.size() - 1;
    }

    void push(const Fill3Op& op) { P.ops.push_back(op); }

    // bc3amr.f90 on a whole buffer with the given physical extent
    bool bc_op(int buf, double hx, double hy, double hz, const double lo[3], const double hi[3],
               int stage) {
        if (H.xperdom && H.yperdom && H.zperdom) return true;
        const Fill3Buf& fb = P.bufs[buf];
        const int n[3] = {fb.nrow, fb.ncol, fb.nfil};
        const double h[3] = {hx, hy, hz};
        const double dlo[3] = {H.xlower, H.ylower, H.zlower};
        const double dhi[3] = {H.xupper, H.yupper, H.zupper};
        Fill3Op op{};
        op.type = F3_BC;
        op.dst = buf;
        op.ni = fb.nrow;
        op.nj = fb.ncol;
        op.nk = fb.nfil;
        op.stage = stage;
        bool any = false;
        for (int d = 0; d < 3; d++) {
            const double marg = h[d] * .01;
            if (lo[d] < dlo[d] - marg) {
                int nl = (int)((dlo[d] + marg - lo[d]) / h[d]);
                int mb = H.mthbc[2 * d];
                if (mb < 1 || mb > 3) return fail("bc3amr: user or unknown boundary condition");
                if (mb == 3 && nv < d + 2) return fail("bc3amr: wall boundary needs the momentum component");
                // a buffer wholly outside the domain on this side has no
                // source plane; the grid's own bc3amr overwrites those cells
                bool fits = (mb == 1 && nl < n[d]) || (mb == 3 && 2 * nl <= n[d]);
                if (nl > 0 && mb != 2 && fits) {
                    op.bc[2 * d] = nl;
                    any = true;
                }
            }
            if (hi[d] > dhi[d] + marg) {
                int nr = (int)((hi[d] - dhi[d] + marg) / h[d]);
                int beg = std::max(n[d] - nr + 1, 1);
                int mb = H.mthbc[2 * d + 1];
                if (mb < 1 || mb > 3) return fail("bc3amr: user or unknown boundary condition");
                if (mb == 3 && nv < d + 2) return fail("bc3amr: wall boundary needs the momentum component");
                bool fits = (mb == 1 && beg >= 2) || (mb == 3 && 2 * beg - 1 - n[d] >= 1);
                if (beg <= n[d] && mb != 2 && fits) {
                    op.bc[2 * d + 1] = beg;
                    any = true;
                }
            }
        }
        if (any) push(op);
        return true;
    }

    // filrecur: fill box [ilo..ihi]x[jlo..jhi]x[klo..khi] of `level` into
    // buffer `buf` starting at its 1-based cell (rst, cst, fst).  Returns the
    // stage after which the box (and, with patchOnly, the buffer) is final.
    int filrecur(int level, int buf, int rst, int cst, int fst, int ilo, int ihi, int jlo,
                 int jhi, int klo, int khi, bool patchOnly) {
        if (!err.empty()) return -1;
        const Level3Geom& G = H.lev[level];
        const int ni = ihi - ilo + 1, nj = jhi - jlo + 1, nk = khi - klo + 1;
        const Fill3Buf fb = P.bufs[buf];
        if (rst < 1 || cst < 1 || fst < 1 || rst - 1 + ni > fb.nrow || cst - 1 + nj > fb.ncol ||
            fst - 1 + nk > fb.nfil) {
            fail("filrecur: box outside its buffer");
            return -1;
        }
        auto at = [&](int i, int j, int k) { return ((size_t)(k - klo) * nj + (j - jlo)) * ni + (i - ilo); };
        std::vector<unsigned char> flag((size_t)ni * nj * nk, 0);
        for (int g = 0; g < (int)G.boxes.size(); g++) {
            const auto& b = G.boxes[g];
            int ixlo = std::max(b[0], ilo), ixhi = std::min(b[1], ihi);
            int jxlo = std::max(b[2], jlo), jxhi = std::min(b[3], jhi);
            int kxlo = std::max(b[4], klo), kxhi = std::min(b[5], khi);
            if (!(ixlo <= ixhi && jxlo <= jxhi && kxlo <= kxhi)) continue;
            Fill3Op op{};
            op.type = F3_COPY;
            op.dst = buf;
            op.src = g;
            op.src_level = level;
            op.di0 = ixlo - ilo + rst - 1;
            op.dj0 = jxlo - jlo + cst - 1;
            op.dk0 = kxlo - klo + fst - 1;
            op.ni = ixhi - ixlo + 1;
            op.nj = jxhi - jxlo + 1;
            op.nk = kxhi - kxlo + 1;
            op.si0 = ixlo - b[0] + NG3;
            op.sj0 = jxlo - b[2] + NG3;
            op.sk0 = kxlo - b[4] + NG3;
            op.stage = 0;
            push(op);
            P.src_levels_mask |= 1u << level;
            for (int k = kxlo; k <= kxhi; k++)
                for (int j = jxlo; j <= jxhi; j++)
                    for (int i = ixlo; i <= ixhi; i++) flag[at(i, j, k)] = 1;
        }
        auto mark = [&](int i0, int i1, int j0, int j1, int k0, int k1) {
            i0 = std::max(i0, ilo); i1 = std::min(i1, ihi);
            j0 = std::max(j0, jlo); j1 = std::min(j1, jhi);
            k0 = std::max(k0, klo); k1 = std::min(k1, khi);
            for (int k = k0; k <= k1; k++)
                for (int j = j0; j <= j1; j++)
                    for (int i = i0; i <= i1; i++) flag[at(i, j, k)] = 1;
        };
        if (khi >= G.kregsz) mark(ilo, ihi, jlo, jhi, std::max(G.kregsz, klo), khi);
        if (klo < 0) mark(ilo, ihi, jlo, jhi, klo, std::min(-1, fst + khi - klo));
        if (jhi >= G.jregsz) mark(ilo, ihi, std::max(G.jregsz, jlo), jhi, klo, khi);
        if (jlo < 0) mark(ilo, ihi, jlo, std::min(-1, cst + jhi - jlo), klo, khi);
        if (ilo < 0) mark(ilo, std::min(-1, rst + ihi - ilo), jlo, jhi, klo, khi);
        if (ihi >= G.iregsz) mark(std::max(G.iregsz, ilo), ihi, jlo, jhi, klo, khi);
        // trimbd: box of the unset cells (global indices)
        bool set = true;
        int il = ihi + 1, ir = ilo - 1, jl = jhi + 1, jr = jlo - 1, kl = khi + 1, kr = klo - 1;
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++)
                    if (!flag[at(i, j, k)]) {
                        set = false;
                        il = std::min(il, i); ir = std::max(ir, i);
                        jl = std::min(jl, j); jr = std::max(jr, j);
                        kl = std::min(kl, k); kr = std::max(kr, k);
                    }
        int ready = 0;
        if (!set) {
            if (level == 1) {
                fail("filrecur: level 1 cells left unset");
                return -1;
            }
            const Level3Geom& C = H.lev[level - 1];
            const int rx = C.ratx, ry = C.raty, rz = C.ratz, ng = NG3;
            int iplo = (il - rx + ng * rx) / rx - ng, iphi = (ir + rx) / rx;
            int jplo = (jl - ry + ng * ry) / ry - ng, jphi = (jr + ry) / ry;
            int kplo = (kl - rz + ng * rz) / rz - ng, kphi = (kr + rz) / rz;
            int nrc = iphi - iplo + 1, ncc = jphi - jplo + 1, nfc = kphi - kplo + 1;
            if (nrc > ihi - ilo + 3 || ncc > jhi - jlo + 3 || nfc > khi - klo + 3) {
                fail("filrecur: coarse scratch too small");
                return -1;
            }
            int cbuf = scratch(nrc, ncc, nfc);
            bool sticks = iplo < 0 || jplo < 0 || kplo < 0 || iphi >= C.iregsz ||
                          jphi >= C.jregsz || kphi >= C.kregsz;
            int rc;
            if ((H.xperdom || H.yperdom || H.zperdom) && sticks)
                rc = prefil(level - 1, cbuf, iplo, iphi, jplo, jphi, kplo, kphi, iplo, jplo, kplo);
            else
                rc = filrecur(level - 1, cbuf, 1, 1, 1, iplo, iphi, jplo, jphi, kplo, kphi, true);
            if (rc < 0) return -1;
            // the interpolation reads coarse cells ic-1..ic+1 etc. of the
            // scratch for every unset fine cell: check them all
            for (int k = klo; k <= khi; k++)
                for (int j = jlo; j <= jhi; j++)
                    for (int i = ilo; i <= ihi; i++) {
                        if (flag[at(i, j, k)]) continue;
                        int ic = (int)std::floor((double)i / (double)rx) - iplo + 1;
                        int jc = (int)std::floor((double)j / (double)ry) - jplo + 1;
                        int kc = (int)std::floor((double)k / (double)rz) - kplo + 1;
                        if (ic - 1 < 1 || ic + 1 > nrc || jc - 1 < 1 || jc + 1 > ncc ||
                            kc - 1 < 1 || kc + 1 > nfc) {
                            fail("filrecur: interpolation stencil outside the coarse patch");
                            return -1;
                        }
                    }
            Fill3Op op{};
            op.type = F3_INTERP;
            op.dst = buf;
            op.src = cbuf;
            op.di0 = rst - 1;
            op.dj0 = cst - 1;
            op.dk0 = fst - 1;
            op.ni = ni;
            op.nj = nj;
            op.nk = nk;
            op.mask_off = (int32_t)P.masks.size();
            P.masks.insert(P.masks.end(), flag.begin(), flag.end());
            op.ilo = ilo; op.jlo = jlo; op.klo = klo;
            op.iplo = iplo; op.jplo = jplo; op.kplo = kplo;
            op.ratx = rx; op.raty = ry; op.ratz = rz;
            op.level = level;
            op.stage = rc + 1;
            push(op);
            ready = rc + 1;
        }
        if (patchOnly) {
            const double lo[3] = {H.xlower + ilo * G.hx, H.ylower + jlo * G.hy, H.zlower + klo * G.hz};
            const double hi[3] = {H.xlower + (ihi + 1) * G.hx, H.ylower + (jhi + 1) * G.hy,
                                  H.zlower + (khi + 1) * G.hz};
            if (!bc_op(buf, G.hx, G.hy, G.hz, lo, hi, ready + 1)) return -1;
            ready = ready + 1;
        }
        return ready;
    }

    // prefilrecur (periodic split); buffer cell (1,1,1) is global (iglo, jglo, kglo)
    int prefil(int level, int buf, int ilo, int ihi, int jlo, int jhi, int klo, int khi, int iglo,
               int jglo, int kglo) {
        if (!err.empty()) return -1;
        const Level3Geom& G = H.lev[level];
        const int lo[3] = {ilo, jlo, klo}, hi[3] = {ihi, jhi, khi};
        const bool per[3] = {H.xperdom, H.yperdom, H.zperdom};
        const int sz[3] = {G.iregsz, G.jregsz, G.kregsz};
        int st[3][3], en[3][3], sh[3][3];
        for (int d = 0; d < 3; d++) {
            if (per[d]) {
                st[d][0] = lo[d]; st[d][1] = 0; st[d][2] = sz[d];
                en[d][0] = -1; en[d][1] = sz[d] - 1; en[d][2] = hi[d];
                sh[d][0] = sz[d]; sh[d][1] = 0; sh[d][2] = -sz[d];
            } else {
                st[d][0] = sz[d]; st[d][1] = lo[d]; st[d][2] = sz[d];
                en[d][0] = -1; en[d][1] = hi[d]; en[d][2] = -1;
                sh[d][0] = sh[d][1] = sh[d][2] = 0;
            }
        }
        int ready = 0;
        for (int a = 0; a < 3; a++) {
            int i1 = std::max(ilo, st[0][a]), i2 = std::min(ihi, en[0][a]);
            if (i1 > i2) continue;
            for (int b = 0; b < 3; b++) {
                int j1 = std::max(jlo, st[1][b]), j2 = std::min(jhi, en[1][b]);
                if (j1 > j2) continue;
                for (int c = 0; c < 3; c++) {
                    int k1 = std::max(klo, st[2][c]), k2 = std::min(khi, en[2][c]);
                    if (k1 > k2) continue;
                    int mi = i2 - i1 + 1, mj = j2 - j1 + 1, mk = k2 - k1 + 1;
                    int tbuf = scratch(mi, mj, mk);
                    int rc = filrecur(level, tbuf, 1, 1, 1, i1 + sh[0][a], i2 + sh[0][a],
                                      j1 + sh[1][b], j2 + sh[1][b], k1 + sh[2][c], k2 + sh[2][c],
                                      true);
                    if (rc < 0) return -1;
                    Fill3Op op{};
                    op.type = F3_COPYBACK;
                    op.dst = buf;
                    op.src = tbuf;
                    op.di0 = i1 - iglo;
                    op.dj0 = j1 - jglo;
                    op.dk0 = k1 - kglo;
                    op.ni = mi;
                    op.nj = mj;
                    op.nk = mk;
                    op.stage = rc + 1;
                    const Fill3Buf& db = P.bufs[buf];
                    if (op.di0 < 0 || op.dj0 < 0 || op.dk0 < 0 || op.di0 + mi > db.nrow ||
                        op.dj0 + mj > db.ncol || op.dk0 + mk > db.nfil) {
                        fail("prefilrecur: piece outside its buffer");
                        return -1;
                    }
                    push(op);
                    ready = std::max(ready, rc + 1);
                }
            }
        }
        return ready;
    }

    // bound.f (3-D) for grid `grid` of `level` into its own storage
    bool bound(int level, int grid) { return bound_into(level, grid, pbuf(level, grid), NG3); }

    // bound.f (3-D) with a ghost ring of width ng into buffer `buf`
    // ((mx + 2 ng) x (my + 2 ng) x (mz + 2 ng), the interior at (ng, ng, ng))
    bool bound_into(int level, int grid, int buf, int ng) {
        const Level3Geom& G = H.lev[level];
        const auto& b = G.boxes[grid];
        const int ilo = b[0], ihi = b[1], jlo = b[2], jhi = b[3], klo = b[4], khi = b[5];
        const int mitot = ihi - ilo + 1 + 2 * ng, mjtot = jhi - jlo + 1 + 2 * ng;
        const int mktot = khi - klo + 1 + 2 * ng;
        const double xlo = ilo * G.hx + H.xlower, xhi = (ihi + 1) * G.hx + H.xlower;
        const double ylo = jlo * G.hy + H.ylower, yhi = (jhi + 1) * G.hy + H.ylower;
        const double zlo = klo * G.hz + H.zlower, zhi = (khi + 1) * G.hz + H.zlower;
        const int ig = ilo - ng, jg = jlo - ng, kg = klo - ng;
        const double xl = xlo - ng * G.hx, xr = xhi + ng * G.hx;
        const double yf = ylo - ng * G.hy, yr = yhi + ng * G.hy;
        const double zb = zlo - ng * G.hz, zt = zhi + ng * G.hz;
        const bool ys = yf < H.ylower || yr > H.yupper;
        const bool xs = xl < H.xlower || xr > H.xupper;
        int r[6];
        if ((H.xperdom && xl < H.xlower) || (H.yperdom && ys))
            r[0] = prefil(level, buf, ilo - ng, ilo - 1, jlo - ng, jhi + ng, klo, khi, ig, jg, kg);
        else
            r[0] = filrecur(level, buf, 1, 1, ng + 1, ilo - ng, ilo - 1, jlo - ng, jhi + ng, klo,
                            khi, false);
        if ((H.xperdom && xr > H.xupper) || (H.yperdom && ys))
            r[1] = prefil(level, buf, ihi + 1, ihi + ng, jlo - ng, jhi + ng, klo, khi, ig, jg, kg);
        else
            r[1] = filrecur(level, buf, mitot - ng + 1, 1, ng + 1, ihi + 1, ihi + ng, jlo - ng,
                            jhi + ng, klo, khi, false);
        if (H.yperdom && yf < H.ylower)
            r[2] = prefil(level, buf, ilo, ihi, jlo - ng, jlo - 1, klo, khi, ig, jg, kg);
        else
            r[2] = filrecur(level, buf, ng + 1, 1, ng + 1, ilo, ihi, jlo - ng, jlo - 1, klo, khi,
                            false);
        if (H.yperdom && yr > H.yupper)
            r[3] = prefil(level, buf, ilo, ihi, jhi + 1, jhi + ng, klo, khi, ig, jg, kg);
        else
            r[3] = filrecur(level, buf, ng + 1, mjtot - ng + 1, ng + 1, ilo, ihi, jhi + 1,
                            jhi + ng, klo, khi, false);
        if ((H.zperdom && zb < H.zlower) || (H.xperdom && xs) || (H.yperdom && ys))
            r[4] = prefil(level, buf, ilo - ng, ihi + ng, jlo - ng, jhi + ng, klo - ng, klo - 1,
                          ig, jg, kg);
        else
            r[4] = filrecur(level, buf, 1, 1, 1, ilo - ng, ihi + ng, jlo - ng, jhi + ng, klo - ng,
                            klo - 1, false);
        if ((H.zperdom && zt > H.zupper) || (H.xperdom && xs) || (H.yperdom && ys))
            r[5] = prefil(level, buf, ilo - ng, ihi + ng, jlo - ng, jhi + ng, khi + 1, khi + ng,
                          ig, jg, kg);
        else
            r[5] = filrecur(level, buf, 1, 1, mktot - ng + 1, ilo - ng, ihi + ng, jlo - ng,
                            jhi + ng, khi + 1, khi + ng, false);
        if (!err.empty()) return false;
        int ready = 0;
        for (int s = 0; s < 6; s++) ready = std::max(ready, r[s]);
        const double lo[3] = {xl, yf, zb}, hi[3] = {xr, yr, zt};
        return bc_op(buf, G.hx, G.hy, G.hz, lo, hi, ready + 1);
    }

    // an FP64 arena buffer (kind 2) laid out like a grid
    int arena(int nrow, int ncol, int nfil, int64_t& arena_elems) {
        Fill3Buf fb{};
        fb.kind = 2;
        fb.patch = -1;
        fb.nrow = nrow;
        fb.ncol = ncol;
        fb.nfil = nfil;
        fb.plane = (int64_t)nrow * ncol * nfil;
        fb.off = arena_elems;
        arena_elems += ((int64_t)nv * fb.plane + 63) / 64 * 64;
        P.bufs.push_back(fb);
        return (int)P.bufs.size() - 1;
    }

    // icall.f (3-D): copy the level's grids (store1) over the box into
    // `buf` with the box corner at 1-based cell (ip, jp, kp)
    void icall(int level, int buf, int ilo, int ihi, int jlo, int jhi, int klo, int khi, int ip,
               int jp, int kp, std::vector<unsigned char>& covered) {
        const Level3Geom& G = H.lev[level];
        const Fill3Buf fb = P.bufs[buf];
        for (int g = 0; g < (int)G.boxes.size(); g++) {
            const auto& b = G.boxes[g];
            int ixlo = std::max(b[0], ilo), ixhi = std::min(b[1], ihi);
            int jxlo = std::max(b[2], jlo), jxhi = std::min(b[3], jhi);
            int kxlo = std::max(b[4], klo), kxhi = std::min(b[5], khi);
            if (!(ixlo <= ixhi && jxlo <= jxhi && kxlo <= kxhi)) continue;
            Fill3Op op{};
            op.type = F3_COPY;
            op.flags = F3F_FORCE_NEW;
            op.dst = buf;
            op.src = g;
            op.src_level = level;
            op.di0 = ixlo - ilo + ip - 1;
            op.dj0 = jxlo - jlo + jp - 1;
            op.dk0 = kxlo - klo + kp - 1;
            op.ni = ixhi - ixlo + 1;
            op.nj = jxhi - jxlo + 1;
            op.nk = kxhi - kxlo + 1;
            op.si0 = ixlo - b[0] + NG3;
            op.sj0 = jxlo - b[2] + NG3;
            op.sk0 = kxlo - b[4] + NG3;
            if (op.di0 < 0 || op.dj0 < 0 || op.dk0 < 0 || op.di0 + op.ni > fb.nrow ||
                op.dj0 + op.nj > fb.ncol || op.dk0 + op.nk > fb.nfil) {
                fail("icall: copy outside its buffer");
                return;
            }
            push(op);
            for (int k = 0; k < op.nk; k++)
                for (int j = 0; j < op.nj; j++)
                    for (int i = 0; i < op.ni; i++)
                        covered[((size_t)(op.dk0 + k) * fb.ncol + (op.dj0 + j)) * fb.nrow + (op.di0 + i)] = 1;
        }
    }

    // saveqc.f (3-D) for grid g of fine level `level`: the coarsened grid
    // grown by one coarse cell, copied from the coarse grids (preicall.f
    // when every direction is periodic), bc3amr, then cstore's ring
    bool saveqc(int level, int g) {
        const Level3Geom& F = H.lev[level];
        const Level3Geom& C = H.lev[level - 1];
        const auto& b = F.boxes[g];
        const int rx = C.ratx, ry = C.raty, rz = C.ratz;
        const int iclo = b[0] / rx - 1, jclo = b[2] / ry - 1, kclo = b[4] / rz - 1;
        const int ichi = (b[1] + 1) / rx, jchi = (b[3] + 1) / ry, kchi = (b[5] + 1) / rz;
        const int nrow = ichi - iclo + 1, ncol = jchi - jclo + 1, nfil = kchi - kclo + 1;
        const int S = scratch(nrow, ncol, nfil);
        std::vector<unsigned char> covered((size_t)nrow * ncol * nfil, 0);
        const bool sticks = iclo < 0 || ichi == C.iregsz || jclo < 0 || jchi == C.jregsz ||
                            kclo < 0 || kchi == C.kregsz;
        if (sticks && H.xperdom && H.yperdom && H.zperdom) {
            const int lo[3] = {iclo, jclo, kclo}, hi[3] = {ichi, jchi, kchi};
            const int sz[3] = {C.iregsz, C.jregsz, C.kregsz};
            int st[3][3], en[3][3], sh[3][3];
            for (int d = 0; d < 3; d++) {
                st[d][0] = lo[d]; st[d][1] = 0; st[d][2] = sz[d];
                en[d][0] = -1; en[d][1] = sz[d] - 1; en[d][2] = hi[d];
                sh[d][0] = sz[d]; sh[d][1] = 0; sh[d][2] = -sz[d];
            }
            for (int a = 0; a < 3; a++) {
                int i1 = std::max(iclo, st[0][a]), i2 = std::min(ichi, en[0][a]);
                if (i1 > i2) continue;
                for (int c = 0; c < 3; c++) {
                    int j1 = std::max(jclo, st[1][c]), j2 = std::min(jchi, en[1][c]);
                    if (j1 > j2) continue;
                    for (int e = 0; e < 3; e++) {
                        int k1 = std::max(kclo, st[2][e]), k2 = std::min(kchi, en[2][e]);
                        if (k1 > k2) continue;
                        icall(level - 1, S, i1 + sh[0][a], i2 + sh[0][a], j1 + sh[1][c],
                              j2 + sh[1][c], k1 + sh[2][e], k2 + sh[2][e], i1 - iclo + 1,
                              j1 - jclo + 1, k1 - kclo + 1, covered);
                    }
                }
            }
        } else {
            icall(level - 1, S, iclo, ichi, jclo, jchi, kclo, kchi, 1, 1, 1, covered);
        }
        if (!err.empty()) return false;
        const double lo[3] = {(b[0] * F.hx + H.xlower) - C.hx, (b[2] * F.hy + H.ylower) - C.hy,
                              (b[4] * F.hz + H.zlower) - C.hz};
        const double hi[3] = {((b[1] + 1) * F.hx + H.xlower) + C.hx,
                              ((b[3] + 1) * F.hy + H.ylower) + C.hy,
                              ((b[5] + 1) * F.hz + H.zlower) + C.hz};
        const size_t nops0 = P.ops.size();
        if (!bc_op(S, C.hx, C.hy, C.hz, lo, hi, 1)) return false;
        // every ring cell cstore reads must come from a coarse grid or from
        // the boundary conditions; one left unset means the fine grid is not
        // properly nested in the coarse level
        std::vector<unsigned char> set = covered;
        if (P.ops.size() > nops0) {
            const Fill3Op& bc = P.ops.back();
            const int n[3] = {nrow, ncol, nfil};
            for (int d = 0; d < 3; d++) {
                int a0 = 0, a1 = 0;
                for (int side = 0; side < 2; side++) {
                    const int v = bc.bc[2 * d + side];
                    if (!v) continue;
                    if (side == 0) { a0 = 0; a1 = v; } else { a0 = v - 1; a1 = n[d]; }
                    for (int k = 0; k < nfil; k++)
                        for (int j = 0; j < ncol; j++)
                            for (int i = 0; i < nrow; i++) {
                                const int c = d == 0 ? i : (d == 1 ? j : k);
                                if (c >= a0 && c < a1) set[((size_t)k * ncol + j) * nrow + i] = 1;
                            }
                }
            }
        }
        auto ring = [&](int i, int j, int k) {
            return set[((size_t)k * ncol + j) * nrow + i] != 0;
        };
        for (int k = 1; k < nfil - 1; k++)
            for (int j = 1; j < ncol - 1; j++)
                if (!ring(0, j, k) || !ring(nrow - 1, j, k))
                    return fail("saveqc: fine grid not properly nested in the coarse level");
        for (int k = 1; k < nfil - 1; k++)
            for (int i = 1; i < nrow - 1; i++)
                if (!ring(i, 0, k) || !ring(i, ncol - 1, k))
                    return fail("saveqc: fine grid not properly nested in the coarse level");
        for (int j = 1; j < ncol - 1; j++)
            for (int i = 1; i < nrow - 1; i++)
                if (!ring(i, j, 0) || !ring(i, j, nfil - 1))
                    return fail("saveqc: fine grid not properly nested in the coarse level");
        Fill3Op rg{};
        rg.type = F3_RING;
        rg.dst = S;
        rg.src = S;
        rg.level = level;
        rg.si0 = g;
        rg.ni = nrow;
        rg.nj = ncol;
        rg.nk = nfil;
        rg.stage = 2;
        push(rg);
        return true;
    }
};

void finish_plan(Fill3Plan& P, int nv) {
    std::stable_sort(P.ops.begin(), P.ops.end(),
                     [](const Fill3Op& a, const Fill3Op& b) { return a.stage < b.stage; });
    P.nstages = P.ops.empty() ? 0 : P.ops.back().stage + 1;
    // split copies and interpolations into items of at most CHUNK elements
    // so that a large ghost slab is spread over several workgroups
    constexpr int CHUNK = 2048;
    P.work.clear();
    P.stage_start.assign(P.nstages + 1, 0);
    for (int o = 0; o < (int)P.ops.size(); o++) {
        const Fill3Op& op = P.ops[o];
        const int64_t ncell = (int64_t)op.ni * op.nj * op.nk;
        int64_t n = 0;
        if (op.type == F3_COPY || op.type == F3_COPYBACK || op.type == F3_COARSEN) n = ncell * nv;
        else if (op.type == F3_INTERP) n = ncell;
        if (n == 0) {
            P.work.push_back(Fill3Work{o, 0, 0, 0});
            P.stage_start[op.stage + 1]++;
            continue;
        }
        for (int64_t a = 0; a < n; a += CHUNK) {
            P.work.push_back(Fill3Work{o, (int32_t)a, (int32_t)std::min<int64_t>(a + CHUNK, n), 0});
            P.stage_start[op.stage + 1]++;
        }
    }
    for (int s = 0; s < P.nstages; s++) P.stage_start[s + 1] += P.stage_start[s];
}

bool periodic_ok(const Hier3Geom& H, std::string& err) {
    const int n = (int)H.xperdom + (int)H.yperdom + (int)H.zperdom;
    if (n != 0 && n != 3) {
        err = "3-D domains must be periodic in all three directions or in none (saveqc.f wraps "
              "only fully periodic domains)";
        return false;
    }
    return true;
}

// setuse.f (3-D): the listbc entries of fine grid `kid` around coarse grid
// box cb, for the (possibly shifted) coarsened fine box
void setuse3(std::vector<std::array<int, 6>>& out, int kid, const std::array<int, 6>& cb,
             int iclo, int ichi, int jclo, int jchi, int kclo, int kchi) {
    const int ilo = cb[0], ihi = cb[1], jlo = cb[2], jhi = cb[3], klo = cb[4], khi = cb[5];
    const int ng = NG3;
    const int ist = iclo - 1, iend = ichi + 1, jst = jclo - 1, jend = jchi + 1;
    const int kst = kclo - 1, kend = kchi + 1;
    const int ni = ichi - iclo + 1, nj = jchi - jclo + 1, nk = kchi - kclo + 1;
    if (!(ist < ilo))
        for (int k = std::max(klo, kclo); k <= std::min(khi, kchi); k++) {
            int lkid = (k - kclo) * nj + std::max(jlo, jclo) - jclo + 1;
            for (int j = std::max(jlo, jclo); j <= std::min(jhi, jchi); j++)
                out.push_back({ist - ilo + ng + 1, j - jlo + ng + 1, k - klo + ng + 1, 3, kid, lkid++});
        }
    if (!(jend > jhi))
        for (int k = std::max(klo, kclo); k <= std::min(khi, kchi); k++) {
            int lkid = nk * nj + (k - kclo) * ni + std::max(ilo, iclo) - iclo + 1;
            for (int i = std::max(ilo, iclo); i <= std::min(ihi, ichi); i++)
                out.push_back({i - ilo + ng + 1, jend - jlo + ng + 1, k - klo + ng + 1, 4, kid, lkid++});
        }
    if (!(iend > ihi))
        for (int k = std::max(klo, kclo); k <= std::min(khi, kchi); k++) {
            int lkid = nk * nj + nk * ni + (k - kclo) * nj + std::max(jlo, jclo) - jclo + 1;
            for (int j = std::max(jlo, jclo); j <= std::min(jhi, jchi); j++)
                out.push_back({iend - ilo + ng + 1, j - jlo + ng + 1, k - klo + ng + 1, 1, kid, lkid++});
        }
    if (!(jst < jlo))
        for (int k = std::max(klo, kclo); k <= std::min(khi, kchi); k++) {
            int lkid = 2 * nk * nj + nk * ni + (k - kclo) * ni + std::max(ilo, iclo) - iclo + 1;
            for (int i = std::max(ilo, iclo); i <= std::min(ihi, ichi); i++)
                out.push_back({i - ilo + ng + 1, jst - jlo + ng + 1, k - klo + ng + 1, 2, kid, lkid++});
        }
    if (!(kst < klo))
        for (int j = std::max(jlo, jclo); j <= std::min(jhi, jchi); j++) {
            int lkid = 2 * nk * nj + 2 * nk * ni + (j - jclo) * ni + std::max(ilo, iclo) - iclo + 1;
            for (int i = std::max(ilo, iclo); i <= std::min(ihi, ichi); i++)
                out.push_back({i - ilo + ng + 1, j - jlo + ng + 1, kst - klo + ng + 1, 6, kid, lkid++});
        }
    if (!(kend > khi))
        for (int j = std::max(jlo, jclo); j <= std::min(jhi, jchi); j++) {
            int lkid = 2 * nk * nj + 2 * nk * ni + nj * ni + (j - jclo) * ni + std::max(ilo, iclo) - iclo + 1;
            for (int i = std::max(ilo, iclo); i <= std::min(ihi, ichi); i++)
                out.push_back({i - ilo + ng + 1, j - jlo + ng + 1, kend - klo + ng + 1, 5, kid, lkid++});
        }
}

}  // namespace

bool plan_fill3_level(const Hier3Geom& H, int level, int nvar, Fill3Plan& P, std::string& err) {
    P = Fill3Plan();
    if (!periodic_ok(H, err)) return false;
    Planner3 pl(H, P, nvar);
    for (int g = 0; g < (int)H.lev[level].boxes.size(); g++)
        if (!pl.bound(level, g)) {
            err = pl.err.empty() ? "3-D ghost-fill planning failed" : pl.err;
            return false;
        }
    finish_plan(P, nvar);
    return true;
}

bool plan_saveqc3(const Hier3Geom& H, int level, int nvar, Fill3Plan& P, std::string& err) {
    P = Fill3Plan();
    if (!periodic_ok(H, err)) return false;
    Planner3 pl(H, P, nvar);
    for (int g = 0; g < (int)H.lev[level].boxes.size(); g++)
        if (!pl.saveqc(level, g)) {
            err = pl.err.empty() ? "3-D saveqc planning failed" : pl.err;
            return false;
        }
    finish_plan(P, nvar);
    return true;
}

bool plan_richardson3(const Hier3Geom& H, int level, int nvar, Rich3Plan& R, std::string& err) {
    R = Rich3Plan();
    if (!periodic_ok(H, err)) return false;
    Fill3Plan& P = R.fill;
    Planner3 pl(H, P, nvar);
    const Level3Geom& G = H.lev[level];
    const int ng = NG3, ngbig = 2 * NG3;
    const int n = (int)G.boxes.size();
    std::vector<int> dub(n);
    for (int g = 0; g < n; g++) {
        const auto& b = G.boxes[g];
        const int mx = b[1] - b[0] + 1, my = b[3] - b[2] + 1, mz = b[5] - b[4] + 1;
        if (mx % 2 || my % 2 || mz % 2) {
            err = "Richardson estimation needs grids of even size (errest.f: nx/2)";
            return false;
        }
        // copysol.f: the old interior into the centre of the wider array
        const int d = pl.arena(mx + 2 * ngbig, my + 2 * ngbig, mz + 2 * ngbig, R.arena_elems);
        dub[g] = d;
        Fill3Op op{};
        op.type = F3_COPY;
        op.flags = F3F_FORCE_OLD;
        op.dst = d;
        op.src = g;
        op.src_level = level;
        op.di0 = op.dj0 = op.dk0 = ngbig;
        op.ni = mx;
        op.nj = my;
        op.nk = mz;
        op.si0 = op.sj0 = op.sk0 = ng;
        op.stage = 0;
        pl.push(op);
        if (!pl.bound_into(level, g, d, ngbig)) {
            err = pl.err.empty() ? "Richardson ghost-fill planning failed" : pl.err;
            return false;
        }
    }
    int last = 0;
    for (const auto& op : P.ops) last = std::max(last, op.stage);
    for (int g = 0; g < n; g++) {
        const auto& b = G.boxes[g];
        const int mx = b[1] - b[0] + 1, my = b[3] - b[2] + 1, mz = b[5] - b[4] + 1;
        // coarsen.f into (mx/2 + 2 ng)^3
        const int c = pl.arena(mx / 2 + 2 * ng, my / 2 + 2 * ng, mz / 2 + 2 * ng, R.arena_elems);
        Fill3Op op{};
        op.type = F3_COARSEN;
        op.dst = c;
        op.src = dub[g];
        op.ni = mx / 2 + 2 * ng;
        op.nj = my / 2 + 2 * ng;
        op.nk = mz / 2 + 2 * ng;
        op.stage = last + 1;
        pl.push(op);
        // spest's locbig: the grid's new values with their ghost cells
        const int f = pl.arena(mx + 2 * ng, my + 2 * ng, mz + 2 * ng, R.arena_elems);
        Fill3Op cp{};
        cp.type = F3_COPY;
        cp.flags = F3F_FORCE_NEW;
        cp.dst = f;
        cp.src = g;
        cp.src_level = level;
        cp.ni = mx + 2 * ng;
        cp.nj = my + 2 * ng;
        cp.nk = mz + 2 * ng;
        cp.stage = 0;
        pl.push(cp);
        Patch3 pc{}, pf{};
        pc.qoff = P.bufs[c].off;
        pc.mx = mx / 2;
        pc.my = my / 2;
        pc.mz = mz / 2;
        pf.qoff = P.bufs[f].off;
        pf.mx = mx;
        pf.my = my;
        pf.mz = mz;
        R.crs.push_back(pc);
        R.fine.push_back(pf);
    }
    finish_plan(P, nvar);
    return true;
}

bool plan_registers3(const Hier3Geom& H, int nvar, Reg3Level* R, std::string& err) {
    if (!periodic_ok(H, err)) return false;
    const int ng = NG3;
    for (int l = 1; l <= H.nlev; l++) R[l] = Reg3Level();
    // fine side: prepf.f register layout and qad's slots
    for (int l = 2; l <= H.nlev; l++) {
        const Level3Geom& C = H.lev[l - 1];
        const auto& boxes = H.lev[l].boxes;
        int64_t off = 0;
        for (int g = 0; g < (int)boxes.size(); g++) {
            const auto& b = boxes[g];
            const int nx = b[1] - b[0] + 1, ny = b[3] - b[2] + 1, nz = b[5] - b[4] + 1;
            if (nx % C.ratx || ny % C.raty || nz % C.ratz || b[0] % C.ratx || b[2] % C.raty ||
                b[4] % C.ratz) {
                err = "fine grids must start and end on coarse cell boundaries";
                return false;
            }
            const int ik = nx / C.ratx, jk = ny / C.raty, kk = nz / C.ratz;
            RegGrid3 rg{};
            rg.reg_off = off;
            rg.lenbc = 2 * (ik * jk + jk * kk + kk * ik);
            R[l].regs.push_back(rg);
            off += 2 * (int64_t)nvar * rg.lenbc;
            // slots in fluxad/qad order: sides 1-4 (k-major), 5-6 (j-major)
            int slot = 0;
            const int outer[6] = {kk, kk, kk, kk, jk, jk};
            const int inner[6] = {jk, ik, jk, ik, ik, ik};
            for (int side = 1; side <= 6; side++)
                for (int a = 1; a <= outer[side - 1]; a++)
                    for (int c = 1; c <= inner[side - 1]; c++)
                        R[l].qad.push_back(Qad3Item{g, side, c, a, ++slot, 0});
        }
        R[l].reg_elems = off;
    }
    // coarse side: prepc.f listbc, fluxsv write flags, upbnd cells, update blocks
    for (int l = 1; l < H.nlev; l++) {
        const Level3Geom& G = H.lev[l];
        const Level3Geom& F = H.lev[l + 1];
        if (F.boxes.empty()) continue;
        const int rx = G.ratx, ry = G.raty, rz = G.ratz;
        const int imx = G.iregsz - 1, jmx = G.jregsz - 1, kmx = G.kregsz - 1;
        std::map<std::pair<int, int>, int> last;   // slot -> index of its last writer
        for (int m = 0; m < (int)G.boxes.size(); m++) {
            const auto& cb = G.boxes[m];
            const int ilo = cb[0], ihi = cb[1], jlo = cb[2], jhi = cb[3], klo = cb[4], khi = cb[5];
            std::vector<std::array<int, 6>> ent;
            for (int kid = 0; kid < (int)F.boxes.size(); kid++) {
                const auto& fb = F.boxes[kid];
                int iclo = fb[0] / rx, jclo = fb[2] / ry, kclo = fb[4] / rz;
                int ichi = fb[1] / rx, jchi = fb[3] / ry, kchi = fb[5] / rz;
                int iplo = std::max(ilo, iclo), jplo = std::max(jlo, jclo), kplo = std::max(klo, kclo);
                int iphi = std::min(ihi, ichi), jphi = std::min(jhi, jchi), kphi = std::min(khi, kchi);
                if (iplo <= iphi + 1 && jplo <= jphi + 1 && kplo <= kphi + 1)
                    setuse3(ent, kid, cb, iclo, ichi, jclo, jchi, kclo, kchi);
                if (H.xperdom && ilo == 0 && ichi == imx)
                    setuse3(ent, kid, cb, iclo - G.iregsz, ichi - G.iregsz, jclo, jchi, kclo, kchi);
                if (H.xperdom && iclo == 0 && ihi == imx)
                    setuse3(ent, kid, cb, iclo + G.iregsz, ichi + G.iregsz, jclo, jchi, kclo, kchi);
                if (H.yperdom && jlo == 0 && jchi == jmx)
                    setuse3(ent, kid, cb, iclo, ichi, jclo - G.jregsz, jchi - G.jregsz, kclo, kchi);
                if (H.yperdom && jclo == 0 && jhi == jmx)
                    setuse3(ent, kid, cb, iclo, ichi, jclo + G.jregsz, jchi + G.jregsz, kclo, kchi);
                if (H.zperdom && klo == 0 && kchi == kmx)
                    setuse3(ent, kid, cb, iclo, ichi, jclo, jchi, kclo - G.kregsz, kchi - G.kregsz);
                if (H.zperdom && kclo == 0 && khi == kmx)
                    setuse3(ent, kid, cb, iclo, ichi, jclo, jchi, kclo + G.kregsz, kchi + G.kregsz);
            }
            const int mitot = ihi - ilo + 1 + 2 * ng, mjtot = jhi - jlo + 1 + 2 * ng;
            const int mktot = khi - klo + 1 + 2 * ng;
            // every listed cell is an interior cell of the coarse grid and
            // every slot lies in its kid's register
            for (const auto& e : ent) {
                if (e[0] < ng + 1 || e[0] > mitot - ng || e[1] < ng + 1 || e[1] > mjtot - ng ||
                    e[2] < ng + 1 || e[2] > mktot - ng) {
                    err = "listbc entry outside its coarse grid's interior";
                    return false;
                }
                if (e[5] < 1 || e[5] > R[l + 1].regs[e[4]].lenbc) {
                    err = "listbc slot outside its fine grid's register";
                    return false;
                }
                Sv3Entry s{m, e[0], e[1], e[2], e[3], e[4], e[5], 1};
                last[{e[4], e[5]}] = (int)R[l].sv.size();
                R[l].sv.push_back(s);
            }
            // upbnd.f: each (cell, side) once, entries in list order
            std::vector<int> iused((size_t)mitot * mjtot * mktot, 0);
            std::map<int64_t, std::vector<Upbnd3Entry>> per_cell;
            std::vector<int64_t> order;
            static const double chsign[6] = {1., -1., -1., 1., 1., -1.};
            for (const auto& e : ent) {
                const int64_t c = ((int64_t)(e[2] - 1) * mjtot + (e[1] - 1)) * mitot + (e[0] - 1);
                const int norm = 1 << (e[3] - 1);
                if ((iused[c] / norm) % 2 == 1) continue;
                iused[c] += norm;
                if (!per_cell.count(c)) order.push_back(c);
                per_cell[c].push_back(Upbnd3Entry{e[4], e[5], chsign[e[3] - 1]});
            }
            for (int64_t c : order) {
                Upbnd3Cell uc{};
                uc.grid = m;
                uc.i = (int)(c % mitot);
                uc.j = (int)((c / mitot) % mjtot);
                uc.k = (int)(c / ((int64_t)mitot * mjtot));
                uc.e0 = (int)R[l].uents.size();
                for (const auto& u : per_cell[c]) R[l].uents.push_back(u);
                uc.e1 = (int)R[l].uents.size();
                R[l].ucells.push_back(uc);
            }
            // update.f: coarse cells covered by each kid
            for (int kid = 0; kid < (int)F.boxes.size(); kid++) {
                const auto& fb = F.boxes[kid];
                int iclo = fb[0] / rx, jclo = fb[2] / ry, kclo = fb[4] / rz;
                int ichi = fb[1] / rx, jchi = fb[3] / ry, kchi = fb[5] / rz;
                int iplo = std::max(ilo, iclo), jplo = std::max(jlo, jclo), kplo = std::max(klo, kclo);
                int iphi = std::min(ihi, ichi), jphi = std::min(jhi, jchi), kphi = std::min(khi, kchi);
                if (iplo > iphi || jplo > jphi || kplo > kphi) continue;
                Avg3Block ab{};
                ab.grid = m;
                ab.kid = kid;
                ab.ci0 = iplo - ilo + ng;
                ab.cj0 = jplo - jlo + ng;
                ab.ck0 = kplo - klo + ng;
                ab.ni = iphi - iplo + 1;
                ab.nj = jphi - jplo + 1;
                ab.nk = kphi - kplo + 1;
                ab.fi0 = iplo * rx - fb[0] + ng;
                ab.fj0 = jplo * ry - fb[2] + ng;
                ab.fk0 = kplo * rz - fb[4] + ng;
                const int fmx = fb[1] - fb[0] + 1 + 2 * ng, fmy = fb[3] - fb[2] + 1 + 2 * ng;
                const int fmz = fb[5] - fb[4] + 1 + 2 * ng;
                if (ab.fi0 < 0 || ab.fj0 < 0 || ab.fk0 < 0 || ab.fi0 + ab.ni * rx > fmx ||
                    ab.fj0 + ab.nj * ry > fmy || ab.fk0 + ab.nk * rz > fmz) {
                    err = "update: fine cells outside their grid";
                    return false;
                }
                ab.cell0 = R[l].avg_cells;
                R[l].avg_cells += (int64_t)ab.ni * ab.nj * ab.nk;
                R[l].avg.push_back(ab);
            }
        }
        for (size_t s = 0; s < R[l].sv.size(); s++) {
            const auto& e = R[l].sv[s];
            R[l].sv[s].write = last[{e.kid, e.lkid}] == (int)s ? 1 : 0;
        }
    }
    return true;
}

}  // namespace amrgpu
