/*
 * orc_amr3.c -- oracle for the 3-D level update around the step (TEST
 * INFRASTRUCTURE ONLY): ghost filling (bound.f, filpatch.f, intfil.f,
 * trimbd.f, prefilp.f, bc3amr.f90), the coarse-fine flux registers
 * (prepf.f, prepc.f, setuse.f, saveqc.f, icall.f, preicall.f, cstore.f,
 * qad.f, fluxsv.f, fluxad.f), refluxing and averaging (upbnd.f, update.f),
 * conck.f and flagging (spest.f's bound + flag2refine.f; errest.f with
 * copysol.f, coarsen.f and errf1.f), all of src/3d.  Plain-C restatements in the reference's
 * evaluation order; arrays in the reference's AoS layout
 * val(nvar, nrow, ncol, nfil), component fastest.
 *
 * Where the reference stops the oracle returns an error code (the codes of
 * orc_amr.c).  Guards that depart from the literal Fortran:
 *   - bc3amr on an array lying wholly outside the domain on one side has no
 *     source plane inside the array (the reference reads past it; the
 *     grid's own bc3amr overwrites those cells afterwards): the side is
 *     skipped, as in the 2-D oracle.
 *   - wall boundaries negate component 2, 3 or 4; with fewer components
 *     (advection) the reference writes outside the array: refused.
 *   - saveqc.f wraps the coarsened patch periodically only when all three
 *     directions are periodic (saveqc.f:71-81); with some but not all
 *     directions periodic it would read unset heap memory for the cells
 *     across the periodic sides: such domains are refused.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

enum { ORC_E_LEVEL1 = 11, ORC_E_SCRATCH = 12, ORC_E_TIME = 13, ORC_E_BC = 14,
       ORC_E_PERIODIC = 15, ORC_E_ALLOC = 16 };

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

/* val(m, i, j, k), 1-based, array nrow x ncol x (any) */
#define V3(a, nv, nrow, ncol, m, i, j, k) \
    (a)[((m) - 1) + (size_t)(nv) * (((i) - 1) + (size_t)(nrow) * (((j) - 1) + (size_t)(ncol) * ((k) - 1)))]

static int partial_periodic(const orc_hier3* h) {
    int n = h->xperdom + h->yperdom + h->zperdom;
    return n != 0 && n != 3;
}

/* ------------------------------------------------------------ bc3amr --- */
/* bc3amr.f90 (3-D): sides x-lo, x-hi, y-lo, y-hi, z-lo, z-hi in turn */
int orc_bc3amr(const orc_params* p, const orc_hier3* h, double* val, int nrow, int ncol,
               int nfil, double hx, double hy, double hz, double xlo_patch, double xhi_patch,
               double ylo_patch, double yhi_patch, double zlo_patch, double zhi_patch) {
    const int nv = p->meqn;
    const double hxmarg = hx * .01, hymarg = hy * .01, hzmarg = hz * .01;
    if (h->xperdom && h->yperdom && h->zperdom) return 0;
#define VV(m, i, j, k) V3(val, nv, nrow, ncol, m, i, j, k)
    /* x-lo: planes i = 1..nxl from nxl+1 (extrapolation) or mirrored (wall) */
    if (xlo_patch < h->xlower - hxmarg) {
        int nxl = (int)((h->xlower + hxmarg - xlo_patch) / hx);
        int mb = h->mthbc[0];
        if ((mb == 1 && nxl >= nrow) || (mb == 3 && 2 * nxl > nrow)) mb = 2;
        if (mb == 3 && nv < 2) return ORC_E_BC;
        if (mb == 1) {
            for (int k = 1; k <= nfil; k++)
                for (int j = 1; j <= ncol; j++)
                    for (int i = 1; i <= nxl; i++)
                        for (int m = 1; m <= nv; m++) VV(m, i, j, k) = VV(m, nxl + 1, j, k);
        } else if (mb == 3) {
            for (int k = 1; k <= nfil; k++)
                for (int j = 1; j <= ncol; j++)
                    for (int i = 1; i <= nxl; i++)
                        for (int m = 1; m <= nv; m++) VV(m, i, j, k) = VV(m, 2 * nxl + 1 - i, j, k);
            for (int k = 1; k <= nfil; k++)
                for (int j = 1; j <= ncol; j++)
                    for (int i = 1; i <= nxl; i++) VV(2, i, j, k) = -VV(2, i, j, k);
        } else if (mb != 2) {
            return ORC_E_BC;
        }
    }
    if (xhi_patch > h->xupper + hxmarg) {
        int nxr = (int)((xhi_patch - h->xupper + hxmarg) / hx);
        int ibeg = imax(nrow - nxr + 1, 1);
        int mb = h->mthbc[1];
        if ((mb == 1 && ibeg < 2) || (mb == 3 && 2 * ibeg - 1 - nrow < 1)) mb = 2;
        if (mb == 3 && nv < 2) return ORC_E_BC;
        if (mb == 1) {
            for (int k = 1; k <= nfil; k++)
                for (int i = ibeg; i <= nrow; i++)
                    for (int j = 1; j <= ncol; j++)
                        for (int m = 1; m <= nv; m++) VV(m, i, j, k) = VV(m, ibeg - 1, j, k);
        } else if (mb == 3) {
            for (int k = 1; k <= nfil; k++)
                for (int i = ibeg; i <= nrow; i++)
                    for (int j = 1; j <= ncol; j++)
                        for (int m = 1; m <= nv; m++) VV(m, i, j, k) = VV(m, 2 * ibeg - 1 - i, j, k);
            for (int k = 1; k <= nfil; k++)
                for (int i = ibeg; i <= nrow; i++)
                    for (int j = 1; j <= ncol; j++) VV(2, i, j, k) = -VV(2, i, j, k);
        } else if (mb != 2) {
            return ORC_E_BC;
        }
    }
    if (ylo_patch < h->ylower - hymarg) {
        int nyl = (int)((h->ylower + hymarg - ylo_patch) / hy);
        int mb = h->mthbc[2];
        if ((mb == 1 && nyl >= ncol) || (mb == 3 && 2 * nyl > ncol)) mb = 2;
        if (mb == 3 && nv < 3) return ORC_E_BC;
        if (mb == 1) {
            for (int k = 1; k <= nfil; k++)
                for (int j = 1; j <= nyl; j++)
                    for (int i = 1; i <= nrow; i++)
                        for (int m = 1; m <= nv; m++) VV(m, i, j, k) = VV(m, i, nyl + 1, k);
        } else if (mb == 3) {
            for (int k = 1; k <= nfil; k++)
                for (int j = 1; j <= nyl; j++)
                    for (int i = 1; i <= nrow; i++)
                        for (int m = 1; m <= nv; m++) VV(m, i, j, k) = VV(m, i, 2 * nyl + 1 - j, k);
            for (int k = 1; k <= nfil; k++)
                for (int j = 1; j <= nyl; j++)
                    for (int i = 1; i <= nrow; i++) VV(3, i, j, k) = -VV(3, i, j, k);
        } else if (mb != 2) {
            return ORC_E_BC;
        }
    }
    if (yhi_patch > h->yupper + hymarg) {
        int nyr = (int)((yhi_patch - h->yupper + hymarg) / hy);
        int jbeg = imax(ncol - nyr + 1, 1);
        int mb = h->mthbc[3];
        if ((mb == 1 && jbeg < 2) || (mb == 3 && 2 * jbeg - 1 - ncol < 1)) mb = 2;
        if (mb == 3 && nv < 3) return ORC_E_BC;
        if (mb == 1) {
            for (int k = 1; k <= nfil; k++)
                for (int j = jbeg; j <= ncol; j++)
                    for (int i = 1; i <= nrow; i++)
                        for (int m = 1; m <= nv; m++) VV(m, i, j, k) = VV(m, i, jbeg - 1, k);
        } else if (mb == 3) {
            for (int k = 1; k <= nfil; k++)
                for (int j = jbeg; j <= ncol; j++)
                    for (int i = 1; i <= nrow; i++)
                        for (int m = 1; m <= nv; m++) VV(m, i, j, k) = VV(m, i, 2 * jbeg - 1 - j, k);
            for (int k = 1; k <= nfil; k++)
                for (int j = jbeg; j <= ncol; j++)
                    for (int i = 1; i <= nrow; i++) VV(3, i, j, k) = -VV(3, i, j, k);
        } else if (mb != 2) {
            return ORC_E_BC;
        }
    }
    if (zlo_patch < h->zlower - hzmarg) {
        int nzl = (int)((h->zlower + hzmarg - zlo_patch) / hz);
        int mb = h->mthbc[4];
        if ((mb == 1 && nzl >= nfil) || (mb == 3 && 2 * nzl > nfil)) mb = 2;
        if (mb == 3 && nv < 4) return ORC_E_BC;
        if (mb == 1) {
            for (int k = 1; k <= nzl; k++)
                for (int j = 1; j <= ncol; j++)
                    for (int i = 1; i <= nrow; i++)
                        for (int m = 1; m <= nv; m++) VV(m, i, j, k) = VV(m, i, j, nzl + 1);
        } else if (mb == 3) {
            for (int k = 1; k <= nzl; k++)
                for (int j = 1; j <= ncol; j++)
                    for (int i = 1; i <= nrow; i++)
                        for (int m = 1; m <= nv; m++) VV(m, i, j, k) = VV(m, i, j, 2 * nzl + 1 - k);
            for (int k = 1; k <= nzl; k++)
                for (int j = 1; j <= ncol; j++)
                    for (int i = 1; i <= nrow; i++) VV(4, i, j, k) = -VV(4, i, j, k);
        } else if (mb != 2) {
            return ORC_E_BC;
        }
    }
    if (zhi_patch > h->zupper + hzmarg) {
        int nzr = (int)((zhi_patch - h->zupper + hzmarg) / hz);
        int kbeg = imax(nfil - nzr + 1, 1);
        int mb = h->mthbc[5];
        if ((mb == 1 && kbeg < 2) || (mb == 3 && 2 * kbeg - 1 - nfil < 1)) mb = 2;
        if (mb == 3 && nv < 4) return ORC_E_BC;
        if (mb == 1) {
            for (int k = kbeg; k <= nfil; k++)
                for (int j = 1; j <= ncol; j++)
                    for (int i = 1; i <= nrow; i++)
                        for (int m = 1; m <= nv; m++) VV(m, i, j, k) = VV(m, i, j, kbeg - 1);
        } else if (mb == 3) {
            for (int k = kbeg; k <= nfil; k++)
                for (int j = 1; j <= ncol; j++)
                    for (int i = 1; i <= nrow; i++)
                        for (int m = 1; m <= nv; m++) VV(m, i, j, k) = VV(m, i, j, 2 * kbeg - 1 - k);
            for (int k = kbeg; k <= nfil; k++)
                for (int j = 1; j <= ncol; j++)
                    for (int i = 1; i <= nrow; i++) VV(4, i, j, k) = -VV(4, i, j, k);
        } else if (mb != 2) {
            return ORC_E_BC;
        }
    }
#undef VV
    return 0;
}

/* ------------------------------------------------------------- trimbd --- */
/* trimbd.f (3-D): set = all cells used; else the smallest box holding the
 * unused cells, r = (il, ir, jl, jr, kl, kr), 1-based in the patch */
static void trimbd3(const signed char* used, int nrow, int ncol, int nfil, int* set, int r[6]) {
    long utot = 0;
    const size_t n = (size_t)nrow * ncol * nfil;
    for (size_t c = 0; c < n; c++) utot += used[c];
    if (utot >= (long)n) { *set = 1; return; }
    *set = 0;
    int il = nrow + 1, ir = 0, jl = ncol + 1, jr = 0, kl = nfil + 1, kr = 0;
    for (int k = 1; k <= nfil; k++)
        for (int j = 1; j <= ncol; j++)
            for (int i = 1; i <= nrow; i++)
                if (!used[(i - 1) + (size_t)nrow * ((j - 1) + (size_t)ncol * (k - 1))]) {
                    il = imin(il, i); ir = imax(ir, i);
                    jl = imin(jl, j); jr = imax(jr, j);
                    kl = imin(kl, k); kr = imax(kr, k);
                }
    r[0] = il; r[1] = ir; r[2] = jl; r[3] = jr; r[4] = kl; r[5] = kr;
}

/* ------------------------------------------------------------- intfil --- */
/* intfil.f (3-D).  flaguse(ilo:ihi, jlo:jhi, klo:khi).  The time tolerance
 * is possk(level)/10 (intfil.f:58-59), not the 2-D code's 1e-4. */
static int intfil3(const orc_params* p, const orc_hier3* h, double* val, int mi, int mj,
                   double time, signed char* flaguse, int nrowst, int ncolst, int nfilst,
                   int ilo, int ihi, int jlo, int jhi, int klo, int khi, int level) {
    const int nv = p->meqn, ng = p->mbc;
    const int fw = ihi - ilo + 1, fh = jhi - jlo + 1;
#define FU(i, j, k) flaguse[((i) - ilo) + (size_t)fw * (((j) - jlo) + (size_t)fh * ((k) - klo))]
    const double dt = h->possk[level];
    const double teps = dt / 10.;
    memset(flaguse, 0, (size_t)fw * fh * (khi - klo + 1));
    /* the grids of the level (for a grid's own strips the reference scans
     * its bndList, a superset of the grids that can meet the strip) */
    for (int g = 0; g < h->ngrids[level]; g++) {
        const int* b = h->boxes[level] + 6 * g;
        int ixlo = imax(b[0], ilo), ixhi = imin(b[1], ihi);
        int jxlo = imax(b[2], jlo), jxhi = imin(b[3], jhi);
        int kxlo = imax(b[4], klo), kxhi = imin(b[5], khi);
        if (!(ixlo <= ixhi && jxlo <= jxhi && kxlo <= kxhi)) continue;
        int mitot = b[1] - b[0] + 1 + 2 * ng, mjtot = b[3] - b[2] + 1 + 2 * ng;
        double alphac = (h->time[level] - time) / dt;
        double alphai = 1.0 - alphac;
        if (alphai < -teps || alphai > 1.0 + teps) return ORC_E_TIME;
        const double* qn = h->qnew[level] + h->qoff[level][g];
        const double* qo = h->qold[level] + h->qoff[level][g];
        int interp = 0;
        const double* src = qn;
        if (fabs(alphai - 1.0) < teps) {
            src = qn;
        } else if (fabs(alphai) < teps) {
            src = qo;
            if (level == h->mxnest) return ORC_E_TIME;
        } else {
            interp = 1;
            if (level == h->mxnest) return ORC_E_TIME;
        }
        for (int k = kxlo; k <= kxhi; k++)
            for (int j = jxlo; j <= jxhi; j++)
                for (int i = ixlo; i <= ixhi; i++) {
                    for (int m = 1; m <= nv; m++) {
                        size_t s = (size_t)(m - 1) + (size_t)nv * ((i - b[0] + ng) +
                                   (size_t)mitot * ((j - b[2] + ng) + (size_t)mjtot * (k - b[4] + ng)));
                        double v = interp ? qn[s] * alphai + qo[s] * alphac : src[s];
                        V3(val, nv, mi, mj, m, i - ilo + nrowst, j - jlo + ncolst, k - klo + nfilst) = v;
                    }
                    FU(i, j, k) = 1;
                }
    }
    /* cells outside the domain are left to the boundary conditions
     * (intfil.f:207-262, the reference's ranges, clipped to the patch) */
    if (khi >= h->kregsz[level])
        for (int k = imax(h->kregsz[level], klo); k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++) FU(i, j, k) = 1;
    if (klo < 0)
        for (int k = klo; k <= imin(imin(-1, nfilst + khi - klo), khi); k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++) FU(i, j, k) = 1;
    if (jhi >= h->jregsz[level])
        for (int k = klo; k <= khi; k++)
            for (int j = imax(h->jregsz[level], jlo); j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++) FU(i, j, k) = 1;
    if (jlo < 0)
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= imin(imin(-1, ncolst + jhi - jlo), jhi); j++)
                for (int i = ilo; i <= ihi; i++) FU(i, j, k) = 1;
    if (ilo < 0)
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= imin(imin(-1, nrowst + ihi - ilo), ihi); i++) FU(i, j, k) = 1;
    if (ihi >= h->iregsz[level])
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = imax(h->iregsz[level], ilo); i <= ihi; i++) FU(i, j, k) = 1;
#undef FU
    return 0;
}

static int prefilrecur3(const orc_params* p, const orc_hier3* h, int level, double* valbig,
                        int mitot, int mjtot, double time, int ilo, int ihi, int jlo, int jhi,
                        int klo, int khi, int iglo, int jglo, int kglo);

/* ----------------------------------------------------------- filrecur --- */
/* filpatch.f (3-D) filrecur: intfil at this level, then the unset box from
 * the coarser level (recursively) with MC-limited slopes in x, y and z */
static int filrecur3(const orc_params* p, const orc_hier3* h, int level, double* valbig,
                     int mitot, int mjtot, double t, int nrowst, int ncolst, int nfilst,
                     int ilo, int ihi, int jlo, int jhi, int klo, int khi, int patchOnly) {
    const int nv = p->meqn, ng = p->mbc;
    int rc = 0;
    const int nrowp = ihi - ilo + 1, ncolp = jhi - jlo + 1, nfilp = khi - klo + 1;
    const double hxf = h->hx[level], hyf = h->hy[level], hzf = h->hz[level];
    signed char* flaguse = malloc((size_t)nrowp * ncolp * nfilp);
    if (!flaguse) return ORC_E_ALLOC;
    rc = intfil3(p, h, valbig, mitot, mjtot, t, flaguse, nrowst, ncolst, nfilst, ilo, ihi, jlo,
                 jhi, klo, khi, level);
    if (rc) { free(flaguse); return rc; }
    int set, un[6] = {0, 0, 0, 0, 0, 0};
    trimbd3(flaguse, nrowp, ncolp, nfilp, &set, un);
    if (!set) {
        if (level == 1) { free(flaguse); return ORC_E_LEVEL1; }
        const int levc = level - 1;
        const double hxc = h->hx[levc], hyc = h->hy[levc], hzc = h->hz[levc];
        /* trimbd's box in global indices */
        int il = un[0] + ilo - 1, ir = un[1] + ilo - 1;
        int jl = un[2] + jlo - 1, jr = un[3] + jlo - 1;
        int kl = un[4] + klo - 1, kr = un[5] + klo - 1;
        const int rx = h->intratx[levc], ry = h->intraty[levc], rz = h->intratz[levc];
        int iplo = (il - rx + ng * rx) / rx - ng, iphi = (ir + rx) / rx;
        int jplo = (jl - ry + ng * ry) / ry - ng, jphi = (jr + ry) / ry;
        int kplo = (kl - rz + ng * rz) / rz - ng, kphi = (kr + rz) / rz;
        double xlc = h->xlower + iplo * hxc, ybc = h->ylower + jplo * hyc, zfc = h->zlower + kplo * hzc;
        int nrowc = iphi - iplo + 1, ncolc = jphi - jplo + 1, nfilc = kphi - kplo + 1;
        if (nrowc > ihi - ilo + 3 || ncolc > jhi - jlo + 3 || nfilc > khi - klo + 3) {
            free(flaguse);
            return ORC_E_SCRATCH;
        }
        double* valcrse = calloc((size_t)nrowc * ncolc * nfilc * nv, sizeof(double));
        if (!valcrse) { free(flaguse); return ORC_E_ALLOC; }
        int sticks = iplo < 0 || jplo < 0 || kplo < 0 || iphi >= h->iregsz[levc] ||
                     jphi >= h->jregsz[levc] || kphi >= h->kregsz[levc];
        if ((h->xperdom || h->yperdom || h->zperdom) && sticks)
            rc = prefilrecur3(p, h, levc, valcrse, nrowc, ncolc, t, iplo, iphi, jplo, jphi, kplo,
                              kphi, iplo, jplo, kplo);
        else
            rc = filrecur3(p, h, levc, valcrse, nrowc, ncolc, t, 1, 1, 1, iplo, iphi, jplo, jphi,
                           kplo, kphi, 1);
        if (rc) { free(valcrse); free(flaguse); return rc; }
        const double ratiox = (double)rx, ratioy = (double)ry, ratioz = (double)rz;
#define VC(n, i, j, k) V3(valcrse, nv, nrowc, ncolc, n, i, j, k)
        for (int iff = 1; iff <= nrowp; iff++) {
            int ic = (int)floor((iff + ilo - 1) / ratiox) - iplo + 1;
            double xcent_coarse = xlc + (ic - .5) * hxc;
            double xcent_fine = h->xlower + (iff - 1 + ilo + .5) * hxf;
            double eta1 = (xcent_fine - xcent_coarse) / hxc;
            for (int jf = 1; jf <= ncolp; jf++) {
                int jc = (int)floor((jf + jlo - 1) / ratioy) - jplo + 1;
                double ycent_coarse = ybc + (jc - .5) * hyc;
                double ycent_fine = h->ylower + (jf - 1 + jlo + .5) * hyf;
                double eta2 = (ycent_fine - ycent_coarse) / hyc;
                for (int kf = 1; kf <= nfilp; kf++) {
                    int kc = (int)floor((kf + klo - 1) / ratioz) - kplo + 1;
                    double zcent_coarse = zfc + (kc - .5) * hzc;
                    double zcent_fine = h->zlower + (kf - 1 + klo + .5) * hzf;
                    double eta3 = (zcent_fine - zcent_coarse) / hzc;
                    if (flaguse[(iff - 1) + (size_t)nrowp * ((jf - 1) + (size_t)ncolp * (kf - 1))]) continue;
                    for (int n = 1; n <= nv; n++) {
                        double valp100 = VC(n, ic + 1, jc, kc), valm100 = VC(n, ic - 1, jc, kc);
                        double valc = VC(n, ic, jc, kc);
                        double valp010 = VC(n, ic, jc + 1, kc), valm010 = VC(n, ic, jc - 1, kc);
                        double valp001 = VC(n, ic, jc, kc + 1), valm001 = VC(n, ic, jc, kc - 1);
                        double dupc = valp100 - valc, dumc = valc - valm100, ducc = valp100 - valm100;
                        double du = fmin(fabs(dupc), fabs(dumc));
                        du = fmin(2.0 * du, .5 * fabs(ducc));
                        double fu = fmax(0.0, copysign(1.0, dupc * dumc));
                        double uslope = du * copysign(1.0, ducc) * fu;
                        double dvpc = valp010 - valc, dvmc = valc - valm010, dvcc = valp010 - valm010;
                        double dv = fmin(fabs(dvpc), fabs(dvmc));
                        dv = fmin(2.0 * dv, .5 * fabs(dvcc));
                        double fv = fmax(0.0, copysign(1.0, dvpc * dvmc));
                        double vslope = dv * copysign(1.0, dvcc) * fv;
                        double dwpc = valp001 - valc, dwmc = valc - valm001, dwcc = valp001 - valm001;
                        double dw = fmin(fabs(dwpc), fabs(dwmc));
                        dw = fmin(2.0 * dw, .5 * fabs(dwcc));
                        double fw = fmax(0.0, copysign(1.0, dwpc * dwmc));
                        double wslope = dw * copysign(1.0, dwcc) * fw;
                        double valint = valc + eta1 * uslope + eta2 * vslope + eta3 * wslope;
                        V3(valbig, nv, mitot, mjtot, n, iff + nrowst - 1, jf + ncolst - 1,
                           kf + nfilst - 1) = valint;
                    }
                }
            }
        }
#undef VC
        free(valcrse);
    }
    free(flaguse);
    if (patchOnly) {
        /* the patch array is the whole of valbig here (nrowst = ... = 1) */
        double xlp = h->xlower + ilo * hxf, xrp = h->xlower + (ihi + 1) * hxf;
        double ybp = h->ylower + jlo * hyf, ytp = h->ylower + (jhi + 1) * hyf;
        double zfp = h->zlower + klo * hzf, zrp = h->zlower + (khi + 1) * hzf;
        int nfil_all = nfilp + nfilst - 1;
        rc = orc_bc3amr(p, h, valbig, mitot, mjtot, nfil_all, hxf, hyf, hzf, xlp, xrp, ybp, ytp,
                        zfp, zrp);
    }
    return rc;
}

/* -------------------------------------------------------- prefilrecur --- */
/* prefilp.f (3-D): the patch split into the (up to 27) pieces inside and
 * across the periodic sides, each filled from its periodic image through a
 * temporary patch and copied back (patchCopyOut) to valbig, whose cell
 * (1,1,1) is global (iglo, jglo, kglo) */
static int prefilrecur3(const orc_params* p, const orc_hier3* h, int level, double* valbig,
                        int mitot, int mjtot, double time, int ilo, int ihi, int jlo, int jhi,
                        int klo, int khi, int iglo, int jglo, int kglo) {
    const int nv = p->meqn;
    int st[3][3], en[3][3], sh[3][3];
    const int lo[3] = {ilo, jlo, klo}, hi[3] = {ihi, jhi, khi};
    const int per[3] = {h->xperdom, h->yperdom, h->zperdom};
    const int sz[3] = {h->iregsz[level], h->jregsz[level], h->kregsz[level]};
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
    for (int a = 0; a < 3; a++) {
        int i1 = imax(ilo, st[0][a]), i2 = imin(ihi, en[0][a]);
        if (i1 > i2) continue;
        for (int b = 0; b < 3; b++) {
            int j1 = imax(jlo, st[1][b]), j2 = imin(jhi, en[1][b]);
            if (j1 > j2) continue;
            for (int c = 0; c < 3; c++) {
                int k1 = imax(klo, st[2][c]), k2 = imin(khi, en[2][c]);
                if (k1 > k2) continue;
                int mi = i2 - i1 + 1, mj = j2 - j1 + 1, mk = k2 - k1 + 1;
                double* vp = calloc((size_t)mi * mj * mk * nv, sizeof(double));
                if (!vp) return ORC_E_ALLOC;
                int rc = filrecur3(p, h, level, vp, mi, mj, time, 1, 1, 1, i1 + sh[0][a],
                                   i2 + sh[0][a], j1 + sh[1][b], j2 + sh[1][b], k1 + sh[2][c],
                                   k2 + sh[2][c], 1);
                if (rc) { free(vp); return rc; }
                int ist = i1 - iglo, jst = j1 - jglo, kst = k1 - kglo;
                for (int k = 1; k <= mk; k++)
                    for (int j = 1; j <= mj; j++)
                        for (int i = 1; i <= mi; i++)
                            for (int m = 1; m <= nv; m++)
                                V3(valbig, nv, mitot, mjtot, m, ist + i, jst + j, kst + k) =
                                    V3(vp, nv, mi, mj, m, i, j, k);
                free(vp);
            }
        }
    }
    return 0;
}

/* -------------------------------------------------------------- bound --- */
/* bound.f (3-D): left, right (all j, interior k), front, rear (interior i,
 * interior k), bottom, top (all i and j), then bc3amr on the whole grid */
int orc_bound3(const orc_params* p, const orc_hier3* h, int level, int grid, double time,
               double* valbig) {
    return orc_bound3_ng(p, h, level, grid, time, valbig, p->mbc);
}

/* bound.f (3-D) with a ghost ring of width ng around the grid (the
 * storage keeps p->mbc = nghost; errest.f calls bound with 2 nghost) */
int orc_bound3_ng(const orc_params* p, const orc_hier3* h, int level, int grid, double time,
                  double* valbig, int ng) {
    if (partial_periodic(h)) return ORC_E_PERIODIC;
    const int* b = h->boxes[level] + 6 * grid;
    const int ilo = b[0], ihi = b[1], jlo = b[2], jhi = b[3], klo = b[4], khi = b[5];
    const int mitot = ihi - ilo + 1 + 2 * ng, mjtot = jhi - jlo + 1 + 2 * ng;
    const int mktot = khi - klo + 1 + 2 * ng;
    const double hx = h->hx[level], hy = h->hy[level], hz = h->hz[level];
    const double xlo = ilo * hx + h->xlower, xhi = (ihi + 1) * hx + h->xlower;
    const double ylo = jlo * hy + h->ylower, yhi = (jhi + 1) * hy + h->ylower;
    const double zlo = klo * hz + h->zlower, zhi = (khi + 1) * hz + h->zlower;
    const int ig = ilo - ng, jg = jlo - ng, kg = klo - ng;
    int rc;
    double xl, xr, yf, yr, zb, zt;
    int xs, ys;
    /* left */
    xl = xlo - ng * hx; yf = ylo - ng * hy; yr = yhi + ng * hy;
    ys = yf < h->ylower || yr > h->yupper;
    if ((h->xperdom && xl < h->xlower) || (h->yperdom && ys))
        rc = prefilrecur3(p, h, level, valbig, mitot, mjtot, time, ilo - ng, ilo - 1, jlo - ng,
                          jhi + ng, klo, khi, ig, jg, kg);
    else
        rc = filrecur3(p, h, level, valbig, mitot, mjtot, time, 1, 1, ng + 1, ilo - ng, ilo - 1,
                       jlo - ng, jhi + ng, klo, khi, 0);
    if (rc) return rc;
    /* right */
    xr = xhi + ng * hx;
    if ((h->xperdom && xr > h->xupper) || (h->yperdom && ys))
        rc = prefilrecur3(p, h, level, valbig, mitot, mjtot, time, ihi + 1, ihi + ng, jlo - ng,
                          jhi + ng, klo, khi, ig, jg, kg);
    else
        rc = filrecur3(p, h, level, valbig, mitot, mjtot, time, mitot - ng + 1, 1, ng + 1,
                       ihi + 1, ihi + ng, jlo - ng, jhi + ng, klo, khi, 0);
    if (rc) return rc;
    /* front */
    if (h->yperdom && yf < h->ylower)
        rc = prefilrecur3(p, h, level, valbig, mitot, mjtot, time, ilo, ihi, jlo - ng, jlo - 1,
                          klo, khi, ig, jg, kg);
    else
        rc = filrecur3(p, h, level, valbig, mitot, mjtot, time, ng + 1, 1, ng + 1, ilo, ihi,
                       jlo - ng, jlo - 1, klo, khi, 0);
    if (rc) return rc;
    /* rear */
    if (h->yperdom && yr > h->yupper)
        rc = prefilrecur3(p, h, level, valbig, mitot, mjtot, time, ilo, ihi, jhi + 1, jhi + ng,
                          klo, khi, ig, jg, kg);
    else
        rc = filrecur3(p, h, level, valbig, mitot, mjtot, time, ng + 1, mjtot - ng + 1, ng + 1,
                       ilo, ihi, jhi + 1, jhi + ng, klo, khi, 0);
    if (rc) return rc;
    /* bottom */
    xl = xlo - ng * hx; xr = xhi + ng * hx;
    zb = zlo - ng * hz; zt = zhi + ng * hz;
    xs = xl < h->xlower || xr > h->xupper;
    if ((h->zperdom && zb < h->zlower) || (h->xperdom && xs) || (h->yperdom && ys))
        rc = prefilrecur3(p, h, level, valbig, mitot, mjtot, time, ilo - ng, ihi + ng, jlo - ng,
                          jhi + ng, klo - ng, klo - 1, ig, jg, kg);
    else
        rc = filrecur3(p, h, level, valbig, mitot, mjtot, time, 1, 1, 1, ilo - ng, ihi + ng,
                       jlo - ng, jhi + ng, klo - ng, klo - 1, 0);
    if (rc) return rc;
    /* top */
    if ((h->zperdom && zt > h->zupper) || (h->xperdom && xs) || (h->yperdom && ys))
        rc = prefilrecur3(p, h, level, valbig, mitot, mjtot, time, ilo - ng, ihi + ng, jlo - ng,
                          jhi + ng, khi + 1, khi + ng, ig, jg, kg);
    else
        rc = filrecur3(p, h, level, valbig, mitot, mjtot, time, 1, 1, mktot - ng + 1, ilo - ng,
                       ihi + ng, jlo - ng, jhi + ng, khi + 1, khi + ng, 0);
    if (rc) return rc;
    return orc_bc3amr(p, h, valbig, mitot, mjtot, mktot, hx, hy, hz, xlo - ng * hx,
                      xhi + ng * hx, ylo - ng * hy, yhi + ng * hy, zlo - ng * hz, zhi + ng * hz);
}

/* ======================================================= registers ====== */

static int push6(int** a, int* n, int* cap, const int e[6]) {
    if (*n + 1 > *cap) {
        int nc = *cap ? 2 * *cap : 64;
        int* na = realloc(*a, (size_t)nc * 6 * sizeof(int));
        if (!na) return ORC_E_ALLOC;
        *a = na;
        *cap = nc;
    }
    memcpy(*a + 6 * (size_t)(*n), e, 6 * sizeof(int));
    (*n)++;
    return 0;
}

/* setuse.f (3-D): the coarse cells around fine grid mkid (coarsened box
 * iclo..kchi) that lie in coarse grid ilo..khi, side by side: left of the
 * fine grid (coarse side 3), rear (4), right (1), front (2), below (6),
 * above (5), with the slot lkid of the fine grid's register */
static int setuse3(int** lst, int* n, int* cap, int mkid, int ng, int ilo, int ihi, int jlo,
                   int jhi, int klo, int khi, int iclo, int ichi, int jclo, int jchi, int kclo,
                   int kchi) {
    const int ist = iclo - 1, iend = ichi + 1, jst = jclo - 1, jend = jchi + 1;
    const int kst = kclo - 1, kend = kchi + 1;
    const int ni = ichi - iclo + 1, nj = jchi - jclo + 1, nk = kchi - kclo + 1;
    int e[6], rc;
    if (!(ist < ilo))
        for (int k = imax(klo, kclo); k <= imin(khi, kchi); k++) {
            int lkid = (k - kclo) * nj + imax(jlo, jclo) - jclo + 1;
            for (int j = imax(jlo, jclo); j <= imin(jhi, jchi); j++) {
                e[0] = ist - ilo + ng + 1; e[1] = j - jlo + ng + 1; e[2] = k - klo + ng + 1;
                e[3] = 3; e[4] = mkid; e[5] = lkid++;
                if ((rc = push6(lst, n, cap, e))) return rc;
            }
        }
    if (!(jend > jhi))
        for (int k = imax(klo, kclo); k <= imin(khi, kchi); k++) {
            int lkid = nk * nj + (k - kclo) * ni + imax(ilo, iclo) - iclo + 1;
            for (int i = imax(ilo, iclo); i <= imin(ihi, ichi); i++) {
                e[0] = i - ilo + ng + 1; e[1] = jend - jlo + ng + 1; e[2] = k - klo + ng + 1;
                e[3] = 4; e[4] = mkid; e[5] = lkid++;
                if ((rc = push6(lst, n, cap, e))) return rc;
            }
        }
    if (!(iend > ihi))
        for (int k = imax(klo, kclo); k <= imin(khi, kchi); k++) {
            int lkid = nk * nj + nk * ni + (k - kclo) * nj + imax(jlo, jclo) - jclo + 1;
            for (int j = imax(jlo, jclo); j <= imin(jhi, jchi); j++) {
                e[0] = iend - ilo + ng + 1; e[1] = j - jlo + ng + 1; e[2] = k - klo + ng + 1;
                e[3] = 1; e[4] = mkid; e[5] = lkid++;
                if ((rc = push6(lst, n, cap, e))) return rc;
            }
        }
    if (!(jst < jlo))
        for (int k = imax(klo, kclo); k <= imin(khi, kchi); k++) {
            int lkid = nk * nj + nk * ni + nk * nj + (k - kclo) * ni + imax(ilo, iclo) - iclo + 1;
            for (int i = imax(ilo, iclo); i <= imin(ihi, ichi); i++) {
                e[0] = i - ilo + ng + 1; e[1] = jst - jlo + ng + 1; e[2] = k - klo + ng + 1;
                e[3] = 2; e[4] = mkid; e[5] = lkid++;
                if ((rc = push6(lst, n, cap, e))) return rc;
            }
        }
    if (!(kst < klo))
        for (int j = imax(jlo, jclo); j <= imin(jhi, jchi); j++) {
            int lkid = 2 * nk * nj + 2 * nk * ni + (j - jclo) * ni + imax(ilo, iclo) - iclo + 1;
            for (int i = imax(ilo, iclo); i <= imin(ihi, ichi); i++) {
                e[0] = i - ilo + ng + 1; e[1] = j - jlo + ng + 1; e[2] = kst - klo + ng + 1;
                e[3] = 6; e[4] = mkid; e[5] = lkid++;
                if ((rc = push6(lst, n, cap, e))) return rc;
            }
        }
    if (!(kend > khi))
        for (int j = imax(jlo, jclo); j <= imin(jhi, jchi); j++) {
            int lkid = 2 * nk * nj + 2 * nk * ni + nj * ni + (j - jclo) * ni + imax(ilo, iclo) - iclo + 1;
            for (int i = imax(ilo, iclo); i <= imin(ihi, ichi); i++) {
                e[0] = i - ilo + ng + 1; e[1] = j - jlo + ng + 1; e[2] = kend - klo + ng + 1;
                e[3] = 5; e[4] = mkid; e[5] = lkid++;
                if ((rc = push6(lst, n, cap, e))) return rc;
            }
        }
    return 0;
}

void orc_free_regs3(orc_regs3* r) {
    for (int l = 0; l < ORC_MAXLV; l++) {
        free(r->lenbc[l]); free(r->roff[l]); free(r->reg[l]);
        free(r->lst_start[l]); free(r->lst_count[l]); free(r->listbc[l]);
    }
    memset(r, 0, sizeof *r);
}

/* prepf.f (3-D) for fine levels, prepc.f (3-D) for coarse levels */
int orc_prep_regs3(const orc_params* p, const orc_hier3* h, orc_regs3* r) {
    const int nv = p->meqn, ng = p->mbc;
    memset(r, 0, sizeof *r);
    if (partial_periodic(h)) return ORC_E_PERIODIC;
    r->nlev = h->nlev;
    for (int l = 2; l <= h->nlev; l++) {
        int n = h->ngrids[l];
        r->lenbc[l] = calloc((size_t)n + 1, sizeof(int));
        r->roff[l] = calloc((size_t)n + 1, sizeof(int64_t));
        if (!r->lenbc[l] || !r->roff[l]) return ORC_E_ALLOC;
        int64_t tot = 0;
        for (int k = 0; k < n; k++) {
            const int* b = h->boxes[l] + 6 * k;
            int ik = (b[1] - b[0] + 1) / h->intratx[l - 1];
            int jk = (b[3] - b[2] + 1) / h->intraty[l - 1];
            int kk = (b[5] - b[4] + 1) / h->intratz[l - 1];
            r->lenbc[l][k] = 2 * (ik * jk + jk * kk + kk * ik);
            r->roff[l][k] = tot;
            tot += 2 * (int64_t)nv * r->lenbc[l][k];
        }
        r->nreg[l] = tot;
        r->reg[l] = calloc((size_t)tot + 1, sizeof(double));
        if (!r->reg[l]) return ORC_E_ALLOC;
    }
    for (int l = 1; l < h->nlev; l++) {
        if (h->ngrids[l + 1] == 0) continue;
        int nc = h->ngrids[l], cap = 0, n = 0;
        int* lst = NULL;
        r->lst_start[l] = calloc((size_t)nc + 1, sizeof(int));
        r->lst_count[l] = calloc((size_t)nc + 1, sizeof(int));
        if (!r->lst_start[l] || !r->lst_count[l]) return ORC_E_ALLOC;
        const int imx = h->iregsz[l] - 1, jmx = h->jregsz[l] - 1, kmx = h->kregsz[l] - 1;
        const int rx = h->intratx[l], ry = h->intraty[l], rz = h->intratz[l];
        const int isz = h->iregsz[l], jsz = h->jregsz[l], ksz = h->kregsz[l];
        for (int m = 0; m < nc; m++) {
            const int* cb = h->boxes[l] + 6 * m;
            const int ilo = cb[0], ihi = cb[1], jlo = cb[2], jhi = cb[3], klo = cb[4], khi = cb[5];
            r->lst_start[l][m] = n;
            for (int kid = 0; kid < h->ngrids[l + 1]; kid++) {
                const int* fb = h->boxes[l + 1] + 6 * kid;
                int iclo = fb[0] / rx, jclo = fb[2] / ry, kclo = fb[4] / rz;
                int ichi = fb[1] / rx, jchi = fb[3] / ry, kchi = fb[5] / rz;
                int iplo = imax(ilo, iclo), jplo = imax(jlo, jclo), kplo = imax(klo, kclo);
                int iphi = imin(ihi, ichi), jphi = imin(jhi, jchi), kphi = imin(khi, kchi);
                int rc = 0;
#define SU(a, b, c, d, e, f) setuse3(&lst, &n, &cap, kid, ng, ilo, ihi, jlo, jhi, klo, khi, a, b, c, d, e, f)
                if (iplo <= iphi + 1 && jplo <= jphi + 1 && kplo <= kphi + 1)
                    rc = SU(iclo, ichi, jclo, jchi, kclo, kchi);
                if (!rc && h->xperdom && ilo == 0 && ichi == imx)
                    rc = SU(iclo - isz, ichi - isz, jclo, jchi, kclo, kchi);
                if (!rc && h->xperdom && iclo == 0 && ihi == imx)
                    rc = SU(iclo + isz, ichi + isz, jclo, jchi, kclo, kchi);
                if (!rc && h->yperdom && jlo == 0 && jchi == jmx)
                    rc = SU(iclo, ichi, jclo - jsz, jchi - jsz, kclo, kchi);
                if (!rc && h->yperdom && jclo == 0 && jhi == jmx)
                    rc = SU(iclo, ichi, jclo + jsz, jchi + jsz, kclo, kchi);
                if (!rc && h->zperdom && klo == 0 && kchi == kmx)
                    rc = SU(iclo, ichi, jclo, jchi, kclo - ksz, kchi - ksz);
                if (!rc && h->zperdom && kclo == 0 && khi == kmx)
                    rc = SU(iclo, ichi, jclo, jchi, kclo + ksz, kchi + ksz);
#undef SU
                if (rc) { free(lst); return rc; }
            }
            r->lst_count[l][m] = n - r->lst_start[l][m];
        }
        r->listbc[l] = lst;
        r->nlist[l] = n;
    }
    return 0;
}

/* icall.f (3-D): the level's grids (store1, interiors) over the box into
 * val, box corner at 1-based (iputst, jputst, kputst) */
static void icall3(const orc_params* p, const orc_hier3* h, double* val, int nrow, int ncol,
                   int ilo, int ihi, int jlo, int jhi, int klo, int khi, int level, int iputst,
                   int jputst, int kputst) {
    const int nv = p->meqn, ng = p->mbc;
    for (int g = 0; g < h->ngrids[level]; g++) {
        const int* b = h->boxes[level] + 6 * g;
        int ixlo = imax(b[0], ilo), ixhi = imin(b[1], ihi);
        int jxlo = imax(b[2], jlo), jxhi = imin(b[3], jhi);
        int kxlo = imax(b[4], klo), kxhi = imin(b[5], khi);
        if (!(ixlo <= ixhi && jxlo <= jxhi && kxlo <= kxhi)) continue;
        int mitot = b[1] - b[0] + 1 + 2 * ng, mjtot = b[3] - b[2] + 1 + 2 * ng;
        const double* q = h->qnew[level] + h->qoff[level][g];
        for (int k = kxlo; k <= kxhi; k++)
            for (int j = jxlo; j <= jxhi; j++)
                for (int i = ixlo; i <= ixhi; i++)
                    for (int m = 1; m <= nv; m++)
                        V3(val, nv, nrow, ncol, m, i - ilo + iputst, j - jlo + jputst, k - klo + kputst) =
                            V3(q, nv, mitot, mjtot, m, i - b[0] + ng + 1, j - b[2] + ng + 1, k - b[4] + ng + 1);
    }
}

/* preicall.f (3-D), used by saveqc only when every direction is periodic */
static void preicall3(const orc_params* p, const orc_hier3* h, double* val, int nrow, int ncol,
                      int ilo, int ihi, int jlo, int jhi, int klo, int khi, int level) {
    const int lo[3] = {ilo, jlo, klo}, hi[3] = {ihi, jhi, khi};
    const int sz[3] = {h->iregsz[level], h->jregsz[level], h->kregsz[level]};
    int st[3][3], en[3][3], sh[3][3];
    for (int d = 0; d < 3; d++) {
        st[d][0] = lo[d]; st[d][1] = 0; st[d][2] = sz[d];
        en[d][0] = -1; en[d][1] = sz[d] - 1; en[d][2] = hi[d];
        sh[d][0] = sz[d]; sh[d][1] = 0; sh[d][2] = -sz[d];
    }
    for (int a = 0; a < 3; a++) {
        int i1 = imax(ilo, st[0][a]), i2 = imin(ihi, en[0][a]);
        if (i1 > i2) continue;
        for (int b = 0; b < 3; b++) {
            int j1 = imax(jlo, st[1][b]), j2 = imin(jhi, en[1][b]);
            if (j1 > j2) continue;
            for (int c = 0; c < 3; c++) {
                int k1 = imax(klo, st[2][c]), k2 = imin(khi, en[2][c]);
                if (k1 > k2) continue;
                icall3(p, h, val, nrow, ncol, i1 + sh[0][a], i2 + sh[0][a], j1 + sh[1][b],
                       j2 + sh[1][b], k1 + sh[2][c], k2 + sh[2][c], level, i1 - ilo + 1,
                       j1 - jlo + 1, k1 - klo + 1);
            }
        }
    }
}

/* saveqc.f (3-D) and cstore.f (3-D) for the grids of fine level `level` */
int orc_saveqc3(const orc_params* p, const orc_hier3* h, orc_regs3* r, int level) {
    const int nv = p->meqn, levc = level - 1;
    const double hxc = h->hx[levc], hyc = h->hy[levc], hzc = h->hz[levc];
    const int rx = h->intratx[levc], ry = h->intraty[levc], rz = h->intratz[levc];
    const int perdom3 = h->xperdom && h->yperdom && h->zperdom;
    if (partial_periodic(h)) return ORC_E_PERIODIC;
    for (int g = 0; g < h->ngrids[level]; g++) {
        const int* b = h->boxes[level] + 6 * g;
        const int lenbc = r->lenbc[level][g];
        int iclo = b[0] / rx - 1, jclo = b[2] / ry - 1, kclo = b[4] / rz - 1;
        int ichi = (b[1] + 1) / rx, jchi = (b[3] + 1) / ry, kchi = (b[5] + 1) / rz;
        int nrow = ichi - iclo + 1, ncol = jchi - jclo + 1, nfil = kchi - kclo + 1;
        double xl = (b[0] * h->hx[level] + h->xlower) - hxc;
        double yf = (b[2] * h->hy[level] + h->ylower) - hyc;
        double zb = (b[4] * h->hz[level] + h->zlower) - hzc;
        double xr = ((b[1] + 1) * h->hx[level] + h->xlower) + hxc;
        double yr = ((b[3] + 1) * h->hy[level] + h->ylower) + hyc;
        double zt = ((b[5] + 1) * h->hz[level] + h->zlower) + hzc;
        size_t nval = (size_t)nrow * ncol * nfil * nv;
        double* val = malloc(nval * sizeof(double));
        if (!val) return ORC_E_ALLOC;
        for (size_t t = 0; t < nval; t++) val[t] = NAN;   /* igetsp: unset */
        int sticks = iclo < 0 || ichi == h->iregsz[levc] || jclo < 0 || jchi == h->jregsz[levc] ||
                     kclo < 0 || kchi == h->kregsz[levc];
        if (sticks && perdom3)
            preicall3(p, h, val, nrow, ncol, iclo, ichi, jclo, jchi, kclo, kchi, levc);
        else
            icall3(p, h, val, nrow, ncol, iclo, ichi, jclo, jchi, kclo, kchi, levc, 1, 1, 1);
        int rc = orc_bc3amr(p, h, val, nrow, ncol, nfil, hxc, hyc, hzc, xl, xr, yf, yr, zb, zt);
        if (rc) { free(val); return rc; }
        double* qc1d = r->reg[level] + r->roff[level][g] + (size_t)nv * lenbc;
        int index = 0;
#define QCS(i, j, k) do { for (int m = 1; m <= nv; m++) qc1d[(size_t)nv * index + m - 1] = \
                              V3(val, nv, nrow, ncol, m, i, j, k); index++; } while (0)
        for (int k = 2; k <= nfil - 1; k++) for (int j = 2; j <= ncol - 1; j++) QCS(1, j, k);
        for (int k = 2; k <= nfil - 1; k++) for (int i = 2; i <= nrow - 1; i++) QCS(i, ncol, k);
        for (int k = 2; k <= nfil - 1; k++) for (int j = 2; j <= ncol - 1; j++) QCS(nrow, j, k);
        for (int k = 2; k <= nfil - 1; k++) for (int i = 2; i <= nrow - 1; i++) QCS(i, 1, k);
        for (int j = 2; j <= ncol - 1; j++) for (int i = 2; i <= nrow - 1; i++) QCS(i, j, 1);
        for (int j = 2; j <= ncol - 1; j++) for (int i = 2; i <= nrow - 1; i++) QCS(i, j, nfil);
#undef QCS
        free(val);
    }
    return 0;
}

/* qad.f (3-D): for each side and each fine plane along the side, a line of
 * Riemann problems between the coarse values (qc1d, repeated r times) and
 * the fine grid's first ghost cells; the fluctuations (amdq + apdq) times
 * face area and dt are added to the coarse segment's register slot */
int orc_qad3(const orc_params* p, const orc_hier3* h, orc_regs3* r, int level, int grid) {
    const int nv = p->meqn, ng = p->mbc, mw = p->mwaves;
    const int* b = h->boxes[level] + 6 * grid;
    const int nr = b[1] - b[0] + 1, nc = b[3] - b[2] + 1, nf = b[5] - b[4] + 1;
    const int mitot = nr + 2 * ng, mjtot = nc + 2 * ng, mktot = nf + 2 * ng;
    const int lrx = h->intratx[level - 1], lry = h->intraty[level - 1], lrz = h->intratz[level - 1];
    const double hx = h->hx[level], hy = h->hy[level], hz = h->hz[level], delt = h->possk[level];
    const int lenbc = r->lenbc[level][grid];
    double* svdflx = r->reg[level] + r->roff[level][grid];
    const double* qc1d = svdflx + (size_t)nv * lenbc;
    const double* vb = h->qnew[level] + h->qoff[level][grid];
    const int nmax = imax(nr, imax(nc, nf));
    const int npos = nmax + 2;
    double* ql = calloc((size_t)nv * npos, sizeof(double));
    double* qr = calloc((size_t)nv * npos, sizeof(double));
    double* wave = calloc((size_t)nv * mw * npos, sizeof(double));
    double* s = calloc((size_t)mw * npos, sizeof(double));
    double* amdq = calloc((size_t)nv * npos, sizeof(double));
    double* apdq = calloc((size_t)nv * npos, sizeof(double));
    if (!ql || !qr || !wave || !s || !amdq || !apdq) return ORC_E_ALLOC;
#define P1(a, m, pos) (a)[((m) - 1) + (size_t)nv * ((pos) - 1)]
#define QB(m, i, j, k) V3(vb, nv, mitot, mjtot, m, i, j, k)
#define QC(m, ix) qc1d[((m) - 1) + (size_t)nv * ((ix) - 1)]
#define SV(m, ix) svdflx[((m) - 1) + (size_t)nv * ((ix) - 1)]
    const int ncrsei = nr / lrx, ncrsej = nc / lry;
    int index0 = 0, influx0 = 0;
    /* rpn3 sees positions p as Fortran index p - ng (arrays from 1 - mbc);
     * with mx = n + 1 - 2 ng it solves interfaces at positions 2..n+1 */
    /* sides 1-4 loop over fine planes k; sides 5-6 over fine rows j */
    for (int side = 1; side <= 6; side++) {
        const int zside = side >= 5;
        const int nout = zside ? nc : nf;               /* planes (rows) */
        const int rout = zside ? lry : lrz;
        const int xline = side == 1 || side == 3;       /* line along j */
        const int n = xline ? nc : nr;
        const int ncrse = xline ? ncrsej : ncrsei;
        const int rl = xline ? lry : lrx;
        const int ixyz = side <= 4 ? (xline ? 1 : 2) : 3;
        /* face area factors in the reference's order: hy hz, hx hz, hx hy */
        const double a1 = xline ? hy : hx, a2 = xline ? hz : (zside ? hy : hz);
        const int coarse_left = side == 1 || side == 4 || side == 5;
        for (int o = ng + 1; o <= ng + nout; o++) {
            int oc = 1 + (o - (ng + 1)) / rout;
            int ocnext = 1 + (o + 1 - (ng + 1)) / rout;
            /* fine ghost cells of this line */
            for (int t = ng + 1; t <= ng + n; t++) {
                int i, j, k;
                switch (side) {
                case 1: i = ng; j = t; k = o; break;
                case 2: i = t; j = mjtot - ng + 1; k = o; break;
                case 3: i = mitot - ng + 1; j = t; k = o; break;
                case 4: i = t; j = ng; k = o; break;
                case 5: i = t; j = o; k = ng; break;
                default: i = t; j = o; k = mktot - ng + 1; break;
                }
                int pos = t - ng + (coarse_left ? 1 : 0);
                for (int m = 1; m <= nv; m++) {
                    if (coarse_left) P1(ql, m, pos) = QB(m, i, j, k);
                    else P1(qr, m, pos) = QB(m, i, j, k);
                }
            }
            int lind = 0, index = index0;
            for (int c = 1; c <= ncrse; c++) {
                index++;
                for (int l = 1; l <= rl; l++) {
                    lind++;
                    for (int m = 1; m <= nv; m++) {
                        if (coarse_left) P1(qr, m, lind) = QC(m, index);
                        else P1(ql, m, lind + 1) = QC(m, index);
                    }
                }
            }
            if (ocnext != oc) index0 = index;
            orc_rpn3(p, ixyz, npos - 2 * ng, nv, mw, ng, n + 1 - 2 * ng, ql, qr, wave, s, amdq, apdq);
            int influx = influx0;
            for (int c = 1; c <= n / rl; c++) {
                influx++;
                int fine = (c - 1) * rl;
                int indexa = fine + 1;
                for (int m = 1; m <= nv; m++)
                    for (int l = 1; l <= rl; l++) {
                        if (coarse_left)
                            SV(m, influx) = SV(m, influx) + P1(amdq, m, indexa + l) * a1 * a2 * delt
                                            + P1(apdq, m, indexa + l) * a1 * a2 * delt;
                        else
                            SV(m, influx) = SV(m, influx) - P1(amdq, m, indexa + l) * a1 * a2 * delt
                                            - P1(apdq, m, indexa + l) * a1 * a2 * delt;
                    }
            }
            if (ocnext != oc) influx0 = influx;
        }
    }
#undef P1
#undef QB
#undef QC
#undef SV
    free(ql); free(qr); free(wave); free(s); free(amdq); free(apdq);
    return 0;
}

/* fluxsv.f (3-D) for coarse grid `grid` of `level`: overwrite each listed
 * register slot with minus the coarse flux times dt and face area */
static void fluxsv3(const orc_params* p, const orc_hier3* h, orc_regs3* r, int level, int grid,
                    const double* fm, const double* fp, const double* gm, const double* gp,
                    const double* hm, const double* hp) {
    const int nv = p->meqn, ng = p->mbc;
    if (!r->listbc[level]) return;
    const int* b = h->boxes[level] + 6 * grid;
    const int mitot = b[1] - b[0] + 1 + 2 * ng, mjtot = b[3] - b[2] + 1 + 2 * ng;
    const double dtc = h->possk[level], hx = h->hx[level], hy = h->hy[level], hz = h->hz[level];
    for (int e = r->lst_start[level][grid]; e < r->lst_start[level][grid] + r->lst_count[level][grid]; e++) {
        const int* L = r->listbc[level] + 6 * (size_t)e;
        int i = L[0], j = L[1], k = L[2], side = L[3], kid = L[4], lkid = L[5];
        double* slot = r->reg[level + 1] + r->roff[level + 1][kid] + (size_t)nv * (lkid - 1);
#define F(a, m, i, j, k) V3(a, nv, mitot, mjtot, m, i, j, k)
        for (int m = 1; m <= nv; m++) {
            double v = 0.0;
            switch (side) {
            case 1: v = -F(fp, m, i, j, k) * dtc * hy * hz; break;
            case 2: v = -F(gm, m, i, j + 1, k) * dtc * hx * hz; break;
            case 3: v = -F(fm, m, i + 1, j, k) * dtc * hy * hz; break;
            case 4: v = -F(gp, m, i, j, k) * dtc * hx * hz; break;
            case 5: v = -F(hp, m, i, j, k) * dtc * hx * hy; break;
            default: v = -F(hm, m, i, j, k + 1) * dtc * hx * hy; break;
            }
            slot[m - 1] = v;
        }
#undef F
    }
}

/* fluxad.f (3-D) for fine grid `grid` of `level`: accumulate the fine
 * fluxes of the r x r faces of each coarse face, times dt and face area */
static void fluxad3(const orc_params* p, const orc_hier3* h, orc_regs3* r, int level, int grid,
                    const double* fm, const double* fp, const double* gm, const double* gp,
                    const double* hm, const double* hp) {
    const int nv = p->meqn, ng = p->mbc;
    const int* b = h->boxes[level] + 6 * grid;
    const int nx = b[1] - b[0] + 1, ny = b[3] - b[2] + 1, nz = b[5] - b[4] + 1;
    const int mitot = nx + 2 * ng, mjtot = ny + 2 * ng, mktot = nz + 2 * ng;
    const int lrx = h->intratx[level - 1], lry = h->intraty[level - 1], lrz = h->intratz[level - 1];
    const int nxc = nx / lrx, nyc = ny / lry, nzc = nz / lrz;
    const double dtf = h->possk[level], dx = h->hx[level], dy = h->hy[level], dz = h->hz[level];
    double* svdflx = r->reg[level] + r->roff[level][grid];
#define SV(m, ix) svdflx[((m) - 1) + (size_t)nv * ((ix) - 1)]
#define F(a, m, i, j, k) V3(a, nv, mitot, mjtot, m, i, j, k)
    int lind = 0;
    for (int k = 1; k <= nzc; k++) {
        int kfine = (k - 1) * lrz + ng;
        for (int j = 1; j <= nyc; j++) {
            int jfine = (j - 1) * lry + ng;
            lind++;
            for (int m = 1; m <= nv; m++)
                for (int kl = 1; kl <= lrz; kl++)
                    for (int jl = 1; jl <= lry; jl++)
                        SV(m, lind) = SV(m, lind) + F(fm, m, ng + 1, jfine + jl, kfine + kl) * dtf * dy * dz;
        }
    }
    for (int k = 1; k <= nzc; k++) {
        int kfine = (k - 1) * lrz + ng;
        for (int i = 1; i <= nxc; i++) {
            int ifine = (i - 1) * lrx + ng;
            lind++;
            for (int m = 1; m <= nv; m++)
                for (int kl = 1; kl <= lrz; kl++)
                    for (int il = 1; il <= lrx; il++)
                        SV(m, lind) = SV(m, lind) + F(gp, m, ifine + il, mjtot - ng + 1, kfine + kl) * dtf * dx * dz;
        }
    }
    for (int k = 1; k <= nzc; k++) {
        int kfine = (k - 1) * lrz + ng;
        for (int j = 1; j <= nyc; j++) {
            int jfine = (j - 1) * lry + ng;
            lind++;
            for (int m = 1; m <= nv; m++)
                for (int kl = 1; kl <= lrz; kl++)
                    for (int jl = 1; jl <= lry; jl++)
                        SV(m, lind) = SV(m, lind) + F(fp, m, mitot - ng + 1, jfine + jl, kfine + kl) * dtf * dy * dz;
        }
    }
    for (int k = 1; k <= nzc; k++) {
        int kfine = (k - 1) * lrz + ng;
        for (int i = 1; i <= nxc; i++) {
            int ifine = (i - 1) * lrx + ng;
            lind++;
            for (int m = 1; m <= nv; m++)
                for (int kl = 1; kl <= lrz; kl++)
                    for (int il = 1; il <= lrx; il++)
                        SV(m, lind) = SV(m, lind) + F(gm, m, ifine + il, ng + 1, kfine + kl) * dtf * dx * dz;
        }
    }
    for (int j = 1; j <= nyc; j++) {
        int jfine = (j - 1) * lry + ng;
        for (int i = 1; i <= nxc; i++) {
            int ifine = (i - 1) * lrx + ng;
            lind++;
            for (int m = 1; m <= nv; m++)
                for (int jl = 1; jl <= lry; jl++)
                    for (int il = 1; il <= lrx; il++)
                        SV(m, lind) = SV(m, lind) + F(hm, m, ifine + il, jfine + jl, ng + 1) * dtf * dx * dy;
        }
    }
    for (int j = 1; j <= nyc; j++) {
        int jfine = (j - 1) * lry + ng;
        for (int i = 1; i <= nxc; i++) {
            int ifine = (i - 1) * lrx + ng;
            lind++;
            for (int m = 1; m <= nv; m++)
                for (int jl = 1; jl <= lry; jl++)
                    for (int il = 1; il <= lrx; il++)
                        SV(m, lind) = SV(m, lind) + F(hp, m, ifine + il, jfine + jl, mktot - ng + 1) * dtf * dx * dy;
        }
    }
    (void)mktot;
#undef SV
#undef F
}

/* advanc.f (3-D): bound every grid, saveqc for the next level, then per
 * grid the new->old copy (level < mxnest), qad, stepgrid, fluxsv, fluxad */
int orc_advanc3(const orc_params* p, orc_hier3* h, orc_regs3* r, int level, double* dtlevnew,
                double* cfl_level) {
    const int nv = p->meqn, ng = p->mbc;
    int rc;
    for (int g = 0; g < h->ngrids[level]; g++) {
        rc = orc_bound3(p, h, level, g, h->time[level], h->qnew[level] + h->qoff[level][g]);
        if (rc) return rc;
    }
    int has_fine = level + 1 <= h->mxnest && level + 1 <= h->nlev && h->ngrids[level + 1] > 0;
    if (has_fine && (rc = orc_saveqc3(p, h, r, level + 1))) return rc;
    double dmin = ORC_RINFINITY, cmax = 0.0;
    for (int g = 0; g < h->ngrids[level]; g++) {
        const int* b = h->boxes[level] + 6 * g;
        int mx = b[1] - b[0] + 1, my = b[3] - b[2] + 1, mz = b[5] - b[4] + 1;
        size_t ntot = (size_t)nv * (mx + 2 * ng) * (my + 2 * ng) * (mz + 2 * ng);
        double* q = h->qnew[level] + h->qoff[level][g];
        if (level < h->mxnest) memcpy(h->qold[level] + h->qoff[level][g], q, ntot * sizeof(double));
        if (level > 1 && (rc = orc_qad3(p, h, r, level, g))) return rc;
        double* f = calloc(6 * ntot, sizeof(double));
        if (!f) return ORC_E_ALLOC;
        double cflgrid, dtnew;
        orc_stepgrid3(p, mx, my, mz, q, h->hx[level], h->hy[level], h->hz[level], h->possk[level],
                      f, f + ntot, f + 2 * ntot, f + 3 * ntot, f + 4 * ntot, f + 5 * ntot,
                      &cflgrid, &dtnew);
        if (has_fine)
            fluxsv3(p, h, r, level, g, f, f + ntot, f + 2 * ntot, f + 3 * ntot, f + 4 * ntot, f + 5 * ntot);
        if (level > 1)
            fluxad3(p, h, r, level, g, f, f + ntot, f + 2 * ntot, f + 3 * ntot, f + 4 * ntot, f + 5 * ntot);
        free(f);
        if (dtnew < dmin) dmin = dtnew;
        if (cflgrid > cmax) cmax = cflgrid;
    }
    h->time[level] = h->time[level] + h->possk[level];
    if (dtlevnew) *dtlevnew = dmin;
    if (cfl_level) *cfl_level = cmax;
    return 0;
}

/* update.f (3-D) with upbnd.f (3-D), no capacity */
int orc_update3(const orc_params* p, const orc_hier3* h, orc_regs3* r, int level) {
    const int nv = p->meqn, ng = p->mbc;
    const int rx = h->intratx[level], ry = h->intraty[level], rz = h->intratz[level];
    const double volume = h->hx[level] * h->hy[level] * h->hz[level];
    static const double chsign[6] = {1., -1., -1., 1., 1., -1.};
    for (int m0 = 0; m0 < h->ngrids[level]; m0++) {
        const int* cb = h->boxes[level] + 6 * m0;
        int ilo = cb[0], ihi = cb[1], jlo = cb[2], jhi = cb[3], klo = cb[4], khi = cb[5];
        int mitot = ihi - ilo + 1 + 2 * ng, mjtot = jhi - jlo + 1 + 2 * ng, mktot = khi - klo + 1 + 2 * ng;
        double* val = h->qnew[level] + h->qoff[level][m0];
        if (r->listbc[level] && r->lst_count[level][m0] > 0) {
            int* iused = calloc((size_t)mitot * mjtot * mktot, sizeof(int));
            if (!iused) return ORC_E_ALLOC;
            for (int e = r->lst_start[level][m0]; e < r->lst_start[level][m0] + r->lst_count[level][m0]; e++) {
                const int* L = r->listbc[level] + 6 * (size_t)e;
                int ic = L[0], jc = L[1], kc = L[2], iside = L[3], kid = L[4], lkid = L[5];
                int norm = 1 << (iside - 1);
                int* u = &iused[(ic - 1) + (size_t)mitot * ((jc - 1) + (size_t)mjtot * (kc - 1))];
                if ((*u / norm) % 2 == 1) continue;
                double sgnm = chsign[iside - 1];
                const double* svd = r->reg[level + 1] + r->roff[level + 1][kid] + (size_t)nv * (lkid - 1);
                for (int m = 1; m <= nv; m++)
                    V3(val, nv, mitot, mjtot, m, ic, jc, kc) =
                        V3(val, nv, mitot, mjtot, m, ic, jc, kc) + sgnm * svd[m - 1] / volume;
                *u += norm;
            }
            free(iused);
        }
        if (level + 1 > h->nlev) continue;
        for (int kid = 0; kid < h->ngrids[level + 1]; kid++) {
            const int* fb = h->boxes[level + 1] + 6 * kid;
            int iclo = fb[0] / rx, jclo = fb[2] / ry, kclo = fb[4] / rz;
            int ichi = fb[1] / rx, jchi = fb[3] / ry, kchi = fb[5] / rz;
            int mi = fb[1] - fb[0] + 1 + 2 * ng, mj = fb[3] - fb[2] + 1 + 2 * ng;
            const double* qf = h->qnew[level + 1] + h->qoff[level + 1][kid];
            int iplo = imax(ilo, iclo), jplo = imax(jlo, jclo), kplo = imax(klo, kclo);
            int iphi = imin(ihi, ichi), jphi = imin(jhi, jchi), kphi = imin(khi, kchi);
            if (iplo > iphi || jplo > jphi || kplo > kphi) continue;
            const double totrat = (double)(rx * ry * rz);
            int iff = iplo * rx - fb[0] + ng + 1;
            for (int i = iplo - ilo + ng + 1; i <= iphi - ilo + ng + 1; i++) {
                int jff = jplo * ry - fb[2] + ng + 1;
                for (int j = jplo - jlo + ng + 1; j <= jphi - jlo + ng + 1; j++) {
                    int kff = kplo * rz - fb[4] + ng + 1;
                    for (int k = kplo - klo + ng + 1; k <= kphi - klo + ng + 1; k++) {
                        for (int m = 1; m <= nv; m++) V3(val, nv, mitot, mjtot, m, i, j, k) = 0.0;
                        for (int kco = 1; kco <= rz; kco++)
                            for (int jco = 1; jco <= ry; jco++)
                                for (int ico = 1; ico <= rx; ico++)
                                    for (int m = 1; m <= nv; m++)
                                        V3(val, nv, mitot, mjtot, m, i, j, k) =
                                            V3(val, nv, mitot, mjtot, m, i, j, k) +
                                            V3(qf, nv, mi, mj, m, iff + ico - 1, jff + jco - 1, kff + kco - 1);
                        for (int m = 1; m <= nv; m++)
                            V3(val, nv, mitot, mjtot, m, i, j, k) = V3(val, nv, mitot, mjtot, m, i, j, k) / totrat;
                        kff += rz;
                    }
                    jff += ry;
                }
                iff += rx;
            }
        }
    }
    return 0;
}

/* conck.f (3-D): hx hy hz times the sum of q(1) over the level's interiors */
double orc_conck3(const orc_params* p, const orc_hier3* h, int level) {
    const int nv = p->meqn, ng = p->mbc;
    double totmass = 0.0;
    for (int g = 0; g < h->ngrids[level]; g++) {
        const int* b = h->boxes[level] + 6 * g;
        int mitot = b[1] - b[0] + 1 + 2 * ng, mjtot = b[3] - b[2] + 1 + 2 * ng;
        int mktot = b[5] - b[4] + 1 + 2 * ng;
        const double* q = h->qnew[level] + h->qoff[level][g];
        for (int k = ng + 1; k <= mktot - ng; k++)
            for (int j = ng + 1; j <= mjtot - ng; j++)
                for (int i = ng + 1; i <= mitot - ng; i++)
                    totmass = totmass + V3(q, nv, mitot, mjtot, 1, i, j, k);
    }
    return totmass * h->hx[level] * h->hy[level] * h->hz[level];
}

/* ===================================================== flagging ======== */

/* spest.f (3-D) for one grid: big = the grid's new values with ghost cells
 * filled at its time (bound), flags(mitot, mjtot, mktot) = goodpt (0), then
 * flag2refine.f's badpt (2) / goodpt on the interior; returns the number
 * of flagged cells in *nflag */
int orc_spest3(const orc_params* p, const orc_hier3* h, int level, int grid, double tolsp,
               double* big, double* flags, int* nflag) {
    const int nv = p->meqn, ng = p->mbc;
    const int* b = h->boxes[level] + 6 * grid;
    const int mx = b[1] - b[0] + 1, my = b[3] - b[2] + 1, mz = b[5] - b[4] + 1;
    const size_t n = (size_t)(mx + 2 * ng) * (my + 2 * ng) * (mz + 2 * ng);
    memcpy(big, h->qnew[level] + h->qoff[level][grid], n * nv * sizeof(double));
    int rc = orc_bound3(p, h, level, grid, h->time[level], big);
    if (rc) return rc;
    for (size_t c = 0; c < n; c++) flags[c] = 0.0;
    int k = orc_flag2refine3(mx, my, mz, ng, nv, tolsp, big, 0.0, 2.0, flags);
    if (nflag) *nflag = k;
    return 0;
}

/* errest.f (3-D) with coarsen.f, copysol.f and errf1.f for one grid:
 * the old values with 2 nghost ghost cells filled at t - possk, coarsened
 * by 2 and stepped by 2 possk on the coarsened grid; `big` (spest's array)
 * stepped by possk on the grid; their difference flags 2x2x2 blocks whose
 * Richardson estimate reaches tol.  rflags(mitot, mjtot, mktot) receives
 * errf1's rctflg component 1 (goodpt 0 / badpt 2), which the reference
 * leaves in the grid's store2 */
int orc_errest3(const orc_params* p, const orc_hier3* h, int level, int grid, double tol,
                int iorder, double* big, double* rflags) {
    const int nv = p->meqn, ng = p->mbc, ngbig = 2 * p->mbc;
    const int* b = h->boxes[level] + 6 * grid;
    const int nx = b[1] - b[0] + 1, ny = b[3] - b[2] + 1, nz = b[5] - b[4] + 1;
    if (nx % 2 || ny % 2 || nz % 2) return ORC_E_SCRATCH;
    const int mitot = nx + 2 * ng, mjtot = ny + 2 * ng, mktot = nz + 2 * ng;
    const int mi2 = nx / 2 + 2 * ng, mj2 = ny / 2 + 2 * ng, mk2 = nz / 2 + 2 * ng;
    const int midub = nx + 2 * ngbig, mjdub = ny + 2 * ngbig, mkdub = nz + 2 * ngbig;
    const double hx = h->hx[level], hy = h->hy[level], hz = h->hz[level];
    const double dt = h->possk[level], time = h->time[level], tpre = time - dt;
    double* dub = calloc((size_t)midub * mjdub * mkdub * nv, sizeof(double));
    double* crs = calloc((size_t)mi2 * mj2 * mk2 * nv, sizeof(double));
    const size_t nf = (size_t)mitot * mjtot * mktot * nv, nc = (size_t)mi2 * mj2 * mk2 * nv;
    double* f = calloc(6 * (nf > nc ? nf : nc), sizeof(double));
    if (!dub || !crs || !f) { free(dub); free(crs); free(f); return ORC_E_ALLOC; }
    /* copysol: old interior into the centre of the wider array */
    const double* old = h->qold[level] + h->qoff[level][grid];
    for (int k = ng + 1; k <= mktot - ng; k++)
        for (int j = ng + 1; j <= mjtot - ng; j++)
            for (int i = ng + 1; i <= mitot - ng; i++)
                for (int m = 1; m <= nv; m++)
                    V3(dub, nv, midub, mjdub, m, i - ng + ngbig, j - ng + ngbig, k - ng + ngbig) =
                        V3(old, nv, mitot, mjtot, m, i, j, k);
    int rc = orc_bound3_ng(p, h, level, grid, tpre, dub, ngbig);
    if (rc) { free(dub); free(crs); free(f); return rc; }
    /* coarsen.f */
    for (int m = 1; m <= nv; m++)
        for (int k = 1; k <= mk2; k++)
            for (int j = 1; j <= mj2; j++)
                for (int i = 1; i <= mi2; i++) {
                    int fi = 2 * (i - 1) + 1, fj = 2 * (j - 1) + 1, fk = 2 * (k - 1) + 1;
                    V3(crs, nv, mi2, mj2, m, i, j, k) =
                        (V3(dub, nv, midub, mjdub, m, fi, fj, fk) + V3(dub, nv, midub, mjdub, m, fi + 1, fj, fk) +
                         V3(dub, nv, midub, mjdub, m, fi, fj + 1, fk) + V3(dub, nv, midub, mjdub, m, fi + 1, fj + 1, fk) +
                         V3(dub, nv, midub, mjdub, m, fi, fj, fk + 1) + V3(dub, nv, midub, mjdub, m, fi + 1, fj, fk + 1) +
                         V3(dub, nv, midub, mjdub, m, fi, fj + 1, fk + 1) + V3(dub, nv, midub, mjdub, m, fi + 1, fj + 1, fk + 1)) / 8.0;
                }
    free(dub);
    double cfl, dtn;
    orc_stepgrid3(p, nx / 2, ny / 2, nz / 2, crs, 2.0 * hx, 2.0 * hy, 2.0 * hz, 2. * dt, f, f + nc,
                  f + 2 * nc, f + 3 * nc, f + 4 * nc, f + 5 * nc, &cfl, &dtn);
    memset(f, 0, 6 * (nf > nc ? nf : nc) * sizeof(double));
    orc_stepgrid3(p, nx, ny, nz, big, hx, hy, hz, dt, f, f + nf, f + 2 * nf, f + 3 * nf, f + 4 * nf,
                  f + 5 * nf, &cfl, &dtn);
    free(f);
    /* errf1.f (3-D) */
    const double order = (double)((1 << (iorder + 1)) - 2);
    for (size_t c = 0; c < (size_t)mitot * mjtot * mktot; c++) rflags[c] = 0.0;
#define RF(i, j, k) rflags[((i) - 1) + (size_t)mitot * (((j) - 1) + (size_t)mjtot * ((k) - 1))]
    int kfine = ng + 1;
    for (int k = ng + 1; k <= mk2 - ng; k++, kfine += 2) {
        int jfine = ng + 1;
        for (int j = ng + 1; j <= mj2 - ng; j++, jfine += 2) {
            int ifine = ng + 1;
            for (int i = ng + 1; i <= mi2 - ng; i++, ifine += 2) {
                double t1 = V3(big, nv, mitot, mjtot, 1, ifine, jfine, kfine);
                double t2 = V3(big, nv, mitot, mjtot, 1, ifine + 1, jfine, kfine);
                double t3 = V3(big, nv, mitot, mjtot, 1, ifine + 1, jfine + 1, kfine);
                double t4 = V3(big, nv, mitot, mjtot, 1, ifine, jfine + 1, kfine);
                double t5 = V3(big, nv, mitot, mjtot, 1, ifine, jfine, kfine + 1);
                double t6 = V3(big, nv, mitot, mjtot, 1, ifine + 1, jfine, kfine + 1);
                double t7 = V3(big, nv, mitot, mjtot, 1, ifine + 1, jfine + 1, kfine + 1);
                double t8 = V3(big, nv, mitot, mjtot, 1, ifine, jfine + 1, kfine + 1);
                double aval = (t1 + t2 + t3 + t4 + t5 + t6 + t7 + t8) / 8.0;
                double est = fabs((aval - V3(crs, nv, mi2, mj2, 1, i, j, k)) / order);
                if (est >= tol)
                    for (int dk = 0; dk < 2; dk++)
                        for (int dj = 0; dj < 2; dj++)
                            for (int di = 0; di < 2; di++) RF(ifine + di, jfine + dj, kfine + dk) = 2.0;
            }
        }
    }
#undef RF
    free(crs);
    return 0;
}
