/*
 * orc_amr.c -- oracle for ghost filling, flux registers, reflux, flagging
 * (TEST INFRASTRUCTURE ONLY).  Plain-C restatements of the reference's
 * Fortran in its evaluation order; arrays in the reference's AoS layout.
 *
 * Where the reference stops (prints and `stop`s) the oracle returns an
 * error code instead:
 *   ORC_E_LEVEL1  filrecur: cells of a level-1 patch left unset (filpatch.f90:168)
 *   ORC_E_SCRATCH filrecur: coarse scratch too small (filpatch.f90:220-227)
 *   ORC_E_TIME    intfil: time not available from a source grid (intfil.f90:197-233)
 *   ORC_E_BC      bc2amr: user (0) or unknown boundary condition (bc2amr.f90:122-154)
 *   ORC_E_SPHERE  spherical domains are not restated
 *
 * One guard departs from the literal Fortran: bc2amr on an array lying wholly
 * outside the domain on one side (a periodic piece of a ghost strip below or
 * above a non-periodic boundary) has no source column/row inside the array;
 * the reference reads past the array there (undefined) and bound's own
 * bc2amr overwrites those cells afterwards.  The oracle skips such a side.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

enum { ORC_E_LEVEL1 = 11, ORC_E_SCRATCH = 12, ORC_E_TIME = 13, ORC_E_BC = 14,
       ORC_E_SPHERE = 15, ORC_E_ALLOC = 16 };

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

/* val(m, i, j), 1-based i in [1, nrow], j in [1, ncol] */
#define VAL(a, nv, nrow, m, i, j) (a)[((m) - 1) + (size_t)(nv) * (((i) - 1) + (size_t)(nrow) * ((j) - 1))]

/* ------------------------------------------------------------- bc2amr --- */
int orc_bc2amr(const orc_params* p, const orc_hier* h, double* val, int nrow,
               int ncol, double hx, double hy, double xlo_patch,
               double xhi_patch, double ylo_patch, double yhi_patch) {
    const int nv = p->meqn;
    double hxmarg = hx * .01, hymarg = hy * .01;                 /* :104-105 */
    if (h->xperdom && (h->yperdom || h->spheredom)) return 0;    /* :109 */
    if (xlo_patch < h->xlower - hxmarg) {                         /* :118 */
        int nxl = (int)((h->xlower + hxmarg - xlo_patch) / hx);
        int mb = h->mthbc[0];
        if ((mb == 1 && nxl >= nrow) || (mb == 3 && 2 * nxl > nrow)) mb = 2;
        switch (mb) {
        case 1:
            for (int j = 1; j <= ncol; j++)
                for (int i = 1; i <= nxl; i++)
                    for (int m = 1; m <= nv; m++)
                        VAL(val, nv, nrow, m, i, j) = VAL(val, nv, nrow, m, nxl + 1, j);
            break;
        case 2: case 4: break;
        case 3:
            for (int j = 1; j <= ncol; j++)
                for (int i = 1; i <= nxl; i++)
                    for (int m = 1; m <= nv; m++)
                        VAL(val, nv, nrow, m, i, j) = VAL(val, nv, nrow, m, 2 * nxl + 1 - i, j);
            for (int j = 1; j <= ncol; j++)
                for (int i = 1; i <= nxl; i++)
                    VAL(val, nv, nrow, 2, i, j) = -VAL(val, nv, nrow, 2, i, j);
            break;
        default: return ORC_E_BC;
        }
    }
    if (xhi_patch > h->xupper + hxmarg) {                         /* :161 */
        int nxr = (int)((xhi_patch - h->xupper + hxmarg) / hx);
        int ibeg = imax(nrow - nxr + 1, 1);
        int mb = h->mthbc[1];
        if ((mb == 1 && ibeg < 2) || (mb == 3 && 2 * ibeg - 1 - nrow < 1)) mb = 2;
        switch (mb) {
        case 1:
            for (int i = ibeg; i <= nrow; i++)
                for (int j = 1; j <= ncol; j++)
                    for (int m = 1; m <= nv; m++)
                        VAL(val, nv, nrow, m, i, j) = VAL(val, nv, nrow, m, ibeg - 1, j);
            break;
        case 2: case 4: break;
        case 3:
            for (int i = ibeg; i <= nrow; i++)
                for (int j = 1; j <= ncol; j++)
                    for (int m = 1; m <= nv; m++)
                        VAL(val, nv, nrow, m, i, j) = VAL(val, nv, nrow, m, 2 * ibeg - 1 - i, j);
            for (int i = ibeg; i <= nrow; i++)
                for (int j = 1; j <= ncol; j++)
                    VAL(val, nv, nrow, 2, i, j) = -VAL(val, nv, nrow, 2, i, j);
            break;
        default: return ORC_E_BC;
        }
    }
    if (ylo_patch < h->ylower - hymarg) {                         /* :207 */
        int nyb = (int)((h->ylower + hymarg - ylo_patch) / hy);
        int mb = h->mthbc[2];
        if ((mb == 1 && nyb >= ncol) || (mb == 3 && 2 * nyb > ncol)) mb = 2;
        switch (mb) {
        case 1:
            for (int j = 1; j <= nyb; j++)
                for (int i = 1; i <= nrow; i++)
                    for (int m = 1; m <= nv; m++)
                        VAL(val, nv, nrow, m, i, j) = VAL(val, nv, nrow, m, i, nyb + 1);
            break;
        case 2: case 4: break;
        case 3:
            for (int j = 1; j <= nyb; j++)
                for (int i = 1; i <= nrow; i++)
                    for (int m = 1; m <= nv; m++)
                        VAL(val, nv, nrow, m, i, j) = VAL(val, nv, nrow, m, i, 2 * nyb + 1 - j);
            for (int j = 1; j <= nyb; j++)
                for (int i = 1; i <= nrow; i++)
                    VAL(val, nv, nrow, 3, i, j) = -VAL(val, nv, nrow, 3, i, j);
            break;
        default: return ORC_E_BC;
        }
    }
    if (yhi_patch > h->yupper + hymarg) {                         /* :253 */
        int nyt = (int)((yhi_patch - h->yupper + hymarg) / hy);
        int jbeg = imax(ncol - nyt + 1, 1);
        int mb = h->mthbc[3];
        if ((mb == 1 && jbeg < 2) || (mb == 3 && 2 * jbeg - 1 - ncol < 1)) mb = 2;
        switch (mb) {
        case 1:
            for (int j = jbeg; j <= ncol; j++)
                for (int i = 1; i <= nrow; i++)
                    for (int m = 1; m <= nv; m++)
                        VAL(val, nv, nrow, m, i, j) = VAL(val, nv, nrow, m, i, jbeg - 1);
            break;
        case 2: case 4: break;
        case 3:
            for (int j = jbeg; j <= ncol; j++)
                for (int i = 1; i <= nrow; i++)
                    for (int m = 1; m <= nv; m++)
                        VAL(val, nv, nrow, m, i, j) = VAL(val, nv, nrow, m, i, 2 * jbeg - 1 - j);
            for (int j = jbeg; j <= ncol; j++)
                for (int i = 1; i <= nrow; i++)
                    VAL(val, nv, nrow, 3, i, j) = -VAL(val, nv, nrow, 3, i, j);
            break;
        default: return ORC_E_BC;
        }
    }
    return 0;
}

/* ------------------------------------------------------------- trimbd --- */
/* trimbd.f90:32-97; used(nrow, ncol) */
static void trimbd(const signed char* used, int nrow, int ncol, int* set, int r[4]) {
#define U(i, j) used[((i) - 1) + (size_t)nrow * ((j) - 1)]
    long utot = 0;
    for (int j = 1; j <= ncol; j++)
        for (int i = 1; i <= nrow; i++) utot += U(i, j);
    if (utot == (long)nrow * ncol) { *set = 1; return; }
    *set = 0;
    int check;
    for (int i = 1; i <= nrow; i++) {
        check = 1;
        for (int j = 1; j <= ncol; j++) check = imin(check, U(i, j));
        r[0] = i;
        if (check == 0) break;
    }
    for (int i = 1; i <= nrow; i++) {
        check = 1;
        for (int j = 1; j <= ncol; j++) check = imin(check, U(nrow - i + 1, j));
        r[1] = nrow - i + 1;
        if (check == 0) break;
    }
    for (int j = 1; j <= ncol; j++) {
        check = 1;
        for (int i = 1; i <= nrow; i++) check = imin(check, U(i, j));
        r[2] = j;
        if (check == 0) break;
    }
    for (int j = 1; j <= ncol; j++) {
        check = 1;
        for (int i = 1; i <= nrow; i++) check = imin(check, U(i, ncol - j + 1));
        r[3] = ncol - j + 1;
        if (check == 0) break;
    }
#undef U
}

/* ------------------------------------------------------------- intfil --- */
/* intfil.f90:98-311.  flaguse(ilo:ihi, jlo:jhi) */
static int intfil(const orc_params* p, const orc_hier* h, double* val, int mi,
                  int mj, double time, signed char* flaguse, int nrowst,
                  int ncolst, int ilo, int ihi, int jlo, int jhi, int level) {
    const int nv = p->meqn, ng = p->mbc;
    const int fw = ihi - ilo + 1;
#define FU(i, j) flaguse[((i) - ilo) + (size_t)fw * ((j) - jlo)]
    const double t_epsilon = 1.0e-4;
    const double dt = h->possk[level];
    (void)mj;
    memset(flaguse, 0, (size_t)fw * (jhi - jlo + 1));
    for (int k = 0; k < h->ngrids[level]; k++) {
        const int* b = h->boxes[level] + 4 * k;
        int imlo = b[0], imhi = b[1], jmlo = b[2], jmhi = b[3];
        int mitot = imhi - imlo + 1 + 2 * ng;
        int ixlo = imax(imlo, ilo), ixhi = imin(imhi, ihi);
        int jxlo = imax(jmlo, jlo), jxhi = imin(jmhi, jhi);
        if (!(ixlo <= ixhi && jxlo <= jxhi)) continue;
        double alphac = (h->time[level] - time) / dt;             /* :194 */
        double alphai = 1.0 - alphac;
        if (alphai < -t_epsilon || alphai > 1.0 + t_epsilon) return ORC_E_TIME;
        const double* qn = h->qnew[level] + h->qoff[level][k];
        const double* qo = h->qold[level] + h->qoff[level][k];
        int interp = 0;
        const double* src = qn;
        if (fabs(alphai - 1.0) < t_epsilon) {
            src = qn;
        } else if (fabs(alphai) < t_epsilon) {
            src = qo;
            if (level == h->mxnest) return ORC_E_TIME;
        } else {
            interp = 1;
            if (level == h->mxnest) return ORC_E_TIME;
        }
        for (int m = 1; m <= nv; m++)
            for (int j = jxlo; j <= jxhi; j++)
                for (int i = ixlo; i <= ixhi; i++) {
                    size_t s = (size_t)(m - 1) + (size_t)nv * ((i - imlo + ng) + (size_t)mitot * (j - jmlo + ng));
                    double v = interp ? qn[s] * alphai + qo[s] * alphac : src[s];
                    VAL(val, nv, mi, m, i - ilo + nrowst, j - jlo + ncolst) = v;
                    FU(i, j) = 1;
                }
    }
    /* cells outside the domain are set by boundary conditions (:275-289);
     * the reference's ranges, clipped to the array */
    if (jhi >= h->jregsz[level])
        for (int j = imax(h->jregsz[level], jlo); j <= jhi; j++)
            for (int i = ilo; i <= ihi; i++) FU(i, j) = 1;
    if (jlo < 0)
        for (int j = jlo; j <= imin(imin(-1, ncolst + jhi - jlo), jhi); j++)
            for (int i = ilo; i <= ihi; i++) FU(i, j) = 1;
    if (ilo < 0)
        for (int j = jlo; j <= jhi; j++)
            for (int i = ilo; i <= imin(imin(-1, nrowst + ihi - ilo), ihi); i++) FU(i, j) = 1;
    if (ihi >= h->iregsz[level])
        for (int j = jlo; j <= jhi; j++)
            for (int i = imax(h->iregsz[level], ilo); i <= ihi; i++) FU(i, j) = 1;
#undef FU
    return 0;
}

static int prefilrecur(const orc_params* p, const orc_hier* h, int level,
                       double* valbig, int mitot, int mjtot, double time,
                       int ilo, int ihi, int jlo, int jhi, int iglo, int jglo);

/* ----------------------------------------------------------- filrecur --- */
/* filpatch.f90:71-355 */
static int filrecur(const orc_params* p, const orc_hier* h, int level,
                    double* valbig, int mitot, int mjtot, double t,
                    int nrowst, int ncolst, int ilo, int ihi, int jlo, int jhi,
                    int patchOnly) {
    const int nv = p->meqn, ng = p->mbc;
    int rc = 0;
    int mitot_patch = ihi - ilo + 1, mjtot_patch = jhi - jlo + 1;
    double dx_fine = h->hx[level], dy_fine = h->hy[level];
    double xlow_fine = h->xlower + ilo * dx_fine;
    double ylow_fine = h->ylower + jlo * dy_fine;
    signed char* flaguse = malloc((size_t)mitot_patch * mjtot_patch);
    if (!flaguse) return ORC_E_ALLOC;
    rc = intfil(p, h, valbig, mitot, mjtot, t, flaguse, nrowst, ncolst, ilo, ihi, jlo, jhi, level);
    if (rc) { free(flaguse); return rc; }
    int set, un[4] = {0, 0, 0, 0};
    trimbd(flaguse, mitot_patch, mjtot_patch, &set, un);
    if (!set) {
        if (level == 1) { free(flaguse); return ORC_E_LEVEL1; }
        double dx_coarse = h->hx[level - 1], dy_coarse = h->hy[level - 1];
        un[0] += ilo - 1; un[1] += ilo - 1; un[2] += jlo - 1; un[3] += jlo - 1;
        int rx = h->intratx[level - 1], ry = h->intraty[level - 1];
        int iplo = (un[0] - rx + ng * rx) / rx - ng;                 /* :205-210 */
        int iphi = (un[1] + rx) / rx;
        int jplo = (un[2] - ry + ng * ry) / ry - ng;
        int jphi = (un[3] + ry) / ry;
        double xlow_coarse = h->xlower + iplo * dx_coarse;
        double ylow_coarse = h->ylower + jplo * dy_coarse;
        int mitot_coarse = iphi - iplo + 1, mjtot_coarse = jphi - jplo + 1;
        if (mitot_coarse > ihi - ilo + 3 || mjtot_coarse > jhi - jlo + 3) {
            free(flaguse);
            return ORC_E_SCRATCH;
        }
        double* valcrse = calloc((size_t)(ihi - ilo + 3) * (jhi - jlo + 3) * nv, sizeof(double));
        if (!valcrse) { free(flaguse); return ORC_E_ALLOC; }
        int sticks = iplo < 0 || jplo < 0 || iphi >= h->iregsz[level - 1] || jphi >= h->jregsz[level - 1];
        if ((h->xperdom || h->yperdom || h->spheredom) && sticks)
            rc = prefilrecur(p, h, level - 1, valcrse, mitot_coarse, mjtot_coarse, t,
                             iplo, iphi, jplo, jphi, iplo, jplo);
        else
            rc = filrecur(p, h, level - 1, valcrse, mitot_coarse, mjtot_coarse, t, 1, 1,
                          iplo, iphi, jplo, jphi, 1);
        if (rc) { free(valcrse); free(flaguse); return rc; }
        double ratiox = (double)rx, ratioy = (double)ry;
#define CI(n, i, j) valcrse[((n) - 1) + (size_t)nv * ((i) - 1) + (size_t)nv * mitot_coarse * ((j) - 1)]
        for (int j_fine = 1; j_fine <= mjtot_patch; j_fine++) {
            int j_coarse = (int)floor((j_fine + jlo - 1) / ratioy) - jplo + 1;
            double ycent_coarse = ylow_coarse + (j_coarse - .5) * dy_coarse;
            double ycent_fine = h->ylower + ((j_fine - 1 + jlo) + .5) * dy_fine;
            double eta2 = (ycent_fine - ycent_coarse) / dy_coarse;
            for (int i_fine = 1; i_fine <= mitot_patch; i_fine++) {
                int i_coarse = (int)floor((i_fine + ilo - 1) / ratiox) - iplo + 1;
                double xcent_coarse = xlow_coarse + (i_coarse - .5) * dx_coarse;
                double xcent_fine = h->xlower + ((i_fine - 1 + ilo) + .5) * dx_fine;
                double eta1 = (xcent_fine - xcent_coarse) / dx_coarse;
                if (flaguse[(i_fine - 1) + (size_t)mitot_patch * (j_fine - 1)] != 0) continue;
                for (int n = 1; n <= nv; n++) {
                    double valp10 = CI(n, i_coarse + 1, j_coarse);
                    double valm10 = CI(n, i_coarse - 1, j_coarse);
                    double valc = CI(n, i_coarse, j_coarse);
                    double valp01 = CI(n, i_coarse, j_coarse + 1);
                    double valm01 = CI(n, i_coarse, j_coarse - 1);
                    double dupc = valp10 - valc, dumc = valc - valm10, ducc = valp10 - valm10;
                    double du = fmin(fabs(dupc), fabs(dumc));
                    du = fmin(2.0 * du, 0.5 * fabs(ducc));
                    double fu = fmax(0.0, copysign(1.0, dupc * dumc));
                    double uslope = du * copysign(1.0, ducc) * fu;
                    double dvpc = valp01 - valc, dvmc = valc - valm01, dvcc = valp01 - valm01;
                    double dv = fmin(fabs(dvpc), fabs(dvmc));
                    dv = fmin(2.0 * dv, 0.5 * fabs(dvcc));
                    double fv = fmax(0.0, copysign(1.0, dvpc * dvmc));
                    double vslope = dv * copysign(1.0, dvcc) * fv;
                    double valint = valc + eta1 * uslope + eta2 * vslope;
                    VAL(valbig, nv, mitot, n, i_fine + nrowst - 1, j_fine + ncolst - 1) = valint;
                }
            }
        }
#undef CI
        free(valcrse);
    }
    free(flaguse);
    if (patchOnly) {
        double xhi_fine = h->xlower + (ihi + 1) * dx_fine;
        double yhi_fine = h->ylower + (jhi + 1) * dy_fine;
        rc = orc_bc2amr(p, h, valbig, mitot, mjtot, dx_fine, dy_fine, xlow_fine, xhi_fine,
                        ylow_fine, yhi_fine);
    }
    return rc;
}

/* -------------------------------------------------------- prefilrecur --- */
/* prefilp.f90:18-214 (periodic wrap; spherical domains not restated).
 * The piece [i1..i2]x[j1..j2] of the patch is filled from the periodic
 * image through a temporary patch and copied to valbig, whose cell (1,1) is
 * global (iglo, jglo) (patchCopyOut, :218-245). */
static int prefilrecur(const orc_params* p, const orc_hier* h, int level,
                       double* valbig, int mitot, int mjtot, double time,
                       int ilo, int ihi, int jlo, int jhi, int iglo, int jglo) {
    const int nv = p->meqn;
    int ist[3], iend[3], ishift[3], jst[3], jend[3], jshift[3];
    if (h->spheredom) return ORC_E_SPHERE;
    (void)mjtot;
    if (h->xperdom) {
        ist[0] = ilo; ist[1] = 0; ist[2] = h->iregsz[level];
        iend[0] = -1; iend[1] = h->iregsz[level] - 1; iend[2] = ihi;
        ishift[0] = h->iregsz[level]; ishift[1] = 0; ishift[2] = -h->iregsz[level];
    } else {
        ist[0] = h->iregsz[level]; ist[1] = ilo; ist[2] = h->iregsz[level];
        iend[0] = -1; iend[1] = ihi; iend[2] = -1;
        ishift[0] = ishift[1] = ishift[2] = 0;
    }
    if (h->yperdom) {
        jst[0] = jlo; jst[1] = 0; jst[2] = h->jregsz[level];
        jend[0] = -1; jend[1] = h->jregsz[level] - 1; jend[2] = jhi;
        jshift[0] = h->jregsz[level]; jshift[1] = 0; jshift[2] = -h->jregsz[level];
    } else {
        jst[0] = h->jregsz[level]; jst[1] = jlo; jst[2] = h->jregsz[level];
        jend[0] = -1; jend[1] = jhi; jend[2] = -1;
        jshift[0] = jshift[1] = jshift[2] = 0;
    }
    for (int i = 0; i < 3; i++) {
        int i1 = imax(ilo, ist[i]), i2 = imin(ihi, iend[i]);
        if (i1 > i2) continue;
        for (int j = 0; j < 3; j++) {
            int j1 = imax(jlo, jst[j]), j2 = imin(jhi, jend[j]);
            if (j1 > j2) continue;
            int mi = i2 - i1 + 1, mj = j2 - j1 + 1;
            double* valPatch = calloc((size_t)mi * mj * nv, sizeof(double));
            if (!valPatch) return ORC_E_ALLOC;
            int rc = filrecur(p, h, level, valPatch, mi, mj, time, 1, 1, i1 + ishift[i],
                              i2 + ishift[i], j1 + jshift[j], j2 + jshift[j], 1);
            if (rc) { free(valPatch); return rc; }
            int ist0 = i1 - iglo + 1, jst0 = j1 - jglo + 1;
            for (int jj = 1; jj <= mj; jj++)
                for (int ii = 1; ii <= mi; ii++)
                    for (int m = 1; m <= nv; m++)
                        VAL(valbig, nv, mitot, m, ist0 + ii - 1, jst0 + jj - 1) =
                            VAL(valPatch, nv, mi, m, ii, jj);
            free(valPatch);
        }
    }
    return 0;
}

/* -------------------------------------------------------------- bound --- */
/* bound.f90:51-158 */
int orc_bound(const orc_params* p, const orc_hier* h, int level, int grid,
              double time, double* valbig) {
    const int ng = p->mbc;
    const int* b = h->boxes[level] + 4 * grid;
    int ilo = b[0], ihi = b[1], jlo = b[2], jhi = b[3];
    int mitot = ihi - ilo + 1 + 2 * ng, mjtot = jhi - jlo + 1 + 2 * ng;
    double hx = h->hx[level], hy = h->hy[level];
    double xleft = ilo * hx + h->xlower, xright = (ihi + 1) * hx + h->xlower;   /* grdfit2.f:98-101 */
    double ybot = jlo * hy + h->ylower, ytop = (jhi + 1) * hy + h->ylower;
    double xloWithGhost = xleft - ng * hx, xhiWithGhost = xright + ng * hx;
    double yloWithGhost = ybot - ng * hy, yhiWithGhost = ytop + ng * hy;
    int rc;
    double xl, xr, yb, yt;
    /* left */
    xl = xleft - ng * hx;
    if (xl < h->xlower && h->xperdom)
        rc = prefilrecur(p, h, level, valbig, mitot, mjtot, time, ilo - ng, ilo - 1, jlo, jhi,
                         ilo - ng, jlo - ng);
    else
        rc = filrecur(p, h, level, valbig, mitot, mjtot, time, 1, ng + 1, ilo - ng, ilo - 1,
                      jlo, jhi, 0);
    if (rc) return rc;
    /* right */
    xr = xright + ng * hx;
    if (xr > h->xupper && h->xperdom)
        rc = prefilrecur(p, h, level, valbig, mitot, mjtot, time, ihi + 1, ihi + ng, jlo, jhi,
                         ilo - ng, jlo - ng);
    else
        rc = filrecur(p, h, level, valbig, mitot, mjtot, time, mitot - ng + 1, ng + 1, ihi + 1,
                      ihi + ng, jlo, jhi, 0);
    if (rc) return rc;
    /* bottom */
    xl = xleft - ng * hx; xr = xright + ng * hx; yb = ybot - ng * hy;
    if ((yb < h->ylower && (h->yperdom || h->spheredom)) ||
        ((xl < h->xlower || xr > h->xupper) && h->xperdom))
        rc = prefilrecur(p, h, level, valbig, mitot, mjtot, time, ilo - ng, ihi + ng, jlo - ng,
                         jlo - 1, ilo - ng, jlo - ng);
    else
        rc = filrecur(p, h, level, valbig, mitot, mjtot, time, 1, 1, ilo - ng, ihi + ng,
                      jlo - ng, jlo - 1, 0);
    if (rc) return rc;
    /* top */
    yt = ytop + ng * hy;
    if ((yt > h->yupper && (h->yperdom || h->spheredom)) ||
        ((xl < h->xlower || xr > h->xupper) && h->xperdom))
        rc = prefilrecur(p, h, level, valbig, mitot, mjtot, time, ilo - ng, ihi + ng, jhi + 1,
                         jhi + ng, ilo - ng, jlo - ng);
    else
        rc = filrecur(p, h, level, valbig, mitot, mjtot, time, 1, mjtot - ng + 1, ilo - ng,
                      ihi + ng, jhi + 1, jhi + ng, 0);
    if (rc) return rc;
    (void)yb;
    return orc_bc2amr(p, h, valbig, mitot, mjtot, hx, hy, xloWithGhost, xhiWithGhost,
                      yloWithGhost, yhiWithGhost);
}
