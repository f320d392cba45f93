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
 *   ORC_E_BC      bc2amr: user (0) or unknown boundary condition (bc2amr.f90:122-154), or a
 *                 solid wall on a system without the velocity component it negates
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
    /* a solid wall negates q(2) on the x-sides and q(3) on the y-sides
     * (:145, :190, :236, :283): a system with fewer components has no wall
     * condition (the reference's loops would write past the patch) */
    if (((h->mthbc[0] == 3 || h->mthbc[1] == 3) && nv < 2) ||
        ((h->mthbc[2] == 3 || h->mthbc[3] == 3) && nv < 3))
        return ORC_E_BC;
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
                  int ncolst, int ilo, int ihi, int jlo, int jhi, int level, int msrc) {
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
        if (msrc >= 0) {
            /* only the bndList neighbours of grid msrc: the other grids that
             * meet msrc grown by nghost (init_bndryList.f:69-97;
             * intfil.f90:153-169) */
            const int* s0 = h->boxes[level] + 4 * msrc;
            if (k == msrc || imax(imlo, s0[0] - ng) > imin(imhi, s0[1] + ng) ||
                imax(jmlo, s0[2] - ng) > imin(jmhi, s0[3] + ng))
                continue;
        }
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
                    int patchOnly, int msrc) {
    const int nv = p->meqn, ng = p->mbc;
    int rc = 0;
    int mitot_patch = ihi - ilo + 1, mjtot_patch = jhi - jlo + 1;
    double dx_fine = h->hx[level], dy_fine = h->hy[level];
    double xlow_fine = h->xlower + ilo * dx_fine;
    double ylow_fine = h->ylower + jlo * dy_fine;
    signed char* flaguse = malloc((size_t)mitot_patch * mjtot_patch);
    if (!flaguse) return ORC_E_ALLOC;
    rc = intfil(p, h, valbig, mitot, mjtot, t, flaguse, nrowst, ncolst, ilo, ihi, jlo, jhi, level, msrc);
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
                          iplo, iphi, jplo, jphi, 1, -1);
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
                              i2 + ishift[i], j1 + jshift[j], j2 + jshift[j], 1, -1);
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
    return orc_bound_ng(p, h, level, grid, time, valbig, p->mbc);
}

/* bound.f90:51-158 with a ghost ring of width ng around the grid (the
 * storage keeps p->mbc = nghost; prepbigstep.f:53 calls bound with 2 nghost) */
int orc_bound_ng(const orc_params* p, const orc_hier* h, int level, int grid,
                 double time, double* valbig, int ng) {
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
                      jlo, jhi, 0, grid);
    if (rc) return rc;
    /* right */
    xr = xright + ng * hx;
    if (xr > h->xupper && h->xperdom)
        rc = prefilrecur(p, h, level, valbig, mitot, mjtot, time, ihi + 1, ihi + ng, jlo, jhi,
                         ilo - ng, jlo - ng);
    else
        rc = filrecur(p, h, level, valbig, mitot, mjtot, time, mitot - ng + 1, ng + 1, ihi + 1,
                      ihi + ng, jlo, jhi, 0, grid);
    if (rc) return rc;
    /* bottom */
    xl = xleft - ng * hx; xr = xright + ng * hx; yb = ybot - ng * hy;
    if ((yb < h->ylower && (h->yperdom || h->spheredom)) ||
        ((xl < h->xlower || xr > h->xupper) && h->xperdom))
        rc = prefilrecur(p, h, level, valbig, mitot, mjtot, time, ilo - ng, ihi + ng, jlo - ng,
                         jlo - 1, ilo - ng, jlo - ng);
    else
        rc = filrecur(p, h, level, valbig, mitot, mjtot, time, 1, 1, ilo - ng, ihi + ng,
                      jlo - ng, jlo - 1, 0, grid);
    if (rc) return rc;
    /* top */
    yt = ytop + ng * hy;
    if ((yt > h->yupper && (h->yperdom || h->spheredom)) ||
        ((xl < h->xlower || xr > h->xupper) && h->xperdom))
        rc = prefilrecur(p, h, level, valbig, mitot, mjtot, time, ilo - ng, ihi + ng, jhi + 1,
                         jhi + ng, ilo - ng, jlo - ng);
    else
        rc = filrecur(p, h, level, valbig, mitot, mjtot, time, 1, mjtot - ng + 1, ilo - ng,
                      ihi + ng, jhi + 1, jhi + ng, 0, grid);
    if (rc) return rc;
    (void)yb;
    return orc_bc2amr(p, h, valbig, mitot, mjtot, hx, hy, xloWithGhost, xhiWithGhost,
                      yloWithGhost, yhiWithGhost);
}

/* =================================================== flux registers ===== */

static int orc_list_push(int** a, int* n, int* cap, const int e[5]) {
    if (*n + 1 > *cap) {
        int nc = *cap ? 2 * *cap : 64;
        int* na = realloc(*a, (size_t)nc * 5 * sizeof(int));
        if (!na) return ORC_E_ALLOC;
        *a = na;
        *cap = nc;
    }
    memcpy(*a + 5 * (size_t)(*n), e, 5 * sizeof(int));
    (*n)++;
    return 0;
}

/* setuse.f:61-213, regular (kflag = 1) case */
static int setuse(int** lst, int* n, int* cap, int mkid, int ng, int ilo, int ihi,
                  int jlo, int jhi, int iclo, int ichi, int jclo, int jchi) {
    int ist = iclo - 1, iend = ichi + 1, jst = jclo - 1, jend = jchi + 1;
    int e[5], rc;
    if (!(ist < ilo)) {                                   /* :106-116 */
        int lkid = imax(jlo, jclo) - jclo + 1;
        for (int j = imax(jlo, jclo); j <= imin(jhi, jchi); j++) {
            e[0] = ist - ilo + ng + 1; e[1] = j - jlo + ng + 1; e[2] = 3; e[3] = mkid; e[4] = lkid++;
            if ((rc = orc_list_push(lst, n, cap, e))) return rc;
        }
    }
    if (!(jend > jhi)) {                                  /* :120-133 */
        int lkid = (jchi - jclo + 1) + imax(ilo, iclo) - iclo + 1;
        for (int i = imax(ilo, iclo); i <= imin(ihi, ichi); i++) {
            e[0] = i - ilo + ng + 1; e[1] = jend - jlo + ng + 1; e[2] = 4; e[3] = mkid; e[4] = lkid++;
            if ((rc = orc_list_push(lst, n, cap, e))) return rc;
        }
    }
    if (!(iend > ihi)) {                                  /* :159-170 */
        int lkid = (ichi - iclo + 1) + (jchi - jclo + 1) + imax(jlo, jclo) - jclo + 1;
        for (int j = imax(jlo, jclo); j <= imin(jhi, jchi); j++) {
            e[0] = iend - ilo + ng + 1; e[1] = j - jlo + ng + 1; e[2] = 1; e[3] = mkid; e[4] = lkid++;
            if ((rc = orc_list_push(lst, n, cap, e))) return rc;
        }
    }
    if (!(jst < jlo)) {                                   /* :175-186 */
        int lkid = 2 * (jchi - jclo + 1) + (ichi - iclo + 1) + imax(ilo, iclo) - iclo + 1;
        for (int i = imax(ilo, iclo); i <= imin(ihi, ichi); i++) {
            e[0] = i - ilo + ng + 1; e[1] = jst - jlo + ng + 1; e[2] = 2; e[3] = mkid; e[4] = lkid++;
            if ((rc = orc_list_push(lst, n, cap, e))) return rc;
        }
    }
    return 0;
}

void orc_free_regs(orc_regs* r) {
    for (int l = 0; l < ORC_MAXLV; l++) {
        free(r->lenbc[l]); free(r->roff[l]); free(r->reg[l]);
        free(r->lst_start[l]); free(r->lst_count[l]); free(r->listbc[l]);
    }
    memset(r, 0, sizeof *r);
}

/* prepf.f:18-53 (fine levels) and prepc.f:11-157 (coarse levels) */
int orc_prep_regs(const orc_params* p, const orc_hier* h, orc_regs* r) {
    const int nv = p->meqn, ng = p->mbc;
    memset(r, 0, sizeof *r);
    r->nlev = h->nlev;
    for (int l = 2; l <= h->nlev; l++) {
        int n = h->ngrids[l];
        r->lenbc[l] = calloc((size_t)n + 1, sizeof(int));
        r->roff[l] = calloc((size_t)n + 1, sizeof(int64_t));
        int64_t tot = 0;
        for (int k = 0; k < n; k++) {
            const int* b = h->boxes[l] + 4 * k;
            int nx = b[1] - b[0] + 1, ny = b[3] - b[2] + 1;
            r->lenbc[l][k] = 2 * (nx / h->intratx[l - 1] + ny / h->intraty[l - 1]);
            r->roff[l][k] = tot;
            tot += (2 * (int64_t)nv + p->maux) * r->lenbc[l][k];   /* prepf.f:33-48 */
        }
        r->nreg[l] = tot;
        r->reg[l] = calloc((size_t)tot + 1, sizeof(double));
        if (!r->reg[l]) return ORC_E_ALLOC;
    }
    for (int l = 1; l < h->nlev; l++) {
        if (h->ngrids[l + 1] == 0) continue;
        int nc = h->ngrids[l];
        int cap = 0, n = 0;
        int* lst = NULL;
        r->lst_start[l] = calloc((size_t)nc + 1, sizeof(int));
        r->lst_count[l] = calloc((size_t)nc + 1, sizeof(int));
        int imax_ = h->iregsz[l] - 1, jmax_ = h->jregsz[l] - 1;
        int rx = h->intratx[l], ry = h->intraty[l];
        for (int m = 0; m < nc; m++) {
            const int* cb = h->boxes[l] + 4 * m;
            int ilo = cb[0], ihi = cb[1], jlo = cb[2], jhi = cb[3];
            r->lst_start[l][m] = n;
            for (int kid = 0; kid < h->ngrids[l + 1]; kid++) {
                const int* fb = h->boxes[l + 1] + 4 * kid;
                int iclo = fb[0] / rx, jclo = fb[2] / ry, ichi = fb[1] / rx, jchi = fb[3] / ry;
                int iplo = imax(ilo, iclo), jplo = imax(jlo, jclo);
                int iphi = imin(ihi, ichi), jphi = imin(jhi, jchi);
                int rc = 0;
                if (iplo <= iphi + 1 && jplo <= jphi + 1)
                    rc = setuse(&lst, &n, &cap, kid, ng, ilo, ihi, jlo, jhi, iclo, ichi, jclo, jchi);
                if (!rc && h->xperdom && ilo == 0 && ichi == imax_)
                    rc = setuse(&lst, &n, &cap, kid, ng, ilo, ihi, jlo, jhi, iclo - h->iregsz[l],
                                ichi - h->iregsz[l], jclo, jchi);
                if (!rc && h->xperdom && iclo == 0 && ihi == imax_)
                    rc = setuse(&lst, &n, &cap, kid, ng, ilo, ihi, jlo, jhi, iclo + h->iregsz[l],
                                ichi + h->iregsz[l], jclo, jchi);
                if (!rc && h->yperdom && jlo == 0 && jchi == jmax_)
                    rc = setuse(&lst, &n, &cap, kid, ng, ilo, ihi, jlo, jhi, iclo, ichi,
                                jclo - h->jregsz[l], jchi - h->jregsz[l]);
                if (!rc && h->yperdom && jclo == 0 && jhi == jmax_)
                    rc = setuse(&lst, &n, &cap, kid, ng, ilo, ihi, jlo, jhi, iclo, ichi,
                                jclo + h->jregsz[l], jchi + h->jregsz[l]);
                if (rc) { free(lst); return rc; }
            }
            r->lst_count[l][m] = n - r->lst_start[l][m];
        }
        r->listbc[l] = lst;
        r->nlist[l] = n;
    }
    return 0;
}

/* icall.f (copy from the level's grids, new time) into val at (iputst,jputst),
 * and their aux values into auxv (maux per cell) when auxv and the level's
 * aux arrays are given */
static void icall_aux(const orc_params* p, const orc_hier* h, double* val, double* auxv,
                      int nrow, int ilo, int ihi, int jlo, int jhi, int level, int iputst,
                      int jputst) {
    const int nv = p->meqn, ng = p->mbc, na = p->maux;
    for (int k = 0; k < h->ngrids[level]; k++) {
        const int* b = h->boxes[level] + 4 * k;
        int ixlo = imax(b[0], ilo), ixhi = imin(b[1], ihi);
        int jxlo = imax(b[2], jlo), jxhi = imin(b[3], jhi);
        if (!(ixlo <= ixhi && jxlo <= jxhi)) continue;
        int mitot = b[1] - b[0] + 1 + 2 * ng;
        const double* q = h->qnew[level] + h->qoff[level][k];
        for (int j = jxlo; j <= jxhi; j++)
            for (int i = ixlo; i <= ixhi; i++)
                for (int m = 1; m <= nv; m++)
                    VAL(val, nv, nrow, m, i - ilo + iputst, j - jlo + jputst) =
                        VAL(q, nv, mitot, m, i - b[0] + ng + 1, j - b[2] + ng + 1);
        if (auxv && na > 0 && h->aux[level]) {
            const double* a = h->aux[level] + h->auxoff[level][k];
            for (int j = jxlo; j <= jxhi; j++)
                for (int i = ixlo; i <= ixhi; i++)
                    for (int m = 1; m <= na; m++)
                        VAL(auxv, na, nrow, m, i - ilo + iputst, j - jlo + jputst) =
                            VAL(a, na, mitot, m, i - b[0] + ng + 1, j - b[2] + ng + 1);
        }
    }
}

static void icall(const orc_params* p, const orc_hier* h, double* val, int nrow, int ilo,
                  int ihi, int jlo, int jhi, int level, int iputst, int jputst) {
    icall_aux(p, h, val, NULL, nrow, ilo, ihi, jlo, jhi, level, iputst, jputst);
}

/* preicall.f (periodic wrap of icall) */
static void preicall_aux(const orc_params* p, const orc_hier* h, double* val, double* auxv,
                         int nrow, int ilo, int ihi, int jlo, int jhi, int level) {
    const int ng = p->mbc;
    int ist[3], iend[3], ishift[3], jst[3], jend[3], jshift[3];
    if (h->xperdom) {
        ist[0] = ilo; ist[1] = 0; ist[2] = h->iregsz[level];
        iend[0] = -1; iend[1] = h->iregsz[level] - 1; iend[2] = ihi;
        ishift[0] = h->iregsz[level]; ishift[1] = 0; ishift[2] = -h->iregsz[level];
    } else {
        ist[0] = h->iregsz[level]; ist[1] = ilo; ist[2] = h->iregsz[level];
        iend[0] = -ng; iend[1] = ihi; iend[2] = -ng;
        ishift[0] = ishift[1] = ishift[2] = 0;
    }
    if (h->yperdom) {
        jst[0] = jlo; jst[1] = 0; jst[2] = h->jregsz[level];
        jend[0] = -1; jend[1] = h->jregsz[level] - 1; jend[2] = jhi;
        jshift[0] = h->jregsz[level]; jshift[1] = 0; jshift[2] = -h->jregsz[level];
    } else {
        jst[0] = h->jregsz[level]; jst[1] = jlo; jst[2] = h->jregsz[level];
        jend[0] = -ng; jend[1] = jhi; jend[2] = -ng;
        jshift[0] = jshift[1] = jshift[2] = 0;
    }
    for (int a = 0; a < 3; a++) {
        int i1 = imax(ilo, ist[a]), i2 = imin(ihi, iend[a]);
        if (i1 > i2) continue;
        for (int b = 0; b < 3; b++) {
            int j1 = imax(jlo, jst[b]), j2 = imin(jhi, jend[b]);
            if (j1 > j2) continue;
            icall_aux(p, h, val, auxv, nrow, i1 + ishift[a], i2 + ishift[a], j1 + jshift[b],
                      j2 + jshift[b], level, i1 - ilo + 1, j1 - jlo + 1);
        }
    }
}

static void preicall(const orc_params* p, const orc_hier* h, double* val, int nrow, int ilo,
                     int ihi, int jlo, int jhi, int level) {
    preicall_aux(p, h, val, NULL, nrow, ilo, ihi, jlo, jhi, level);
}

/* saveqc.f:13-109 and cstore.f:4-63 */
int orc_saveqc(const orc_params* p, const orc_hier* h, orc_regs* r, int level) {
    const int nv = p->meqn;
    const int levc = level - 1;
    const double hxc = h->hx[levc], hyc = h->hy[levc];
    const int rx = h->intratx[levc], ry = h->intraty[levc];
    for (int k = 0; k < h->ngrids[level]; k++) {
        const int* b = h->boxes[level] + 4 * k;
        int ilo = b[0], ihi = b[1], jlo = b[2], jhi = b[3];
        int nx = ihi - ilo + 1, ny = jhi - jlo + 1;
        int lenbc = 2 * (nx / rx + ny / ry);
        int iclo = ilo / rx - 1, jclo = jlo / ry - 1, ichi = (ihi + 1) / rx, jchi = (jhi + 1) / ry;
        int nrow = ichi - iclo + 1, ncol = jchi - jclo + 1;
        double xl = (ilo * h->hx[level] + h->xlower) - hxc;
        double yb = (jlo * h->hy[level] + h->ylower) - hyc;
        double xr = ((ihi + 1) * h->hx[level] + h->xlower) + hxc;
        double yt = ((jhi + 1) * h->hy[level] + h->ylower) + hyc;
        double* val = malloc((size_t)nrow * ncol * nv * sizeof(double));
        if (!val) return ORC_E_ALLOC;
        for (size_t t = 0; t < (size_t)nrow * ncol * nv; t++) val[t] = NAN;   /* igetsp: unset */
        /* the aux part of the temporary, NEEDS_TO_BE_SET until copied
         * (saveqc.f:62-64); where the patch sticks out of a non-periodic
         * domain the reference then calls the application's setaux
         * (saveqc.f:88-90), which the oracle does not restate: those ring
         * cells lie outside the domain and their register slots are never
         * read back (upbnd only corrects coarse cells of the domain) */
        const int na = p->maux;
        double* auxv = NULL;
        if (na > 0) {
            auxv = malloc((size_t)nrow * ncol * na * sizeof(double));
            if (!auxv) { free(val); return ORC_E_ALLOC; }
            for (size_t t = 0; t < (size_t)nrow * ncol * na; t++) auxv[t] = (double)10.e33f;
        }
        int sticks = iclo < 0 || ichi == h->iregsz[levc] || jclo < 0 || jchi == h->jregsz[levc];
        if (sticks && (h->xperdom || h->yperdom || h->spheredom))
            preicall_aux(p, h, val, auxv, nrow, iclo, ichi, jclo, jchi, levc);
        else
            icall_aux(p, h, val, auxv, nrow, iclo, ichi, jclo, jchi, levc, 1, 1);
        int rc = orc_bc2amr(p, h, val, nrow, ncol, hxc, hyc, xl, xr, yb, yt);
        if (rc) { free(val); free(auxv); return rc; }
        double* qc1d = r->reg[level] + r->roff[level][k] + (size_t)nv * lenbc;
        int index = 0;
#define QC(m, ix) qc1d[((m) - 1) + (size_t)nv * ((ix) - 1)]
        for (int j = 2; j <= ncol - 1; j++) { index++; for (int m = 1; m <= nv; m++) QC(m, index) = VAL(val, nv, nrow, m, 1, j); }
        for (int i = 2; i <= nrow - 1; i++) { index++; for (int m = 1; m <= nv; m++) QC(m, index) = VAL(val, nv, nrow, m, i, ncol); }
        for (int j = 2; j <= ncol - 1; j++) { index++; for (int m = 1; m <= nv; m++) QC(m, index) = VAL(val, nv, nrow, m, nrow, j); }
        for (int i = 2; i <= nrow - 1; i++) { index++; for (int m = 1; m <= nv; m++) QC(m, index) = VAL(val, nv, nrow, m, i, 1); }
#undef QC
        if (na > 0) {                                            /* cstore.f, auxc1d */
            double* auxc1d = r->reg[level] + r->roff[level][k] + 2 * (size_t)nv * lenbc;
            int ix = 0;
#define AC(m, i) auxc1d[((m) - 1) + (size_t)na * ((i) - 1)]
            for (int j = 2; j <= ncol - 1; j++) { ix++; for (int m = 1; m <= na; m++) AC(m, ix) = VAL(auxv, na, nrow, m, 1, j); }
            for (int i = 2; i <= nrow - 1; i++) { ix++; for (int m = 1; m <= na; m++) AC(m, ix) = VAL(auxv, na, nrow, m, i, ncol); }
            for (int j = 2; j <= ncol - 1; j++) { ix++; for (int m = 1; m <= na; m++) AC(m, ix) = VAL(auxv, na, nrow, m, nrow, j); }
            for (int i = 2; i <= nrow - 1; i++) { ix++; for (int m = 1; m <= na; m++) AC(m, ix) = VAL(auxv, na, nrow, m, i, 1); }
#undef AC
        }
        free(val);
        free(auxv);
    }
    return 0;
}

/* filval.f90:3-187 (naux = 0, mcapa = 0) with intcopy.f and preintcopy.f:
 * the values of a new grid of `level` whose box is (ilo, ihi, jlo, jhi).
 * The grids of `level` in h are the OLD ones (gfixup.f:60-98 fills the new
 * grids before it frees them); level-1 holds the coarse grids the new grid
 * nests in.  val is the new grid's AoS array (nvar, nx + 2ng, ny + 2ng).
 * Component 1 of every cell is first NEEDS_TO_BE_SET (filval.f90:140),
 * then the old grids are copied over the grid and its ghost ring, and the
 * interior cells still flagged get the min-mod interpolation of the coarse
 * patch grown by one coarse cell (intcopy of store1, bc2amr).  Cells of the
 * ghost ring no old grid covers keep the flag in component 1 and what val
 * held before in the others. */
int orc_filval(const orc_params* p, const orc_hier* h, int level, const int* box, double time,
               double* val) {
    (void)time;                                                   /* bc2amr's user BCs only */
    const int nv = p->meqn, ng = p->mbc;
    const double needs = (double)10.e33f;                         /* amr_module.f90:170 */
    const int levc = level - 1;
    const int rx = h->intratx[levc], ry = h->intraty[levc];       /* :44-45 */
    const double dx = h->hx[level], dy = h->hy[level];
    const double dxc = dx * rx, dyc = dy * ry;                    /* :46-47 */
    const int ilo = box[0], ihi = box[1], jlo = box[2], jhi = box[3];
    const int nx = ihi - ilo + 1, ny = jhi - jlo + 1;
    const int mitot = nx + 2 * ng, mjtot = ny + 2 * ng;
    const int mic = nx / rx + 2, mjc = ny / ry + 2;               /* gfixup.f:83-84 */
    const double xleft = ilo * dx + h->xlower, xright = (ihi + 1) * dx + h->xlower; /* grdfit2.f:98-101 */
    const double ybot = jlo * dy + h->ylower, ytop = (jhi + 1) * dy + h->ylower;
    const double xl = xleft - dxc, xr = xright + dxc, yb = ybot - dyc, yt = ytop + dyc;  /* :48-51 */
    const int iclo = ilo / rx - 1, jclo = jlo / ry - 1;           /* :59-62 */
    const int ichi = (ihi + 1) / rx - 1 + 1, jchi = (jhi + 1) / ry - 1 + 1;
    const int sxf = ilo < 0 || ihi >= h->iregsz[level], syf = jlo < 0 || jhi >= h->jregsz[level];
    const int sxc = iclo < 0 || ichi >= h->iregsz[levc], syc = jclo < 0 || jchi >= h->jregsz[levc];
    if (h->spheredom) return ORC_E_SPHERE;
    double* valc = malloc((size_t)nv * mic * mjc * sizeof(double));
    if (!valc) return ORC_E_ALLOC;
    for (size_t t = 0; t < (size_t)nv * mic * mjc; t++) valc[t] = NAN;   /* unset local */
    if ((h->xperdom && sxc) || (h->yperdom && syc))               /* :69-74 */
        preicall(p, h, valc, mic, iclo, ichi, jclo, jchi, levc);
    else
        icall(p, h, valc, mic, iclo, ichi, jclo, jchi, levc, 1, 1);
    int rc = orc_bc2amr(p, h, valc, mic, mjc, dxc, dyc, xl, xr, yb, yt);   /* :98-99 */
    if (rc) { free(valc); return rc; }
    for (int j = 1; j <= mjtot; j++)                              /* :140 */
        for (int i = 1; i <= mitot; i++) VAL(val, nv, mitot, 1, i, j) = needs;
    if ((h->xperdom && sxf) || (h->yperdom && syf))               /* :141-147 */
        preicall(p, h, val, mitot, ilo - ng, ihi + ng, jlo - ng, jhi + ng, level);
    else
        icall(p, h, val, mitot, ilo - ng, ihi + ng, jlo - ng, jhi + ng, level, 1, 1);
    unsigned char* unset = malloc((size_t)mitot * mjtot);         /* setflags, :148 */
    if (!unset) { free(valc); return ORC_E_ALLOC; }
    for (int j = 1; j <= mjtot; j++)
        for (int i = 1; i <= mitot; i++)
            unset[(size_t)(j - 1) * mitot + (i - 1)] = VAL(val, nv, mitot, 1, i, j) == needs;
    for (int j = 2; j <= mjc - 1; j++)                            /* :154-188 */
        for (int i = 2; i <= mic - 1; i++)
            for (int m = 1; m <= nv; m++) {
                const double vc = VAL(valc, nv, mic, m, i, j);
                double s1p = VAL(valc, nv, mic, m, i + 1, j) - vc;
                double s1m = vc - VAL(valc, nv, mic, m, i - 1, j);
                double slopex = fmin(fabs(s1p), fabs(s1m)) *
                                copysign(1.0, VAL(valc, nv, mic, m, i + 1, j) - VAL(valc, nv, mic, m, i - 1, j));
                if (s1m * s1p <= 0.0) slopex = 0.0;
                s1p = VAL(valc, nv, mic, m, i, j + 1) - vc;
                s1m = vc - VAL(valc, nv, mic, m, i, j - 1);
                double slopey = fmin(fabs(s1p), fabs(s1m)) *
                                copysign(1.0, VAL(valc, nv, mic, m, i, j + 1) - VAL(valc, nv, mic, m, i, j - 1));
                if (s1m * s1p <= 0.0) slopey = 0.0;
                for (int jco = 1; jco <= ry; jco++) {
                    const double yoff = ((double)jco - 0.5) / ry - 0.5;
                    const int jfine = (j - 2) * ry + ng + jco;
                    for (int ico = 1; ico <= rx; ico++) {
                        const double xoff = ((double)ico - 0.5) / rx - 0.5;
                        const int ifine = (i - 2) * rx + ng + ico;
                        if (unset[(size_t)(jfine - 1) * mitot + (ifine - 1)])
                            VAL(val, nv, mitot, m, ifine, jfine) = vc + xoff * slopex + yoff * slopey;
                    }
                }
            }
    free(unset);
    free(valc);
    return 0;
}

/* init_bndryList.f:50-107 (makeBndryList) for grid k of `level`: the other
 * grids of the level whose box meets grid k's box grown by nghost, in the
 * order its linked list holds them (each found grid is inserted at the
 * front, so the last one found comes first).  Writes the 0-based grid
 * indices into out (room for ngrids - 1) and returns their count. */
int orc_bndry_list(const orc_params* p, const orc_hier* h, int level, int k, int* out) {
    const int ng = p->mbc;
    const int* b = h->boxes[level] + 4 * k;
    const int ilo = b[0] - ng, ihi = b[1] + ng, jlo = b[2] - ng, jhi = b[3] + ng;
    int n = 0;
    for (int s = 0; s < h->ngrids[level]; s++) {
        if (s == k) continue;
        const int* c = h->boxes[level] + 4 * s;
        if (imax(c[0], ilo) <= imin(c[1], ihi) && imax(c[2], jlo) <= imin(c[3], jhi)) {
            for (int t = n; t > 0; t--) out[t] = out[t - 1];
            out[0] = s;
            n++;
        }
    }
    return n;
}

/* qad.f:9-360 (no aux, no source terms) */
int orc_qad(const orc_params* p, const orc_hier* h, orc_regs* r, int level, int grid) {
    const int nv = p->meqn, ng = p->mbc, mw = p->mwaves, na = p->maux;
    const int* b = h->boxes[level] + 4 * grid;
    const int nr = b[1] - b[0] + 1, nc = b[3] - b[2] + 1;
    const int mitot = nr + 2 * ng, mjtot = nc + 2 * ng;
    const int lrx = h->intratx[level - 1], lry = h->intraty[level - 1];
    const double hx = h->hx[level], hy = h->hy[level], delt = h->possk[level];
    const int lenbc = r->lenbc[level][grid];
    double* svdflx = r->reg[level] + r->roff[level][grid];
    const double* qc1d = svdflx + (size_t)nv * lenbc;
    const double* auxc1d = svdflx + 2 * (size_t)nv * lenbc;
    const double* valbig = h->qnew[level] + h->qoff[level][grid];
    const double* aux = (na > 0 && h->aux[level]) ? h->aux[level] + h->auxoff[level][grid] : NULL;
    const int npos = (nr > nc ? nr : nc) + 2;          /* positions 1..npos */
    const int nab = na > 0 ? na : 1;
    double* ql = calloc((size_t)nv * npos, sizeof(double));
    double* qr = calloc((size_t)nv * npos, sizeof(double));
    double* wave = calloc((size_t)nv * mw * npos, sizeof(double));
    double* s = calloc((size_t)mw * npos, sizeof(double));
    double* amdq = calloc((size_t)nv * npos, sizeof(double));
    double* apdq = calloc((size_t)nv * npos, sizeof(double));
    double* auxl = calloc((size_t)nab * npos, sizeof(double));
    double* auxr = calloc((size_t)nab * npos, sizeof(double));
    if (!ql || !qr || !wave || !s || !amdq || !apdq || !auxl || !auxr) return ORC_E_ALLOC;
#define P1(a, m, pos) (a)[((m) - 1) + (size_t)nv * ((pos) - 1)]
#define PA(a, m, pos) (a)[((m) - 1) + (size_t)na * ((pos) - 1)]
#define QB(m, i, j) VAL(valbig, nv, mitot, m, i, j)
#define AB(m, i, j) VAL(aux, na, mitot, m, i, j)
#define QC(m, ix) qc1d[((m) - 1) + (size_t)nv * ((ix) - 1)]
#define AC(m, ix) auxc1d[((m) - 1) + (size_t)na * ((ix) - 1)]
#define SV(m, ix) svdflx[((m) - 1) + (size_t)nv * ((ix) - 1)]
    /* rpn2 is called with mbc = nghost and mx = n+1-2*nghost, arrays indexed
     * from position 1 (= Fortran index 1-mbc): interfaces at positions 2..n+1;
     * aux as qad.f:81-91 (side 1), :160-185 (side 2), :228-255 (side 3),
     * :300-330 (side 4): the fine grid's face value for xleft / yleft, else
     * the cell's, and the coarse ring's auxc1d on the coarse side */
    int index = 0, influx = 0;
    /* side 1 */
    for (int j = ng + 1; j <= mjtot - ng; j++) {
        if (aux)
            for (int m = 1; m <= na; m++)
                PA(auxl, m, j - ng + 1) = p->auxtype[m - 1] == 1 ? AB(m, ng + 1, j) : AB(m, ng, j);
        for (int m = 1; m <= nv; m++) P1(ql, m, j - ng + 1) = QB(m, ng, j);
    }
    {
        int lind = 0;
        for (int jc = 1; jc <= nc / lry; jc++) {
            index++;
            for (int l = 1; l <= lry; l++) {
                lind++;
                if (aux) for (int m = 1; m <= na; m++) PA(auxr, m, lind) = AC(m, index);
                for (int m = 1; m <= nv; m++) P1(qr, m, lind) = QC(m, index);
            }
        }
    }
    orc_rpn2(p, 1, npos - 2 * ng, nv, mw, na, ng, nc + 1 - 2 * ng, ql, qr, auxl, auxr, wave, s, amdq, apdq);
    for (int j = 1; j <= nc / lry; j++) {
        influx++;
        int jfine = (j - 1) * lry;
        for (int m = 1; m <= nv; m++)
            for (int l = 1; l <= lry; l++)
                SV(m, influx) = SV(m, influx) + P1(amdq, m, jfine + l + 1) * hy * delt
                                + P1(apdq, m, jfine + l + 1) * hy * delt;
    }
    if (mjtot != 2 * ng + 1) {
        /* side 2 */
        for (int i = ng + 1; i <= mitot - ng; i++) {
            if (aux) for (int m = 1; m <= na; m++) PA(auxr, m, i - ng) = AB(m, i, mjtot - ng + 1);
            for (int m = 1; m <= nv; m++) P1(qr, m, i - ng) = QB(m, i, mjtot - ng + 1);
        }
        int lind = 0;
        for (int ic = 1; ic <= nr / lrx; ic++) {
            index++;
            for (int l = 1; l <= lrx; l++) {
                lind++;
                if (aux)
                    for (int m = 1; m <= na; m++)
                        PA(auxl, m, lind + 1) = p->auxtype[m - 1] == 2
                            ? AB(m, (ic - 1) * lrx + ng + l, mjtot - ng + 1) : AC(m, index);
                for (int m = 1; m <= nv; m++) P1(ql, m, lind + 1) = QC(m, index);
            }
        }
        orc_rpn2(p, 2, npos - 2 * ng, nv, mw, na, ng, nr + 1 - 2 * ng, ql, qr, auxl, auxr, wave, s, amdq, apdq);
        for (int i = 1; i <= nr / lrx; i++) {
            influx++;
            int ifine = (i - 1) * lrx;
            for (int m = 1; m <= nv; m++)
                for (int l = 1; l <= lrx; l++)
                    SV(m, influx) = SV(m, influx) - P1(amdq, m, ifine + l + 1) * hx * delt
                                    - P1(apdq, m, ifine + l + 1) * hx * delt;
        }
    }
    /* side 3 */
    for (int j = ng + 1; j <= mjtot - ng; j++) {
        if (aux) for (int m = 1; m <= na; m++) PA(auxr, m, j - ng) = AB(m, mitot - ng + 1, j);
        for (int m = 1; m <= nv; m++) P1(qr, m, j - ng) = QB(m, mitot - ng + 1, j);
    }
    {
        int lind = 0;
        for (int jc = 1; jc <= nc / lry; jc++) {
            index++;
            for (int l = 1; l <= lry; l++) {
                lind++;
                if (aux)
                    for (int m = 1; m <= na; m++)
                        PA(auxl, m, lind + 1) = p->auxtype[m - 1] == 1
                            ? AB(m, mitot - ng + 1, (jc - 1) * lry + ng + l) : AC(m, index);
                for (int m = 1; m <= nv; m++) P1(ql, m, lind + 1) = QC(m, index);
            }
        }
    }
    orc_rpn2(p, 1, npos - 2 * ng, nv, mw, na, ng, nc + 1 - 2 * ng, ql, qr, auxl, auxr, wave, s, amdq, apdq);
    for (int j = 1; j <= nc / lry; j++) {
        influx++;
        int jfine = (j - 1) * lry;
        for (int m = 1; m <= nv; m++)
            for (int l = 1; l <= lry; l++)
                SV(m, influx) = SV(m, influx) - P1(amdq, m, jfine + l + 1) * hy * delt
                                - P1(apdq, m, jfine + l + 1) * hy * delt;
    }
    if (mjtot != 2 * ng + 1) {
        /* side 4 */
        for (int i = ng + 1; i <= mitot - ng; i++) {
            if (aux)
                for (int m = 1; m <= na; m++)
                    PA(auxl, m, i - ng + 1) = p->auxtype[m - 1] == 2 ? AB(m, i, ng + 1) : AB(m, i, ng);
            for (int m = 1; m <= nv; m++) P1(ql, m, i - ng + 1) = QB(m, i, ng);
        }
        int lind = 0;
        for (int ic = 1; ic <= nr / lrx; ic++) {
            index++;
            for (int l = 1; l <= lrx; l++) {
                lind++;
                if (aux) for (int m = 1; m <= na; m++) PA(auxr, m, lind) = AC(m, index);
                for (int m = 1; m <= nv; m++) P1(qr, m, lind) = QC(m, index);
            }
        }
        orc_rpn2(p, 2, npos - 2 * ng, nv, mw, na, ng, nr + 1 - 2 * ng, ql, qr, auxl, auxr, wave, s, amdq, apdq);
        for (int i = 1; i <= nr / lrx; i++) {
            influx++;
            int ifine = (i - 1) * lrx;
            for (int m = 1; m <= nv; m++)
                for (int l = 1; l <= lrx; l++)
                    SV(m, influx) = SV(m, influx) + P1(amdq, m, ifine + l + 1) * hx * delt
                                    + P1(apdq, m, ifine + l + 1) * hx * delt;
        }
    }
#undef P1
#undef PA
#undef QB
#undef AB
#undef QC
#undef AC
#undef SV
    free(ql); free(qr); free(wave); free(s); free(amdq); free(apdq); free(auxl); free(auxr);
    return 0;
}

/* fluxsv.f:10-109 for coarse grid `grid` of `level` */
static void fluxsv(const orc_params* p, const orc_hier* h, orc_regs* r, int level, int grid,
                   const double* fm, const double* fp, const double* gm, const double* gp) {
    const int nv = p->meqn, ng = p->mbc;
    if (!r->listbc[level]) return;
    const int* b = h->boxes[level] + 4 * grid;
    const int mitot = b[1] - b[0] + 1 + 2 * ng;
    const double dtc = h->possk[level], hx = h->hx[level], hy = h->hy[level];
    for (int e = r->lst_start[level][grid]; e < r->lst_start[level][grid] + r->lst_count[level][grid]; e++) {
        const int* L = r->listbc[level] + 5 * (size_t)e;
        int i = L[0], j = L[1], side = L[2], kid = L[3], lkid = L[4];
        double* slot = r->reg[level + 1] + r->roff[level + 1][kid] + (size_t)nv * (lkid - 1);
        for (int m = 1; m <= nv; m++) {
            double v = 0.0;
            if (side == 1) v = -VAL(fp, nv, mitot, m, i, j) * dtc * hy;
            if (side == 2) v = -VAL(gm, nv, mitot, m, i, j + 1) * dtc * hx;
            if (side == 3) v = -VAL(fm, nv, mitot, m, i + 1, j) * dtc * hy;
            if (side == 4) v = -VAL(gp, nv, mitot, m, i, j) * dtc * hx;
            slot[m - 1] = v;
        }
    }
}

/* fluxad.f:7-96 for fine grid `grid` of `level` */
static void fluxad(const orc_params* p, const orc_hier* h, orc_regs* r, int level, int grid,
                   const double* fm, const double* fp, const double* gm, const double* gp) {
    const int nv = p->meqn, ng = p->mbc;
    const int* b = h->boxes[level] + 4 * grid;
    const int nx = b[1] - b[0] + 1, ny = b[3] - b[2] + 1;
    const int mitot = nx + 2 * ng, mjtot = ny + 2 * ng;
    const int lrx = h->intratx[level - 1], lry = h->intraty[level - 1];
    const int nyc = ny / lry, nxc = nx / lrx;
    const double dtf = h->possk[level], dx = h->hx[level], dy = h->hy[level];
    double* svdflx = r->reg[level] + r->roff[level][grid];
#define SV(m, ix) svdflx[((m) - 1) + (size_t)nv * ((ix) - 1)]
    int lind = 0;
    for (int j = 1; j <= nyc; j++) {
        lind++;
        int jfine = (j - 1) * lry + ng;
        for (int m = 1; m <= nv; m++)
            for (int l = 1; l <= lry; l++)
                SV(m, lind) = SV(m, lind) + VAL(fm, nv, mitot, m, ng + 1, jfine + l) * dtf * dy;
    }
    for (int i = 1; i <= nxc; i++) {
        lind++;
        int ifine = (i - 1) * lrx + ng;
        for (int m = 1; m <= nv; m++)
            for (int l = 1; l <= lrx; l++)
                SV(m, lind) = SV(m, lind) + VAL(gp, nv, mitot, m, ifine + l, mjtot - ng + 1) * dtf * dx;
    }
    for (int j = 1; j <= nyc; j++) {
        lind++;
        int jfine = (j - 1) * lry + ng;
        for (int m = 1; m <= nv; m++)
            for (int l = 1; l <= lry; l++)
                SV(m, lind) = SV(m, lind) + VAL(fp, nv, mitot, m, mitot - ng + 1, jfine + l) * dtf * dy;
    }
    for (int i = 1; i <= nxc; i++) {
        lind++;
        int ifine = (i - 1) * lrx + ng;
        for (int m = 1; m <= nv; m++)
            for (int l = 1; l <= lrx; l++)
                SV(m, lind) = SV(m, lind) + VAL(gm, nv, mitot, m, ifine + l, ng + 1) * dtf * dx;
    }
#undef SV
}

int orc_advanc(const orc_params* p, orc_hier* h, orc_regs* r, int level, double* dtlevnew,
               double* cfl_level) {
    const int nv = p->meqn, ng = p->mbc;
    int rc;
    for (int k = 0; k < h->ngrids[level]; k++) {                    /* advanc.f:60-84 */
        rc = orc_bound(p, h, level, k, h->time[level], h->qnew[level] + h->qoff[level][k]);
        if (rc) return rc;
    }
    int has_fine = level + 1 <= h->mxnest && level + 1 <= h->nlev && h->ngrids[level + 1] > 0;
    if (has_fine && (rc = orc_saveqc(p, h, r, level + 1))) return rc; /* :92-96 */
    double dmin = ORC_RINFINITY, cmax = 0.0;
    for (int k = 0; k < h->ngrids[level]; k++) {                    /* :106-129 */
        const int* b = h->boxes[level] + 4 * k;
        int mx = b[1] - b[0] + 1, my = b[3] - b[2] + 1;
        size_t ntot = (size_t)nv * (mx + 2 * ng) * (my + 2 * ng);
        double* q = h->qnew[level] + h->qoff[level][k];
        if (level < h->mxnest) memcpy(h->qold[level] + h->qoff[level][k], q, ntot * sizeof(double));
        if (level > 1 && (rc = orc_qad(p, h, r, level, k))) return rc;
        double* f = malloc(4 * ntot * sizeof(double));
        if (!f) return ORC_E_ALLOC;
        double cflgrid, dtnew;
        const double* aux = (p->maux > 0 && h->aux[level]) ? h->aux[level] + h->auxoff[level][k] : NULL;
        orc_stepgrid(p, mx, my, q, aux, h->hx[level], h->hy[level], h->possk[level], f, f + ntot,
                     f + 2 * ntot, f + 3 * ntot, &cflgrid, &dtnew);
        if (has_fine) fluxsv(p, h, r, level, k, f, f + ntot, f + 2 * ntot, f + 3 * ntot);
        if (level > 1) fluxad(p, h, r, level, k, f, f + ntot, f + 2 * ntot, f + 3 * ntot);
        free(f);
        if (dtnew < dmin) dmin = dtnew;
        if (cflgrid > cmax) cmax = cflgrid;
    }
    h->time[level] = h->time[level] + h->possk[level];              /* :289 */
    if (dtlevnew) *dtlevnew = dmin;
    if (cfl_level) *cfl_level = cmax;
    return 0;
}

/* update.f:16-189 with upbnd.f:18-127 (no capacity) */
int orc_update(const orc_params* p, const orc_hier* h, orc_regs* r, int level) {
    const int nv = p->meqn, ng = p->mbc;
    const int rx = h->intratx[level], ry = h->intraty[level];
    const double area = h->hx[level] * h->hy[level];
    for (int m0 = 0; m0 < h->ngrids[level]; m0++) {
        const int* cb = h->boxes[level] + 4 * m0;
        int ilo = cb[0], ihi = cb[1], jlo = cb[2], jhi = cb[3];
        int mitot = ihi - ilo + 1 + 2 * ng, mjtot = jhi - jlo + 1 + 2 * ng;
        double* val = h->qnew[level] + h->qoff[level][m0];
        const int capa = p->mcapa > 0 && h->aux[level];
        const double* caux = capa ? h->aux[level] + h->auxoff[level][m0] : NULL;
        if (r->listbc[level]) {                                    /* upbnd */
            int* iused = calloc((size_t)mitot * mjtot, sizeof(int));
            if (!iused) return ORC_E_ALLOC;
            for (int e = r->lst_start[level][m0]; e < r->lst_start[level][m0] + r->lst_count[level][m0]; e++) {
                const int* L = r->listbc[level] + 5 * (size_t)e;
                int icrse = L[0], jcrse = L[1], iside = L[2], kid = L[3], lkid = L[4];
                int norm = 1 << (iside - 1);
                int* u = &iused[(icrse - 1) + (size_t)mitot * (jcrse - 1)];
                if ((*u / norm) % 2 == 1) continue;                 /* :79-83 */
                double sgnm = (iside == 2 || iside == 3 || iside == 6) ? -1.0 : 1.0;
                if (capa) sgnm = sgnm / VAL(caux, p->maux, mitot, p->mcapa, icrse, jcrse);   /* :104-108 */
                const double* svd = r->reg[level + 1] + r->roff[level + 1][kid] + (size_t)nv * (lkid - 1);
                for (int m = 1; m <= nv; m++)
                    VAL(val, nv, mitot, m, icrse, jcrse) = VAL(val, nv, mitot, m, icrse, jcrse) + sgnm * svd[m - 1] / area;
                *u += norm;
            }
            free(iused);
        }
        if (level + 1 > h->nlev) continue;
        for (int kid = 0; kid < h->ngrids[level + 1]; kid++) {     /* :95-177 */
            const int* fb = h->boxes[level + 1] + 4 * kid;
            int iclo = fb[0] / rx, jclo = fb[2] / ry, ichi = fb[1] / rx, jchi = fb[3] / ry;
            int mi = fb[1] - fb[0] + 1 + 2 * ng;
            const double* qf = h->qnew[level + 1] + h->qoff[level + 1][kid];
            const double* faux = capa && h->aux[level + 1] ? h->aux[level + 1] + h->auxoff[level + 1][kid] : NULL;
            int iplo = imax(ilo, iclo), jplo = imax(jlo, jclo), iphi = imin(ihi, ichi), jphi = imin(jhi, jchi);
            if (iplo > iphi || jplo > jphi) continue;
            double totrat = (double)(rx * ry);
            for (int i = iplo - ilo + ng + 1; i <= iphi - ilo + ng + 1; i++) {
                int iff = iplo * rx - fb[0] + ng + 1 + (i - (iplo - ilo + ng + 1)) * rx;
                for (int j = jplo - jlo + ng + 1; j <= jphi - jlo + ng + 1; j++) {
                    int jff = jplo * ry - fb[2] + ng + 1 + (j - (jplo - jlo + ng + 1)) * ry;
                    for (int m = 1; m <= nv; m++) VAL(val, nv, mitot, m, i, j) = 0.0;
                    if (!faux) {
                        for (int jco = 1; jco <= ry; jco++)
                            for (int ico = 1; ico <= rx; ico++)
                                for (int m = 1; m <= nv; m++)
                                    VAL(val, nv, mitot, m, i, j) = VAL(val, nv, mitot, m, i, j)
                                        + VAL(qf, nv, mi, m, iff + ico - 1, jff + jco - 1);
                        for (int m = 1; m <= nv; m++)
                            VAL(val, nv, mitot, m, i, j) = VAL(val, nv, mitot, m, i, j) / totrat;
                    } else {                                       /* update.f:147-161 */
                        for (int jco = 1; jco <= ry; jco++)
                            for (int ico = 1; ico <= rx; ico++) {
                                const double cf = VAL(faux, p->maux, mi, p->mcapa, iff + ico - 1, jff + jco - 1);
                                for (int m = 1; m <= nv; m++)
                                    VAL(val, nv, mitot, m, i, j) = VAL(val, nv, mitot, m, i, j)
                                        + VAL(qf, nv, mi, m, iff + ico - 1, jff + jco - 1) * cf;
                            }
                        for (int m = 1; m <= nv; m++)
                            VAL(val, nv, mitot, m, i, j) = VAL(val, nv, mitot, m, i, j) /
                                (totrat * VAL(caux, p->maux, mitot, p->mcapa, i, j));
                    }
                }
            }
        }
    }
    return 0;
}

double orc_conck(const orc_params* p, const orc_hier* h, int level) {
    const int nv = p->meqn, ng = p->mbc;
    double totmass = 0.0;
    for (int k = 0; k < h->ngrids[level]; k++) {
        const int* b = h->boxes[level] + 4 * k;
        int mitot = b[1] - b[0] + 1 + 2 * ng, mjtot = b[3] - b[2] + 1 + 2 * ng;
        const double* q = h->qnew[level] + h->qoff[level][k];
        if (p->mcapa > 0 && h->aux[level]) {                       /* conck.f:56-60 */
            const double* a = h->aux[level] + h->auxoff[level][k];
            for (int j = ng + 1; j <= mjtot - ng; j++)
                for (int i = ng + 1; i <= mitot - ng; i++)
                    totmass = totmass + VAL(q, nv, mitot, 1, i, j) * VAL(a, p->maux, mitot, p->mcapa, i, j);
        } else {
            for (int j = ng + 1; j <= mjtot - ng; j++)
                for (int i = ng + 1; i <= mitot - ng; i++) totmass = totmass + VAL(q, nv, mitot, 1, i, j);
        }
    }
    return totmass * h->hx[level] * h->hy[level];
}
