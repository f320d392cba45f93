/* orc_flag.c -- test infrastructure: CPU restatement of the cell flagging
 * of one grid (flagregions2.f90:40-120 followed by flag2refine2.f90:53-124,
 * as flagger.f:131-164 calls them).  Pinned against the reference's own
 * Fortran (compiled into oracle/_ref) on random grids and regions. */
#include <math.h>
#include <stdlib.h>
#include "orc.h"

enum { ORC_E_ALLOC = 16, ORC_E_REGION = 17 };

#define FLAG(f, i, j) (f)[((i) + mbuff - 1) + (size_t)(mx + 2 * mbuff) * ((j) + mbuff - 1)]

/* flagregions2.f90:40-120.  regions: nreg records of 8 doubles in the order
 * of regions.data (regions_module.f90): min_level, max_level, t_low, t_hi,
 * x_low, x_hi, y_low, y_hi.  Returns ORC_E_REGION if a cell's minimum level
 * exceeds its maximum level (the reference stops there). */
int orc_flagregions2(int mx, int my, int mbuff, double xlower, double ylower, double dx,
                     double dy, int level, double t, const double* regions, int nreg,
                     double* amrflags) {
    int* minl = calloc((size_t)mx * my, sizeof(int));
    int* maxl = calloc((size_t)mx * my, sizeof(int));
    if (!minl || !maxl) { free(minl); free(maxl); return ORC_E_ALLOC; }
    const double xupper = xlower + mx * dx, yupper = ylower + my * dy;
    for (int m = 0; m < nreg; m++) {
        const double* r = regions + 8 * m;
        int rmin = (int)r[0], rmax = (int)r[1];
        double t_low = r[2], t_hi = r[3], x_low = r[4], x_hi = r[5], y_low = r[6], y_hi = r[7];
        if (t < t_low || t > t_hi) continue;
        if (xlower >= x_hi || xupper <= x_low) continue;
        int i1 = (int)floor((x_low - xlower) / dx) + 1; if (i1 < 1) i1 = 1;
        int i2 = (int)floor((x_hi - xlower) / dx) + 1; if (i2 > mx) i2 = mx;
        if (ylower >= y_hi || yupper <= y_low) continue;
        int j1 = (int)floor((y_low - ylower) / dy) + 1; if (j1 < 1) j1 = 1;
        int j2 = (int)floor((y_hi - ylower) / dy) + 1; if (j2 > my) j2 = my;
        for (int j = j1; j <= j2; j++)
            for (int i = i1; i <= i2; i++) {
                int* a = &minl[(i - 1) + (size_t)mx * (j - 1)];
                int* b = &maxl[(i - 1) + (size_t)mx * (j - 1)];
                if (rmin > *a) *a = rmin;
                if (rmax > *b) *b = rmax;
            }
    }
    int rc = 0;
    for (int j = 1; j <= my && !rc; j++)
        for (int i = 1; i <= mx; i++) {
            int a = minl[(i - 1) + (size_t)mx * (j - 1)], b = maxl[(i - 1) + (size_t)mx * (j - 1)];
            if (a > b) { rc = ORC_E_REGION; break; }
            if (b != 0) {
                if (level < a) FLAG(amrflags, i, j) = ORC_DOFLAG;
                else if (level >= b) FLAG(amrflags, i, j) = ORC_DONTFLAG;
            }
        }
    free(minl);
    free(maxl);
    return rc;
}

/* flag2refine2.f90:96-122: undivided differences of every component */
void orc_flag2refine2(int mx, int my, int mbc, int mbuff, int meqn, double tolsp, const double* q,
                      double* amrflags) {
    const int mitot = mx + 2 * mbc;
#define Q(m, i, j) q[(m) + (size_t)meqn * (((i) + mbc - 1) + (size_t)mitot * ((j) + mbc - 1))]
    for (int j = 1; j <= my; j++)
        for (int i = 1; i <= mx; i++) {
            if (FLAG(amrflags, i, j) != ORC_UNSET) continue;
            for (int m = 0; m < meqn; m++) {
                double dqi = fabs(Q(m, i + 1, j) - Q(m, i - 1, j));
                double dqj = fabs(Q(m, i, j + 1) - Q(m, i, j - 1));
                double dq = 0.0;
                if (dqi > dq) dq = dqi;
                if (dqj > dq) dq = dqj;
                if (dq > tolsp) { FLAG(amrflags, i, j) = ORC_DOFLAG; break; }
            }
        }
#undef Q
}

/* flagger.f:56-164 for every grid of `level` without Richardson estimation:
 * flags(mx+2 mbuff, my+2 mbuff) per grid at flags + foff[k], set UNSET,
 * then regions, then gradients of the new values (old at the initial time).
 * The caller has filled the ghost cells (bound) of the array used. */
int orc_flag_level(const orc_params* p, const orc_hier* h, int level, double tolsp,
                   const double* regions, int nreg, int use_old, int mbuff, double* flags,
                   const long* foff) {
    const int nv = p->meqn, ng = p->mbc;
    for (int k = 0; k < h->ngrids[level]; k++) {
        const int* b = h->boxes[level] + 4 * k;
        int mx = b[1] - b[0] + 1, my = b[3] - b[2] + 1;
        double* f = flags + foff[k];
        for (size_t c = 0; c < (size_t)(mx + 2 * mbuff) * (my + 2 * mbuff); c++) f[c] = ORC_UNSET;
        double xleft = h->xlower + b[0] * h->hx[level], ybot = h->ylower + b[2] * h->hy[level];
        int rc = orc_flagregions2(mx, my, mbuff, xleft, ybot, h->hx[level], h->hy[level], level,
                                  h->time[level], regions, nreg, f);
        if (rc) return rc;
        const double* q = (use_old ? h->qold[level] : h->qnew[level]) + h->qoff[level][k];
        orc_flag2refine2(mx, my, ng, mbuff, nv, tolsp, q, f);
    }
    return 0;
}
