/* orc_flag.c -- test infrastructure: CPU restatement of the cell flagging
 * (flagregions2.f90:40-120, flag2refine2.f90:53-124 and the Richardson
 * estimate of errest.f / errf1.f, as flagger.f:56-164 calls them).  Pinned
 * against the reference's own Fortran (compiled into oracle/_ref). */
#include <math.h>
#include <stdlib.h>
#include <string.h>
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

/* errest.f:4-63 for grid k of `level`: prepbigstep.f:4-69 (the old values,
 * copysol.f into an array with 2 nghost ghost cells, bound at the old time
 * with that width, coarsen.f by 2, one stepgrid of 2 dt on 2 h),
 * prepregstep.f:4-32 (one stepgrid of dt on `big`, the new values with
 * their ghost cells, or the old ones at the initial time) and errf1.f:4-155.
 * errf1 indexes the flags from nghost+1 although they carry mbuff border
 * cells (flagger.f:121-123), so for mbuff > nghost its marks sit
 * mbuff - nghost cells down and left of the coarse cell they test; this is
 * reproduced as written. */
int orc_errest(const orc_params* p, const orc_hier* h, int level, int k, const double* big,
               double tol, int iorder, int mbuff, double* flags) {
    const int nv = p->meqn, ng = p->mbc, ng2 = 2 * ng;
    const int* b = h->boxes[level] + 4 * k;
    const int nx = b[1] - b[0] + 1, ny = b[3] - b[2] + 1;
    const int mitot = nx + 2 * ng, mjtot = ny + 2 * ng;
    const int mi2tot = nx / 2 + 2 * ng, mj2tot = ny / 2 + 2 * ng;
    const int midub = nx + 4 * ng, mjdub = ny + 4 * ng;
    const int mibuff = nx + 2 * mbuff;
    const double hx = h->hx[level], hy = h->hy[level];
    const double hx2 = 2.0 * hx, hy2 = 2.0 * hy;
    const double dt = h->possk[level], dt2 = 2.0 * dt;
    const double time = h->time[level], tpre = time - dt;
    double* valdub = calloc((size_t)nv * midub * mjdub, sizeof(double));
    double* valbgc = calloc((size_t)nv * mi2tot * mj2tot, sizeof(double));
    double* fine = malloc((size_t)nv * mitot * mjtot * sizeof(double));
    double* f = calloc((size_t)4 * nv * mitot * mjtot, sizeof(double));
    if (!valdub || !valbgc || !fine || !f) {
        free(valdub); free(valbgc); free(fine); free(f);
        return ORC_E_ALLOC;
    }
    /* copysol.f:8-13 (old values, interior only) */
    const double* old = h->qold[level] + h->qoff[level][k];
    for (int j = ng + 1; j <= mjtot - ng; j++)
        for (int i = ng + 1; i <= mitot - ng; i++)
            for (int m = 0; m < nv; m++)
                valdub[m + (size_t)nv * ((i - ng + ng2 - 1) + (size_t)midub * (j - ng + ng2 - 1))] =
                    old[m + (size_t)nv * ((i - 1) + (size_t)mitot * (j - 1))];
    int rc = orc_bound_ng(p, h, level, k, tpre, valdub, ng2);
    if (rc) { free(valdub); free(valbgc); free(fine); free(f); return rc; }
    /* coarsen.f:15-30 */
    for (int j = 1; j <= mj2tot; j++) {
        int jf = 2 * (j - 1) + 1;
        for (int i = 1; i <= mi2tot; i++) {
            int ifn = 2 * (i - 1) + 1;
            for (int m = 0; m < nv; m++) {
#define VD(ii, jj) valdub[m + (size_t)nv * (((ii) - 1) + (size_t)midub * ((jj) - 1))]
                valbgc[m + (size_t)nv * ((i - 1) + (size_t)mi2tot * (j - 1))] =
                    (VD(ifn, jf) + VD(ifn + 1, jf) + VD(ifn, jf + 1) + VD(ifn + 1, jf + 1)) / 4.0;
#undef VD
            }
        }
    }
    double cfl, dtnew;
    size_t n2 = (size_t)nv * mi2tot * mj2tot;
    orc_stepgrid(p, nx / 2, ny / 2, valbgc, NULL, hx2, hy2, dt2, f, f + n2, f + 2 * n2, f + 3 * n2,
                 &cfl, &dtnew);
    memcpy(fine, big, (size_t)nv * mitot * mjtot * sizeof(double));
    size_t n1 = (size_t)nv * mitot * mjtot;
    orc_stepgrid(p, nx, ny, fine, NULL, hx, hy, dt, f, f + n1, f + 2 * n1, f + 3 * n1, &cfl, &dtnew);
    /* errf1.f:33-137 */
    const double order = (double)((1 << (iorder + 1)) - 2);
#define FL(ii, jj) flags[((ii) - 1) + (size_t)mibuff * ((jj) - 1)]
#define RF(ii, jj) fine[(size_t)nv * (((ii) - 1) + (size_t)mitot * ((jj) - 1))]
#define RC(ii, jj) valbgc[(size_t)nv * (((ii) - 1) + (size_t)mi2tot * ((jj) - 1))]
    int jfine = ng + 1;
    for (int j = ng + 1; j <= mj2tot - ng; j++) {
        int ifine = ng + 1;
        for (int i = ng + 1; i <= mi2tot - ng; i++) {
            double rflag = ORC_UNSET;
            if (FL(ifine, jfine) == ORC_UNSET || FL(ifine + 1, jfine) == ORC_UNSET ||
                FL(ifine, jfine + 1) == ORC_UNSET || FL(ifine + 1, jfine + 1) == ORC_UNSET) {
                double term1 = RF(ifine, jfine), term2 = RF(ifine + 1, jfine);
                double term3 = RF(ifine + 1, jfine + 1), term4 = RF(ifine, jfine + 1);
                double aval = (term1 + term2 + term3 + term4) / 4.0;
                double est = fabs((aval - RC(i, j)) / order);
                if (est >= tol) rflag = ORC_DOFLAG;
            }
            RC(i, j) = rflag;
            ifine += 2;
        }
        jfine += 2;
    }
    jfine = ng + 1;
    for (int j = ng + 1; j <= mj2tot - ng; j++) {
        int ifine = ng + 1;
        for (int i = ng + 1; i <= mi2tot - ng; i++) {
            if (RC(i, j) == ORC_DOFLAG) {
                FL(ifine, jfine) = ORC_DOFLAG;
                FL(ifine + 1, jfine) = ORC_DOFLAG;
                FL(ifine, jfine + 1) = ORC_DOFLAG;
                FL(ifine + 1, jfine + 1) = ORC_DOFLAG;
            }
            ifine += 2;
        }
        jfine += 2;
    }
#undef FL
#undef RF
#undef RC
    free(valdub); free(valbgc); free(fine); free(f);
    return 0;
}

/* flagger.f:56-164 for every grid of `level` without Richardson estimation:
 * flags(mx+2 mbuff, my+2 mbuff) per grid at flags + foff[k], set UNSET,
 * then regions, then gradients of the new values (old at the initial time).
 * The caller has filled the ghost cells (bound) of the array used. */
int orc_flag_level(const orc_params* p, const orc_hier* h, int level, double tolsp,
                   const double* regions, int nreg, int use_old, int mbuff, double* flags,
                   const long* foff, int flag_gradient, int flag_richardson, double tol,
                   int iorder) {
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
        if (flag_gradient) orc_flag2refine2(mx, my, ng, mbuff, nv, tolsp, q, f);
        /* errest runs while any flag is UNSET; the mbuff border always is */
        if (flag_richardson && (rc = orc_errest(p, h, level, k, q, tol, iorder, mbuff, f)))
            return rc;
    }
    return 0;
}
