/*
 * orc_sweep.c -- oracle for the wave-propagation sweep (TEST INFRASTRUCTURE
 * ONLY).  Plain-C restatement of limiter/flux2/step2/stepgrid in the
 * reference's evaluation order; arrays use the reference's Fortran layouts.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "orc.h"

static inline double dmin1(double a, double b) { return a < b ? a : b; }
static inline double dmax1(double a, double b) { return a > b ? a : b; }

/* 1-D slice helpers: Fortran index i in [1-mbc, maxm+mbc] */
#define P1(i) ((i) + mbc - 1)
#define V1(a, m, i) (a)[((m) - 1) + meqn * P1(i)]
#define WV(m, mw, i) wave[((m) - 1) + meqn * (((mw) - 1) + mwaves * P1(i))]
#define SV(mw, i) s[((mw) - 1) + mwaves * P1(i)]
#define GADD(a, m, i, k) (a)[((m) - 1) + meqn * (P1(i) + L1 * ((k) - 1))]
#define DT1(i) dtdx1d[P1(i)]

/* inlinelimiter.f:4-108.  Waves limited in place for i = 1..mx+1; dotr is
 * carried from one interface to the next exactly as in the reference loop
 * (:41-48), upwind ratio chosen by sign(s) (:52-56). */
void orc_limiter(int maxm, int meqn, int mwaves, int mbc, int mx, double* wave,
                 const double* s, const int* mthlim) {
    (void)maxm;
    for (int mw = 1; mw <= mwaves; mw++) {
        if (mthlim[mw - 1] == 0) continue;
        double dotr = 0.0;
        for (int i = 0; i <= mx + 1; i++) {
            double wnorm2 = 0.0;
            double dotl = dotr;
            dotr = 0.0;
            for (int m = 1; m <= meqn; m++) {
                wnorm2 = wnorm2 + WV(m, mw, i) * WV(m, mw, i);
                dotr = dotr + WV(m, mw, i) * WV(m, mw, i + 1);
            }
            if (i == 0) continue;
            if (wnorm2 == 0.0) continue;
            double r = (SV(mw, i) > 0.0) ? dotl / wnorm2 : dotr / wnorm2;
            double wlimitr;
            switch (mthlim[mw - 1]) {
            case 1: /* minmod */
                wlimitr = dmax1(0.0, dmin1(1.0, r));
                break;
            case 2: /* superbee */
                wlimitr = dmax1(dmax1(0.0, dmin1(1.0, 2.0 * r)), dmin1(2.0, r));
                break;
            case 3: /* van Leer */
                wlimitr = (r + fabs(r)) / (1.0 + fabs(r));
                break;
            case 4: { /* monotonized centred */
                double c = (1.0 + r) / 2.0;
                wlimitr = dmax1(0.0, dmin1(dmin1(c, 2.0), 2.0 * r));
                break;
            }
            case 5: /* Beam-Warming */
                wlimitr = r;
                break;
            default:
                wlimitr = 1.0;
            }
            for (int m = 1; m <= meqn; m++) WV(m, mw, i) = wlimitr * WV(m, mw, i);
        }
    }
}

/* flux2.f:25-237, one slice.  mbc = p->mbc; arrays sized for maxm. */
void orc_flux2(const orc_params* p, int ixy, int maxm, int mx,
               const double* q1d, const double* dtdx1d, const double* aux1,
               const double* aux2, const double* aux3, double* faddm,
               double* faddp, double* gaddm, double* gaddp, double* cfl1d,
               double* wave, double* s, double* amdq, double* apdq,
               double* cqxx, double* bmasdq, double* bpasdq) {
    const int meqn = p->meqn, mwaves = p->mwaves, mbc = p->mbc;
    const int L1 = maxm + 2 * mbc;
    int limit = 0;
    for (int mw = 1; mw <= mwaves; mw++)
        if (p->mthlim[mw - 1] > 0) limit = 1;                /* :102-105 */

    for (int i = 1 - mbc; i <= mx + mbc; i++)                 /* :110-119 */
        for (int m = 1; m <= meqn; m++) {
            V1(faddm, m, i) = 0.0;
            V1(faddp, m, i) = 0.0;
            GADD(gaddm, m, i, 1) = 0.0;
            GADD(gaddp, m, i, 1) = 0.0;
            GADD(gaddm, m, i, 2) = 0.0;
            GADD(gaddp, m, i, 2) = 0.0;
        }

    orc_rpn2(p, ixy, maxm, meqn, mwaves, p->maux, mbc, mx, q1d, q1d, aux2,
             aux2, wave, s, amdq, apdq);                       /* :125 */

    for (int i = 1; i <= mx + 1; i++)                         /* :129-133 */
        for (int m = 1; m <= meqn; m++) {
            V1(faddp, m, i) = V1(faddp, m, i) - V1(apdq, m, i);
            V1(faddm, m, i) = V1(faddm, m, i) + V1(amdq, m, i);
        }

    double cfl = 0.0;                                          /* :136-143 */
    for (int mw = 1; mw <= mwaves; mw++)
        for (int i = 1; i <= mx + 1; i++)
            cfl = dmax1(dmax1(cfl, DT1(i) * SV(mw, i)), -DT1(i - 1) * SV(mw, i));
    *cfl1d = cfl;

    if (p->method[2] != 1) {
        if (limit) orc_limiter(maxm, meqn, mwaves, mbc, mx, wave, s, p->mthlim);
        for (int i = 1; i <= mx + 1; i++) {                   /* :153-180 */
            double dtdxave = 0.5 * (DT1(i - 1) + DT1(i));
            for (int m = 1; m <= meqn; m++) {
                double c = 0.0;
                for (int mw = 1; mw <= mwaves; mw++) {
                    double abs_sign = p->use_fwaves
                                          ? copysign(1.0, SV(mw, i))
                                          : fabs(SV(mw, i));
                    c = c + abs_sign * (1.0 - fabs(SV(mw, i)) * dtdxave)
                                * WV(m, mw, i);
                }
                V1(cqxx, m, i) = c;
                V1(faddm, m, i) = V1(faddm, m, i) + 0.5 * c;
                V1(faddp, m, i) = V1(faddp, m, i) + 0.5 * c;
            }
        }
    }

    if (p->method[3] == 0) return;                             /* :185 */

    if (p->method[2] > 1 && p->method[3] == 2)                 /* :187-194 */
        for (int i = 1; i <= mx + 1; i++)
            for (int m = 1; m <= meqn; m++) {
                V1(amdq, m, i) = V1(amdq, m, i) + V1(cqxx, m, i);
                V1(apdq, m, i) = V1(apdq, m, i) - V1(cqxx, m, i);
            }

    orc_rpt2(p, ixy, 1, maxm, meqn, mwaves, p->maux, mbc, mx, q1d, q1d, aux1,
             aux2, aux3, amdq, bmasdq, bpasdq);                /* :202 */
    for (int i = 1; i <= mx + 1; i++)                         /* :207-216 */
        for (int m = 1; m <= meqn; m++) {
            double gupdate = 0.5 * DT1(i - 1) * V1(bmasdq, m, i);
            GADD(gaddm, m, i - 1, 1) = GADD(gaddm, m, i - 1, 1) - gupdate;
            GADD(gaddp, m, i - 1, 1) = GADD(gaddp, m, i - 1, 1) - gupdate;
            gupdate = 0.5 * DT1(i - 1) * V1(bpasdq, m, i);
            GADD(gaddm, m, i - 1, 2) = GADD(gaddm, m, i - 1, 2) - gupdate;
            GADD(gaddp, m, i - 1, 2) = GADD(gaddp, m, i - 1, 2) - gupdate;
        }

    orc_rpt2(p, ixy, 2, maxm, meqn, mwaves, p->maux, mbc, mx, q1d, q1d, aux1,
             aux2, aux3, apdq, bmasdq, bpasdq);                /* :219 */
    for (int i = 1; i <= mx + 1; i++)                         /* :224-233 */
        for (int m = 1; m <= meqn; m++) {
            double gupdate = 0.5 * DT1(i) * V1(bmasdq, m, i);
            GADD(gaddm, m, i, 1) = GADD(gaddm, m, i, 1) - gupdate;
            GADD(gaddp, m, i, 1) = GADD(gaddp, m, i, 1) - gupdate;
            gupdate = 0.5 * DT1(i) * V1(bpasdq, m, i);
            GADD(gaddm, m, i, 2) = GADD(gaddm, m, i, 2) - gupdate;
            GADD(gaddp, m, i, 2) = GADD(gaddp, m, i, 2) - gupdate;
        }
}

/* 2-D patch helpers: q(m, i, j), i in [1-mbc, mx+mbc], j in [1-mbc, my+mbc] */
#define P2(a, nv, m, i, j) \
    (a)[((m) - 1) + (nv) * (((i) + mbc - 1) + mitot * ((j) + mbc - 1))]

/* step2.f90:7-169 */
void orc_step2(const orc_params* p, int mx, int my, const double* qold,
               const double* aux, double dx, double dy, double dt,
               double* cflgrid, double* fm, double* fp, double* gm,
               double* gp) {
    const int meqn = p->meqn, mwaves = p->mwaves, maux = p->maux, mbc = p->mbc;
    const int mitot = mx + 2 * mbc, mjtot = my + 2 * mbc;
    const int maxm = mx > my ? mx : my;
    const int L1 = maxm + 2 * mbc;
    const int naux1 = maux > 0 ? maux : 1;

    double* faddm = calloc((size_t)meqn * L1, sizeof(double));
    double* faddp = calloc((size_t)meqn * L1, sizeof(double));
    double* gaddm = calloc((size_t)meqn * L1 * 2, sizeof(double));
    double* gaddp = calloc((size_t)meqn * L1 * 2, sizeof(double));
    double* q1d = calloc((size_t)meqn * L1, sizeof(double));
    double* aux1 = calloc((size_t)naux1 * L1, sizeof(double));
    double* aux2 = calloc((size_t)naux1 * L1, sizeof(double));
    double* aux3 = calloc((size_t)naux1 * L1, sizeof(double));
    double* dtdx1d = calloc((size_t)L1, sizeof(double));
    double* dtdy1d = calloc((size_t)L1, sizeof(double));
    double* wave = calloc((size_t)meqn * mwaves * L1, sizeof(double));
    double* s = calloc((size_t)mwaves * L1, sizeof(double));
    double* amdq = calloc((size_t)meqn * L1, sizeof(double));
    double* apdq = calloc((size_t)meqn * L1, sizeof(double));
    double* cqxx = calloc((size_t)meqn * L1, sizeof(double));
    double* bmadq = calloc((size_t)meqn * L1, sizeof(double));
    double* bpadq = calloc((size_t)meqn * L1, sizeof(double));

    double cfl = 0.0;
    double dtdx = dt / dx, dtdy = dt / dy;
    size_t ntot = (size_t)meqn * mitot * mjtot;
    memset(fm, 0, ntot * sizeof(double));                      /* :79-82 */
    memset(fp, 0, ntot * sizeof(double));
    memset(gm, 0, ntot * sizeof(double));
    memset(gp, 0, ntot * sizeof(double));

    /* x-sweeps :86-124 */
    for (int j = 0; j <= my + 1; j++) {
        for (int i = 1 - mbc; i <= mx + mbc; i++)
            for (int m = 1; m <= meqn; m++)
                V1(q1d, m, i) = P2(qold, meqn, m, i, j);
        for (int i = 1 - mbc; i <= mx + mbc; i++)
            dtdx1d[P1(i)] = p->mcapa > 0
                                ? dtdx / P2(aux, maux, p->mcapa, i, j)
                                : dtdx;
        if (p->mcapa == 0)
            for (int k = 0; k < L1; k++) dtdx1d[k] = dtdx;
        if (maux > 0)
            for (int i = 1 - mbc; i <= mx + mbc; i++)
                for (int ma = 1; ma <= maux; ma++) {
                    aux1[(ma - 1) + maux * P1(i)] =
                        P2(aux, maux, ma, i, j - 1 < 1 - mbc ? 1 - mbc : j - 1);
                    aux2[(ma - 1) + maux * P1(i)] = P2(aux, maux, ma, i, j);
                    aux3[(ma - 1) + maux * P1(i)] =
                        P2(aux, maux, ma, i, j + 1 > my + mbc ? my + mbc : j + 1);
                }
        double cfl1d;
        orc_flux2(p, 1, maxm, mx, q1d, dtdx1d, aux1, aux2, aux3, faddm, faddp,
                  gaddm, gaddp, &cfl1d, wave, s, amdq, apdq, cqxx, bmadq,
                  bpadq);
        cfl = dmax1(cfl, cfl1d);
        for (int i = 1; i <= mx + 1; i++)
            for (int m = 1; m <= meqn; m++) {
                P2(fm, meqn, m, i, j) = P2(fm, meqn, m, i, j) + V1(faddm, m, i);
                P2(fp, meqn, m, i, j) = P2(fp, meqn, m, i, j) + V1(faddp, m, i);
                P2(gm, meqn, m, i, j) = P2(gm, meqn, m, i, j) + GADD(gaddm, m, i, 1);
                P2(gp, meqn, m, i, j) = P2(gp, meqn, m, i, j) + GADD(gaddp, m, i, 1);
                P2(gm, meqn, m, i, j + 1) = P2(gm, meqn, m, i, j + 1) + GADD(gaddm, m, i, 2);
                P2(gp, meqn, m, i, j + 1) = P2(gp, meqn, m, i, j + 1) + GADD(gaddp, m, i, 2);
            }
    }

    /* y-sweeps :129-166 */
    for (int i = 0; i <= mx + 1; i++) {
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int m = 1; m <= meqn; m++)
                V1(q1d, m, j) = P2(qold, meqn, m, i, j);
        for (int j = 1 - mbc; j <= my + mbc; j++)
            dtdy1d[P1(j)] = p->mcapa > 0
                                ? dtdy / P2(aux, maux, p->mcapa, i, j)
                                : dtdy;
        if (p->mcapa == 0)
            for (int k = 0; k < L1; k++) dtdy1d[k] = dtdy;
        if (maux > 0)
            for (int j = 1 - mbc; j <= my + mbc; j++)
                for (int ma = 1; ma <= maux; ma++) {
                    aux1[(ma - 1) + maux * P1(j)] =
                        P2(aux, maux, ma, i - 1 < 1 - mbc ? 1 - mbc : i - 1, j);
                    aux2[(ma - 1) + maux * P1(j)] = P2(aux, maux, ma, i, j);
                    aux3[(ma - 1) + maux * P1(j)] =
                        P2(aux, maux, ma, i + 1 > mx + mbc ? mx + mbc : i + 1, j);
                }
        double cfl1d;
        orc_flux2(p, 2, maxm, my, q1d, dtdy1d, aux1, aux2, aux3, faddm, faddp,
                  gaddm, gaddp, &cfl1d, wave, s, amdq, apdq, cqxx, bmadq,
                  bpadq);
        cfl = dmax1(cfl, cfl1d);
        for (int j = 1; j <= my + 1; j++)
            for (int m = 1; m <= meqn; m++) {
                P2(gm, meqn, m, i, j) = P2(gm, meqn, m, i, j) + V1(faddm, m, j);
                P2(gp, meqn, m, i, j) = P2(gp, meqn, m, i, j) + V1(faddp, m, j);
                P2(fm, meqn, m, i, j) = P2(fm, meqn, m, i, j) + GADD(gaddm, m, j, 1);
                P2(fp, meqn, m, i, j) = P2(fp, meqn, m, i, j) + GADD(gaddp, m, j, 1);
                P2(fm, meqn, m, i + 1, j) = P2(fm, meqn, m, i + 1, j) + GADD(gaddm, m, j, 2);
                P2(fp, meqn, m, i + 1, j) = P2(fp, meqn, m, i + 1, j) + GADD(gaddp, m, j, 2);
            }
    }
    *cflgrid = cfl;

    free(faddm); free(faddp); free(gaddm); free(gaddp); free(q1d);
    free(aux1); free(aux2); free(aux3); free(dtdx1d); free(dtdy1d);
    free(wave); free(s); free(amdq); free(apdq); free(cqxx);
    free(bmadq); free(bpadq);
}

/* stepgrid.f:129-212: step2, conservative update (:154-173), dtnew. */
void orc_stepgrid(const orc_params* p, int mx, int my, double* q,
                  const double* aux, double dx, double dy, double dt,
                  double* fm, double* fp, double* gm, double* gp,
                  double* cflgrid, double* dtnew) {
    const int meqn = p->meqn, mbc = p->mbc, maux = p->maux;
    const int mitot = mx + 2 * mbc;
    orc_step2(p, mx, my, q, aux, dx, dy, dt, cflgrid, fm, fp, gm, gp);
    double dtdx = dt / dx, dtdy = dt / dy;
    for (int j = 1; j <= my; j++)
        for (int i = 1; i <= mx; i++)
            for (int m = 1; m <= meqn; m++) {
                if (p->mcapa == 0)
                    P2(q, meqn, m, i, j) =
                        P2(q, meqn, m, i, j)
                        - dtdx * (P2(fm, meqn, m, i + 1, j) - P2(fp, meqn, m, i, j))
                        - dtdy * (P2(gm, meqn, m, i, j + 1) - P2(gp, meqn, m, i, j));
                else
                    P2(q, meqn, m, i, j) =
                        P2(q, meqn, m, i, j)
                        - (dtdx * (P2(fm, meqn, m, i + 1, j) - P2(fp, meqn, m, i, j))
                           + dtdy * (P2(gm, meqn, m, i, j + 1) - P2(gp, meqn, m, i, j)))
                              / P2(aux, maux, p->mcapa, i, j);
            }
    if (*cflgrid > 0.0)
        *dtnew = dt * p->cfl / *cflgrid;
    else
        *dtnew = ORC_RINFINITY;
}
