/*
 * orc_sweep3.c -- oracle of the 3-D sweep (TEST INFRASTRUCTURE ONLY).
 *
 * CPU restatement of src/3d/step3.f:3-486 (x-, y- and z-sweeps and the
 * accumulation of the transverse terms into the six flux arrays),
 * src/3d/flux3.f:4-633 (normal solve, CFL, limiter, second-order
 * correction, the transverse solves rpt3 of A-dq / A+dq and of the
 * correction cqxx, and the double-transverse solves rptt3 for
 * method(3) = 10 m3 + m4) and the update and dt of src/3d/stepgrid.f:79-230.
 * The limiter is inlinelimiter.f (identical in src/2d and src/3d), shared
 * with the 2-D oracle.  Arithmetic follows the Fortran's order; built with
 * -ffp-contract=off.
 *
 * The 3-D Riemann solvers (rpn3, rpt3, rptt3) are clawpack/riemann's,
 * absent from the reference tree; the ones here restate the published
 * algorithms (LeVeque 2002, ch. 18-21) in the same argument convention:
 * constant-velocity advection and the Euler equations with the 3-wave Roe
 * split (acoustic waves u -+ c, one wave at u carrying the entropy and both
 * shear jumps) and the Harten-Hyman entropy fix of the 2-D solver.  The
 * transverse solvers split a fluctuation along velocity direction mv
 * (icoor = 2) or mw (icoor = 3) of the sweep's cyclic triple (mu, mv, mw)
 * with the interface's Roe average.  Their arithmetic is pinned by closed
 * forms (tests/test_oracle_euler_eigen.py): every wave is the spectral part
 * of the jump for its eigenvalue of the Roe matrix (the complex-step flux
 * Jacobian at the Roe average), A-dq / A+dq and each rpt3 / rptt3 split
 * are the spectral splits, and the entropy fix satisfies its two-speed
 * identity and reflection symmetry.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define POS(i) ((i) + mbc - 1)
#define Q1(a, m, i) (a)[((m) - 1) + meqn * POS(i)]
#define W3(m, mw, i) wave[((m) - 1) + meqn * (((mw) - 1) + mwaves * POS(i))]
#define S3(mw, i) s[((mw) - 1) + mwaves * POS(i)]

static inline double dmin1(double a, double b) { return a < b ? a : b; }
static inline double dmax1(double a, double b) { return a > b ? a : b; }

/* velocity components (1-based) of the cyclic triple of sweep ixyz */
static void triple(int ixyz, int* mu, int* mv, int* mw) {
    if (ixyz == 1) { *mu = 2; *mv = 3; *mw = 4; }
    else if (ixyz == 2) { *mu = 3; *mv = 4; *mw = 2; }
    else { *mu = 4; *mv = 2; *mw = 3; }
}

typedef struct {
    double u[5];            /* u[m] for velocity component m = 2, 3, 4 */
    double enth, a, g1a2, euv, u2;
} roe3;

/* Roe average of two Euler states (rho, rho u, rho v, rho w, E) */
static roe3 roe3_average(double gamma1, const double* L, const double* R) {
    roe3 r;
    double rhsqrtl = sqrt(L[0]), rhsqrtr = sqrt(R[0]);
    double pl = gamma1 * (L[4] - 0.5 * (L[1] * L[1] + L[2] * L[2] + L[3] * L[3]) / L[0]);
    double pr = gamma1 * (R[4] - 0.5 * (R[1] * R[1] + R[2] * R[2] + R[3] * R[3]) / R[0]);
    double rhsq2 = rhsqrtl + rhsqrtr;
    for (int m = 2; m <= 4; m++) r.u[m] = (L[m - 1] / rhsqrtl + R[m - 1] / rhsqrtr) / rhsq2;
    r.enth = ((L[4] + pl) / rhsqrtl + (R[4] + pr) / rhsqrtr) / rhsq2;
    r.u2 = r.u[2] * r.u[2] + r.u[3] * r.u[3] + r.u[4] * r.u[4];
    double a2 = gamma1 * (r.enth - 0.5 * r.u2);
    r.a = sqrt(a2);
    r.g1a2 = gamma1 / a2;
    r.euv = r.enth - r.u2;
    r.u[0] = r.u[1] = 0.0;
    return r;
}

/* split d (5 components) along velocity direction md with the other two
 * velocity directions o1, o2: three waves wv[3][5] and speeds sp[3] */
static void euler3_waves(const roe3* r, int md, int o1, int o2, const double* d, double wv[3][5],
                         double sp[3]) {
    double ud = r->u[md], a = r->a;
    double a4 = r->g1a2 * (r->euv * d[0] + r->u[2] * d[1] + r->u[3] * d[2] + r->u[4] * d[3] - d[4]);
    double a2 = d[o1 - 1] - r->u[o1] * d[0];
    double a3 = d[o2 - 1] - r->u[o2] * d[0];
    double a5 = (d[md - 1] + (a - ud) * d[0] - a * a4) / (2.0 * a);
    double a1 = d[0] - a4 - a5;
    wv[0][0] = a1;
    wv[0][md - 1] = a1 * (ud - a);
    wv[0][o1 - 1] = a1 * r->u[o1];
    wv[0][o2 - 1] = a1 * r->u[o2];
    wv[0][4] = a1 * (r->enth - ud * a);
    sp[0] = ud - a;
    wv[1][0] = a4;
    wv[1][md - 1] = a4 * ud;
    wv[1][o1 - 1] = a4 * r->u[o1] + a2;
    wv[1][o2 - 1] = a4 * r->u[o2] + a3;
    wv[1][4] = a4 * 0.5 * r->u2 + a2 * r->u[o1] + a3 * r->u[o2];
    sp[1] = ud;
    wv[2][0] = a5;
    wv[2][md - 1] = a5 * (ud + a);
    wv[2][o1 - 1] = a5 * r->u[o1];
    wv[2][o2 - 1] = a5 * r->u[o2];
    wv[2][4] = a5 * (r->enth + ud * a);
    sp[2] = ud + a;
}

static inline double e3_pressure(double gamma1, const double* q) {
    return gamma1 * (q[4] - 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0]);
}

void orc_rpn3(const orc_params* p, int ixyz, int maxm, int meqn, int mwaves, int mbc, int mx,
              const double* ql, const double* qr, double* wave, double* s, double* amdq,
              double* apdq) {
    (void)maxm;
    int mu, mv, mw;
    triple(ixyz, &mu, &mv, &mw);
    if (p->solver == ORC_ADVECTION) {
        double sp = ixyz == 1 ? p->ubar : (ixyz == 2 ? p->vbar : p->wbar);
        for (int i = 2 - mbc; i <= mx + mbc; i++) {
            W3(1, 1, i) = Q1(ql, 1, i) - Q1(qr, 1, i - 1);
            S3(1, i) = sp;
            Q1(amdq, 1, i) = dmin1(sp, 0.0) * W3(1, 1, i);
            Q1(apdq, 1, i) = dmax1(sp, 0.0) * W3(1, 1, i);
        }
        return;
    }
    const double gamma = p->gamma, gamma1 = p->gamma - 1.0;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        const double* L = &Q1(qr, 1, i - 1);
        const double* R = &Q1(ql, 1, i);
        roe3 r = roe3_average(gamma1, L, R);
        double d[5], wv[3][5], sp[3];
        for (int m = 0; m < 5; m++) d[m] = R[m] - L[m];
        euler3_waves(&r, mu, mv, mw, d, wv, sp);
        for (int w = 0; w < 3; w++) {
            for (int m = 1; m <= 5; m++) W3(m, w + 1, i) = wv[w][m - 1];
            S3(w + 1, i) = sp[w];
        }
        double am[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (!p->efix) {
            for (int w = 1; w <= 3; w++)
                for (int m = 1; m <= 5; m++) am[m - 1] = am[m - 1] + dmin1(S3(w, i), 0.0) * W3(m, w, i);
        } else {
            /* Harten-Hyman fix of the acoustic waves (LeVeque 2002 sec. 15.3.5) */
            double cim1 = sqrt(gamma * e3_pressure(gamma1, L) / L[0]);
            double s0 = L[mu - 1] / L[0] - cim1;
            if (!(s0 >= 0.0 && S3(1, i) > 0.0)) {
                double q1[5];
                for (int m = 0; m < 5; m++) q1[m] = L[m] + W3(m + 1, 1, i);
                double c1 = sqrt(gamma * e3_pressure(gamma1, q1) / q1[0]);
                double s1 = q1[mu - 1] / q1[0] - c1;
                double sfract;
                if (s0 < 0.0 && s1 > 0.0) sfract = s0 * (s1 - S3(1, i)) / (s1 - s0);
                else if (S3(1, i) < 0.0) sfract = S3(1, i);
                else sfract = 0.0;
                for (int m = 1; m <= 5; m++) am[m - 1] = sfract * W3(m, 1, i);
                if (S3(2, i) < 0.0) {
                    for (int m = 1; m <= 5; m++) am[m - 1] = am[m - 1] + S3(2, i) * W3(m, 2, i);
                    double ci = sqrt(gamma * e3_pressure(gamma1, R) / R[0]);
                    double s3 = R[mu - 1] / R[0] + ci;
                    double q2[5];
                    for (int m = 0; m < 5; m++) q2[m] = R[m] - W3(m + 1, 3, i);
                    double c2 = sqrt(gamma * e3_pressure(gamma1, q2) / q2[0]);
                    double s2 = q2[mu - 1] / q2[0] + c2;
                    double sf3 = 0.0;
                    int add3 = 1;
                    if (s2 < 0.0 && s3 > 0.0) sf3 = s2 * (s3 - S3(3, i)) / (s3 - s2);
                    else if (S3(3, i) < 0.0) sf3 = S3(3, i);
                    else add3 = 0;
                    if (add3)
                        for (int m = 1; m <= 5; m++) am[m - 1] = am[m - 1] + sf3 * W3(m, 3, i);
                }
            }
        }
        for (int m = 1; m <= 5; m++) {
            double df = 0.0;
            for (int w = 1; w <= 3; w++) df = df + S3(w, i) * W3(m, w, i);
            Q1(amdq, m, i) = am[m - 1];
            if (p->efix) {
                Q1(apdq, m, i) = df - am[m - 1];
            } else {
                double ap = 0.0;
                for (int w = 1; w <= 3; w++) ap = ap + dmax1(S3(w, i), 0.0) * W3(m, w, i);
                Q1(apdq, m, i) = ap;
            }
        }
    }
}

/* split asdq along direction icoor (2: mv, 3: mw) of sweep ixyz */
static void split3(const orc_params* p, int ixyz, int icoor, int meqn, int mbc, int mx,
                   const double* ql, const double* qr, const double* asdq, double* bm,
                   double* bp) {
    int mu, mv, mw;
    triple(ixyz, &mu, &mv, &mw);
    int md = icoor == 2 ? mv : mw;
    if (p->solver == ORC_ADVECTION) {
        double sp = md == 2 ? p->ubar : (md == 3 ? p->vbar : p->wbar);
        double spm = dmin1(sp, 0.0), spp = dmax1(sp, 0.0);
        for (int i = 2 - mbc; i <= mx + mbc; i++) {
            Q1(bm, 1, i) = spm * Q1(asdq, 1, i);
            Q1(bp, 1, i) = spp * Q1(asdq, 1, i);
        }
        return;
    }
    int o1 = md == mv ? mw : mu, o2 = md == mv ? mu : mv;
    const double gamma1 = p->gamma - 1.0;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        const double* L = &Q1(qr, 1, i - 1);
        const double* R = &Q1(ql, 1, i);
        roe3 r = roe3_average(gamma1, L, R);
        double wv[3][5], sp[3];
        euler3_waves(&r, md, o1, o2, &Q1(asdq, 1, i), wv, sp);
        for (int m = 1; m <= 5; m++) {
            double a = 0.0, b = 0.0;
            for (int w = 0; w < 3; w++) {
                a = a + dmin1(sp[w], 0.0) * wv[w][m - 1];
                b = b + dmax1(sp[w], 0.0) * wv[w][m - 1];
            }
            Q1(bm, m, i) = a;
            Q1(bp, m, i) = b;
        }
    }
}

void orc_rpt3(const orc_params* p, int ixyz, int icoor, int imp, int maxm, int meqn, int mwaves,
              int mbc, int mx, const double* ql, const double* qr, const double* asdq,
              double* bmasdq, double* bpasdq) {
    (void)imp; (void)maxm; (void)mwaves;
    split3(p, ixyz, icoor, meqn, mbc, mx, ql, qr, asdq, bmasdq, bpasdq);
}

void orc_rptt3(const orc_params* p, int ixyz, int icoor, int imp, int impt, int maxm, int meqn,
               int mwaves, int mbc, int mx, const double* ql, const double* qr,
               const double* bsasdq, double* cmbsasdq, double* cpbsasdq) {
    (void)imp; (void)impt; (void)maxm; (void)mwaves;
    split3(p, ixyz, icoor, meqn, mbc, mx, ql, qr, bsasdq, cmbsasdq, cpbsasdq);
}

/* ------------------------------------------------------------ flux3 ----- */
/* gadd/hadd(meqn, 1-mbc:maxm+mbc, 2, -1:1) */
#define GA(a, m, i, s, o) (a)[((m) - 1) + (size_t)meqn * (POS(i) + (size_t)n1 * (((s) - 1) + 2 * ((o) + 1)))]

void orc_flux3(const orc_params* p, int ixyz, int maxm, int meqn, int mwaves, int mbc, int mx,
               const double* q1d, const double* dtdx1d, double dtdy, double dtdz, double* faddm,
               double* faddp, double* gadd, double* hadd, double* cfl1d) {
    const int n1 = maxm + 2 * mbc;
    const size_t nq = (size_t)meqn * n1;
    double* buf = calloc(nq * 31 + (size_t)mwaves * n1 * (meqn + 1), sizeof(double));
    double* amdq = buf;            double* apdq = amdq + nq;       double* cqxx = apdq + nq;
    double* bmamdq = cqxx + nq;    double* bmapdq = bmamdq + nq;   double* bpamdq = bmapdq + nq;
    double* bpapdq = bpamdq + nq;  double* cmamdq = bpapdq + nq;   double* cmapdq = cmamdq + nq;
    double* cpamdq = cmapdq + nq;  double* cpapdq = cpamdq + nq;   double* cmamdq2 = cpapdq + nq;
    double* cmapdq2 = cmamdq2 + nq; double* cpamdq2 = cmapdq2 + nq; double* cpapdq2 = cpamdq2 + nq;
    double* bmcqxxm = cpapdq2 + nq; double* bmcqxxp = bmcqxxm + nq; double* bpcqxxm = bmcqxxp + nq;
    double* bpcqxxp = bpcqxxm + nq; double* cmcqxxm = bpcqxxp + nq; double* cmcqxxp = cmcqxxm + nq;
    double* cpcqxxm = cmcqxxp + nq; double* cpcqxxp = cpcqxxm + nq; double* bmcmamdq = cpcqxxp + nq;
    double* bmcmapdq = bmcmamdq + nq; double* bpcmamdq = bmcmapdq + nq; double* bpcmapdq = bpcmamdq + nq;
    double* bmcpamdq = bpcmapdq + nq; double* bmcpapdq = bmcpamdq + nq; double* bpcpamdq = bmcpapdq + nq;
    double* bpcpapdq = bpcpamdq + nq;
    double* wave = bpcpapdq + nq;
    double* s = wave + (size_t)meqn * mwaves * n1;
    (void)bmcpamdq;
#define D1(a, m, i) (a)[((m) - 1) + (size_t)meqn * POS(i)]
#define T1(i) dtdx1d[POS(i)]
    int limit = 0;
    for (int w = 0; w < mwaves; w++)
        if (p->mthlim[w] > 0) limit = 1;
    for (int m = 1; m <= meqn; m++)
        for (int i = 1 - mbc; i <= mx + mbc; i++) {
            D1(faddm, m, i) = 0.0;
            D1(faddp, m, i) = 0.0;
            for (int sd = 1; sd <= 2; sd++)
                for (int o = -1; o <= 1; o++) {
                    GA(gadd, m, i, sd, o) = 0.0;
                    GA(hadd, m, i, sd, o) = 0.0;
                }
        }
    int m3, m4;
    if (p->method[3] < 0) { m3 = -1; m4 = 0; }
    else { m3 = p->method[3] / 10; m4 = p->method[3] - 10 * m3; }
    orc_rpn3(p, ixyz, maxm, meqn, mwaves, mbc, mx, q1d, q1d, wave, s, amdq, apdq);
    for (int i = 1; i <= mx + 1; i++)
        for (int m = 1; m <= meqn; m++) {
            D1(faddp, m, i) = D1(faddp, m, i) - D1(apdq, m, i);
            D1(faddm, m, i) = D1(faddm, m, i) + D1(amdq, m, i);
        }
    *cfl1d = 0.0;
    for (int w = 1; w <= mwaves; w++)
        for (int i = 1; i <= mx + 1; i++)
            *cfl1d = dmax1(dmax1(*cfl1d, T1(i) * S3(w, i)), -T1(i - 1) * S3(w, i));
    if (p->method[2] != 1) {
        if (limit) orc_limiter(maxm, meqn, mwaves, mbc, mx, wave, s, p->mthlim);
        for (int i = 1; i <= mx + 1; i++) {
            double dtdxave = 0.5 * (T1(i - 1) + T1(i));
            for (int m = 1; m <= meqn; m++) {
                D1(cqxx, m, i) = 0.0;
                for (int w = 1; w <= mwaves; w++) {
                    double fac = p->use_fwaves ? (S3(w, i) >= 0.0 ? 1.0 : -1.0) : fabs(S3(w, i));
                    D1(cqxx, m, i) = D1(cqxx, m, i) + 0.5 * fac * (1.0 - fabs(S3(w, i)) * dtdxave) * W3(m, w, i);
                }
                D1(faddm, m, i) = D1(faddm, m, i) + D1(cqxx, m, i);
                D1(faddp, m, i) = D1(faddp, m, i) + D1(cqxx, m, i);
            }
        }
    }
    if (m3 <= 0) { free(buf); return; }
    orc_rpt3(p, ixyz, 2, 1, maxm, meqn, mwaves, mbc, mx, q1d, q1d, amdq, bmamdq, bpamdq);
    orc_rpt3(p, ixyz, 2, 2, maxm, meqn, mwaves, mbc, mx, q1d, q1d, apdq, bmapdq, bpapdq);
    orc_rpt3(p, ixyz, 3, 1, maxm, meqn, mwaves, mbc, mx, q1d, q1d, amdq, cmamdq, cpamdq);
    orc_rpt3(p, ixyz, 3, 2, maxm, meqn, mwaves, mbc, mx, q1d, q1d, apdq, cmapdq, cpapdq);
    if (m3 == 2) {
        /* no aux: imp = 0 and the m/p copies (flux3.f:173-186) */
        orc_rpt3(p, ixyz, 2, 0, maxm, meqn, mwaves, mbc, mx, q1d, q1d, cqxx, bmcqxxm, bpcqxxm);
        orc_rpt3(p, ixyz, 3, 0, maxm, meqn, mwaves, mbc, mx, q1d, q1d, cqxx, cmcqxxm, cpcqxxm);
        for (int i = 0; i <= mx + 2; i++)
            for (int m = 1; m <= meqn; m++) {
                D1(bmcqxxp, m, i) = D1(bmcqxxm, m, i);
                D1(bpcqxxp, m, i) = D1(bpcqxxm, m, i);
                D1(cmcqxxp, m, i) = D1(cmcqxxm, m, i);
                D1(cpcqxxp, m, i) = D1(cpcqxxm, m, i);
            }
    }
    if (m4 == 1) {
        for (int i = 0; i <= mx + 2; i++)
            for (int m = 1; m <= meqn; m++) {
                D1(cpapdq2, m, i) = D1(cpapdq, m, i);
                D1(cpamdq2, m, i) = D1(cpamdq, m, i);
                D1(cmapdq2, m, i) = D1(cmapdq, m, i);
                D1(cmamdq2, m, i) = D1(cmamdq, m, i);
            }
    } else if (m4 == 2) {
        for (int i = 0; i <= mx + 2; i++)
            for (int m = 1; m <= meqn; m++) {
                D1(cpapdq2, m, i) = D1(cpapdq, m, i) - 3.0 * D1(cpcqxxp, m, i);
                D1(cpamdq2, m, i) = D1(cpamdq, m, i) + 3.0 * D1(cpcqxxm, m, i);
                D1(cmapdq2, m, i) = D1(cmapdq, m, i) - 3.0 * D1(cmcqxxp, m, i);
                D1(cmamdq2, m, i) = D1(cmamdq, m, i) + 3.0 * D1(cmcqxxm, m, i);
            }
    }
    if (m4 > 0) {
        orc_rptt3(p, ixyz, 2, 2, 2, maxm, meqn, mwaves, mbc, mx, q1d, q1d, cpapdq2, bmcpapdq, bpcpapdq);
        orc_rptt3(p, ixyz, 2, 1, 2, maxm, meqn, mwaves, mbc, mx, q1d, q1d, cpamdq2, bmcpamdq, bpcpamdq);
        orc_rptt3(p, ixyz, 2, 2, 1, maxm, meqn, mwaves, mbc, mx, q1d, q1d, cmapdq2, bmcmapdq, bpcmapdq);
        orc_rptt3(p, ixyz, 2, 1, 1, maxm, meqn, mwaves, mbc, mx, q1d, q1d, cmamdq2, bmcmamdq, bpcmamdq);
    }
    const double sixth = 1.0 / 6.0;
    for (int i = 1; i <= mx + 1; i++)
        for (int m = 1; m <= meqn; m++) {
            GA(gadd, m, i - 1, 1, 0) = GA(gadd, m, i - 1, 1, 0) - 0.5 * T1(i - 1) * D1(bmamdq, m, i);
            GA(gadd, m, i - 1, 2, 0) = GA(gadd, m, i - 1, 2, 0) - 0.5 * T1(i - 1) * D1(bpamdq, m, i);
            GA(gadd, m, i, 1, 0) = GA(gadd, m, i, 1, 0) - 0.5 * T1(i) * D1(bmapdq, m, i);
            GA(gadd, m, i, 2, 0) = GA(gadd, m, i, 2, 0) - 0.5 * T1(i) * D1(bpapdq, m, i);
            if (m4 > 0) {
                GA(gadd, m, i, 2, 0) = GA(gadd, m, i, 2, 0)
                    + sixth * T1(i - 1) * dtdz * (D1(bpcpapdq, m, i) - D1(bpcmapdq, m, i));
                GA(gadd, m, i, 1, 0) = GA(gadd, m, i, 1, 0)
                    + sixth * T1(i) * dtdz * (D1(bmcpapdq, m, i) - D1(bmcmapdq, m, i));
                GA(gadd, m, i, 2, 1) = GA(gadd, m, i, 2, 1) - sixth * T1(i) * dtdz * D1(bpcpapdq, m, i);
                GA(gadd, m, i, 1, 1) = GA(gadd, m, i, 1, 1) - sixth * T1(i) * dtdz * D1(bmcpapdq, m, i);
                GA(gadd, m, i, 2, -1) = GA(gadd, m, i, 2, -1) + sixth * T1(i) * dtdz * D1(bpcmapdq, m, i);
                GA(gadd, m, i, 1, -1) = GA(gadd, m, i, 1, -1) + sixth * T1(i) * dtdz * D1(bmcmapdq, m, i);
                GA(gadd, m, i - 1, 2, 0) = GA(gadd, m, i - 1, 2, 0)
                    + sixth * T1(i - 1) * dtdz * (D1(bpcpamdq, m, i) - D1(bpcmamdq, m, i));
                GA(gadd, m, i - 1, 1, 0) = GA(gadd, m, i - 1, 1, 0)
                    + sixth * T1(i - 1) * dtdz * (D1(bmcpamdq, m, i) - D1(bmcmamdq, m, i));
                GA(gadd, m, i - 1, 2, 1) = GA(gadd, m, i - 1, 2, 1) - sixth * T1(i - 1) * dtdz * D1(bpcpamdq, m, i);
                GA(gadd, m, i - 1, 1, 1) = GA(gadd, m, i - 1, 1, 1) - sixth * T1(i - 1) * dtdz * D1(bmcpamdq, m, i);
                GA(gadd, m, i - 1, 2, -1) = GA(gadd, m, i - 1, 2, -1) + sixth * T1(i - 1) * dtdz * D1(bpcmamdq, m, i);
                GA(gadd, m, i - 1, 1, -1) = GA(gadd, m, i - 1, 1, -1) + sixth * T1(i - 1) * dtdz * D1(bmcmamdq, m, i);
            }
            if (m3 < 2) continue;
            GA(gadd, m, i, 2, 0) = GA(gadd, m, i, 2, 0) + T1(i) * D1(bpcqxxp, m, i);
            GA(gadd, m, i, 1, 0) = GA(gadd, m, i, 1, 0) + T1(i) * D1(bmcqxxp, m, i);
            GA(gadd, m, i - 1, 2, 0) = GA(gadd, m, i - 1, 2, 0) - T1(i - 1) * D1(bpcqxxm, m, i);
            GA(gadd, m, i - 1, 1, 0) = GA(gadd, m, i - 1, 1, 0) - T1(i - 1) * D1(bmcqxxm, m, i);
        }
    if (m4 == 2) {
        for (int i = 0; i <= mx + 2; i++)
            for (int m = 1; m <= meqn; m++) {
                D1(bpapdq, m, i) = D1(bpapdq, m, i) - 3.0 * D1(bpcqxxp, m, i);
                D1(bpamdq, m, i) = D1(bpamdq, m, i) + 3.0 * D1(bpcqxxm, m, i);
                D1(bmapdq, m, i) = D1(bmapdq, m, i) - 3.0 * D1(bmcqxxp, m, i);
                D1(bmamdq, m, i) = D1(bmamdq, m, i) + 3.0 * D1(bmcqxxm, m, i);
            }
    }
    if (m4 > 0) {
        orc_rptt3(p, ixyz, 3, 2, 2, maxm, meqn, mwaves, mbc, mx, q1d, q1d, bpapdq, bmcpapdq, bpcpapdq);
        orc_rptt3(p, ixyz, 3, 1, 2, maxm, meqn, mwaves, mbc, mx, q1d, q1d, bpamdq, bmcpamdq, bpcpamdq);
        orc_rptt3(p, ixyz, 3, 2, 1, maxm, meqn, mwaves, mbc, mx, q1d, q1d, bmapdq, bmcmapdq, bpcmapdq);
        orc_rptt3(p, ixyz, 3, 1, 1, maxm, meqn, mwaves, mbc, mx, q1d, q1d, bmamdq, bmcmamdq, bpcmamdq);
    }
    for (int i = 1; i <= mx + 1; i++)
        for (int m = 1; m <= meqn; m++) {
            GA(hadd, m, i - 1, 1, 0) = GA(hadd, m, i - 1, 1, 0) - 0.5 * T1(i - 1) * D1(cmamdq, m, i);
            GA(hadd, m, i - 1, 2, 0) = GA(hadd, m, i - 1, 2, 0) - 0.5 * T1(i - 1) * D1(cpamdq, m, i);
            GA(hadd, m, i, 1, 0) = GA(hadd, m, i, 1, 0) - 0.5 * T1(i) * D1(cmapdq, m, i);
            GA(hadd, m, i, 2, 0) = GA(hadd, m, i, 2, 0) - 0.5 * T1(i) * D1(cpapdq, m, i);
            if (m4 > 0) {
                GA(hadd, m, i, 2, 0) = GA(hadd, m, i, 2, 0)
                    + sixth * T1(i) * dtdy * (D1(bpcpapdq, m, i) - D1(bpcmapdq, m, i));
                GA(hadd, m, i, 1, 0) = GA(hadd, m, i, 1, 0)
                    + sixth * T1(i) * dtdy * (D1(bmcpapdq, m, i) - D1(bmcmapdq, m, i));
                GA(hadd, m, i, 2, 1) = GA(hadd, m, i, 2, 1) - sixth * T1(i) * dtdy * D1(bpcpapdq, m, i);
                GA(hadd, m, i, 1, 1) = GA(hadd, m, i, 1, 1) - sixth * T1(i) * dtdy * D1(bmcpapdq, m, i);
                GA(hadd, m, i, 2, -1) = GA(hadd, m, i, 2, -1) + sixth * T1(i) * dtdy * D1(bpcmapdq, m, i);
                GA(hadd, m, i, 1, -1) = GA(hadd, m, i, 1, -1) + sixth * T1(i) * dtdy * D1(bmcmapdq, m, i);
                GA(hadd, m, i - 1, 2, 0) = GA(hadd, m, i - 1, 2, 0)
                    + sixth * T1(i - 1) * dtdy * (D1(bpcpamdq, m, i) - D1(bpcmamdq, m, i));
                GA(hadd, m, i - 1, 1, 0) = GA(hadd, m, i - 1, 1, 0)
                    + sixth * T1(i - 1) * dtdy * (D1(bmcpamdq, m, i) - D1(bmcmamdq, m, i));
                GA(hadd, m, i - 1, 2, 1) = GA(hadd, m, i - 1, 2, 1) - sixth * T1(i - 1) * dtdy * D1(bpcpamdq, m, i);
                GA(hadd, m, i - 1, 1, 1) = GA(hadd, m, i - 1, 1, 1) - sixth * T1(i - 1) * dtdy * D1(bmcpamdq, m, i);
                GA(hadd, m, i - 1, 2, -1) = GA(hadd, m, i - 1, 2, -1) + sixth * T1(i - 1) * dtdy * D1(bpcmamdq, m, i);
                GA(hadd, m, i - 1, 1, -1) = GA(hadd, m, i - 1, 1, -1) + sixth * T1(i - 1) * dtdy * D1(bmcmamdq, m, i);
            }
            if (m3 < 2) continue;
            GA(hadd, m, i, 2, 0) = GA(hadd, m, i, 2, 0) + T1(i) * D1(cpcqxxp, m, i);
            GA(hadd, m, i, 1, 0) = GA(hadd, m, i, 1, 0) + T1(i) * D1(cmcqxxp, m, i);
            GA(hadd, m, i - 1, 2, 0) = GA(hadd, m, i - 1, 2, 0) - T1(i - 1) * D1(cpcqxxm, m, i);
            GA(hadd, m, i - 1, 1, 0) = GA(hadd, m, i - 1, 1, 0) - T1(i - 1) * D1(cmcqxxm, m, i);
        }
    free(buf);
#undef D1
#undef T1
}

/* ------------------------------------------------------------ step3 ----- */
/* 3-D arrays (meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc) */
#define A3(a, m, i, j, k) (a)[((m) - 1) + (size_t)meqn * (((i) + mbc - 1) + (size_t)mit * (((j) + mbc - 1) + (size_t)mjt * ((k) + mbc - 1)))]

void orc_step3(const orc_params* p, int mx, int my, int mz, const double* qold, double dx,
               double dy, double dz, double dt, double* fm, double* fp, double* gm, double* gp,
               double* hm, double* hp, double* cflgrid) {
    const int meqn = p->meqn, mwaves = p->mwaves, mbc = p->mbc;
    const int mit = mx + 2 * mbc, mjt = my + 2 * mbc, mkt = mz + 2 * mbc;
    int maxm = mx > my ? mx : my;
    if (mz > maxm) maxm = mz;
    const int n1 = maxm + 2 * mbc;
    const size_t ntot = (size_t)meqn * mit * mjt * mkt;
    double* q1d = calloc((size_t)meqn * n1, sizeof(double));
    double* dtdx1d = calloc(n1, sizeof(double));
    double* dtdy1d = calloc(n1, sizeof(double));
    double* dtdz1d = calloc(n1, sizeof(double));
    double* faddm = calloc((size_t)meqn * n1, sizeof(double));
    double* faddp = calloc((size_t)meqn * n1, sizeof(double));
    double* gadd = calloc((size_t)meqn * n1 * 6, sizeof(double));
    double* hadd = calloc((size_t)meqn * n1 * 6, sizeof(double));
    *cflgrid = 0.0;
    const double dtdx = dt / dx, dtdy = dt / dy, dtdz = dt / dz;
    for (size_t t = 0; t < ntot; t++) fm[t] = fp[t] = gm[t] = gp[t] = hm[t] = hp[t] = 0.0;
    for (int i = 0; i < n1; i++) { dtdx1d[i] = dtdx; dtdy1d[i] = dtdy; dtdz1d[i] = dtdz; }
    double cfl1d;
#define F1(a, m, i) (a)[((m) - 1) + (size_t)meqn * ((i) + mbc - 1)]
#define GG(a, m, i, s, o) (a)[((m) - 1) + (size_t)meqn * (((i) + mbc - 1) + (size_t)n1 * (((s) - 1) + 2 * ((o) + 1)))]
    /* x-sweeps (step3.f:140-195) */
    for (int k = 0; k <= mz + 1; k++)
        for (int j = 0; j <= my + 1; j++) {
            for (int i = 1 - mbc; i <= mx + mbc; i++)
                for (int m = 1; m <= meqn; m++) F1(q1d, m, i) = A3(qold, m, i, j, k);
            orc_flux3(p, 1, maxm, meqn, mwaves, mbc, mx, q1d, dtdx1d, dtdy, dtdz, faddm, faddp, gadd,
                      hadd, &cfl1d);
            *cflgrid = dmax1(*cflgrid, cfl1d);
            for (int i = 1; i <= mx + 1; i++)
                for (int m = 1; m <= meqn; m++) {
                    A3(fm, m, i, j, k) = A3(fm, m, i, j, k) + F1(faddm, m, i);
                    A3(fp, m, i, j, k) = A3(fp, m, i, j, k) + F1(faddp, m, i);
                    for (int o = -1; o <= 1; o++) {
                        A3(gm, m, i, j, k + o) = A3(gm, m, i, j, k + o) + GG(gadd, m, i, 1, o);
                        A3(gp, m, i, j, k + o) = A3(gp, m, i, j, k + o) + GG(gadd, m, i, 1, o);
                    }
                    for (int o = -1; o <= 1; o++) {
                        A3(gm, m, i, j + 1, k + o) = A3(gm, m, i, j + 1, k + o) + GG(gadd, m, i, 2, o);
                        A3(gp, m, i, j + 1, k + o) = A3(gp, m, i, j + 1, k + o) + GG(gadd, m, i, 2, o);
                    }
                    for (int o = -1; o <= 1; o++) {
                        A3(hm, m, i, j + o, k) = A3(hm, m, i, j + o, k) + GG(hadd, m, i, 1, o);
                        A3(hp, m, i, j + o, k) = A3(hp, m, i, j + o, k) + GG(hadd, m, i, 1, o);
                    }
                    for (int o = -1; o <= 1; o++) {
                        A3(hm, m, i, j + o, k + 1) = A3(hm, m, i, j + o, k + 1) + GG(hadd, m, i, 2, o);
                        A3(hp, m, i, j + o, k + 1) = A3(hp, m, i, j + o, k + 1) + GG(hadd, m, i, 2, o);
                    }
                }
        }
    /* y-sweeps (step3.f:197-259) */
    for (int k = 0; k <= mz + 1; k++)
        for (int i = 0; i <= mx + 1; i++) {
            for (int j = 1 - mbc; j <= my + mbc; j++)
                for (int m = 1; m <= meqn; m++) F1(q1d, m, j) = A3(qold, m, i, j, k);
            orc_flux3(p, 2, maxm, meqn, mwaves, mbc, my, q1d, dtdy1d, dtdz, dtdx, faddm, faddp, gadd,
                      hadd, &cfl1d);
            *cflgrid = dmax1(*cflgrid, cfl1d);
            for (int j = 1; j <= my + 1; j++)
                for (int m = 1; m <= meqn; m++) {
                    A3(gm, m, i, j, k) = A3(gm, m, i, j, k) + F1(faddm, m, j);
                    A3(gp, m, i, j, k) = A3(gp, m, i, j, k) + F1(faddp, m, j);
                    for (int o = -1; o <= 1; o++) {
                        A3(hm, m, i + o, j, k) = A3(hm, m, i + o, j, k) + GG(gadd, m, j, 1, o);
                        A3(hp, m, i + o, j, k) = A3(hp, m, i + o, j, k) + GG(gadd, m, j, 1, o);
                    }
                    for (int o = -1; o <= 1; o++) {
                        A3(hm, m, i + o, j, k + 1) = A3(hm, m, i + o, j, k + 1) + GG(gadd, m, j, 2, o);
                        A3(hp, m, i + o, j, k + 1) = A3(hp, m, i + o, j, k + 1) + GG(gadd, m, j, 2, o);
                    }
                    for (int o = -1; o <= 1; o++) {
                        A3(fm, m, i, j, k + o) = A3(fm, m, i, j, k + o) + GG(hadd, m, j, 1, o);
                        A3(fp, m, i, j, k + o) = A3(fp, m, i, j, k + o) + GG(hadd, m, j, 1, o);
                    }
                    for (int o = -1; o <= 1; o++) {
                        A3(fm, m, i + 1, j, k + o) = A3(fm, m, i + 1, j, k + o) + GG(hadd, m, j, 2, o);
                        A3(fp, m, i + 1, j, k + o) = A3(fp, m, i + 1, j, k + o) + GG(hadd, m, j, 2, o);
                    }
                }
        }
    /* z-sweeps (step3.f:261-321) */
    for (int j = 0; j <= my + 1; j++)
        for (int i = 0; i <= mx + 1; i++) {
            for (int k = 1 - mbc; k <= mz + mbc; k++)
                for (int m = 1; m <= meqn; m++) F1(q1d, m, k) = A3(qold, m, i, j, k);
            orc_flux3(p, 3, maxm, meqn, mwaves, mbc, mz, q1d, dtdz1d, dtdx, dtdy, faddm, faddp, gadd,
                      hadd, &cfl1d);
            *cflgrid = dmax1(*cflgrid, cfl1d);
            for (int k = 1; k <= mz + 1; k++)
                for (int m = 1; m <= meqn; m++) {
                    A3(hm, m, i, j, k) = A3(hm, m, i, j, k) + F1(faddm, m, k);
                    A3(hp, m, i, j, k) = A3(hp, m, i, j, k) + F1(faddp, m, k);
                    for (int o = -1; o <= 1; o++) {
                        A3(fm, m, i, j + o, k) = A3(fm, m, i, j + o, k) + GG(gadd, m, k, 1, o);
                        A3(fp, m, i, j + o, k) = A3(fp, m, i, j + o, k) + GG(gadd, m, k, 1, o);
                    }
                    for (int o = -1; o <= 1; o++) {
                        A3(fm, m, i + 1, j + o, k) = A3(fm, m, i + 1, j + o, k) + GG(gadd, m, k, 2, o);
                        A3(fp, m, i + 1, j + o, k) = A3(fp, m, i + 1, j + o, k) + GG(gadd, m, k, 2, o);
                    }
                    for (int o = -1; o <= 1; o++) {
                        A3(gm, m, i + o, j, k) = A3(gm, m, i + o, j, k) + GG(hadd, m, k, 1, o);
                        A3(gp, m, i + o, j, k) = A3(gp, m, i + o, j, k) + GG(hadd, m, k, 1, o);
                    }
                    for (int o = -1; o <= 1; o++) {
                        A3(gm, m, i + o, j + 1, k) = A3(gm, m, i + o, j + 1, k) + GG(hadd, m, k, 2, o);
                        A3(gp, m, i + o, j + 1, k) = A3(gp, m, i + o, j + 1, k) + GG(hadd, m, k, 2, o);
                    }
                }
        }
    free(q1d); free(dtdx1d); free(dtdy1d); free(dtdz1d);
    free(faddm); free(faddp); free(gadd); free(hadd);
#undef F1
#undef GG
}

/* stepgrid.f (3-D):147-178 and :205-212 (no capacity) */
void orc_stepgrid3(const orc_params* p, int mx, int my, int mz, double* q, double dx, double dy,
                   double dz, double dt, double* fm, double* fp, double* gm, double* gp, double* hm,
                   double* hp, double* cflgrid, double* dtnew) {
    const int meqn = p->meqn, mbc = p->mbc;
    const int mit = mx + 2 * mbc, mjt = my + 2 * mbc;
    orc_step3(p, mx, my, mz, q, dx, dy, dz, dt, fm, fp, gm, gp, hm, hp, cflgrid);
    const double dtdx = dt / dx, dtdy = dt / dy, dtdz = dt / dz;
    for (int k = 1; k <= mz; k++)
        for (int j = 1; j <= my; j++)
            for (int i = 1; i <= mx; i++)
                for (int m = 1; m <= meqn; m++)
                    A3(q, m, i, j, k) = A3(q, m, i, j, k)
                        - dtdx * (A3(fm, m, i + 1, j, k) - A3(fp, m, i, j, k))
                        - dtdy * (A3(gm, m, i, j + 1, k) - A3(gp, m, i, j, k))
                        - dtdz * (A3(hm, m, i, j, k + 1) - A3(hp, m, i, j, k));
    *dtnew = *cflgrid > 0.0 ? dt * p->cfl / *cflgrid : ORC_RINFINITY;
}

/* ------------------------------------------------------- flag2refine ----- */
/* src/3d/flag2refine.f:3-92 with the library's allowflag (allowflag.f: always
 * true) as spest.f:71-84 calls it: amrflags(1-mbc:mx+mbc, 1-mbc:my+mbc,
 * 1-mbc:mz+mbc) starts at goodpt everywhere; every interior cell gets
 * DONTFLAG, or DOFLAG when the largest centred difference of any component in
 * any direction exceeds tolsp.  Returns the number of DOFLAG cells. */
int orc_flag2refine3(int mx, int my, int mz, int mbc, int meqn, double tolsp, const double* q,
                     double dontflag, double doflag, double* amrflags) {
    const int mit = mx + 2 * mbc, mjt = my + 2 * mbc;
#define F3(i, j, k) amrflags[((i) + mbc - 1) + (size_t)mit * (((j) + mbc - 1) + (size_t)mjt * ((k) + mbc - 1))]
    int nflag = 0;
    for (int k = 1; k <= mz; k++)
        for (int j = 1; j <= my; j++)
            for (int i = 1; i <= mx; i++) {
                F3(i, j, k) = dontflag;
                double dq = 0.0;
                for (int m = 1; m <= meqn; m++) {
                    double dqi = fabs(A3(q, m, i + 1, j, k) - A3(q, m, i - 1, j, k));
                    double dqj = fabs(A3(q, m, i, j + 1, k) - A3(q, m, i, j - 1, k));
                    double dqk = fabs(A3(q, m, i, j, k + 1) - A3(q, m, i, j, k - 1));
                    dq = dmax1(dmax1(dmax1(dq, dqi), dqj), dqk);
                }
                if (dq > tolsp) {
                    F3(i, j, k) = doflag;
                    nflag++;
                }
            }
#undef F3
    return nflag;
}

/* conck.f (3-D):47-60: the running sum of q(1) over the interior of one
 * grid, in the reference's loop order (k, j, i), continued from acc (the
 * sum of the grids before it); conck.f:77 then scales by hx hy hz */
double orc_mass3(int mx, int my, int mz, int mbc, int meqn, const double* q, double acc) {
    const int mit = mx + 2 * mbc, mjt = my + 2 * mbc;
    double s = acc;
    for (int k = 1; k <= mz; k++)
        for (int j = 1; j <= my; j++)
            for (int i = 1; i <= mx; i++) s = s + A3(q, 1, i, j, k);
    return s;
}
