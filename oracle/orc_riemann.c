/*
 * orc_riemann.c -- oracle Riemann solvers (TEST INFRASTRUCTURE ONLY).
 *
 * The reference calls rpn2/rpt2 through dummy-procedure arguments
 * (flux2.f:125-126, 202-204, 219-221) and by global name from qad.f:129.
 * The solvers themselves live in clawpack/riemann@48a46d3, which the
 * reference tree does not contain; what follows restates the published
 * algorithms (LeVeque 2002) with the argument meaning of that interface:
 *   interface i sits between cell i-1 and cell i; its left state is
 *   qr(:,i-1) and its right state is ql(:,i); AMRClaw passes the same slice
 *   for both (flux2.f:125), so ql == qr == q1d there.
 * Loops run over i = 2-mbc .. mx+mbc as in the library solvers.
 *
 * What pins them (the reference holds no solver source or fixture): the
 * advection and acoustics solvers by exact shift, slab and eigen identities
 * (tests/test_oracle_sweep.py); the Euler rpn2 / rpt2 by closed forms
 * (tests/test_oracle_euler_eigen.py): each wave is the part of the jump in
 * its eigenspace of the Roe matrix (complex-step Jacobian of the flux at the
 * Roe average), the shear and entropy waves have the textbook eigenvector
 * forms, A-dq / A+dq and B-asdq / B+asdq are the spectral splits, and the
 * Harten-Hyman fix is pinned by its two-speed identity, continuity at the
 * sonic points and reflection symmetry (1-wave branch vs 4-wave branch).
 */
#include <math.h>
#include "orc.h"

#define POS(i) ((i) + mbc - 1)          /* 0-based slot of Fortran index i */
#define Q(a, m, i) (a)[((m) - 1) + meqn * POS(i)]
#define W(m, mw, i) wave[((m) - 1) + meqn * (((mw) - 1) + mwaves * POS(i))]
#define S(mw, i) s[((mw) - 1) + mwaves * POS(i)]

static inline double dmin1(double a, double b) { return a < b ? a : b; }
static inline double dmax1(double a, double b) { return a > b ? a : b; }

/* ---- constant-velocity advection q_t + ubar q_x + vbar q_y = 0 -----------
 * LeVeque 2002 sec. 20.2: one wave, speed ubar (ixy=1) or vbar (ixy=2). */
static void rpn2_advection(const orc_params* p, int ixy, int meqn, int mwaves,
                           int mbc, int mx, const double* ql, const double* qr,
                           double* wave, double* s, double* amdq,
                           double* apdq) {
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        W(1, 1, i) = Q(ql, 1, i) - Q(qr, 1, i - 1);
        S(1, i) = (ixy == 1) ? p->ubar : p->vbar;
        Q(amdq, 1, i) = dmin1(S(1, i), 0.0) * W(1, 1, i);
        Q(apdq, 1, i) = dmax1(S(1, i), 0.0) * W(1, 1, i);
    }
}

static void rpt2_advection(const orc_params* p, int ixy, int meqn, int mbc,
                           int mx, const double* asdq, double* bmasdq,
                           double* bpasdq) {
    double stran = (ixy == 1) ? p->vbar : p->ubar;
    double stranm = dmin1(stran, 0.0);
    double stranp = dmax1(stran, 0.0);
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        Q(bmasdq, 1, i) = stranm * Q(asdq, 1, i);
        Q(bpasdq, 1, i) = stranp * Q(asdq, 1, i);
    }
}

/* ---- variable-coefficient advection q_t + (u q)_x + (v q)_y = 0 ----------
 * clawpack/riemann's rpn2_vc_advection / rpt2_vc_advection (the solvers of
 * the reference's tests/advection_2d_annulus, Makefile:31-32): the edge
 * velocities are aux fields, aux(1) the x-velocity at the left edge of a
 * cell and aux(2) the y-velocity at its bottom edge.  One wave, Δq, moving
 * at the velocity of the interface; each fluctuation goes downwind
 * (LeVeque 2002 sec. 9.5 for the normal solve, sec. 21.2 for the transverse
 * split with the velocities of the edges the fluctuation crosses). */
#define AUX(a, ma, i) (a)[((ma) - 1) + maux * POS(i)]
static void rpn2_vcadv(int ixy, int meqn, int mwaves, int maux, int mbc, int mx,
                       const double* ql, const double* qr, const double* auxl, double* wave,
                       double* s, double* amdq, double* apdq) {
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        W(1, 1, i) = Q(ql, 1, i) - Q(qr, 1, i - 1);
        S(1, i) = AUX(auxl, ixy, i);            /* the velocity at the edge i - 1/2 */
        Q(amdq, 1, i) = dmin1(AUX(auxl, ixy, i), 0.0) * W(1, 1, i);
        Q(apdq, 1, i) = dmax1(AUX(auxl, ixy, i), 0.0) * W(1, 1, i);
    }
}

static void rpt2_vcadv(int ixy, int imp, int meqn, int maux, int mbc, int mx,
                       const double* aux2, const double* aux3, const double* asdq,
                       double* bmasdq, double* bpasdq) {
    const int kv = 3 - ixy;                    /* the transverse velocity */
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        const int i1 = i - 2 + imp;            /* the cell the fluctuation entered */
        Q(bmasdq, 1, i) = dmin1(AUX(aux2, kv, i1), 0.0) * Q(asdq, 1, i);
        Q(bpasdq, 1, i) = dmax1(AUX(aux3, kv, i1), 0.0) * Q(asdq, 1, i);
    }
}
#undef AUX

/* ---- linear acoustics, q = (p, u, v), density rho, bulk modulus K -------
 * LeVeque 2002 sec. 3.10 / 21.2: waves r1 = (-Z,1,0) at -c, r2 = (Z,1,0)
 * at +c (normal velocity component mu), impedance Z = rho c. */
static void rpn2_acoustics(const orc_params* p, int ixy, int meqn, int mwaves,
                           int mbc, int mx, const double* ql, const double* qr,
                           double* wave, double* s, double* amdq,
                           double* apdq) {
    int mu = (ixy == 1) ? 2 : 3, mv = (ixy == 1) ? 3 : 2;
    double zz = p->zz, cc = p->cc;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        double d1 = Q(ql, 1, i) - Q(qr, 1, i - 1);
        double d2 = Q(ql, mu, i) - Q(qr, mu, i - 1);
        double a1 = (-d1 + zz * d2) / (2.0 * zz);
        double a2 = (d1 + zz * d2) / (2.0 * zz);
        W(1, 1, i) = -a1 * zz;
        W(mu, 1, i) = a1;
        W(mv, 1, i) = 0.0;
        S(1, i) = -cc;
        W(1, 2, i) = a2 * zz;
        W(mu, 2, i) = a2;
        W(mv, 2, i) = 0.0;
        S(2, i) = cc;
    }
    for (int m = 1; m <= meqn; m++)
        for (int i = 2 - mbc; i <= mx + mbc; i++) {
            Q(amdq, m, i) = S(1, i) * W(m, 1, i);
            Q(apdq, m, i) = S(2, i) * W(m, 2, i);
        }
}

static void rpt2_acoustics(const orc_params* p, int ixy, int meqn, int mbc,
                           int mx, const double* asdq, double* bmasdq,
                           double* bpasdq) {
    int mu = (ixy == 1) ? 2 : 3, mv = (ixy == 1) ? 3 : 2;
    double zz = p->zz, cc = p->cc;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        double a1 = (-Q(asdq, 1, i) + zz * Q(asdq, mv, i)) / (2.0 * zz);
        double a2 = (Q(asdq, 1, i) + zz * Q(asdq, mv, i)) / (2.0 * zz);
        Q(bmasdq, 1, i) = cc * a1 * zz;
        Q(bmasdq, mu, i) = 0.0;
        Q(bmasdq, mv, i) = -cc * a1;
        Q(bpasdq, 1, i) = cc * a2 * zz;
        Q(bpasdq, mu, i) = 0.0;
        Q(bpasdq, mv, i) = cc * a2;
    }
}

/* ---- Euler equations, q = (rho, rho u, rho v, E), Roe solver ------------
 * LeVeque 2002 sec. 15.3.2-15.3.3 (Roe averages, eigen-decomposition of the
 * jump into 4 waves: acoustic u-a, shear, entropy, acoustic u+a) and
 * sec. 15.3.5 (Harten-Hyman entropy fix for transonic rarefactions). */
typedef struct { double u, v, enth, a, g1a2, euv, u2v2; } roe_avg;

static roe_avg roe_average(double gamma1, int mu, int mv, const double* L,
                           const double* R) {
    roe_avg r;
    double rhsqrtl = sqrt(L[0]);
    double rhsqrtr = sqrt(R[0]);
    double pl = gamma1 * (L[3] - 0.5 * (L[1] * L[1] + L[2] * L[2]) / L[0]);
    double pr = gamma1 * (R[3] - 0.5 * (R[1] * R[1] + R[2] * R[2]) / R[0]);
    double rhsq2 = rhsqrtl + rhsqrtr;
    r.u = (L[mu - 1] / rhsqrtl + R[mu - 1] / rhsqrtr) / rhsq2;
    r.v = (L[mv - 1] / rhsqrtl + R[mv - 1] / rhsqrtr) / rhsq2;
    r.enth = ((L[3] + pl) / rhsqrtl + (R[3] + pr) / rhsqrtr) / rhsq2;
    r.u2v2 = r.u * r.u + r.v * r.v;
    double a2 = gamma1 * (r.enth - 0.5 * r.u2v2);
    r.a = sqrt(a2);
    r.g1a2 = gamma1 / a2;
    r.euv = r.enth - r.u2v2;
    return r;
}

static inline double euler_pressure(double gamma1, double rho, double mx_,
                                    double my_, double en) {
    return gamma1 * (en - 0.5 * (mx_ * mx_ + my_ * my_) / rho);
}

static void rpn2_euler(const orc_params* p, int ixy, int meqn, int mwaves,
                       int mbc, int mx, const double* ql, const double* qr,
                       double* wave, double* s, double* amdq, double* apdq) {
    int mu = (ixy == 1) ? 2 : 3, mv = (ixy == 1) ? 3 : 2;
    double gamma = p->gamma, gamma1 = p->gamma - 1.0;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        const double* L = &Q(qr, 1, i - 1);
        const double* R = &Q(ql, 1, i);
        roe_avg r = roe_average(gamma1, mu, mv, L, R);
        double u = r.u, v = r.v, enth = r.enth, a = r.a;
        double d1 = R[0] - L[0];
        double d2 = R[mu - 1] - L[mu - 1];
        double d3 = R[mv - 1] - L[mv - 1];
        double d4 = R[3] - L[3];
        double a3 = r.g1a2 * (r.euv * d1 + u * d2 + v * d3 - d4);
        double a2 = d3 - v * d1;
        double a4 = (d2 + (a - u) * d1 - a * a3) / (2.0 * a);
        double a1 = d1 - a3 - a4;
        /* acoustic 1-wave */
        W(1, 1, i) = a1;
        W(mu, 1, i) = a1 * (u - a);
        W(mv, 1, i) = a1 * v;
        W(4, 1, i) = a1 * (enth - u * a);
        S(1, i) = u - a;
        /* shear wave */
        W(1, 2, i) = 0.0;
        W(mu, 2, i) = 0.0;
        W(mv, 2, i) = a2;
        W(4, 2, i) = a2 * v;
        S(2, i) = u;
        /* entropy wave */
        W(1, 3, i) = a3;
        W(mu, 3, i) = a3 * u;
        W(mv, 3, i) = a3 * v;
        W(4, 3, i) = a3 * 0.5 * r.u2v2;
        S(3, i) = u;
        /* acoustic 4-wave */
        W(1, 4, i) = a4;
        W(mu, 4, i) = a4 * (u + a);
        W(mv, 4, i) = a4 * v;
        W(4, 4, i) = a4 * (enth + u * a);
        S(4, i) = u + a;

        double am[4] = {0.0, 0.0, 0.0, 0.0};
        if (!p->efix) {
            for (int mw = 1; mw <= 4; mw++)
                for (int m = 1; m <= 4; m++)
                    am[m - 1] = am[m - 1] + dmin1(S(mw, i), 0.0) * W(m, mw, i);
        } else {
            /* left-going part with Harten-Hyman fix (sec. 15.3.5) */
            double rhoim1 = L[0];
            double pim1 = euler_pressure(gamma1, L[0], L[1], L[2], L[3]);
            double cim1 = sqrt(gamma * pim1 / rhoim1);
            double s0 = L[mu - 1] / rhoim1 - cim1;   /* u-c left of 1-wave */
            if (!(s0 >= 0.0 && S(1, i) > 0.0)) {
                double rho1 = L[0] + W(1, 1, i);
                double rhou1 = L[mu - 1] + W(mu, 1, i);
                double rhov1 = L[mv - 1] + W(mv, 1, i);
                double en1 = L[3] + W(4, 1, i);
                double p1 = euler_pressure(gamma1, rho1, rhou1, rhov1, en1);
                double c1 = sqrt(gamma * p1 / rho1);
                double s1 = rhou1 / rho1 - c1;      /* u-c right of 1-wave */
                double sfract;
                if (s0 < 0.0 && s1 > 0.0)
                    sfract = s0 * (s1 - S(1, i)) / (s1 - s0);
                else if (S(1, i) < 0.0)
                    sfract = S(1, i);
                else
                    sfract = 0.0;
                for (int m = 1; m <= 4; m++) am[m - 1] = sfract * W(m, 1, i);
                /* shear and entropy waves travel at u */
                if (S(2, i) < 0.0) {
                    for (int m = 1; m <= 4; m++)
                        am[m - 1] = am[m - 1] + S(2, i) * W(m, 2, i)
                                    + S(3, i) * W(m, 3, i);
                    /* acoustic 4-wave */
                    double rhoi = R[0];
                    double pi_ = euler_pressure(gamma1, R[0], R[1], R[2], R[3]);
                    double ci = sqrt(gamma * pi_ / rhoi);
                    double s3 = R[mu - 1] / rhoi + ci; /* u+c right of 4-wave */
                    double rho2 = R[0] - W(1, 4, i);
                    double rhou2 = R[mu - 1] - W(mu, 4, i);
                    double rhov2 = R[mv - 1] - W(mv, 4, i);
                    double en2 = R[3] - W(4, 4, i);
                    double p2 = euler_pressure(gamma1, rho2, rhou2, rhov2, en2);
                    double c2 = sqrt(gamma * p2 / rho2);
                    double s2 = rhou2 / rho2 + c2;     /* u+c left of 4-wave */
                    double sf4 = 0.0;
                    int add4 = 1;
                    if (s2 < 0.0 && s3 > 0.0)
                        sf4 = s2 * (s3 - S(4, i)) / (s3 - s2);
                    else if (S(4, i) < 0.0)
                        sf4 = S(4, i);
                    else
                        add4 = 0;                  /* 4-wave right-going */
                    if (add4)
                        for (int m = 1; m <= 4; m++)
                            am[m - 1] = am[m - 1] + sf4 * W(m, 4, i);
                }
            }
        }
        /* A+dq = (sum_p s_p W_p) - A-dq */
        for (int m = 1; m <= 4; m++) {
            double df = 0.0;
            for (int mw = 1; mw <= 4; mw++) df = df + S(mw, i) * W(m, mw, i);
            Q(amdq, m, i) = am[m - 1];
            if (p->efix)
                Q(apdq, m, i) = df - am[m - 1];
            else {
                double ap = 0.0;
                for (int mw = 1; mw <= 4; mw++)
                    ap = ap + dmax1(S(mw, i), 0.0) * W(m, mw, i);
                Q(apdq, m, i) = ap;
            }
        }
    }
}

/* transverse solver: split asdq into the eigenvectors of the Jacobian in the
 * other direction at the same Roe average (LeVeque 2002 sec. 21.6, 21.7) */
static void rpt2_euler(const orc_params* p, int ixy, int meqn, int mbc,
                       int mx, const double* ql, const double* qr,
                       const double* asdq, double* bmasdq, double* bpasdq) {
    int mu = (ixy == 1) ? 2 : 3, mv = (ixy == 1) ? 3 : 2;
    double gamma1 = p->gamma - 1.0;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        const double* L = &Q(qr, 1, i - 1);
        const double* R = &Q(ql, 1, i);
        roe_avg r = roe_average(gamma1, mu, mv, L, R);
        double u = r.u, v = r.v, enth = r.enth, a = r.a;
        double A1 = Q(asdq, 1, i), Amu = Q(asdq, mu, i), Amv = Q(asdq, mv, i),
               A4 = Q(asdq, 4, i);
        double a3 = r.g1a2 * (r.euv * A1 + u * Amu + v * Amv - A4);
        double a2 = Amu - u * A1;
        double a4 = (Amv + (a - v) * A1 - a * a3) / (2.0 * a);
        double a1 = A1 - a3 - a4;
        double wb[4][4], sb[4];
        /* [wave][component 1, mu, mv, 4] stored as component index m-1 */
        wb[0][0] = a1; wb[0][mu - 1] = a1 * u; wb[0][mv - 1] = a1 * (v - a);
        wb[0][3] = a1 * (enth - v * a); sb[0] = v - a;
        wb[1][0] = a3; wb[1][mu - 1] = a3 * u; wb[1][mv - 1] = a3 * v;
        wb[1][3] = a3 * 0.5 * r.u2v2; sb[1] = v;
        wb[2][0] = 0.0; wb[2][mu - 1] = a2; wb[2][mv - 1] = 0.0;
        wb[2][3] = a2 * u; sb[2] = v;
        wb[3][0] = a4; wb[3][mu - 1] = a4 * u; wb[3][mv - 1] = a4 * (v + a);
        wb[3][3] = a4 * (enth + v * a); sb[3] = v + a;
        for (int m = 1; m <= 4; m++) {
            double bm = 0.0, bp = 0.0;
            for (int mw = 0; mw < 4; mw++) {
                bm = bm + dmin1(sb[mw], 0.0) * wb[mw][m - 1];
                bp = bp + dmax1(sb[mw], 0.0) * wb[mw][m - 1];
            }
            Q(bmasdq, m, i) = bm;
            Q(bpasdq, m, i) = bp;
        }
    }
}

/* ---- shallow water (h, hu, hv), Roe split without entropy fix ----------
 * LeVeque 2002 sec. 15.3.3 (Roe averages: u, v weighted by sqrt(h), gravity
 * wave speed from the mean depth) and sec. 21.6 (transverse split with the
 * same averages).  Not one of the reference's example physics: the parity
 * case of the library's solver plug-in contract. */
typedef struct { double u, v, c; } sw_avg;

static sw_avg sw_average(double grav, int mu, int mv, const double* L, const double* R) {
    sw_avg r;
    double hsl = sqrt(L[0]), hsr = sqrt(R[0]);
    double hsq = hsl + hsr;
    r.u = (L[mu - 1] / hsl + R[mu - 1] / hsr) / hsq;
    r.v = (L[mv - 1] / hsl + R[mv - 1] / hsr) / hsq;
    r.c = sqrt(grav * 0.5 * (L[0] + R[0]));
    return r;
}

static void rpn2_shallow(const orc_params* p, int ixy, int meqn, int mwaves, int mbc, int mx,
                         const double* ql, const double* qr, double* wave, double* s,
                         double* amdq, double* apdq) {
    int mu = (ixy == 1) ? 2 : 3, mv = (ixy == 1) ? 3 : 2;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        const double* L = &Q(qr, 1, i - 1);
        const double* R = &Q(ql, 1, i);
        sw_avg r = sw_average(p->grav, mu, mv, L, R);
        double d1 = R[0] - L[0], d2 = R[mu - 1] - L[mu - 1], d3 = R[mv - 1] - L[mv - 1];
        double a1 = ((r.u + r.c) * d1 - d2) / (2.0 * r.c);
        double a2 = d3 - r.v * d1;
        double a3 = (d2 - (r.u - r.c) * d1) / (2.0 * r.c);
        W(1, 1, i) = a1; W(mu, 1, i) = a1 * (r.u - r.c); W(mv, 1, i) = a1 * r.v;
        S(1, i) = r.u - r.c;
        W(1, 2, i) = 0.0; W(mu, 2, i) = 0.0; W(mv, 2, i) = a2;
        S(2, i) = r.u;
        W(1, 3, i) = a3; W(mu, 3, i) = a3 * (r.u + r.c); W(mv, 3, i) = a3 * r.v;
        S(3, i) = r.u + r.c;
        for (int m = 1; m <= 3; m++) {
            double am = 0.0, ap = 0.0;
            for (int mw = 1; mw <= 3; mw++) {
                am = am + dmin1(S(mw, i), 0.0) * W(m, mw, i);
                ap = ap + dmax1(S(mw, i), 0.0) * W(m, mw, i);
            }
            Q(amdq, m, i) = am;
            Q(apdq, m, i) = ap;
        }
    }
    (void)meqn; (void)mwaves;
}

static void rpt2_shallow(const orc_params* p, int ixy, int meqn, int mbc, int mx,
                         const double* ql, const double* qr, const double* asdq,
                         double* bmasdq, double* bpasdq) {
    int mu = (ixy == 1) ? 2 : 3, mv = (ixy == 1) ? 3 : 2;
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        const double* L = &Q(qr, 1, i - 1);
        const double* R = &Q(ql, 1, i);
        sw_avg r = sw_average(p->grav, mu, mv, L, R);
        double A1 = Q(asdq, 1, i), Amu = Q(asdq, mu, i), Amv = Q(asdq, mv, i);
        double a1 = ((r.v + r.c) * A1 - Amv) / (2.0 * r.c);
        double a2 = Amu - r.u * A1;
        double a3 = (Amv - (r.v - r.c) * A1) / (2.0 * r.c);
        double wb[3][3], sb[3];
        wb[0][0] = a1; wb[0][mu - 1] = a1 * r.u; wb[0][mv - 1] = a1 * (r.v - r.c); sb[0] = r.v - r.c;
        wb[1][0] = 0.0; wb[1][mu - 1] = a2; wb[1][mv - 1] = 0.0; sb[1] = r.v;
        wb[2][0] = a3; wb[2][mu - 1] = a3 * r.u; wb[2][mv - 1] = a3 * (r.v + r.c); sb[2] = r.v + r.c;
        for (int m = 1; m <= 3; m++) {
            double bm = 0.0, bp = 0.0;
            for (int mw = 0; mw < 3; mw++) {
                bm = bm + dmin1(sb[mw], 0.0) * wb[mw][m - 1];
                bp = bp + dmax1(sb[mw], 0.0) * wb[mw][m - 1];
            }
            Q(bmasdq, m, i) = bm;
            Q(bpasdq, m, i) = bp;
        }
    }
}

void orc_rpn2(const orc_params* p, int ixy, int maxm, int meqn, int mwaves,
              int maux, int mbc, int mx, const double* ql, const double* qr,
              const double* auxl, const double* auxr, double* wave, double* s,
              double* amdq, double* apdq) {
    (void)maxm; (void)auxr;
    switch (p->solver) {
    case ORC_ADVECTION:
        rpn2_advection(p, ixy, meqn, mwaves, mbc, mx, ql, qr, wave, s, amdq,
                       apdq);
        break;
    case ORC_ACOUSTICS:
        rpn2_acoustics(p, ixy, meqn, mwaves, mbc, mx, ql, qr, wave, s, amdq,
                       apdq);
        break;
    case ORC_EULER:
        rpn2_euler(p, ixy, meqn, mwaves, mbc, mx, ql, qr, wave, s, amdq,
                   apdq);
        break;
    case ORC_SHALLOW:
        rpn2_shallow(p, ixy, meqn, mwaves, mbc, mx, ql, qr, wave, s, amdq, apdq);
        break;
    case ORC_VCADV:
        rpn2_vcadv(ixy, meqn, mwaves, maux, mbc, mx, ql, qr, auxl, wave, s, amdq, apdq);
        break;
    }
}

void orc_rpt2(const orc_params* p, int ixy, int imp, int maxm, int meqn,
              int mwaves, int maux, int mbc, int mx, const double* ql,
              const double* qr, const double* aux1, const double* aux2,
              const double* aux3, const double* asdq, double* bmasdq,
              double* bpasdq) {
    (void)maxm; (void)mwaves; (void)aux1;
    switch (p->solver) {
    case ORC_ADVECTION:
        rpt2_advection(p, ixy, meqn, mbc, mx, asdq, bmasdq, bpasdq);
        break;
    case ORC_ACOUSTICS:
        rpt2_acoustics(p, ixy, meqn, mbc, mx, asdq, bmasdq, bpasdq);
        break;
    case ORC_EULER:
        rpt2_euler(p, ixy, meqn, mbc, mx, ql, qr, asdq, bmasdq, bpasdq);
        break;
    case ORC_SHALLOW:
        rpt2_shallow(p, ixy, meqn, mbc, mx, ql, qr, asdq, bmasdq, bpasdq);
        break;
    case ORC_VCADV:
        rpt2_vcadv(ixy, imp, meqn, maux, mbc, mx, aux2, aux3, asdq, bmasdq, bpasdq);
        break;
    }
}

void orc_flux(const orc_params* p, int ixy, const double* q, double* f) {
    switch (p->solver) {
    case ORC_ADVECTION:
        f[0] = ((ixy == 1) ? p->ubar : p->vbar) * q[0];
        break;
    case ORC_ACOUSTICS: {
        /* p_t + K (u_x + v_y) = 0, u_t + p_x / rho = 0, v_t + p_y / rho = 0 */
        int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
        f[0] = p->bulk * q[mu];
        f[mu] = q[0] / p->rho;
        f[mv] = 0.0;
        break;
    }
    case ORC_EULER: {
        int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
        double g1 = p->gamma - 1.0;
        double pr = g1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
        double un = q[mu] / q[0];
        f[0] = q[mu];
        f[mu] = q[mu] * un + pr;
        f[mv] = q[mv] * un;
        f[3] = (q[3] + pr) * un;
        break;
    }
    case ORC_SHALLOW: {
        int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
        double un = q[mu] / q[0];
        f[0] = q[mu];
        f[mu] = q[mu] * un + 0.5 * p->grav * q[0] * q[0];
        f[mv] = q[mv] * un;
        break;
    }
    }
}
