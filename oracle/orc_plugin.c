/*
 * orc_plugin.c -- by-reference entry points for the oracle Riemann solvers
 * (TEST INFRASTRUCTURE ONLY).
 *
 * The reference's Fortran calls rpn2/rpt2 with every argument by reference
 * (flux2.f:125, qad.f:129) and reads physics constants from common /cparam/.
 * The Fortran harness in oracle/ref forwards its rpn2/rpt2 to these
 * functions; the constants come from the params installed with
 * orc_set_plugin(), the analogue of filling /cparam/ in setprob.
 */
#include <math.h>

#include "orc.h"

static orc_params g_plugin;

void orc_set_plugin(const orc_params* p) { g_plugin = *p; }

/* The physics of a run of the reference's own driver (oracle/driver/app.f90
 * calls this from its setprob): the solver and the /cparam/ constants, with
 * the acoustic impedance and sound speed derived as the library derives
 * them (cc = sqrt(K / rho), zz = rho cc). */
void orc_driver_physics(const int* solver, const double* ubar, const double* vbar,
                        const double* rho, const double* bulk, const double* gamma,
                        const int* efix) {
    orc_params p = {0};
    static const int nv[6] = {0, 1, 3, 4, 3, 1}, nw[6] = {0, 1, 2, 4, 3, 1};
    p.solver = *solver;
    p.meqn = (*solver >= 1 && *solver <= 5) ? nv[*solver] : 0;
    p.mwaves = (*solver >= 1 && *solver <= 5) ? nw[*solver] : 0;
    p.mbc = 2;
    p.ubar = *ubar;
    p.vbar = *vbar;
    p.rho = *rho;
    p.bulk = *bulk;
    p.cc = sqrt(*bulk / *rho);
    p.zz = *rho * p.cc;
    p.gamma = *gamma;
    p.efix = *efix;
    p.grav = 1.0;
    g_plugin = p;
}

void orc_rpn2_byref(const int* ixy, const int* maxm, const int* meqn,
                    const int* mwaves, const int* maux, const int* mbc,
                    const int* mx, const double* ql, const double* qr,
                    const double* auxl, const double* auxr, double* wave,
                    double* s, double* amdq, double* apdq) {
    orc_rpn2(&g_plugin, *ixy, *maxm, *meqn, *mwaves, *maux, *mbc, *mx, ql, qr,
             auxl, auxr, wave, s, amdq, apdq);
}

void orc_rpt2_byref(const int* ixy, const int* imp, const int* maxm,
                    const int* meqn, const int* mwaves, const int* maux,
                    const int* mbc, const int* mx, const double* ql,
                    const double* qr, const double* aux1, const double* aux2,
                    const double* aux3, const double* asdq, double* bmasdq,
                    double* bpasdq) {
    orc_rpt2(&g_plugin, *ixy, *imp, *maxm, *meqn, *mwaves, *maux, *mbc, *mx,
             ql, qr, aux1, aux2, aux3, asdq, bmasdq, bpasdq);
}

void orc_rpn3_byref(const int* ixyz, const int* maxm, const int* meqn, const int* mwaves,
                    const int* maux, const int* mbc, const int* mx, const double* ql,
                    const double* qr, const double* auxl, const double* auxr, double* wave,
                    double* s, double* amdq, double* apdq) {
    (void)maux; (void)auxl; (void)auxr;
    orc_rpn3(&g_plugin, *ixyz, *maxm, *meqn, *mwaves, *mbc, *mx, ql, qr, wave, s, amdq, apdq);
}

void orc_rpt3_byref(const int* ixyz, const int* icoor, const int* imp, const int* maxm,
                    const int* meqn, const int* mwaves, const int* maux, const int* mbc,
                    const int* mx, const double* ql, const double* qr, const double* aux1,
                    const double* aux2, const double* aux3, const double* asdq, double* bmasdq,
                    double* bpasdq) {
    (void)maux; (void)aux1; (void)aux2; (void)aux3;
    orc_rpt3(&g_plugin, *ixyz, *icoor, *imp, *maxm, *meqn, *mwaves, *mbc, *mx, ql, qr, asdq,
             bmasdq, bpasdq);
}

void orc_rptt3_byref(const int* ixyz, const int* icoor, const int* imp, const int* impt,
                     const int* maxm, const int* meqn, const int* mwaves, const int* maux,
                     const int* mbc, const int* mx, const double* ql, const double* qr,
                     const double* aux1, const double* aux2, const double* aux3,
                     const double* bsasdq, double* cmbsasdq, double* cpbsasdq) {
    (void)maux; (void)aux1; (void)aux2; (void)aux3;
    orc_rptt3(&g_plugin, *ixyz, *icoor, *imp, *impt, *maxm, *meqn, *mwaves, *mbc, *mx, ql, qr,
              bsasdq, cmbsasdq, cpbsasdq);
}
