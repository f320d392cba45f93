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
#include "orc.h"

static orc_params g_plugin;

void orc_set_plugin(const orc_params* p) { g_plugin = *p; }

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
