/*
 * orc.h -- CPU oracle for the AMRClaw per-level patch update.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (amrclaw_amd/) may include,
 * link or call this code: it is the checker that tests/, the smoke entry point
 * and the CPU-baseline leg of bench.py compare the HIP path against.
 *
 * The oracle is a plain-C restatement of the reference's FP64 Fortran, in the
 * reference's own evaluation order and with the reference's array layout
 * (AoS, component fastest: q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc)).  Every function
 * cites the reference file:line it follows.  Build with -ffp-contract=off so
 * that no multiply-add is fused (the reference is built the same way).
 *
 * Riemann solvers: the reference links rpn2/rpt2 from clawpack/riemann
 * (commit 48a46d3, see tests/acoustics_2d_radial/regression_data/
 * claw_git_status.txt), which is not part of the reference tree.  The solvers
 * here restate the published algorithms (LeVeque, "Finite Volume Methods for
 * Hyperbolic Problems", 2002: advection ch. 20, acoustics ch. 3/21, Euler Roe
 * solver with Harten-Hyman entropy fix sec. 15.3) behind the reference's own
 * rpn2/rpt2 calling convention (flux2.f:125,202,219; rpn2from1d.f:5-17).
 */
#ifndef AMR_ORACLE_H
#define AMR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* solver ids */
enum { ORC_ADVECTION = 1, ORC_ACOUSTICS = 2, ORC_EULER = 3,
       /* a physics the library does not build in: the parity case of its
        * solver plug-in contract */
       ORC_SHALLOW = 4,
       /* variable-coefficient advection with edge velocities in aux(1:2)
        * (the annulus test's rpn2_vc_advection / rpt2_vc_advection) */
       ORC_VCADV = 5 };

/* Method parameters: the fields of amr_module.f90 that the hot path reads
 * (method(7), mthlim(:), mwaves, mcapa, use_fwaves, cfl, cflv1) plus the
 * physics parameters the reference passes through common /cparam/. */
typedef struct {
    int solver;        /* ORC_* */
    int meqn, mwaves, maux, mbc;
    int method[8];     /* 1-based like the Fortran: method[1..7] */
    int mthlim[10];    /* 0-based: mthlim[mw-1] */
    int mcapa;         /* 0 or the aux index (1-based) of the capacity */
    int use_fwaves;
    double cfl;        /* cfl_desired (amr_module cfl)   */
    double cflv1;      /* cfl_max     (amr_module cflv1) */
    /* /cparam/ */
    double ubar, vbar;              /* advection            */
    double rho, bulk, cc, zz;       /* acoustics            */
    double gamma;                   /* Euler                */
    int efix;                       /* Euler entropy fix    */
    double wbar;                    /* 3-D advection, z velocity */
    double grav;                    /* shallow water: gravity    */
    /* amr_module auxtype(1:maux): 0 "center", 1 "xleft", 2 "yleft",
     * 3 "capacity" (qad.f:81-91 reads the fine grid's face values for
     * xleft / yleft) */
    int auxtype[8];
} orc_params;

/* ---- Riemann solvers, Fortran calling convention over a 1-D slice -------
 * All arrays are Fortran column-major with first index = component and
 * second index i in [1-mbc, maxm+mbc]; pointers address element (1, 1-mbc).
 * Interface i uses qr(:,i-1) (left state) and ql(:,i) (right state). */
void orc_rpn2(const orc_params* p, int ixy, int maxm, int meqn, int mwaves,
              int maux, int mbc, int mx, const double* ql, const double* qr,
              const double* auxl, const double* auxr, double* wave, double* s,
              double* amdq, double* apdq);
void orc_rpt2(const orc_params* p, int ixy, int imp, int maxm, int meqn,
              int mwaves, int maux, int mbc, int mx, const double* ql,
              const double* qr, const double* aux1, const double* aux2,
              const double* aux3, const double* asdq, double* bmasdq,
              double* bpasdq);

/* Physical flux of one state in direction ixy (used only by closed-form
 * checks of the Roe property; the reference has no such routine). */
void orc_flux(const orc_params* p, int ixy, const double* q, double* f);

/* ---- sweep ---------------------------------------------------------------
 * limiter   : inlinelimiter.f:4-108
 * flux2     : flux2.f:25-237
 * step2     : step2.f90:7-169
 * stepgrid  : stepgrid.f:24-234 (update, cfl, dtnew; b4step2/src2 are the
 *             library no-ops)                                              */
void orc_limiter(int maxm, int meqn, int mwaves, int mbc, int mx, double* wave,
                 const double* s, const int* mthlim);
void orc_flux2(const orc_params* p, int ixy, int maxm, int mx,
               const double* q1d, const double* dtdx1d, const double* aux1,
               const double* aux2, const double* aux3, double* faddm,
               double* faddp, double* gaddm, double* gaddp, double* cfl1d,
               double* wave, double* s, double* amdq, double* apdq,
               double* cqxx, double* bmasdq, double* bpasdq);
/* qold, aux, fm, fp, gm, gp: (meqn|maux, mx+2mbc, my+2mbc) */
void orc_step2(const orc_params* p, int mx, int my, const double* qold,
               const double* aux, double dx, double dy, double dt,
               double* cflgrid, double* fm, double* fp, double* gm,
               double* gp);
/* q updated in place (interior only); fluxes returned; returns cflgrid and
 * dtnew like stepgrid.f:206-212 (rinfinity = 10.e32 when cflgrid == 0). */
void orc_stepgrid(const orc_params* p, int mx, int my, double* q,
                  const double* aux, double dx, double dy, double dt,
                  double* fm, double* fp, double* gm, double* gp,
                  double* cflgrid, double* dtnew);

/* ---- hierarchy description (the parts of amr_module the hot path reads) --
 * Levels are numbered 1..nlev like the reference.  Grid k of level l has the
 * global index box boxes[l][4k..4k+3] = (ilo, ihi, jlo, jhi) and AoS FP64
 * arrays (nvar, mx+2ng, my+2ng) at qnew[l] + qoff[l][k] (store1) and
 * qold[l] + qoff[l][k] (store2).  All grids of a level share the time
 * time[l] (rnode(timemult)) and step possk[l]; hx/hy = hxposs/hyposs;
 * iregsz/jregsz = level size in cells; intratx/intraty[l] = ratio from
 * level l to level l+1. */
#define ORC_MAXLV 12
typedef struct {
    int nlev, mxnest;
    int ngrids[ORC_MAXLV];
    const int* boxes[ORC_MAXLV];
    double* qnew[ORC_MAXLV];
    double* qold[ORC_MAXLV];
    const int64_t* qoff[ORC_MAXLV];
    double hx[ORC_MAXLV], hy[ORC_MAXLV], possk[ORC_MAXLV], time[ORC_MAXLV];
    int iregsz[ORC_MAXLV], jregsz[ORC_MAXLV];
    int intratx[ORC_MAXLV], intraty[ORC_MAXLV];
    double xlower, xupper, ylower, yupper;
    int xperdom, yperdom, spheredom;
    int mthbc[4];
    /* aux arrays (maux > 0): grid k of level l has (maux, mx+2ng, my+2ng)
     * at aux[l] + auxoff[l][k] (alloc(node(storeaux))); NULL: none */
    double* aux[ORC_MAXLV];
    const int64_t* auxoff[ORC_MAXLV];
} orc_hier;

/* ghost filling, bound.f90:51-158 (with filrecur, intfil, trimbd,
 * prefilrecur, bc2amr).  Fills the ghost cells of grid `grid` of `level`
 * in valbig (the grid's AoS array) at time `time`.  Returns 0, or a nonzero
 * code where the reference would stop (errors listed in orc_amr.c). */
int orc_bound(const orc_params* p, const orc_hier* h, int level, int grid,
              double time, double* valbig);
/* bc2amr.f90:84-296 on an enlarged patch val(meqn, nrow, ncol) */
int orc_bc2amr(const orc_params* p, const orc_hier* h, double* val, int nrow,
               int ncol, double hx, double hy, double xlo_patch,
               double xhi_patch, double ylo_patch, double yhi_patch);

/* ---- coarse-fine flux registers (prepf.f, prepc.f, setuse.f) ------------
 * Fine side, levels 2..nlev: grid k's block at reg[l] + roff[l][k] holds
 * svdflx(nvar, lenbc) then qc1d(nvar, lenbc), lenbc = 2(nx/r + ny/r)
 * (prepf.f:35-48).  Coarse side, levels 1..nlev-1: grid k's listbc entries
 * (i, j, side, kid, lkid) are listbc[l][5e..5e+4] for e in
 * [lst_start[l][k], lst_start[l][k] + lst_count[l][k]) (setuse.f:61-213);
 * kid is the 0-based grid index on level l+1, i/j local 1-based with ghosts. */
typedef struct {
    int nlev;
    int* lenbc[ORC_MAXLV];
    int64_t* roff[ORC_MAXLV];
    double* reg[ORC_MAXLV];
    int64_t nreg[ORC_MAXLV];
    int* lst_start[ORC_MAXLV];
    int* lst_count[ORC_MAXLV];
    int* listbc[ORC_MAXLV];
    int nlist[ORC_MAXLV];
} orc_regs;

int orc_prep_regs(const orc_params* p, const orc_hier* h, orc_regs* r);
void orc_free_regs(orc_regs* r);
/* saveqc.f:13-109 for the grids of fine level `level` */
int orc_saveqc(const orc_params* p, const orc_hier* h, orc_regs* r, int level);
/* filval.f90:3-187 (naux = 0) for a new grid of `level` with box (ilo, ihi,
 * jlo, jhi); the grids of `level` in h are the old ones */
int orc_filval(const orc_params* p, const orc_hier* h, int level, const int* box, double time,
               double* val);
/* init_bndryList.f:50-107: grid k's neighbour list in its linked-list order */
int orc_bndry_list(const orc_params* p, const orc_hier* h, int level, int k, int* out);
/* qad.f:9-360 for grid `grid` of fine level `level` (uses its new array) */
int orc_qad(const orc_params* p, const orc_hier* h, orc_regs* r, int level, int grid);
/* advanc.f:10-142 with par_advanc :170-292 (no gauges): bound every grid,
 * saveqc for the next level, then per grid new->old copy (level < mxnest),
 * qad, stepgrid, fluxsv, fluxad, time += possk.  Returns dtlevnew and the
 * level's max Courant number. */
int orc_advanc(const orc_params* p, orc_hier* h, orc_regs* r, int level,
               double* dtlevnew, double* cfl_level);
/* update.f:16-189 with upbnd.f:18-127 */
int orc_update(const orc_params* p, const orc_hier* h, orc_regs* r, int level);
/* conck.f:6-73: hx*hy*sum of q(1) over the interiors of the level's grids */
double orc_conck(const orc_params* p, const orc_hier* h, int level);

int orc_bound_ng(const orc_params* p, const orc_hier* h, int level, int grid,
                 double time, double* valbig, int ng);

/* 3-D sweep (orc_sweep3.c): arrays (meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc,
 * 1-mbc:mz+mbc), component fastest */
void orc_rpn3(const orc_params* p, int ixyz, int maxm, int meqn, int mwaves, int mbc, int mx,
              const double* ql, const double* qr, double* wave, double* s, double* amdq,
              double* apdq);
void orc_rpt3(const orc_params* p, int ixyz, int icoor, int imp, int maxm, int meqn, int mwaves,
              int mbc, int mx, const double* ql, const double* qr, const double* asdq,
              double* bmasdq, double* bpasdq);
void orc_rptt3(const orc_params* p, int ixyz, int icoor, int imp, int impt, int maxm, int meqn,
               int mwaves, int mbc, int mx, const double* ql, const double* qr,
               const double* bsasdq, double* cmbsasdq, double* cpbsasdq);
void orc_flux3(const orc_params* p, int ixyz, int maxm, int meqn, int mwaves, int mbc, int mx,
               const double* q1d, const double* dtdx1d, double dtdy, double dtdz, double* faddm,
               double* faddp, double* gadd, double* hadd, double* cfl1d);
void orc_step3(const orc_params* p, int mx, int my, int mz, const double* qold, double dx,
               double dy, double dz, double dt, double* fm, double* fp, double* gm, double* gp,
               double* hm, double* hp, double* cflgrid);
void orc_stepgrid3(const orc_params* p, int mx, int my, int mz, double* q, double dx, double dy,
                   double dz, double dt, double* fm, double* fp, double* gm, double* gp, double* hm,
                   double* hp, double* cflgrid, double* dtnew);
/* 3-D flagging and mass (orc_sweep3.c): flag2refine.f, conck.f */
int orc_flag2refine3(int mx, int my, int mz, int mbc, int meqn, double tolsp, const double* q,
                     double dontflag, double doflag, double* amrflags);
double orc_mass3(int mx, int my, int mz, int mbc, int meqn, const double* q, double acc);

/* ---- 3-D level update around the step (orc_amr3.c, src/3d) ---------------
 * As orc_hier with a third direction: boxes[l][6k..6k+5] = (ilo, ihi, jlo,
 * jhi, klo, khi); AoS arrays (nvar, mx+2ng, my+2ng, mz+2ng); intratz/kregsz
 * and hz; mthbc(1:6) = x-lo, x-hi, y-lo, y-hi, z-lo, z-hi; periodic in a
 * direction when its flag is set (all three or none, see orc_amr3.c). */
typedef struct {
    int nlev, mxnest;
    int ngrids[ORC_MAXLV];
    const int* boxes[ORC_MAXLV];
    double* qnew[ORC_MAXLV];
    double* qold[ORC_MAXLV];
    const int64_t* qoff[ORC_MAXLV];
    double hx[ORC_MAXLV], hy[ORC_MAXLV], hz[ORC_MAXLV], possk[ORC_MAXLV], time[ORC_MAXLV];
    int iregsz[ORC_MAXLV], jregsz[ORC_MAXLV], kregsz[ORC_MAXLV];
    int intratx[ORC_MAXLV], intraty[ORC_MAXLV], intratz[ORC_MAXLV];
    double xlower, xupper, ylower, yupper, zlower, zupper;
    int xperdom, yperdom, zperdom;
    int mthbc[6];
} orc_hier3;

/* registers of a 3-D hierarchy: lenbc = 2 (ik jk + jk kk + kk ik)
 * (prepf.f); listbc entries (i, j, k, side, kid, lkid) (setuse.f) */
typedef struct {
    int nlev;
    int* lenbc[ORC_MAXLV];
    int64_t* roff[ORC_MAXLV];
    double* reg[ORC_MAXLV];
    int64_t nreg[ORC_MAXLV];
    int* lst_start[ORC_MAXLV];
    int* lst_count[ORC_MAXLV];
    int* listbc[ORC_MAXLV];
    int nlist[ORC_MAXLV];
} orc_regs3;

int orc_bc3amr(const orc_params* p, const orc_hier3* h, double* val, int nrow, int ncol,
               int nfil, double hx, double hy, double hz, double xlo_patch, double xhi_patch,
               double ylo_patch, double yhi_patch, double zlo_patch, double zhi_patch);
int orc_bound3(const orc_params* p, const orc_hier3* h, int level, int grid, double time,
               double* valbig);
int orc_prep_regs3(const orc_params* p, const orc_hier3* h, orc_regs3* r);
void orc_free_regs3(orc_regs3* r);
int orc_saveqc3(const orc_params* p, const orc_hier3* h, orc_regs3* r, int level);
int orc_qad3(const orc_params* p, const orc_hier3* h, orc_regs3* r, int level, int grid);
int orc_advanc3(const orc_params* p, orc_hier3* h, orc_regs3* r, int level, double* dtlevnew,
                double* cfl_level);
int orc_update3(const orc_params* p, const orc_hier3* h, orc_regs3* r, int level);
double orc_conck3(const orc_params* p, const orc_hier3* h, int level);
int orc_bound3_ng(const orc_params* p, const orc_hier3* h, int level, int grid, double time,
                  double* valbig, int ng);
/* spest.f / errest.f + errf1.f (3-D) for one grid (orc_amr3.c) */
int orc_spest3(const orc_params* p, const orc_hier3* h, int level, int grid, double tolsp,
               double* big, double* flags, int* nflag);
int orc_errest3(const orc_params* p, const orc_hier3* h, int level, int grid, double tol,
                int iorder, double* big, double* rflags);

/* cell flagging (orc_flag.c) */
#define ORC_UNSET (-1.0)
#define ORC_DONTFLAG 0.0
#define ORC_DOFLAG 2.0
int orc_flagregions2(int mx, int my, int mbuff, double xlower, double ylower, double dx,
                     double dy, int level, double t, const double* regions, int nreg,
                     double* amrflags);
void orc_flag2refine2(int mx, int my, int mbc, int mbuff, int meqn, double tolsp, const double* q,
                      double* amrflags);
int orc_flag_level(const orc_params* p, const orc_hier* h, int level, double tolsp,
                   const double* regions, int nreg, int use_old, int mbuff, double* flags,
                   const long* foff, int flag_gradient, int flag_richardson, double tol,
                   int iorder);
int orc_errest(const orc_params* p, const orc_hier* h, int level, int k, const double* big,
               double tol, int iorder, int mbuff, double* flags);

/* amr_module.f90 declares these real(kind=8) parameters from default-real
 * (single precision) literals 10.e32 and 10.e33, so their values are the
 * float roundings widened to double. */
#define ORC_RINFINITY ((double)10.e32f)
#define ORC_NEEDS_TO_BE_SET ((double)10.e33f)

#ifdef __cplusplus
}
#endif
#endif
