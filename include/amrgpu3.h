/*
 * amrgpu3.h -- C ABI of the 3-D patch step (libamrgpu.so).
 *
 * Drop-in for the per-grid call of the reference's 3-D advanc
 * (src/3d/advanc.f calls stepgrid for each grid of a level after bound):
 * the batch below holds every grid of a level on the device and advances all
 * of them with one launch per sweep direction, step3.f:3-486 + flux3.f:4-633
 * + stepgrid.f (3-D):79-230 for each grid.  Arrays cross the boundary in the
 * reference's layout: FP64, q(nvar, 1-nghost:mx+nghost, 1-nghost:my+nghost,
 * 1-nghost:mz+nghost), component fastest.  Ghost cells must already be
 * filled (the 3-D ghost fill stays with the caller).
 *
 * Every function returns 0 on success and a nonzero code (the AMRGPU_ERR_*
 * codes of amrgpu.h) on failure; the message is available from
 * amrgpu3_last_error().  Not thread-safe.
 */
#ifndef AMRGPU3_H
#define AMRGPU3_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run parameters: the amr_module fields the 3-D step reads (claw.data,
 * amr.data) and the solver constants of common /cparam/. */
typedef struct amrgpu3_config {
    int solver;          /* AMRGPU_ADVECTION (1 eqn, 1 wave) or AMRGPU_EULER
                            (5 eqns, 3 waves)                                  */
    int nvar;            /* num_eqn                                           */
    int mwaves;          /* num_waves                                         */
    int naux;            /* must be 0: no solver here reads aux               */
    int nghost;          /* must be 2                                         */
    int method[7];       /* method(1:7): [1] = order, [2] = 10 m3 + m4
                            (flux3.f:205-211; m4 = 2 needs m3 = 2)            */
    int mthlim[8];       /* limiter per wave (inlinelimiter.f)                */
    int mcapa;           /* must be 0                                         */
    int use_fwaves;
    double cfl;          /* cfl_desired                                       */
    double cflv1;        /* cfl_max                                           */
    double ubar, vbar, wbar;   /* advection velocities                       */
    double gamma;        /* Euler                                             */
    int efix;            /* Euler entropy fix                                 */
    int precision;       /* AMRGPU_FP64 or AMRGPU_FP32 (bulk arithmetic)     */
} amrgpu3_config;

/* Initialise (or re-initialise) the 3-D step on the current HIP device. */
int amrgpu3_init(const amrgpu3_config* cfg);

/* Release the 3-D device state. */
int amrgpu3_finalize(void);

/* Copy the last error message of an amrgpu3_* call. */
int amrgpu3_last_error(char* buf, int len);

/* Define the batch: npatch grids of mx[p] x my[p] x mz[p] interior cells
 * (the grids of a level, node(ndilo..ndkhi)); allocates their solution and
 * flux arrays on the device.  plane_block: planes per workgroup of a sweep
 * (0: default 4); a smaller block gives more workgroups, a larger one fewer
 * recomputed planes. */
int amrgpu3_set_patches(int npatch, const int* mx, const int* my, const int* mz,
                        int plane_block);

/* Copy one grid's q (with ghost cells) to / from the device. */
int amrgpu3_put_patch(int patch, const double* q);
int amrgpu3_get_patch(int patch, double* q);

/* One step of every grid of the batch: stepgrid.f (3-D) for each, in place.
 * cflgrid[npatch] receives each grid's Courant number (may be NULL). */
int amrgpu3_step_patches(double dx, double dy, double dz, double dt, double* cflgrid);

/* How amrgpu3_step_patches runs an FP32 batch.  0: the x-, y- and z-sweep
 * kernels accumulate the six flux arrays in HBM in step3.f's order, then the
 * update (the flux arrays stay readable through amrgpu3_get_fluxes).  1 (the
 * FP32 default): one fused kernel per tile of at most 13 cells per
 * direction, every face kept on chip and each face value added straight to
 * the two cells it separates; the same interface arithmetic, summed in
 * another order (agrees with mode 0 to FP32 rounding); no flux arrays.  FP64
 * runs mode 0 only (bit for bit the reference). No reference counterpart. */
int amrgpu3_set_step_mode(int mode);

/* The six flux arrays of the last step of one grid, in stepgrid's shape
 * (nvar, mx+4, my+4, mz+4).  Only faces of interior cells are written:
 * fm/fp at x-faces 1..mx+1 of rows 1..my, 1..mz; gm/gp, hm/hp likewise. */
int amrgpu3_get_fluxes(int patch, double* fm, double* fp, double* gm, double* gp,
                       double* hm, double* hp);

/* Single-grid step, src/3d/stepgrid.f:79-230: q_in -> q_out (ghost cells of
 * q_out = those of q_in), optional flux outputs (all six or NULL), cflgrid
 * and dtnew = dt cfl / cflgrid (rinfinity when cflgrid = 0).  Replaces the
 * batch defined by amrgpu3_set_patches. */
int amrgpu3_stepgrid(int mx, int my, int mz, const double* q_in, double* q_out, double dx,
                     double dy, double dz, double dt, double* fm, double* fp, double* gm,
                     double* gp, double* hm, double* hp, double* cflgrid, double* dtnew);

/* Flag every grid of the batch for refinement: flag2refine.f (3-D):3 as
 * spest.f:71-84 calls it with the library's allowflag (always true), on the
 * grids' current solution (ghost cells filled).  Flags are FP64 per grid,
 * (mx+4, my+4, mz+4): interior cells 2 (DOFLAG, badpt) where the largest
 * centred difference of any component in any direction exceeds tolsp, else
 * 0 (DONTFLAG); the border stays 0 (goodpt).  nflagged (may be NULL)
 * receives the number of flagged cells. */
int amrgpu3_flag_patches(double tolsp, int64_t* nflagged);
int amrgpu3_get_flags(int patch, double* amrflags);

/* conck.f (3-D):6: hx hy hz times the sum of q(1) over the interiors of the
 * batch's grids, in the reference's sequential order (grids, then k, j, i):
 * the reference's FP64 total bit for bit. */
int amrgpu3_conck(double hx, double hy, double hz, double* totmass);

/* ---- Hierarchy: the level update around the step (src/3d advanc.f) -----
 * Drop-in for the reference's 3-D advanc(level, nvar, dtlevnew, ...) and
 * update(level, nvar, naux) (tick.f calls both by name): the grids of every
 * level live on the device with both time levels; advance_level fills
 * every grid's ghost cells (bound.f, filpatch.f, intfil.f, prefilp.f,
 * bc3amr.f90), saves the coarse perimeter of the next level (saveqc.f,
 * cstore.f), copies new to old, adds qad.f's coarse-fine Riemann fluxes,
 * steps every grid (stepgrid.f) and saves fluxes in the registers
 * (fluxsv.f, fluxad.f); update_level refluxes (upbnd.f) and averages
 * (update.f).  Periodic domains must be periodic in all three directions
 * (saveqc.f wraps only then); mthbc 0 (user) and 4 are refused. */

/* Levels 1..nlev: index-space sizes, ratios to the next finer level
 * (intratx/y/z), mesh widths (hxposs/hyposs/hzposs); domain = (xlower,
 * xupper, ylower, yupper, zlower, zupper); mthbc(1:6).  Drops every level. */
int amrgpu3_set_geometry(int nlev, const int* iregsz, const int* jregsz, const int* kregsz,
                         const int* ratx, const int* raty, const int* ratz, const double* hx,
                         const double* hy, const double* hz, const double* domain,
                         const int* mthbc);
/* The grids of a level: boxes[6g..6g+5] = node(ndilo, ndihi, ndjlo, ndjhi,
 * ndklo, ndkhi); allocates their storage (new and old) on the device. */
int amrgpu3_set_level(int level, int ngrids, const int* boxes);
/* Planes per workgroup of the levels' sweeps, applied by later
 * amrgpu3_set_level calls (no reference counterpart: a launch shape).  0
 * (default): per level, 1 (FP32) or 2 (FP64) for levels of few grids, the
 * batch default for levels of 4096 planes or more. */
int amrgpu3_set_plane_block(int plane_block);
/* rnode(timemult) of the level's grids and possk(level). */
int amrgpu3_set_time(int level, double time, double possk);
int amrgpu3_get_time(int level, double* time);
/* Copy a level's new (which = 0) or old (1) solutions to / from the
 * device: grid g is the reference's alloc(node(store1|store2)) array
 * q(nvar, mx+4, my+4, mz+4) at alloc_base + offsets[g]. */
int amrgpu3_put_level(int level, int which, const double* alloc_base, const int64_t* offsets);
int amrgpu3_get_level(int level, int which, double* alloc_base, const int64_t* offsets);
/* bound.f (3-D) for every grid of the level at `time`. */
int amrgpu3_bound_level(int level, double time);
/* advanc.f (3-D): dtlevnew = min over grids of possk*cfl/cflgrid, the
 * level's largest Courant number; advances the level's time by possk. */
int amrgpu3_advance_level(int level, double* dtlevnew, double* cfl_level);
/* update.f (3-D) with upbnd.f: onto `level` from level+1. */
int amrgpu3_update_level(int level);
/* conck.f (3-D): hx hy hz times the sum of q(1) over the level's grids, in
 * the reference's sequential order (bit for bit). */
int amrgpu3_conck_level(int level, double* totmass);
/* Register block of fine grid `grid` of `level` (prepf.f): svdflx(nvar,
 * lenbc) then qc1d(nvar, lenbc); lenbc = 2 (ik jk + jk kk + kk ik).  With
 * svdflx_qc1d NULL only lenbc is returned. */
int amrgpu3_get_registers(int level, int grid, double* svdflx_qc1d, int* lenbc);
/* Flagging of the level's grids, replacing the per-grid loops of spest.f
 * (3-D):56-84 and errest.f (3-D):68-216.  spest: every grid's ghost cells
 * at the level's time (bound.f), then flag2refine.f into a per-grid FP64
 * array (mx+4, my+4, mz+4) of badpt (2) / goodpt (0).  With flag_richardson
 * (levels below the finest): errest.f's copysol of the old values, bound
 * with 2 nghost ghost cells at time - possk, coarsen.f, the two FP64
 * stepgrid calls (2 possk on the coarsened grid, possk on the grid's new
 * values) and errf1.f into a second array, which the reference leaves in
 * component 1 of store2 and the device keeps apart (store2 is unchanged).
 * nflag / nrich (may be NULL) receive the flagged cell counts.  Grids
 * need even sizes for the Richardson estimate. */
int amrgpu3_flag_level(int level, double tolsp, int flag_richardson, double tol, int iorder,
                       int64_t* nflag, int64_t* nrich);
/* The two flag arrays of grid `grid` (either pointer may be NULL). */
int amrgpu3_get_level_flags(int level, int grid, double* amrflags, double* rflags);

/* Device timing of the step's kernels: amrgpu3_set_timing(1) resets and
 * starts it; kind 0, 1, 2 = x-, y-, z-sweep, 3 = update. */
int amrgpu3_set_timing(int on);
int amrgpu3_kernel_time(int kind, double* ms, int64_t* launches);
int amrgpu3_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* AMRGPU3_H */
