/*
 * amrgpu.h -- C ABI of the MI355X level-update library (libamrgpu.so).
 *
 * Drop-in boundary for AMRClaw's per-level patch update.  The reference
 * driver (tick.f) calls advanc/update/flagger; a Fortran ISO_C_BINDING shim
 * replaces those routines and calls the entry points below.  All arguments
 * are plain C types; arrays are host pointers unless stated otherwise.
 * Solution arrays cross the boundary in the reference's layout: FP64,
 * q(nvar, 1-nghost:mx+nghost, 1-nghost:my+nghost), component fastest
 * (update.f:31).  Indices are the reference's global 0-based cell indices of
 * a level (node(ndilo..ndjhi, mptr), amr_module.f90:100-112).
 *
 * Every function returns 0 on success and a nonzero code on failure; the
 * message is available from amrgpu_last_error().  No C++ exception crosses
 * the boundary.  The library is not thread-safe: call it from one host
 * thread (the shim replaces the reference's OpenMP patch loops, advanc.f:60,
 * 106, with device parallelism).
 */
#ifndef AMRGPU_H
#define AMRGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    AMRGPU_OK = 0,
    AMRGPU_ERR_ARG = 1,       /* invalid argument                          */
    AMRGPU_ERR_STATE = 2,     /* call out of order (e.g. before init)      */
    AMRGPU_ERR_HIP = 3,       /* HIP runtime error                          */
    AMRGPU_ERR_UNSUPPORTED = 4
};

/* AMRGPU_VCADV: variable-coefficient advection q_t + u q_x + v q_y = 0 with
 * the edge velocities in aux(1) (x-velocity at a cell's left edge) and aux(2)
 * (y-velocity at its bottom edge) -- the rpn2/rpt2 of the reference's
 * tests/advection_2d_annulus; needs naux >= 2 and FP64 or FP32 precision. */
enum { AMRGPU_ADVECTION = 1, AMRGPU_ACOUSTICS = 2, AMRGPU_EULER = 3, AMRGPU_VCADV = 4 };
/* Precision of a run.  FP64: the reference's arithmetic, bit for bit.
 * FP32: storage and wave/flux arithmetic in FP32; registers, coarse
 * interpolation, averaging and totals FP64.  MIXED: wave/flux arithmetic in
 * FP32 on FP64 storage, in flux form (both cells of an interface take
 * f(q_l) + A-dq) with the update in FP64, so that totals are conserved to FP64
 * rounding; everything else as FP64. */
enum { AMRGPU_FP64 = 0, AMRGPU_FP32 = 1, AMRGPU_MIXED = 2 };

/* Run parameters: the amr_module.f90 fields the level update reads, filled
 * by the shim from claw.data/amr.data (amr2.f90:152-392), plus the physics
 * constants the reference's solvers read from common /cparam/. */
typedef struct amrgpu_config {
    int solver;          /* AMRGPU_ADVECTION / _ACOUSTICS / _EULER, or the id
                            of a registered plug-in solver                    */
    int nvar;            /* num_eqn                                           */
    int mwaves;          /* num_waves                                         */
    int naux;            /* num_aux: aux values per cell (0..8)                */
    int nghost;          /* num_ghost (must be 2)                             */
    int method[7];       /* method(1:7): [1]=order, [2]=transverse            */
    int mthlim[8];       /* limiter per wave: 0 none 1 minmod 2 superbee
                            3 van Leer 4 MC 5 Beam-Warming                    */
    int mcapa;           /* capa_index: the capacity's aux component (1-based),
                            0 for none                                        */
    int use_fwaves;
    double cfl;          /* cfl_desired                                       */
    double cflv1;        /* cfl_max                                           */
    int mthbc[4];        /* bc_lower/upper: 1 extrap 2 periodic 3 wall        */
    double xlower, xupper, ylower, yupper;
    double ubar, vbar;   /* advection velocities                              */
    double rho, bulk;    /* acoustics                                         */
    double gamma;        /* Euler                                             */
    int efix;            /* Euler entropy fix                                 */
    int precision;       /* AMRGPU_FP64 or AMRGPU_FP32 (bulk arithmetic)     */
    int auxtype[8];      /* aux_type per component (amr_module auxtype):
                            AMRGPU_AUX_CENTER, _XLEFT, _YLEFT, _CAPACITY; qad.f
                            takes the fine grid's face value for xleft /
                            yleft components                                  */
} amrgpu_config;
enum { AMRGPU_AUX_CENTER = 0, AMRGPU_AUX_XLEFT = 1, AMRGPU_AUX_YLEFT = 2, AMRGPU_AUX_CAPACITY = 3 };

/* Initialise (or re-initialise) the library on the current HIP device.
 * Replaces the parameter set-up the reference does in amr2.f90:148-392. */
int amrgpu_init(const amrgpu_config* cfg);

/* Release all device memory and state. */
int amrgpu_finalize(void);

/* Copy the last error message (NUL-terminated, truncated to len). */
int amrgpu_last_error(char* buf, int len);

/* Tile-size caps of the sweep (autotuner knob); 0 keeps the current value. */
int amrgpu_set_tiling(int txmax, int tymax);
/* Occupancy bounds (autotuner knobs, no counterpart in the reference):
 * 5, 6 or 7 set the workgroups per CU the register allocation of the
 * AMRGPU_MIXED tile sweep is bounded for (default 7; the other precisions
 * ignore it); 2 or 3 set the waves per SIMD of the march sweep with FP32
 * arithmetic (AMRGPU_LAUNCH_MARCH; default 2, no spill; 3 spills part of
 * its state); 0 restores both defaults. */
int amrgpu_set_occupancy(int wgs);

/* Single-patch time step: stepgrid.f:24 (step2 + conservative update + dtnew)
 * on one grid with ghost cells already filled.  q_in/q_out: (nvar, mx+4,
 * my+4) AoS FP64; fm/fp/gm/gp: same shape, may all be NULL (stepgrid's flux
 * outputs, only interior-row x-edges and interior-column y-edges are set);
 * returns cflgrid and dtnew (rinfinity when cflgrid == 0). */
int amrgpu_stepgrid(int mx, int my, const double* q_in, double* q_out,
                    double dx, double dy, double dt, double* fm, double* fp,
                    double* gm, double* gp, double* cflgrid, double* dtnew);

/* amrgpu_stepgrid on device-resident arrays without allocating: no
 * counterpart in the reference.  For callers that cannot call hipMalloc
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  q_in / q_out are DEVICE pointers to the reference's layout,
 * AoS FP64 (nvar, mx+4, my+4), ghost cells of q_in filled; q_out receives
 * the updated interior and q_in's ghost cells, exactly as amrgpu_stepgrid's
 * q_out (same kernels, same arithmetic).  cflgrid / dtnew are host
 * pointers, returned after the stream has drained.  The work buffers (the
 * grid's descriptor and tile table, its Courant maximum and two SoA copies
 * of the grid) live in one static device array of
 * AMRGPU_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends: grids
 * whose two copies need more than that (in FP64 storage with nvar = 4,
 * about (mx+4)(my+4) <= 260,000 cells) or more than AMRGPU_NOALLOC_MAX_TILES
 * tiles are refused with AMRGPU_ERR_ARG.  Every call on a GPU shares that
 * GPU's array, so calls that overlap on one GPU (other host threads, other
 * streams) race; the array keeps the last call's contents.  Call amrgpu_init
 * on the device first (it looks up the array's address there). */
#define AMRGPU_NOALLOC_SCRATCH_BYTES (16u << 20)
#define AMRGPU_NOALLOC_MAX_TILES 4096
int amrgpu_stepgrid_noalloc(int mx, int my, const double* q_in, double* q_out,
                            double dx, double dy, double dt, double* cflgrid,
                            double* dtnew);

/* Level geometry of the whole run (domain.f:92-106, amr2.f90:455): for
 * levels 1..mxnest the index-space size iregsz/jregsz, the refinement ratio
 * to the next level intratx/intraty and the mesh widths hxposs/hyposs
 * (arrays indexed from 0 for level 1).  The domain and boundary conditions
 * come from amrgpu_config (periodic directions: mthbc 2 on both sides,
 * amr2.f90:279-280). */
int amrgpu_set_geometry(int mxnest, const int* iregsz, const int* jregsz,
                        const int* intratx, const int* intraty,
                        const double* hxposs, const double* hyposs);

/* Time of every grid of a level (rnode(timemult)) and its step possk. */
int amrgpu_set_time(int level, double time, double possk);

/* Fill the ghost cells of every grid of a level at time `time`:
 * bound.f90:51-158 for all grids (advanc.f:60-84), including same-level
 * copies, time-interpolated coarse values and limited linear coarse-fine
 * interpolation through coarser levels (filpatch.f90:71-355), periodic
 * wrap (prefilp.f90:18) and physical boundary conditions (bc2amr.f90:84).
 * The fill program is planned on the host at the first call after a
 * layout change and replayed on the device afterwards. */
int amrgpu_bound_level(int level, double time);

/* One time step of every grid of a level, the body of advanc.f:16-305:
 * bound at the level's time; saveqc.f:13-109 for the next finer level (its
 * registers receive the coarse values around each fine grid); the sweep of
 * stepgrid.f for all grids with fluxsv.f (the coarse fluxes through the
 * edges of the finer grids into their registers) and, for a fine level,
 * qad.f + fluxad.f (the Riemann fix-up and the fine fluxes along the grid
 * perimeters); the level's time advances by possk.  Returns the smallest
 * new time step over the level's grids (stepgrid.f:206-212,
 * advanc.f:124-126) and the largest Courant number.  The register layout
 * (prepf.f/prepc.f) is rebuilt after any change of the hierarchy. */
int amrgpu_advance_level(int level, double* dtlevnew, double* cfl_level);

/* update.f:16-189 for coarse level `level`: upbnd.f's conservative reflux
 * from the registers of level+1 into the coarse cells next to the fine
 * grids, then the coarse cells under level+1 set to the mean of the fine
 * cells they cover. */
int amrgpu_update_level(int level);

/* conck.f:36-63: sum of component 1 over the interior of all grids of
 * `level`, times hx*hy.  Summed on the device in the reference's sequential
 * order (grids in level order, then j, then i), so the total is the
 * reference's bit for bit.  A diagnostic: one wave, latency-bound. */
int amrgpu_conck(int level, double* totmass);

/* Lower-left corners of the grids of a level, rnode(cornxlo/cornylo)
 * (grdfit2.f:98-101, birect.f:67), used where the reference computes
 * positions from them (flagregions2.f90:61-92).  Default after
 * amrgpu_set_level: ilo*hx + xlower, jlo*hy + ylower. */
int amrgpu_set_corners(int level, const double* xlo, const double* ylo);

/* Flag the cells of every grid of `level` as flagger.f:56-164 does: the
 * flags (mx+2*mbuff) x (my+2*mbuff) of each grid (FP64: -1 UNSET,
 * 0 DONTFLAG, 2 DOFLAG; amr_module.f90:163-168) are set UNSET, then
 * flagregions2.f90:40-120 applies the regions (nregions records of 8
 * doubles in regions.data order: min_level, max_level, t_low, t_hi, x_low,
 * x_hi, y_low, y_hi), then, with flag_gradient, flag2refine2.f90:96-122
 * flags UNSET interior cells whose undivided difference exceeds tolsp,
 * then, with flag_richardson, errest.f (prepbigstep/prepregstep/errf1:
 * the old values coarsened by 2 and advanced 2 dt against the values
 * advanced dt, component 1, est = |mean - coarse| / (2^(iorder+1) - 2)
 * >= tol marks the 2x2 block; not on level mxnest, which keeps no old
 * values).  Like flagger.f:98, the new values are used after a ghost fill
 * at the level's time unless start_time + possk equals that time exactly
 * (then the old values).  With flags_base non-NULL each grid's flags are
 * copied to flags_base + offsets[k] (node(storeflags)); *nflagged
 * (optional) gets the number of DOFLAG cells. */
int amrgpu_flag_level(int level, int flag_gradient, double tolsp, int flag_richardson,
                      double tol, int iorder, double start_time, int nregions,
                      const double* regions, int mbuff, double* flags_base,
                      const int64_t* offsets, int64_t* nflagged);

/* Sweep launch shape: AMRGPU_LAUNCH_BATCHED (default) puts all tiles of
 * all grids of a level into one launch, one workgroup per tile;
 * AMRGPU_LAUNCH_PERSISTENT launches as many workgroups as the device holds
 * and lets them take the level's tiles (largest first) from a device
 * counter; AMRGPU_LAUNCH_PER_PATCH issues one launch per grid, the plain
 * per-patch GPU path, kept as the comparison baseline;
 * AMRGPU_LAUNCH_MARCH runs the sweep of a level that has no finer level as
 * column strips marching up the rows with every quantity in registers (one
 * Riemann solve per interface instead of the tiles' halo re-solves; levels
 * with a finer level keep the tile sweep).  Results are identical (FP64) or
 * agree to FP32 rounding.  No reference counterpart. */
enum { AMRGPU_LAUNCH_BATCHED = 0, AMRGPU_LAUNCH_PER_PATCH = 1, AMRGPU_LAUNCH_PERSISTENT = 2,
       AMRGPU_LAUNCH_MARCH = 3 };
int amrgpu_set_launch_mode(int mode);

/* With on != 0, amrgpu_advance_level captures its launches (ghost fill,
 * saveqc, sweep, register fix-up) into a hipGraph on first use and replays
 * it while the layout, the step, the buffer roles and the time-interpolation
 * weights are the same (between regrids, every step of a fixed-dt run
 * repeats one of a few such keys).  No reference counterpart. */
int amrgpu_set_graphs(int on);

/* Per-kernel-class timing with HIP events on the library's stream (not in
 * graph mode).  amrgpu_set_timing(1) resets and starts it; amrgpu_kernel_time
 * returns the summed milliseconds and the number of timed launches of one
 * class.  No reference counterpart. */
enum { AMRGPU_KT_SWEEP = 0, AMRGPU_KT_FILL = 1, AMRGPU_KT_REGISTERS = 2, AMRGPU_KT_UPDATE = 3,
       AMRGPU_KT_FLAG = 4, AMRGPU_KT_COUNT = 5 };
int amrgpu_set_timing(int on);
int amrgpu_kernel_time(int kind, double* ms, int64_t* launches);

/* The registers of grid `grid` of fine level `level`: svdflx(nvar, lenbc)
 * followed by qc1d(nvar, lenbc), component fastest, as the reference's
 * node(ffluxptr) block (prepf.f:33-48).  For tests and debugging. */
int amrgpu_get_registers(int level, int grid, double* svdflx_qc1d, int* lenbc);

/* Ghost-fill program inspection (host only; needs no device and no init).
 * Plans bound for every grid of `level` of the hierarchy described by the
 * arguments (levels 1..nlev; grid boxes concatenated level by level as
 * (ilo, ihi, jlo, jhi); per-level arrays indexed from 0 for level 1;
 * domain = xlower, xupper, ylower, yupper) and reports its size; the
 * records are then copied out with amrgpu_plan_fill_get.  Operation types:
 * 0 copy/time-blend from a grid, 1 coarse-fine interpolation, 2 boundary
 * conditions on a whole buffer, 3 copy of a periodic piece (4, the saved
 * coarse ring of the flux registers, only appears in saveqc programs).  Buffers of
 * kind 0 are grids of the level store, kind 1 FP64 scratch arrays. */
typedef struct amrgpu_fill_op {
    int type, dst, src, src_level;
    int di0, dj0, ni, nj;
    int si0, sj0;
    int mask_off;
    int ilo, jlo, iplo, jplo, ratx, raty, level;
    int bc[4];
    int stage;
    int flags;      /* 1: copy the new-time values, no time blend (icall.f) */
    int pad;
} amrgpu_fill_op;
typedef struct amrgpu_fill_buf {
    int kind, level, patch, nrow, ncol, plane, pad;
    int64_t off;
} amrgpu_fill_buf;
int amrgpu_plan_fill(int nlev, const int* ngrids, const int* boxes,
                     const int* iregsz, const int* jregsz, const int* intratx,
                     const int* intraty, const double* hxposs,
                     const double* hyposs, const double* domain,
                     const int* mthbc, int nvar, int level, int* nops,
                     int* nbufs, int64_t* nmask, int64_t* nscratch);
int amrgpu_plan_fill_get(amrgpu_fill_op* ops, amrgpu_fill_buf* bufs,
                         unsigned char* masks);

/* The aux arrays of every grid of `level` (alloc(node(storeaux)): naux
 * values per cell, component fastest, ghost cells included, as setaux
 * leaves them), at 0-based element offsets offsets[k] of alloc_base.  The
 * library keeps them until the level's layout changes (set them again after
 * every amrgpu_set_level of the level: amrgpu_advance_level refuses a step
 * that would read a level's aux arrays not set since).  The capacity (mcapa) scales dtdx
 * in the sweep and the update (step2.f90:91-95, stepgrid.f:166-172), the
 * reflux (upbnd.f:104-108), the averaging (update.f:147-161) and conck
 * (conck.f:56-60); solvers that read aux get the interface's cells' values. */
int amrgpu_put_level_aux(int level, const double* alloc_base, const int64_t* offsets);

/* Regrid of one level on the device (gfixup.f:33-98 with filval.f90:3-188,
 * naux = 0): the grids of `level` are replaced by the ngrids new boxes, and
 * the new grids are filled as filval fills them: the old grids' new-time
 * values copied over each new grid and its ghost ring (intcopy.f /
 * preintcopy.f), and the min-mod interpolation of the coarse patch grown
 * by one coarse cell (from level-1's current grids, with bc2amr) into the
 * interior cells no old grid covers.  Level-1 must hold the (already
 * regridded) coarse grids the new ones nest in, as regrid builds levels
 * coarse to fine.  The level's time becomes `time` (rnode(timemult) of the
 * new grids); ghost cells no old grid covers are left for the next bound.
 * The old grids are released; registers and fill programs are planned anew
 * at the next step, as after amrgpu_set_level. */
int amrgpu_regrid_level(int level, int ngrids, const int* ilo, const int* ihi, const int* jlo,
                        const int* jhi, double time);

/* The neighbour lists of makeBndryList (init_bndryList.f:50-107; host
 * only, needs no device and no init): for each of the n grids of one level
 * (boxes as (ilo, ihi, jlo, jhi)), the other grids whose box meets its box
 * grown by two ghost cells, in the order the reference's linked list holds
 * them (descending grid index), found through a bucket index over the
 * level's index space (iregsz x jregsz) instead of the reference's
 * all-pairs scan.  counts[k] receives grid k's count and items (room for
 * cap entries) the lists one after another; *total is the full number of
 * entries, and AMRGPU_ERR_ARG is returned when it exceeds cap. */
int amrgpu_bndry_lists(int n, const int* boxes, int iregsz, int jregsz, int* counts, int* items,
                       int64_t cap, int64_t* total);

/* Register-list inspection (host only; needs no device and no init): the
 * listbc entries (prepc.f:47-154 with setuse.f:61-213) of grid `grid` of
 * coarse level `level` for the hierarchy described as in amrgpu_plan_fill,
 * five ints each (i, j counted from the first ghost cell, side, kid
 * numbered from 0 within level+1, slot), at most maxn written; *n is the
 * full count. */
int amrgpu_plan_listbc(int nlev, const int* ngrids, const int* boxes,
                       const int* iregsz, const int* jregsz, const int* intratx,
                       const int* intraty, const int* mthbc, int nvar, int level,
                       int grid, int* list5, int maxn, int* n);

/* The coarse side of the flux registers of grid `grid` of `level` as the
 * caller's prepc.f built it (node(cfluxptr), prepc.f:47-154): n entries of
 * five ints, (i, j) of the coarse cell counted from the first ghost cell,
 * the side 1..4 facing the fine grid, the fine grid as its 0-based index on
 * level+1 (amrgpu_set_level order) and its register slot.  upbnd adds the
 * entries of a coarse cell in list order (upbnd.f:59-113), so handing the
 * reference's lists over keeps FP64 sums bit for bit when the reference's
 * kid order differs from its grid list (regrid.f builds the lists before
 * arrangeGrids sorts the grids).  Without this call the library builds the
 * lists itself, visiting the finer grids in their index order.  Either all
 * grids of a level have a list (possibly empty) or none; amrgpu_set_level
 * of the level or of the next finer level discards them. */
int amrgpu_set_listbc(int level, int grid, int n, const int* list5);

/* Define the grids of one level (after regrid): global index boxes
 * [ilo..ihi] x [jlo..jhi] and mesh widths hx, hy (hxposs/hyposs).
 * Replaces the layout bookkeeping of regrid.f:72-118 on the device side;
 * allocates the level's two solution buffers (new and old time). */
int amrgpu_set_level(int level, int ngrids, const int* ilo, const int* ihi,
                     const int* jlo, const int* jhi, double hx, double hy);

/* Upload / download one grid's solution (which: 0 = new time, store1;
 * 1 = old time, store2), AoS FP64 (nvar, mx+4, my+4) including ghosts. */
int amrgpu_put_grid(int level, int grid, int which, const double* q);
int amrgpu_get_grid(int level, int grid, int which, double* q);

/* Upload / download every grid of a level at once from the reference's heap
 * layout: grid p's AoS FP64 block (nvar, mx+4, my+4) starts at
 * alloc_base[offsets[p]] (the shim passes alloc and node(store1|store2,
 * mptr)-1 for the level's grids in amrgpu_set_level order; amr_module.f90
 * alloc, store1/store2).  which: 0 = new time, 1 = old time. */
int amrgpu_put_level(int level, int which, const double* alloc_base,
                     const int64_t* offsets);
int amrgpu_get_level(int level, int which, double* alloc_base,
                     const int64_t* offsets);

/* Advance every grid of a level by dt from its current ghost-filled
 * state: the stepgrid part of par_advanc (advanc.f:170-292) for all grids in
 * one launch.  new -> old, updated interior -> new.  Returns the level's
 * max Courant number (cfl_level) and min suggested dt (dtlevnew,
 * advanc.f:124-126). */
int amrgpu_sweep_level(int level, double dt, double* cfl_level,
                       double* dtlevnew);

/* ---- Gauges ----------------------------------------------------------------
 * The reference samples every gauge from its best source grid (setbestsrc:
 * the finest grid containing it, gauges_module.f90:255-302) before that
 * grid's step, by bilinear interpolation of the time-t values with the
 * ghost cells filled (update_gauges, gauges_module.f90:313-452, called from
 * par_advanc, advanc.f:249-253).  Here amrgpu_advance_level samples the
 * gauges whose source lies on the level it advances, on the device, with
 * the same arithmetic (FP64 records equal the reference's bit for bit),
 * and keeps the records until they are fetched.
 *
 * amrgpu_set_gauges: locations, time windows and minimum time increments
 * (gauges.data); replaces any previous set.  amrgpu_set_gauge_sources: per
 * gauge the level and the grid (index within amrgpu_set_level's order) of
 * its source, level 0 for none; call it after every layout change, as the
 * reference calls setbestsrc after regrid (tick.f:203).
 * amrgpu_get_gauge_records: the records since the last call, oldest first
 * per gauge: gauge[r] (0-based), level[r], and vals[r*(nvar+1) ...] = the
 * grid time then the nvar interpolated components; fails if more than
 * maxrec are waiting or records were lost. */
int amrgpu_set_gauges(int n, const double* x, const double* y, const double* t_start,
                      const double* t_end, const double* min_dt);
int amrgpu_set_gauge_sources(const int* level, const int* grid);
int amrgpu_get_gauge_records(int maxrec, int* nrec, int* gauge, int* level, double* vals);

/* ---- Riemann-solver plug-ins ---------------------------------------------
 * The reference passes the application's rpn2/rpt2 as external procedures
 * (step2.f90:7,26; flux2.f:25-28,125,202,219) and calls rpn2 by name from
 * qad.f:129, so an application supplies its own solvers at link time.  Here
 * an application writes its solver as a device functor with the same
 * argument meaning (include/amrgpu_solver.hpp documents the contract) in a
 * HIP translation unit of its own; AMRGPU_DEFINE_SOLVER there instantiates
 * the library's sweep and qad kernels for it and exports a function that
 * fills this table.  The application registers the table and selects the
 * solver by the id returned, in amrgpu_config.solver (with its nvar and
 * mwaves).  Builds with -ffp-contract=off reproduce the oracle's FP64
 * arithmetic bit for bit, as the library's own FP64 units do.
 *
 * The function pointers take the library's launch-argument blocks as
 * opaque pointers and return a hipError_t value (0 = success).  The table
 * records the sizes of those blocks as its translation unit saw them; the
 * library refuses a table built against other headers. */
#define AMRGPU_SOLVER_ABI 1
#define AMRGPU_SOLVER_PARAMS_MAX 16
#define AMRGPU_SOLVER_PLUGIN0 100      /* id of the first registered solver */
typedef struct amrgpu_solver_ops {
    int abi;                 /* AMRGPU_SOLVER_ABI                              */
    int nvar, mwaves;        /* meqn and mwaves of the solver                  */
    int sweep_args_bytes;    /* layout check of the launch-argument blocks     */
    int qad_args_bytes;
    int physics_bytes;
    const char* name;
    int (*sweep_f64)(const void* args, size_t lds, void* stream);
    int (*sweep_f32)(const void* args, size_t lds, void* stream);
    int (*sweep_mixed)(const void* args, size_t lds, void* stream, int wgs);
    int (*qad)(int esize, const void* args, const void* physics, void* stream);
} amrgpu_solver_ops;

/* Register a plug-in solver (no counterpart call in the reference: there the
 * linker binds rpn2/rpt2).  The table is copied; *solver_id receives
 * AMRGPU_SOLVER_PLUGIN0 + k for the k-th registration.  At most 16 solvers;
 * registration survives amrgpu_finalize. */
int amrgpu_register_solver(const amrgpu_solver_ops* ops, int* solver_id);

/* The constants a plug-in solver reads (Physics.user[0..n-1] on the device),
 * the counterpart of the /cparam/ common block its setprob fills.  Taken by
 * the next amrgpu_init, and at once by a library already initialised.
 * n <= AMRGPU_SOLVER_PARAMS_MAX. */
int amrgpu_set_solver_params(int n, const double* params);

/* Device timing on the library's stream: record event `slot` (0..15) and
 * read the elapsed milliseconds between two recorded slots. */
int amrgpu_event_record(int slot);
int amrgpu_event_elapsed(int slot_a, int slot_b, float* ms);
int amrgpu_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif
