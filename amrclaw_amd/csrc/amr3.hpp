// amr3.hpp -- device programs of the 3-D level update around the step:
// ghost-cell fill programs (bound.f and its recursion) as data, the
// coarse-fine flux registers (saveqc/cstore, qad, fluxsv, fluxad) and
// refluxing/averaging (upbnd, update), all of src/3d.
//
// The host planner (amr3_plan.cpp) replays bound -> filrecur -> intfil /
// trimbd -> (recursive coarse patch) -> interpolation -> bc3amr and
// prefilrecur's periodic split on the index boxes of the hierarchy once per
// regrid, and emits operations on buffers (the level's grids and FP64
// scratch arrays for the coarse patches valcrse and periodic pieces
// valPatch), each tagged with a dependency stage.  At every level step one
// launch per stage executes all operations of that stage, copies and
// interpolations split into workgroup items of at most 2048 elements,
// boundary conditions one workgroup each, in FP64 and in the reference's
// order.
#pragma once
#include <cstdint>
#include "sweep3.hpp"

namespace amrgpu {

constexpr int AMR3_MAXLV = 12;

enum Fill3OpType : int32_t {
    F3_COPY = 0,      // intfil: copy / time-blend from a grid (intfil.f:140-180)
    F3_INTERP = 1,    // filrecur: limited linear interpolation (filpatch.f:160-260)
    F3_BC = 2,        // bc3amr on a whole buffer (bc3amr.f90)
    F3_COPYBACK = 3,  // prefilrecur: patchCopyOut (prefilp.f)
    F3_RING = 4,      // cstore: the coarse perimeter of a buffer -> qc1d (cstore.f)
    F3_COARSEN = 5    // coarsen.f: 2x2x2 means of a buffer into another (errest.f)
};
enum : int32_t {
    F3F_FORCE_NEW = 1,   // icall: store1, no time blend
    F3F_FORCE_OLD = 2    // copysol of store2 (errest.f)
};

// A destination/scratch array, structure of arrays: cell (i, j, k), 0-based
// with ghosts, component m at off + m*plane + (k*ncol + j)*nrow + i.
struct Fill3Buf {
    int32_t kind;          // 0: grid `patch` of level `level`; 1: FP64 scratch;
                           // 2: FP64 arena read afterwards by other kernels
    int32_t level, patch;
    int32_t nrow, ncol, nfil;
    int64_t plane;
    int64_t off;           // kind 1: element offset in the scratch arena
};

struct Fill3Op {
    int32_t type, dst, src, src_level;
    int32_t di0, dj0, dk0;       // destination box corner (0-based cells of dst)
    int32_t ni, nj, nk;          // box size
    int32_t si0, sj0, sk0;       // COPY: matching cell of the source grid (0-based, ghosts)
    int32_t mask_off;            // INTERP: ni*nj*nk flaguse bytes, i fastest
    int32_t ilo, jlo, klo;       // INTERP: global index of box cell (0,0,0), fine level
    int32_t iplo, jplo, kplo;    // INTERP: global index of coarse cell (0,0,0) of src
    int32_t ratx, raty, ratz;    // INTERP: refinement ratios coarse -> fine
    int32_t level;               // INTERP: fine level; RING: fine level
    int32_t bc[6];               // BC: nxl, ibeg, nyl, jbeg, nzl, kbeg (0: side not applied)
    int32_t stage, flags;
};

enum Time3Mode : int32_t { T3_NEW = 0, T3_OLD = 1, T3_BLEND = 2 };

// one workgroup's share of an operation: elements [lo, hi) of its linear
// range (copies: cells x components; interpolation: cells); a boundary
// condition or ring op is one item
struct Fill3Work {
    int32_t op, lo, hi, pad;
};

// Per-grid register layout of a fine level (prepf.f): svdflx(nv, lenbc)
// at reg_off, qc1d(nv, lenbc) after it.
struct RegGrid3 {
    int64_t reg_off;
    int32_t lenbc, pad;
};

struct Fill3Args {
    const Fill3Work* work;       // this launch's items
    const Fill3Op* ops;          // all ops of the program
    const Fill3Buf* bufs;
    double* scratch;
    const unsigned char* masks;
    void* lvl_new[AMR3_MAXLV];
    void* lvl_old[AMR3_MAXLV];
    const Patch3* lvl_patches[AMR3_MAXLV];
    const RegGrid3* lvl_regs[AMR3_MAXLV];
    double* lvl_reg[AMR3_MAXLV];
    double alphai[AMR3_MAXLV], alphac[AMR3_MAXLV];
    int32_t tmode[AMR3_MAXLV];
    double hx[AMR3_MAXLV], hy[AMR3_MAXLV], hz[AMR3_MAXLV];
    double xlower, ylower, zlower;
    int32_t mthbc[6];
    int32_t nv, nwork;
    double* arena;               // kind-2 buffers
};

// fluxsv.f: coarse cell (I, J, K) (1-based with ghosts) of grid `grid`,
// side 1..6, register slot lkid (1-based) of fine grid kid; `write` = 0 for
// an entry whose slot a later entry overwrites (the reference's last write
// wins)
struct Sv3Entry {
    int32_t grid, I, J, K, side, kid, lkid, write;
};

// qad.f: one register slot of one side of a fine grid: coarse segment (c1,
// c2) (1-based along the side's two directions) and slot index
struct Qad3Item {
    int32_t grid, side, c1, c2, slot, pad;
};

// upbnd.f: one coarse cell and its ordered register slots
struct Upbnd3Cell {
    int32_t grid, i, j, k;     // 0-based cell of the grid array (with ghosts)
    int32_t e0, e1;            // entries [e0, e1)
};
struct Upbnd3Entry {
    int32_t kid, lkid;
    double sgnm;
};

// update.f: coarse cells [ci0, ci0+ni) x [cj0, ..) x [ck0, ..) of `grid` (0-based
// with ghosts) are the means of r^3 fine cells of `kid` from (fi0, fj0, fk0)
struct Avg3Block {
    int32_t grid, kid;
    int32_t ci0, cj0, ck0, ni, nj, nk;
    int32_t fi0, fj0, fk0, pad;
    int64_t cell0;             // first cell of this block in the launch's cell numbering
};

struct Reg3Args {
    int32_t nv;
    int32_t esize;             // storage element size (8 FP64, 4 FP32)
    int32_t merged;            // FP32 step: fp and gp hold fp - fm, gp - gm
    int32_t pad;
    // fine level (qad / fluxad)
    const Patch3* fpatches;
    const RegGrid3* fregs;
    double* freg;
    const void* fq;            // fine level, new time (ghosts filled)
    const void* fface;         // fine level's flux blocks (after its step)
    int32_t ratx, raty, ratz;  // intratx/y/z(level-1)
    int32_t nitems;
    const Qad3Item* qad;
    double hx, hy, hz, delt;
    // coarse level (fluxsv / upbnd / averaging)
    const Patch3* cpatches;
    const void* cface;
    void* cq;
    const Sv3Entry* sv;
    int32_t nsv;
    int32_t ncells;
    const Upbnd3Cell* ucells;
    const Upbnd3Entry* uents;
    const Avg3Block* avg;
    int32_t navg;
    int32_t cratx, craty, cratz;  // intratx/y/z(level)
    int64_t navg_cells;
    double chx, chy, chz, cdt;
    double volume;
};

}  // namespace amrgpu
