// fill.hpp -- ghost-cell fill programs (bound.f90 and its recursion) as data.
//
// The host planner replays bound -> filrecur -> intfil/trimbd ->
// (recursive coarse patch) -> interpolation -> bc2amr, and prefilrecur's
// periodic split, on the index boxes of the hierarchy once per regrid.  It
// emits a list of operations on "buffers" (the level's patches and FP64
// scratch arrays for the coarse patches valcrse and periodic pieces
// valPatch), each tagged with a dependency stage.  At every level step one
// launch per stage executes all operations of that stage, one workgroup per
// operation.  Values are computed in FP64 in the reference's order.
#pragma once
#include <cstdint>
#include "types.hpp"

namespace amrgpu {

constexpr int FILL_MAXLV = 12;

enum FillOpType : int32_t {
    FOP_COPY = 0,      // intfil: copy / time-blend from a grid (intfil.f90:237-259)
    FOP_INTERP = 1,    // filrecur: limited linear interpolation (filpatch.f90:260-327)
    FOP_BC = 2,        // bc2amr on a whole buffer (bc2amr.f90:84-296)
    FOP_COPYBACK = 3,  // prefilrecur: patchCopyOut (prefilp.f90:218-245)
    FOP_RING = 4,      // cstore: coarse perimeter ring of a buffer -> qc1d (cstore.f:4-63)
    FOP_COARSEN = 5,   // coarsen.f: 2x2 means of a buffer into another (prepbigstep.f:58)
    FOP_FILVAL = 6     // filval.f90:154-188: min-mod interpolation of a coarse patch into
                       // the interior cells of a new grid no old grid covered
};
// the FillArgs level slot that holds the old grids of a level being
// regridded (gfixup.f fills the new grids while the old ones still exist)
constexpr int FILL_OLDSLOT = FILL_MAXLV - 1;
// FillOp.flags bits (COPY)
enum : int32_t {
    FOPF_FORCE_NEW = 1,   // icall copies store1, no time blend (icall.f)
    FOPF_FORCE_OLD = 2,   // copysol of store2 (prepbigstep.f:35)
    FOPF_BIG = 4          // flagger's locbig: store1, or store2 when FillArgs.big_old (flagger.f:98-117)
};

// A destination/scratch array, SoA over components: cell (i, j), 0-based
// with ghosts, component m at off + m*plane + j*nrow + i.
struct FillBuf {
    int32_t kind;      // 0: patch `patch` of level `level` (level storage); 1: FP64 scratch;
                       // 2: FP64 arena (FillArgs.arena) read afterwards by other kernels
    int32_t level;
    int32_t patch;
    int32_t nrow, ncol, plane;
    int32_t pad;
    int64_t off;       // kind 1: element offset in the scratch arena
};

struct FillOp {
    int32_t type;
    int32_t dst;              // destination buffer
    int32_t src;              // COPY: source grid of src_level; INTERP/COPYBACK: source buffer
    int32_t src_level;        // COPY: level of the source grid
    int32_t di0, dj0, ni, nj; // destination rectangle (0-based cells of dst)
    int32_t si0, sj0;         // COPY: matching cell of the source grid (0-based, with ghosts)
    int32_t mask_off;         // INTERP: offset of the ni*nj flaguse bytes (row-major)
    int32_t ilo, jlo;         // INTERP: global index of rectangle cell (0,0), fine level
    int32_t iplo, jplo;       // INTERP: global index of coarse cell (0,0) of src buffer
    int32_t ratx, raty;       // INTERP: refinement ratio coarse -> fine
    int32_t level;            // INTERP: fine level
    int32_t bc[4];            // BC: nxl, ibeg, nyb, jbeg (0 = side not applied)
    int32_t stage;
    int32_t flags;            // FOPF_* bits
    int32_t pad;
};

enum TimeMode : int32_t { TIME_NEW = 0, TIME_OLD = 1, TIME_BLEND = 2 };

struct FillArgs {
    const FillOp* ops;
    const FillBuf* bufs;
    double* scratch;
    const unsigned char* masks;
    void* lvl_new[FILL_MAXLV];           // level storage, new time (store1)
    void* lvl_old[FILL_MAXLV];           // old time (store2)
    const PatchDesc* lvl_patches[FILL_MAXLV];
    double* lvl_reg[FILL_MAXLV];          // register arenas (FOP_RING)
    double alphai[FILL_MAXLV], alphac[FILL_MAXLV];
    int32_t tmode[FILL_MAXLV];
    double hx[FILL_MAXLV], hy[FILL_MAXLV];
    double xlower, ylower;
    int32_t mthbc[4];
    int32_t nv;
    int32_t nops;
    double* arena;                        // kind-2 buffers
    int32_t big_old;                      // FOPF_BIG copies read store2
    int32_t pad;
    // zeroed by workgroup 0 of the launch (the level's per-grid Courant
    // maxima, cleared here for the sweep that follows: one launch fewer)
    unsigned long long* clear;
    int64_t nclear;
    // aux mode (a saveqc program run a second time for the aux values,
    // saveqc.f:60-90 and cstore.f): copies read the levels' aux arrays
    // (FP64, naux components, patch p at off * auxk), scratch buffer b's
    // component c starts at scratch + bufs[b].off * auxk + c * plane, rings
    // write auxc1d; boundary conditions leave aux alone
    const double* lvl_aux[FILL_MAXLV];
    int32_t aux_mode, naux, auxk, pad2;
};

}  // namespace amrgpu
