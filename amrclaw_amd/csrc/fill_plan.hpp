// fill_plan.hpp -- hierarchy geometry and the ghost-fill planner interface.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "fill.hpp"

namespace amrgpu {

// Index-space description of one level (the parts of amr_module's
// node/hxposs/iregsz/intratx tables the fill needs).
struct LevelGeom {
    std::vector<std::array<int, 4>> boxes;   // (ilo, ihi, jlo, jhi) per grid
    int iregsz = 0, jregsz = 0;              // level size in cells
    int ratx = 2, raty = 2;                  // ratio to the next finer level
    double hx = 0, hy = 0;
    // listbc of every grid as the caller's prepc.f built it (i, j counted
    // from the first ghost cell, side, kid as its 0-based index on the next
    // finer level, slot); empty: the planner builds its own (setuse.f order
    // over the finer grids in their index order)
    std::vector<std::vector<std::array<int, 5>>> listbc;
};

// A bucket index over the boxes of one level: the grids a rectangle can
// meet, in ascending grid order, without scanning every grid of the level
// (the reference scans them all: intfil.f90:153-156 with msrc = -1,
// init_bndryList.f:50-107 for the neighbour lists).  Buckets are squares of
// about the mean box size over the level's index space.
struct BoxIndex {
    int bs = 1, nbx = 0, nby = 0;
    std::vector<int> start, items;   // CSR: grids of bucket b in items[start[b]..start[b+1])
    void build(const std::vector<std::array<int, 4>>& boxes, int iregsz, int jregsz);
    // grids whose box meets [ilo..ihi] x [jlo..jhi], ascending
    void query(int ilo, int ihi, int jlo, int jhi, std::vector<int>& out) const;
    bool empty() const { return nbx == 0; }
};

struct HierGeom {
    int nlev = 0;
    LevelGeom lev[FILL_MAXLV];
    double xlower = 0, xupper = 1, ylower = 0, yupper = 1;
    bool xperdom = false, yperdom = false, spheredom = false;
    int mthbc[4] = {1, 1, 1, 1};
    int naux = 0;                            // aux values per cell (register arenas hold auxc1d)
};

struct FillPlan {
    std::vector<FillBuf> bufs;
    std::vector<FillOp> ops;                 // sorted by stage
    std::vector<unsigned char> masks;
    std::vector<int> stage_start;            // ops of stage s: [stage_start[s], stage_start[s+1])
    int nstages = 0;
    int64_t scratch_elems = 0;
    uint32_t src_levels_mask = 0;            // levels read as time sources
};

bool plan_fill_level(const HierGeom& H, int level, int nvar, FillPlan& P, std::string& err);

// errest.f for every grid of `level`: the fill program builds, in an FP64
// arena, the coarsened old solution of each grid (copysol of store2,
// bound with 2 nghost ghost cells at the old time, coarsen.f) and the copy
// of the grid's new (or old) values with ghost cells; crs / fine describe
// those arrays as patches (offsets into the arena) for the two stepgrid
// calls of prepbigstep.f / prepregstep.f.
struct RichPlan {
    FillPlan fill;
    std::vector<PatchDesc> crs, fine;
    int64_t arena_elems = 0;
};
bool plan_richardson(const HierGeom& H, int level, int nvar, RichPlan& R, std::string& err);
// filval.f90 (gfixup.f:60-98) for the new grids of `level` (H.lev[level]),
// the old grids of the level described in H.lev[FILL_OLDSLOT]
bool plan_filval(const HierGeom& H, int level, int nvar, FillPlan& P, std::string& err);
// saveqc.f for the grids of fine level `level` (coarse level `level`-1)
bool plan_saveqc(const HierGeom& H, int level, int nvar, FillPlan& P, std::string& err);

}  // namespace amrgpu
