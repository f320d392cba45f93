// amr3_plan.hpp -- host planning of the 3-D ghost-fill programs and of the
// coarse-fine registers of a 3-D hierarchy.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "amr3.hpp"

namespace amrgpu {

struct Level3Geom {
    std::vector<std::array<int, 6>> boxes;   // (ilo, ihi, jlo, jhi, klo, khi) per grid
    int iregsz = 0, jregsz = 0, kregsz = 0;  // level size in cells
    int ratx = 2, raty = 2, ratz = 2;        // ratios to the next finer level
    double hx = 0, hy = 0, hz = 0;
};

struct Hier3Geom {
    int nlev = 0;
    Level3Geom lev[AMR3_MAXLV];
    double xlower = 0, xupper = 1, ylower = 0, yupper = 1, zlower = 0, zupper = 1;
    bool xperdom = false, yperdom = false, zperdom = false;
    int mthbc[6] = {1, 1, 1, 1, 1, 1};
};

struct Fill3Plan {
    std::vector<Fill3Buf> bufs;
    std::vector<Fill3Op> ops;                // sorted by stage
    std::vector<unsigned char> masks;
    std::vector<Fill3Work> work;             // workgroup items, by stage
    std::vector<int> stage_start;            // items of stage s: [stage_start[s], stage_start[s+1])
    int nstages = 0;
    int64_t scratch_elems = 0;
    uint32_t src_levels_mask = 0;            // levels read as time sources
};

// bound.f (3-D) for every grid of `level`
bool plan_fill3_level(const Hier3Geom& H, int level, int nvar, Fill3Plan& P, std::string& err);
// saveqc.f (3-D) for the grids of fine level `level`
bool plan_saveqc3(const Hier3Geom& H, int level, int nvar, Fill3Plan& P, std::string& err);

// errest.f (3-D) for every grid of `level`: the fill program builds, in an
// FP64 arena, the coarsened old solution of each grid (copysol of store2,
// bound with 2 nghost ghost cells at t - possk, coarsen.f) and the copy of
// the grid's new values with ghost cells (spest's locbig); crs / fine
// describe those arrays as grids (qoff into the arena) for the two
// stepgrid calls of errest.f
struct Rich3Plan {
    Fill3Plan fill;
    std::vector<Patch3> crs, fine;
    int64_t arena_elems = 0;
};
bool plan_richardson3(const Hier3Geom& H, int level, int nvar, Rich3Plan& R, std::string& err);

struct Reg3Level {
    // fine side (level > 1): prepf.f
    std::vector<RegGrid3> regs;
    int64_t reg_elems = 0;
    std::vector<Qad3Item> qad;
    // coarse side (level has a finer level): prepc.f / setuse.f
    std::vector<Sv3Entry> sv;
    std::vector<Upbnd3Cell> ucells;
    std::vector<Upbnd3Entry> uents;
    std::vector<Avg3Block> avg;
    int64_t avg_cells = 0;
};

bool plan_registers3(const Hier3Geom& H, int nvar, Reg3Level* R, std::string& err);

}  // namespace amrgpu
