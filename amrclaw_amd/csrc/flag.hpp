// flag.hpp -- device arguments of the cell flagging kernel.
#pragma once
#include <cstdint>
#include "types.hpp"

namespace amrgpu {

// amr_module.f90:163-168
constexpr double FLAG_UNSET = -1.0, FLAG_DONTFLAG = 0.0, FLAG_DOFLAG = 2.0;

// one region clipped to one grid (flagregions2.f90:66-92): 1-based cells
// i1..i2, j1..j2 of the grid and the region's level bounds
struct RegionRect {
    int32_t i1, i2, j1, j2, min_level, max_level;
};

struct FlagArgs {
    const PatchDesc* patches;
    const void* q;                  // new (or old at the initial time), ghosts filled
    double* flags;
    const int64_t* foff;            // element offset of each grid's flags
    const RegionRect* rects;        // nrect per grid (empty rects: i1 > i2)
    unsigned long long* counters;   // [0] DOFLAG cells, [1] bit 0: minlevel > maxlevel
    int32_t nrect;
    int32_t nv;
    int32_t mbuff;
    int32_t level;
    double tolsp;
};

// errf1.f:33-137 over every grid of a level: the coarse cells of the
// 2 dt step (crs, in `crs_q`) against the 2x2 means of the dt step (fine,
// in `fine_q`), component 1, FP64
struct ErrfArgs {
    const PatchDesc* crs;
    const PatchDesc* fine;
    const double* crs_q;
    const double* fine_q;
    double* flags;
    const int64_t* foff;
    unsigned long long* counters;   // [0] DOFLAG cells (incremented for new marks)
    int32_t mbuff;
    int32_t pad;
    double tol;
    double order;                   // 2^(iorder+1) - 2
};

}  // namespace amrgpu
