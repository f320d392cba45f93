// gauges.hpp -- gauge sampling on the device (gauges_module.f90:313-452).
#pragma once
#include <cstdint>
#include "types.hpp"

namespace amrgpu {

// One gauge: its location and time window (gauges_module.f90 gauge_type),
// the grid it samples (setbestsrc: the finest grid containing it) with that
// grid's lower corner including the ghost cells (xlow, ylow: advanc.f:229-
// 230) and mesh widths, and the time of its last record.
struct GaugeDev {
    double x, y, t_start, t_end, min_dt, last_time;
    double xlow, ylow, hx, hy;
    int32_t level, grid;        // grid: index within its level, -1 none
};

struct GaugeArgs {
    GaugeDev* gauges;
    int32_t n;
    int32_t level;
    double tgrid;               // rnode(timemult) of the level's grids
    const PatchDesc* patches;
    const void* q;              // the level's values at tgrid, ghost cells filled
    int32_t nv;
    // records: meta[2 r] = gauge, meta[2 r + 1] = level; vals[r (nv + 1)] =
    // tgrid then the nv interpolated components
    int32_t* meta;
    double* vals;
    unsigned int* count;        // [0] records written, [1] records refused (full or out of range)
    int32_t cap;
};

}  // namespace amrgpu
