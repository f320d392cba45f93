// regs_plan.hpp -- host planning of the coarse-fine registers of a hierarchy.
#pragma once
#include <string>
#include <vector>

#include "fill_plan.hpp"
#include "regs.hpp"

namespace amrgpu {

struct RegLevel {
    // fine side (level > 1): register layout, prepf.f:33-48
    std::vector<int64_t> reg_off;
    std::vector<int32_t> lenbc;
    int64_t reg_elems = 0;
    // coarse side (level has a finer level): listbc of every grid in the
    // reference's order (prepc.f:47-154), entries with fluxsv duplicates
    // (the same kid slot reached twice) marked
    std::vector<SvEntry> sv;
    std::vector<int32_t> sv_grid_start;   // ngrids + 1
    std::vector<unsigned char> sv_write;  // 1: this entry writes its slot in fluxsv
    std::vector<UpbndCell> ucells;
    std::vector<UpbndEntry> uents;
    std::vector<AvgBlock> avg;
    std::vector<int64_t> avg_start;       // nblocks + 1 cell offsets
    // fine side: qad work items
    std::vector<QadItem> qad;
};

bool plan_registers(const HierGeom& H, int nvar, RegLevel* R, std::string& err);


}  // namespace amrgpu
