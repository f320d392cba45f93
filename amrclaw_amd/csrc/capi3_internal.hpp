// capi3_internal.hpp -- state of the 3-D library (capi3.cpp) shared with the
// 3-D hierarchy code (capi3_amr.cpp).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>

#include "../../include/amrgpu3.h"
#include "sweep3.hpp"

namespace amrgpu {

// device arrays of one step over a batch of grids
struct Step3IO {
    void* q;                      // solutions (compute precision)
    void* face;                   // flux blocks
    const Patch3* d_patches;
    Block3* d_blocks[4];          // x, y, z sweeps, update
    int nblocks[4];
    int npatch;
    unsigned long long* d_cfl;    // per-grid Courant number (double bits)
};

bool ctx3_inited();
const amrgpu3_config& ctx3_cfg();
size_t ctx3_esize();
hipStream_t ctx3_stream();
int ctx3_plane_block();
int ctx3_fail(int code, const char* fmt, ...);
hipError_t ctx3_h2d(void* dst, const void* src, size_t n);
hipError_t ctx3_d2h(void* dst, const void* src, size_t n);
// stepgrid.f (3-D) for every grid of the batch: clears d_cfl, launches the
// x-, y- and z-sweeps and the conservative update; esize 8 runs the FP64
// kernels whatever the library's precision (0: the library's)
int ctx3_step(const Step3IO& io, double dx, double dy, double dz, double dt, int esize = 0);
// the work lists of a batch: per sweep (grid, plane block), largest first;
// update and totals (grid, 256 cells)
void ctx3_plan_blocks(const std::vector<Patch3>& patches, int plane_block,
                      std::vector<Block3> out[4]);
// releases the hierarchy state of capi3_amr.cpp (amrgpu3_finalize)
void amr3_release();

}  // namespace amrgpu
