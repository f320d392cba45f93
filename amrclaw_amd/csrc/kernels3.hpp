// kernels3.hpp -- host-side launchers of the 3-D step kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "sweep3.hpp"
#include "sweep3_tile.hpp"

namespace amrgpu {
// one sweep phase (a.phase = 1, 2, 3) over a.blocks
hipError_t launch_sweep3_f64(int solver, const Sweep3Args& a, hipStream_t st);
hipError_t launch_sweep3_f32(int solver, const Sweep3Args& a, hipStream_t st);
// the fused FP32 step over a tile list (every face on chip): qin -> qout
hipError_t launch_sweep3_tile_f32(int solver, const Tile3Args& a, size_t lds, hipStream_t st);
// conservative update of the cells listed in blocks (256 per block)
hipError_t launch_update3_f64(int nv, double* q, const double* face, const Patch3* patches,
                              const Block3* blocks, int nblocks, double dtdx, double dtdy,
                              double dtdz, hipStream_t st);
hipError_t launch_update3_f32(int nv, float* q, const float* face, const Patch3* patches,
                              const Block3* blocks, int nblocks, double dtdx, double dtdy,
                              double dtdz, hipStream_t st);
// 3-D flag2refine (interior cells of the blocks' grids) and conck partials
hipError_t launch_flag3(int esize, int nv, const void* q, double* flags, const int64_t* foff,
                        const Patch3* patches, const Block3* blocks, int nblocks, double tolsp,
                        unsigned long long* nflag, hipStream_t st);
hipError_t launch_conck3(int esize, const void* q, const Patch3* patches, int npatch, double* out,
                         hipStream_t st);
}  // namespace amrgpu
