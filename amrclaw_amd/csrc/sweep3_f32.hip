// FP32 instantiations of the 3-D step: wave and flux arithmetic in FP32 with
// hardware reciprocal and square root.
#include "sweep3_launch.hpp"
#include "sweep3_tile.hpp"

namespace amrgpu {

hipError_t launch_sweep3_f32(int solver, const Sweep3Args& a, hipStream_t st) {
    return sweep3_launch<float, FastMathF32>(solver, a, st);
}

hipError_t launch_update3_f32(int nv, float* q, const float* face, const Patch3* patches,
                              const Block3* blocks, int nblocks, double dtdx, double dtdy,
                              double dtdz, hipStream_t st) {
    return update3_launch<float>(nv, q, face, patches, blocks, nblocks, dtdx, dtdy, dtdz, st);
}

template <class Sol>
static hipError_t tile3_go(const Tile3Args& a, size_t lds, hipStream_t st) {
    static size_t attr_lds = 0;
    if (lds > attr_lds) {
        hipError_t e = hipFuncSetAttribute((const void*)sweep3_tile_kernel<Sol>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr_lds = lds;
    }
    if (a.ntiles == 0) return hipSuccess;
    int most = 0;
    for (int x = 0; x < TILE3_XCDS; x++) most = a.xcd_count[x] > most ? a.xcd_count[x] : most;
    hipLaunchKernelGGL((sweep3_tile_kernel<Sol>), dim3(most * TILE3_XCDS), dim3(TILE3_LANES), lds,
                       st, a);
    return hipGetLastError();
}

hipError_t launch_sweep3_tile_f32(int solver, const Tile3Args& a, size_t lds, hipStream_t st) {
    switch (solver) {
    case SOLVER_ADVECTION: return tile3_go<Advection3<float, FastMathF32>>(a, lds, st);
    case SOLVER_EULER: return tile3_go<Euler3<float, FastMathF32>>(a, lds, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace amrgpu
