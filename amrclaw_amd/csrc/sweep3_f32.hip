// FP32 instantiations of the 3-D step: wave and flux arithmetic in FP32 with
// hardware reciprocal and square root.
#include "sweep3_launch.hpp"

namespace amrgpu {

hipError_t launch_sweep3_f32(int solver, const Sweep3Args& a, hipStream_t st) {
    return sweep3_launch<float, FastMathF32>(solver, a, st);
}

hipError_t launch_update3_f32(int nv, float* q, const float* face, const Patch3* patches,
                              const Block3* blocks, int nblocks, double dtdx, double dtdy,
                              double dtdz, hipStream_t st) {
    return update3_launch<float>(nv, q, face, patches, blocks, nblocks, dtdx, dtdy, dtdz, st);
}

}  // namespace amrgpu
