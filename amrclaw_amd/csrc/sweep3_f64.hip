// FP64 instantiations of the 3-D step: IEEE division and square root, built
// without FMA contraction (bit-for-bit the reference's FP64 arithmetic).
#include "sweep3_launch.hpp"

namespace amrgpu {

hipError_t launch_sweep3_f64(int solver, const Sweep3Args& a, hipStream_t st) {
    return sweep3_launch<double, ExactMath<double>>(solver, a, st);
}

hipError_t launch_update3_f64(int nv, double* q, const double* face, const Patch3* patches,
                              const Block3* blocks, int nblocks, double dtdx, double dtdy,
                              double dtdz, hipStream_t st) {
    return update3_launch<double>(nv, q, face, patches, blocks, nblocks, dtdx, dtdy, dtdz, st);
}

}  // namespace amrgpu
