// FP32 instantiations of the level sweep: wave and flux arithmetic in FP32
// with hardware reciprocal and square root, on FP32 storage (the FP32 path)
// or FP64 storage (the mixed path: flux form, FP64 update).
#include <algorithm>
#include "sweep_launch.hpp"

namespace amrgpu {

hipError_t launch_sweep_f32(int solver, const SweepArgs& a, size_t lds, hipStream_t st) {
    return sweep_dispatch<float, FastMathF32, float>(solver, a, lds, st, 0);
}

hipError_t launch_sweep_mixed(int solver, const SweepArgs& a, size_t lds, hipStream_t st,
                              int wgs) {
    return sweep_dispatch<float, FastMathF32, double>(solver, a, lds, st, wgs);
}

}  // namespace amrgpu
