// FP32 instantiations of the level sweep: wave and flux arithmetic in FP32
// with hardware reciprocal and square root.
#include "sweep.hpp"
#include "kernels.hpp"

namespace amrgpu {

hipError_t launch_sweep_f32(int solver, const SweepArgs& a, size_t lds, hipStream_t st) {
    dim3 grid(a.ntiles), block(256);
    static bool attr_set = false;
    if (!attr_set) {
        // allow tiles whose LDS footprint exceeds the 64 KiB default
        const int max_lds = 160 * 1024;
        hipFuncSetAttribute((const void*)sweep_kernel<Advection<float, FastMathF32>, float, float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        hipFuncSetAttribute((const void*)sweep_kernel<Acoustics<float, FastMathF32>, float, float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        hipFuncSetAttribute((const void*)sweep_kernel<EulerRoe<float, FastMathF32>, float, float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        attr_set = true;
    }
    if (a.ntiles == 0) return hipSuccess;   // (a call with no tiles only sets the attributes)
    switch (solver) {
    case SOLVER_ADVECTION:
        hipLaunchKernelGGL((sweep_kernel<Advection<float, FastMathF32>, float, float>), grid, block, lds, st, a);
        break;
    case SOLVER_ACOUSTICS:
        hipLaunchKernelGGL((sweep_kernel<Acoustics<float, FastMathF32>, float, float>), grid, block, lds, st, a);
        break;
    case SOLVER_EULER:
        hipLaunchKernelGGL((sweep_kernel<EulerRoe<float, FastMathF32>, float, float>), grid, block, lds, st, a);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Mixed path: FP32 wave arithmetic on FP64 storage, flux form, FP64 update.
hipError_t launch_sweep_mixed(int solver, const SweepArgs& a, size_t lds, hipStream_t st) {
    dim3 grid(a.ntiles), block(256);
    static bool attr_set = false;
    if (!attr_set) {
        const int max_lds = 160 * 1024;
        hipFuncSetAttribute((const void*)sweep_kernel<Advection<float, FastMathF32>, float, double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        hipFuncSetAttribute((const void*)sweep_kernel<Acoustics<float, FastMathF32>, float, double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        hipFuncSetAttribute((const void*)sweep_kernel<EulerRoe<float, FastMathF32>, float, double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        attr_set = true;
    }
    if (a.ntiles == 0) return hipSuccess;
    switch (solver) {
    case SOLVER_ADVECTION:
        hipLaunchKernelGGL((sweep_kernel<Advection<float, FastMathF32>, float, double>), grid, block, lds, st, a);
        break;
    case SOLVER_ACOUSTICS:
        hipLaunchKernelGGL((sweep_kernel<Acoustics<float, FastMathF32>, float, double>), grid, block, lds, st, a);
        break;
    case SOLVER_EULER:
        hipLaunchKernelGGL((sweep_kernel<EulerRoe<float, FastMathF32>, float, double>), grid, block, lds, st, a);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace amrgpu
