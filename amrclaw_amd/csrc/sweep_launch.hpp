// sweep_launch.hpp -- launchers of the 2-D sweep for one compute precision T,
// arithmetic policy Mth and storage precision S (included by sweep_f64.hip
// and sweep_f32.hip, which differ in their compile flags).
#pragma once
#include <algorithm>
#include "kernels.hpp"
#include "sweep.hpp"

namespace amrgpu {

template <class Sol, typename T, typename S, int WGS, int LX = 0, int LY = 0>
static hipError_t sweep_go(const SweepArgs& a, size_t lds, hipStream_t st) {
    static bool attr_set = false;
    if (!attr_set) {
        // allow tiles whose LDS footprint exceeds the 64 KiB default: the
        // CU's 160 KiB less the kernel's static LDS
        const void* fns[3] = {(const void*)sweep_kernel<Sol, T, S, WGS>,
                              (const void*)sweep_persistent_kernel<Sol, T, S, WGS>,
                              (const void*)sweep_kernel<Sol, T, S, WGS, LX, LY>};
        for (const void* f : fns) {
            hipFuncAttributes fa{};
            hipError_t e = hipFuncGetAttributes(&fa, f);
            if (e == hipSuccess)
                e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        160 * 1024 - (int)fa.sharedSizeBytes);
            if (e != hipSuccess) return e;
        }
        attr_set = true;
    }
    if (a.ntiles == 0) return hipSuccess;   // (a call with no tiles only sets the attributes)
    if (!a.next) {
        hipLaunchKernelGGL((sweep_kernel<Sol, T, S, WGS, LX, LY>), dim3(a.ntiles), dim3(256), lds,
                           st, a);
        return hipGetLastError();
    }
    // persistent: as many workgroups as fit on the device at this LDS size
    static size_t grid_lds = 0;
    static int grid = 0;
    if (grid == 0 || grid_lds != lds) {
        int per_cu = 0, dev = 0, ncu = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, (const void*)sweep_persistent_kernel<Sol, T, S, WGS>, 256, lds);
        if (e == hipSuccess) e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        grid = std::max(1, per_cu) * std::max(1, ncu);
        grid_lds = lds;
    }
    const int g = std::min(grid, a.ntiles);
    hipLaunchKernelGGL((sweep_persistent_kernel<Sol, T, S, WGS>), dim3(g), dim3(256), lds, st, a);
    return hipGetLastError();
}

template <typename T, class Mth, typename S, int WGS, int LX = 0, int LY = 0>
static hipError_t sweep_dispatch_w(int solver, const SweepArgs& a, size_t lds, hipStream_t st) {
    switch (solver) {
    case SOLVER_ADVECTION: return sweep_go<Advection<T, Mth>, T, S, WGS>(a, lds, st);
    case SOLVER_ACOUSTICS: return sweep_go<Acoustics<T, Mth>, T, S, WGS>(a, lds, st);
    case SOLVER_EULER: return sweep_go<EulerRoe<T, Mth>, T, S, WGS, LX, LY>(a, lds, st);
    default: return hipErrorInvalidValue;
    }
}

// the occupancy bound: the mixed path (FP32 arithmetic on FP64 storage, one
// edge array per direction, about 20 KB of LDS per tile) comes in 5, 6 and 7
// workgroups per CU (0: 7, measured best on the batched Euler level); the
// FP32 path's two edge arrays per direction keep it at 5 by LDS; the FP64
// path is bounded for 4 (128 VGPRs for Euler with a 40-byte spill, 3 %
// faster than its unbounded 137 VGPRs at 3)
template <typename T, class Mth, typename S>
static hipError_t sweep_dispatch(int solver, const SweepArgs& a, size_t lds, hipStream_t st,
                                 int wgs) {
    if constexpr (sizeof(T) == 4 && sizeof(S) == 8) {
        if (wgs == 5) return sweep_dispatch_w<T, Mth, S, 5>(solver, a, lds, st);
        if (wgs == 6) return sweep_dispatch_w<T, Mth, S, 6>(solver, a, lds, st);
        // the tuned tile caps of the Euler level: LDS strides fixed at
        // compile time when every tile fits 28x12 (a.layout == 1)
        if (a.layout == 1) return sweep_dispatch_w<T, Mth, S, 7, 28, 12>(solver, a, lds, st);
        return sweep_dispatch_w<T, Mth, S, 7>(solver, a, lds, st);
    } else if constexpr (sizeof(T) == 4) {
        (void)wgs;
        if (a.layout == 1) return sweep_dispatch_w<T, Mth, S, 5, 28, 12>(solver, a, lds, st);
        return sweep_dispatch_w<T, Mth, S, 5>(solver, a, lds, st);
    } else {
        (void)wgs;
        if (a.layout == 2) return sweep_dispatch_w<T, Mth, S, 4, 12, 12>(solver, a, lds, st);
        return sweep_dispatch_w<T, Mth, S, 4>(solver, a, lds, st);
    }
}

}  // namespace amrgpu
