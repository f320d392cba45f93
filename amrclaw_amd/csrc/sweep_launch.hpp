// sweep_launch.hpp -- launchers of the 2-D sweep for one Riemann solver, one
// compute precision T with arithmetic policy Mth and one storage precision S.
// Included by sweep_f64.hip and sweep_f32.hip (the library's own solvers,
// which differ in their compile flags) and by a plug-in solver's translation
// unit (include/amrgpu_solver.hpp).
#pragma once
#include <algorithm>
#include "kernels.hpp"
#include "sweep.hpp"
#include "sweep_march.hpp"

namespace amrgpu {

template <class Sol, typename T, typename S, int WGS, int LX = 0, int LY = 0, bool STD = false>
static hipError_t sweep_go(const SweepArgs& a, size_t lds, hipStream_t st) {
    static bool attr_set = false;
    if (!attr_set) {
        // allow tiles whose LDS footprint exceeds the 64 KiB default: the
        // CU's 160 KiB less the kernel's static LDS
        const void* fns[2] = {(const void*)sweep_kernel<Sol, T, S, WGS, LX, LY, STD>,
                              (const void*)sweep_persistent_kernel<Sol, T, S, WGS, LX, LY, STD>};
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
        hipLaunchKernelGGL((sweep_kernel<Sol, T, S, WGS, LX, LY, STD>), dim3(a.ntiles), dim3(256),
                           lds, st, a);
        return hipGetLastError();
    }
    // persistent: as many workgroups as fit on the device at this LDS size
    static size_t grid_lds = 0;
    static int grid = 0;
    if (grid == 0 || grid_lds != lds) {
        int per_cu = 0, dev = 0, ncu = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, (const void*)sweep_persistent_kernel<Sol, T, S, WGS, LX, LY, STD>, 256, lds);
        if (e == hipSuccess) e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        grid = std::max(1, per_cu) * std::max(1, ncu);
        grid_lds = lds;
    }
    const int g = std::min(grid, a.ntiles);
    hipLaunchKernelGGL((sweep_persistent_kernel<Sol, T, S, WGS, LX, LY, STD>), dim3(g), dim3(256),
                       lds, st, a);
    return hipGetLastError();
}

// The run-time layout or, when FIXED and every tile of the level fits the
// layout's extents (a.layout: 1 = 28x12 for FP32 arithmetic, 2 = 12x12 for
// FP64), the instantiation with compile-time LDS strides.
template <class Sol, typename T, typename S, int WGS, bool FIXED>
static hipError_t sweep_layout(const SweepArgs& a, size_t lds, hipStream_t st) {
    if constexpr (FIXED) {
        if (sizeof(T) == 4 && a.layout == 1) {
            // the standard method's instantiation too (its attributes are
            // set by the same attributes-only call)
            if (a.ntiles == 0 && !a.waves) {
                hipError_t e = sweep_go<Sol, T, S, WGS, 28, 12, true>(a, lds, st);
                if (e != hipSuccess) return e;
            }
            if (std_method(a.meth)) return sweep_go<Sol, T, S, WGS, 28, 12, true>(a, lds, st);
            return sweep_go<Sol, T, S, WGS, 28, 12>(a, lds, st);
        }
        if (sizeof(T) == 8 && a.layout == 2) return sweep_go<Sol, T, S, WGS, 12, 12>(a, lds, st);
    }
    return sweep_go<Sol, T, S, WGS>(a, lds, st);
}

// The sweep of one solver family SolT<T, Mth>.  Occupancy bound: the mixed
// path (FP32 arithmetic on FP64 storage, one edge array per direction, about
// 20 KB of LDS per tile) comes in 5, 6 and 7 workgroups per CU (0: 7,
// measured best on the batched Euler level); the FP32 path's two edge arrays
// per direction keep it at 5 by LDS; the FP64 path is bounded for 4 (128
// VGPRs for Euler with a 40-byte spill, 3 % faster than its unbounded 137
// VGPRs at 3).  FIXED: instantiate the fixed-layout variants too (the Euler
// solver, whose tuned caps they match).
// The march sweep (a.waves set): four independent waves per workgroup, no
// LDS.  Its register bound: 2 waves per SIMD (256 VGPRs, the default), or 3
// with FP32 arithmetic (wgs == 3: 168 VGPRs, part of the state spilled).
// A level with aux arrays (a capacity or a solver that reads aux): the tile
// sweep with the run-time layout, per-lane aux loads and the capacity
// update (FP64 or FP32 storage; the mixed path has no aux variant).
template <class Sol, typename T, typename S, int WGS>
static hipError_t sweep_aux_go(const SweepArgs& a, size_t lds, hipStream_t st) {
    static bool attr_set = false;
    const void* f = (const void*)sweep_kernel<Sol, T, S, WGS, 0, 0, false, true>;
    if (!attr_set) {
        hipFuncAttributes fa{};
        hipError_t e = hipFuncGetAttributes(&fa, f);
        if (e == hipSuccess)
            e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    160 * 1024 - (int)fa.sharedSizeBytes);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    if (a.ntiles == 0) return hipSuccess;
    hipLaunchKernelGGL((sweep_kernel<Sol, T, S, WGS, 0, 0, false, true>), dim3(a.ntiles), dim3(256),
                       lds, st, a);
    return hipGetLastError();
}

template <class Sol, typename T, typename S, int MWGS, bool STD = false>
static hipError_t march_go(const SweepArgs& a, hipStream_t st) {
    if (a.nwaves == 0) return hipSuccess;
    hipLaunchKernelGGL((sweep_march_kernel<Sol, T, S, MWGS, STD>), dim3((a.nwaves + 3) / 4),
                       dim3(256), 0, st, a);
    return hipGetLastError();
}

template <template <typename, class> class SolT, typename T, class Mth, typename S, bool FIXED>
static hipError_t sweep_solver(const SweepArgs& a, size_t lds, hipStream_t st, int wgs) {
    using Sol = SolT<T, Mth>;
    if constexpr (sizeof(T) == sizeof(S)) {
        if (a.aux) return sweep_aux_go<Sol, T, S, sizeof(T) == 8 ? 4 : 5>(a, lds, st);
        if (a.ntiles == 0 && !a.waves) {   // attributes only: the aux variant's too
            hipError_t e = sweep_aux_go<Sol, T, S, sizeof(T) == 8 ? 4 : 5>(a, lds, st);
            if (e != hipSuccess) return e;
        }
    }
    if constexpr (sol_naux<Sol>::value > 0) {
        // a solver that reads aux runs only in the aux sweep
        (void)lds; (void)st; (void)wgs;
        return a.ntiles == 0 && !a.waves ? hipSuccess : hipErrorInvalidValue;
    } else {
    if (a.aux) return hipErrorInvalidValue;
    if (a.waves) {
        if constexpr (sizeof(T) == 4) {
            if constexpr (FIXED) {
                if (std_method(a.meth))
                    return wgs == 3 ? march_go<Sol, T, S, 3, true>(a, st) : march_go<Sol, T, S, 2, true>(a, st);
            }
            if (wgs == 3) return march_go<Sol, T, S, 3>(a, st);
        }
        return march_go<Sol, T, S, 2>(a, st);
    }
    if constexpr (sizeof(T) == 4 && sizeof(S) == 8) {
        if (wgs == 5) return sweep_layout<Sol, T, S, 5, FIXED>(a, lds, st);
        if (wgs == 6) return sweep_layout<Sol, T, S, 6, FIXED>(a, lds, st);
        return sweep_layout<Sol, T, S, 7, FIXED>(a, lds, st);
    } else if constexpr (sizeof(T) == 4) {
        (void)wgs;
        return sweep_layout<Sol, T, S, 5, FIXED>(a, lds, st);
    } else {
        (void)wgs;
        return sweep_layout<Sol, T, S, 4, FIXED>(a, lds, st);
    }
    }
}

// the library's own solvers, by id
template <typename T, class Mth, typename S>
static hipError_t sweep_dispatch(int solver, const SweepArgs& a, size_t lds, hipStream_t st,
                                 int wgs) {
    switch (solver) {
    case SOLVER_ADVECTION: return sweep_solver<Advection, T, Mth, S, false>(a, lds, st, wgs);
    case SOLVER_ACOUSTICS: return sweep_solver<Acoustics, T, Mth, S, false>(a, lds, st, wgs);
    case SOLVER_EULER: return sweep_solver<EulerRoe, T, Mth, S, true>(a, lds, st, wgs);
    case SOLVER_VCADV: return sweep_solver<VcAdvection, T, Mth, S, false>(a, lds, st, wgs);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace amrgpu
