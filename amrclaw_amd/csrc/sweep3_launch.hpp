// sweep3_launch.hpp -- launchers of the 3-D step kernels for one compute
// precision T with arithmetic policy Mth (included by sweep3_f64.hip and
// sweep3_f32.hip, which differ in their compile flags).
#pragma once
#include "kernels3.hpp"
#include "sweep3.hpp"

namespace amrgpu {

template <class Sol, typename T, int D>
static hipError_t sweep3_go_d(const Sweep3Args& a, hipStream_t st) {
    constexpr size_t lds = sweep3_lds_bytes<Sol::NV, Sol::NW, T>();
    static_assert(lds <= 160 * 1024, "exchange buffer exceeds the LDS of a CU");
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)sweep3_phase_kernel<Sol, T, D>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    if (a.nblocks == 0) return hipSuccess;
    hipLaunchKernelGGL((sweep3_phase_kernel<Sol, T, D>), dim3(a.nblocks), dim3(SWEEP3_LANES), lds,
                       st, a);
    return hipGetLastError();
}

template <class Sol, typename T>
static hipError_t sweep3_go(const Sweep3Args& a, hipStream_t st) {
    switch (a.phase) {
    case 1: return sweep3_go_d<Sol, T, 1>(a, st);
    case 2: return sweep3_go_d<Sol, T, 2>(a, st);
    case 3: return sweep3_go_d<Sol, T, 3>(a, st);
    default: return hipErrorInvalidValue;
    }
}

template <typename T, class Mth>
static hipError_t sweep3_launch(int solver, const Sweep3Args& a, hipStream_t st) {
    switch (solver) {
    case SOLVER_ADVECTION: return sweep3_go<Advection3<T, Mth>, T>(a, st);
    case SOLVER_EULER: return sweep3_go<Euler3<T, Mth>, T>(a, st);
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
static hipError_t update3_launch(int nv, T* q, const T* face, const Patch3* patches,
                                 const Block3* blocks, int nblocks, double dtdx, double dtdy,
                                 double dtdz, hipStream_t st) {
    if (nblocks == 0) return hipSuccess;
    dim3 grid(nblocks), block(256);
    switch (nv) {
    case 1:
        hipLaunchKernelGGL((update3_kernel<1, T>), grid, block, 0, st, q, face, patches, blocks,
                           dtdx, dtdy, dtdz);
        break;
    case 5:
        hipLaunchKernelGGL((update3_kernel<5, T>), grid, block, 0, st, q, face, patches, blocks,
                           dtdx, dtdy, dtdz);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace amrgpu
