// kernels.hpp -- host-side launchers of the device kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "sweep.hpp"

namespace amrgpu {
hipError_t launch_sweep_f64(int solver, const SweepArgs& a, size_t lds, hipStream_t st);
hipError_t launch_sweep_f32(int solver, const SweepArgs& a, size_t lds, hipStream_t st);
}  // namespace amrgpu
