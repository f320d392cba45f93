// level.hpp -- device-side descriptors shared by the kernels and the host.
#pragma once
#include <cstdint>
#include "riemann.hpp"
#include "types.hpp"

namespace amrgpu {

// Method switches of the reference (amr_module: method(2), method(3),
// mthlim(:), use_fwaves) and the physics plugin constants.
struct MethodArgs {
    Physics ph;
    int32_t order;        // method(2)
    int32_t transverse;   // method(3)
    int32_t use_fwaves;
    int32_t mwaves;
    int32_t mthlim[NWMAX];
};

}  // namespace amrgpu
