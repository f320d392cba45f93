// FP64 instantiations of the level sweep.  Built with -ffp-contract=off:
// no multiply-add is fused, so results match the reference bit for bit.
#include <algorithm>
#include "sweep_launch.hpp"

namespace amrgpu {

hipError_t launch_sweep_f64(int solver, const SweepArgs& a, size_t lds, hipStream_t st) {
    return sweep_dispatch<double, ExactMath<double>, double>(solver, a, lds, st, 0);
}

}  // namespace amrgpu
