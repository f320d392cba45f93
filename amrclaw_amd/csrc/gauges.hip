// gauges.hip -- update_gauges (gauges_module.f90:313-452) for the gauges
// whose best source grid lies on the level being advanced: bilinear
// interpolation of the grid's time-t values (ghost cells filled), with the
// reference's index arithmetic and evaluation order (built without
// contraction, so FP64 records equal the reference's bit for bit).
#include <hip/hip_runtime.h>
#include "gauges.hpp"

namespace amrgpu {

template <typename S>
__global__ void gauge_kernel(GaugeArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    GaugeDev& G = a.gauges[i];
    if (G.level != a.level || G.grid < 0) return;
    const double tgrid = a.tgrid;
    if (tgrid < G.t_start || tgrid > G.t_end) return;              // :345-347
    if (tgrid - G.last_time < G.min_dt) return;                    // :350-352
    const PatchDesc pd = a.patches[G.grid];
    const int mitot = pd.mitot, mjtot = pd.plane / pd.mitot;
    // :356-372: the cell whose centre is just below-left of the gauge
    const int iindex = (int)(.5 + (G.x - G.xlow) / G.hx);
    const int jindex = (int)(.5 + (G.y - G.ylow) / G.hy);
    if (iindex < NGHOST || iindex > mitot - NGHOST || jindex < NGHOST || jindex > mjtot - NGHOST) {
        atomicAdd(&a.count[1], 1u);   // the reference prints an error here
        return;
    }
    const double xcent = G.xlow + (iindex - .5) * G.hx;
    const double ycent = G.ylow + (jindex - .5) * G.hy;
    const double xoff = (G.x - xcent) / G.hx;
    const double yoff = (G.y - ycent) / G.hy;
    const unsigned int r = atomicAdd(&a.count[0], 1u);
    if (r >= (unsigned int)a.cap) {
        atomicAdd(&a.count[1], 1u);
        return;
    }
    const S* q = static_cast<const S*>(a.q) + pd.off;
    // 1-based (iindex, jindex) of q(nvar, mitot, mjtot) at patch-local
    // index (jindex-1)*mitot + (iindex-1)
    const int64_t c00 = (int64_t)(jindex - 1) * mitot + (iindex - 1);
    a.meta[2 * (int64_t)r] = i;
    a.meta[2 * (int64_t)r + 1] = a.level;
    double* v = a.vals + (int64_t)r * (a.nv + 1);
    v[0] = tgrid;
    for (int m = 0; m < a.nv; m++) {
        const S* qm = q + (int64_t)m * pd.plane;
        double var = (1.0 - xoff) * (1.0 - yoff) * double(qm[c00]) +
                     xoff * (1.0 - yoff) * double(qm[c00 + 1]) +
                     (1.0 - xoff) * yoff * double(qm[c00 + mitot]) +
                     xoff * yoff * double(qm[c00 + mitot + 1]);   // :374-388
        if (fabs(var) < 1e-90) var = 0.0;                          // :410-412
        v[1 + m] = var;
    }
    G.last_time = tgrid;                                           // :437
}

hipError_t launch_gauges(int esize, const GaugeArgs& a, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    const dim3 grid((a.n + 63) / 64), block(64);
    if (esize == 8) hipLaunchKernelGGL(gauge_kernel<double>, grid, block, 0, st, a);
    else hipLaunchKernelGGL(gauge_kernel<float>, grid, block, 0, st, a);
    return hipGetLastError();
}

}  // namespace amrgpu
