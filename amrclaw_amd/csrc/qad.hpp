// qad.hpp -- the fine-side register kernel of a level step (qad + fluxad),
// templated on the Riemann solver: the library instantiates it for its own
// solvers (regs.hip) and a plug-in solver's translation unit for its solver
// (include/amrgpu_solver.hpp).  FP64 arithmetic in the reference's order:
// build with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include "regs.hpp"
#include "sweep.hpp"

namespace amrgpu {

// The fine-side register work of one level step, one register slot per
// thread, in the reference's order for that slot: qad.f:9-360 (the r Riemann
// problems between the saved coarse value and the fine grid's first ghost
// cells along one coarse segment, before the step) and then fluxad.f:36-92
// (the r fine fluxes of the step through that segment).  `q` is the level's
// time-t array with its ghost cells; the fluxes come from the perimeter
// block the sweep wrote.
template <class Sol, typename S>
__global__ __launch_bounds__(256) void finereg_kernel(QadArgs a, Physics ph) {
    constexpr int NV = Sol::NV, NW = Sol::NW;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nitems) return;
    const QadItem it = a.items[t];
    const PatchDesc pd = a.patches[it.grid];
    const S* q = static_cast<const S*>(a.q) + pd.off;
    double* sv = a.reg + pd.reg_off;
    const double* qc = sv + (int64_t)NV * pd.lenbc;
    const double* per = a.reg + perim_offset(pd, NV);
    const int r = (it.side == 1 || it.side == 3) ? a.raty : a.ratx;
    const double h = (it.side == 1 || it.side == 3) ? a.hy : a.hx;
    const int ixy = (it.side == 1 || it.side == 3) ? 1 : 2;
    double acc[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) acc[m] = sv[(int64_t)(it.slot - 1) * NV + m];
    double qcv[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) qcv[m] = qc[(int64_t)(it.slot - 1) * NV + m];
    for (int l = 1; l <= (a.flux_form ? 0 : r); l++) {
        const int f = (it.seg - 1) * r + l;        // fine row (sides 1,3) or column (2,4)
        int I, J;
        if (it.side == 1) { I = 0; J = f; }
        else if (it.side == 2) { I = f; J = pd.my + 1; }
        else if (it.side == 3) { I = pd.mx + 1; J = f; }
        else { I = f; J = 0; }
        const int64_t li = (int64_t)(J + 1) * pd.mitot + (I + 1);
        double qf[NV];
#pragma unroll
        for (int m = 0; m < NV; m++) qf[m] = (double)q[(int64_t)m * pd.plane + li];
        // interface states: coarse is left on sides 1 and 4, right on 2 and 3
        const double* ql = (it.side == 1 || it.side == 4) ? qcv : qf;
        const double* qr = (it.side == 1 || it.side == 4) ? qf : qcv;
        // aux of the two cells (qad.f:81-91, 160-185, 228-255, 300-330): the
        // coarse ring's value on the coarse side, on the fine side the ghost
        // cell's, or for an xleft / yleft component the fine face between
        // them (sides 1 and 4 take the first interior cell's; on sides 2
        // and 3 the face value is the ghost cell's, and it replaces the
        // coarse ring's value)
        constexpr int SNA = sol_naux<Sol>::value, NA = SNA > 0 ? SNA : 1;
        double al[NA], ar[NA];
        if constexpr (SNA > 0) {
            const double* fa = a.aux + pd.off * a.auxk;
            const double* ac = a.reg + aux_ring_offset(pd, NV) + (int64_t)(it.slot - 1) * a.naux;
            const int64_t lin = (int64_t)(J + 1) * pd.mitot + (I + 1);
            // the face cell on sides 1 (I = 1) and 4 (J = 1)
            const int64_t lface = it.side == 1 ? lin + 1 : (it.side == 4 ? lin + pd.mitot : lin);
#pragma unroll
            for (int c = 0; c < SNA; c++) {
                const double vf = fa[(int64_t)c * pd.plane + lin];
                const double vface = fa[(int64_t)c * pd.plane + lface];
                const bool face = (a.auxtype[c] == 1 && (it.side == 1 || it.side == 3)) ||
                                  (a.auxtype[c] == 2 && (it.side == 2 || it.side == 4));
                if (it.side == 1 || it.side == 4) {
                    al[c] = ac[c];
                    ar[c] = face ? vface : vf;
                } else {
                    al[c] = vf;
                    ar[c] = face ? vf : ac[c];
                }
            }
        }
        typename Sol::Cache cache = Sol::prep(ph, ixy, ql, qr);
        double wave[NW][NV], s[NW], amdq[NV], apdq[NV];
        Sol::rpn2(ph, ixy, ql, qr, SNA > 0 ? al : nullptr, SNA > 0 ? ar : nullptr, cache, wave, s,
                  amdq, apdq);
#pragma unroll
        for (int m = 0; m < NV; m++) {
            if (it.side == 1 || it.side == 4)
                acc[m] = acc[m] + amdq[m] * h * a.delt + apdq[m] * h * a.delt;
            else
                acc[m] = acc[m] - amdq[m] * h * a.delt - apdq[m] * h * a.delt;
        }
    }
    // fluxad: perimeter side blocks 0 left, 1 top, 2 right, 3 bottom
    const int64_t side_off[4] = {0, (int64_t)NV * pd.my, (int64_t)NV * (pd.my + pd.mx),
                                 (int64_t)NV * (2 * pd.my + pd.mx)};
    const double* ps = per + side_off[it.side - 1];
    for (int l = 1; l <= r; l++) {
        const int f = (it.seg - 1) * r + l;
#pragma unroll
        for (int m = 0; m < NV; m++) acc[m] = acc[m] + ps[(int64_t)(f - 1) * NV + m] * a.delt * h;
    }
#pragma unroll
    for (int m = 0; m < NV; m++) sv[(int64_t)(it.slot - 1) * NV + m] = acc[m];
}

template <class Sol>
hipError_t finereg_launch(int esize, const QadArgs& a, const Physics& ph, hipStream_t st) {
    if (a.nitems <= 0) return hipSuccess;
    const dim3 grid((a.nitems + 255) / 256);
    if (esize == 8)
        hipLaunchKernelGGL((finereg_kernel<Sol, double>), grid, dim3(256), 0, st, a, ph);
    else
        hipLaunchKernelGGL((finereg_kernel<Sol, float>), grid, dim3(256), 0, st, a, ph);
    return hipGetLastError();
}

}  // namespace amrgpu
