// regs.hip -- coarse-fine register kernels: qad + fluxad (the fine side of
// the registers), upbnd (reflux into coarse cells), the averaging of fine
// cells onto the coarse cells they cover and the conservation total.  FP64 arithmetic in the
// reference's order; built with -ffp-contract=off.
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
        typename Sol::Cache cache = Sol::prep(ph, ixy, ql, qr);
        double wave[NW][NV], s[NW], amdq[NV], apdq[NV];
        Sol::rpn2(ph, ixy, ql, qr, nullptr, nullptr, cache, wave, s, amdq, apdq);
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

// upbnd.f:62-124: one coarse cell per thread, its register slots in list order
template <typename S>
__global__ __launch_bounds__(256) void upbnd_kernel(UpdateArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ncells) return;
    const UpbndCell c = a.cells[t];
    const PatchDesc pd = a.patches[c.grid];
    S* q = static_cast<S*>(a.q) + pd.off;
    const int64_t li = (int64_t)c.j * pd.mitot + c.i;
    for (int m = 0; m < a.nv; m++) {
        double v = (double)q[(int64_t)m * pd.plane + li];
        for (int e = c.e0; e < c.e1; e++) {
            const UpbndEntry en = a.entries[e];
            const double* svd = a.reg + a.fine_patches[en.kid].reg_off + (int64_t)a.nv * (en.lkid - 1);
            v = v + en.sgnm * svd[m] / a.area;
        }
        q[(int64_t)m * pd.plane + li] = (S)v;
    }
}

// update.f:95-177: coarse cell = mean of its r x r fine cells
template <typename S>
__global__ __launch_bounds__(256) void average_kernel(UpdateArgs a, const int64_t* block_start) {
    // one thread per coarse cell; block_start[b] = first cell of block b
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= block_start[a.nblocks]) return;
    int lo = 0, hi = a.nblocks - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (block_start[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const AvgBlock b = a.blocks[lo];
    const int64_t c = t - block_start[lo];
    const int ci = (int)(c / b.nj), cj = (int)(c - (int64_t)ci * b.nj);  // i-major like update.f:122-123
    const PatchDesc pc = a.patches[b.grid];
    const PatchDesc pf = a.fine_patches[b.kid];
    S* q = static_cast<S*>(a.q) + pc.off;
    const S* qf = static_cast<const S*>(a.qf) + pf.off;
    const int64_t lic = (int64_t)(b.cj0 + cj) * pc.mitot + (b.ci0 + ci);
    const int fi = b.fi0 + ci * a.ratx, fj = b.fj0 + cj * a.raty;
    const double totrat = (double)(a.ratx * a.raty);
    for (int m = 0; m < a.nv; m++) {
        double v = 0.0;
        for (int jco = 0; jco < a.raty; jco++)
            for (int ico = 0; ico < a.ratx; ico++)
                v = v + (double)qf[(int64_t)m * pf.plane + (int64_t)(fj + jco) * pf.mitot + (fi + ico)];
        q[(int64_t)m * pc.plane + lic] = (S)(v / totrat);
    }
}

// a lane's double, read by the whole wave (lane index uniform)
__device__ __forceinline__ double lane_value(double v, int k) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// conck.f:36-63 in the reference's order: one wave walks the grids in level
// order, rows j, then columns i; each lane loads one cell of a 64-cell chunk
// of the row (coalesced) and the chunk is added lane by lane in ascending i
// by a loop uniform over the wave, so the total is the reference's
// sequential FP64 sum bit for bit.  A diagnostic: one wave, latency-bound.
template <typename S>
__global__ __launch_bounds__(64) void conck_kernel(const PatchDesc* patches, int npatch,
                                                   const void* qv, double* out) {
    const int lane = threadIdx.x;
    double tot = 0.0;
    for (int p = 0; p < npatch; p++) {
        const PatchDesc pd = patches[p];
        const S* q = static_cast<const S*>(qv) + pd.off;
        for (int j = 0; j < pd.my; j++) {
            const S* row = q + (int64_t)(j + NGHOST) * pd.mitot + NGHOST;
            for (int i0 = 0; i0 < pd.mx; i0 += 64) {
                const int n = min(64, pd.mx - i0);
                const double v = lane < n ? (double)row[i0 + lane] : 0.0;
                for (int k = 0; k < n; k++) tot = tot + lane_value(v, k);
            }
        }
    }
    if (lane == 0) out[0] = tot;
}

hipError_t launch_conck(int esize, const PatchDesc* patches, int npatch, const void* q,
                        double* out, hipStream_t st) {
    if (npatch <= 0) return hipSuccess;
    if (esize == 8)
        hipLaunchKernelGGL(conck_kernel<double>, dim3(1), dim3(64), 0, st, patches, npatch, q, out);
    else
        hipLaunchKernelGGL(conck_kernel<float>, dim3(1), dim3(64), 0, st, patches, npatch, q, out);
    return hipGetLastError();
}

template <class Sol>
static void finereg_launch(int esize, dim3 grid, const QadArgs& a, const Physics& ph, hipStream_t st) {
    if (esize == 8)
        hipLaunchKernelGGL((finereg_kernel<Sol, double>), grid, dim3(256), 0, st, a, ph);
    else
        hipLaunchKernelGGL((finereg_kernel<Sol, float>), grid, dim3(256), 0, st, a, ph);
}

hipError_t launch_finereg(int solver, int esize, const QadArgs& a, const Physics& ph, hipStream_t st) {
    if (a.nitems <= 0) return hipSuccess;
    dim3 grid((a.nitems + 255) / 256);
    switch (solver) {
    case SOLVER_ADVECTION: finereg_launch<Advection<double>>(esize, grid, a, ph, st); break;
    case SOLVER_ACOUSTICS: finereg_launch<Acoustics<double>>(esize, grid, a, ph, st); break;
    case SOLVER_EULER: finereg_launch<EulerRoe<double>>(esize, grid, a, ph, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_update(int esize, const UpdateArgs& a, const int64_t* d_block_start,
                         int64_t navg_cells, hipStream_t st) {
    if (a.ncells > 0) {
        dim3 grid((a.ncells + 255) / 256), block(256);
        if (esize == 8)
            hipLaunchKernelGGL(upbnd_kernel<double>, grid, block, 0, st, a);
        else
            hipLaunchKernelGGL(upbnd_kernel<float>, grid, block, 0, st, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (a.nblocks > 0 && navg_cells > 0) {
        dim3 grid((unsigned)((navg_cells + 255) / 256)), block(256);
        if (esize == 8)
            hipLaunchKernelGGL(average_kernel<double>, grid, block, 0, st, a, d_block_start);
        else
            hipLaunchKernelGGL(average_kernel<float>, grid, block, 0, st, a, d_block_start);
    }
    return hipGetLastError();
}

}  // namespace amrgpu
