// regs.hip -- coarse-fine register kernels: qad + fluxad (the fine side of
// the registers), upbnd (reflux into coarse cells), the averaging of fine
// cells onto the coarse cells they cover and the conservation total.  FP64 arithmetic in the
// reference's order; built with -ffp-contract=off.
#include <hip/hip_runtime.h>
#include "qad.hpp"
#include "regs.hpp"
#include "sweep.hpp"

namespace amrgpu {

// upbnd.f:62-124: one coarse cell per thread, its register slots in list order
template <typename S>
__global__ __launch_bounds__(256) void upbnd_kernel(UpdateArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ncells) return;
    const UpbndCell c = a.cells[t];
    const PatchDesc pd = a.patches[c.grid];
    S* q = static_cast<S*>(a.q) + pd.off;
    const int64_t li = (int64_t)c.j * pd.mitot + c.i;
    // capacity: the sign divided by the coarse cell's capacity (upbnd.f:104-108)
    const double capa = a.mcapa > 0 ? a.aux_c[pd.off * a.auxk + (int64_t)(a.mcapa - 1) * pd.plane + li] : 1.0;
    for (int m = 0; m < a.nv; m++) {
        double v = (double)q[(int64_t)m * pd.plane + li];
        for (int e = c.e0; e < c.e1; e++) {
            const UpbndEntry en = a.entries[e];
            const double* svd = a.reg + a.fine_patches[en.kid].reg_off + (int64_t)a.nv * (en.lkid - 1);
            const double sgnm = a.mcapa > 0 ? en.sgnm / capa : en.sgnm;
            v = v + sgnm * svd[m] / a.area;
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
    if (a.mcapa > 0) {
        // update.f:147-161: the capacity-weighted mean
        const double* capf = a.aux_f + pf.off * a.auxk + (int64_t)(a.mcapa - 1) * pf.plane;
        const double capc = a.aux_c[pc.off * a.auxk + (int64_t)(a.mcapa - 1) * pc.plane + lic];
        for (int m = 0; m < a.nv; m++) {
            double v = 0.0;
            for (int jco = 0; jco < a.raty; jco++)
                for (int ico = 0; ico < a.ratx; ico++) {
                    const int64_t lf = (int64_t)(fj + jco) * pf.mitot + (fi + ico);
                    v = v + (double)qf[(int64_t)m * pf.plane + lf] * capf[lf];
                }
            q[(int64_t)m * pc.plane + lic] = (S)(v / (totrat * capc));
        }
        return;
    }
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
                                                   const void* qv, double* out, const double* aux,
                                                   int mcapa, int auxk) {
    const int lane = threadIdx.x;
    double tot = 0.0;
    for (int p = 0; p < npatch; p++) {
        const PatchDesc pd = patches[p];
        const S* q = static_cast<const S*>(qv) + pd.off;
        // capacity: q(1) * capa (conck.f:56-60)
        const double* ca = mcapa > 0 ? aux + pd.off * auxk + (int64_t)(mcapa - 1) * pd.plane : nullptr;
        for (int j = 0; j < pd.my; j++) {
            const S* row = q + (int64_t)(j + NGHOST) * pd.mitot + NGHOST;
            const double* crow = ca ? ca + (int64_t)(j + NGHOST) * pd.mitot + NGHOST : nullptr;
            for (int i0 = 0; i0 < pd.mx; i0 += 64) {
                const int n = min(64, pd.mx - i0);
                const double v = lane < n ? (crow ? (double)row[i0 + lane] * crow[i0 + lane]
                                                  : (double)row[i0 + lane]) : 0.0;
                for (int k = 0; k < n; k++) tot = tot + lane_value(v, k);
            }
        }
    }
    if (lane == 0) out[0] = tot;
}

hipError_t launch_conck(int esize, const PatchDesc* patches, int npatch, const void* q,
                        double* out, hipStream_t st, const double* aux, int mcapa, int auxk) {
    if (npatch <= 0) return hipSuccess;
    if (esize == 8)
        hipLaunchKernelGGL(conck_kernel<double>, dim3(1), dim3(64), 0, st, patches, npatch, q, out,
                           aux, mcapa, auxk);
    else
        hipLaunchKernelGGL(conck_kernel<float>, dim3(1), dim3(64), 0, st, patches, npatch, q, out,
                           aux, mcapa, auxk);
    return hipGetLastError();
}

hipError_t launch_finereg(int solver, int esize, const QadArgs& a, const Physics& ph, hipStream_t st) {
    switch (solver) {
    case SOLVER_ADVECTION: return finereg_launch<Advection<double>>(esize, a, ph, st);
    case SOLVER_ACOUSTICS: return finereg_launch<Acoustics<double>>(esize, a, ph, st);
    case SOLVER_EULER: return finereg_launch<EulerRoe<double>>(esize, a, ph, st);
    case SOLVER_VCADV: return finereg_launch<VcAdvection<double>>(esize, a, ph, st);
    default: return hipErrorInvalidValue;
    }
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
