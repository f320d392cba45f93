// flag3.hip -- 3-D cell flagging and mass total over the grids of a batch.
//
// flag3_kernel: flag2refine.f (3-D):56-88 as spest.f:71-84 calls it, one
// thread per interior cell: the largest centred difference of any component
// in any direction against tolsp, DOFLAG (2) or DONTFLAG (0) into the grid's
// FP64 flag array (mx+4, my+4, mz+4), whose ghost border stays goodpt (0).
// conck3_kernel: conck.f (3-D):47-60, the sum of q(1) over the interiors of
// the grids in the reference's sequential order (one wave).
#include "kernels3.hpp"

namespace amrgpu {

template <int NV, typename T>
__global__ __launch_bounds__(256) void flag3_kernel(const T* q, double* flags, const int64_t* foff,
                                                    const Patch3* patches, const Block3* blocks,
                                                    double tolsp, unsigned long long* nflag) {
    const Block3 blk = blocks[blockIdx.x];
    const int c = blk.pb + (int)threadIdx.x;
    bool hit = false;
    if (c < blk.pe) {
        const Patch3 pt = patches[blk.patch];
        const int mx = pt.mx, my = pt.my;
        const int i = c % mx + 1, t = c / mx, j = t % my + 1, k = t / my + 1;
        const int64_t mit = mx + 4, mij = mit * (my + 4), plane = mij * (pt.mz + 4);
        const int64_t lc = (int64_t)(k + 1) * mij + (j + 1) * mit + (i + 1);
        const T* qc = q + pt.qoff + lc;
        double dq = 0.0;
#pragma unroll
        for (int m = 0; m < NV; m++) {
            const T* qm = qc + m * plane;
            const double dqi = fabs(double(qm[1]) - double(qm[-1]));
            const double dqj = fabs(double(qm[mit]) - double(qm[-mit]));
            const double dqk = fabs(double(qm[mij]) - double(qm[-mij]));
            dq = fmax(fmax(fmax(dq, dqi), dqj), dqk);
        }
        hit = dq > tolsp;
        flags[foff[blk.patch] + lc] = hit ? 2.0 : 0.0;
    }
    // one atomic per wave
    const unsigned long long b = __ballot(hit);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(nflag, (unsigned long long)__popcll(b));
}

// a lane's double, read by the whole wave (lane index uniform)
__device__ __forceinline__ double lane_value3(double v, int k) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// conck.f (3-D):47-60 in the reference's order: one wave walks the grids in
// order, planes k, rows j, columns i in 64-cell chunks loaded by the lanes
// and added lane by lane: the sequential FP64 sum bit for bit
template <typename T>
__global__ __launch_bounds__(64) void conck3_kernel(const T* q, const Patch3* patches, int npatch,
                                                    double* out) {
    const int lane = threadIdx.x;
    double tot = 0.0;
    for (int p = 0; p < npatch; p++) {
        const Patch3 pt = patches[p];
        const int64_t mit = pt.mx + 4, mij = mit * (pt.my + 4);
        for (int k = 0; k < pt.mz; k++)
            for (int j = 0; j < pt.my; j++) {
                const T* row = q + pt.qoff + (int64_t)(k + 2) * mij + (int64_t)(j + 2) * mit + 2;
                for (int i0 = 0; i0 < pt.mx; i0 += 64) {
                    const int n = min(64, pt.mx - i0);
                    const double v = lane < n ? double(row[i0 + lane]) : 0.0;
                    for (int c = 0; c < n; c++) tot = tot + lane_value3(v, c);
                }
            }
    }
    if (lane == 0) out[0] = tot;
}

template <typename T>
static hipError_t flag3_go(int nv, const T* q, double* flags, const int64_t* foff,
                           const Patch3* patches, const Block3* blocks, int nblocks, double tolsp,
                           unsigned long long* nflag, hipStream_t st) {
    if (nblocks == 0) return hipSuccess;
    dim3 grid(nblocks), block(256);
    switch (nv) {
    case 1:
        hipLaunchKernelGGL((flag3_kernel<1, T>), grid, block, 0, st, q, flags, foff, patches,
                           blocks, tolsp, nflag);
        break;
    case 5:
        hipLaunchKernelGGL((flag3_kernel<5, T>), grid, block, 0, st, q, flags, foff, patches,
                           blocks, tolsp, nflag);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_flag3(int esize, int nv, const void* q, double* flags, const int64_t* foff,
                        const Patch3* patches, const Block3* blocks, int nblocks, double tolsp,
                        unsigned long long* nflag, hipStream_t st) {
    return esize == 8 ? flag3_go(nv, static_cast<const double*>(q), flags, foff, patches, blocks,
                                 nblocks, tolsp, nflag, st)
                      : flag3_go(nv, static_cast<const float*>(q), flags, foff, patches, blocks,
                                 nblocks, tolsp, nflag, st);
}

hipError_t launch_conck3(int esize, const void* q, const Patch3* patches, int npatch, double* out,
                         hipStream_t st) {
    if (npatch <= 0) return hipSuccess;
    if (esize == 8)
        hipLaunchKernelGGL((conck3_kernel<double>), dim3(1), dim3(64), 0, st,
                           static_cast<const double*>(q), patches, npatch, out);
    else
        hipLaunchKernelGGL((conck3_kernel<float>), dim3(1), dim3(64), 0, st,
                           static_cast<const float*>(q), patches, npatch, out);
    return hipGetLastError();
}

}  // namespace amrgpu
