// flag3.hip -- 3-D cell flagging and mass total over the grids of a batch.
//
// flag3_kernel: flag2refine.f (3-D):56-88 as spest.f:71-84 calls it, one
// thread per interior cell: the largest centred difference of any component
// in any direction against tolsp, DOFLAG (2) or DONTFLAG (0) into the grid's
// FP64 flag array (mx+4, my+4, mz+4), whose ghost border stays goodpt (0).
// conck3_kernel: conck.f (3-D):47-60, the sum of q(1) over a block of 256
// interior cells of one grid (one partial per block; the host adds them).
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

template <typename T>
__global__ __launch_bounds__(256) void conck3_kernel(const T* q, const Patch3* patches,
                                                     const Block3* blocks, double* partial) {
    __shared__ double red[256];
    const Block3 blk = blocks[blockIdx.x];
    const int c = blk.pb + (int)threadIdx.x;
    double v = 0.0;
    if (c < blk.pe) {
        const Patch3 pt = patches[blk.patch];
        const int mx = pt.mx, my = pt.my;
        const int i = c % mx + 1, t = c / mx, j = t % my + 1, k = t / my + 1;
        const int64_t mit = mx + 4, mij = mit * (my + 4);
        v = double(q[pt.qoff + (int64_t)(k + 1) * mij + (j + 1) * mit + (i + 1)]);
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
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

hipError_t launch_conck3(int esize, const void* q, const Patch3* patches, const Block3* blocks,
                         int nblocks, double* partial, hipStream_t st) {
    if (nblocks == 0) return hipSuccess;
    dim3 grid(nblocks), block(256);
    if (esize == 8)
        hipLaunchKernelGGL((conck3_kernel<double>), grid, block, 0, st,
                           static_cast<const double*>(q), patches, blocks, partial);
    else
        hipLaunchKernelGGL((conck3_kernel<float>), grid, block, 0, st,
                           static_cast<const float*>(q), patches, blocks, partial);
    return hipGetLastError();
}

}  // namespace amrgpu
