// flag.hip -- cell flagging of every grid of a level: flagger.f:124-164
// with flagregions2.f90:40-120 and flag2refine2.f90:53-124.  Flags are the
// reference's FP64 values (UNSET -1, DONTFLAG 0, DOFLAG 2) on the grid grown
// by mbuff cells (flagger.f:121-123); the buffer ring stays UNSET.
#include <hip/hip_runtime.h>
#include "flag.hpp"

namespace amrgpu {

// one workgroup per grid; thread-strided over the grid's flag array
template <typename S>
__global__ __launch_bounds__(256) void flag_kernel(FlagArgs a) {
    const int g = blockIdx.x;
    const PatchDesc pd = a.patches[g];
    const S* q = static_cast<const S*>(a.q) + pd.off;
    double* f = a.flags + a.foff[g];
    const int mb = a.mbuff, mib = pd.mx + 2 * mb, n = mib * (pd.my + 2 * mb);
    const RegionRect* rr = a.rects + (int64_t)g * a.nrect;
    int count = 0, bad = 0;
    for (int c = threadIdx.x; c < n; c += 256) {
        const int jb = c / mib, ib = c - jb * mib;
        const int i = ib - mb + 1, j = jb - mb + 1;          // 1-based interior indices
        double v = FLAG_UNSET;
        if (i >= 1 && i <= pd.mx && j >= 1 && j <= pd.my) {
            int minl = 0, maxl = 0;
            for (int r = 0; r < a.nrect; r++)
                if (i >= rr[r].i1 && i <= rr[r].i2 && j >= rr[r].j1 && j <= rr[r].j2) {
                    minl = max(minl, rr[r].min_level);
                    maxl = max(maxl, rr[r].max_level);
                }
            if (minl > maxl) bad = 1;
            if (maxl != 0) {
                if (a.level < minl) v = FLAG_DOFLAG;
                else if (a.level >= maxl) v = FLAG_DONTFLAG;
            }
            if (v == FLAG_UNSET) {
                // q(m, i, j): 0-based with ghosts at (i+1, j+1)
                const int64_t li = (int64_t)(j + 1) * pd.mitot + (i + 1);
                for (int m = 0; m < a.nv; m++) {
                    const S* qm = q + (int64_t)m * pd.plane;
                    const double dqi = fabs((double)qm[li + 1] - (double)qm[li - 1]);
                    const double dqj = fabs((double)qm[li + pd.mitot] - (double)qm[li - pd.mitot]);
                    double dq = 0.0;
                    if (dqi > dq) dq = dqi;
                    if (dqj > dq) dq = dqj;
                    if (dq > a.tolsp) {
                        v = FLAG_DOFLAG;
                        break;
                    }
                }
            }
            count += v == FLAG_DOFLAG;
        }
        f[c] = v;
    }
    // flagged-cell count and region error, one atomic per wave
    for (int off = 32; off >= 1; off >>= 1) {
        count += __shfl_xor(count, off, 64);
        bad |= __shfl_xor(bad, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (count) atomicAdd(a.counters, (unsigned long long)count);
        if (bad) atomicOr(reinterpret_cast<unsigned int*>(a.counters + 1), 1u);
    }
}

// errf1.f:33-137, one coarse cell per thread.  Its 2x2 block of flags is
// read and written by that thread only, so the two passes of errf1 (estimate
// every cell, then mark) merge into one.  As errf1 does, the block of coarse
// cell (i, j) is addressed from flag index nghost + 1 of the
// (mx + 2 mbuff) x (my + 2 mbuff) array.
__global__ __launch_bounds__(256) void errf1_kernel(ErrfArgs a) {
    const int g = blockIdx.x;
    const PatchDesc pc = a.crs[g], pf = a.fine[g];
    const double* rc = a.crs_q + pc.off;             // component 1 plane
    const double* rf = a.fine_q + pf.off;
    double* fl = a.flags + a.foff[g];
    const int mib = pf.mx + 2 * a.mbuff;
    const int n = pc.mx * pc.my;
    int count = 0;
    for (int c = threadIdx.x; c < n; c += 256) {
        const int jc = c / pc.mx, ic = c - jc * pc.mx;
        const int ifine = NGHOST + 1 + 2 * ic, jfine = NGHOST + 1 + 2 * jc;   // 1-based
        double* f00 = fl + (ifine - 1) + (int64_t)mib * (jfine - 1);
        double* f01 = f00 + mib;
        if (!(f00[0] == FLAG_UNSET || f00[1] == FLAG_UNSET || f01[0] == FLAG_UNSET ||
              f01[1] == FLAG_UNSET))
            continue;
        const double* r0 = rf + (int64_t)(jfine - 1) * pf.mitot + (ifine - 1);
        const double* r1 = r0 + pf.mitot;
        const double term1 = r0[0], term2 = r0[1], term3 = r1[1], term4 = r1[0];
        const double aval = (term1 + term2 + term3 + term4) / 4.0;
        const double cv = rc[(int64_t)(jc + NGHOST) * pc.mitot + (ic + NGHOST)];
        const double est = fabs((aval - cv) / a.order);
        if (est >= a.tol) {
            count += (f00[0] != FLAG_DOFLAG) + (f00[1] != FLAG_DOFLAG) +
                     (f01[0] != FLAG_DOFLAG) + (f01[1] != FLAG_DOFLAG);
            f00[0] = FLAG_DOFLAG;
            f00[1] = FLAG_DOFLAG;
            f01[0] = FLAG_DOFLAG;
            f01[1] = FLAG_DOFLAG;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) count += __shfl_xor(count, off, 64);
    if ((threadIdx.x & 63) == 0 && count) atomicAdd(a.counters, (unsigned long long)count);
}

hipError_t launch_errf1(const ErrfArgs& a, int ngrids, hipStream_t st) {
    if (ngrids <= 0) return hipSuccess;
    hipLaunchKernelGGL(errf1_kernel, dim3(ngrids), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_flag(int esize, const FlagArgs& a, int ngrids, hipStream_t st) {
    if (ngrids <= 0) return hipSuccess;
    if (esize == 8)
        hipLaunchKernelGGL(flag_kernel<double>, dim3(ngrids), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(flag_kernel<float>, dim3(ngrids), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace amrgpu
