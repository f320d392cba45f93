// util.hip -- small device helpers used by the host library.
#include <hip/hip_runtime.h>
#include <cstdint>

namespace amrgpu {

// Fill kernel used instead of hipMemsetAsync so that the sequence stays
// valid inside captured graphs.
__global__ void fill_u64_kernel(unsigned long long* p, unsigned long long v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_u64(unsigned long long* p, unsigned long long v, int64_t n,
                           hipStream_t st) {
    if (n <= 0) return hipSuccess;
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(blocks), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

// AoS (the reference's q(nvar, mitot, mjtot), component fastest, FP64) to
// the level's SoA storage and back.  One thread per (patch, cell); the
// staging buffer holds the patches back to back at stage_off[p].
template <typename S>
__global__ void aos_to_soa_kernel(const double* __restrict__ stage, const int64_t* __restrict__ stage_off,
                                  const int64_t* __restrict__ off, const int32_t* __restrict__ plane,
                                  const int64_t* __restrict__ cell_start, int npatch, int nv,
                                  int64_t ncell, S* __restrict__ out) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    int lo = 0, hi = npatch - 1;          // patch owning global cell t
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (cell_start[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int64_t c = t - cell_start[lo];
    const double* src = stage + stage_off[lo] + c * nv;
    S* dst = out + off[lo] + c;
    for (int m = 0; m < nv; m++) dst[(int64_t)m * plane[lo]] = (S)src[m];
}

template <typename S>
__global__ void soa_to_aos_kernel(double* __restrict__ stage, const int64_t* __restrict__ stage_off,
                                  const int64_t* __restrict__ off, const int32_t* __restrict__ plane,
                                  const int64_t* __restrict__ cell_start, int npatch, int nv,
                                  int64_t ncell, const S* __restrict__ in) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    int lo = 0, hi = npatch - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (cell_start[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int64_t c = t - cell_start[lo];
    double* dst = stage + stage_off[lo] + c * nv;
    const S* src = in + off[lo] + c;
    for (int m = 0; m < nv; m++) dst[m] = (double)src[(int64_t)m * plane[lo]];
}

hipError_t launch_aos_soa(bool to_soa, int esize, double* stage, const int64_t* stage_off,
                          const int64_t* off, const int32_t* plane, const int64_t* cell_start,
                          int npatch, int nv, int64_t ncell, void* level, hipStream_t st) {
    if (ncell <= 0) return hipSuccess;
    dim3 grid((unsigned)((ncell + 255) / 256)), block(256);
    if (to_soa) {
        if (esize == 8)
            hipLaunchKernelGGL(aos_to_soa_kernel<double>, grid, block, 0, st, stage, stage_off, off,
                               plane, cell_start, npatch, nv, ncell, (double*)level);
        else
            hipLaunchKernelGGL(aos_to_soa_kernel<float>, grid, block, 0, st, stage, stage_off, off,
                               plane, cell_start, npatch, nv, ncell, (float*)level);
    } else {
        if (esize == 8)
            hipLaunchKernelGGL(soa_to_aos_kernel<double>, grid, block, 0, st, stage, stage_off, off,
                               plane, cell_start, npatch, nv, ncell, (const double*)level);
        else
            hipLaunchKernelGGL(soa_to_aos_kernel<float>, grid, block, 0, st, stage, stage_off, off,
                               plane, cell_start, npatch, nv, ncell, (const float*)level);
    }
    return hipGetLastError();
}

}  // namespace amrgpu
