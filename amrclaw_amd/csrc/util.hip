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

}  // namespace amrgpu
