// noalloc.hip -- the work space of amrgpu_stepgrid_noalloc: one static
// device array per GPU, holding the patch descriptor, the Courant maximum,
// the tile table and the two structure-of-arrays copies of the grid the
// sweep reads and writes.  Nothing here allocates device memory.
//
// Byte layout of the array (noalloc.hpp gives the offsets):
//   [0, 48)        PatchDesc of the grid (element offset 0)
//   [64, 72)       per-patch Courant maximum (double bits, atomicMax target)
//   [256, ...)     TileDesc[NOALLOC_MAX_TILES]
//   [DATA, ...)    q at time t (SoA, storage precision), then q at t + dt
#include <hip/hip_runtime.h>
#include "noalloc.hpp"

namespace amrgpu {

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// The array's device address on each GPU (the array's name is a host-side
// stand-in that must never be passed to a kernel or a copy).
constexpr int NOALLOC_MAXDEV = 64;
static void* g_scratch[NOALLOC_MAXDEV] = {};

hipError_t noalloc_init() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= NOALLOC_MAXDEV) return hipErrorInvalidDevice;
    if (g_scratch[dev]) return hipSuccess;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = p;
    return hipSuccess;
}

unsigned char* noalloc_scratch() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAXDEV) return nullptr;
    return static_cast<unsigned char*>(g_scratch[dev]);
}

// The descriptor of the one grid and its tiles: tile t covers x-tile
// t % ntx and y-tile t / ntx, tw x th cells except at the grid's upper
// edges (the tiling split_axis gives the host-planned levels); the Courant
// maximum starts at zero.
__global__ void noalloc_setup_kernel(unsigned char* base, int mx, int my, int tw, int th, int ntx,
                                     int nty) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        PatchDesc pd{};
        pd.off = 0;
        pd.mx = mx;
        pd.my = my;
        pd.mitot = mx + 2 * NGHOST;
        pd.plane = pd.mitot * (my + 2 * NGHOST);
        pd.ilo = 0;
        pd.ihi = mx - 1;
        pd.jlo = 0;
        pd.jhi = my - 1;
        *reinterpret_cast<PatchDesc*>(base + NOALLOC_PATCH_OFF) = pd;
        *reinterpret_cast<unsigned long long*>(base + NOALLOC_CFL_OFF) = 0ull;
    }
    if (t < ntx * nty) {
        const int ix = t % ntx, iy = t / ntx;
        TileDesc d{};
        d.patch = 0;
        d.i0 = (int16_t)(ix * tw);
        d.j0 = (int16_t)(iy * th);
        d.tx = (int16_t)min(tw, mx - ix * tw);
        d.ty = (int16_t)min(th, my - iy * th);
        d.wx = (int16_t)(d.tx + 3 <= 16 ? 16 : (d.tx + 3 <= 32 ? 32 : 64));
        d.wy = (int16_t)(d.ty + 3 <= 16 ? 16 : (d.ty + 3 <= 32 ? 32 : 64));
        reinterpret_cast<TileDesc*>(base + NOALLOC_TILES_OFF)[t] = d;
    }
}

// one grid, the reference's AoS FP64 q(nv, mitot, mjtot) <-> SoA (S)
template <typename S>
__global__ void noalloc_to_soa_kernel(const double* __restrict__ aos, S* __restrict__ soa, int nv,
                                      int plane) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= plane) return;
    for (int m = 0; m < nv; m++) soa[(int64_t)m * plane + c] = (S)aos[(int64_t)c * nv + m];
}

template <typename S>
__global__ void noalloc_to_aos_kernel(const S* __restrict__ soa, double* __restrict__ aos, int nv,
                                      int plane) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= plane) return;
    for (int m = 0; m < nv; m++) aos[(int64_t)c * nv + m] = (double)soa[(int64_t)m * plane + c];
}

hipError_t noalloc_setup(unsigned char* base, int mx, int my, int tw, int th, int ntx, int nty,
                         hipStream_t st) {
    const int n = ntx * nty;
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3((n + 255) / 256), dim3(256), 0, st, base, mx, my,
                       tw, th, ntx, nty);
    return hipGetLastError();
}

hipError_t noalloc_transpose(bool to_soa, int esize, const double* aos_in, double* aos_out,
                             void* soa, int nv, int plane, hipStream_t st) {
    const dim3 grid((plane + 255) / 256), block(256);
    if (to_soa) {
        if (esize == 8)
            hipLaunchKernelGGL(noalloc_to_soa_kernel<double>, grid, block, 0, st, aos_in,
                               (double*)soa, nv, plane);
        else
            hipLaunchKernelGGL(noalloc_to_soa_kernel<float>, grid, block, 0, st, aos_in,
                               (float*)soa, nv, plane);
    } else {
        if (esize == 8)
            hipLaunchKernelGGL(noalloc_to_aos_kernel<double>, grid, block, 0, st,
                               (const double*)soa, aos_out, nv, plane);
        else
            hipLaunchKernelGGL(noalloc_to_aos_kernel<float>, grid, block, 0, st,
                               (const float*)soa, aos_out, nv, plane);
    }
    return hipGetLastError();
}

}  // namespace amrgpu
