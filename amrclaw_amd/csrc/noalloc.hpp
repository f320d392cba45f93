// noalloc.hpp -- host-side interface of the static work space behind
// amrgpu_stepgrid_noalloc (noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include "amrgpu.h"
#include "types.hpp"

namespace amrgpu {

constexpr size_t NOALLOC_SCRATCH_BYTES = AMRGPU_NOALLOC_SCRATCH_BYTES;
constexpr size_t NOALLOC_PATCH_OFF = 0;
constexpr size_t NOALLOC_CFL_OFF = 64;
constexpr size_t NOALLOC_TILES_OFF = 256;
constexpr int NOALLOC_MAX_TILES = AMRGPU_NOALLOC_MAX_TILES;
constexpr size_t NOALLOC_DATA_OFF =
    (NOALLOC_TILES_OFF + (size_t)NOALLOC_MAX_TILES * sizeof(TileDesc) + 255) / 256 * 256;
constexpr size_t NOALLOC_DATA_BYTES = NOALLOC_SCRATCH_BYTES - NOALLOC_DATA_OFF;

// Look up the array's device address on the current GPU (once per GPU).
hipError_t noalloc_init();
// The current GPU's address of the array, or nullptr before noalloc_init.
unsigned char* noalloc_scratch();
// Write the grid's descriptor and tile table and clear its Courant maximum.
hipError_t noalloc_setup(unsigned char* base, int mx, int my, int tw, int th, int ntx, int nty,
                         hipStream_t st);
// AoS FP64 (the reference's layout) -> SoA of the storage precision
// (to_soa, from aos_in) or back (into aos_out).
hipError_t noalloc_transpose(bool to_soa, int esize, const double* aos_in, double* aos_out,
                             void* soa, int nv, int plane, hipStream_t st);

}  // namespace amrgpu
