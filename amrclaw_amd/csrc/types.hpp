// types.hpp -- device-independent descriptors shared by kernels and host code.
#pragma once
#include <cstdint>

namespace amrgpu {

constexpr int NWMAX = 8;     // max waves per Riemann problem
constexpr int NGHOST = 2;    // the reference's nghost (amr_module nghost = 2)

// One grid (patch) of a level.  Storage is structure-of-arrays per patch:
// component m, cell (I, J) with 1-based interior indices I in [1-ng, mx+ng],
// J in [1-ng, my+ng] lives at  off + m*plane + (J+ng-1)*mitot + (I+ng-1).
// The reference stores the same patch AoS, component fastest
// (update.f:31; intfil.f90:293-297); the C-ABI converts at the boundary.
struct PatchDesc {
    int64_t off;        // element offset of component 0
    int32_t mx, my;     // interior cells
    int32_t mitot;      // mx + 2*ng (row stride)
    int32_t plane;      // mitot * mjtot (component stride)
    int32_t ilo, jlo;   // global index of the first interior cell (level space)
    int32_t ihi, jhi;
    int64_t reg_off;    // fine grids: offset of svdflx(nvar, lenbc) in the level's
                        // register arena; qc1d(nvar, lenbc) follows it
    int32_t lenbc;      // 2 (mx/r + my/r) coarse segments around the grid (prepf.f:35-39)
    int32_t pad;
};

// fluxsv.f entry of a coarse grid: its interior cell (I, J) (1-based),
// side 1..4 of that cell, fine grid `kid` of the next level and the slot
// `lkid` (1-based) of the kid's register (setuse.f:61-213)
struct SvEntry {
    int32_t I, J, side, kid, lkid, pad;
};

// One workgroup's piece of a patch for the sweep.  Interior cells
// i0+1..i0+tx, j0+1..j0+ty of patch `patch`.  wx/wy: lanes per x-row /
// y-column segment (16, 32 or 64, >= tx+3 / ty+3).
struct TileDesc {
    int32_t patch;
    int16_t i0, j0, tx, ty;
    int16_t wx, wy;
    int32_t reg_lo, reg_hi;   // coarse-side register entries of this tile
    int32_t flags;            // TILE_* bits
    int32_t pad;
};
enum { TILE_FLUXAD_L = 1, TILE_FLUXAD_R = 2, TILE_FLUXAD_B = 4, TILE_FLUXAD_T = 8 };

// The march sweep's work unit (sweep_march.hpp): one 64-lane wave holding
// nseg strips of segw lanes each (16, 32 or 64), all of grids my rows high.
// Strip s covers the interior columns x0+1 .. x0+w (1-based, w <= segw - 4)
// of grid `patch`.
struct MarchSeg {
    int32_t patch;
    int16_t x0, w;
};
struct MarchWave {
    MarchSeg seg[4];
    int16_t segw, nseg, my, pad;
};

}  // namespace amrgpu
