// regs.hpp -- device programs of the coarse-fine flux registers and update.
#pragma once
#include <cstdint>
#include "types.hpp"

namespace amrgpu {

// upbnd.f: one coarse cell and the ordered register slots added to it
struct UpbndCell {
    int32_t grid;           // coarse grid
    int32_t i, j;           // 0-based cell of the grid array (with ghosts)
    int32_t e0, e1;         // its entries [e0, e1)
};
struct UpbndEntry {
    int32_t kid, lkid;      // fine grid and 1-based register slot
    double sgnm;            // +1 or -1 (upbnd.f:89-95)
};

// update.f:95-177: coarse cells [ci0, ci0+ni) x [cj0, cj0+nj) of grid
// `grid` (0-based with ghosts) are averages of the fine cells of grid `kid`
// starting at (fi0, fj0) (0-based with ghosts), r x r per coarse cell
struct AvgBlock {
    int32_t grid, kid;
    int32_t ci0, cj0, ni, nj;
    int32_t fi0, fj0;
};

// qad.f: one coarse segment (register slot) of one side of a fine grid
struct QadItem {
    int32_t grid;
    int32_t side;           // 1 left, 2 top, 3 right, 4 bottom
    int32_t seg;            // 1-based segment along the side
    int32_t slot;           // 1-based register slot (influx)
};

struct QadArgs {
    const QadItem* items;
    int32_t nitems;
    int32_t nv;
    const PatchDesc* patches;
    const void* q;          // the level's new-time storage (ghosts filled)
    double* reg;            // the level's register arena
    int32_t ratx, raty;     // intratx/intraty(level-1)
    double hx, hy, delt;
    // 1: the level's fluxes are in flux form (the mixed path), so fm/fp at
    // the grid edges are the fluxes themselves and qad's Riemann problems,
    // which turn fluctuations into fluxes, are skipped
    int32_t flux_form;
    // aux arrays of the level (FP64, naux components, patch p at
    // patches[p].off * auxk) and their types (AMRGPU_AUX_*); naux 0: none
    int32_t naux;
    const double* aux;
    int32_t auxk;
    int32_t auxtype[8];
    int32_t pad;
};

struct UpdateArgs {
    const UpbndCell* cells;
    int32_t ncells;
    int32_t nv;
    const UpbndEntry* entries;
    const AvgBlock* blocks;
    int32_t nblocks;
    int32_t ratx, raty;
    const PatchDesc* patches;       // coarse level
    const PatchDesc* fine_patches;  // fine level
    void* q;                        // coarse level, new time
    const void* qf;                 // fine level, new time
    const double* reg;              // fine level register arena
    double area;                    // hxposs*hyposs of the coarse level
    // capacity (mcapa > 0): the coarse and fine levels' aux arrays (FP64,
    // patch p at off * auxk), the capacity's component (1-based)
    const double* aux_c;
    const double* aux_f;
    int32_t mcapa, auxk;
};

}  // namespace amrgpu
