// sweep.hpp -- fused, batched wave-propagation sweep over the tiles of a level.
//
// One workgroup (256 lanes) advances one tile of one patch:
//   1. stage the tile plus its 2-cell halo in LDS (compute precision T);
//   2. x-phase: one lane per interface, one 16/32/64-lane segment per row
//      (rows j0..j0+ty+1); normal solve, CFL, wave limiter (neighbour waves
//      by cross-lane shuffle), second-order correction, both transverse
//      solves; F fluxes to LDS, transverse G increments accumulated into LDS;
//   3. y-phase: the same per column, producing G fluxes and the transverse F
//      increments;
//   4. conservative update of the interior into the output buffer and the
//      per-patch CFL maximum.
// This is step2.f90:7-169 + flux2.f:25-237 + inlinelimiter.f:4-108 +
// stepgrid.f:146-173 for every patch of a level in one launch.
//
// Accumulation order follows the reference exactly, so an FP64 instantiation
// built without FMA contraction reproduces it bit for bit:
//   fm(i,j) = (F_x(i,j) + Gy2(col i-1)(j)) + Gy1(col i)(j)   (step2.f90:117,161,163)
//   gm(i,j) = (Gx2(row j-1)(i) + Gx1(row j)(i)) + F_y(i,j)   (step2.f90:119,121,159)
//   gadd(k,1) = (0 - 0.5 dtdx B(A-dq)(k+1)) - 0.5 dtdx B(A+dq)(k)  (flux2.f:207-233)
// Row (column) contributions of different rows are applied in two
// barrier-separated steps, rows in ascending order, as the reference's
// j (i) loop does.
#pragma once
#include "lanes.hpp"
#include "level.hpp"

namespace amrgpu {

struct SweepArgs {
    const void* qin;                // level buffer, time t (ghosts filled)
    void* qout;                     // level buffer, receives the interior at t+dt
    const PatchDesc* patches;
    const TileDesc* tiles;
    int32_t ntiles;
    int32_t layout;        // fixed LDS layout every tile fits: 1 28x12 (FP32 arithmetic), 2 12x12 (FP64); 0 none
    double dt, dx, dy;
    MethodArgs meth;
    unsigned long long* cfl_patch;  // per-patch max Courant number (double bits)
    double* flux_out;               // optional: fm,fp,gm,gp per patch (tests)
    const int64_t* flux_off;        // element offset of each patch in flux_out
    // coarse-fine registers (FP64).  reg: this level's arena when the level
    // is a fine level; each tile on a grid edge stores that edge's fluxes in
    // the grid's perimeter block (after svdflx and qc1d).  sv / reg_fine /
    // fine_patches: when a finer level exists, the fluxsv entries of each
    // tile (TileDesc.reg_lo..reg_hi) write -flux*dt*h into the finer level's
    // register slots.
    double* reg;
    double* reg_fine;
    const PatchDesc* fine_patches;
    const SvEntry* sv;
    // persistent launch when non-null: the tile counter, zeroed before the
    // launch
    unsigned long long* next;
    // march sweep (sweep_march.hpp) when non-null: its wave table; the tile
    // fields are then unused
    const MarchWave* waves;
    int32_t nwaves;
    // aux arrays of the level (FP64, naux components per cell; patch p's
    // component a of cell li at aux[patches[p].off * auxk + a * plane + li])
    // and the capacity component (1-based, 0: none, step2.f90:91-95)
    const double* aux;
    int32_t naux, auxk, mcapa, pad2;
};

// Perimeter flux block of a fine grid (FP64, after svdflx and qc1d):
// side 0 left fm at x-edge 1 (rows 1..my), side 1 top gp at y-edge my+1
// (columns 1..mx), side 2 right fp at x-edge mx+1, side 3 bottom gm at
// y-edge 1; component fastest.  These are the fluxes fluxad.f:36-92 adds.
__host__ __device__ inline int64_t perim_offset(const PatchDesc& pd, int nv) {
    return pd.reg_off + 2 * (int64_t)nv * pd.lenbc;
}
// auxc1d(naux, lenbc) of a fine grid when its level has aux arrays: the
// coarse ring's aux values cstore.f saves, after the perimeter block
__host__ __device__ inline int64_t aux_ring_offset(const PatchDesc& pd, int nv) {
    return perim_offset(pd, nv) + 2 * (int64_t)nv * (pd.mx + pd.my);
}


// limiter function of inlinelimiter.f:58-94
template <class Mth, typename T>
__device__ __forceinline__ T phi_limiter(int mth, T r) {
    switch (mth) {
    case 1: return dmax1(T(0), dmin1(T(1), r));
    case 2: return dmax1(dmax1(T(0), dmin1(T(1), T(2) * r)), dmin1(T(2), r));
    case 3: return Mth::div(r + fabs(r), T(1) + fabs(r));
    case 4: {
        T c = (T(1) + r) / T(2);
        return dmax1(T(0), dmin1(dmin1(c, T(2)), T(2) * r));
    }
    case 5: return r;
    default: return T(1);
    }
}

// Results of one interface lane of a segment.
// The aux values one lane's interface needs when its level has aux arrays
// (a capacity, or a solver that reads aux): the solver's components of the
// cells left and right of the interface and of each one's neighbours across
// the slice (below / above for an x-slice, left / right for a y-slice: the
// aux1 / aux3 entries rpt2 reads at i-2+imp, step2.f90:98-102), and the
// capacity-scaled dt/dx of the two cells (dtdx1d, step2.f90:91-95).
template <typename T, int NA> struct LaneAux {
    T l[NA], r[NA];
    T lm[NA], lp[NA], rm[NA], rp[NA];
    T dtl, dtr;
};
template <class Sol> constexpr int lane_naux() {
    return sol_naux<Sol>::value > 0 ? sol_naux<Sol>::value : 1;
}

template <class Sol, typename T> struct LaneResult {
    T fadm[Sol::NV], fadp[Sol::NV];   // faddm / faddp at this interface
    T g1[Sol::NV], g2[Sol::NV];       // gadd(.,1) / gadd(.,2) of the cell right of it
    T cfl;
};

// flux2.f for one interface.  `inner` marks interfaces 1..n+1 of the slice
// (the ones flux2 limits and uses); lanes outside still run the solver on
// clamped, valid states so the shuffles see defined values.
//
// FLUXFORM (the mixed path: FP32 waves, FP64 storage): both cells of the
// interface take the same increment G = A-dq (+ the second-order and
// transverse terms) instead of A-dq and -A+dq; the update adds the FP64
// physical fluxes, so the interface flux is F = f(q_l) + G for both cells.
// In exact arithmetic f(q_r) - F = A+dq (the Riemann solvers' conservation
// property), so the scheme is the same; in finite precision the cells'
// updates telescope and the totals are conserved to FP64 rounding although
// the waves are FP32, and F carries only G's FP32 error, which scales with
// the jump, not with f.
//
// STD: the method is known to be the standard one (std_method: second
// order, both transverse corrections, MC limiter on every wave, no
// f-waves), so its switches are compile-time constants; same arithmetic.
//
// AUX: the level has aux arrays; ax holds the lane's values (dtdx is then
// unused: the two cells' dtdx1d replace it where flux2.f reads them).
template <class Sol, typename T, bool FLUXFORM = false, bool STD = false, bool AUX = false>
__device__ __forceinline__ void flux2_lane(const MethodArgs& M, int ixy, const T* ql,
                                           const T* qr, bool inner, int w, T dtdx,
                                           LaneResult<Sol, T>& o,
                                           const LaneAux<T, lane_naux<Sol>()>* ax = nullptr) {
    constexpr int NV = Sol::NV, NW = Sol::NW;
    constexpr bool SAUX = AUX && sol_naux<Sol>::value > 0;   // the solver reads aux
    const T dtl = AUX ? ax->dtl : dtdx, dtr = AUX ? ax->dtr : dtdx;
    // FP32 arithmetic (no bit-exactness requirement) leaves out the terms of
    // the wave components that are zero by construction and the additions
    // of zero that the reference's sums start from
    constexpr bool SKIPZ = sizeof(T) == 4;
    typename Sol::Cache cache = Sol::prep(M.ph, ixy, ql, qr);
    T wave[NW][NV], s[NW], amdq[NV], apdq[NV];
    Sol::rpn2(M.ph, ixy, ql, qr, SAUX ? ax->l : nullptr, SAUX ? ax->r : nullptr, cache, wave, s,
              amdq, apdq);

    if (FLUXFORM) {
#pragma unroll
        for (int m = 0; m < NV; m++) {
            o.fadm[m] = SKIPZ ? amdq[m] : T(0) + amdq[m];
            o.fadp[m] = o.fadm[m];
        }
    } else {
#pragma unroll
        for (int m = 0; m < NV; m++) {
            o.fadm[m] = SKIPZ ? amdq[m] : T(0) + amdq[m];
            o.fadp[m] = SKIPZ ? -apdq[m] : T(0) - apdq[m];
        }
    }
    T cfl = T(0);
    if (inner) {
#pragma unroll
        for (int mw = 0; mw < NW; mw++) cfl = dmax1(dmax1(cfl, dtr * s[mw]), -dtl * s[mw]);
    }
    o.cfl = cfl;

    T cqxx[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) cqxx[m] = T(0);
    if (STD || M.order != 1) {
#pragma unroll
        for (int mw = 0; mw < NW; mw++) {
            T wn2 = T(0), dotr = T(0);
#pragma unroll
            for (int m = 0; m < NV; m++) {
                if (SKIPZ && Sol::wave_zero(ixy, mw, m)) continue;
                T wnext = lane_down(wave[mw][m], w);
                wn2 = wn2 + wave[mw][m] * wave[mw][m];
                dotr = dotr + wave[mw][m] * wnext;
            }
            T dotl = lane_up(dotr, w);
            const int mth = STD ? 4 : M.mthlim[mw];
            if (mth != 0 && inner && wn2 != T(0)) {
                using Mth = typename Sol::Math;
                T r = Mth::ratio((s[mw] > T(0)) ? dotl : dotr, wn2);
                T wl = phi_limiter<Mth>(mth, r);
#pragma unroll
                for (int m = 0; m < NV; m++)
                    if (!(SKIPZ && Sol::wave_zero(ixy, mw, m))) wave[mw][m] = wl * wave[mw][m];
            }
        }
        const T dtdxave = T(0.5) * (dtl + dtr);
#pragma unroll
        for (int m = 0; m < NV; m++) {
            T c = T(0);
#pragma unroll
            for (int mw = 0; mw < NW; mw++) {
                if (SKIPZ && Sol::wave_zero(ixy, mw, m)) continue;
                T abs_sign = (!STD && M.use_fwaves) ? copysign(T(1), s[mw]) : fabs(s[mw]);
                c = c + abs_sign * (T(1) - fabs(s[mw]) * dtdxave) * wave[mw][m];
            }
            cqxx[m] = c;
            o.fadm[m] = o.fadm[m] + T(0.5) * c;
            o.fadp[m] = o.fadp[m] + T(0.5) * c;
        }
    }

    if (STD || M.transverse != 0) {
        if (STD || (M.order > 1 && M.transverse == 2)) {
#pragma unroll
            for (int m = 0; m < NV; m++) {
                amdq[m] = amdq[m] + cqxx[m];
                apdq[m] = apdq[m] - cqxx[m];
            }
        }
        T bm1[NV], bp1[NV], bm2[NV], bp2[NV];
        Sol::rpt2(M.ph, ixy, 1, ql, qr, SAUX ? ax->lm : nullptr, SAUX ? ax->l : nullptr,
                  SAUX ? ax->lp : nullptr, cache, amdq, bm1, bp1);
        Sol::rpt2(M.ph, ixy, 2, ql, qr, SAUX ? ax->rm : nullptr, SAUX ? ax->r : nullptr,
                  SAUX ? ax->rp : nullptr, cache, apdq, bm2, bp2);
        // the increments land in the cell right of the interface: its dtdx1d
        const T hdt = T(0.5) * dtr;
#pragma unroll
        for (int m = 0; m < NV; m++) {
            T nbm = lane_down(bm1[m], w);   // imp=1 result of interface k+1
            T nbp = lane_down(bp1[m], w);
            o.g1[m] = (SKIPZ ? -(hdt * nbm) : T(0) - hdt * nbm) - hdt * bm2[m];
            o.g2[m] = (SKIPZ ? -(hdt * nbp) : T(0) - hdt * nbp) - hdt * bp2[m];
        }
    } else {
#pragma unroll
        for (int m = 0; m < NV; m++) { o.g1[m] = T(0); o.g2[m] = T(0); }
    }
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v) {
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = dmax1(v, __shfl_xor(v, off, 64));
    return v;
}

// The standard method of the run configurations (method(2) = 2,
// method(3) = 2, MC limiter on every wave, no f-waves): the sweeps have
// instantiations with these switches fixed at compile time.
__host__ __device__ inline bool std_method(const MethodArgs& M) {
    if (M.order != 2 || M.transverse != 2 || M.use_fwaves) return false;
    for (int mw = 0; mw < M.mwaves && mw < NWMAX; mw++)
        if (M.mthlim[mw] != 4) return false;
    return true;
}

// LDS bytes needed for a tile of tx * ty cells (compute type T)
__host__ __device__ inline int sweep_qs_stride(int tx) { return (tx + 4) | 1; }
// LDS of one tile: the staged states and the edge increments; `merged` (the
// mixed path, where both cells of an edge take the same increment) keeps one
// array per direction instead of the minus / plus pair
template <typename T>
__host__ __device__ inline size_t sweep_lds_bytes(int nv, int tx, int ty, bool merged = false) {
    const size_t k = merged ? 1 : 2;
    size_t qs = (size_t)nv * (ty + 4) * sweep_qs_stride(tx);
    size_t f = k * (size_t)nv * ty * (tx + 1);
    size_t g = k * (size_t)nv * (ty + 1) * tx;
    return (qs + f + g) * sizeof(T);
}

// one tile of the sweep, run by one workgroup.  LX, LY: compile-time extents
// of the LDS layout (the tile caps; every tile fits in them), so that the
// LDS strides are constants; 0: the tile's own extents
template <class Sol, typename T, typename S, int LX = 0, int LY = 0, bool STD = false,
          bool AUX = false>
__device__ __forceinline__ void sweep_tile(const SweepArgs& a, const int tile,
                                           unsigned char* smem_raw) {
    constexpr int NV = Sol::NV;
    // compute precision below the storage precision: the mixed path (flux
    // form, update in the storage precision)
    constexpr bool MIXED = sizeof(T) < sizeof(S);
    T* smem = reinterpret_cast<T*>(smem_raw);

    const TileDesc td = a.tiles[tile];
    const PatchDesc pd = a.patches[td.patch];
    const int tx = td.tx, ty = td.ty;
    const int sx = LX ? LX : tx, sy = LY ? LY : ty;   // layout extents
    const int qsx = sweep_qs_stride(sx), qsy = sy + 4;
    T* qs = smem;                          // [NV][qsy][qsx], cell (k,r) at [r+1][k+1]
    T* fm = qs + NV * qsy * qsx;           // [NV][sy][sx+1]  x-edge e=1..tx+1, row j
    // mixed path: fp == fm and gp == gm (flux form), one array each
    T* fp = MIXED ? fm : fm + NV * sy * (sx + 1);
    T* gm = fp + NV * sy * (sx + 1);       // [NV][sy+1][sx]  y-edge e=1..ty+1, col k
    T* gp = MIXED ? gm : gm + NV * (sy + 1) * sx;
    const int fstr = sx + 1;               // row stride of fm/fp
    const int fpl = sy * (sx + 1);         // component stride of fm/fp
    const int gsx = sx;                    // row stride of gm/gp
    const int gpl = (sy + 1) * sx;         // component stride of gm/gp

    const int tid = threadIdx.x;
    const S* qin = static_cast<const S*>(a.qin) + pd.off;
    // aux (AUX): component c of the patch's ghost-origin cell (J, I), and a
    // lane's LaneAux for the interface between cells (JL, IL) and (JR, IR)
    // of a slice in direction ixy
    constexpr int NA = lane_naux<Sol>(), SNA = sol_naux<Sol>::value;
    const double* auxp = AUX ? a.aux + pd.off * a.auxk : nullptr;
    auto auxv = [&](int c, int J, int I) -> T {
        return T(auxp[(int64_t)c * pd.plane + (int64_t)J * pd.mitot + I]);
    };
    auto fill_aux = [&](LaneAux<T, NA>& ax, int ixy, int JL, int IL, int JR, int IR, T dt) {
        const int dJ = ixy == 1 ? 1 : 0, dI = ixy == 1 ? 0 : 1;
#pragma unroll
        for (int c = 0; c < SNA; c++) {
            ax.l[c] = auxv(c, JL, IL);
            ax.r[c] = auxv(c, JR, IR);
            ax.lm[c] = auxv(c, JL - dJ, IL - dI);
            ax.lp[c] = auxv(c, JL + dJ, IL + dI);
            ax.rm[c] = auxv(c, JR - dJ, IR - dI);
            ax.rp[c] = auxv(c, JR + dJ, IR + dI);
        }
        ax.dtl = a.mcapa > 0 ? dt / auxv(a.mcapa - 1, JL, IL) : dt;   // step2.f90:91-95
        ax.dtr = a.mcapa > 0 ? dt / auxv(a.mcapa - 1, JR, IR) : dt;
    };
    // mixed path: FP64 physical flux of the stored state at patch-local
    // linear index li (x-flux for ixy = 1, y-flux for 2)
    auto phys_flux = [&](int ixy, int64_t li, double* f) {
        double qd[NV];
#pragma unroll
        for (int m = 0; m < NV; m++) qd[m] = double(qin[(int64_t)m * pd.plane + li]);
        Sol::template Rebind<double>::flux(a.meth.ph, ixy, qd, f);
    };

    // 1. stage tile + halo; zero gm (the x-phase accumulates into it)
    {
        const int w4 = tx + 4, rows = ty + 4, n = NV * rows * w4;
        for (int idx = tid; idx < n; idx += 256) {
            int m = idx / (rows * w4);
            int rem = idx - m * rows * w4;
            int rr = rem / w4, kk = rem - rr * w4;
            int li = (td.j0 + rr) * pd.mitot + (td.i0 + kk);   // patch-local, ghost-origin
            qs[(m * qsy + rr) * qsx + kk] = T(qin[(int64_t)m * pd.plane + li]);
        }
        for (int idx = tid; idx < NV * gpl; idx += 256) gm[idx] = T(0);
    }
    __syncthreads();

    const double dtdx_d = a.dt / a.dx, dtdy_d = a.dt / a.dy;
    const T dtdx = T(dtdx_d), dtdy = T(dtdy_d);
    T cfl = T(0);

    // 2. x-phase: segment = row r in [0, ty+1], lane = interface c in [0, tx+2]
    {
        const int w = td.wx, rows_per = 256 / w;
        const int seg = tid / w, c = tid - seg * w;
        const int npass = (ty + 2 + rows_per - 1) / rows_per;
        for (int pass = 0; pass < npass; pass++) {
            const int r = pass * rows_per + seg;
            const bool live = (r <= ty + 1) && (c <= tx + 2);
            const int rc = r <= ty + 1 ? r : ty + 1;
            const int cc = c <= tx + 2 ? c : tx + 2;
            T ql[NV], qr[NV];
#pragma unroll
            for (int m = 0; m < NV; m++) {
                ql[m] = qs[(m * qsy + rc + 1) * qsx + cc];       // cell c-1
                qr[m] = qs[(m * qsy + rc + 1) * qsx + cc + 1];   // cell c
            }
            const bool inner = live && c >= 1 && c <= tx + 1;
            LaneResult<Sol, T> o;
            LaneAux<T, NA> ax;
            if constexpr (AUX)
                fill_aux(ax, 1, td.j0 + rc + 1, td.i0 + cc, td.j0 + rc + 1, td.i0 + cc + 1, dtdx);
            flux2_lane<Sol, T, MIXED, STD, AUX>(a.meth, 1, ql, qr, inner, w, dtdx, o, &ax);
            cfl = dmax1(cfl, o.cfl);
            if (inner && r >= 1 && r <= ty) {
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    fm[m * fpl + (r - 1) * fstr + (c - 1)] = o.fadm[m];
                    if (!MIXED) fp[m * fpl + (r - 1) * fstr + (c - 1)] = o.fadp[m];
                }
            }
            const bool cell = live && c >= 1 && c <= tx;       // cell k = c
            // step A: gadd(.,2) of row r goes to edge r+1
            if ((STD || a.meth.transverse != 0) && cell && r <= ty) {
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    T* g = &gm[m * gpl + r * gsx + (c - 1)];
                    *g = *g + o.g2[m];
                }
            }
            __syncthreads();
            // step B: gadd(.,1) of row r goes to edge r
            if ((STD || a.meth.transverse != 0) && cell && r >= 1) {
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    T* g = &gm[m * gpl + (r - 1) * gsx + (c - 1)];
                    *g = *g + o.g1[m];
                }
            }
            __syncthreads();
        }
    }

    // 3. y-phase: segment = column c in [0, tx+1], lane = interface e in [0, ty+2]
    {
        const int w = td.wy, cols_per = 256 / w;
        const int seg = tid / w, e = tid - seg * w;
        const int npass = (tx + 2 + cols_per - 1) / cols_per;
        for (int pass = 0; pass < npass; pass++) {
            const int c = pass * cols_per + seg;
            const bool live = (c <= tx + 1) && (e <= ty + 2);
            const int cc = c <= tx + 1 ? c : tx + 1;
            const int ee = e <= ty + 2 ? e : ty + 2;
            T ql[NV], qr[NV];
#pragma unroll
            for (int m = 0; m < NV; m++) {
                ql[m] = qs[(m * qsy + ee) * qsx + cc + 1];       // cell (c, e-1)
                qr[m] = qs[(m * qsy + ee + 1) * qsx + cc + 1];   // cell (c, e)
            }
            const bool inner = live && e >= 1 && e <= ty + 1;
            LaneResult<Sol, T> o;
            LaneAux<T, NA> ax;
            if constexpr (AUX)
                fill_aux(ax, 2, td.j0 + ee, td.i0 + cc + 1, td.j0 + ee + 1, td.i0 + cc + 1, dtdy);
            flux2_lane<Sol, T, MIXED, STD, AUX>(a.meth, 2, ql, qr, inner, w, dtdy, o, &ax);
            cfl = dmax1(cfl, o.cfl);
            if (inner && c >= 1 && c <= tx) {
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    const int gi = m * gpl + (e - 1) * gsx + (c - 1);
                    T xg = gm[gi];
                    gm[gi] = xg + o.fadm[m];
                    if (!MIXED) gp[gi] = xg + o.fadp[m];
                }
            }
            const bool cell = live && e >= 1 && e <= ty;       // cell j = e
            // step A: gadd(.,2) of column c goes to x-edge c+1
            if ((STD || a.meth.transverse != 0) && cell && c <= tx) {
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    const int fi = m * fpl + (e - 1) * fstr + c;
                    fm[fi] = fm[fi] + o.g2[m];
                    if (!MIXED) fp[fi] = fp[fi] + o.g2[m];
                }
            }
            __syncthreads();
            // step B: gadd(.,1) of column c goes to x-edge c
            if ((STD || a.meth.transverse != 0) && cell && c >= 1) {
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    const int fi = m * fpl + (e - 1) * fstr + (c - 1);
                    fm[fi] = fm[fi] + o.g1[m];
                    if (!MIXED) fp[fi] = fp[fi] + o.g1[m];
                }
            }
            __syncthreads();
        }
    }

    // 4. conservative update (stepgrid.f:154-165) and CFL
    {
        S* qout = static_cast<S*>(a.qout) + pd.off;
        const int n = tx * ty;
        for (int idx = tid; idx < n; idx += 256) {
            const int j = idx / tx, k = idx - j * tx;   // 0-based interior of tile
            const int li = (td.j0 + j + 2) * pd.mitot + (td.i0 + k + 2);
            double fxc[NV], fxw[NV], fyc[NV], fys[NV];
            if constexpr (MIXED) {
                phys_flux(1, li, fxc);
                phys_flux(1, li - 1, fxw);
                phys_flux(2, li, fyc);
                phys_flux(2, li - pd.mitot, fys);
            }
#pragma unroll
            for (int m = 0; m < NV; m++) {
                if constexpr (MIXED) {
                    // update in FP64 from the stored state: F(i+1) - F(i) =
                    // (f(q_i) - f(q_i-1)) + (G(i+1) - G(i)), f in FP64
                    const double q = qin[(int64_t)m * pd.plane + li];
                    const double dx_ = (fxc[m] - fxw[m]) + (double(fm[m * fpl + j * fstr + (k + 1)]) -
                                                            double(fp[m * fpl + j * fstr + k]));
                    const double dy_ = (fyc[m] - fys[m]) + (double(gm[m * gpl + (j + 1) * gsx + k]) -
                                                            double(gp[m * gpl + j * gsx + k]));
                    qout[(int64_t)m * pd.plane + li] = q - dtdx_d * dx_ - dtdy_d * dy_;
                } else {
                    T q = qs[(m * qsy + j + 2) * qsx + k + 2];
                    T dx_ = fm[m * fpl + j * fstr + (k + 1)] - fp[m * fpl + j * fstr + k];
                    T dy_ = gm[m * gpl + (j + 1) * gsx + k] - gp[m * gpl + j * gsx + k];
                    if (AUX && a.mcapa > 0)                 // stepgrid.f:166-172
                        qout[(int64_t)m * pd.plane + li] =
                            S(q - (dtdx * dx_ + dtdy * dy_) / auxv(a.mcapa - 1, td.j0 + j + 2, td.i0 + k + 2));
                    else
                        qout[(int64_t)m * pd.plane + li] = S(q - dtdx * dx_ - dtdy * dy_);
                }
            }
        }
        // ghost cells on the patch edges keep their time-t values in the new
        // array, as stepgrid.f:154-165 updates q in place over the interior
        // only; each ghost cell is written by the one tile whose range,
        // extended by the ghost width on its patch-edge sides, holds it
        {
            const int xlo = td.i0 == 0 ? -2 : 0, xhi = td.i0 + tx == pd.mx ? tx + 2 : tx;
            const int ylo = td.j0 == 0 ? -2 : 0, yhi = td.j0 + ty == pd.my ? ty + 2 : ty;
            if (xlo != 0 || xhi != tx || ylo != 0 || yhi != ty) {
                const int w = xhi - xlo, hgt = yhi - ylo;
                for (int idx = tid; idx < w * hgt; idx += 256) {
                    const int j = ylo + idx / w, k = xlo + idx % w;
                    if (k >= 0 && k < tx && j >= 0 && j < ty) continue;
                    const int li = (td.j0 + j + 2) * pd.mitot + (td.i0 + k + 2);
#pragma unroll
                    for (int m = 0; m < NV; m++)
                        qout[(int64_t)m * pd.plane + li] =
                            MIXED ? qin[(int64_t)m * pd.plane + li]
                                  : S(qs[(m * qsy + j + 2) * qsx + k + 2]);
                }
            }
        }
        if (a.flux_out) {
            // fm,fp,gm,gp of the patch, each (NV, my+4, mx+4) like the
            // reference's flux arrays: x-edge e of row j -> fm(e, j)
            double* fo = a.flux_out + a.flux_off[td.patch];
            const int64_t fpln = (int64_t)pd.plane;
            const int64_t arr = NV * fpln;
            const int fcells = ty * (tx + 1), gcells = (ty + 1) * tx;
            for (int idx = tid; idx < NV * fcells; idx += 256) {
                int m = idx / fcells, rem = idx - m * fcells, j = rem / (tx + 1), e = rem - j * (tx + 1);
                int64_t li = (int64_t)(td.j0 + j + 2) * pd.mitot + (td.i0 + e + 2) + m * fpln;
                const int lx = m * fpl + j * fstr + e;
                fo[li] = double(fm[lx]);
                fo[arr + li] = double(fp[lx]);
            }
            for (int idx = tid; idx < NV * gcells; idx += 256) {
                int m = idx / gcells, rem = idx - m * gcells, e = rem / tx, k = rem - e * tx;
                int64_t li = (int64_t)(td.j0 + e + 2) * pd.mitot + (td.i0 + k + 2) + m * fpln;
                const int lx = m * gpl + e * gsx + k;
                fo[2 * arr + li] = double(gm[lx]);
                fo[3 * arr + li] = double(gp[lx]);
            }
        }
    }
    // 5. coarse-fine registers (reads the final fluxes in LDS only)
    if (a.reg) {
        // mixed path: the flux through an edge is f(state left of / below
        // it) + G (FP64 f), the same value the cells' update implies
        double* per = a.reg + perim_offset(pd, NV);
        auto flux_at = [&](const T* g, int gstride, int m, int ixy, int64_t li_left) -> double {
            double v = double(g[m * gstride]);
            if constexpr (MIXED) {
                double f[NV];
                phys_flux(ixy, li_left, f);
                v = f[m] + v;
            }
            return v;
        };
        if (td.i0 == 0)
            for (int idx = tid; idx < NV * ty; idx += 256) {
                const int j = idx / NV, m = idx - j * NV;
                per[(int64_t)(td.j0 + j) * NV + m] =
                    flux_at(fm + j * fstr, fpl, m, 1, (int64_t)(td.j0 + j + 2) * pd.mitot + 1);
            }
        if (td.i0 + tx == pd.mx)
            for (int idx = tid; idx < NV * ty; idx += 256) {
                const int j = idx / NV, m = idx - j * NV;
                per[(int64_t)NV * (pd.my + pd.mx) + (int64_t)(td.j0 + j) * NV + m] =
                    flux_at(fp + j * fstr + tx, fpl, m, 1, (int64_t)(td.j0 + j + 2) * pd.mitot + pd.mx + 1);
            }
        if (td.j0 + ty == pd.my)
            for (int idx = tid; idx < NV * tx; idx += 256) {
                const int k = idx / NV, m = idx - k * NV;
                per[(int64_t)NV * pd.my + (int64_t)(td.i0 + k) * NV + m] =
                    flux_at(gp + ty * gsx + k, gpl, m, 2, (int64_t)(pd.my + 1) * pd.mitot + (td.i0 + k + 2));
            }
        if (td.j0 == 0)
            for (int idx = tid; idx < NV * tx; idx += 256) {
                const int k = idx / NV, m = idx - k * NV;
                per[(int64_t)NV * (2 * pd.my + pd.mx) + (int64_t)(td.i0 + k) * NV + m] =
                    flux_at(gm + k, gpl, m, 2, (int64_t)1 * pd.mitot + (td.i0 + k + 2));
            }
    }
    if (a.sv) {
        // fluxsv.f:43-75: slot = -flux * dtc * h of the coarse cell's side
        const double dtc = a.dt;
        for (int e = td.reg_lo + tid; e < td.reg_hi; e += 256) {
            const SvEntry s = a.sv[e];
            const int k = s.I - 1 - td.i0, j = s.J - 1 - td.j0;
            double* slot = a.reg_fine + a.fine_patches[s.kid].reg_off + (int64_t)NV * (s.lkid - 1);
            // the flux through the side of the coarse cell that faces the
            // fine grid; side chosen once, then one strided read per component
            const T* src;
            int stride;
            double h;
            if (s.side == 1) { src = fp + j * fstr + k; stride = fpl; h = a.dy; }
            else if (s.side == 2) { src = gm + (j + 1) * gsx + k; stride = gpl; h = a.dx; }
            else if (s.side == 3) { src = fm + j * fstr + k + 1; stride = fpl; h = a.dy; }
            else { src = gp + j * gsx + k; stride = gpl; h = a.dx; }
            if constexpr (MIXED) {
                // F = f(state left of / below the edge) + G: sides 1 and 4
                // look at the coarse cell's left / lower neighbour
                const int64_t lc = (int64_t)(s.J + 1) * pd.mitot + (s.I + 1);
                const int64_t ll = s.side == 1 ? lc - 1 : (s.side == 4 ? lc - pd.mitot : lc);
                double f[NV];
                phys_flux((s.side == 1 || s.side == 3) ? 1 : 2, ll, f);
                for (int m = 0; m < NV; m++) slot[m] = -(f[m] + double(src[m * stride])) * dtc * h;
            } else {
                for (int m = 0; m < NV; m++) slot[m] = -double(src[m * stride]) * dtc * h;
            }
        }
    }
    cfl = wave_max(cfl);
    if ((tid & 63) == 0) atomic_max_nonneg(&a.cfl_patch[td.patch], double(cfl));
}

// batched launch: one workgroup per tile of the tile table.  WGS: the
// workgroups per CU the register allocation is bounded for (a tuning knob:
// the mixed path is instantiated for 5, 6 and 7)
template <class Sol, typename T, typename S, int WGS, int LX = 0, int LY = 0, bool STD = false,
          bool AUX = false>
__global__ __launch_bounds__(256, WGS) void sweep_kernel(SweepArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    sweep_tile<Sol, T, S, LX, LY, STD, AUX>(a, blockIdx.x, smem_raw);
}

// persistent launch: a grid sized to fill the device takes tiles from the
// table in order (largest first) through the counter a.next, which starts
// at zero; every workgroup leaves once the counter has passed the table
template <class Sol, typename T, typename S, int WGS, int LX = 0, int LY = 0, bool STD = false>
__global__ __launch_bounds__(256, WGS) void sweep_persistent_kernel(SweepArgs a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int tile;
    for (;;) {
        if (threadIdx.x == 0) tile = (int)atomicAdd(a.next, 1ull);
        __syncthreads();
        const int t = tile;
        if (t >= a.ntiles) break;   // the same value in every lane
        sweep_tile<Sol, T, S, LX, LY, STD>(a, t, smem_raw);
        __syncthreads();            // LDS and `tile` are reused by the next tile
    }
}

}  // namespace amrgpu
