// sweep_march.hpp -- the level sweep as column strips marching up the rows,
// every quantity in registers (no barriers; LDS only as per-lane storage).
//
// The work unit is one 64-lane wave holding up to four strips (segments of
// 16, 32 or 64 lanes) of grids of the same height my.  A strip covers the
// interior columns x0+1 .. x0+W of its grid (W <= segment width - 4); lane
// l of the segment owns strip cell k = l - 1 (k = -1 .. W+2, the strip plus
// its two-cell halo on either side) and the x-interface between cells k-1
// and k.  The wave marches J = -1 .. my+1 and at each step
//   1. loads row J+1 of its column (storage precision; the ghost cells of
//      the output keep their time-t values, written here by one lane each),
//   2. solves y-interface J+1 (rows J, J+1: rpn2 and the Roe cache only),
//   3. finishes y-interface J: the wave limiter against interfaces J-1 and
//      J+1 (their waves passed down the march), the second-order correction,
//      both transverse solves,
//   4. solves x-row J completely (flux2 for one slice: the limiter's and the
//      transverse terms' neighbours come from the adjacent lanes by DPP),
//   5. completes the x-fluxes of row J-1 and the y-flux of edge J, and
//   6. updates row J-1.
// Every interface of a strip is solved once (the tile sweep re-solves the
// rows and columns of each tile's halo), the transverse increments of both
// directions meet in registers, and each cell of the grid is read from
// memory once (the mixed update's FP64 rows come back from a per-lane ring
// in LDS).
//
// The fluxes are accumulated in the reference's order (step2.f90:117-166,
// flux2.f:202-233), the same sums as the tile sweep (sweep.hpp):
//   fm(i,j) = (F_x(i,j) + Gy2(col i-1)(j)) + Gy1(col i)(j)
//   gm(i,j) = ((0 + Gx2(row j-1)(i)) + Gx1(row j)(i)) + F_y(i,j)
//   gadd(k,1) = (0 - 0.5 dtdx B(A-dq)(k+1)) - 0.5 dtdx B(A+dq)(k)
// so an FP64 instantiation built without contraction reproduces the
// reference bit for bit.
#pragma once
#include "sweep.hpp"

namespace amrgpu {

// flux2.f after rpn2 for one interface whose neighbours' waves are known
// through the dot products dotl (with the interface before) and dotr (with
// the one after): CFL, limiter, second-order correction, transverse solves.
// Returns the fluctuations fadm / fadp and the raw transverse results
// B(A-dq) (bm1, bp1) and B(A+dq) (bm2, bp2); the caller assembles gadd from
// them and the neighbouring interfaces' results.  Same arithmetic as
// flux2_lane.
template <class Sol, typename T, bool FLUXFORM, bool STD = false>
__device__ __forceinline__ void flux2_finish(const MethodArgs& M, const int ixy, const T* ql,
                                             const T* qr, const typename Sol::Cache& cache,
                                             T wave[Sol::NW][Sol::NV], const T s[Sol::NW],
                                             T amdq[Sol::NV], T apdq[Sol::NV],
                                             const T dotl[Sol::NW], const T dotr[Sol::NW],
                                             bool inner, T dtdx, T fadm[Sol::NV],
                                             T fadp[Sol::NV], T& cfl, T bm1[Sol::NV],
                                             T bp1[Sol::NV], T bm2[Sol::NV], T bp2[Sol::NV]) {
    constexpr int NV = Sol::NV, NW = Sol::NW;
    constexpr bool SKIPZ = sizeof(T) == 4;
#pragma unroll
    for (int m = 0; m < NV; m++) {
        fadm[m] = SKIPZ ? amdq[m] : T(0) + amdq[m];
        fadp[m] = FLUXFORM ? fadm[m] : (SKIPZ ? -apdq[m] : T(0) - apdq[m]);
    }
    if (inner) {
#pragma unroll
        for (int mw = 0; mw < NW; mw++) cfl = dmax1(dmax1(cfl, dtdx * s[mw]), -dtdx * s[mw]);
    }
    T cqxx[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) cqxx[m] = T(0);
    if (STD || M.order != 1) {
#pragma unroll
        for (int mw = 0; mw < NW; mw++) {
            T wn2 = T(0);
#pragma unroll
            for (int m = 0; m < NV; m++) {
                if (SKIPZ && Sol::wave_zero(ixy, mw, m)) continue;
                wn2 = wn2 + wave[mw][m] * wave[mw][m];
            }
            const int mth = STD ? 4 : M.mthlim[mw];
            if (mth != 0 && inner && wn2 != T(0)) {
                using Mth = typename Sol::Math;
                T r = Mth::ratio((s[mw] > T(0)) ? dotl[mw] : dotr[mw], wn2);
                T wl = phi_limiter<Mth>(mth, r);
#pragma unroll
                for (int m = 0; m < NV; m++)
                    if (!(SKIPZ && Sol::wave_zero(ixy, mw, m))) wave[mw][m] = wl * wave[mw][m];
            }
        }
        const T dtdxave = T(0.5) * (dtdx + dtdx);
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
            fadm[m] = fadm[m] + T(0.5) * c;
            fadp[m] = fadp[m] + T(0.5) * c;
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
        Sol::rpt2(M.ph, ixy, 1, ql, qr, nullptr, nullptr, nullptr, cache, amdq, bm1, bp1);
        Sol::rpt2(M.ph, ixy, 2, ql, qr, nullptr, nullptr, nullptr, cache, apdq, bm2, bp2);
    }
}

// dot products of the waves of two successive interfaces, per family (the
// limiter's numerator, inlinelimiter.f:36-50; flux2_lane's dotr)
template <class Sol, typename T>
__device__ __forceinline__ void wave_dots(const int ixy, const T a[Sol::NW][Sol::NV],
                                          const T b[Sol::NW][Sol::NV], T d[Sol::NW]) {
    constexpr bool SKIPZ = sizeof(T) == 4;
#pragma unroll
    for (int mw = 0; mw < Sol::NW; mw++) {
        T x = T(0);
#pragma unroll
        for (int m = 0; m < Sol::NV; m++) {
            if (SKIPZ && Sol::wave_zero(ixy, mw, m)) continue;
            x = x + a[mw][m] * b[mw][m];
        }
        d[mw] = x;
    }
}

template <class Sol, typename T, typename S, int WGS, bool STD = false>
__global__ __launch_bounds__(256, WGS) void sweep_march_kernel(SweepArgs a) {
    constexpr int NV = Sol::NV, NW = Sol::NW;
    constexpr bool MIXED = sizeof(T) < sizeof(S);
    constexpr bool SKIPZ = sizeof(T) == 4;
    using Cache = typename Sol::Cache;
    const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wv >= a.nwaves) return;   // the whole wave leaves
    // fields read from the table directly (a local copy of the wave's
    // record, indexed by segment, would live in scratch memory)
    const MarchWave* D = a.waves + wv;
    const int lane = threadIdx.x & 63, segw = D->segw;
    const int sidx = lane / segw, l = lane - sidx * segw;
    const bool seg_live = sidx < D->nseg;
    const MarchSeg sg = D->seg[seg_live ? sidx : 0];
    const PatchDesc pd = a.patches[sg.patch];
    const int W = sg.w, x0 = sg.x0, my = D->my;
    const int k = min(l - 1, W + 2);                 // strip cell of this lane
    const bool live = seg_live && l <= W + 3;
    const int col = x0 + k;                          // patch-local 1-based column
    const bool xs_inner = live && l >= 2 && l <= W + 2;   // x-interfaces 1..W+1
    const bool x_cell = live && k >= 1 && k <= W;         // interior columns
    const bool y_col = live && k >= 0 && k <= W + 1;      // columns the y-sweep needs
    const bool ghost_cols = live && ((x0 == 0 && k <= 0) || (x0 + W == pd.mx && k >= W + 1));
    const int64_t plane = pd.plane;
    const int mitot = pd.mitot;
    const S* qin = static_cast<const S*>(a.qin) + pd.off;
    S* qout = static_cast<S*>(a.qout) + pd.off;
    const double dtdx_d = a.dt / a.dx, dtdy_d = a.dt / a.dy;
    const T dtdx = T(dtdx_d), dtdy = T(dtdy_d);
    const T hdy = T(0.5) * dtdy;
    const bool trans = STD || a.meth.transverse != 0;

    // The mixed update needs rows in storage precision two steps after they
    // were loaded: each lane keeps its last three rows in a ring in LDS
    // (written and read by the same lane only, so no barrier), which costs
    // LDS instructions instead of a second read of the row from memory.
    __shared__ S ring[MIXED ? 3 : 1][NV][256];
    const int tid = threadIdx.x;
    auto slot = [](int R) { return (R + 3) % 3; };   // R >= -1

    // row R of this lane's column in storage precision: issued a step
    // before the row is needed, so the load's latency overlaps a step's work
    auto fetch_row = [&](int R, S p[NV]) {
        const int64_t li = (int64_t)(R + 1) * mitot + (col + 1);
#pragma unroll
        for (int m = 0; m < NV; m++) p[m] = qin[m * plane + li];
    };
    // row R, fetched, in compute precision; the ghost cells of the output
    // keep their time-t values (stepgrid.f updates the interior only), each
    // written by one lane
    auto accept_row = [&](int R, const S p[NV], T q[NV]) {
        const int64_t li = (int64_t)(R + 1) * mitot + (col + 1);
#pragma unroll
        for (int m = 0; m < NV; m++) q[m] = T(p[m]);
        if constexpr (MIXED) {
#pragma unroll
            for (int m = 0; m < NV; m++) ring[slot(R)][m][tid] = p[m];
        }
        if (ghost_cols || (x_cell && (R < 1 || R > my))) {
#pragma unroll
            for (int m = 0; m < NV; m++) qout[m * plane + li] = MIXED ? p[m] : S(q[m]);
        }
    };
    // row R (loaded at most two steps before) in storage precision
    auto stored_row = [&](int R, S p[NV]) {
#pragma unroll
        for (int m = 0; m < NV; m++) p[m] = ring[slot(R)][m][tid];
    };
    // FP64 physical flux of a stored state (the mixed path's flux form)
    auto phys = [&](int ixy, const S p[NV], double f[NV]) {
        double qd[NV];
#pragma unroll
        for (int m = 0; m < NV; m++) qd[m] = double(p[m]);
        Sol::template Rebind<double>::flux(a.meth.ph, ixy, qd, f);
    };

    T qa[NV], qb[NV], qc[NV];   // compute-precision rows J-1, J, J+1
    // y-interface J after rpn2 (rows J-1, J)
    Cache cJ{};
    T wJ[NW][NV], sJ[NW], amJ[NV], apJ[NV], dotlJ[NW];
    // carried from step J-1
    T bm2p[NV], bp2p[NV];      // y-interface J-1: B(A+dq) (its imp = 2 solve)
    T fxm_p[NV], fxp_p[NV];    // x-row J-1: faddm / faddp
    T gx2_p[NV];               // x-row J-1: gadd(., 2), bound for y-edge J
    T gym_p[NV], gyp_p[NV];    // y-edge J-1: its complete flux (gm / gp)
    double fy_p[NV];           // mixed: f_y(q(J-2))
#pragma unroll
    for (int m = 0; m < NV; m++) {
        bm2p[m] = bp2p[m] = fxm_p[m] = fxp_p[m] = gx2_p[m] = gym_p[m] = gyp_p[m] = T(0);
        fy_p[m] = 0.0;
    }
#pragma unroll
    for (int mw = 0; mw < NW; mw++) dotlJ[mw] = T(0);
    T cfl = T(0);
    double* per = a.reg ? a.reg + perim_offset(pd, NV) : nullptr;

    S pn[NV];                   // the next row, in flight
    fetch_row(-1, pn);
    accept_row(-1, pn, qb);
    fetch_row(0, pn);
    for (int J = -1; J <= my + 1; J++) {
        accept_row(J + 1, pn, qc);
        if (J + 2 <= my + 2) fetch_row(J + 2, pn);   // rows -1 .. my+2 exist

        // 4. (first, while the y-sweep's values of this step are not yet
        // live) x-row J (J >= 0): the whole flux2 slice, neighbours by DPP
        LaneResult<Sol, T> o;
        if (J >= 0) {
            T ql[NV];
#pragma unroll
            for (int m = 0; m < NV; m++) ql[m] = lane_up(qb[m], 0);   // cell k-1
            flux2_lane<Sol, T, MIXED, STD>(a.meth, 1, ql, qb, xs_inner, 0, dtdx, o);
            cfl = dmax1(cfl, o.cfl);
        }

        // 2. y-interface J+1 (rows J, J+1): Roe cache and rpn2
        Cache cN = Sol::prep(a.meth.ph, 2, qb, qc);
        T wN[NW][NV], sN[NW], amN[NV], apN[NV];
        Sol::rpn2(a.meth.ph, 2, qb, qc, nullptr, nullptr, cN, wN, sN, amN, apN);
        T dotr[NW];
        wave_dots<Sol, T>(2, wJ, wN, dotr);   // garbage at J = -1 (unused)

        // 3. finish y-interface J (J >= 1): F_y(J), B(A-dq)(J), B(A+dq)(J)
        T fym[NV], fyp[NV], bm1[NV], bp1[NV], bm2[NV], bp2[NV];
        const bool yfin = J >= 1;
        if (yfin) {
            flux2_finish<Sol, T, MIXED, STD>(a.meth, 2, qa, qb, cJ, wJ, sJ, amJ, apJ, dotlJ, dotr,
                                        y_col, dtdy, fym, fyp, cfl, bm1, bp1, bm2, bp2);
        }
        // gadd of row J-1 from the y-sweep (rows 1..my): Gy1 -> its x-edge i,
        // Gy2 -> x-edge i+1
        T gy1[NV], gy2[NV];
        if (trans && J >= 2) {
#pragma unroll
            for (int m = 0; m < NV; m++) {
                gy1[m] = (SKIPZ ? -(hdy * bm1[m]) : T(0) - hdy * bm1[m]) - hdy * bm2p[m];
                gy2[m] = (SKIPZ ? -(hdy * bp1[m]) : T(0) - hdy * bp1[m]) - hdy * bp2p[m];
            }
        }

        // 5a. the x-fluxes of row J-1 (rows 1..my are needed)
        T fxm[NV], fxp[NV];
        if (J >= 2) {
#pragma unroll
            for (int m = 0; m < NV; m++) {
                fxm[m] = fxm_p[m];
                fxp[m] = fxp_p[m];
                if (trans) {
                    const T g2l = lane_up(gy2[m], 0);   // column i-1's gadd(., 2)
                    fxm[m] = (fxm[m] + g2l) + gy1[m];
                    if (!MIXED) fxp[m] = (fxp[m] + g2l) + gy1[m];
                }
                if (MIXED) fxp[m] = fxm[m];   // flux form: one flux per edge
            }
        }
        // 5b. the y-flux of edge J (J = 1..my+1): x-rows J-1 and J, then F_y(J)
        T gym[NV], gyp[NV];
        if (yfin) {
#pragma unroll
            for (int m = 0; m < NV; m++) {
                T g = T(0);
                if (trans) g = (g + gx2_p[m]) + o.g1[m];
                gym[m] = g + fym[m];
                gyp[m] = MIXED ? gym[m] : g + fyp[m];
            }
        }

        // 6. update row J-1 (rows 1..my) and the perimeter fluxes
        double fyc[NV];
        S pa[NV];                                // row J-1, storage precision
        if (MIXED && J >= 1) {
            stored_row(J - 1, pa);
            phys(2, pa, fyc);                    // f_y(q(J-1))
        }
        if (J >= 2) {
            const int R = J - 1;
            const int64_t li = (int64_t)(R + 1) * mitot + (col + 1);
            T fxm_r[NV];
#pragma unroll
            for (int m = 0; m < NV; m++) fxm_r[m] = lane_down(fxm[m], 0);   // x-edge i+1
            double fxc[NV], fxw[NV];
            if constexpr (MIXED) {
                phys(1, pa, fxc);
#pragma unroll
                for (int m = 0; m < NV; m++) fxw[m] = lane_up(fxc[m], 0);
            }
            if (x_cell) {
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    if constexpr (MIXED) {
                        const double q = pa[m];
                        const double dx_ = (fxc[m] - fxw[m]) + (double(fxm_r[m]) - double(fxp[m]));
                        const double dy_ = (fyc[m] - fy_p[m]) + (double(gym[m]) - double(gyp_p[m]));
                        qout[m * plane + li] = q - dtdx_d * dx_ - dtdy_d * dy_;
                    } else {
                        const T q = qa[m];
                        const T dx_ = fxm_r[m] - fxp[m];
                        const T dy_ = gym[m] - gyp_p[m];
                        qout[m * plane + li] = S(q - dtdx * dx_ - dtdy * dy_);
                    }
                }
            }
            if (per) {
                // fluxes through the grid's edges (fluxad.f:36-92): left
                // x-edge 1, right x-edge mx+1, component fastest
                if (live && x0 == 0 && k == 1) {
#pragma unroll
                    for (int m = 0; m < NV; m++)
                        per[(int64_t)(R - 1) * NV + m] = MIXED ? fxw[m] + double(fxm[m]) : double(fxm[m]);
                }
                if (live && x0 + W == pd.mx && k == W + 1) {
#pragma unroll
                    for (int m = 0; m < NV; m++)
                        per[(int64_t)NV * (pd.my + pd.mx) + (int64_t)(R - 1) * NV + m] =
                            MIXED ? fxw[m] + double(fxp[m]) : double(fxp[m]);
                }
            }
        }
        if (per && x_cell && (J == 1 || J == my + 1)) {
            // bottom y-edge 1 (gm) and top y-edge my+1 (gp): f_y of the row
            // below the edge plus G in the flux form
            const int64_t base = J == 1 ? (int64_t)NV * (2 * pd.my + pd.mx) : (int64_t)NV * pd.my;
#pragma unroll
            for (int m = 0; m < NV; m++) {
                const T g = J == 1 ? gym[m] : gyp[m];
                per[base + (int64_t)(col - 1) * NV + m] = MIXED ? fyc[m] + double(g) : double(g);
            }
        }

        // carry to step J+1
#pragma unroll
        for (int m = 0; m < NV; m++) {
            if (J >= 0) {
                fxm_p[m] = o.fadm[m];
                fxp_p[m] = o.fadp[m];
                gx2_p[m] = o.g2[m];
            }
            if (yfin) {
                bm2p[m] = bm2[m];
                bp2p[m] = bp2[m];
                gym_p[m] = gym[m];
                gyp_p[m] = gyp[m];
            }
            if (MIXED && J >= 1) fy_p[m] = fyc[m];
            qa[m] = qb[m];
            qb[m] = qc[m];
            amJ[m] = amN[m];
            apJ[m] = apN[m];
        }
#pragma unroll
        for (int mw = 0; mw < NW; mw++) {
            dotlJ[mw] = dotr[mw];
            sJ[mw] = sN[mw];
#pragma unroll
            for (int m = 0; m < NV; m++) wJ[mw][m] = wN[mw][m];
        }
        cJ = cN;
    }
    // Courant maximum of each strip's grid: reduce within the segment
    for (int off = segw >> 1; off >= 1; off >>= 1) cfl = dmax1(cfl, __shfl_xor(cfl, off, 64));
    if (seg_live && l == 0) atomic_max_nonneg(&a.cfl_patch[sg.patch], double(cfl));
}

}  // namespace amrgpu
