// sweep3_tile.hpp -- the 3-D step fused over a block of cells, with every
// face on chip (FP32 storage and arithmetic).
//
// The phase kernels of sweep3.hpp reproduce step3.f's accumulation into the
// six flux arrays, which live in HBM.  Here one workgroup owns a tile of at
// most TILE3_MAX cells per direction of one grid and runs the x-, y- and
// z-sweeps and the update for it alone:
//   * the tile and two cells on every side (the cells the tile's lines read,
//     ghost cells of the grid where the tile reaches its edge) are staged in
//     LDS once;
//   * each sweep solves the lines through the tile and one line on every
//     side of it across the sweep (their transverse increments reach the
//     tile's cells), tr + 1 lanes per line, one lane per interface:
//     interfaces 0 .. tr of a tile with tr cells along the line (lanes 0
//     and tr also solve interfaces -1 and tr+1, whose waves the limiter
//     needs);
//   * instead of accumulating faces and differencing them afterwards, every
//     face contribution goes at once to the two cells it separates, in an
//     LDS array dq of the tile's cells: a face value F between cells a < b
//     enters the update as q(a) -= (dt/h) F, q(b) += (dt/h) F (stepgrid.f
//     (3-D):150-173 with fm = fp = F for the transverse increments of
//     step3.f:221-250, 279-310), the normal fluxes as
//     q(c) -= dt/h (faddm(c+1) - faddp(c));
//   * the lines are scheduled so that no two waves of a round touch the same
//     cell and the lanes of a wave add in ordered steps (tile3_sweep): no
//     atomics, the sums are formed in a fixed order and the result is
//     deterministic.
// The per-interface arithmetic (rpn3, limiter, cqxx, the rpt3 / rptt3
// splits and their 1/2, 1/6 weights, flux3.f:220-628) is that of the phase
// kernel; only the order in which each cell's increments are summed differs
// from the reference, so the result agrees with it to FP32 rounding.  The
// grid's ghost cells are read, never written: the step reads q from one
// buffer and writes the updated interiors into another.
#pragma once
#include <cstdint>
#include "lanes.hpp"
#include "sweep3.hpp"

namespace amrgpu {

constexpr int TILE3_MAX = 15;        // cells per direction of a tile (15: lines of 16 lanes fill a wave)
constexpr int TILE3_LANES = 256;     // four waves
constexpr int TILE3_XCDS = 8;        // workgroups go to the XCDs in turn
constexpr int TILE3_WG_PER_CU = 2;   // workgroups a CU holds (LDS and registers sized for it)

// a tile: grid `patch`, 0-based first interior cell (i0, j0, k0), extents
struct Tile3 {
    int32_t patch;
    int16_t i0, j0, k0, tx, ty, tz, pad;
};

struct Tile3Args {
    const void* qin;                // solutions at time t (grid layout of Patch3)
    void* qout;                     // interiors at t + dt (ghost cells untouched)
    const Patch3* patches;
    const Tile3* tiles;
    int32_t ntiles;
    int32_t nv;
    // the tiles of XCD x are tiles[xcd_start[x] ..+ xcd_count[x]]: workgroup
    // b runs on XCD b % 8 and takes its (b / 8)-th tile
    int32_t xcd_start[TILE3_XCDS], xcd_count[TILE3_XCDS];
    double dtd[3];                  // dt/dx, dt/dy, dt/dz
    Method3 meth;
    unsigned long long* cfl_patch;  // per-patch max Courant number (double bits)
};

// LDS of a tile workgroup: the staged box and the increments
template <int NV>
__host__ __device__ constexpr size_t tile3_lds_bytes(int bx, int by, int bz, int cells) {
    return ((size_t)NV * bx * by * bz + (size_t)NV * cells) * sizeof(float);
}

// Wave-synchronous ordering: the LDS accesses of one wave before this point
// are complete and visible to its other lanes after it (as __syncwarp).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// idx = tid + TILE3_LANES i written as digits (x, y, z, m) of the radices
// (nx, ny, nz): stepped by TILE3_LANES with one carry per digit instead of
// three integer divisions per element (a digit plus its step plus a carry
// stays below twice its radix, so one subtraction brings it back)
struct Digits3 {
    int x, y, z, m, sx, sy, sz, sm, nx, ny, nz;
    __device__ __forceinline__ Digits3(int idx, int nx_, int ny_, int nz_) : nx(nx_), ny(ny_), nz(nz_) {
        int r = idx;
        x = r % nx; r /= nx; y = r % ny; r /= ny; z = r % nz; m = r / nz;
        r = TILE3_LANES;
        sx = r % nx; r /= nx; sy = r % ny; r /= ny; sz = r % nz; sm = r / nz;
    }
    __device__ __forceinline__ void step() {
        x += sx;
        int c = x >= nx;
        x -= c ? nx : 0;
        y += sy + c;
        c = y >= ny;
        y -= c ? ny : 0;
        z += sz + c;
        c = z >= nz;
        z -= c ? nz : 0;
        m += sm + c;
    }
};

// one sweep of a tile (D: 1 x, 2 y, 3 z, a template parameter so that every
// component index of the solver is a compile-time constant).
//
// Schedule: the lines of the sweep form rows of constant P (P = -1 .. tp),
// each of tq + 2 lines Q = -1 .. tq.  A wave takes a chunk of up to
// lpw = 64 / (tr + 1) consecutive lines of one row, tr + 1 lanes per line.
// The rows are taken in three classes ((P + 1) mod 3); in a round the four
// waves take chunks of distinct rows of one class, which lie three or more
// rows apart, so the cells their lines touch (P - 1 .. P + 1) are disjoint.
// Within a wave the lines of a chunk touch common cells; each lane adds its
// increments to the nine cells (dQ, dP) around its own in nine ordered
// steps, in each of which the wave's lanes touch distinct cells.  Rounds are
// separated by barriers.
template <class Sol, int D>
__device__ __forceinline__ void tile3_sweep(const Tile3Args& a, float* qs, float* dq, int nbox,
                                            int ncell, int tx, int ty, int tz, int bx, int by,
                                            int wave, int wl, float& cfl) {
    using T = float;
    constexpr int NV = Sol::NV, NW = Sol::NW;
    constexpr int d = D;
    const Method3& M = a.meth;
    const bool trans = M.m3 > 0, cq = M.m3 >= 2, tt = M.m4 > 0;
    bool limit = false;
    for (int w = 0; w < NW; w++) limit = limit || M.mthlim[w] > 0;
    // axes (r, Q, P): x (x, y, z), y (y, x, z), z (z, x, y); LDS strides
    const int sx = 1, sy = bx, sz = bx * by;
    const int tr = d == 1 ? tx : (d == 2 ? ty : tz);
    const int tq = d == 1 ? ty : tx;
    const int tp = d == 3 ? ty : tz;
    const int rs = d == 1 ? sx : (d == 2 ? sy : sz);
    const int qsd = d == 1 ? sy : sx;
    const int psd = d == 3 ? sy : sz;
    // dq strides (tile cells, x fastest)
    const int drs = d == 1 ? 1 : (d == 2 ? tx : tx * ty);
    const int dqs = d == 1 ? tx : 1;
    const int dps = d == 3 ? tx : tx * ty;
    const T dtdx = T(a.dtd[d - 1]);
    const T dtdy_p = T(a.dtd[d % 3]);
    const T dtdz_p = T(a.dtd[(d + 1) % 3]);
    const T dtQ = T(a.dtd[d == 1 ? 1 : 0]);     // across Q (type A faces)
    const T dtP = T(a.dtd[d == 3 ? 1 : 2]);     // across P (type B faces)
    const bool g_to_a = d != 2;
    const T half_t = T(0.5) * dtdx;
    const T sixth = T(1) / T(6);
    const T g6 = sixth * dtdx * dtdz_p;
    const T h6 = sixth * dtdx * dtdy_p;

    // a line's lanes take its interfaces 0 .. tr; the limiter's outer
    // neighbours (interfaces -1 and tr+1, waves only) are solved by lanes 0
    // and tr themselves
    const int L = tr + 1;                      // lanes of a line
    const int lpw = 64 / L;                    // lines of a chunk
    const int li = wl / L;                     // this lane's line in the chunk
    const int e = wl - li * L;                 // interface e: between cells e-1, e
    const int K = (tq + 2 + lpw - 1) / lpw;    // chunks per row
#pragma unroll 1
    for (int c3 = 0; c3 < 3; c3++) {
        const int nrows = (tp + 4 - c3) / 3;   // rows P = -1 + c3 + 3 i
        const int wpr = nrows < 4 ? nrows : 4;
        const int npairs = nrows * K;
        const int nrounds = (npairs + wpr - 1) / wpr;
        // pair = rd * wpr + wave as (row, chunk) = (pair % nrows, pair / nrows),
        // stepped by wpr <= nrows a round (one carry)
        int prow = wave < wpr ? wave : 0, pchunk = 0;
#pragma unroll 1
        for (int rd = 0; rd < nrounds; rd++) {
            const int pair = rd * wpr + wave;
            const bool wlive = wave < wpr && pair < npairs;
            const int P = -1 + c3 + 3 * (wlive ? prow : 0);
            const int Qr = -1 + (wlive ? pchunk : 0) * lpw + li;
            const bool live = wlive && li < lpw && Qr <= tq;
            const int Q = live ? Qr : -1;
            const int ec = live ? e : 0;
            // cells e-1 and e of the line (local -2 .. t+1)
            const int b0 = 2 * (1 + sy + sz) + Q * qsd + P * psd + (ec - 1) * rs;
            T ql[NV], qrr[NV];
#pragma unroll
            for (int m = 0; m < NV; m++) {
                ql[m] = qs[m * nbox + b0];
                qrr[m] = qs[m * nbox + b0 + rs];
            }
            const typename Sol::Cache cache = Sol::prep(M.ph, ql, qrr);
            T wave_[NW][NV], s[NW], amdq[NV], apdq[NV];
            Sol::rpn3(M.ph, d, ql, qrr, cache, wave_, s, amdq, apdq);
            // Courant number of interfaces 1 .. n+1 of every line
            // (ghost lines included), each counted by some tile
            if (live && e >= 0 && e <= tr) {
#pragma unroll
                for (int w = 0; w < NW; w++) cfl = dmax1(dmax1(cfl, dtdx * s[w]), -dtdx * s[w]);
            }
            T fadm[NV], fadp[NV], cqxx[NV];
#pragma unroll
            for (int m = 0; m < NV; m++) {
                fadm[m] = T(0) + amdq[m];
                fadp[m] = T(0) - apdq[m];
                cqxx[m] = T(0);
            }
            if (M.order != 1) {
                if (limit) {   // inlinelimiter.f on unlimited neighbour waves
                    T wd[NW][NV], wu[NW][NV];
#pragma unroll
                    for (int w = 0; w < NW; w++)
#pragma unroll
                        for (int m = 0; m < NV; m++) {
                            wd[w][m] = lane_down(wave_[w][m], L);
                            wu[w][m] = lane_up(wave_[w][m], L);
                        }
                    if (live && (e == 0 || e == tr)) {
                        // the outer neighbour of the line's end: interface -1
                        // (cells -2, -1) for lane 0, tr+1 (cells tr, tr+1) for
                        // lane tr (both when tr = 0 cannot happen: tr >= 1)
                        const int bx0 = e == 0 ? b0 - rs : b0 + rs;
                        T xl[NV], xr[NV];
#pragma unroll
                        for (int m = 0; m < NV; m++) {
                            xl[m] = qs[m * nbox + bx0];
                            xr[m] = qs[m * nbox + bx0 + rs];
                        }
                        const typename Sol::Cache cx = Sol::prep(M.ph, xl, xr);
                        T wx[NW][NV], sx_[NW], amx[NV], apx[NV];
                        Sol::rpn3(M.ph, d, xl, xr, cx, wx, sx_, amx, apx);
#pragma unroll
                        for (int w = 0; w < NW; w++)
#pragma unroll
                            for (int m = 0; m < NV; m++) {
                                if (e == 0) wu[w][m] = wx[w][m];
                                else wd[w][m] = wx[w][m];
                            }
                    }
                    if (live && e >= 0 && e <= tr) {
#pragma unroll
                        for (int w = 0; w < NW; w++) {
                            const int mth = M.mthlim[w];
                            if (mth == 0) continue;
                            T wn2 = T(0), dotr = T(0), dotl = T(0);
#pragma unroll
                            for (int m = 0; m < NV; m++) {
                                wn2 = wn2 + wave_[w][m] * wave_[w][m];
                                dotr = dotr + wave_[w][m] * wd[w][m];
                                dotl = dotl + wu[w][m] * wave_[w][m];
                            }
                            if (wn2 == T(0)) continue;
                            using Mth = typename Sol::Math;
                            T rr = Mth::ratio(s[w] > T(0) ? dotl : dotr, wn2);
                            T wlim = phi_limiter3<Mth>(mth, rr);
#pragma unroll
                            for (int m = 0; m < NV; m++) wave_[w][m] = wlim * wave_[w][m];
                        }
                    }
                }
                const T dtdxave = T(0.5) * (dtdx + dtdx);
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    T c = T(0);
#pragma unroll
                    for (int w = 0; w < NW; w++) {
                        const T fac = M.use_fwaves ? (s[w] >= T(0) ? T(1) : T(-1)) : fabs(s[w]);
                        c = c + T(0.5) * fac * (T(1) - fabs(s[w]) * dtdxave) * wave_[w][m];
                    }
                    cqxx[m] = c;
                    fadm[m] = fadm[m] + c;
                    fadp[m] = fadp[m] + c;
                }
            }
            // lane e completes cell c = e-1 of the line (tile cells 0 .. t-1):
            // acc[(dQ+1) + 3 (dP+1)] collects its increments to cell
            // (c, Q+dQ, P+dP)
            T acc[9][NV];
#pragma unroll
            for (int k = 0; k < 9; k++)
#pragma unroll
                for (int m = 0; m < NV; m++) acc[k][m] = T(0);
            {
                // normal fluxes: q(c) -= dt/dx (faddm(c+1) - faddp(c))
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    const T fpl = lane_up(fadp[m], L);
                    acc[4][m] = T(0) - dtdx * (fadm[m] - fpl);
                }
            }
            if (trans) {
                // the splits both families use, once: of cqxx, A+dq and A-dq
                // along icoor 2 and 3 (index 0: 2, 1: 3; [0] down-, [1] up-going)
                T Cs[2][2][NV], As[2][2][NV], Ms[2][2][NV];
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    if (cq) Sol::split(M.ph, d, 2 + i, ql, qrr, cache, cqxx, Cs[i][0], Cs[i][1]);
                    Sol::split(M.ph, d, 2 + i, ql, qrr, cache, apdq, As[i][0], As[i][1]);
                    Sol::split(M.ph, d, 2 + i, ql, qrr, cache, amdq, Ms[i][0], Ms[i][1]);
                }
#pragma unroll
                for (int fam = 0; fam < 2; fam++) {
                    // fam 0 = gadd: primary split along icoor 2; fam 1 = hadd
                    const int ip = fam, is = 1 - fam;   // primary, secondary
                    const int icp = 2 + ip;
                    const T w6 = fam == 0 ? g6 : h6;
                    T tmp[NV], rm[NV], rp[NV], sm[NV], sp[NV];
                    // right-cell partials of interface e (cell e): slot
                    // k = (s-1)*3 + (o+1), flux3.f:450-505
                    T v[6][NV];
#pragma unroll
                    for (int k = 0; k < 6; k++)
#pragma unroll
                        for (int m = 0; m < NV; m++) v[k][m] = T(0);
#pragma unroll
                    for (int m = 0; m < NV; m++) {
                        v[1][m] = T(0) - half_t * As[ip][0][m];
                        v[4][m] = T(0) - half_t * As[ip][1][m];
                    }
                    if (tt) {
#pragma unroll
                        for (int m = 0; m < NV; m++)
                            tmp[m] = M.m4 == 2 ? As[is][1][m] - T(3) * Cs[is][1][m] : As[is][1][m];
                        Sol::split(M.ph, d, icp, ql, qrr, cache, tmp, rm, rp);
#pragma unroll
                        for (int m = 0; m < NV; m++)
                            tmp[m] = M.m4 == 2 ? As[is][0][m] - T(3) * Cs[is][0][m] : As[is][0][m];
                        Sol::split(M.ph, d, icp, ql, qrr, cache, tmp, sm, sp);
#pragma unroll
                        for (int m = 0; m < NV; m++) {
                            v[4][m] = v[4][m] + w6 * (rp[m] - sp[m]);
                            v[1][m] = v[1][m] + w6 * (rm[m] - sm[m]);
                            v[5][m] = T(0) - w6 * rp[m];
                            v[2][m] = T(0) - w6 * rm[m];
                            v[3][m] = T(0) + w6 * sp[m];
                            v[0][m] = T(0) + w6 * sm[m];
                        }
                    }
                    if (cq) {
#pragma unroll
                        for (int m = 0; m < NV; m++) {
                            v[4][m] = v[4][m] + dtdx * Cs[ip][1][m];
                            v[1][m] = v[1][m] + dtdx * Cs[ip][0][m];
                        }
                    }
                    // cell e-1: the left neighbour's right-cell partials,
                    // then this interface's left-cell increments (A-dq side)
#pragma unroll
                    for (int k = 0; k < 6; k++)
#pragma unroll
                        for (int m = 0; m < NV; m++) v[k][m] = lane_up(v[k][m], L);
#pragma unroll
                    for (int m = 0; m < NV; m++) {
                        v[1][m] = v[1][m] - half_t * Ms[ip][0][m];
                        v[4][m] = v[4][m] - half_t * Ms[ip][1][m];
                    }
                    if (tt) {
#pragma unroll
                        for (int m = 0; m < NV; m++)
                            tmp[m] = M.m4 == 2 ? Ms[is][1][m] + T(3) * Cs[is][1][m] : Ms[is][1][m];
                        Sol::split(M.ph, d, icp, ql, qrr, cache, tmp, rm, rp);
#pragma unroll
                        for (int m = 0; m < NV; m++)
                            tmp[m] = M.m4 == 2 ? Ms[is][0][m] + T(3) * Cs[is][0][m] : Ms[is][0][m];
                        Sol::split(M.ph, d, icp, ql, qrr, cache, tmp, sm, sp);
#pragma unroll
                        for (int m = 0; m < NV; m++) {
                            v[4][m] = v[4][m] + w6 * (rp[m] - sp[m]);
                            v[1][m] = v[1][m] + w6 * (rm[m] - sm[m]);
                            v[5][m] = v[5][m] - w6 * rp[m];
                            v[2][m] = v[2][m] - w6 * rm[m];
                            v[3][m] = v[3][m] + w6 * sp[m];
                            v[0][m] = v[0][m] + w6 * sm[m];
                        }
                    }
                    if (cq) {
#pragma unroll
                        for (int m = 0; m < NV; m++) {
                            v[4][m] = v[4][m] - dtdx * Cs[ip][1][m];
                            v[1][m] = v[1][m] - dtdx * Cs[ip][0][m];
                        }
                    }
                    // slot (s, o) is the face Q+s-1 of cell row P+o (type
                    // A, across Q) or the face P+s-1 of cell row Q+o (type
                    // B, across P); a face enters the cells on its two sides
                    const bool to_a = (fam == 0) == g_to_a;
                    const T h = to_a ? dtQ : dtP;
#pragma unroll
                    for (int o = -1; o <= 1; o++) {
                        if (o != 0 && !tt) continue;
#pragma unroll
                        for (int s2 = 1; s2 <= 2; s2++) {
                            const int k = (s2 - 1) * 3 + o + 1;
                            // lower cell (dQ, dP) and upper cell of the face
                            const int kl = to_a ? (s2 - 1) + 3 * (o + 1) : (o + 1) + 3 * (s2 - 1);
                            const int ku = to_a ? s2 + 3 * (o + 1) : (o + 1) + 3 * s2;
#pragma unroll
                            for (int m = 0; m < NV; m++) {
                                acc[kl][m] = acc[kl][m] - h * v[k][m];
                                acc[ku][m] = acc[ku][m] + h * v[k][m];
                            }
                        }
                    }
                }
            }
            // nine ordered steps: in each, the wave's lanes add to distinct
            // cells (its lines share P and differ in Q)
            const int c = e - 1;
            const bool own = live && c >= 0 && c < tr;
            const int dbase = Q * dqs + P * dps + c * drs;
#pragma unroll
            for (int k = 0; k < 9; k++) {
                const int dQ = k % 3 - 1, dP = k / 3 - 1;
                const bool in = own && Q + dQ >= 0 && Q + dQ < tq && P + dP >= 0 && P + dP < tp;
                if (in) {
                    const int di = dbase + dQ * dqs + dP * dps;
#pragma unroll
                    for (int m = 0; m < NV; m++) {
                        T* x = &dq[m * ncell + di];
                        *x = *x + acc[k][m];
                    }
                }
                wave_sync();
            }
            __syncthreads();   // the next round's waves may touch these cells
            prow += wpr;
            const int carry = prow >= nrows;
            prow -= carry ? nrows : 0;
            pchunk += carry;
        }
    }
}

template <class Sol>
__global__ __launch_bounds__(TILE3_LANES, TILE3_WG_PER_CU) void sweep3_tile_kernel(Tile3Args a) {
    using T = float;
    constexpr int NV = Sol::NV;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int xcd = blockIdx.x % TILE3_XCDS, k = blockIdx.x / TILE3_XCDS;
    if (k >= a.xcd_count[xcd]) return;   // (the whole workgroup: no barrier reached)
    const Tile3 td = a.tiles[a.xcd_start[xcd] + k];
    const Patch3 pt = a.patches[td.patch];
    const int tx = td.tx, ty = td.ty, tz = td.tz;
    const int bx = tx + 4, by = ty + 4, bz = tz + 4;
    const int nbox = bx * by * bz, ncell = tx * ty * tz;
    T* qs = reinterpret_cast<T*>(smem_raw);   // [NV][bz][by][bx], local -2 .. t+1
    T* dq = qs + NV * nbox;                   // [NV][tz][ty][tx]
    const int tid = threadIdx.x;

    const int64_t mit = pt.mx + 4, mij = mit * (pt.my + 4);
    const int64_t plane = mij * (pt.mz + 4);
    const T* qin = static_cast<const T*>(a.qin) + pt.qoff;
    // array index of local cell (-2, -2, -2): 1-based cell i0 + 1 - 2 sits
    // at index i0 + 1 - 2 + 1 = i0 (two ghost cells)
    const int64_t g0 = (int64_t)td.k0 * mij + (int64_t)td.j0 * mit + td.i0;
    {
        Digits3 u(tid, bx, by, bz);
        for (int idx = tid; idx < NV * nbox; idx += TILE3_LANES, u.step())
            qs[idx] = qin[u.m * plane + g0 + (int64_t)u.z * mij + (int64_t)u.y * mit + u.x];
    }
    for (int idx = tid; idx < NV * ncell; idx += TILE3_LANES) dq[idx] = T(0);
    __syncthreads();

    const int wave = tid / 64, wl = tid - wave * 64;
    T cfl = T(0);

    tile3_sweep<Sol, 1>(a, qs, dq, nbox, ncell, tx, ty, tz, bx, by, wave, wl, cfl);
    tile3_sweep<Sol, 2>(a, qs, dq, nbox, ncell, tx, ty, tz, bx, by, wave, wl, cfl);
    tile3_sweep<Sol, 3>(a, qs, dq, nbox, ncell, tx, ty, tz, bx, by, wave, wl, cfl);

    // q(t + dt) = q(t) + dq on the tile's cells
    T* qout = static_cast<T*>(a.qout) + pt.qoff;
    {
        Digits3 u(tid, tx, ty, tz);
        for (int idx = tid; idx < NV * ncell; idx += TILE3_LANES, u.step()) {
            const T v = qs[u.m * nbox + (u.z + 2) * bx * by + (u.y + 2) * bx + (u.x + 2)] + dq[idx];
            qout[u.m * plane + g0 + (int64_t)(u.z + 2) * mij + (int64_t)(u.y + 2) * mit + (u.x + 2)] = v;
        }
    }
    double cd = double(cfl);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double o2 = __shfl_xor(cd, off, 64);
        cd = cd > o2 ? cd : o2;
    }
    if ((tid & 63) == 0) atomic_max_nonneg3(&a.cfl_patch[td.patch], cd);
}

}  // namespace amrgpu
