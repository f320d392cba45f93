// sweep3.hpp -- batched 3-D wave-propagation step over the patches of a level.
//
// step3.f:3-486 + flux3.f:4-633 + stepgrid.f (3-D):147-212 for every patch
// of a batch, in four launches:
//   phase 1, 2, 3: the x-, y- and z-sweeps (step3.f:140-195, :197-259,
//                  :261-321).  One workgroup per (patch, block of planes)
//                  marches through the planes P of its block in ascending
//                  order; in each plane it solves every line Q of the patch
//                  (ghost lines included, as step3 does), in rounds of
//                  256 / (n + 3) lines, one lane per interface.
//   update:        q -= dt/dx (fm(i+1) - fp(i)) + ... (stepgrid.f:150-173).
// The six flux arrays of each patch live in HBM, interior faces only.
//
// Phase d's lines run along direction r; Q indexes the lines of a plane and
// P the planes (x: r,Q,P = i,j,k; y: j,i,k; z: k,i,j).  flux3 turns an
// interface into the normal fluxes faddm/faddp and two families of
// transverse increments, gadd and hadd, each indexed (cell, s = 1|2, o =
// -1|0|1).  step3 adds one family into the faces normal to Q ("type A":
// face Q+s-1, cell P+o) and the other into the faces normal to P ("type B":
// cell Q+o, face P+s-1):
//   x: A <- gadd (gm/gp), B <- hadd (hm/hp)
//   y: A <- hadd (fm/fp), B <- gadd (hm/hp)
//   z: A <- gadd (fm/fp), B <- hadd (gm/gp)
// The reference visits lines in (P, Q) order, so every face receives its
// increments in ascending (P, Q) of the line they come from.  Here the
// planes are marched in that order and, within a round, the increments are
// applied in barrier-separated steps by Q-offset (+1, then 0, then -1): the
// same order, so an FP64 build without FMA contraction reproduces the
// reference bit for bit.  Within a line, gadd(cell i) is the sum of the
// increments interface i makes to its right cell, then those of interface
// i+1 to its left cell (flux3.f:450-505, 573-628): each lane passes its
// right-cell partial sums to the next lane through LDS.
//
// Faces that start at zero and have received only identical increments are
// kept once: during the x-sweep gm = gp and hm = hp (step3.f:221-250), so
// the x-sweep accumulates G in gm and H in hm; the y-sweep forms
// gm = G + faddm, gp = G + faddp and keeps accumulating H in hm; the z-sweep
// forms hm = H + faddm, hp = H + faddp.
#pragma once
#include <cstdint>
#include "riemann3.hpp"

namespace amrgpu {

constexpr int NWMAX3 = 4;
constexpr int SWEEP3_LANES = 256;

struct Patch3 {
    int64_t qoff;     // element offset of component 0 of q: [m][k][j][i], 2 ghosts
    int64_t foff;     // element offset of the patch's flux block
    int32_t mx, my, mz, pad;
};

// a work item: patch and a block [pb, pe) of planes (phases) or of linear
// cell indices (update)
struct Block3 {
    int32_t patch, pb, pe, pad;
};

struct Method3 {
    Physics3 ph;
    int32_t order;       // method(2)
    int32_t m3, m4;      // method(3) = 10 m3 + m4 (m3 = -1 when method(3) < 0)
    int32_t use_fwaves;
    int32_t mthlim[NWMAX3];
};

struct Sweep3Args {
    void* q;                        // patch solutions (compute precision T)
    void* face;                     // flux blocks
    const Patch3* patches;
    const Block3* blocks;
    int32_t nblocks;
    int32_t phase;                  // 1, 2, 3 (x, y, z)
    double dtd[3];                  // dt/dx, dt/dy, dt/dz
    Method3 meth;
    unsigned long long* cfl_patch;  // per-patch max Courant number (double bits)
};

// flux block of a patch: fm, fp [NV][mz][my][mx+1]; gm, gp [NV][mz][my+1][mx];
// hm, hp [NV][mz+1][my][mx] (interior faces, 1-based face and cell indices)
struct Faces3 {
    int64_t fm, fp, gm, gp, hm, hp;
    int64_t sx, sy, sz;   // component strides
};
__host__ __device__ inline Faces3 faces3(const Patch3& p, int nv) {
    Faces3 f;
    f.sx = (int64_t)(p.mx + 1) * p.my * p.mz;
    f.sy = (int64_t)p.mx * (p.my + 1) * p.mz;
    f.sz = (int64_t)p.mx * p.my * (p.mz + 1);
    f.fm = p.foff;
    f.fp = f.fm + nv * f.sx;
    f.gm = f.fp + nv * f.sx;
    f.gp = f.gm + nv * f.sy;
    f.hm = f.gp + nv * f.sy;
    f.hp = f.hm + nv * f.sz;
    return f;
}
__host__ __device__ inline int64_t faces3_elems(const Patch3& p, int nv) {
    Faces3 f = faces3(p, nv);
    return 2 * nv * (f.sx + f.sy + f.sz);
}
__host__ __device__ inline int64_t q3_elems(const Patch3& p, int nv) {
    return (int64_t)nv * (p.mx + 4) * (p.my + 4) * (p.mz + 4);
}

// linear index of 1-based coordinates (a, b, c) in the r, Q, P directions
struct Lin3 {
    int64_t base, sa, sb, sc;
    __device__ __forceinline__ int64_t at(int a, int b, int c) const {
        return base + sa * a + sb * b + sc * c;
    }
};
__device__ __forceinline__ Lin3 lin3(int64_t sa, int64_t sb, int64_t sc) {
    return Lin3{-(sa + sb + sc), sa, sb, sc};
}

__device__ __forceinline__ void atomic_max_nonneg3(unsigned long long* p, double v) {
    atomicMax(p, (unsigned long long)__double_as_longlong(v + 0.0));
}

template <class Mth, typename T>
__device__ __forceinline__ T phi_limiter3(int mth, T r) {
    switch (mth) {   // inlinelimiter.f:58-94
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

// LDS of a phase workgroup: the exchange buffer of one round (slots x 256
// lanes) and the staged states of the round's lines (y-, z-sweeps): at most
// lines * (n + 4) <= 256 + lines <= 320 cells, lines <= 64
template <int NV, int NW>
__host__ __device__ constexpr int sweep3_xs_slots() {
    return NW * NV > 6 * NV ? NW * NV : 6 * NV;
}
constexpr int SWEEP3_STAGE_CELLS = 320;
template <int NV, int NW, typename T>
__host__ __device__ constexpr size_t sweep3_lds_bytes() {
    return ((size_t)sweep3_xs_slots<NV, NW>() * SWEEP3_LANES + (size_t)NV * SWEEP3_STAGE_CELLS) *
           sizeof(T);
}

// D: the phase (1 x, 2 y, 3 z), a template parameter so that every
// component index of the solvers is a compile-time constant
template <class Sol, typename T, int D>
__global__ __launch_bounds__(SWEEP3_LANES, sizeof(T) == 4 ? 3 : 1) void sweep3_phase_kernel(Sweep3Args a) {
    constexpr int NV = Sol::NV, NW = Sol::NW;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* xs = reinterpret_cast<T*>(smem_raw);   // [slot][256]
    T* qst = xs + sweep3_xs_slots<NV, NW>() * SWEEP3_LANES;   // [NV][n+4][lines]
    const Method3& M = a.meth;

    const Block3 blk = a.blocks[blockIdx.x];
    const Patch3 pt = a.patches[blk.patch];
    constexpr int d = D;
    // SPLIT (FP64): fm and fp (gm and gp) accumulate the same transverse
    // increments separately, as the reference does.  Merged (FP32 bulk): the
    // increments go to fm (gm) only and fp (gp) holds faddp - faddm, so
    // fp = fm + (faddp - faddm) up to rounding: half the face traffic.
    constexpr bool SPLIT = sizeof(T) == 8;
    const int mx = pt.mx, my = pt.my, mz = pt.mz;
    const int64_t mit = mx + 4, mij = mit * (my + 4);
    const Faces3 F = faces3(pt, NV);
    T* q = static_cast<T*>(a.q) + pt.qoff;
    T* face = static_cast<T*>(a.face);

    // geometry of the phase (see the header comment)
    int n, NQ, NP;
    int64_t qr_s, qq_s, qp_s;           // q strides along r, Q, P
    Lin3 N, A, B;                       // normal faces, type A, type B
    T *n1, *n2, *a1, *a2, *b1, *b2;     // arrays (a2/b2 null: single)
    int64_t ncs, acs, bcs;
    bool g_to_a;
    const int64_t xs_j = mx + 1, xs_k = (int64_t)my * (mx + 1);   // x-face strides
    const int64_t ys_J = mx, ys_k = (int64_t)(my + 1) * mx;       // y-face
    const int64_t zs_j = mx, zs_K = (int64_t)my * mx;             // z-face
    if (d == 1) {
        n = mx; NQ = my; NP = mz;
        qr_s = 1; qq_s = mit; qp_s = mij;
        N = lin3(1, xs_j, xs_k); n1 = face + F.fm; n2 = face + F.fp; ncs = F.sx;
        A = lin3(1, ys_J, ys_k); a1 = face + F.gm; a2 = nullptr; acs = F.sy;
        B = lin3(1, zs_j, zs_K); b1 = face + F.hm; b2 = nullptr; bcs = F.sz;
        g_to_a = true;
    } else if (d == 2) {
        n = my; NQ = mx; NP = mz;
        qr_s = mit; qq_s = 1; qp_s = mij;
        N = lin3(ys_J, 1, ys_k); n1 = face + F.gm; n2 = face + F.gp; ncs = F.sy;
        A = lin3(xs_j, 1, xs_k); a1 = face + F.fm; a2 = SPLIT ? face + F.fp : nullptr; acs = F.sx;
        B = lin3(zs_j, 1, zs_K); b1 = face + F.hm; b2 = nullptr; bcs = F.sz;
        g_to_a = false;
    } else {
        n = mz; NQ = mx; NP = my;
        qr_s = mij; qq_s = 1; qp_s = mit;
        N = lin3(zs_K, 1, zs_j); n1 = face + F.hm; n2 = face + F.hp; ncs = F.sz;
        A = lin3(xs_k, 1, xs_j); a1 = face + F.fm; a2 = SPLIT ? face + F.fp : nullptr; acs = F.sx;
        B = lin3(ys_k, 1, ys_J); b1 = face + F.gm; b2 = SPLIT ? face + F.gp : nullptr; bcs = F.sy;
        g_to_a = true;
    }
    const T dtdx = T(a.dtd[d - 1]);
    const T dtdy_p = T(a.dtd[d % 3]);         // flux3's dtdy argument
    const T dtdz_p = T(a.dtd[(d + 1) % 3]);   // flux3's dtdz argument
    const T half_t = T(0.5) * dtdx;
    const T sixth = T(1) / T(6);
    const T g6 = sixth * dtdx * dtdz_p;       // gadd's 1/6-weighted factor
    const T h6 = sixth * dtdx * dtdy_p;       // hadd's
    const bool trans = M.m3 > 0;
    const bool cq = M.m3 >= 2;
    const bool tt = M.m4 > 0;
    bool limit = false;
    for (int w = 0; w < NW; w++) limit = limit || M.mthlim[w] > 0;

    const int L = n + 3;                      // lanes per line: interfaces 0..n+2
    const int lines_per = SWEEP3_LANES / L;
    const int tid = threadIdx.x;
    const int qi = tid / L, r = tid - qi * L;
    const int nrounds = (NQ + 2 + lines_per - 1) / lines_per;
    const int pe_face = blk.pe == NP + 1 ? NP + 2 : blk.pe;   // owned type-B faces
    T cfl = T(0);
    if (D == 1 && trans) {
        // G and H start at zero (step3.f:134-139): clear the faces this
        // workgroup owns (k in [pb, pe) of gm; K in [pb, pe_face) of hm)
        const int64_t g0 = (int64_t)(blk.pb - 1) * (my + 1) * mx, g1 = (int64_t)(blk.pe - 1) * (my + 1) * mx;
        const int64_t h0 = (int64_t)(blk.pb - 1) * my * mx, h1 = (int64_t)(pe_face - 1) * my * mx;
        for (int m = 0; m < NV; m++) {
            for (int64_t e = g0 + tid; e < g1; e += SWEEP3_LANES) a1[m * acs + e] = T(0);
            for (int64_t e = h0 + tid; e < h1; e += SWEEP3_LANES) b1[m * bcs + e] = T(0);
        }
        __syncthreads();
    }

    for (int P = blk.pb - 1; P <= blk.pe; P++) {
        for (int rd = 0; rd < nrounds; rd++) {
            const int Q = rd * lines_per + qi;
            const bool live = qi < lines_per && Q <= NQ + 1;
            const bool inner = live && r >= 1 && r <= n + 1;
            const int Qc = live ? Q : 0, rc = live ? r : 1;
            // lines of this round (the last round may hold fewer)
            const int nl = min(lines_per, NQ + 2 - rd * lines_per);
            const int64_t plane = (int64_t)(pt.mx + 4) * (pt.my + 4) * (pt.mz + 4);
            T ql[NV], qrr[NV];   // states left (cell r-1) and right (cell r) of interface r
            if (D == 1) {
                // lanes run along x: direct loads coalesce.  1-based cell x
                // sits at array index x + 1 (two ghost cells)
                const int64_t c0 = (int64_t)rc * qr_s + (int64_t)(Qc + 1) * qq_s +
                                   (int64_t)(P + 1) * qp_s;
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    ql[m] = q[m * plane + c0];
                    qrr[m] = q[m * plane + c0 + qr_s];
                }
            } else {
                // lines run along y or z: stage the round's cells through LDS
                // with consecutive lanes on consecutive lines (= consecutive x)
                const int rows = n + 4, ncell = nl * rows;
                const int64_t base = (int64_t)(rd * lines_per + 1) * qq_s + (int64_t)(P + 1) * qp_s;
                __syncthreads();   // previous round's readers are done
                for (int idx = tid; idx < NV * ncell; idx += SWEEP3_LANES) {
                    const int m = idx / ncell, rem = idx - m * ncell;
                    const int row = rem / nl, qq = rem - row * nl;
                    qst[idx] = q[m * plane + base + (int64_t)row * qr_s + (int64_t)qq * qq_s];
                }
                __syncthreads();
                const int qq = live ? qi : 0;
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    ql[m] = qst[(m * rows + rc) * nl + qq];
                    qrr[m] = qst[(m * rows + rc + 1) * nl + qq];
                }
            }
            const typename Sol::Cache cache = Sol::prep(M.ph, ql, qrr);
            T wave[NW][NV], s[NW], amdq[NV], apdq[NV];
            Sol::rpn3(M.ph, d, ql, qrr, cache, wave, s, amdq, apdq);
            if (inner) {
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
                    __syncthreads();
#pragma unroll
                    for (int w = 0; w < NW; w++)
#pragma unroll
                        for (int m = 0; m < NV; m++) xs[(w * NV + m) * SWEEP3_LANES + tid] = wave[w][m];
                    __syncthreads();
                    if (inner) {
#pragma unroll
                        for (int w = 0; w < NW; w++) {
                            const int mth = M.mthlim[w];
                            if (mth == 0) continue;
                            T wn2 = T(0), dotr = T(0), dotl = T(0);
#pragma unroll
                            for (int m = 0; m < NV; m++) {
                                const T* x = &xs[(w * NV + m) * SWEEP3_LANES + tid];
                                wn2 = wn2 + wave[w][m] * wave[w][m];
                                dotr = dotr + wave[w][m] * x[1];
                                dotl = dotl + x[-1] * wave[w][m];
                            }
                            if (wn2 == T(0)) continue;
                            using Mth = typename Sol::Math;
                            T rr = Mth::ratio(s[w] > T(0) ? dotl : dotr, wn2);
                            T wl = phi_limiter3<Mth>(mth, rr);
#pragma unroll
                            for (int m = 0; m < NV; m++) wave[w][m] = wl * wave[w][m];
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
                        c = c + T(0.5) * fac * (T(1) - fabs(s[w]) * dtdxave) * wave[w][m];
                    }
                    cqxx[m] = c;
                    fadm[m] = fadm[m] + c;
                    fadp[m] = fadp[m] + c;
                }
            }
            // normal fluxes of the owned faces (plane P in [pb, pe), line Q in
            // [1, NQ]).  y-, z-sweeps: through LDS to lanes across the lines,
            // so that the face accesses run along x
            if (D == 1) {
                if (inner && Q >= 1 && Q <= NQ && P >= blk.pb && P < blk.pe) {
                    const int64_t idx = N.at(r, Q, P);
#pragma unroll
                    for (int m = 0; m < NV; m++) {
                        n1[idx + m * ncs] = fadm[m];
                        n2[idx + m * ncs] = SPLIT ? fadp[m] : fadp[m] - fadm[m];
                    }
                }
            } else {
                __syncthreads();   // xs free
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    xs[m * SWEEP3_LANES + tid] = fadm[m];
                    xs[(NV + m) * SWEEP3_LANES + tid] = fadp[m];
                }
                __syncthreads();
                const int Q0 = rd * lines_per;
                for (int t = tid; t < nl * (n + 1); t += SWEEP3_LANES) {
                    const int qq = t % nl, rr = t / nl + 1, Qn = Q0 + qq;
                    if (Qn < 1 || Qn > NQ || P < blk.pb || P >= blk.pe) continue;
                    const int64_t idx = N.at(rr, Qn, P);
                    const int src = qq * L + rr;
#pragma unroll
                    for (int m = 0; m < NV; m++) {
                        const int64_t k = idx + m * ncs;
                        const T x0 = trans ? n1[k] : T(0);
                        const T fm_ = xs[m * SWEEP3_LANES + src];
                        const T fp_ = xs[(NV + m) * SWEEP3_LANES + src];
                        n1[k] = x0 + fm_;
                        // y-sweep, merged form: gp keeps faddp - faddm (the
                        // z-sweep adds its increments to gm only)
                        n2[k] = (SPLIT || D == 3) ? x0 + fp_ : fp_ - fm_;
                    }
                }
            }
            if (!trans) continue;

            // ---- transverse solves, two families (flux3.f:302-628) ----
            // fam 0 = gadd: primary split along icoor 2 (rpt3 of A-+dq and of
            // cqxx), rptt3 along 2 of the icoor-3 splits; fam 1 = hadd: the
            // same with 2 and 3 exchanged.  The splits one family shares with
            // the other are recomputed (same inputs, same results) rather than
            // held, which keeps the FP64 build within the register file.
#pragma unroll
            for (int fam = 0; fam < 2; fam++) {
                const int icp = fam == 0 ? 2 : 3;   // primary direction
                const int ics = fam == 0 ? 3 : 2;   // the other one
                const T w6 = fam == 0 ? g6 : h6;
                T xcm[NV], xcp[NV], scm[NV], scp[NV];   // cqxx split along icp, ics
                T pm[NV], pp[NV], sm[NV], sp[NV], tmp[NV], rm[NV], rp[NV];
                if (cq) {
                    Sol::split(M.ph, d, icp, ql, qrr, cache, cqxx, xcm, xcp);
                    if (tt) Sol::split(M.ph, d, ics, ql, qrr, cache, cqxx, scm, scp);
                }
                // right-cell partials of this interface: slot k = (s-1)*3 + (o+1),
                // statement order of flux3.f:450-505 (A+dq side)
                T v[6][NV];
                Sol::split(M.ph, d, icp, ql, qrr, cache, apdq, pm, pp);
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    v[1][m] = T(0) - half_t * pm[m];
                    v[4][m] = T(0) - half_t * pp[m];
                }
                if (tt) {
                    Sol::split(M.ph, d, ics, ql, qrr, cache, apdq, sm, sp);
#pragma unroll
                    for (int m = 0; m < NV; m++) tmp[m] = M.m4 == 2 ? sp[m] - T(3) * scp[m] : sp[m];
                    Sol::split(M.ph, d, icp, ql, qrr, cache, tmp, rm, rp);      // "p" part
#pragma unroll
                    for (int m = 0; m < NV; m++) tmp[m] = M.m4 == 2 ? sm[m] - T(3) * scm[m] : sm[m];
                    Sol::split(M.ph, d, icp, ql, qrr, cache, tmp, sm, sp);      // "m" part
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
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    if (cq) {
                        v[4][m] = v[4][m] + dtdx * xcp[m];
                        v[1][m] = v[1][m] + dtdx * xcm[m];
                    }
                }
                __syncthreads();   // xs free
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    if (!tt && k != 1 && k != 4) continue;
#pragma unroll
                    for (int m = 0; m < NV; m++) xs[(k * NV + m) * SWEEP3_LANES + tid] = v[k][m];
                }
                __syncthreads();
                // finish cell r-1: the left neighbour's partials, then this
                // interface's left-cell increments (A-dq side)
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    if (!tt && k != 1 && k != 4) continue;
#pragma unroll
                    for (int m = 0; m < NV; m++)
                        v[k][m] = xs[(k * NV + m) * SWEEP3_LANES + tid - (tid > 0)];
                }
                Sol::split(M.ph, d, icp, ql, qrr, cache, amdq, pm, pp);
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    v[1][m] = v[1][m] - half_t * pm[m];
                    v[4][m] = v[4][m] - half_t * pp[m];
                }
                if (tt) {
                    Sol::split(M.ph, d, ics, ql, qrr, cache, amdq, sm, sp);
#pragma unroll
                    for (int m = 0; m < NV; m++) tmp[m] = M.m4 == 2 ? sp[m] + T(3) * scp[m] : sp[m];
                    Sol::split(M.ph, d, icp, ql, qrr, cache, tmp, rm, rp);
#pragma unroll
                    for (int m = 0; m < NV; m++) tmp[m] = M.m4 == 2 ? sm[m] + T(3) * scm[m] : sm[m];
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
#pragma unroll
                for (int m = 0; m < NV; m++) {
                    if (cq) {
                        v[4][m] = v[4][m] - dtdx * xcp[m];
                        v[1][m] = v[1][m] - dtdx * xcm[m];
                    }
                }
                // ---- apply to the faces (step3.f:221-250 and analogues) ----
                // The round's finished cells go to LDS; then one lane per
                // target face of the round (targets one line beyond the
                // round's on each side) loads the face once per plane offset,
                // adds this round's increments in ascending line order and
                // stores it: the reference's order with a single access per
                // face.  Consecutive lanes take consecutive x (cells c for the
                // x-sweep, lines for the y-, z-sweeps), so the accesses coalesce.
                __syncthreads();   // the partial sums have been read
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    if (!tt && k != 1 && k != 4) continue;
#pragma unroll
                    for (int m = 0; m < NV; m++) xs[(k * NV + m) * SWEEP3_LANES + tid] = v[k][m];
                }
                __syncthreads();
                const int Q0 = rd * lines_per;
                const bool to_a = (fam == 0) == g_to_a;
                const int ntq = nl + 2, ntarget = ntq * n;
                // increment slot k of the cell cc of round line `line`
                auto inc = [&](int line, int cc, int k, int m) -> T {
                    return xs[(k * NV + m) * SWEEP3_LANES + (line - Q0) * L + cc + 1];
                };
                for (int t = tid; t < ntarget; t += SWEEP3_LANES) {
                    int qq, cc;
                    if (D == 1) { cc = t % n + 1; qq = t / n; }
                    else { qq = t % ntq; cc = t / ntq + 1; }
                    const int Qt = Q0 - 1 + qq;
                    if (to_a) {
                        // face Qt normal to Q: line Qt-1 (s = 2), then line Qt (s = 1)
                        const bool lo = Qt - 1 >= Q0 && Qt - 1 < Q0 + nl;
                        const bool hi = Qt >= Q0 && Qt < Q0 + nl;
                        if (Qt < 1 || Qt > NQ + 1 || !(lo || hi)) continue;
#pragma unroll
                        for (int o = -1; o <= 1; o++) {
                            if (o != 0 && !tt) continue;
                            const int Pc = P + o;
                            if (Pc < blk.pb || Pc >= blk.pe) continue;
                            const int64_t idx = A.at(cc, Qt, Pc);
#pragma unroll
                            for (int m = 0; m < NV; m++) {
                                const int64_t e = idx + m * acs;
                                T x1 = a1[e], x2 = a2 ? a2[e] : T(0);
                                if (lo) {
                                    const T w = inc(Qt - 1, cc, 3 + o + 1, m);
                                    x1 = x1 + w;
                                    x2 = x2 + w;
                                }
                                if (hi) {
                                    const T w = inc(Qt, cc, o + 1, m);
                                    x1 = x1 + w;
                                    x2 = x2 + w;
                                }
                                a1[e] = x1;
                                if (a2) a2[e] = x2;
                            }
                        }
                    } else {
                        // face P+s-1 normal to P, cell Qt: lines Qt-1 (o = +1),
                        // Qt (o = 0), Qt+1 (o = -1)
                        if (Qt < 1 || Qt > NQ) continue;
                        bool any = false;
#pragma unroll
                        for (int o = 1; o >= -1; o--)
                            any = any || ((o == 0 || tt) && Qt - o >= Q0 && Qt - o < Q0 + nl);
                        if (!any) continue;
#pragma unroll
                        for (int s2 = 1; s2 <= 2; s2++) {
                            const int Pf = P + s2 - 1;
                            if (Pf < blk.pb || Pf >= pe_face) continue;
                            const int64_t idx = B.at(cc, Qt, Pf);
#pragma unroll
                            for (int m = 0; m < NV; m++) {
                                const int64_t e = idx + m * bcs;
                                T x1 = b1[e], x2 = b2 ? b2[e] : T(0);
#pragma unroll
                                for (int o = 1; o >= -1; o--) {
                                    const int line = Qt - o;
                                    if (!((o == 0 || tt) && line >= Q0 && line < Q0 + nl)) continue;
                                    const T w = inc(line, cc, (s2 - 1) * 3 + o + 1, m);
                                    x1 = x1 + w;
                                    x2 = x2 + w;
                                }
                                b1[e] = x1;
                                if (b2) b2[e] = x2;
                            }
                        }
                    }
                }
            }
        }
    }
    // per-patch Courant number: wave maximum, one atomic per wave
    double cd = double(cfl);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double o2 = __shfl_xor(cd, off, 64);
        cd = cd > o2 ? cd : o2;
    }
    if ((tid & 63) == 0) atomic_max_nonneg3(&a.cfl_patch[blk.patch], cd);
}

// conservative update, stepgrid.f (3-D):150-173 (no capacity), in place
template <int NV, typename T>
__global__ __launch_bounds__(256) void update3_kernel(T* q, const T* face, const Patch3* patches,
                                                      const Block3* blocks, double dtdx_d,
                                                      double dtdy_d, double dtdz_d) {
    const Block3 blk = blocks[blockIdx.x];
    const int cidx = blk.pb + (int)threadIdx.x;
    if (cidx >= blk.pe) return;
    const Patch3 pt = patches[blk.patch];
    const int mx = pt.mx, my = pt.my;
    const int i = cidx % mx + 1, t = cidx / mx, j = t % my + 1, k = t / my + 1;
    const Faces3 F = faces3(pt, NV);
    const int64_t mit = mx + 4, mij = mit * (my + 4);
    const int64_t plane = mij * (pt.mz + 4);
    T* qc = q + pt.qoff + (k + 1) * mij + (j + 1) * mit + (i + 1);
    const int64_t xf = ((int64_t)(k - 1) * my + (j - 1)) * (mx + 1) + (i - 1);
    const int64_t yf = ((int64_t)(k - 1) * (my + 1) + (j - 1)) * mx + (i - 1);
    const int64_t zf = ((int64_t)(k - 1) * my + (j - 1)) * mx + (i - 1);
    const T dtdx = T(dtdx_d), dtdy = T(dtdy_d), dtdz = T(dtdz_d);
#pragma unroll
    for (int m = 0; m < NV; m++) {
        const T fmr = face[F.fm + m * F.sx + xf + 1];
        const T gmr = face[F.gm + m * F.sy + yf + mx];
        T fpl = face[F.fp + m * F.sx + xf], gpl = face[F.gp + m * F.sy + yf];
        if (sizeof(T) != 8) {   // merged form: fp, gp hold faddp - faddm
            fpl = face[F.fm + m * F.sx + xf] + fpl;
            gpl = face[F.gm + m * F.sy + yf] + gpl;
        }
        const T hmr = face[F.hm + m * F.sz + zf + (int64_t)mx * my], hpl = face[F.hp + m * F.sz + zf];
        T v = qc[m * plane];
        v = v - dtdx * (fmr - fpl) - dtdy * (gmr - gpl) - dtdz * (hmr - hpl);
        qc[m * plane] = v;
    }
}

}  // namespace amrgpu
