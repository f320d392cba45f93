// amr3.hip -- kernels of the 3-D level update around the step: one stage of
// a ghost-fill program, the coarse-fine registers (qad, fluxsv, fluxad) and
// refluxing / averaging (upbnd, update).  Built with -ffp-contract=off: the
// time blend, the interpolation, the register sums and the averages are
// evaluated as the reference's FP64 code (src/3d).
#include <hip/hip_runtime.h>
#include "amr3.hpp"
#include "riemann3.hpp"

namespace amrgpu {

namespace {

__device__ __forceinline__ int64_t qplane3(const Patch3& p) {
    return (int64_t)(p.mx + 4) * (p.my + 4) * (p.mz + 4);
}

template <typename S>
struct Buf3 {
    S* lvl;          // kind 0
    double* scr;     // kind 1
    int64_t off, plane;
    int32_t nrow, ncol;
    __device__ int64_t at(int m, int i, int j, int k) const {
        return off + (int64_t)m * plane + ((int64_t)k * ncol + j) * nrow + i;
    }
    __device__ double load(int m, int i, int j, int k) const {
        const int64_t a = at(m, i, j, k);
        return lvl ? (double)lvl[a] : scr[a];
    }
    __device__ void store(int m, int i, int j, int k, double v) const {
        const int64_t a = at(m, i, j, k);
        if (lvl) lvl[a] = (S)v; else scr[a] = v;
    }
};

template <typename S>
__device__ Buf3<S> buf3(const Fill3Args& a, int b) {
    const Fill3Buf fb = a.bufs[b];
    Buf3<S> r;
    if (fb.kind == 0) {
        r.lvl = static_cast<S*>(a.lvl_new[fb.level]);
        r.scr = nullptr;
        r.off = a.lvl_patches[fb.level][fb.patch].qoff;
    } else {
        r.lvl = nullptr;
        r.scr = fb.kind == 1 ? a.scratch : a.arena;
        r.off = fb.off;
    }
    r.plane = fb.plane;
    r.nrow = fb.nrow;
    r.ncol = fb.ncol;
    return r;
}

// MC-limited slope of filpatch.f (3-D):183-207
__device__ __forceinline__ double mc_slope(double vp, double vc, double vm) {
    const double dpc = vp - vc, dmc = vc - vm, dcc = vp - vm;
    double d = fmin(fabs(dpc), fabs(dmc));
    d = fmin(2.0 * d, .5 * fabs(dcc));
    const double f = fmax(0.0, copysign(1.0, dpc * dmc));
    return d * copysign(1.0, dcc) * f;
}

}  // namespace

template <typename S>
__global__ __launch_bounds__(256) void fill3_stage_kernel(Fill3Args a) {
    const Fill3Work w = a.work[blockIdx.x];
    const Fill3Op op = a.ops[w.op];
    const int tid = threadIdx.x, nv = a.nv;
    const Buf3<S> dst = buf3<S>(a, op.dst);
    const int64_t ncell = (int64_t)op.ni * op.nj * op.nk;

    if (op.type == F3_COPY) {
        const Patch3 sp = a.lvl_patches[op.src_level][op.src];
        const S* qn = static_cast<const S*>(a.lvl_new[op.src_level]) + sp.qoff;
        const S* qo = static_cast<const S*>(a.lvl_old[op.src_level]) + sp.qoff;
        const int mode = (op.flags & F3F_FORCE_NEW) ? T3_NEW
                         : (op.flags & F3F_FORCE_OLD) ? T3_OLD : a.tmode[op.src_level];
        const double ai = a.alphai[op.src_level], ac = a.alphac[op.src_level];
        const int64_t smit = sp.mx + 4, smij = smit * (sp.my + 4), spl = qplane3(sp);
        for (int64_t t = w.lo + tid; t < w.hi; t += 256) {
            const int m = (int)(t / ncell);
            const int64_t c = t - m * ncell;
            const int di = (int)(c % op.ni), r = (int)(c / op.ni);
            const int dj = r % op.nj, dk = r / op.nj;
            const int64_t s = m * spl + (op.sk0 + dk) * smij + (op.sj0 + dj) * smit + (op.si0 + di);
            double v;
            if (mode == T3_NEW) v = (double)qn[s];
            else if (mode == T3_OLD) v = (double)qo[s];
            else v = (double)qn[s] * ai + (double)qo[s] * ac;
            dst.store(m, op.di0 + di, op.dj0 + dj, op.dk0 + dk, v);
        }
    } else if (op.type == F3_INTERP) {
        const Buf3<S> crs = buf3<S>(a, op.src);
        const int lf = op.level, lc = op.level - 1;
        const double hxf = a.hx[lf], hyf = a.hy[lf], hzf = a.hz[lf];
        const double hxc = a.hx[lc], hyc = a.hy[lc], hzc = a.hz[lc];
        const double xlc = a.xlower + op.iplo * hxc, ybc = a.ylower + op.jplo * hyc;
        const double zfc = a.zlower + op.kplo * hzc;
        const double ratiox = (double)op.ratx, ratioy = (double)op.raty, ratioz = (double)op.ratz;
        const unsigned char* mask = a.masks + op.mask_off;
        for (int64_t c = w.lo + tid; c < w.hi; c += 256) {
            if (mask[c]) continue;
            const int iff = (int)(c % op.ni) + 1, r = (int)(c / op.ni);
            const int jf = r % op.nj + 1, kf = r / op.nj + 1;   // 1-based in the box
            const int ic = (int)floor((double)(iff + op.ilo - 1) / ratiox) - op.iplo + 1;
            const double eta1 = (a.xlower + ((double)(iff - 1 + op.ilo) + .5) * hxf -
                                 (xlc + ((double)ic - .5) * hxc)) / hxc;
            const int jc = (int)floor((double)(jf + op.jlo - 1) / ratioy) - op.jplo + 1;
            const double eta2 = (a.ylower + ((double)(jf - 1 + op.jlo) + .5) * hyf -
                                 (ybc + ((double)jc - .5) * hyc)) / hyc;
            const int kc = (int)floor((double)(kf + op.klo - 1) / ratioz) - op.kplo + 1;
            const double eta3 = (a.zlower + ((double)(kf - 1 + op.klo) + .5) * hzf -
                                 (zfc + ((double)kc - .5) * hzc)) / hzc;
            const int i0 = ic - 1, j0 = jc - 1, k0 = kc - 1;   // 0-based
            for (int n = 0; n < nv; n++) {
                const double valc = crs.load(n, i0, j0, k0);
                const double us = mc_slope(crs.load(n, i0 + 1, j0, k0), valc, crs.load(n, i0 - 1, j0, k0));
                const double vs = mc_slope(crs.load(n, i0, j0 + 1, k0), valc, crs.load(n, i0, j0 - 1, k0));
                const double ws = mc_slope(crs.load(n, i0, j0, k0 + 1), valc, crs.load(n, i0, j0, k0 - 1));
                const double valint = valc + eta1 * us + eta2 * vs + eta3 * ws;
                dst.store(n, op.di0 + iff - 1, op.dj0 + jf - 1, op.dk0 + kf - 1, valint);
            }
        }
    } else if (op.type == F3_COPYBACK) {
        const Buf3<S> src = buf3<S>(a, op.src);
        for (int64_t t = w.lo + tid; t < w.hi; t += 256) {
            const int m = (int)(t / ncell);
            const int64_t c = t - m * ncell;
            const int di = (int)(c % op.ni), r = (int)(c / op.ni);
            const int dj = r % op.nj, dk = r / op.nj;
            dst.store(m, op.di0 + di, op.dj0 + dj, op.dk0 + dk, src.load(m, di, dj, dk));
        }
    } else if (op.type == F3_RING) {
        // cstore.f (3-D): sides 1 (i = 1), 2 (j = ncol), 3 (i = nrow),
        // 4 (j = 1) over interior (k, j|i); 5 (k = 1), 6 (k = nfil) over (j, i)
        const Buf3<S> src = buf3<S>(a, op.src);
        const RegGrid3 rg = a.lvl_regs[op.level][op.si0];
        double* qc1d = a.lvl_reg[op.level] + rg.reg_off + (int64_t)nv * rg.lenbc;
        const int nr = op.ni - 2, nc = op.nj - 2, nf = op.nk - 2;
        const int s1 = nf * nc, s2 = nf * nr, s5 = nc * nr;
        const int len = 2 * (s1 + s2 + s5);
        for (int t = tid; t < len * nv; t += 256) {
            const int idx = t / nv, m = t - idx * nv;
            int i, j, k, x = idx;
            if (x < s1) { k = 1 + x / nc; j = 1 + x % nc; i = 0; }
            else if ((x -= s1) < s2) { k = 1 + x / nr; i = 1 + x % nr; j = op.nj - 1; }
            else if ((x -= s2) < s1) { k = 1 + x / nc; j = 1 + x % nc; i = op.ni - 1; }
            else if ((x -= s1) < s2) { k = 1 + x / nr; i = 1 + x % nr; j = 0; }
            else if ((x -= s2) < s5) { j = 1 + x / nr; i = 1 + x % nr; k = 0; }
            else { x -= s5; j = 1 + x / nr; i = 1 + x % nr; k = op.nk - 1; }
            qc1d[(int64_t)idx * nv + m] = src.load(m, i, j, k);
        }
    } else if (op.type == F3_COARSEN) {
        // coarsen.f (3-D): dst(i, j, k) = mean of src(2i..2i+1, 2j.., 2k..)
        const Buf3<S> src = buf3<S>(a, op.src);
        for (int64_t t = w.lo + tid; t < w.hi; t += 256) {
            const int m = (int)(t / ncell);
            const int64_t c = t - m * ncell;
            const int i = (int)(c % op.ni), r = (int)(c / op.ni);
            const int j = r % op.nj, k = r / op.nj;
            const int fi = 2 * i, fj = 2 * j, fk = 2 * k;
            const double v = (src.load(m, fi, fj, fk) + src.load(m, fi + 1, fj, fk) +
                              src.load(m, fi, fj + 1, fk) + src.load(m, fi + 1, fj + 1, fk) +
                              src.load(m, fi, fj, fk + 1) + src.load(m, fi + 1, fj, fk + 1) +
                              src.load(m, fi, fj + 1, fk + 1) + src.load(m, fi + 1, fj + 1, fk + 1)) / 8.0;
            dst.store(m, i, j, k, v);
        }
    } else if (op.type == F3_BC) {
        // bc3amr.f90: x-lo, x-hi, y-lo, y-hi, z-lo, z-hi in turn, each
        // complete before the next (later sides read cells earlier ones set)
        const int n[3] = {op.ni, op.nj, op.nk};
        for (int s = 0; s < 6; s++) {
            const int v = op.bc[s];
            const int d = s / 2, mb = a.mthbc[s];
            if (v > 0 && (mb == 1 || mb == 3)) {
                const int lo = (s & 1) ? v - 1 : 0;              // first cell set (0-based)
                const int cnt = (s & 1) ? n[d] - v + 1 : v;
                const int64_t other = (int64_t)n[0] * n[1] * n[2] / n[d];
                const int neg = d + 1;                           // component negated by a wall
                for (int64_t t = tid; t < other * cnt; t += 256) {
                    const int c = (int)(t % cnt);
                    const int64_t o = t / cnt;
                    const int dst_c = lo + c;
                    int src_c;
                    if (mb == 1) src_c = (s & 1) ? v - 2 : v;
                    else src_c = (s & 1) ? 2 * v - 1 - (dst_c + 1) - 1 : 2 * v - 1 - dst_c;
                    int ci[3], cs[3];
                    // the two other directions from o (lower direction fastest)
                    const int da = d == 0 ? 1 : 0, db = d == 2 ? 1 : 2;
                    ci[da] = (int)(o % n[da]);
                    ci[db] = (int)(o / n[da]);
                    cs[da] = ci[da];
                    cs[db] = ci[db];
                    ci[d] = dst_c;
                    cs[d] = src_c;
                    for (int m = 0; m < nv; m++) {
                        double val = dst.load(m, cs[0], cs[1], cs[2]);
                        if (mb == 3 && m == neg) val = -val;
                        dst.store(m, ci[0], ci[1], ci[2], val);
                    }
                }
            }
            __syncthreads();
        }
    }
}

// qad.f (3-D): one register slot per thread, the Riemann problems of its
// r x r fine faces in the reference's order (fine plane, then position
// along the line), fluctuations times the face area and dt added to svdflx
template <class Sol, typename S>
__global__ __launch_bounds__(128) void qad3_kernel(Reg3Args a, Physics3 ph) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nitems) return;
    constexpr int NV = Sol::NV, NW = Sol::NW;
    const Qad3Item it = a.qad[t];
    const Patch3 pt = a.fpatches[it.grid];
    const RegGrid3 rg = a.fregs[it.grid];
    double* svd = a.freg + rg.reg_off + (int64_t)NV * (it.slot - 1);
    const double* qc = a.freg + rg.reg_off + (int64_t)NV * rg.lenbc + (int64_t)NV * (it.slot - 1);
    const S* q = static_cast<const S*>(a.fq) + pt.qoff;
    const int64_t mit = pt.mx + 4, mij = mit * (pt.my + 4), pl = qplane3(pt);
    const int side = it.side;
    const bool xline = side == 1 || side == 3, zside = side >= 5;
    const int rl = xline ? a.raty : a.ratx;               // along the line
    const int ro = zside ? a.raty : a.ratz;               // across the planes
    const int ixyz = side <= 4 ? (xline ? 1 : 2) : 3;
    const double a1 = xline ? a.hy : a.hx, a2 = xline ? a.hz : (zside ? a.hy : a.hz);
    const bool coarse_left = side == 1 || side == 4 || side == 5;
    double cst[NV];
    for (int m = 0; m < NV; m++) cst[m] = qc[m];
    double acc[NV];
    for (int m = 0; m < NV; m++) acc[m] = svd[m];
    for (int o = 1; o <= ro; o++) {
        const int oc = (it.c2 - 1) * ro + o;                // fine plane (1-based interior)
        for (int l = 1; l <= rl; l++) {
            const int lc = (it.c1 - 1) * rl + l;            // fine cell along the line
            int i, j, k;                                    // 0-based with ghosts
            switch (side) {
            case 1: i = 1; j = lc + 1; k = oc + 1; break;
            case 2: i = lc + 1; j = pt.my + 2; k = oc + 1; break;
            case 3: i = pt.mx + 2; j = lc + 1; k = oc + 1; break;
            case 4: i = lc + 1; j = 1; k = oc + 1; break;
            case 5: i = lc + 1; j = oc + 1; k = 1; break;
            default: i = lc + 1; j = oc + 1; k = pt.mz + 2; break;
            }
            double fine[NV];
            for (int m = 0; m < NV; m++) fine[m] = (double)q[m * pl + k * mij + j * mit + i];
            const double* L = coarse_left ? cst : fine;
            const double* R = coarse_left ? fine : cst;
            typename Sol::Cache cache = Sol::prep(ph, L, R);
            double wave[NW][NV], s[NW], amdq[NV], apdq[NV];
            Sol::rpn3(ph, ixyz, L, R, cache, wave, s, amdq, apdq);
            for (int m = 0; m < NV; m++) {
                if (coarse_left)
                    acc[m] = acc[m] + amdq[m] * a1 * a2 * a.delt + apdq[m] * a1 * a2 * a.delt;
                else
                    acc[m] = acc[m] - amdq[m] * a1 * a2 * a.delt - apdq[m] * a1 * a2 * a.delt;
            }
        }
    }
    for (int m = 0; m < NV; m++) svd[m] = acc[m];
}

namespace {

// true flux of a face array value (the FP32 step keeps fp - fm, gp - gm)
template <typename T>
__device__ __forceinline__ double face_p(const T* face, int64_t pm, int64_t pp, bool merged) {
    if (merged) return (double)(T)(face[pm] + face[pp]);
    return (double)face[pp];
}

}  // namespace

// fluxsv.f (3-D): each entry overwrites its register slot with minus the
// coarse flux times dt and face area (entries a later one overwrites skip)
template <typename T>
__global__ __launch_bounds__(256) void fluxsv3_kernel(Reg3Args a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nsv * a.nv) return;
    const int e = t / a.nv, m = t - e * a.nv;
    const Sv3Entry sv = a.sv[e];
    if (!sv.write) return;
    const Patch3 pt = a.cpatches[sv.grid];
    const Faces3 F = faces3(pt, a.nv);
    const T* face = static_cast<const T*>(a.cface);
    const int mx = pt.mx, my = pt.my;
    const int i = sv.I - 2, j = sv.J - 2, k = sv.K - 2;   // interior cell, 1-based
    auto xf = [&](int fi, int fj, int fk) { return m * F.sx + ((int64_t)(fk - 1) * my + (fj - 1)) * (mx + 1) + (fi - 1); };
    auto yf = [&](int fi, int fj, int fk) { return m * F.sy + ((int64_t)(fk - 1) * (my + 1) + (fj - 1)) * mx + (fi - 1); };
    auto zf = [&](int fi, int fj, int fk) { return m * F.sz + ((int64_t)(fk - 1) * my + (fj - 1)) * mx + (fi - 1); };
    const bool merged = a.merged != 0;
    double v;
    switch (sv.side) {
    case 1: v = -face_p(face, F.fm + xf(i, j, k), F.fp + xf(i, j, k), merged) * a.cdt * a.chy * a.chz; break;
    case 2: v = -(double)face[F.gm + yf(i, j + 1, k)] * a.cdt * a.chx * a.chz; break;
    case 3: v = -(double)face[F.fm + xf(i + 1, j, k)] * a.cdt * a.chy * a.chz; break;
    case 4: v = -face_p(face, F.gm + yf(i, j, k), F.gp + yf(i, j, k), merged) * a.cdt * a.chx * a.chz; break;
    case 5: v = -(double)face[F.hp + zf(i, j, k)] * a.cdt * a.chx * a.chy; break;
    default: v = -(double)face[F.hm + zf(i, j, k + 1)] * a.cdt * a.chx * a.chy; break;
    }
    const RegGrid3 rg = a.fregs[sv.kid];
    a.freg[rg.reg_off + (int64_t)a.nv * (sv.lkid - 1) + m] = v;
}

// fluxad.f (3-D): one (slot, component) per thread, the fine fluxes of the
// slot's r x r faces summed in the reference's order
template <typename T>
__global__ __launch_bounds__(256) void fluxad3_kernel(Reg3Args a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nitems * a.nv) return;
    const int e = t / a.nv, m = t - e * a.nv;
    const Qad3Item it = a.qad[e];
    const Patch3 pt = a.fpatches[it.grid];
    const RegGrid3 rg = a.fregs[it.grid];
    const Faces3 F = faces3(pt, a.nv);
    const T* face = static_cast<const T*>(a.fface);
    const int mx = pt.mx, my = pt.my, mz = pt.mz;
    const bool merged = a.merged != 0;
    double* svd = a.freg + rg.reg_off + (int64_t)a.nv * (it.slot - 1) + m;
    double acc = *svd;
    const int side = it.side;
    const bool xline = side == 1 || side == 3, zside = side >= 5;
    const int r1 = xline ? a.raty : a.ratx;               // inner (fastest) direction
    const int r2 = zside ? a.raty : a.ratz;               // outer
    const double f1 = xline ? a.hy : a.hx, f2 = xline ? a.hz : (zside ? a.hy : a.hz);
    for (int o = 1; o <= r2; o++) {
        const int oc = (it.c2 - 1) * r2 + o;
        for (int l = 1; l <= r1; l++) {
            const int lc = (it.c1 - 1) * r1 + l;
            double v;
            switch (side) {
            case 1: v = (double)face[F.fm + m * F.sx + ((int64_t)(oc - 1) * my + (lc - 1)) * (mx + 1)]; break;
            case 2: {
                const int64_t x = m * F.sy + ((int64_t)(oc - 1) * (my + 1) + my) * mx + (lc - 1);
                v = face_p(face, F.gm + x, F.gp + x, merged);
                break;
            }
            case 3: {
                const int64_t x = m * F.sx + ((int64_t)(oc - 1) * my + (lc - 1)) * (mx + 1) + mx;
                v = face_p(face, F.fm + x, F.fp + x, merged);
                break;
            }
            case 4: v = (double)face[F.gm + m * F.sy + ((int64_t)(oc - 1) * (my + 1)) * mx + (lc - 1)]; break;
            case 5: v = (double)face[F.hm + m * F.sz + ((int64_t)(oc - 1)) * mx + (lc - 1)]; break;
            default: v = (double)face[F.hp + m * F.sz + ((int64_t)mz * my + (oc - 1)) * mx + (lc - 1)]; break;
            }
            acc = acc + v * a.delt * f1 * f2;
        }
    }
    *svd = acc;
}

// upbnd.f (3-D): one coarse cell per thread, its register slots in order
template <typename S>
__global__ __launch_bounds__(256) void upbnd3_kernel(Reg3Args a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ncells) return;
    const Upbnd3Cell c = a.ucells[t];
    const Patch3 pt = a.cpatches[c.grid];
    S* q = static_cast<S*>(a.cq) + pt.qoff;
    const int64_t mit = pt.mx + 4, mij = mit * (pt.my + 4), pl = qplane3(pt);
    const int64_t cell = c.k * mij + c.j * mit + c.i;
    for (int m = 0; m < a.nv; m++) {
        double v = (double)q[m * pl + cell];
        for (int e = c.e0; e < c.e1; e++) {
            const Upbnd3Entry u = a.uents[e];
            const RegGrid3 rg = a.fregs[u.kid];
            const double s = a.freg[rg.reg_off + (int64_t)a.nv * (u.lkid - 1) + m];
            v = v + u.sgnm * s / a.volume;
            if (sizeof(S) == 4) v = (double)(S)v;   // each update stored in the grid's precision
        }
        q[m * pl + cell] = (S)v;
    }
}

// update.f (3-D):129-160: coarse cell = mean of its r^3 fine cells
template <typename S>
__global__ __launch_bounds__(256) void avg3_kernel(Reg3Args a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.navg_cells) return;
    int lo = 0, hi = a.navg - 1;
    while (lo < hi) {   // block holding cell t
        const int mid = (lo + hi + 1) / 2;
        if (a.avg[mid].cell0 <= t) lo = mid; else hi = mid - 1;
    }
    const Avg3Block b = a.avg[lo];
    const int64_t c = t - b.cell0;
    const int ci = (int)(c % b.ni), r = (int)(c / b.ni);
    const int cj = r % b.nj, ck = r / b.nj;
    const Patch3 cp = a.cpatches[b.grid];
    const Patch3 fp = a.fpatches[b.kid];
    S* q = static_cast<S*>(a.cq) + cp.qoff;
    const S* qf = static_cast<const S*>(a.fq) + fp.qoff;
    const int64_t cmit = cp.mx + 4, cmij = cmit * (cp.my + 4), cpl = qplane3(cp);
    const int64_t fmit = fp.mx + 4, fmij = fmit * (fp.my + 4), fpl = qplane3(fp);
    const int64_t ccell = (b.ck0 + ck) * cmij + (b.cj0 + cj) * cmit + (b.ci0 + ci);
    const int fi = b.fi0 + ci * a.cratx, fj = b.fj0 + cj * a.craty, fk = b.fk0 + ck * a.cratz;
    const double totrat = (double)(a.cratx * a.craty * a.cratz);
    for (int m = 0; m < a.nv; m++) {
        double v = 0.0;
        for (int kco = 0; kco < a.cratz; kco++)
            for (int jco = 0; jco < a.craty; jco++)
                for (int ico = 0; ico < a.cratx; ico++)
                    v = v + (double)qf[m * fpl + (fk + kco) * fmij + (fj + jco) * fmit + (fi + ico)];
        q[m * cpl + ccell] = (S)(v / totrat);
    }
}

// errf1.f (3-D): one coarsened interior cell per thread; the Richardson
// estimate |mean of the 8 fine values - coarse value| / (2^(iorder+1) - 2)
// of component 1 flags (badpt = 2) its 2x2x2 fine cells in rflags (zeroed
// before the launch); fine and coarse solutions in the FP64 arena
__global__ __launch_bounds__(256) void errf1_3_kernel(const double* arena, const Patch3* fine,
                                                      const Patch3* crs, const int64_t* foff,
                                                      const int64_t* cstart, int ngrids,
                                                      double tol, double order, double* rflags,
                                                      unsigned long long* nflag) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cstart[ngrids]) return;
    int lo = 0, hi = ngrids - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (cstart[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const int g = lo;
    const Patch3 pf = fine[g], pc = crs[g];
    const int64_t c = t - cstart[g];
    const int i = (int)(c % pc.mx), r = (int)(c / pc.mx);
    const int j = r % pc.my, k = r / pc.my;          // 0-based interior coarse cell
    const int64_t fmit = pf.mx + 4, fmij = fmit * (pf.my + 4);
    const int64_t cmit = pc.mx + 4, cmij = cmit * (pc.my + 4);
    const double* qf = arena + pf.qoff;              // component 1 = plane 0
    const double* qc = arena + pc.qoff;
    const int fi = 2 + 2 * i, fj = 2 + 2 * j, fk = 2 + 2 * k;   // 0-based with ghosts
    auto F = [&](int a, int b, int d) { return qf[(int64_t)d * fmij + (int64_t)b * fmit + a]; };
    const double t1 = F(fi, fj, fk), t2 = F(fi + 1, fj, fk), t3 = F(fi + 1, fj + 1, fk);
    const double t4 = F(fi, fj + 1, fk), t5 = F(fi, fj, fk + 1), t6 = F(fi + 1, fj, fk + 1);
    const double t7 = F(fi + 1, fj + 1, fk + 1), t8 = F(fi, fj + 1, fk + 1);
    const double aval = (t1 + t2 + t3 + t4 + t5 + t6 + t7 + t8) / 8.0;
    const double est = fabs((aval - qc[(int64_t)(k + 2) * cmij + (int64_t)(j + 2) * cmit + (i + 2)]) / order);
    if (est >= tol) {
        double* rf = rflags + foff[g];
        for (int dk = 0; dk < 2; dk++)
            for (int dj = 0; dj < 2; dj++)
                for (int di = 0; di < 2; di++)
                    rf[(int64_t)(fk + dk) * fmij + (int64_t)(fj + dj) * fmit + (fi + di)] = 2.0;
        atomicAdd(nflag, 8ull);
    }
}

// ---------------------------------------------------------------- launchers
hipError_t launch_errf1_3(const double* arena, const Patch3* fine, const Patch3* crs,
                          const int64_t* foff, const int64_t* cstart, int ngrids, int64_t ncells,
                          double tol, double order, double* rflags, unsigned long long* nflag,
                          hipStream_t st) {
    if (ncells <= 0) return hipSuccess;
    hipLaunchKernelGGL(errf1_3_kernel, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, st,
                       arena, fine, crs, foff, cstart, ngrids, tol, order, rflags, nflag);
    return hipGetLastError();
}

hipError_t launch_fill3_stage(int esize, const Fill3Args& a, hipStream_t st) {
    if (a.nwork <= 0) return hipSuccess;
    if (esize == 8)
        hipLaunchKernelGGL(fill3_stage_kernel<double>, dim3(a.nwork), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(fill3_stage_kernel<float>, dim3(a.nwork), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_qad3(int solver, const Reg3Args& a, const Physics3& ph, hipStream_t st) {
    if (a.nitems <= 0) return hipSuccess;
    const dim3 grid((a.nitems + 127) / 128), block(128);
    if (solver == 1) {
        if (a.esize == 8)
            hipLaunchKernelGGL((qad3_kernel<Advection3<double>, double>), grid, block, 0, st, a, ph);
        else
            hipLaunchKernelGGL((qad3_kernel<Advection3<double>, float>), grid, block, 0, st, a, ph);
    } else {
        if (a.esize == 8)
            hipLaunchKernelGGL((qad3_kernel<Euler3<double>, double>), grid, block, 0, st, a, ph);
        else
            hipLaunchKernelGGL((qad3_kernel<Euler3<double>, float>), grid, block, 0, st, a, ph);
    }
    return hipGetLastError();
}

hipError_t launch_fluxsv3(const Reg3Args& a, hipStream_t st) {
    const int n = a.nsv * a.nv;
    if (n <= 0) return hipSuccess;
    const dim3 grid((n + 255) / 256), block(256);
    if (a.esize == 8) hipLaunchKernelGGL(fluxsv3_kernel<double>, grid, block, 0, st, a);
    else hipLaunchKernelGGL(fluxsv3_kernel<float>, grid, block, 0, st, a);
    return hipGetLastError();
}

hipError_t launch_fluxad3(const Reg3Args& a, hipStream_t st) {
    const int n = a.nitems * a.nv;
    if (n <= 0) return hipSuccess;
    const dim3 grid((n + 255) / 256), block(256);
    if (a.esize == 8) hipLaunchKernelGGL(fluxad3_kernel<double>, grid, block, 0, st, a);
    else hipLaunchKernelGGL(fluxad3_kernel<float>, grid, block, 0, st, a);
    return hipGetLastError();
}

hipError_t launch_update3_level(const Reg3Args& a, hipStream_t st) {
    if (a.ncells > 0) {
        const dim3 grid((a.ncells + 255) / 256), block(256);
        if (a.esize == 8) hipLaunchKernelGGL(upbnd3_kernel<double>, grid, block, 0, st, a);
        else hipLaunchKernelGGL(upbnd3_kernel<float>, grid, block, 0, st, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (a.navg_cells > 0) {
        const dim3 grid((unsigned)((a.navg_cells + 255) / 256)), block(256);
        if (a.esize == 8) hipLaunchKernelGGL(avg3_kernel<double>, grid, block, 0, st, a);
        else hipLaunchKernelGGL(avg3_kernel<float>, grid, block, 0, st, a);
    }
    return hipGetLastError();
}

}  // namespace amrgpu
