// fill.hip -- executes one stage of a ghost-cell fill program.
// Built with -ffp-contract=off: the time blend and the coarse-fine
// interpolation are evaluated exactly as the reference's FP64 code.
#include <hip/hip_runtime.h>
#include "fill.hpp"
#include "sweep.hpp"

namespace amrgpu {

template <typename S>
struct BufRef {
    S* lvl;          // kind 0
    double* scr;     // kind 1
    int64_t off;
    int32_t nrow, plane;
    __device__ double load(int m, int i, int j) const {
        int64_t k = off + (int64_t)m * plane + (int64_t)j * nrow + i;
        return lvl ? (double)lvl[k] : scr[k];
    }
    __device__ void store(int m, int i, int j, double v) const {
        int64_t k = off + (int64_t)m * plane + (int64_t)j * nrow + i;
        if (lvl) lvl[k] = (S)v; else scr[k] = v;
    }
};

template <typename S>
__device__ BufRef<S> bufref(const FillArgs& a, int b) {
    const FillBuf fb = a.bufs[b];
    BufRef<S> r;
    if (fb.kind == 0) {
        r.lvl = static_cast<S*>(a.lvl_new[fb.level]);
        r.scr = nullptr;
        r.off = a.lvl_patches[fb.level][fb.patch].off;
    } else {
        r.lvl = nullptr;
        r.scr = fb.kind == 1 ? a.scratch : a.arena;
        r.off = fb.off;
    }
    r.nrow = fb.nrow;
    r.plane = fb.plane;
    return r;
}

// the aux-mode copy of a saveqc program: aux values of the coarse grids
// into the coarse patch (icall.f's aux part) and the patch's ring into
// auxc1d (cstore.f); NEEDS_TO_BE_SET stays where no coarse grid reached
__device__ void fill_aux_op(const FillArgs& a, const FillOp& op, int tid) {
    const int na = a.naux;
    if (op.type == FOP_COPY) {
        const PatchDesc sp = a.lvl_patches[op.src_level][op.src];
        const double* sa = a.lvl_aux[op.src_level] + sp.off * a.auxk;
        const FillBuf fb = a.bufs[op.dst];
        double* da = a.scratch + fb.off * a.auxk;
        const int ncell = op.ni * op.nj;
        for (int t = tid; t < ncell * na; t += 256) {
            const int c = t / ncell, r = t - c * ncell;
            const int dj = r / op.ni, di = r - dj * op.ni;
            da[(int64_t)c * fb.plane + (int64_t)(op.dj0 + dj) * fb.nrow + (op.di0 + di)] =
                sa[(int64_t)c * sp.plane + (int64_t)(op.sj0 + dj) * sp.mitot + (op.si0 + di)];
        }
    } else if (op.type == FOP_RING) {
        const FillBuf fb = a.bufs[op.src];
        const double* sa = a.scratch + fb.off * a.auxk;
        const PatchDesc pf = a.lvl_patches[op.level][op.si0];
        double* ac1d = a.lvl_reg[op.level] + aux_ring_offset(pf, a.nv);
        const int nrow = op.ni, ncol = op.nj;
        const int n1 = ncol - 2, n2 = nrow - 2;
        const int len = 2 * (n1 + n2);
        for (int t = tid; t < len * na; t += 256) {
            const int idx = t / na, c = t - idx * na;
            int i, j;
            if (idx < n1) { i = 0; j = 1 + idx; }
            else if (idx < n1 + n2) { i = 1 + idx - n1; j = ncol - 1; }
            else if (idx < 2 * n1 + n2) { i = nrow - 1; j = 1 + idx - n1 - n2; }
            else { i = 1 + idx - 2 * n1 - n2; j = 0; }
            ac1d[(int64_t)idx * na + c] = sa[(int64_t)c * fb.plane + (int64_t)j * nrow + i];
        }
    }
}

// t = tid + 256 k written as digits (i, j, m) of the radices (ni, nj):
// stepped by 256 with one carry per digit instead of two integer divisions
// per element (a digit plus its step plus a carry stays below twice its
// radix, so one subtraction brings it back)
struct Digits2 {
    int i, j, m, si, sj, sm, ni, nj;
    __device__ __forceinline__ Digits2(int t, int ni_, int nj_)
        : ni(ni_ > 0 ? ni_ : 1), nj(nj_ > 0 ? nj_ : 1) {   // (an empty op runs no element)
        i = t % ni; t /= ni; j = t % nj; m = t / nj;
        int r = 256;
        si = r % ni; r /= ni; sj = r % nj; sm = r / nj;
    }
    __device__ __forceinline__ void step() {
        i += si;
        int c = i >= ni;
        i -= c ? ni : 0;
        j += sj + c;
        c = j >= nj;
        j -= c ? nj : 0;
        m += sm + c;
    }
};

template <typename S>
__global__ __launch_bounds__(256) void fill_stage_kernel(FillArgs a) {
    const FillOp op = a.ops[blockIdx.x];
    const int tid = threadIdx.x, nv = a.nv;
    if (a.clear && blockIdx.x == 0)
        for (int64_t t = tid; t < a.nclear; t += 256) a.clear[t] = 0ull;
    if (a.aux_mode) {
        fill_aux_op(a, op, tid);
        return;
    }
    const BufRef<S> dst = bufref<S>(a, op.dst);
    const int ncell = op.ni * op.nj;

    if (op.type == FOP_COPY) {
        const PatchDesc sp = a.lvl_patches[op.src_level][op.src];
        const S* qn = static_cast<const S*>(a.lvl_new[op.src_level]) + sp.off;
        const S* qo = static_cast<const S*>(a.lvl_old[op.src_level]) + sp.off;
        int mode = a.tmode[op.src_level];
        if (op.flags & FOPF_FORCE_NEW) mode = TIME_NEW;
        else if (op.flags & FOPF_FORCE_OLD) mode = TIME_OLD;
        else if (op.flags & FOPF_BIG) mode = a.big_old ? TIME_OLD : TIME_NEW;
        const double ai = a.alphai[op.src_level], ac = a.alphac[op.src_level];
        Digits2 u(tid, op.ni, op.nj);
        for (int t = tid; t < ncell * nv; t += 256, u.step()) {
            const int m = u.m, dj = u.j, di = u.i;
            const int64_t s = (int64_t)m * sp.plane + (int64_t)(op.sj0 + dj) * sp.mitot + (op.si0 + di);
            double v;
            if (mode == TIME_NEW) v = (double)qn[s];
            else if (mode == TIME_OLD) v = (double)qo[s];
            else v = (double)qn[s] * ai + (double)qo[s] * ac;
            dst.store(m, op.di0 + di, op.dj0 + dj, v);
        }
    } else if (op.type == FOP_INTERP) {
        const BufRef<S> crs = bufref<S>(a, op.src);
        const double dxf = a.hx[op.level], dyf = a.hy[op.level];
        const double dxc = a.hx[op.level - 1], dyc = a.hy[op.level - 1];
        const double xlow_coarse = a.xlower + op.iplo * dxc;
        const double ylow_coarse = a.ylower + op.jplo * dyc;
        const double ratiox = (double)op.ratx, ratioy = (double)op.raty;
        const unsigned char* mask = a.masks + op.mask_off;
        Digits2 u(tid, op.ni, 1 << 30);   // (i, j) of cell c; j does not wrap
        for (int c = tid; c < ncell; c += 256, u.step()) {
            if (mask[c]) continue;
            const int jf = u.j + 1, if_ = u.i + 1;   // 1-based
            const int j_coarse = (int)floor((double)(jf + op.jlo - 1) / ratioy) - op.jplo + 1;
            const double ycent_coarse = ylow_coarse + ((double)j_coarse - .5) * dyc;
            const double ycent_fine = a.ylower + ((double)(jf - 1 + op.jlo) + .5) * dyf;
            const double eta2 = (ycent_fine - ycent_coarse) / dyc;
            const int i_coarse = (int)floor((double)(if_ + op.ilo - 1) / ratiox) - op.iplo + 1;
            const double xcent_coarse = xlow_coarse + ((double)i_coarse - .5) * dxc;
            const double xcent_fine = a.xlower + ((double)(if_ - 1 + op.ilo) + .5) * dxf;
            const double eta1 = (xcent_fine - xcent_coarse) / dxc;
            const int ic = i_coarse - 1, jc = j_coarse - 1;                  // 0-based
            for (int n = 0; n < nv; n++) {
                const double valp10 = crs.load(n, ic + 1, jc);
                const double valm10 = crs.load(n, ic - 1, jc);
                const double valc = crs.load(n, ic, jc);
                const double valp01 = crs.load(n, ic, jc + 1);
                const double valm01 = crs.load(n, ic, jc - 1);
                const double dupc = valp10 - valc, dumc = valc - valm10, ducc = valp10 - valm10;
                double du = fmin(fabs(dupc), fabs(dumc));
                du = fmin(2.0 * du, 0.5 * fabs(ducc));
                const double fu = fmax(0.0, copysign(1.0, dupc * dumc));
                const double uslope = du * copysign(1.0, ducc) * fu;
                const double dvpc = valp01 - valc, dvmc = valc - valm01, dvcc = valp01 - valm01;
                double dv = fmin(fabs(dvpc), fabs(dvmc));
                dv = fmin(2.0 * dv, 0.5 * fabs(dvcc));
                const double fv = fmax(0.0, copysign(1.0, dvpc * dvmc));
                const double vslope = dv * copysign(1.0, dvcc) * fv;
                const double valint = valc + eta1 * uslope + eta2 * vslope;
                dst.store(n, op.di0 + if_ - 1, op.dj0 + jf - 1, valint);
            }
        }
    } else if (op.type == FOP_FILVAL) {
        // src: the coarse patch grown by one coarse cell (ni x nj); dst: the
        // new grid; mask: its interior cells, row by row, 1 where an old
        // grid supplied the value (the reference's setflags test)
        const BufRef<S> crs = bufref<S>(a, op.src);
        const int rx = op.ratx, ry = op.raty, mic = op.ni, mjc = op.nj;
        const int ncx = mic - 2, ncc = ncx * (mjc - 2), nx = ncx * rx;
        const unsigned char* mask = a.masks + op.mask_off;
        for (int t = tid; t < ncc * nv; t += 256) {
            const int m = t / ncc, c = t - m * ncc;
            const int jc = c / ncx + 1, ic = c - (jc - 1) * ncx + 1;   // 0-based, 1..mic-2
            const double vc = crs.load(m, ic, jc);
            const double vxp = crs.load(m, ic + 1, jc), vxm = crs.load(m, ic - 1, jc);
            const double vyp = crs.load(m, ic, jc + 1), vym = crs.load(m, ic, jc - 1);
            double s1p = vxp - vc, s1m = vc - vxm;
            double slopex = fmin(fabs(s1p), fabs(s1m)) * copysign(1.0, vxp - vxm);
            if (s1m * s1p <= 0.0) slopex = 0.0;
            s1p = vyp - vc;
            s1m = vc - vym;
            double slopey = fmin(fabs(s1p), fabs(s1m)) * copysign(1.0, vyp - vym);
            if (s1m * s1p <= 0.0) slopey = 0.0;
            for (int jco = 1; jco <= ry; jco++) {
                const double yoff = ((double)jco - 0.5) / (double)ry - 0.5;
                const int jf = (jc - 1) * ry + jco - 1;                    // interior, 0-based
                for (int ico = 1; ico <= rx; ico++) {
                    const double xoff = ((double)ico - 0.5) / (double)rx - 0.5;
                    const int if_ = (ic - 1) * rx + ico - 1;
                    if (!mask[(int64_t)jf * nx + if_])
                        dst.store(m, NGHOST + if_, NGHOST + jf, vc + xoff * slopex + yoff * slopey);
                }
            }
        }
    } else if (op.type == FOP_COPYBACK) {
        const BufRef<S> src = bufref<S>(a, op.src);
        for (int t = tid; t < ncell * nv; t += 256) {
            const int m = t / ncell, c = t - m * ncell;
            const int dj = c / op.ni, di = c - dj * op.ni;
            dst.store(m, op.di0 + di, op.dj0 + dj, src.load(m, di, dj));
        }
    } else if (op.type == FOP_RING) {
        // cstore.f: sides 1 (left column), 2 (top row), 3 (right column),
        // 4 (bottom row) of the coarsened patch, corners excluded
        const BufRef<S> src = bufref<S>(a, op.src);
        const PatchDesc pf = a.lvl_patches[op.level][op.si0];
        double* qc1d = a.lvl_reg[op.level] + pf.reg_off + (int64_t)nv * pf.lenbc;
        const int nrow = op.ni, ncol = op.nj;
        const int n1 = ncol - 2, n2 = nrow - 2;
        const int len = 2 * (n1 + n2);
        for (int t = tid; t < len * nv; t += 256) {
            const int idx = t / nv, m = t - idx * nv;
            int i, j;
            if (idx < n1) { i = 0; j = 1 + idx; }
            else if (idx < n1 + n2) { i = 1 + idx - n1; j = ncol - 1; }
            else if (idx < 2 * n1 + n2) { i = nrow - 1; j = 1 + idx - n1 - n2; }
            else { i = 1 + idx - 2 * n1 - n2; j = 0; }
            qc1d[(int64_t)idx * nv + m] = src.load(m, i, j);
        }
    } else if (op.type == FOP_COARSEN) {
        // coarsen.f:15-30: dst(i, j) = mean of src(2i..2i+1, 2j..2j+1)
        const BufRef<S> src = bufref<S>(a, op.src);
        for (int t = tid; t < ncell * nv; t += 256) {
            const int m = t / ncell, c = t - m * ncell;
            const int j = c / op.ni, i = c - j * op.ni;
            const double v = (src.load(m, 2 * i, 2 * j) + src.load(m, 2 * i + 1, 2 * j) +
                              src.load(m, 2 * i, 2 * j + 1) + src.load(m, 2 * i + 1, 2 * j + 1)) / 4.0;
            dst.store(m, i, j, v);
        }
    } else if (op.type == FOP_BC) {
        // sides in the reference's order, each complete before the next
        const int nrow = op.ni, ncol = op.nj;
        // left: columns 0..nxl-1 from column nxl (extrapolation) or mirrored (wall)
        const int nxl = op.bc[0];
        if (nxl > 0 && (a.mthbc[0] == 1 || a.mthbc[0] == 3)) {
            for (int t = tid; t < nxl * ncol; t += 256) {
                const int j = t / nxl, i = t - j * nxl;
                const int isrc = a.mthbc[0] == 1 ? nxl : 2 * nxl - 1 - i;
                for (int m = 0; m < nv; m++) {
                    double v = dst.load(m, isrc, j);
                    if (a.mthbc[0] == 3 && m == 1) v = -v;
                    dst.store(m, i, j, v);
                }
            }
        }
        __syncthreads();
        const int ibeg = op.bc[1];              // 1-based first column set
        if (ibeg > 0 && (a.mthbc[1] == 1 || a.mthbc[1] == 3)) {
            const int nr = nrow - ibeg + 1;
            for (int t = tid; t < nr * ncol; t += 256) {
                const int j = t / nr, i = ibeg - 1 + (t - j * nr);
                const int isrc = a.mthbc[1] == 1 ? ibeg - 2 : 2 * ibeg - 1 - (i + 1) - 1;
                for (int m = 0; m < nv; m++) {
                    double v = dst.load(m, isrc, j);
                    if (a.mthbc[1] == 3 && m == 1) v = -v;
                    dst.store(m, i, j, v);
                }
            }
        }
        __syncthreads();
        const int nyb = op.bc[2];
        if (nyb > 0 && (a.mthbc[2] == 1 || a.mthbc[2] == 3)) {
            for (int t = tid; t < nyb * nrow; t += 256) {
                const int j = t / nrow, i = t - j * nrow;
                const int jsrc = a.mthbc[2] == 1 ? nyb : 2 * nyb - 1 - j;
                for (int m = 0; m < nv; m++) {
                    double v = dst.load(m, i, jsrc);
                    if (a.mthbc[2] == 3 && m == 2) v = -v;
                    dst.store(m, i, j, v);
                }
            }
        }
        __syncthreads();
        const int jbeg = op.bc[3];
        if (jbeg > 0 && (a.mthbc[3] == 1 || a.mthbc[3] == 3)) {
            const int nc = ncol - jbeg + 1;
            for (int t = tid; t < nc * nrow; t += 256) {
                const int j = jbeg - 1 + t / nrow, i = t - (t / nrow) * nrow;
                const int jsrc = a.mthbc[3] == 1 ? jbeg - 2 : 2 * jbeg - 1 - (j + 1) - 1;
                for (int m = 0; m < nv; m++) {
                    double v = dst.load(m, i, jsrc);
                    if (a.mthbc[3] == 3 && m == 2) v = -v;
                    dst.store(m, i, j, v);
                }
            }
        }
    }
}

hipError_t launch_fill_stage(int esize, const FillArgs& a, hipStream_t st) {
    if (a.nops <= 0) return hipSuccess;
    if (esize == 8)
        hipLaunchKernelGGL(fill_stage_kernel<double>, dim3(a.nops), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(fill_stage_kernel<float>, dim3(a.nops), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace amrgpu
