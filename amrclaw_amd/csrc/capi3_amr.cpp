// capi3_amr.cpp -- the 3-D level update around the step on a hierarchy held
// on the device: the C ABI of include/amrgpu3.h's hierarchy section.
//
// Per level: the grids' new and old solutions (compute precision, structure
// of arrays per grid), the sweep work lists, the ghost-fill program of
// bound.f, the saveqc program of the level as a fine level, and the
// coarse-fine registers (FP64).  One face arena, sized for the largest
// level, holds the six flux arrays of the level being advanced until its
// fluxsv/fluxad have run.  advance_level follows advanc.f (3-D): bound every
// grid, saveqc for the next level, new -> old copy, qad, the step, fluxsv,
// fluxad, time += possk.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/amrgpu.h"
#include "../../include/amrgpu3.h"
#include "amr3_plan.hpp"
#include "capi3_internal.hpp"
#include "kernels3.hpp"

namespace amrgpu {
hipError_t launch_fill3_stage(int esize, const Fill3Args& a, hipStream_t st);
hipError_t launch_qad3(int solver, const Reg3Args& a, const Physics3& ph, hipStream_t st);
hipError_t launch_fluxsv3(const Reg3Args& a, hipStream_t st);
hipError_t launch_fluxad3(const Reg3Args& a, hipStream_t st);
hipError_t launch_update3_level(const Reg3Args& a, hipStream_t st);
hipError_t launch_errf1_3(const double* arena, const Patch3* fine, const Patch3* crs,
                          const int64_t* foff, const int64_t* cstart, int ngrids, int64_t ncells,
                          double tol, double order, double* rflags, unsigned long long* nflag,
                          hipStream_t st);
}  // namespace amrgpu

using namespace amrgpu;

namespace {

const double RINF = (double)10.e32f;   // amr_module's rinfinity

struct DevFill {
    Fill3Plan plan;
    Fill3Work* d_work = nullptr;
    Fill3Op* d_ops = nullptr;
    Fill3Buf* d_bufs = nullptr;
    unsigned char* d_masks = nullptr;
};

struct Lev3 {
    std::vector<std::array<int, 6>> boxes;
    std::vector<Patch3> patches;
    Patch3* d_patches = nullptr;
    std::vector<Block3> blocks[4];
    Block3* d_blocks[4] = {};
    void* d_new = nullptr;
    void* d_old = nullptr;
    int64_t q_elems = 0, face_elems = 0;
    double time = 0.0, possk = 0.0;
    unsigned long long* d_cfl = nullptr;
    std::vector<unsigned long long> h_cfl;
    double* d_partial = nullptr;
    DevFill fill, saveqc;
    Reg3Level reg;
    RegGrid3* d_regs = nullptr;
    double* d_reg = nullptr;
    // layout generation (a new value at every amrgpu3_set_level of this
    // level) and the generations d_reg was laid out for (this level's and
    // the next coarser level's), with its size
    uint64_t gen = 0, reg_gen = 0, reg_gen_coarse = 0;
    int64_t reg_elems = 0;
    Qad3Item* d_qad = nullptr;
    Sv3Entry* d_sv = nullptr;
    Upbnd3Cell* d_ucells = nullptr;
    Upbnd3Entry* d_uents = nullptr;
    Avg3Block* d_avg = nullptr;
    // flagging (flag2refine, errf1) of the level's grids
    double* d_flags = nullptr;
    double* d_rflags = nullptr;
    int64_t* d_foff = nullptr;
    std::vector<int64_t> foff;
};

struct Amr3 {
    bool geom = false;
    Hier3Geom H;
    Lev3 lev[AMR3_MAXLV];
    void* d_face = nullptr;
    int64_t face_cap = 0;
    double* d_scratch = nullptr;
    int64_t scratch_cap = 0;
    bool planned = false;        // ghost-fill programs
    bool planned_regs = false;   // saveqc programs and registers
    uint64_t gen_counter = 0;    // source of Lev3::gen values
    int plane_block = 0;         // sweep plane block of the levels (0: per level)
};

Amr3 A;

#define HIPCHKA(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ctx3_fail(AMRGPU_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

template <typename P>
void dfree(P*& p) {
    if (p) hipFree((void*)p);
    p = nullptr;
}

void free_fill(DevFill& f) {
    dfree(f.d_work);
    dfree(f.d_ops);
    dfree(f.d_bufs);
    dfree(f.d_masks);
    f.plan = Fill3Plan();
}

void free_plans(Lev3& L) {
    free_fill(L.fill);
    free_fill(L.saveqc);
    dfree(L.d_regs);
    dfree(L.d_reg);
    dfree(L.d_qad);
    dfree(L.d_sv);
    dfree(L.d_ucells);
    dfree(L.d_uents);
    dfree(L.d_avg);
    L.reg = Reg3Level();
}

void free_level(Lev3& L) {
    free_plans(L);
    dfree(L.d_patches);
    for (auto& b : L.d_blocks) dfree(b);
    dfree(L.d_new);
    dfree(L.d_old);
    dfree(L.d_cfl);
    dfree(L.d_partial);
    dfree(L.d_flags);
    dfree(L.d_rflags);
    dfree(L.d_foff);
    L = Lev3();
}

int need_geom() {
    if (!ctx3_inited()) return ctx3_fail(AMRGPU_ERR_STATE, "amrgpu3_init has not been called");
    if (!A.geom) return ctx3_fail(AMRGPU_ERR_STATE, "amrgpu3_set_geometry has not been called");
    return AMRGPU_OK;
}

int check_level(int level) {
    int rc = need_geom();
    if (rc) return rc;
    if (level < 1 || level > A.H.nlev) return ctx3_fail(AMRGPU_ERR_ARG, "level %d out of range", level);
    return AMRGPU_OK;
}

template <typename T>
int upload(T*& d, const std::vector<T>& h) {
    dfree(d);
    HIPCHKA(hipMalloc((void**)&d, std::max<size_t>(1, h.size()) * sizeof(T)));
    if (!h.empty()) HIPCHKA(ctx3_h2d(d, h.data(), h.size() * sizeof(T)));
    return AMRGPU_OK;
}

int upload_fill(DevFill& f) {
    int rc = upload(f.d_ops, f.plan.ops);
    if (!rc) rc = upload(f.d_work, f.plan.work);
    if (!rc) rc = upload(f.d_bufs, f.plan.bufs);
    if (!rc) rc = upload(f.d_masks, f.plan.masks);
    if (!rc && f.plan.scratch_elems > A.scratch_cap) {
        dfree(A.d_scratch);
        HIPCHKA(hipMalloc((void**)&A.d_scratch, f.plan.scratch_elems * sizeof(double)));
        A.scratch_cap = f.plan.scratch_elems;
    }
    return rc;
}

// ghost-fill programs of every level (after a layout change)
int ensure_fill_plans() {
    if (A.planned) return AMRGPU_OK;
    const int nv = ctx3_cfg().nvar, nlev = A.H.nlev;
    std::string err;
    for (int l = 1; l <= nlev; l++) {
        Lev3& L = A.lev[l];
        free_fill(L.fill);
        if (!plan_fill3_level(A.H, l, nv, L.fill.plan, err))
            return ctx3_fail(AMRGPU_ERR_ARG, "level %d ghost fill: %s", l, err.c_str());
        int rc = upload_fill(L.fill);
        if (rc) return rc;
    }
    A.planned = true;
    return AMRGPU_OK;
}

// saveqc programs and registers of every level: these need a properly
// nested hierarchy (the ghost fill alone does not)
int ensure_plans() {
    int rc = ensure_fill_plans();
    if (rc || A.planned_regs) return rc;
    const int nv = ctx3_cfg().nvar, nlev = A.H.nlev;
    std::string err;
    for (int l = 2; l <= nlev; l++) {
        Lev3& L = A.lev[l];
        free_fill(L.saveqc);
        if (!plan_saveqc3(A.H, l, nv, L.saveqc.plan, err))
            return ctx3_fail(AMRGPU_ERR_ARG, "level %d saveqc: %s", l, err.c_str());
        rc = upload_fill(L.saveqc);
        if (rc) return rc;
    }
    Reg3Level R[AMR3_MAXLV];
    if (!plan_registers3(A.H, nv, R, err)) return ctx3_fail(AMRGPU_ERR_ARG, "registers: %s", err.c_str());
    for (int l = 1; l <= nlev; l++) {
        Lev3& L = A.lev[l];
        L.reg = R[l];
        rc = upload(L.d_regs, L.reg.regs);
        if (!rc) rc = upload(L.d_qad, L.reg.qad);
        if (!rc) rc = upload(L.d_sv, L.reg.sv);
        if (!rc) rc = upload(L.d_ucells, L.reg.ucells);
        if (!rc) rc = upload(L.d_uents, L.reg.uents);
        if (!rc) rc = upload(L.d_avg, L.reg.avg);
        if (rc) return rc;
        // a regrid replaces the levels above lbase only (regrid.f): a fine
        // level whose grids and coarser grids are unchanged keeps what its
        // registers have accumulated in the current coarse step, as the
        // reference keeps node(ffluxptr) of such grids; the others start
        // from zero (prepf.f)
        const bool keep = l >= 2 && L.d_reg && L.reg_gen == L.gen &&
                          L.reg_gen_coarse == A.lev[l - 1].gen && L.reg_elems == L.reg.reg_elems;
        if (!keep) {
            dfree(L.d_reg);
            const int64_t n = std::max<int64_t>(1, L.reg.reg_elems);
            HIPCHKA(hipMalloc((void**)&L.d_reg, n * sizeof(double)));
            HIPCHKA(hipMemsetAsync(L.d_reg, 0, n * sizeof(double), ctx3_stream()));   // prepf.f
        }
        L.reg_gen = L.gen;
        L.reg_gen_coarse = l >= 2 ? A.lev[l - 1].gen : 0;
        L.reg_elems = L.reg.reg_elems;
    }
    HIPCHKA(hipStreamSynchronize(ctx3_stream()));
    A.planned_regs = true;
    return AMRGPU_OK;
}

// intfil.f (3-D):95-150 time selection for the source levels of a fill at
// time t: new, old or the linear blend, with the tolerance possk/10
int fill_args(int level, double t, const Fill3Plan& P, const DevFill& D, Fill3Args& a,
              double* arena = nullptr) {
    const amrgpu3_config& c = ctx3_cfg();
    a = Fill3Args{};
    a.arena = arena;
    a.ops = D.d_ops;
    a.bufs = D.d_bufs;
    a.scratch = A.d_scratch;
    a.masks = D.d_masks;
    a.nv = c.nvar;
    a.xlower = A.H.xlower;
    a.ylower = A.H.ylower;
    a.zlower = A.H.zlower;
    for (int s = 0; s < 6; s++) a.mthbc[s] = A.H.mthbc[s];
    for (int l = 1; l <= A.H.nlev; l++) {
        const Lev3& L = A.lev[l];
        a.lvl_new[l] = L.d_new;
        a.lvl_old[l] = L.d_old;
        a.lvl_patches[l] = L.d_patches;
        a.lvl_regs[l] = L.d_regs;
        a.lvl_reg[l] = L.d_reg;
        a.hx[l] = A.H.lev[l].hx;
        a.hy[l] = A.H.lev[l].hy;
        a.hz[l] = A.H.lev[l].hz;
        a.tmode[l] = T3_NEW;
        a.alphai[l] = 1.0;
        a.alphac[l] = 0.0;
        if (!(P.src_levels_mask & (1u << l))) continue;
        const double dt = L.possk, teps = dt / 10.;
        const double alphac = (L.time - t) / dt, alphai = 1.0 - alphac;
        if (alphai < -teps || alphai > 1.0 + teps)
            return ctx3_fail(AMRGPU_ERR_STATE, "time %g not available from level %d (at %g)", t, l, L.time);
        a.alphai[l] = alphai;
        a.alphac[l] = alphac;
        if (std::fabs(alphai - 1.0) < teps) {
            a.tmode[l] = T3_NEW;
        } else {
            if (l == A.H.nlev)
                return ctx3_fail(AMRGPU_ERR_STATE, "level %d keeps no old time (finest level)", l);
            a.tmode[l] = std::fabs(alphai) < teps ? T3_OLD : T3_BLEND;
        }
    }
    (void)level;
    return AMRGPU_OK;
}

int run_fill(int level, double t, const DevFill& D, double* arena = nullptr) {
    const Fill3Plan& P = D.plan;
    if (P.ops.empty()) return AMRGPU_OK;
    Fill3Args a;
    int rc = fill_args(level, t, P, D, a, arena);
    if (rc) return rc;
    for (int s = 0; s < P.nstages; s++) {
        a.work = D.d_work + P.stage_start[s];
        a.nwork = P.stage_start[s + 1] - P.stage_start[s];
        HIPCHKA(launch_fill3_stage((int)ctx3_esize(), a, ctx3_stream()));
    }
    return AMRGPU_OK;
}

Reg3Args reg_args(int fine_level, int coarse_level) {
    const amrgpu3_config& c = ctx3_cfg();
    Reg3Args a{};
    a.nv = c.nvar;
    a.esize = (int)ctx3_esize();
    a.merged = ctx3_esize() == 4;
    if (fine_level >= 2 && fine_level <= A.H.nlev) {
        const Lev3& F = A.lev[fine_level];
        const Level3Geom& C = A.H.lev[fine_level - 1];
        a.fpatches = F.d_patches;
        a.fregs = F.d_regs;
        a.freg = F.d_reg;
        a.fq = F.d_new;
        a.fface = A.d_face;
        a.ratx = C.ratx;
        a.raty = C.raty;
        a.ratz = C.ratz;
        a.nitems = (int)F.reg.qad.size();
        a.qad = F.d_qad;
        a.hx = A.H.lev[fine_level].hx;
        a.hy = A.H.lev[fine_level].hy;
        a.hz = A.H.lev[fine_level].hz;
        a.delt = F.possk;
    }
    if (coarse_level >= 1 && coarse_level < A.H.nlev) {
        const Lev3& L = A.lev[coarse_level];
        const Level3Geom& G = A.H.lev[coarse_level];
        a.cpatches = L.d_patches;
        a.cface = A.d_face;
        a.cq = L.d_new;
        a.sv = L.d_sv;
        a.nsv = (int)L.reg.sv.size();
        a.ncells = (int)L.reg.ucells.size();
        a.ucells = L.d_ucells;
        a.uents = L.d_uents;
        a.avg = L.d_avg;
        a.navg = (int)L.reg.avg.size();
        a.navg_cells = L.reg.avg_cells;
        a.cratx = G.ratx;
        a.craty = G.raty;
        a.cratz = G.ratz;
        a.chx = G.hx;
        a.chy = G.hy;
        a.chz = G.hz;
        a.cdt = L.possk;
        a.volume = G.hx * G.hy * G.hz;
    }
    return a;
}

Physics3 physics() {
    const amrgpu3_config& c = ctx3_cfg();
    Physics3 ph{};
    ph.ubar = c.ubar;
    ph.vbar = c.vbar;
    ph.wbar = c.wbar;
    ph.gamma = c.gamma;
    ph.efix = c.efix;
    return ph;
}

bool has_fine(int level) { return level < A.H.nlev && !A.lev[level + 1].boxes.empty(); }

}  // namespace

namespace amrgpu {
void amr3_release() {
    if (ctx3_stream()) hipStreamSynchronize(ctx3_stream());
    for (auto& L : A.lev) free_level(L);
    dfree(A.d_face);
    dfree(A.d_scratch);
    A = Amr3();
}
}  // namespace amrgpu

extern "C" {

int amrgpu3_set_geometry(int nlev, const int* iregsz, const int* jregsz, const int* kregsz,
                         const int* ratx, const int* raty, const int* ratz, const double* hx,
                         const double* hy, const double* hz, const double* domain,
                         const int* mthbc) {
    if (!ctx3_inited()) return ctx3_fail(AMRGPU_ERR_STATE, "amrgpu3_init has not been called");
    if (nlev < 1 || nlev >= AMR3_MAXLV)
        return ctx3_fail(AMRGPU_ERR_ARG, "nlev must be in 1..%d", AMR3_MAXLV - 1);
    if (!iregsz || !jregsz || !kregsz || !ratx || !raty || !ratz || !hx || !hy || !hz || !domain ||
        !mthbc)
        return ctx3_fail(AMRGPU_ERR_ARG, "null geometry array");
    for (int s = 0; s < 6; s++)
        if (mthbc[s] < 1 || mthbc[s] > 3)
            return ctx3_fail(AMRGPU_ERR_UNSUPPORTED, "mthbc(%d) = %d: only 1 (extrapolation), "
                             "2 (periodic) and 3 (wall) are supported", s + 1, mthbc[s]);
    const int nv = ctx3_cfg().nvar;
    for (int s = 0; s < 6; s++)
        if (mthbc[s] == 3 && nv < s / 2 + 2)
            return ctx3_fail(AMRGPU_ERR_UNSUPPORTED, "wall boundaries need a momentum component");
    {
        // the same geometry again (a driver hands it over after every
        // regrid): keep the levels, their plans and their registers
        const Hier3Geom& O = A.H;
        bool same = A.geom && O.nlev == nlev;
        for (int l = 1; same && l <= nlev; l++) {
            const Level3Geom& G = O.lev[l];
            same = G.iregsz == iregsz[l - 1] && G.jregsz == jregsz[l - 1] &&
                   G.kregsz == kregsz[l - 1] && G.ratx == ratx[l - 1] && G.raty == raty[l - 1] &&
                   G.ratz == ratz[l - 1] && G.hx == hx[l - 1] && G.hy == hy[l - 1] &&
                   G.hz == hz[l - 1];
        }
        same = same && O.xlower == domain[0] && O.xupper == domain[1] && O.ylower == domain[2] &&
               O.yupper == domain[3] && O.zlower == domain[4] && O.zupper == domain[5];
        for (int k = 0; same && k < 6; k++) same = O.mthbc[k] == mthbc[k];
        if (same) return AMRGPU_OK;
    }
    amr3_release();
    Hier3Geom& H = A.H;
    H.nlev = nlev;
    for (int l = 1; l <= nlev; l++) {
        Level3Geom& G = H.lev[l];
        G.iregsz = iregsz[l - 1];
        G.jregsz = jregsz[l - 1];
        G.kregsz = kregsz[l - 1];
        G.ratx = ratx[l - 1];
        G.raty = raty[l - 1];
        G.ratz = ratz[l - 1];
        G.hx = hx[l - 1];
        G.hy = hy[l - 1];
        G.hz = hz[l - 1];
        if (G.iregsz < 1 || G.jregsz < 1 || G.kregsz < 1 || G.ratx < 1 || G.raty < 1 ||
            G.ratz < 1 || !(G.hx > 0) || !(G.hy > 0) || !(G.hz > 0))
            return ctx3_fail(AMRGPU_ERR_ARG, "bad geometry of level %d", l);
    }
    H.xlower = domain[0];
    H.xupper = domain[1];
    H.ylower = domain[2];
    H.yupper = domain[3];
    H.zlower = domain[4];
    H.zupper = domain[5];
    for (int s = 0; s < 6; s++) H.mthbc[s] = mthbc[s];
    // amr3.f90: a direction is periodic when both its sides are
    H.xperdom = mthbc[0] == 2 && mthbc[1] == 2;
    H.yperdom = mthbc[2] == 2 && mthbc[3] == 2;
    H.zperdom = mthbc[4] == 2 && mthbc[5] == 2;
    if ((mthbc[0] == 2) != (mthbc[1] == 2) || (mthbc[2] == 2) != (mthbc[3] == 2) ||
        (mthbc[4] == 2) != (mthbc[5] == 2))
        return ctx3_fail(AMRGPU_ERR_ARG, "periodic boundaries come in pairs");
    const int np = (int)H.xperdom + (int)H.yperdom + (int)H.zperdom;
    if (np != 0 && np != 3)
        return ctx3_fail(AMRGPU_ERR_UNSUPPORTED, "3-D domains must be periodic in all three "
                         "directions or in none (saveqc.f wraps only fully periodic domains)");
    A.geom = true;
    return AMRGPU_OK;
}

int amrgpu3_set_level(int level, int ngrids, const int* boxes) {
    int rc = check_level(level);
    if (rc) return rc;
    if (ngrids < 0 || (ngrids > 0 && !boxes)) return ctx3_fail(AMRGPU_ERR_ARG, "bad grid boxes");
    Level3Geom& G = A.H.lev[level];
    Lev3& L = A.lev[level];
    const double t = L.time, pk = L.possk;
    free_level(L);
    L.time = t;
    L.possk = pk;
    L.gen = ++A.gen_counter;
    G.boxes.clear();
    const int nv = ctx3_cfg().nvar;
    int64_t qo = 0, fo = 0;
    for (int g = 0; g < ngrids; g++) {
        std::array<int, 6> b;
        for (int s = 0; s < 6; s++) b[s] = boxes[6 * g + s];
        if (b[0] < 0 || b[1] >= G.iregsz || b[2] < 0 || b[3] >= G.jregsz || b[4] < 0 ||
            b[5] >= G.kregsz || b[1] < b[0] || b[3] < b[2] || b[5] < b[4])
            return ctx3_fail(AMRGPU_ERR_ARG, "grid %d of level %d outside the level", g, level);
        Patch3 p{};
        p.mx = b[1] - b[0] + 1;
        p.my = b[3] - b[2] + 1;
        p.mz = b[5] - b[4] + 1;
        if (std::max(p.mx, std::max(p.my, p.mz)) + 3 > SWEEP3_LANES)
            return ctx3_fail(AMRGPU_ERR_UNSUPPORTED, "grid %d: a side exceeds %d cells", g,
                             SWEEP3_LANES - 3);
        p.qoff = qo;
        p.foff = fo;
        qo += (q3_elems(p, nv) + 63) / 64 * 64;
        fo += (faces3_elems(p, nv) + 63) / 64 * 64;
        G.boxes.push_back(b);
        L.boxes.push_back(b);
        L.patches.push_back(p);
    }
    L.q_elems = qo;
    L.face_elems = fo;
    // planes per sweep workgroup: a level of few grids needs small blocks to
    // fill the chip (measured on the radial run configuration: FP32 best at
    // 1, FP64 at 2); many grids keep the batch's block
    int planes = 0;
    for (const auto& p : L.patches) planes += p.mz;
    int pb = A.plane_block;
    if (pb <= 0) pb = planes >= 4096 ? ctx3_plane_block() : (ctx3_esize() == 4 ? 1 : 2);
    ctx3_plan_blocks(L.patches, pb, L.blocks);
    const size_t es = ctx3_esize();
    rc = upload(L.d_patches, L.patches);
    for (int k = 0; k < 4 && !rc; k++) rc = upload(L.d_blocks[k], L.blocks[k]);
    if (rc) return rc;
    HIPCHKA(hipMalloc(&L.d_new, std::max<int64_t>(1, qo) * es));
    HIPCHKA(hipMalloc(&L.d_old, std::max<int64_t>(1, qo) * es));
    HIPCHKA(hipMemsetAsync(L.d_new, 0, std::max<int64_t>(1, qo) * es, ctx3_stream()));
    HIPCHKA(hipMemsetAsync(L.d_old, 0, std::max<int64_t>(1, qo) * es, ctx3_stream()));
    HIPCHKA(hipMalloc((void**)&L.d_cfl, std::max(1, ngrids) * sizeof(unsigned long long)));
    HIPCHKA(hipMalloc((void**)&L.d_partial,
                      std::max<size_t>(1, L.blocks[3].size()) * sizeof(double)));
    L.h_cfl.assign(ngrids, 0);
    if (fo > A.face_cap) {
        dfree(A.d_face);
        HIPCHKA(hipMalloc(&A.d_face, fo * es));
        HIPCHKA(hipMemsetAsync(A.d_face, 0, fo * es, ctx3_stream()));
        A.face_cap = fo;
    }
    HIPCHKA(hipStreamSynchronize(ctx3_stream()));
    A.planned = false;
    A.planned_regs = false;
    return AMRGPU_OK;
}

int amrgpu3_set_plane_block(int plane_block) {
    if (!ctx3_inited()) return ctx3_fail(AMRGPU_ERR_STATE, "amrgpu3_init has not been called");
    if (plane_block < 0) return ctx3_fail(AMRGPU_ERR_ARG, "plane_block must be >= 0");
    A.plane_block = plane_block;
    return AMRGPU_OK;
}

int amrgpu3_set_time(int level, double time, double possk) {
    int rc = check_level(level);
    if (rc) return rc;
    if (!(possk > 0)) return ctx3_fail(AMRGPU_ERR_ARG, "possk must be positive");
    A.lev[level].time = time;
    A.lev[level].possk = possk;
    return AMRGPU_OK;
}

int amrgpu3_get_time(int level, double* time) {
    int rc = check_level(level);
    if (rc) return rc;
    if (!time) return ctx3_fail(AMRGPU_ERR_ARG, "null time");
    *time = A.lev[level].time;
    return AMRGPU_OK;
}

// grid g's AoS q(nvar, mx+4, my+4, mz+4) at alloc_base + offsets[g]
int amrgpu3_put_level(int level, int which, const double* alloc_base, const int64_t* offsets) {
    int rc = check_level(level);
    if (rc) return rc;
    if (!alloc_base || !offsets) return ctx3_fail(AMRGPU_ERR_ARG, "null arrays");
    Lev3& L = A.lev[level];
    const int nv = ctx3_cfg().nvar;
    const size_t es = ctx3_esize();
    std::vector<char> h(std::max<int64_t>(1, L.q_elems) * es, 0);
    for (size_t g = 0; g < L.patches.size(); g++) {
        const Patch3& p = L.patches[g];
        const int64_t cells = (int64_t)(p.mx + 4) * (p.my + 4) * (p.mz + 4);
        const double* q = alloc_base + offsets[g];
        for (int64_t c = 0; c < cells; c++)
            for (int m = 0; m < nv; m++) {
                const int64_t d = p.qoff + m * cells + c;
                if (es == 8) reinterpret_cast<double*>(h.data())[d] = q[c * nv + m];
                else reinterpret_cast<float*>(h.data())[d] = (float)q[c * nv + m];
            }
    }
    HIPCHKA(ctx3_h2d(which == 0 ? L.d_new : L.d_old, h.data(), L.q_elems * es));
    return AMRGPU_OK;
}

int amrgpu3_get_level(int level, int which, double* alloc_base, const int64_t* offsets) {
    int rc = check_level(level);
    if (rc) return rc;
    if (!alloc_base || !offsets) return ctx3_fail(AMRGPU_ERR_ARG, "null arrays");
    Lev3& L = A.lev[level];
    const int nv = ctx3_cfg().nvar;
    const size_t es = ctx3_esize();
    std::vector<char> h(std::max<int64_t>(1, L.q_elems) * es);
    HIPCHKA(ctx3_d2h(h.data(), which == 0 ? L.d_new : L.d_old, L.q_elems * es));
    for (size_t g = 0; g < L.patches.size(); g++) {
        const Patch3& p = L.patches[g];
        const int64_t cells = (int64_t)(p.mx + 4) * (p.my + 4) * (p.mz + 4);
        double* q = alloc_base + offsets[g];
        for (int64_t c = 0; c < cells; c++)
            for (int m = 0; m < nv; m++) {
                const int64_t d = p.qoff + m * cells + c;
                q[c * nv + m] = es == 8 ? reinterpret_cast<const double*>(h.data())[d]
                                        : (double)reinterpret_cast<const float*>(h.data())[d];
            }
    }
    return AMRGPU_OK;
}

int amrgpu3_bound_level(int level, double time) {
    int rc = check_level(level);
    if (rc) return rc;
    rc = ensure_fill_plans();
    if (rc) return rc;
    rc = run_fill(level, time, A.lev[level].fill);
    if (rc) return rc;
    HIPCHKA(hipStreamSynchronize(ctx3_stream()));
    return AMRGPU_OK;
}

int amrgpu3_advance_level(int level, double* dtlevnew, double* cfl_level) {
    int rc = check_level(level);
    if (rc) return rc;
    rc = ensure_plans();
    if (rc) return rc;
    Lev3& L = A.lev[level];
    const Level3Geom& G = A.H.lev[level];
    const hipStream_t st = ctx3_stream();
    if (!(L.possk > 0)) return ctx3_fail(AMRGPU_ERR_STATE, "level %d has no time step", level);
    // bound every grid (advanc.f (3-D):53-64)
    rc = run_fill(level, L.time, L.fill);
    if (rc) return rc;
    // saveqc for the next level
    if (has_fine(level)) {
        rc = run_fill(level + 1, L.time, A.lev[level + 1].saveqc);
        if (rc) return rc;
    }
    if (L.patches.empty()) {
        L.time = L.time + L.possk;
        if (dtlevnew) *dtlevnew = RINF;
        if (cfl_level) *cfl_level = 0.0;
        return AMRGPU_OK;
    }
    // old = new for levels below the finest (par_advanc)
    if (level < A.H.nlev)
        HIPCHKA(hipMemcpyAsync(L.d_old, L.d_new, L.q_elems * ctx3_esize(), hipMemcpyDeviceToDevice, st));
    if (level > 1) {
        const Reg3Args ra = reg_args(level, 0);
        HIPCHKA(launch_qad3(ctx3_cfg().solver, ra, physics(), st));
    }
    Step3IO io{};
    io.q = L.d_new;
    io.face = A.d_face;
    io.d_patches = L.d_patches;
    for (int k = 0; k < 4; k++) {
        io.d_blocks[k] = L.d_blocks[k];
        io.nblocks[k] = (int)L.blocks[k].size();
    }
    io.npatch = (int)L.patches.size();
    io.d_cfl = L.d_cfl;
    rc = ctx3_step(io, G.hx, G.hy, G.hz, L.possk);
    if (rc) return rc;
    if (has_fine(level)) {
        Reg3Args ra = reg_args(level + 1, level);
        HIPCHKA(launch_fluxsv3(ra, st));
    }
    if (level > 1) {
        const Reg3Args ra = reg_args(level, 0);
        HIPCHKA(launch_fluxad3(ra, st));
    }
    HIPCHKA(ctx3_d2h(L.h_cfl.data(), L.d_cfl, L.patches.size() * sizeof(unsigned long long)));
    double dmin = RINF, cmax = 0.0;
    const double cfl = ctx3_cfg().cfl;
    for (size_t g = 0; g < L.patches.size(); g++) {
        double c;
        std::memcpy(&c, &L.h_cfl[g], sizeof c);
        const double dtnew = c > 0.0 ? L.possk * cfl / c : RINF;
        dmin = std::min(dmin, dtnew);
        cmax = std::max(cmax, c);
    }
    L.time = L.time + L.possk;
    if (dtlevnew) *dtlevnew = dmin;
    if (cfl_level) *cfl_level = cmax;
    return AMRGPU_OK;
}

int amrgpu3_update_level(int level) {
    int rc = check_level(level);
    if (rc) return rc;
    rc = ensure_plans();
    if (rc) return rc;
    if (!has_fine(level)) return AMRGPU_OK;
    const Reg3Args ra = reg_args(level + 1, level);
    HIPCHKA(launch_update3_level(ra, ctx3_stream()));
    return AMRGPU_OK;
}

int amrgpu3_conck_level(int level, double* totmass) {
    int rc = check_level(level);
    if (rc) return rc;
    if (!totmass) return ctx3_fail(AMRGPU_ERR_ARG, "null totmass");
    Lev3& L = A.lev[level];
    const int np = (int)L.patches.size();
    double s = 0.0;   // the sequential sum over the level (conck.f (3-D):47-60)
    if (np) {
        HIPCHKA(launch_conck3((int)ctx3_esize(), L.d_new, L.d_patches, np, L.d_partial,
                              ctx3_stream()));
        HIPCHKA(ctx3_d2h(&s, L.d_partial, sizeof(double)));
    }
    const Level3Geom& G = A.H.lev[level];
    *totmass = s * G.hx * G.hy * G.hz;
    return AMRGPU_OK;
}

int amrgpu3_flag_level(int level, double tolsp, int flag_richardson, double tol, int iorder,
                       int64_t* nflag, int64_t* nrich) {
    int rc = check_level(level);
    if (rc) return rc;
    rc = ensure_fill_plans();
    if (rc) return rc;
    Lev3& L = A.lev[level];
    const hipStream_t st = ctx3_stream();
    const int n = (int)L.patches.size();
    // per-grid flag arrays (mx+4, my+4, mz+4), goodpt = 0 outside the interior
    if (!L.d_flags) {
        L.foff.assign(n + 1, 0);
        for (int g = 0; g < n; g++)
            L.foff[g + 1] = L.foff[g] + (int64_t)(L.patches[g].mx + 4) * (L.patches[g].my + 4) *
                                            (L.patches[g].mz + 4);
        const int64_t tot = std::max<int64_t>(1, L.foff[n]);
        HIPCHKA(hipMalloc((void**)&L.d_flags, tot * sizeof(double)));
        HIPCHKA(hipMalloc((void**)&L.d_rflags, tot * sizeof(double)));
        rc = upload(L.d_foff, L.foff);
        if (rc) return rc;
    }
    const int64_t tot = std::max<int64_t>(1, L.foff[n]);
    HIPCHKA(hipMemsetAsync(L.d_flags, 0, tot * sizeof(double), st));
    HIPCHKA(hipMemsetAsync(L.d_rflags, 0, tot * sizeof(double), st));
    unsigned long long* d_cnt = nullptr;
    HIPCHKA(hipMalloc((void**)&d_cnt, 2 * sizeof(unsigned long long)));
    HIPCHKA(hipMemsetAsync(d_cnt, 0, 2 * sizeof(unsigned long long), st));
    // spest.f (3-D): ghost cells at the grids' time, then flag2refine
    rc = run_fill(level, L.time, L.fill);
    if (rc) { hipFree(d_cnt); return rc; }
    if (n)
        HIPCHKA(launch_flag3((int)ctx3_esize(), ctx3_cfg().nvar, L.d_new, L.d_flags, L.d_foff,
                             L.d_patches, L.d_blocks[3], (int)L.blocks[3].size(), tolsp, d_cnt, st));
    if (flag_richardson && n) {
        // errest.f (3-D) + errf1.f
        if (level == A.H.nlev) {
            hipFree(d_cnt);
            return ctx3_fail(AMRGPU_ERR_ARG, "Richardson estimation on the finest level needs its "
                             "old time, which the finest level does not keep");
        }
        Rich3Plan R;
        std::string err;
        if (!plan_richardson3(A.H, level, ctx3_cfg().nvar, R, err)) {
            hipFree(d_cnt);
            return ctx3_fail(AMRGPU_ERR_ARG, "level %d Richardson: %s", level, err.c_str());
        }
        DevFill D;
        D.plan = R.fill;
        double* arena = nullptr;
        void* face = nullptr;
        Patch3* d_pf = nullptr;
        Patch3* d_pc = nullptr;
        Block3* d_bf[4] = {};
        Block3* d_bc[4] = {};
        int64_t* d_cstart = nullptr;
        auto cleanup = [&]() {
            hipStreamSynchronize(st);
            free_fill(D);
            hipFree(arena);
            hipFree(face);
            hipFree(d_pf);
            hipFree(d_pc);
            for (int k = 0; k < 4; k++) {
                hipFree(d_bf[k]);
                hipFree(d_bc[k]);
            }
            hipFree(d_cstart);
            hipFree(d_cnt);
        };
        const int nv = ctx3_cfg().nvar;
        int64_t fo = 0;
        for (auto& p : R.crs) {
            p.foff = fo;
            fo += (faces3_elems(p, nv) + 63) / 64 * 64;
        }
        int64_t fo2 = 0;
        for (auto& p : R.fine) {
            p.foff = fo2;
            fo2 += (faces3_elems(p, nv) + 63) / 64 * 64;
        }
        std::vector<Block3> bf[4], bc[4];
        ctx3_plan_blocks(R.fine, 2, bf);
        ctx3_plan_blocks(R.crs, 2, bc);
        std::vector<int64_t> cstart(n + 1, 0);
        for (int g = 0; g < n; g++)
            cstart[g + 1] = cstart[g] + (int64_t)R.crs[g].mx * R.crs[g].my * R.crs[g].mz;
        hipError_t e = hipMalloc((void**)&arena, std::max<int64_t>(1, R.arena_elems) * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&face, std::max<int64_t>(1, std::max(fo, fo2)) * sizeof(double));
        if (e != hipSuccess) {
            cleanup();
            return ctx3_fail(AMRGPU_ERR_HIP, "Richardson buffers: %s", hipGetErrorString(e));
        }
        rc = upload_fill(D);
        if (!rc) rc = upload(d_pf, R.fine);
        if (!rc) rc = upload(d_pc, R.crs);
        for (int k = 0; k < 4 && !rc; k++) {
            rc = upload(d_bf[k], bf[k]);
            if (!rc) rc = upload(d_bc[k], bc[k]);
        }
        if (!rc) rc = upload(d_cstart, cstart);
        // the fill at t - possk: copysol of the old values, bound with 2
        // nghost, coarsen; the copy of the new values with their ghost cells
        if (!rc) rc = run_fill(level, L.time - L.possk, D, arena);
        const Level3Geom& G = A.H.lev[level];
        if (!rc) {
            Step3IO io{};
            io.q = arena;
            io.face = face;
            io.d_patches = d_pc;
            for (int k = 0; k < 4; k++) {
                io.d_blocks[k] = d_bc[k];
                io.nblocks[k] = (int)bc[k].size();
            }
            io.npatch = n;
            io.d_cfl = L.d_cfl;
            rc = ctx3_step(io, 2.0 * G.hx, 2.0 * G.hy, 2.0 * G.hz, 2. * L.possk, 8);
        }
        if (!rc) {
            Step3IO io{};
            io.q = arena;
            io.face = face;
            io.d_patches = d_pf;
            for (int k = 0; k < 4; k++) {
                io.d_blocks[k] = d_bf[k];
                io.nblocks[k] = (int)bf[k].size();
            }
            io.npatch = n;
            io.d_cfl = L.d_cfl;
            rc = ctx3_step(io, G.hx, G.hy, G.hz, L.possk, 8);
        }
        if (!rc) {
            const double order = (double)((1 << (iorder + 1)) - 2);
            e = launch_errf1_3(arena, d_pf, d_pc, L.d_foff, d_cstart, n, cstart[n], tol, order,
                               L.d_rflags, d_cnt + 1, st);
            if (e != hipSuccess) rc = ctx3_fail(AMRGPU_ERR_HIP, "errf1: %s", hipGetErrorString(e));
        }
        unsigned long long cnt[2] = {0, 0};
        if (!rc) {
            e = ctx3_d2h(cnt, d_cnt, sizeof cnt);
            if (e != hipSuccess) rc = ctx3_fail(AMRGPU_ERR_HIP, "counts: %s", hipGetErrorString(e));
        }
        cleanup();
        if (rc) return rc;
        if (nflag) *nflag = (int64_t)cnt[0];
        if (nrich) *nrich = (int64_t)cnt[1];
        return AMRGPU_OK;
    }
    unsigned long long cnt[2] = {0, 0};
    hipError_t e = ctx3_d2h(cnt, d_cnt, sizeof cnt);
    hipFree(d_cnt);
    if (e != hipSuccess) return ctx3_fail(AMRGPU_ERR_HIP, "counts: %s", hipGetErrorString(e));
    if (nflag) *nflag = (int64_t)cnt[0];
    if (nrich) *nrich = 0;
    return AMRGPU_OK;
}

int amrgpu3_get_level_flags(int level, int grid, double* amrflags, double* rflags) {
    int rc = check_level(level);
    if (rc) return rc;
    Lev3& L = A.lev[level];
    if (!L.d_flags) return ctx3_fail(AMRGPU_ERR_STATE, "amrgpu3_flag_level has not been called");
    if (grid < 0 || grid >= (int)L.patches.size())
        return ctx3_fail(AMRGPU_ERR_ARG, "grid %d out of range", grid);
    const size_t n = (size_t)(L.foff[grid + 1] - L.foff[grid]);
    if (amrflags) HIPCHKA(ctx3_d2h(amrflags, L.d_flags + L.foff[grid], n * sizeof(double)));
    if (rflags) HIPCHKA(ctx3_d2h(rflags, L.d_rflags + L.foff[grid], n * sizeof(double)));
    return AMRGPU_OK;
}

int amrgpu3_get_registers(int level, int grid, double* svdflx_qc1d, int* lenbc) {
    int rc = check_level(level);
    if (rc) return rc;
    if (level < 2) return ctx3_fail(AMRGPU_ERR_ARG, "level 1 has no registers");
    rc = ensure_plans();
    if (rc) return rc;
    Lev3& L = A.lev[level];
    if (grid < 0 || grid >= (int)L.reg.regs.size())
        return ctx3_fail(AMRGPU_ERR_ARG, "grid %d out of range", grid);
    const RegGrid3 rg = L.reg.regs[grid];
    if (lenbc) *lenbc = rg.lenbc;
    if (svdflx_qc1d)
        HIPCHKA(ctx3_d2h(svdflx_qc1d, L.d_reg + rg.reg_off,
                         2 * (size_t)ctx3_cfg().nvar * rg.lenbc * sizeof(double)));
    return AMRGPU_OK;
}

}  // extern "C"
