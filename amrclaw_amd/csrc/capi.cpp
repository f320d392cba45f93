// capi.cpp -- host side of libamrgpu: the C ABI of include/amrgpu.h.
//
// Owns the device copies of every level (patch table, tile table, two
// solution buffers), plans the tiling of each level and launches the kernels
// on one stream.  Mirrors the reference's level bookkeeping (node/rnode,
// listOfGrids; amr_module.f90:30-151) for the data the level update needs.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/amrgpu.h"
#include "fill_plan.hpp"
#include "gauges.hpp"
#include "kernels.hpp"
#include "noalloc.hpp"
#include "regs_plan.hpp"

namespace amrgpu {
hipError_t launch_fill_u64(unsigned long long* p, unsigned long long v, int64_t n,
                           hipStream_t st);
hipError_t launch_aos_soa(bool to_soa, int esize, double* stage, const int64_t* stage_off,
                          const int64_t* off, const int32_t* plane, const int64_t* cell_start,
                          int npatch, int nv, int64_t ncell, void* level, hipStream_t st);
hipError_t launch_fill_stage(int esize, const FillArgs& a, hipStream_t st);
hipError_t launch_gauges(int esize, const GaugeArgs& a, hipStream_t st);
}

using namespace amrgpu;

namespace {

constexpr int MAXLV = 12;          // reference maxlv = 10 (+ scratch slot 0)
constexpr int NEVENTS = 16;
// rinfinity = 10.e32 declared from a default-real literal (amr_module.f90)
const double RINFINITY = (double)10.e32f;

struct Level {
    int ngrids = 0;
    double hx = 0, hy = 0;
    std::vector<PatchDesc> patches;
    std::vector<TileDesc> tiles;
    PatchDesc* d_patches = nullptr;
    TileDesc* d_tiles = nullptr;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;                    // buf[cur]: new time (store1)
    int64_t nelem = 0;
    unsigned long long* d_cfl = nullptr;
    size_t lds = 0;
    int layout = 0;   // fixed LDS layout of the Euler sweep every tile fits (1 28x12, 2 12x12)
    // layout generation (a new value at every amrgpu_set_level of this
    // level) and the generations the register arena d_reg was laid out for:
    // this level's, the next coarser level's and the geometry's
    uint64_t gen = 0;
    uint64_t reg_gen = 0, reg_gen_coarse = 0, reg_geom = 0;
    int64_t reg_elems = 0;
    // listbc of every grid as the caller built it (amrgpu_set_listbc), kid
    // as an index on the next finer level; empty: planned here
    std::vector<std::vector<std::array<int, 5>>> listbc;
    std::vector<char> listbc_set;
    // march sweep (launch mode AMRGPU_LAUNCH_MARCH): its wave table
    std::vector<MarchWave> mwaves;
    MarchWave* d_mwaves = nullptr;
    // bulk transfer tables: patch element offsets, planes, first cell
    int64_t* d_off = nullptr;
    int32_t* d_plane = nullptr;
    int64_t* d_cell_start = nullptr;
    int64_t ncell_total = 0;       // cells incl. ghosts over all patches
    // ghost-fill program (bound) of this level
    bool plan_valid = false;
    FillPlan plan;
    FillOp* d_ops = nullptr;
    FillBuf* d_bufs = nullptr;
    unsigned char* d_masks = nullptr;
    double* d_scratch = nullptr;
    // time of the level's grids (rnode(timemult)) and its step (possk)
    double time = 0.0, possk = 0.0;
    // coarse-fine registers: this level's arena (fine side) and the device
    // programs of fluxsv / upbnd / averaging (coarse side) and qad (fine side)
    double* d_reg = nullptr;
    SvEntry* d_sv = nullptr;
    QadItem* d_qad = nullptr;
    int nqad = 0;
    UpbndCell* d_ucells = nullptr;
    UpbndEntry* d_uents = nullptr;
    int nucells = 0;
    AvgBlock* d_avg = nullptr;
    int64_t* d_avg_start = nullptr;
    int navg = 0;
    int64_t navg_cells = 0;
    // saveqc program filling this level's qc1d from the next coarser level
    FillPlan sq;
    FillOp* d_sq_ops = nullptr;
    FillBuf* d_sq_bufs = nullptr;
    double* d_sq_scratch = nullptr;
    double* d_sq_aux = nullptr;          // the saveqc program's aux scratch (aux mode)
    // aux arrays of the grids (FP64, see Context::auxk), set by
    // amrgpu_put_level_aux
    double* d_aux = nullptr;
    bool aux_set = false;               // amrgpu_put_level_aux since the last layout
    double* d_partial = nullptr;        // conck total (one double; sized one per grid)
    // grid corners rnode(cornxlo/cornylo) (default ilo*hx + xlower)
    std::vector<double> cxlo, cylo;
    // flagging: FP64 flags of every grid, region rectangles, counters
    double* d_flags = nullptr;
    int64_t flags_cap = 0;
    int64_t* d_foff = nullptr;
    int foff_mbuff = -1;
    std::vector<int64_t> foff;
    RegionRect* d_rects = nullptr;
    int64_t rects_cap = 0;
    unsigned long long* d_counters = nullptr;
    // Richardson estimate (errest.f): its fill program, the two FP64 arenas
    // (inputs of the two stepgrid calls and their results) and the tiles
    bool rich_valid = false;
    RichPlan rich;
    FillOp* d_rops = nullptr;
    FillBuf* d_rbufs = nullptr;
    unsigned char* d_rmasks = nullptr;
    double* d_rscratch = nullptr;
    double* d_rin = nullptr;
    double* d_rout = nullptr;
    PatchDesc* d_rcrs = nullptr;
    PatchDesc* d_rfine = nullptr;
    std::vector<TileDesc> rtiles_c, rtiles_f;
    TileDesc* d_rtiles_c = nullptr;
    TileDesc* d_rtiles_f = nullptr;
    size_t rlds_c = 0, rlds_f = 0;
    unsigned long long* d_rcfl = nullptr;
    // per-patch launch mode: tiles grouped by patch
    std::vector<int> patch_tile_start;
    // captured level steps, keyed by everything their launch arguments hold
    struct Graph {
        std::vector<double> key;
        hipGraphExec_t exec = nullptr;
        uint64_t last_use = 0;
    };
    std::vector<Graph> graphs;
};

// Plug-in solvers (amrgpu_register_solver): the application's tables, ids
// AMRGPU_SOLVER_PLUGIN0 + k; registrations and constants outlive finalize.
constexpr int NPLUGIN = 16;
amrgpu_solver_ops g_plugins[NPLUGIN];
int g_nplugins = 0;
double g_solver_params[AMRGPU_SOLVER_PARAMS_MAX] = {};

struct Context {
    bool inited = false;
    amrgpu_config cfg{};
    const amrgpu_solver_ops* plugin = nullptr;   // the run's solver when it is a plug-in
    MethodArgs meth{};
    size_t esize = 8;
    bool mixed = false;   // AMRGPU_MIXED: FP32 sweep on FP64 storage, flux form
    // aux arrays: a level's FP64 aux array holds grid p's component c at
    // patches[p].off * auxk + c * plane (auxk = ceil(naux / nvar) keeps the
    // grids' blocks apart); 0 without aux
    int auxk = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[NEVENTS] = {};
    int txmax = 28, tymax = 28;
    int sweep_wgs = 0;   // occupancy bound of the mixed sweep (0: its default)
    int march_wgs = 0;   // occupancy bound of the march sweep (0: its default)
    Level levels[MAXLV];
    int mxnest = 0;
    int iregsz[MAXLV] = {}, jregsz[MAXLV] = {}, ratx[MAXLV] = {}, raty[MAXLV] = {};
    double hxposs[MAXLV] = {}, hyposs[MAXLV] = {};
    unsigned long long* h_cfl = nullptr;   // pinned: the per-level-step Courant read-back
    int h_cfl_cap = 0;
    bool regs_valid = false;
    int launch_mode = AMRGPU_LAUNCH_BATCHED;   // or per patch, or persistent
    unsigned long long* d_next = nullptr;       // tile counter of the persistent sweep
    bool timing = false;                        // per-kernel-class HIP events
    std::vector<std::pair<hipEvent_t, hipEvent_t>> tev[AMRGPU_KT_COUNT];
    size_t tev_used[AMRGPU_KT_COUNT] = {};
    bool graphs = false;                        // replay advance_level as hipGraphs
    uint64_t layout_version = 1;                // bumped by every layout change
    uint64_t gen_counter = 0;                   // source of Level::gen values
    // gauges (amrgpu_set_gauges): their state on the device and the records
    // sampled since the last amrgpu_get_gauge_records
    std::vector<GaugeDev> gauges;
    GaugeDev* d_gauges = nullptr;
    int32_t* d_gmeta = nullptr;
    double* d_gvals = nullptr;
    unsigned int* d_gcount = nullptr;           // [0] records, [1] refused
    int gcap = 0;
    uint64_t geom_gen = 1;                      // bumped by every change of the geometry
};

Context g;

// The sweep of the run's solver in one arithmetic: 0 FP64, 1 FP32 storage,
// 2 FP32 arithmetic on FP64 storage (mixed); wgs: the mixed occupancy bound.
hipError_t run_sweep(int prec, const SweepArgs& x, size_t lds, int wgs) {
    if (g.plugin) {
        const int e = prec == 0 ? g.plugin->sweep_f64(&x, lds, g.stream)
                    : prec == 1 ? g.plugin->sweep_f32(&x, lds, g.stream)
                                : g.plugin->sweep_mixed(&x, lds, g.stream, wgs);
        return (hipError_t)e;
    }
    if (prec == 2) return launch_sweep_mixed(g.cfg.solver, x, lds, g.stream, wgs);
    return prec == 0 ? launch_sweep_f64(g.cfg.solver, x, lds, g.stream)
                     : launch_sweep_f32(g.cfg.solver, x, lds, g.stream);
}

int run_precision() { return g.mixed ? 2 : (g.esize == 8 ? 0 : 1); }

// qad + fluxad of a fine level (the solver's rpn2 at the coarse-fine edges)
hipError_t run_qad(const QadArgs& a) {
    if (g.plugin) return (hipError_t)g.plugin->qad((int)g.esize, &a, &g.meth.ph, g.stream);
    return launch_finereg(g.cfg.solver, (int)g.esize, a, g.meth.ph, g.stream);
}
std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                            \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess)                                                   \
            return fail(AMRGPU_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// Every host <-> device transfer and clear goes through the library's
// stream, so it is ordered after the kernels already queued there (the
// stream does not synchronize with the null stream).
hipError_t h2d(void* dst, const void* src, size_t n) {
    if (n == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, g.stream);
    return e != hipSuccess ? e : hipStreamSynchronize(g.stream);
}
hipError_t d2h(void* dst, const void* src, size_t n) {
    if (n == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, g.stream);
    return e != hipSuccess ? e : hipStreamSynchronize(g.stream);
}
hipError_t dzero(void* p, size_t n) { return n ? hipMemsetAsync(p, 0, n, g.stream) : hipSuccess; }
// wait for the queued work before freeing or overwriting what it may use
void quiesce() {
    if (g.stream) hipStreamSynchronize(g.stream);
}

// per-kernel-class timing: an event pair around each launch on the stream
void tbegin(int kind) {
    if (!g.timing) return;
    auto& v = g.tev[kind];
    size_t& u = g.tev_used[kind];
    if (u == v.size()) {
        std::pair<hipEvent_t, hipEvent_t> p{};
        if (hipEventCreate(&p.first) != hipSuccess || hipEventCreate(&p.second) != hipSuccess) return;
        v.push_back(p);
    }
    hipEventRecord(v[u].first, g.stream);
}
void tend(int kind) {
    if (!g.timing) return;
    auto& v = g.tev[kind];
    size_t& u = g.tev_used[kind];
    if (u < v.size()) {
        hipEventRecord(v[u].second, g.stream);
        u++;
    }
}

void free_rich(Level& L) {
    quiesce();
    void* ps[] = {L.d_rops, L.d_rbufs, L.d_rmasks, L.d_rscratch, L.d_rin, L.d_rout, L.d_rcrs,
                  L.d_rfine, L.d_rtiles_c, L.d_rtiles_f, L.d_rcfl};
    for (void* p : ps)
        if (p) hipFree(p);
    L.d_rops = nullptr; L.d_rbufs = nullptr; L.d_rmasks = nullptr; L.d_rscratch = nullptr;
    L.d_rin = nullptr; L.d_rout = nullptr; L.d_rcrs = nullptr; L.d_rfine = nullptr;
    L.d_rtiles_c = nullptr; L.d_rtiles_f = nullptr; L.d_rcfl = nullptr;
    L.rich = RichPlan();
    L.rtiles_c.clear();
    L.rtiles_f.clear();
    L.rich_valid = false;
}

void free_plan(Level& L) {
    quiesce();
    free_rich(L);
    if (L.d_ops) hipFree(L.d_ops);
    if (L.d_bufs) hipFree(L.d_bufs);
    if (L.d_masks) hipFree(L.d_masks);
    if (L.d_scratch) hipFree(L.d_scratch);
    L.d_ops = nullptr;
    L.d_bufs = nullptr;
    L.d_masks = nullptr;
    L.d_scratch = nullptr;
    L.plan = FillPlan();
    L.plan_valid = false;
}

void free_regs(Level& L) {
    quiesce();
    void* ps[] = {L.d_reg, L.d_sv, L.d_qad, L.d_ucells, L.d_uents, L.d_avg, L.d_avg_start,
                  L.d_sq_ops, L.d_sq_bufs, L.d_sq_scratch, L.d_sq_aux};
    for (void* p : ps)
        if (p) hipFree(p);
    L.d_sq_aux = nullptr;
    L.d_reg = nullptr;
    L.d_sv = nullptr;
    L.d_qad = nullptr;
    L.d_ucells = nullptr;
    L.d_uents = nullptr;
    L.d_avg = nullptr;
    L.d_avg_start = nullptr;
    L.d_sq_ops = nullptr;
    L.d_sq_bufs = nullptr;
    L.d_sq_scratch = nullptr;
    L.nqad = L.nucells = L.navg = 0;
    L.navg_cells = 0;
    L.sq = FillPlan();
    for (auto& t : L.tiles) t.reg_lo = t.reg_hi = 0;
    for (auto& p : L.patches) p.reg_off = 0, p.lenbc = 0;
}

void free_graphs(Level& L) {
    quiesce();
    for (auto& gr : L.graphs)
        if (gr.exec) hipGraphExecDestroy(gr.exec);
    L.graphs.clear();
}

void free_level(Level& L) {
    quiesce();
    free_graphs(L);
    free_plan(L);
    free_regs(L);
    if (L.d_partial) hipFree(L.d_partial);
    void* fl[] = {L.d_flags, L.d_foff, L.d_rects, L.d_counters};
    for (void* p : fl)
        if (p) hipFree(p);
    if (L.d_patches) hipFree(L.d_patches);
    if (L.d_tiles) hipFree(L.d_tiles);
    if (L.d_mwaves) hipFree(L.d_mwaves);
    for (auto& b : L.buf)
        if (b) hipFree(b);
    if (L.d_cfl) hipFree(L.d_cfl);
    if (L.d_aux) hipFree(L.d_aux);
    if (L.d_off) hipFree(L.d_off);
    if (L.d_plane) hipFree(L.d_plane);
    if (L.d_cell_start) hipFree(L.d_cell_start);
    double t = L.time, pk = L.possk;
    L = Level();
    L.time = t;
    L.possk = pk;
}

void free_gauges() {
    quiesce();
    void* ps[] = {g.d_gauges, g.d_gmeta, g.d_gvals, g.d_gcount};
    for (void* p : ps)
        if (p) hipFree(p);
    g.d_gauges = nullptr;
    g.d_gmeta = nullptr;
    g.d_gvals = nullptr;
    g.d_gcount = nullptr;
    g.gauges.clear();
    g.gcap = 0;
}

int seg_width(int n) { return n + 3 <= 16 ? 16 : (n + 3 <= 32 ? 32 : 64); }

// split n cells into tiles of at most cap cells, boundaries on multiples of 4
// (so that the r fine faces of a coarse register segment never straddle two
// tiles for refinement ratios 1, 2, 4)
std::vector<std::pair<int, int>> split_axis(int n, int cap) {
    std::vector<std::pair<int, int>> out;
    int k = (n + cap - 1) / cap;
    int t = (n + k - 1) / k;
    t = (t + 3) / 4 * 4;
    if (t > cap) t = std::max(4, cap / 4 * 4);
    for (int s = 0; s < n; s += t) out.push_back({s, std::min(t, n - s)});
    return out;
}

// The tiling of a level for tile caps (txmax, tymax) and a launch mode:
// the tile table, the LDS bytes of its largest tile and the fixed-layout
// variant every tile fits.  Pure: nothing of the level changes, so a caller
// can refuse caps before any level has been re-planned.
struct Tiling {
    std::vector<TileDesc> tiles;
    std::vector<int> patch_tile_start;
    size_t lds = 0;
    int layout = 0;
};

Tiling make_tiling(const Level& L, int txmax, int tymax, int launch_mode) {
    Tiling T;
    const int nv = g.cfg.nvar;
    for (int p = 0; p < L.ngrids; p++) {
        const PatchDesc& pd = L.patches[p];
        auto xs = split_axis(pd.mx, txmax), ys = split_axis(pd.my, tymax);
        for (auto& y : ys)
            for (auto& x : xs) {
                TileDesc t{};
                t.patch = p;
                t.i0 = (int16_t)x.first; t.tx = (int16_t)x.second;
                t.j0 = (int16_t)y.first; t.ty = (int16_t)y.second;
                t.wx = (int16_t)seg_width(t.tx);
                t.wy = (int16_t)seg_width(t.ty);
                T.tiles.push_back(t);
                size_t b = (g.esize == 8 && !g.mixed) ? sweep_lds_bytes<double>(nv, t.tx, t.ty)
                                                      : sweep_lds_bytes<float>(nv, t.tx, t.ty, g.mixed);
                T.lds = std::max(T.lds, b);
            }
    }
    // the Euler sweeps have variants with their LDS laid out for the tuned
    // tile caps (constant strides), used when every tile fits: 28x12 for
    // FP32 arithmetic (mixed and FP32 storage), 12x12 for FP64
    const bool f32 = g.mixed || g.esize == 4;
    const int lx = f32 ? 28 : 12, ly = 12;
    T.layout = (g.cfg.solver == SOLVER_EULER && !T.tiles.empty()) ? (f32 ? 1 : 2) : 0;
    for (const auto& t : T.tiles)
        if (t.tx > lx || t.ty > ly) T.layout = 0;
    if (T.layout)
        T.lds = std::max(T.lds, f32 ? sweep_lds_bytes<float>(nv, lx, ly, g.mixed)
                                    : sweep_lds_bytes<double>(nv, lx, ly));
    if (launch_mode == AMRGPU_LAUNCH_PER_PATCH) {
        // one launch per patch: the patch's tiles are contiguous
        T.patch_tile_start.assign(L.ngrids + 1, 0);
        for (const auto& t : T.tiles) T.patch_tile_start[t.patch + 1]++;
        for (int p = 0; p < L.ngrids; p++) T.patch_tile_start[p + 1] += T.patch_tile_start[p];
    } else {
        // largest tiles first (the reference orders grids by cost, regrid.f:132-169)
        std::stable_sort(T.tiles.begin(), T.tiles.end(), [](const TileDesc& a, const TileDesc& b) {
            return a.tx * a.ty > b.tx * b.ty;
        });
    }
    return T;
}

// The march sweep's waves (sweep_march.hpp).  Each grid is cut into strips
// of at most 60 columns; a strip of w columns takes a segment of 16, 32 or
// 64 lanes (w + 4 lanes: the strip and its halo), and the cut of each grid
// width minimises the lanes (fewest strips among equal cuts).  Segments of
// grids of the same height and the same width are packed into waves (four
// 16-lane, two 32-lane or one 64-lane segment per wave), which march
// together; waves are ordered tallest first (the longest marches start
// first, as the reference orders grids by cost, regrid.f:132-169).
int march_lanes(int w) { return w + 4 <= 16 ? 16 : (w + 4 <= 32 ? 32 : 64); }

std::vector<int> march_cut(int mx) {
    static std::vector<std::vector<int>> memo;
    if ((int)memo.size() <= mx) memo.resize(mx + 1);
    if (!memo[mx].empty() || mx == 0) return memo[mx];
    std::vector<int> cost(mx + 1, 0), nstrip(mx + 1, 0), last(mx + 1, 0);
    for (int n = 1; n <= mx; n++) {
        cost[n] = INT32_MAX;
        for (int w = 1; w <= std::min(60, n); w++) {
            const int c = cost[n - w] + march_lanes(w), ns = nstrip[n - w] + 1;
            if (c < cost[n] || (c == cost[n] && ns < nstrip[n])) {
                cost[n] = c;
                nstrip[n] = ns;
                last[n] = w;
            }
        }
    }
    std::vector<int> cut;
    for (int n = mx; n > 0; n -= last[n]) cut.push_back(last[n]);
    std::sort(cut.begin(), cut.end(), std::greater<int>());
    memo[mx] = cut;
    return cut;
}

std::vector<MarchWave> plan_march(const Level& L) {
    // (my, segment lanes) -> strips
    std::map<std::pair<int, int>, std::vector<MarchSeg>> groups;
    for (int p = 0; p < L.ngrids; p++) {
        const PatchDesc& pd = L.patches[p];
        int x = 0;
        for (int w : march_cut(pd.mx)) {
            groups[{pd.my, march_lanes(w)}].push_back(MarchSeg{p, (int16_t)x, (int16_t)w});
            x += w;
        }
    }
    std::vector<MarchWave> waves;
    for (auto& kv : groups) {
        const int my = kv.first.first, sw = kv.first.second, per = 64 / sw;
        const auto& segs = kv.second;
        for (size_t i = 0; i < segs.size(); i += per) {
            MarchWave mw{};
            mw.segw = (int16_t)sw;
            mw.my = (int16_t)my;
            mw.nseg = (int16_t)std::min<size_t>(per, segs.size() - i);
            for (int s = 0; s < mw.nseg; s++) mw.seg[s] = segs[i + s];
            waves.push_back(mw);
        }
    }
    std::stable_sort(waves.begin(), waves.end(),
                     [](const MarchWave& a, const MarchWave& b) { return a.my > b.my; });
    return waves;
}

constexpr size_t LDS_MAX = 160 * 1024;

// Re-tile a level with the current caps and launch mode.  A tiling whose
// largest tile needs more LDS than a CU has is refused before anything of
// the level changes.
int plan_level(Level& L) {
    Tiling T = make_tiling(L, g.txmax, g.tymax, g.launch_mode);
    if (T.lds > LDS_MAX)
        return fail(AMRGPU_ERR_ARG, "tile caps %dx%d need %zu bytes of LDS", g.txmax, g.tymax, T.lds);
    g.layout_version++;
    free_graphs(L);
    g.regs_valid = false;
    quiesce();
    L.tiles = std::move(T.tiles);
    L.patch_tile_start = std::move(T.patch_tile_start);
    L.lds = T.lds;
    L.layout = T.layout;
    if (L.d_tiles) hipFree(L.d_tiles);
    L.d_tiles = nullptr;
    if (!L.tiles.empty()) {
        HIPCHK(hipMalloc(&L.d_tiles, L.tiles.size() * sizeof(TileDesc)));
        HIPCHK(h2d(L.d_tiles, L.tiles.data(), L.tiles.size() * sizeof(TileDesc)));
    }
    if (L.d_mwaves) hipFree(L.d_mwaves);
    L.d_mwaves = nullptr;
    L.mwaves.clear();
    if (g.launch_mode == AMRGPU_LAUNCH_MARCH && L.ngrids > 0) {
        L.mwaves = plan_march(L);
        HIPCHK(hipMalloc(&L.d_mwaves, L.mwaves.size() * sizeof(MarchWave)));
        HIPCHK(h2d(L.d_mwaves, L.mwaves.data(), L.mwaves.size() * sizeof(MarchWave)));
    }
    return AMRGPU_OK;
}

// Whether every defined level can be tiled with caps (txmax, tymax) in a
// launch mode; the caller changes nothing when one cannot.
int check_tiling(int txmax, int tymax, int launch_mode) {
    for (const auto& L : g.levels)
        if (L.ngrids) {
            Tiling T = make_tiling(L, txmax, tymax, launch_mode);
            if (T.lds > LDS_MAX)
                return fail(AMRGPU_ERR_ARG, "tile caps %dx%d need %zu bytes of LDS", txmax, tymax,
                            T.lds);
        }
    return AMRGPU_OK;
}

int check_level(int level) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    if (level < 0 || level >= MAXLV) return fail(AMRGPU_ERR_ARG, "level %d out of range", level);
    return AMRGPU_OK;
}

int define_level(int level, int ngrids, const int* ilo, const int* ihi, const int* jlo,
                 const int* jhi, double hx, double hy) {
    Level& L = g.levels[level];
    free_level(L);
    // fill programs of this and finer levels read this level's layout
    for (int l = level; l < MAXLV; l++) free_plan(g.levels[l]);
    for (auto& Lv : g.levels) free_graphs(Lv);
    g.regs_valid = false;
    g.layout_version++;
    L.ngrids = ngrids;
    L.hx = hx;
    L.hy = hy;
    L.gen = ++g.gen_counter;
    L.listbc.assign(ngrids, {});
    L.listbc_set.assign(ngrids, 0);
    if (level >= 2) {
        // the coarser level's lists name this level's grids by index
        Level& C = g.levels[level - 1];
        C.listbc.assign(C.ngrids, {});
        C.listbc_set.assign(C.ngrids, 0);
    }
    const int ng = NGHOST, nv = g.cfg.nvar;
    int64_t off = 0;
    for (int p = 0; p < ngrids; p++) {
        PatchDesc pd{};
        pd.ilo = ilo[p]; pd.ihi = ihi[p]; pd.jlo = jlo[p]; pd.jhi = jhi[p];
        pd.mx = ihi[p] - ilo[p] + 1;
        pd.my = jhi[p] - jlo[p] + 1;
        if (pd.mx < 1 || pd.my < 1)
            return fail(AMRGPU_ERR_ARG, "grid %d of level %d is empty", p, level);
        if (pd.mx > 30000 || pd.my > 30000)
            return fail(AMRGPU_ERR_ARG, "grid %d of level %d too large", p, level);
        pd.mitot = pd.mx + 2 * ng;
        pd.plane = pd.mitot * (pd.my + 2 * ng);
        pd.off = off;
        off += (int64_t)nv * pd.plane;
        off = (off + 31) / 32 * 32;   // 256-byte aligned patch starts
        L.patches.push_back(pd);
    }
    L.nelem = off;
    for (int p = 0; p < ngrids; p++) {
        L.cxlo.push_back(L.patches[p].ilo * hx + g.cfg.xlower);    // grdfit2.f:98-101
        L.cylo.push_back(L.patches[p].jlo * hy + g.cfg.ylower);
    }
    if (ngrids > 0) {
        HIPCHK(hipMalloc(&L.d_patches, ngrids * sizeof(PatchDesc)));
        HIPCHK(h2d(L.d_patches, L.patches.data(), ngrids * sizeof(PatchDesc)));
        for (auto& b : L.buf) {
            HIPCHK(hipMalloc(&b, (size_t)L.nelem * g.esize));
            HIPCHK(dzero(b, (size_t)L.nelem * g.esize));
        }
        HIPCHK(hipMalloc(&L.d_cfl, ngrids * sizeof(unsigned long long)));
        HIPCHK(hipMalloc(&L.d_partial, ngrids * sizeof(double)));
        if (g.auxk > 0) {
            HIPCHK(hipMalloc(&L.d_aux, (size_t)L.nelem * g.auxk * sizeof(double)));
            HIPCHK(dzero(L.d_aux, (size_t)L.nelem * g.auxk * sizeof(double)));
            L.aux_set = false;
        }
        std::vector<int64_t> off(ngrids), cs(ngrids);
        std::vector<int32_t> pl(ngrids);
        int64_t c = 0;
        for (int p = 0; p < ngrids; p++) {
            off[p] = L.patches[p].off;
            pl[p] = L.patches[p].plane;
            cs[p] = c;
            c += L.patches[p].plane;
        }
        L.ncell_total = c;
        HIPCHK(hipMalloc(&L.d_off, ngrids * sizeof(int64_t)));
        HIPCHK(hipMalloc(&L.d_plane, ngrids * sizeof(int32_t)));
        HIPCHK(hipMalloc(&L.d_cell_start, ngrids * sizeof(int64_t)));
        HIPCHK(h2d(L.d_off, off.data(), ngrids * sizeof(int64_t)));
        HIPCHK(h2d(L.d_plane, pl.data(), ngrids * sizeof(int32_t)));
        HIPCHK(h2d(L.d_cell_start, cs.data(), ngrids * sizeof(int64_t)));
    }
    int rc = plan_level(L);
    if (rc) free_level(L);   // refused tiling: leave the level without grids
    return rc;
}

// AoS FP64 host <-> SoA device (storage precision)
int put_grid(Level& L, int grid, int which, const double* q) {
    const PatchDesc& pd = L.patches[grid];
    const int nv = g.cfg.nvar;
    const int64_t n = (int64_t)nv * pd.plane;
    void* dst = (char*)L.buf[which == 0 ? L.cur : 1 - L.cur] + pd.off * g.esize;
    if (g.esize == 8) {
        std::vector<double> h(n);
        for (int64_t c = 0; c < pd.plane; c++)
            for (int m = 0; m < nv; m++) h[m * pd.plane + c] = q[c * nv + m];
        HIPCHK(h2d(dst, h.data(), n * 8));
    } else {
        std::vector<float> h(n);
        for (int64_t c = 0; c < pd.plane; c++)
            for (int m = 0; m < nv; m++) h[m * pd.plane + c] = (float)q[c * nv + m];
        HIPCHK(h2d(dst, h.data(), n * 4));
    }
    return AMRGPU_OK;
}

int get_grid(Level& L, int grid, int which, double* q) {
    const PatchDesc& pd = L.patches[grid];
    const int nv = g.cfg.nvar;
    const int64_t n = (int64_t)nv * pd.plane;
    const void* src = (const char*)L.buf[which == 0 ? L.cur : 1 - L.cur] + pd.off * g.esize;
    HIPCHK(hipStreamSynchronize(g.stream));
    if (g.esize == 8) {
        std::vector<double> h(n);
        HIPCHK(d2h(h.data(), src, n * 8));
        for (int64_t c = 0; c < pd.plane; c++)
            for (int m = 0; m < nv; m++) q[c * nv + m] = h[m * pd.plane + c];
    } else {
        std::vector<float> h(n);
        HIPCHK(d2h(h.data(), src, n * 4));
        for (int64_t c = 0; c < pd.plane; c++)
            for (int m = 0; m < nv; m++) q[c * nv + m] = (double)h[m * pd.plane + c];
    }
    return AMRGPU_OK;
}

// Bulk transfer of a whole level between the reference's heap layout
// (alloc(loc + ...), one AoS FP64 block per grid at offsets[p]) and the
// device.  Grids are gathered into one staging buffer, copied in one
// transfer and transposed on the device.
int xfer_level(Level& L, bool to_dev, int which, double* alloc_base, const int64_t* offsets) {
    if (L.ngrids == 0) return AMRGPU_OK;
    const int nv = g.cfg.nvar;
    std::vector<int64_t> soff(L.ngrids);
    int64_t tot = 0;
    for (int p = 0; p < L.ngrids; p++) {
        soff[p] = tot;
        tot += (int64_t)nv * L.patches[p].plane;
    }
    double* h_stage = nullptr;
    double* d_stage = nullptr;
    int64_t* d_soff = nullptr;
    HIPCHK(hipHostMalloc((void**)&h_stage, tot * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipMalloc(&d_stage, tot * sizeof(double)));
    HIPCHK(hipMalloc(&d_soff, L.ngrids * sizeof(int64_t)));
    HIPCHK(h2d(d_soff, soff.data(), L.ngrids * sizeof(int64_t)));
    void* lvl = L.buf[which == 0 ? L.cur : 1 - L.cur];
    int rc = AMRGPU_OK;
    hipError_t e = hipSuccess;
    if (to_dev) {
        for (int p = 0; p < L.ngrids; p++)
            std::memcpy(h_stage + soff[p], alloc_base + offsets[p],
                        (size_t)nv * L.patches[p].plane * sizeof(double));
        e = hipMemcpyAsync(d_stage, h_stage, tot * sizeof(double), hipMemcpyHostToDevice, g.stream);
        if (e == hipSuccess)
            e = launch_aos_soa(true, (int)g.esize, d_stage, d_soff, L.d_off, L.d_plane,
                               L.d_cell_start, L.ngrids, nv, L.ncell_total, lvl, g.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
    } else {
        e = launch_aos_soa(false, (int)g.esize, d_stage, d_soff, L.d_off, L.d_plane,
                           L.d_cell_start, L.ngrids, nv, L.ncell_total, lvl, g.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(h_stage, d_stage, tot * sizeof(double), hipMemcpyDeviceToHost, g.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
        if (e == hipSuccess)
            for (int p = 0; p < L.ngrids; p++)
                std::memcpy(alloc_base + offsets[p], h_stage + soff[p],
                            (size_t)nv * L.patches[p].plane * sizeof(double));
    }
    if (e != hipSuccess) rc = fail(AMRGPU_ERR_HIP, "level transfer: %s", hipGetErrorString(e));
    hipFree(d_soff);
    hipFree(d_stage);
    hipHostFree(h_stage);
    return rc;
}

// hierarchy geometry for the planner
HierGeom make_geom() {
    HierGeom H;
    H.nlev = 0;
    for (int l = 1; l < MAXLV; l++)
        if (g.levels[l].ngrids > 0 || l <= g.mxnest) H.nlev = l;
    for (int l = 1; l <= H.nlev; l++) {
        LevelGeom& G = H.lev[l];
        for (const auto& pd : g.levels[l].patches) G.boxes.push_back({pd.ilo, pd.ihi, pd.jlo, pd.jhi});
        G.iregsz = g.iregsz[l];
        G.jregsz = g.jregsz[l];
        G.ratx = g.ratx[l];
        G.raty = g.raty[l];
        G.hx = g.hxposs[l];
        G.hy = g.hyposs[l];
        const Level& L = g.levels[l];
        bool any = false;
        for (char c : L.listbc_set) any = any || c;
        if (any) G.listbc = L.listbc;
    }
    H.xlower = g.cfg.xlower;
    H.xupper = g.cfg.xupper;
    H.ylower = g.cfg.ylower;
    H.yupper = g.cfg.yupper;
    H.xperdom = g.cfg.mthbc[0] == 2 && g.cfg.mthbc[1] == 2;    // amr2.f90:279-280
    H.yperdom = g.cfg.mthbc[2] == 2 && g.cfg.mthbc[3] == 2;
    H.spheredom = false;
    for (int k = 0; k < 4; k++) H.mthbc[k] = g.cfg.mthbc[k];
    H.naux = g.cfg.naux;
    return H;
}

int ensure_plan(int level) {
    Level& L = g.levels[level];
    if (L.plan_valid) return AMRGPU_OK;
    if (g.mxnest < level) return fail(AMRGPU_ERR_STATE, "amrgpu_set_geometry must cover level %d", level);
    free_plan(L);
    HierGeom H = make_geom();
    std::string err;
    if (!plan_fill_level(H, level, g.cfg.nvar, L.plan, err))
        return fail(AMRGPU_ERR_ARG, "level %d ghost fill: %s", level, err.c_str());
    const FillPlan& P = L.plan;
    if (!P.ops.empty()) {
        HIPCHK(hipMalloc(&L.d_ops, P.ops.size() * sizeof(FillOp)));
        HIPCHK(h2d(L.d_ops, P.ops.data(), P.ops.size() * sizeof(FillOp)));
    }
    if (!P.bufs.empty()) {
        HIPCHK(hipMalloc(&L.d_bufs, P.bufs.size() * sizeof(FillBuf)));
        HIPCHK(h2d(L.d_bufs, P.bufs.data(), P.bufs.size() * sizeof(FillBuf)));
    }
    if (!P.masks.empty()) {
        HIPCHK(hipMalloc(&L.d_masks, P.masks.size()));
        HIPCHK(h2d(L.d_masks, P.masks.data(), P.masks.size()));
    }
    if (P.scratch_elems > 0) HIPCHK(hipMalloc(&L.d_scratch, P.scratch_elems * sizeof(double)));
    L.plan_valid = true;
    return AMRGPU_OK;
}

// intfil's choice of new / old / blended source values (intfil.f90:194-234)
int time_weights(double t, FillArgs& a, uint32_t mask) {
    const double t_epsilon = 1.0e-4;
    for (int l = 1; l < MAXLV; l++) {
        if (!(mask & (1u << l))) continue;
        const Level& S = g.levels[l];
        if (!(S.possk > 0)) return fail(AMRGPU_ERR_STATE, "level %d has no time step (amrgpu_set_time)", l);
        double alphac = (S.time - t) / S.possk;
        double alphai = 1.0 - alphac;
        if (alphai < -t_epsilon || alphai > 1.0 + t_epsilon)
            return fail(AMRGPU_ERR_ARG, "time %.16e not available from level %d (time %.16e)", t, l, S.time);
        a.alphai[l] = alphai;
        a.alphac[l] = alphac;
        if (std::fabs(alphai - 1.0) < t_epsilon) {
            a.tmode[l] = TIME_NEW;
        } else {
            if (l == g.mxnest)
                return fail(AMRGPU_ERR_ARG, "finest level %d keeps no old-time values", l);
            a.tmode[l] = std::fabs(alphai) < t_epsilon ? TIME_OLD : TIME_BLEND;
        }
    }
    return AMRGPU_OK;
}

// launch the stages of a fill program (ghost fill or saveqc) at time t
// clear / nclear: an array the first stage's launch zeroes on the way (the
// level's Courant maxima before its sweep); *cleared reports whether it ran
int run_fill(const FillPlan& P, const FillOp* d_ops, const FillBuf* d_bufs,
             const unsigned char* d_masks, double* d_scratch, double t,
             unsigned long long* clear = nullptr, int64_t nclear = 0, bool* cleared = nullptr) {
    if (cleared) *cleared = false;
    FillArgs a{};
    int rc = time_weights(t, a, P.src_levels_mask);
    if (rc) return rc;
    a.bufs = d_bufs;
    a.scratch = d_scratch;
    a.masks = d_masks;
    for (int l = 1; l < MAXLV && l < FILL_MAXLV; l++) {
        Level& S = g.levels[l];
        a.lvl_new[l] = S.buf[S.cur];
        a.lvl_old[l] = S.buf[1 - S.cur];
        a.lvl_patches[l] = S.d_patches;
        a.lvl_reg[l] = S.d_reg;
        a.hx[l] = g.hxposs[l];
        a.hy[l] = g.hyposs[l];
    }
    a.xlower = g.cfg.xlower;
    a.ylower = g.cfg.ylower;
    for (int k = 0; k < 4; k++) a.mthbc[k] = g.cfg.mthbc[k];
    a.nv = g.cfg.nvar;
    tbegin(AMRGPU_KT_FILL);
    for (int s = 0; s < P.nstages; s++) {
        a.ops = d_ops + P.stage_start[s];
        a.nops = P.stage_start[s + 1] - P.stage_start[s];
        const bool clr = clear && nclear > 0 && a.nops > 0 && (!cleared || !*cleared);
        a.clear = clr ? clear : nullptr;
        a.nclear = clr ? nclear : 0;
        hipError_t e = launch_fill_stage((int)g.esize, a, g.stream);
        if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "fill stage %d: %s", s, hipGetErrorString(e));
        if (clr && cleared) *cleared = true;
        if (clr && !cleared) clear = nullptr;
    }
    tend(AMRGPU_KT_FILL);
    return AMRGPU_OK;
}

int bound_level(int level, double t) {
    Level& L = g.levels[level];
    if (L.ngrids == 0) return AMRGPU_OK;
    int rc = ensure_plan(level);
    if (rc) return rc;
    return run_fill(L.plan, L.d_ops, L.d_bufs, L.d_masks, L.d_scratch, t);
}

template <typename T>
int upload(T** dst, const std::vector<T>& v) {
    quiesce();
    if (*dst) hipFree(*dst);
    *dst = nullptr;
    if (v.empty()) return AMRGPU_OK;
    HIPCHK(hipMalloc(dst, v.size() * sizeof(T)));
    HIPCHK(h2d(*dst, v.data(), v.size() * sizeof(T)));
    return AMRGPU_OK;
}

// Register layout and programs of the whole hierarchy (prepf.f / prepc.f
// at every regrid).  Arenas start at zero, as prepf.f:44-48 leaves them.
int ensure_regs() {
    if (g.regs_valid) return AMRGPU_OK;
    if (g.mxnest < 1) return fail(AMRGPU_ERR_STATE, "amrgpu_set_geometry has not been called");
    for (int l = 1; l < MAXLV; l++) {
        const Level& L = g.levels[l];
        int nset = 0;
        for (char c : L.listbc_set) nset += c ? 1 : 0;
        if (nset > 0 && nset < L.ngrids)
            return fail(AMRGPU_ERR_STATE, "amrgpu_set_listbc: %d of the %d grids of level %d have "
                        "a list", nset, L.ngrids, l);
    }
    HierGeom H = make_geom();
    const int nv = g.cfg.nvar;
    std::vector<RegLevel> R(FILL_MAXLV);
    std::string err;
    if (!plan_registers(H, nv, R.data(), err)) return fail(AMRGPU_ERR_ARG, "registers: %s", err.c_str());
    for (int l = 1; l <= H.nlev && l < MAXLV; l++) {
        Level& L = g.levels[l];
        // a regrid replaces the levels above lbase only (regrid.f:72-118):
        // the registers of a fine level whose grids, coarser grids and
        // geometry are unchanged keep what they have accumulated so far in
        // the current coarse step (the reference keeps node(ffluxptr) of
        // such grids); the others start from zero (prepf.f)
        const bool keep = l >= 2 && L.d_reg && L.reg_gen == L.gen &&
                          L.reg_gen_coarse == g.levels[l - 1].gen && L.reg_geom == g.geom_gen &&
                          L.reg_elems == R[l].reg_elems;
        double* kept = keep ? L.d_reg : nullptr;
        if (keep) L.d_reg = nullptr;
        free_regs(L);
        L.reg_elems = 0;
        if (L.ngrids == 0) continue;
        const RegLevel& Rl = R[l];
        int rc;
        if (l >= 2 && !Rl.reg_off.empty()) {
            for (int p = 0; p < L.ngrids; p++) {
                L.patches[p].reg_off = Rl.reg_off[p];
                L.patches[p].lenbc = Rl.lenbc[p];
            }
            if (kept) {
                L.d_reg = kept;
            } else {
                HIPCHK(hipMalloc(&L.d_reg, Rl.reg_elems * sizeof(double)));
                HIPCHK(dzero(L.d_reg, Rl.reg_elems * sizeof(double)));
            }
            L.reg_elems = Rl.reg_elems;
            L.reg_gen = L.gen;
            L.reg_gen_coarse = g.levels[l - 1].gen;
            L.reg_geom = g.geom_gen;
            if ((rc = upload(&L.d_qad, Rl.qad))) return rc;
            L.nqad = (int)Rl.qad.size();
            if (!plan_saveqc(H, l, nv, L.sq, err))
                return fail(AMRGPU_ERR_ARG, "level %d saveqc: %s", l, err.c_str());
            if ((rc = upload(&L.d_sq_ops, L.sq.ops))) return rc;
            if ((rc = upload(&L.d_sq_bufs, L.sq.bufs))) return rc;
            if (L.sq.scratch_elems > 0)
                HIPCHK(hipMalloc(&L.d_sq_scratch, L.sq.scratch_elems * sizeof(double)));
            // the aux mode's scratch too, here and not in the level step (which
            // a hipGraph may capture)
            if (g.auxk > 0 && L.sq.scratch_elems > 0)
                HIPCHK(hipMalloc(&L.d_sq_aux, L.sq.scratch_elems * g.auxk * sizeof(double)));
        }
        if (!Rl.sv_grid_start.empty()) {
            // fluxsv entries that write their slot, grouped by the tile that
            // owns their coarse cell
            std::vector<std::vector<int>> tiles_of(L.ngrids);
            for (int t = 0; t < (int)L.tiles.size(); t++) tiles_of[L.tiles[t].patch].push_back(t);
            std::vector<std::vector<SvEntry>> per_tile(L.tiles.size());
            for (int m = 0; m < L.ngrids; m++)
                for (int e = Rl.sv_grid_start[m]; e < Rl.sv_grid_start[m + 1]; e++) {
                    if (!Rl.sv_write[e]) continue;
                    const SvEntry& s = Rl.sv[e];
                    int owner = -1;
                    for (int t : tiles_of[m]) {
                        const TileDesc& td = L.tiles[t];
                        if (s.I > td.i0 && s.I <= td.i0 + td.tx && s.J > td.j0 && s.J <= td.j0 + td.ty) {
                            owner = t;
                            break;
                        }
                    }
                    if (owner < 0)
                        return fail(AMRGPU_ERR_ARG, "register entry (%d,%d) outside grid %d of level %d",
                                    s.I, s.J, m, l);
                    per_tile[owner].push_back(s);
                }
            std::vector<SvEntry> flat;
            for (size_t t = 0; t < L.tiles.size(); t++) {
                L.tiles[t].reg_lo = (int32_t)flat.size();
                flat.insert(flat.end(), per_tile[t].begin(), per_tile[t].end());
                L.tiles[t].reg_hi = (int32_t)flat.size();
            }
            if ((rc = upload(&L.d_sv, flat))) return rc;
            if ((rc = upload(&L.d_ucells, Rl.ucells))) return rc;
            if ((rc = upload(&L.d_uents, Rl.uents))) return rc;
            L.nucells = (int)Rl.ucells.size();
            if ((rc = upload(&L.d_avg, Rl.avg))) return rc;
            if ((rc = upload(&L.d_avg_start, Rl.avg_start))) return rc;
            L.navg = (int)Rl.avg.size();
            L.navg_cells = Rl.avg_start.empty() ? 0 : Rl.avg_start.back();
        }
        HIPCHK(h2d(L.d_patches, L.patches.data(), L.ngrids * sizeof(PatchDesc)));
        HIPCHK(h2d(L.d_tiles, L.tiles.data(), L.tiles.size() * sizeof(TileDesc)));
    }
    g.regs_valid = true;
    return AMRGPU_OK;
}

struct RegLink {
    double* reg = nullptr;
    double* reg_fine = nullptr;
    const PatchDesc* fine_patches = nullptr;
    const SvEntry* sv = nullptr;
};

// the sweep of every grid of a level; buf[1-cur] receives the new values
// (the caller swaps the roles afterwards)
int launch_sweep(Level& L, double dt, double* flux_out_dev, const int64_t* flux_off_dev,
                 const RegLink& rl = RegLink(), bool cfl_cleared = false) {
    if (L.ngrids == 0) return AMRGPU_OK;
    if (!cfl_cleared) HIPCHK(launch_fill_u64(L.d_cfl, 0ull, L.ngrids, g.stream));
    SweepArgs a{};
    a.qin = L.buf[L.cur];
    a.qout = L.buf[1 - L.cur];
    a.patches = L.d_patches;
    a.tiles = L.d_tiles;
    a.ntiles = (int)L.tiles.size();
    a.layout = L.layout;
    a.dt = dt;
    a.dx = L.hx;
    a.dy = L.hy;
    a.meth = g.meth;
    a.cfl_patch = L.d_cfl;
    a.flux_out = flux_out_dev;
    a.flux_off = flux_off_dev;
    a.reg = rl.reg;
    a.reg_fine = rl.reg_fine;
    a.fine_patches = rl.fine_patches;
    a.sv = rl.sv;
    if (g.auxk > 0) {
        a.aux = L.d_aux;
        a.naux = g.cfg.naux;
        a.auxk = g.auxk;
        a.mcapa = g.cfg.mcapa;
    }
    auto launch = [&](const SweepArgs& x) { return run_sweep(run_precision(), x, L.lds, g.sweep_wgs); };
    hipError_t e = hipSuccess;
    tbegin(AMRGPU_KT_SWEEP);
    if (g.launch_mode == AMRGPU_LAUNCH_MARCH && L.d_mwaves && !flux_out_dev && !rl.sv && !a.aux) {
        // the march sweep writes no per-edge flux arrays and no fluxsv
        // entries: a level with a finer level, or a call asking for the
        // flux arrays, takes the tile sweep
        SweepArgs x = a;
        x.waves = L.d_mwaves;
        x.nwaves = (int)L.mwaves.size();
        e = run_sweep(run_precision(), x, 0, g.march_wgs);
    } else if (g.launch_mode == AMRGPU_LAUNCH_PER_PATCH && !L.patch_tile_start.empty()) {
        for (int p = 0; p < L.ngrids && e == hipSuccess; p++) {
            SweepArgs x = a;
            x.tiles = L.d_tiles + L.patch_tile_start[p];
            x.ntiles = L.patch_tile_start[p + 1] - L.patch_tile_start[p];
            if (x.ntiles > 0) e = launch(x);
        }
    } else if (g.launch_mode == AMRGPU_LAUNCH_PERSISTENT) {
        // a kernel, not a memset, clears the counter (safe under graph capture)
        e = launch_fill_u64(g.d_next, 0ull, 1, g.stream);
        SweepArgs x = a;
        x.next = g.d_next;
        if (e == hipSuccess) e = launch(x);
    } else {
        e = launch(a);
    }
    tend(AMRGPU_KT_SWEEP);
    if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "sweep launch: %s", hipGetErrorString(e));
    return AMRGPU_OK;
}

// the updated interior is now in buf[1-cur]; buf[cur] keeps time t (with
// its ghost cells) and becomes the old-time array (store2)
int sweep(Level& L, double dt, double* flux_out_dev, const int64_t* flux_off_dev,
          const RegLink& rl = RegLink()) {
    int rc = launch_sweep(L, dt, flux_out_dev, flux_off_dev, rl);
    if (rc == AMRGPU_OK && L.ngrids > 0) L.cur = 1 - L.cur;
    return rc;
}

int level_cfl(Level& L, double dt, double* cfl_level, double* dtlevnew) {
    if (L.ngrids > g.h_cfl_cap) {
        if (g.h_cfl) hipHostFree(g.h_cfl);
        g.h_cfl = nullptr;
        g.h_cfl_cap = 0;
        HIPCHK(hipHostMalloc((void**)&g.h_cfl, (size_t)L.ngrids * sizeof(unsigned long long),
                             hipHostMallocDefault));
        g.h_cfl_cap = L.ngrids;
    }
    if (L.ngrids > 0)
        HIPCHK(hipMemcpyAsync(g.h_cfl, L.d_cfl, L.ngrids * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    double cmax = 0.0, dmin = RINFINITY;
    for (int p = 0; p < L.ngrids; p++) {
        double c;
        unsigned long long b = g.h_cfl[p];
        std::memcpy(&c, &b, 8);
        cmax = std::max(cmax, c);
        double dtnew = c > 0.0 ? dt * g.cfg.cfl / c : RINFINITY;   // stepgrid.f:206-212
        dmin = std::min(dmin, dtnew);                              // advanc.f:124-126
    }
    if (cfl_level) *cfl_level = cmax;
    if (dtlevnew) *dtlevnew = dmin;
    return AMRGPU_OK;
}

// The aux part of saveqc.f / cstore.f for fine level `level`: the saveqc
// program run again in aux mode, the coarse grids' aux values copied into
// the coarse patches (NEEDS_TO_BE_SET where none reaches: the reference
// then calls the application's setaux, amr_module's default of which sets
// nothing) and each patch's ring stored in the grid's auxc1d.
int run_saveqc_aux(int level) {
    Level& F = g.levels[level];
    const FillPlan& P = F.sq;
    if (P.ops.empty()) return AMRGPU_OK;
    const int64_t n = P.scratch_elems * g.auxk;
    if (!F.d_sq_aux && n > 0)
        return fail(AMRGPU_ERR_STATE, "level %d: the saveqc aux scratch was not planned", level);
    if (n > 0) {
        const double needs = (double)10.e33f;                  // amr_module.f90:170
        unsigned long long bits;
        std::memcpy(&bits, &needs, sizeof bits);
        HIPCHK(launch_fill_u64(reinterpret_cast<unsigned long long*>(F.d_sq_aux), bits, n, g.stream));
    }
    FillArgs a{};
    a.bufs = F.d_sq_bufs;
    a.scratch = F.d_sq_aux;
    for (int l = 1; l < MAXLV && l < FILL_MAXLV; l++) {
        Level& S = g.levels[l];
        a.lvl_patches[l] = S.d_patches;
        a.lvl_reg[l] = S.d_reg;
        a.lvl_aux[l] = S.d_aux;
    }
    a.nv = g.cfg.nvar;
    a.aux_mode = 1;
    a.naux = g.cfg.naux;
    a.auxk = g.auxk;
    for (int s = 0; s < P.nstages; s++) {
        a.ops = F.d_sq_ops + P.stage_start[s];
        a.nops = P.stage_start[s + 1] - P.stage_start[s];
        hipError_t e = launch_fill_stage((int)g.esize, a, g.stream);
        if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "saveqc aux: %s", hipGetErrorString(e));
    }
    return AMRGPU_OK;
}

// With aux arrays: every level lo..hi with grids has had them handed over
// since its layout last changed (a cleared capacity would divide by zero)
int aux_ready(int lo, int hi) {
    if (g.auxk == 0) return AMRGPU_OK;
    for (int l = std::max(lo, 1); l <= hi && l < MAXLV && l <= g.mxnest; l++)
        if (g.levels[l].ngrids > 0 && !g.levels[l].aux_set)
            return fail(AMRGPU_ERR_STATE, "level %d has no aux arrays since its layout changed "
                                          "(amrgpu_put_level_aux)", l);
    return AMRGPU_OK;
}

// advanc.f:16-305 for one level: ghost fill, saveqc of the finer level,
// the sweep with fluxsv / fluxad's perimeter fluxes, qad + fluxad, time
// the launches of one level step (bound, saveqc of the finer level, sweep,
// qad + fluxad), issued on the stream with the current state; no host sync
int launch_level_step(int level) {
    Level& L = g.levels[level];
    int rc = ensure_plan(level);
    if (rc) return rc;
    bool cfl_cleared = false;   // the fill's first launch zeroes the Courant maxima
    rc = run_fill(L.plan, L.d_ops, L.d_bufs, L.d_masks, L.d_scratch, L.time, L.d_cfl,
                  L.ngrids, &cfl_cleared);
    if (rc) return rc;
    const bool has_fine = level + 1 <= g.mxnest && level + 1 < MAXLV && g.levels[level + 1].ngrids > 0;
    RegLink rl;
    if (has_fine) {
        Level& F = g.levels[level + 1];
        rc = run_fill(F.sq, F.d_sq_ops, F.d_sq_bufs, nullptr, F.d_sq_scratch, L.time);
        if (rc) return rc;
        if (g.auxk > 0) {
            rc = run_saveqc_aux(level + 1);
            if (rc) return rc;
        }
        rl.reg_fine = F.d_reg;
        rl.fine_patches = F.d_patches;
        rl.sv = L.d_sv;
    }
    if (level > 1) rl.reg = L.d_reg;
    rc = launch_sweep(L, L.possk, nullptr, nullptr, rl, cfl_cleared);
    if (rc) return rc;
    if (level > 1 && L.nqad > 0) {
        QadArgs a{};
        a.items = L.d_qad;
        a.nitems = L.nqad;
        a.nv = g.cfg.nvar;
        a.patches = L.d_patches;
        a.q = L.buf[L.cur];              // time t with its ghost cells
        a.reg = L.d_reg;
        a.ratx = g.ratx[level - 1];
        a.raty = g.raty[level - 1];
        a.hx = L.hx;
        a.hy = L.hy;
        a.delt = L.possk;
        a.flux_form = g.mixed ? 1 : 0;
        if (g.auxk > 0) {
            a.naux = g.cfg.naux;
            a.aux = L.d_aux;
            a.auxk = g.auxk;
            for (int c = 0; c < 8; c++) a.auxtype[c] = g.cfg.auxtype[c];
        }
        tbegin(AMRGPU_KT_REGISTERS);
        hipError_t e = run_qad(a);
        tend(AMRGPU_KT_REGISTERS);
        if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "register fix-up: %s", hipGetErrorString(e));
    }
    return AMRGPU_OK;
}

// Everything the launch arguments of a level step depend on besides the
// layout: the step, the buffer roles of every level and the time weights
// of the source levels of the ghost fill.
std::vector<double> step_key(int level) {
    Level& L = g.levels[level];
    std::vector<double> k;
    k.push_back((double)g.layout_version);
    k.push_back(L.possk);
    double mask = 0;
    for (int l = 1; l < MAXLV; l++) mask += g.levels[l].cur ? (double)(1 << l) : 0.0;
    k.push_back(mask);
    FillArgs a{};
    if (time_weights(L.time, a, L.plan.src_levels_mask) == AMRGPU_OK)
        for (int l = 1; l < MAXLV && l < FILL_MAXLV; l++)
            if (L.plan.src_levels_mask & (1u << l)) {
                k.push_back(a.alphai[l]);
                k.push_back(a.alphac[l]);
                k.push_back((double)a.tmode[l]);
            }
    return k;
}

// replay (or capture, on first use of a key) the level step as a hipGraph
int graph_level_step(int level) {
    Level& L = g.levels[level];
    int rc = ensure_plan(level);
    if (rc) return rc;
    for (int l = 1; l < MAXLV; l++)
        if (L.plan.src_levels_mask & (1u << l)) {
            FillArgs a{};
            rc = time_weights(L.time, a, L.plan.src_levels_mask);   // report unavailable times now
            if (rc) return rc;
            break;
        }
    static uint64_t clock = 0;
    std::vector<double> key = step_key(level);
    for (auto& gr : L.graphs)
        if (gr.key == key) {
            gr.last_use = ++clock;
            HIPCHK(hipGraphLaunch(gr.exec, g.stream));
            return AMRGPU_OK;
        }
    const bool timing = g.timing;
    g.timing = false;          // no event records inside a capture
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
    rc = launch_level_step(level);
    hipError_t e = hipStreamEndCapture(g.stream, &graph);
    g.timing = timing;
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    Level::Graph gr;
    gr.key = key;
    gr.last_use = ++clock;
    e = hipGraphInstantiate(&gr.exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    // keep the 8 most recently used steps of the level
    if (L.graphs.size() >= 8) {
        auto old = std::min_element(L.graphs.begin(), L.graphs.end(),
                                    [](const Level::Graph& a, const Level::Graph& b) {
                                        return a.last_use < b.last_use;
                                    });
        quiesce();
        hipGraphExecDestroy(old->exec);
        L.graphs.erase(old);
    }
    L.graphs.push_back(gr);
    HIPCHK(hipGraphLaunch(L.graphs.back().exec, g.stream));
    return AMRGPU_OK;
}

// advanc.f:16-305 for one level: ghost fill, saveqc of the finer level,
// the sweep with fluxsv / fluxad's perimeter fluxes, qad + fluxad, time
int advance_level(int level, double* dtlevnew, double* cfl_level) {
    Level& L = g.levels[level];
    int rc = ensure_regs();
    if (rc) return rc;
    if (L.ngrids == 0) {
        if (dtlevnew) *dtlevnew = RINFINITY;
        if (cfl_level) *cfl_level = 0.0;
        return AMRGPU_OK;
    }
    if (!(L.possk > 0)) return fail(AMRGPU_ERR_STATE, "level %d has no time step (amrgpu_set_time)", level);
    if ((rc = aux_ready(level - 1, level + 1))) return rc;   // it, the one it fills from, its finer
    rc = g.graphs ? graph_level_step(level) : launch_level_step(level);
    if (rc) return rc;
    if (!g.gauges.empty()) {
        // advanc.f:249-253: the gauges of this level's grids sample the
        // time-t values after their ghost cells are filled (buf[cur], which
        // the sweep only read)
        GaugeArgs a{};
        a.gauges = g.d_gauges;
        a.n = (int)g.gauges.size();
        a.level = level;
        a.tgrid = L.time;
        a.patches = L.d_patches;
        a.q = L.buf[L.cur];
        a.nv = g.cfg.nvar;
        a.meta = g.d_gmeta;
        a.vals = g.d_gvals;
        a.count = g.d_gcount;
        a.cap = g.gcap;
        hipError_t e = launch_gauges((int)g.esize, a, g.stream);
        if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "gauges: %s", hipGetErrorString(e));
    }
    L.cur = 1 - L.cur;                                           // new <-> old roles
    L.time = L.time + L.possk;                                   // advanc.f:289
    if (cfl_level || dtlevnew) return level_cfl(L, L.possk, cfl_level, dtlevnew);
    return AMRGPU_OK;
}

// update.f:16-189: reflux (upbnd) and average the finer level onto `level`
int update_level(int level) {
    Level& C = g.levels[level];
    if (level + 1 >= MAXLV || level + 1 > g.mxnest) return AMRGPU_OK;
    Level& F = g.levels[level + 1];
    if (C.ngrids == 0 || F.ngrids == 0) return AMRGPU_OK;
    int rc = aux_ready(level, level + 1);
    if (rc) return rc;
    if ((rc = ensure_regs())) return rc;
    UpdateArgs a{};
    a.cells = C.d_ucells;
    a.ncells = C.nucells;
    a.nv = g.cfg.nvar;
    a.entries = C.d_uents;
    a.blocks = C.d_avg;
    a.nblocks = C.navg;
    a.ratx = g.ratx[level];
    a.raty = g.raty[level];
    a.patches = C.d_patches;
    a.fine_patches = F.d_patches;
    a.q = C.buf[C.cur];
    a.qf = F.buf[F.cur];
    a.reg = F.d_reg;
    a.area = C.hx * C.hy;
    if (g.auxk > 0 && g.cfg.mcapa > 0) {
        a.aux_c = C.d_aux;
        a.aux_f = F.d_aux;
        a.mcapa = g.cfg.mcapa;
        a.auxk = g.auxk;
    }
    tbegin(AMRGPU_KT_UPDATE);
    hipError_t e = launch_update((int)g.esize, a, C.d_avg_start, C.navg_cells, g.stream);
    tend(AMRGPU_KT_UPDATE);
    if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "update: %s", hipGetErrorString(e));
    return AMRGPU_OK;
}

// FP64 tiles for the stepgrid calls of the Richardson estimate
int plan_f64_tiles(const std::vector<PatchDesc>& P, std::vector<TileDesc>& tiles, size_t& lds) {
    tiles.clear();
    lds = 0;
    const int nv = g.cfg.nvar, txmax = 28, tymax = 12;
    for (int p = 0; p < (int)P.size(); p++) {
        auto xs = split_axis(P[p].mx, txmax), ys = split_axis(P[p].my, tymax);
        for (auto& y : ys)
            for (auto& x : xs) {
                TileDesc t{};
                t.patch = p;
                t.i0 = (int16_t)x.first; t.tx = (int16_t)x.second;
                t.j0 = (int16_t)y.first; t.ty = (int16_t)y.second;
                t.wx = (int16_t)seg_width(t.tx);
                t.wy = (int16_t)seg_width(t.ty);
                tiles.push_back(t);
                lds = std::max(lds, sweep_lds_bytes<double>(nv, t.tx, t.ty));
            }
    }
    return AMRGPU_OK;
}

int ensure_rich(int level) {
    Level& L = g.levels[level];
    if (L.rich_valid) return AMRGPU_OK;
    free_rich(L);
    HierGeom H = make_geom();
    std::string err;
    if (!plan_richardson(H, level, g.cfg.nvar, L.rich, err))
        return fail(AMRGPU_ERR_ARG, "level %d Richardson estimate: %s", level, err.c_str());
    const FillPlan& P = L.rich.fill;
    int rc;
    std::vector<FillOp> ops = P.ops;
    if ((rc = upload(&L.d_rops, ops))) return rc;
    if ((rc = upload(&L.d_rbufs, P.bufs))) return rc;
    if ((rc = upload(&L.d_rmasks, P.masks))) return rc;
    if (P.scratch_elems > 0) HIPCHK(hipMalloc(&L.d_rscratch, P.scratch_elems * sizeof(double)));
    HIPCHK(hipMalloc(&L.d_rin, std::max<int64_t>(L.rich.arena_elems, 1) * sizeof(double)));
    HIPCHK(hipMalloc(&L.d_rout, std::max<int64_t>(L.rich.arena_elems, 1) * sizeof(double)));
    if ((rc = upload(&L.d_rcrs, L.rich.crs))) return rc;
    if ((rc = upload(&L.d_rfine, L.rich.fine))) return rc;
    plan_f64_tiles(L.rich.crs, L.rtiles_c, L.rlds_c);
    plan_f64_tiles(L.rich.fine, L.rtiles_f, L.rlds_f);
    if ((rc = upload(&L.d_rtiles_c, L.rtiles_c))) return rc;
    if ((rc = upload(&L.d_rtiles_f, L.rtiles_f))) return rc;
    HIPCHK(hipMalloc(&L.d_rcfl, std::max(L.ngrids, 1) * sizeof(unsigned long long)));
    L.rich_valid = true;
    return AMRGPU_OK;
}

// errest.f:4-63 for every grid of `level` after the gradient flags
int richardson(int level, bool big_old, int mbuff, double tol, int iorder) {
    Level& L = g.levels[level];
    int rc = ensure_rich(level);
    if (rc) return rc;
    // prepbigstep.f: copysol, bound at tpre = time - dt with 2 nghost,
    // coarsen; flagger's locbig copy
    const double tpre = L.time - L.possk;
    FillArgs a{};
    rc = time_weights(tpre, a, L.rich.fill.src_levels_mask);
    if (rc) return rc;
    a.bufs = L.d_rbufs;
    a.scratch = L.d_rscratch;
    a.masks = L.d_rmasks;
    a.arena = L.d_rin;
    a.big_old = big_old ? 1 : 0;
    for (int l = 1; l < MAXLV && l < FILL_MAXLV; l++) {
        Level& S = g.levels[l];
        a.lvl_new[l] = S.buf[S.cur];
        a.lvl_old[l] = S.buf[1 - S.cur];
        a.lvl_patches[l] = S.d_patches;
        a.lvl_reg[l] = S.d_reg;
        a.hx[l] = g.hxposs[l];
        a.hy[l] = g.hyposs[l];
    }
    a.xlower = g.cfg.xlower;
    a.ylower = g.cfg.ylower;
    for (int k = 0; k < 4; k++) a.mthbc[k] = g.cfg.mthbc[k];
    a.nv = g.cfg.nvar;
    tbegin(AMRGPU_KT_FLAG);
    for (int s = 0; s < L.rich.fill.nstages; s++) {
        a.ops = L.d_rops + L.rich.fill.stage_start[s];
        a.nops = L.rich.fill.stage_start[s + 1] - L.rich.fill.stage_start[s];
        hipError_t e = launch_fill_stage((int)g.esize, a, g.stream);
        if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "Richardson fill stage %d: %s", s, hipGetErrorString(e));
    }
    // the 2 dt step on 2 h (prepbigstep.f:61-64) and the dt step
    // (prepregstep.f:24-27), FP64, into the output arena
    SweepArgs sa{};
    sa.qin = L.d_rin;
    sa.qout = L.d_rout;
    sa.meth = g.meth;
    sa.cfl_patch = L.d_rcfl;
    sa.patches = L.d_rcrs;
    sa.tiles = L.d_rtiles_c;
    sa.ntiles = (int)L.rtiles_c.size();
    sa.dt = 2.0 * L.possk;
    sa.dx = 2.0 * L.hx;
    sa.dy = 2.0 * L.hy;
    hipError_t e = run_sweep(0, sa, L.rlds_c, 0);
    if (e == hipSuccess) {
        sa.patches = L.d_rfine;
        sa.tiles = L.d_rtiles_f;
        sa.ntiles = (int)L.rtiles_f.size();
        sa.dt = L.possk;
        sa.dx = L.hx;
        sa.dy = L.hy;
        e = run_sweep(0, sa, L.rlds_f, 0);
    }
    if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "Richardson steps: %s", hipGetErrorString(e));
    ErrfArgs ea{};
    ea.crs = L.d_rcrs;
    ea.fine = L.d_rfine;
    ea.crs_q = L.d_rout;
    ea.fine_q = L.d_rout;
    ea.flags = L.d_flags;
    ea.foff = L.d_foff;
    ea.counters = L.d_counters;
    ea.mbuff = mbuff;
    ea.tol = tol;
    ea.order = (double)((1 << (iorder + 1)) - 2);                 // errf1.f:35
    e = launch_errf1(ea, L.ngrids, g.stream);
    tend(AMRGPU_KT_FLAG);
    if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "errf1: %s", hipGetErrorString(e));
    return AMRGPU_OK;
}

// flagger.f:56-164 for every grid of `level`
int flag_level(int level, bool gradient, double tolsp, bool rich, double tol, int iorder,
               double start_time, int nreg, const double* regions, int mbuff, double* flags_base,
               const int64_t* offsets, int64_t* nflagged) {
    Level& L = g.levels[level];
    if (nflagged) *nflagged = 0;
    if (L.ngrids == 0) return AMRGPU_OK;
    const double t = L.time;
    // flagger.f:98: the exact test chooses the new values (after bound) or,
    // at the initial time, the old ones whose ghost cells are already set
    const bool use_old = !(start_time + L.possk != t);
    int rc;
    if (!use_old && (rc = bound_level(level, t))) return rc;
    if (L.foff_mbuff != mbuff) {
        quiesce();
        L.foff.assign(L.ngrids + 1, 0);
        for (int p = 0; p < L.ngrids; p++)
            L.foff[p + 1] = L.foff[p] + (int64_t)(L.patches[p].mx + 2 * mbuff) * (L.patches[p].my + 2 * mbuff);
        if (L.foff.back() > L.flags_cap) {
            if (L.d_flags) hipFree(L.d_flags);
            L.d_flags = nullptr;
            HIPCHK(hipMalloc(&L.d_flags, L.foff.back() * sizeof(double)));
            L.flags_cap = L.foff.back();
        }
        if (!L.d_foff) HIPCHK(hipMalloc(&L.d_foff, (L.ngrids + 1) * sizeof(int64_t)));
        HIPCHK(h2d(L.d_foff, L.foff.data(), (L.ngrids + 1) * sizeof(int64_t)));
        L.foff_mbuff = mbuff;
    }
    // flagregions2.f90:61-92: each region clipped to each grid, on the host
    // with the reference's arithmetic
    std::vector<RegionRect> rects((size_t)L.ngrids * std::max(nreg, 1));
    for (int p = 0; p < L.ngrids; p++) {
        const PatchDesc& pd = L.patches[p];
        const double dx = L.hx, dy = L.hy, xlower = L.cxlo[p], ylower = L.cylo[p];
        const double xupper = xlower + pd.mx * dx, yupper = ylower + pd.my * dy;
        for (int m = 0; m < nreg; m++) {
            const double* r = regions + 8 * m;
            RegionRect& R = rects[(size_t)p * nreg + m];
            R = RegionRect{1, 0, 1, 0, 0, 0};
            if (t < r[2] || t > r[3]) continue;
            if (xlower >= r[5] || xupper <= r[4]) continue;
            if (ylower >= r[7] || yupper <= r[6]) continue;
            R.i1 = std::max((int)std::floor((r[4] - xlower) / dx) + 1, 1);
            R.i2 = std::min((int)std::floor((r[5] - xlower) / dx) + 1, pd.mx);
            R.j1 = std::max((int)std::floor((r[6] - ylower) / dy) + 1, 1);
            R.j2 = std::min((int)std::floor((r[7] - ylower) / dy) + 1, pd.my);
            R.min_level = (int)r[0];
            R.max_level = (int)r[1];
        }
    }
    quiesce();   // the previous call's rectangles may still be in use
    if ((int64_t)rects.size() > L.rects_cap) {
        if (L.d_rects) hipFree(L.d_rects);
        L.d_rects = nullptr;
        HIPCHK(hipMalloc(&L.d_rects, rects.size() * sizeof(RegionRect)));
        L.rects_cap = (int64_t)rects.size();
    }
    if (nreg > 0)
        HIPCHK(hipMemcpyAsync(L.d_rects, rects.data(), rects.size() * sizeof(RegionRect),
                              hipMemcpyHostToDevice, g.stream));
    if (!L.d_counters) HIPCHK(hipMalloc(&L.d_counters, 2 * sizeof(unsigned long long)));
    HIPCHK(launch_fill_u64(L.d_counters, 0ull, 2, g.stream));
    FlagArgs a{};
    a.patches = L.d_patches;
    a.q = use_old ? L.buf[1 - L.cur] : L.buf[L.cur];
    a.flags = L.d_flags;
    a.foff = L.d_foff;
    a.rects = L.d_rects;
    a.counters = L.d_counters;
    a.nrect = nreg;
    a.nv = g.cfg.nvar;
    a.mbuff = mbuff;
    a.level = level;
    // without gradient flagging no cell passes: flag2refine2 is not called
    a.tolsp = gradient ? tolsp : HUGE_VAL;
    tbegin(AMRGPU_KT_FLAG);
    hipError_t e = launch_flag((int)g.esize, a, L.ngrids, g.stream);
    tend(AMRGPU_KT_FLAG);
    if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "flag: %s", hipGetErrorString(e));
    // errest runs while any flag is UNSET (flagger.f:166-176); the mbuff
    // border always is
    if (rich && (rc = richardson(level, use_old, mbuff, tol, iorder))) return rc;
    unsigned long long cnt[2];
    HIPCHK(hipMemcpyAsync(cnt, L.d_counters, sizeof cnt, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    if (cnt[1] & 1ull)
        return fail(AMRGPU_ERR_ARG, "regions: minlevel > maxlevel in a cell of level %d (flagregions2.f90:96-100)", level);
    if (nflagged) *nflagged = (int64_t)cnt[0];
    if (flags_base) {
        std::vector<double> h(L.foff.back());
        HIPCHK(d2h(h.data(), L.d_flags, h.size() * sizeof(double)));
        for (int p = 0; p < L.ngrids; p++)
            std::memcpy(flags_base + offsets[p], h.data() + L.foff[p],
                        (size_t)(L.foff[p + 1] - L.foff[p]) * sizeof(double));
    }
    return AMRGPU_OK;
}

}  // namespace

extern "C" {

int amrgpu_init(const amrgpu_config* cfg) {
    if (!cfg) return fail(AMRGPU_ERR_ARG, "cfg is NULL");
    if (g.inited) amrgpu_finalize();
    const amrgpu_solver_ops* plugin = nullptr;
    int nv_need, nw_need;
    if (cfg->solver >= 1 && cfg->solver <= 4) {
        const int nv_expect[5] = {0, 1, 3, 4, 1};
        const int nw_expect[5] = {0, 1, 2, 4, 1};
        nv_need = nv_expect[cfg->solver];
        nw_need = nw_expect[cfg->solver];
    } else if (cfg->solver >= AMRGPU_SOLVER_PLUGIN0 &&
               cfg->solver < AMRGPU_SOLVER_PLUGIN0 + g_nplugins) {
        plugin = &g_plugins[cfg->solver - AMRGPU_SOLVER_PLUGIN0];
        nv_need = plugin->nvar;
        nw_need = plugin->mwaves;
    } else {
        return fail(AMRGPU_ERR_ARG, "unknown solver %d", cfg->solver);
    }
    if (cfg->nvar != nv_need)
        return fail(AMRGPU_ERR_ARG, "solver %d needs nvar=%d", cfg->solver, nv_need);
    if (cfg->mwaves != nw_need)
        return fail(AMRGPU_ERR_ARG, "solver %d needs mwaves=%d", cfg->solver, nw_need);
    if (cfg->nghost != NGHOST) return fail(AMRGPU_ERR_UNSUPPORTED, "nghost must be 2");
    if (cfg->naux < 0 || cfg->naux > 8 || cfg->mcapa < 0 || cfg->mcapa > cfg->naux)
        return fail(AMRGPU_ERR_ARG, "naux=%d mcapa=%d: need 0 <= mcapa <= naux <= 8", cfg->naux,
                    cfg->mcapa);
    if (cfg->solver == SOLVER_VCADV && cfg->naux < 2)
        return fail(AMRGPU_ERR_ARG, "the variable-coefficient advection solver reads aux(1:2)");
    if (plugin && cfg->naux > 0)
        return fail(AMRGPU_ERR_UNSUPPORTED, "aux arrays with a plug-in solver");
    if (cfg->naux > 0 && cfg->precision == AMRGPU_MIXED)
        return fail(AMRGPU_ERR_UNSUPPORTED,
                    "aux arrays / capacity run in FP64 or FP32 (the mixed flux form needs a "
                    "state-only flux)");
    for (int k = 0; k < cfg->naux; k++)
        if (cfg->auxtype[k] < AMRGPU_AUX_CENTER || cfg->auxtype[k] > AMRGPU_AUX_CAPACITY)
            return fail(AMRGPU_ERR_ARG, "auxtype(%d)=%d invalid", k + 1, cfg->auxtype[k]);
    // a solid wall negates q(2) on the x-sides and q(3) on the y-sides
    // (bc2amr.f90:145, 190, 236, 283)
    if (((cfg->mthbc[0] == 3 || cfg->mthbc[1] == 3) && cfg->nvar < 2) ||
        ((cfg->mthbc[2] == 3 || cfg->mthbc[3] == 3) && cfg->nvar < 3))
        return fail(AMRGPU_ERR_ARG, "a solid wall negates q(2) (x-sides) or q(3) (y-sides): "
                                    "nvar=%d has no such component", cfg->nvar);
    if ((cfg->mthbc[0] == 2) != (cfg->mthbc[1] == 2) || (cfg->mthbc[2] == 2) != (cfg->mthbc[3] == 2))
        return fail(AMRGPU_ERR_ARG, "periodic boundaries must be periodic on both sides");
    for (int k = 0; k < 4; k++)
        if (cfg->mthbc[k] < 1 || cfg->mthbc[k] > 3)
            return fail(AMRGPU_ERR_UNSUPPORTED, "mthbc(%d)=%d: supported are 1 extrap, 2 periodic, 3 wall",
                        k + 1, cfg->mthbc[k]);
    if (cfg->precision != AMRGPU_FP64 && cfg->precision != AMRGPU_FP32 &&
        cfg->precision != AMRGPU_MIXED)
        return fail(AMRGPU_ERR_ARG, "precision must be AMRGPU_FP64, AMRGPU_FP32 or AMRGPU_MIXED");
    if (cfg->method[1] < 1 || cfg->method[1] > 2 || cfg->method[2] < 0 || cfg->method[2] > 2)
        return fail(AMRGPU_ERR_ARG, "method(2)=%d method(3)=%d invalid", cfg->method[1], cfg->method[2]);
    for (int mw = 0; mw < cfg->mwaves; mw++)
        if (cfg->mthlim[mw] < 0 || cfg->mthlim[mw] > 5)
            return fail(AMRGPU_ERR_ARG, "mthlim(%d)=%d invalid", mw + 1, cfg->mthlim[mw]);
    g = Context();
    g.cfg = *cfg;
    g.plugin = plugin;
    g.esize = cfg->precision == AMRGPU_FP32 ? 4 : 8;
    g.mixed = cfg->precision == AMRGPU_MIXED;
    g.auxk = cfg->naux > 0 ? (cfg->naux + cfg->nvar - 1) / cfg->nvar : 0;
    MethodArgs& M = g.meth;
    M.ph.ubar = cfg->ubar;
    M.ph.vbar = cfg->vbar;
    M.ph.rho = cfg->rho;
    M.ph.bulk = cfg->bulk;
    M.ph.cc = cfg->rho > 0 ? std::sqrt(cfg->bulk / cfg->rho) : 0.0;   // setprob: cc = sqrt(K/rho)
    M.ph.zz = cfg->rho * M.ph.cc;                                     // zz = rho*cc
    M.ph.gamma = cfg->gamma;
    M.ph.efix = cfg->efix;
    for (int k = 0; k < AMRGPU_SOLVER_PARAMS_MAX; k++) M.ph.user[k] = g_solver_params[k];
    M.order = cfg->method[1];
    M.transverse = cfg->method[2];
    M.use_fwaves = cfg->use_fwaves;
    M.mwaves = cfg->mwaves;
    for (int mw = 0; mw < NWMAX; mw++) M.mthlim[mw] = mw < cfg->mwaves ? cfg->mthlim[mw] : 0;
    HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    HIPCHK(hipMalloc(&g.d_next, sizeof(unsigned long long)));
    for (auto& e : g.ev) HIPCHK(hipEventCreate(&e));
    {
        // set the kernels' LDS limits now, outside any stream capture
        SweepArgs none{};
        for (int lay : {0, 1, 2}) {   // run-time and fixed LDS layouts
            none.layout = lay;
            HIPCHK(run_sweep(0, none, 0, 0));
            HIPCHK(run_sweep(1, none, 0, 0));
            for (int w : {5, 6, 7}) HIPCHK(run_sweep(2, none, 0, w));
        }
    }
    // the device address of amrgpu_stepgrid_noalloc's static work space on
    // this GPU
    HIPCHK(noalloc_init());
    g.txmax = 28;
    g.tymax = (g.esize == 8 && !g.mixed) ? 12 : 28;
    g.inited = true;
    return AMRGPU_OK;
}

int amrgpu_register_solver(const amrgpu_solver_ops* ops, int* solver_id) {
    if (!ops || !solver_id) return fail(AMRGPU_ERR_ARG, "NULL solver table or id");
    if (ops->abi != AMRGPU_SOLVER_ABI)
        return fail(AMRGPU_ERR_ARG, "solver table ABI %d, library %d", ops->abi, AMRGPU_SOLVER_ABI);
    if (ops->sweep_args_bytes != (int)sizeof(SweepArgs) || ops->qad_args_bytes != (int)sizeof(QadArgs) ||
        ops->physics_bytes != (int)sizeof(Physics))
        return fail(AMRGPU_ERR_ARG, "solver table built against other headers (argument blocks "
                    "%d/%d/%d bytes, library %d/%d/%d)", ops->sweep_args_bytes, ops->qad_args_bytes,
                    ops->physics_bytes, (int)sizeof(SweepArgs), (int)sizeof(QadArgs),
                    (int)sizeof(Physics));
    if (ops->nvar < 1 || ops->mwaves < 1 || ops->mwaves > NWMAX)
        return fail(AMRGPU_ERR_ARG, "solver table: nvar %d, mwaves %d (at most %d waves)", ops->nvar,
                    ops->mwaves, NWMAX);
    if (!ops->sweep_f64 || !ops->sweep_f32 || !ops->sweep_mixed || !ops->qad)
        return fail(AMRGPU_ERR_ARG, "solver table has a NULL launcher");
    if (g_nplugins >= NPLUGIN) return fail(AMRGPU_ERR_ARG, "at most %d plug-in solvers", NPLUGIN);
    g_plugins[g_nplugins] = *ops;
    *solver_id = AMRGPU_SOLVER_PLUGIN0 + g_nplugins++;
    return AMRGPU_OK;
}

int amrgpu_set_solver_params(int n, const double* params) {
    if (n < 0 || n > AMRGPU_SOLVER_PARAMS_MAX || (n > 0 && !params))
        return fail(AMRGPU_ERR_ARG, "solver params: n = %d (0..%d)", n, AMRGPU_SOLVER_PARAMS_MAX);
    for (int k = 0; k < AMRGPU_SOLVER_PARAMS_MAX; k++) g_solver_params[k] = k < n ? params[k] : 0.0;
    if (g.inited) {
        for (int k = 0; k < AMRGPU_SOLVER_PARAMS_MAX; k++) g.meth.ph.user[k] = g_solver_params[k];
        for (auto& L : g.levels) free_graphs(L);   // captured launches hold the old constants
    }
    return AMRGPU_OK;
}

int amrgpu_finalize(void) {
    if (!g.inited) return AMRGPU_OK;
    hipStreamSynchronize(g.stream);
    free_gauges();
    for (auto& L : g.levels) free_level(L);
    for (auto& v : g.tev)
        for (auto& p : v) {
            hipEventDestroy(p.first);
            hipEventDestroy(p.second);
        }
    for (auto& e : g.ev)
        if (e) hipEventDestroy(e);
    hipFree(g.d_next);
    if (g.h_cfl) hipHostFree(g.h_cfl);
    hipStreamDestroy(g.stream);
    g = Context();
    return AMRGPU_OK;
}

int amrgpu_last_error(char* buf, int len) {
    if (!buf || len <= 0) return AMRGPU_ERR_ARG;
    std::snprintf(buf, (size_t)len, "%s", g_err.c_str());
    return AMRGPU_OK;
}

int amrgpu_set_occupancy(int wgs) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    if (wgs != 0 && wgs != 2 && wgs != 3 && (wgs < 5 || wgs > 7))
        return fail(AMRGPU_ERR_ARG,
                    "occupancy bound must be 5, 6 or 7 workgroups per CU for the tile sweep, "
                    "2 or 3 waves per SIMD for the march sweep (0: defaults)");
    if (wgs == 0) {
        g.sweep_wgs = 0;
        g.march_wgs = 0;
    } else if (wgs <= 3) {
        g.march_wgs = wgs;
    } else {
        g.sweep_wgs = wgs;
    }
    for (auto& L : g.levels) free_graphs(L);   // captured launches name the kernel
    return AMRGPU_OK;
}

int amrgpu_set_tiling(int txmax, int tymax) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    if (txmax < 0 || tymax < 0 || txmax > 61 || tymax > 61)
        return fail(AMRGPU_ERR_ARG, "tile caps must be in [4, 61]");
    const int tx = txmax ? std::max(4, txmax) : g.txmax, ty = tymax ? std::max(4, tymax) : g.tymax;
    int rc0 = check_tiling(tx, ty, g.launch_mode);
    if (rc0) return rc0;
    g.txmax = tx;
    g.tymax = ty;
    for (auto& L : g.levels)
        if (L.ngrids) {
            int rc = plan_level(L);
            if (rc) return rc;
        }
    return AMRGPU_OK;
}

int amrgpu_stepgrid(int mx, int my, const double* q_in, double* q_out, double dx, double dy,
                    double dt, double* fm, double* fp, double* gm, double* gp, double* cflgrid,
                    double* dtnew) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    if (!q_in || !q_out || mx < 1 || my < 1) return fail(AMRGPU_ERR_ARG, "bad stepgrid arguments");
    if (g.auxk > 0)
        return fail(AMRGPU_ERR_UNSUPPORTED, "stepgrid takes no aux array: use a level (put_level_aux)");
    int ilo = 0, ihi = mx - 1, jlo = 0, jhi = my - 1;
    int rc = define_level(0, 1, &ilo, &ihi, &jlo, &jhi, dx, dy);
    if (rc) return rc;
    Level& L = g.levels[0];
    rc = put_grid(L, 0, 0, q_in);
    if (rc) return rc;
    const bool want_flux = fm && fp && gm && gp;
    const int nv = g.cfg.nvar;
    const int64_t plane = L.patches[0].plane, arr = nv * plane;
    double* d_flux = nullptr;
    int64_t* d_off = nullptr;
    if (want_flux) {
        HIPCHK(hipMalloc(&d_flux, 4 * arr * sizeof(double)));
        HIPCHK(dzero(d_flux, 4 * arr * sizeof(double)));
        HIPCHK(hipMalloc(&d_off, sizeof(int64_t)));
        HIPCHK(dzero(d_off, sizeof(int64_t)));
    }
    // ghost cells of the output keep the input values, as in stepgrid.f
    rc = put_grid(L, 0, 1, q_in);
    if (rc) return rc;
    rc = sweep(L, dt, d_flux, d_off);
    if (rc) return rc;
    rc = get_grid(L, 0, 0, q_out);
    if (rc) return rc;
    rc = level_cfl(L, dt, cflgrid, dtnew);
    if (rc) return rc;
    if (want_flux) {
        std::vector<double> h(4 * arr);
        HIPCHK(d2h(h.data(), d_flux, 4 * arr * sizeof(double)));
        double* outs[4] = {fm, fp, gm, gp};
        for (int k = 0; k < 4; k++)
            for (int64_t c = 0; c < plane; c++)
                for (int m = 0; m < nv; m++) outs[k][c * nv + m] = h[k * arr + m * plane + c];
        hipFree(d_flux);
        hipFree(d_off);
    }
    free_level(L);
    return AMRGPU_OK;
}

int amrgpu_stepgrid_noalloc(int mx, int my, const double* q_in, double* q_out, double dx,
                            double dy, double dt, double* cflgrid, double* dtnew) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    if (!q_in || !q_out || mx < 1 || my < 1 || !(dx > 0) || !(dy > 0))
        return fail(AMRGPU_ERR_ARG, "bad stepgrid_noalloc arguments");
    if (g.auxk > 0)
        return fail(AMRGPU_ERR_UNSUPPORTED, "stepgrid_noalloc takes no aux array");
    unsigned char* base = noalloc_scratch();
    if (!base)
        return fail(AMRGPU_ERR_STATE, "no work space on this device: call amrgpu_init on it first");
    const int nv = g.cfg.nvar;
    const int64_t plane = (int64_t)(mx + 2 * NGHOST) * (my + 2 * NGHOST);
    const int64_t bytes = ((int64_t)nv * plane * (int64_t)g.esize + 255) / 256 * 256;
    if (mx > 30000 || my > 30000 || 2 * bytes > (int64_t)NOALLOC_DATA_BYTES)
        return fail(AMRGPU_ERR_ARG,
                    "grid %dx%d needs %lld bytes of work space; amrgpu_stepgrid_noalloc holds %lld",
                    mx, my, (long long)(2 * bytes), (long long)NOALLOC_DATA_BYTES);
    // the tiling amrgpu_stepgrid's level planner gives one grid
    const auto xs = split_axis(mx, g.txmax), ys = split_axis(my, g.tymax);
    const int ntx = (int)xs.size(), nty = (int)ys.size(), tw = xs[0].second, th = ys[0].second;
    if ((int64_t)ntx * nty > NOALLOC_MAX_TILES)
        return fail(AMRGPU_ERR_ARG, "grid %dx%d has %d tiles at caps %dx%d; at most %d", mx, my,
                    ntx * nty, g.txmax, g.tymax, NOALLOC_MAX_TILES);
    const bool f32 = g.mixed || g.esize == 4;
    size_t lds = f32 ? sweep_lds_bytes<float>(nv, tw, th, g.mixed) : sweep_lds_bytes<double>(nv, tw, th);
    const int lx = f32 ? 28 : 12, ly = 12;
    int layout = 0;
    if (g.cfg.solver == SOLVER_EULER && tw <= lx && th <= ly) {
        layout = f32 ? 1 : 2;
        lds = std::max(lds, f32 ? sweep_lds_bytes<float>(nv, lx, ly, g.mixed)
                                : sweep_lds_bytes<double>(nv, lx, ly));
    }
    if (lds > LDS_MAX)
        return fail(AMRGPU_ERR_ARG, "tile caps %dx%d need %zu bytes of LDS", g.txmax, g.tymax, lds);
    unsigned char* qa = base + NOALLOC_DATA_OFF;
    unsigned char* qb = qa + bytes;
    HIPCHK(noalloc_setup(base, mx, my, tw, th, ntx, nty, g.stream));
    HIPCHK(noalloc_transpose(true, (int)g.esize, q_in, nullptr, qa, nv, (int)plane, g.stream));
    // ghost cells of the output keep the input values, as in stepgrid.f
    HIPCHK(noalloc_transpose(true, (int)g.esize, q_in, nullptr, qb, nv, (int)plane, g.stream));
    SweepArgs a{};
    a.qin = qa;
    a.qout = qb;
    a.patches = reinterpret_cast<const PatchDesc*>(base + NOALLOC_PATCH_OFF);
    a.tiles = reinterpret_cast<const TileDesc*>(base + NOALLOC_TILES_OFF);
    a.ntiles = ntx * nty;
    a.layout = layout;
    a.dt = dt;
    a.dx = dx;
    a.dy = dy;
    a.meth = g.meth;
    a.cfl_patch = reinterpret_cast<unsigned long long*>(base + NOALLOC_CFL_OFF);
    hipError_t e = run_sweep(run_precision(), a, lds, g.sweep_wgs);
    if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "sweep launch: %s", hipGetErrorString(e));
    HIPCHK(noalloc_transpose(false, (int)g.esize, nullptr, q_out, qb, nv, (int)plane, g.stream));
    unsigned long long cb = 0;
    HIPCHK(hipMemcpyAsync(&cb, base + NOALLOC_CFL_OFF, sizeof cb, hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    double c;
    std::memcpy(&c, &cb, 8);
    if (cflgrid) *cflgrid = c;
    if (dtnew) *dtnew = c > 0.0 ? dt * g.cfg.cfl / c : RINFINITY;   // stepgrid.f:206-212
    return AMRGPU_OK;
}

int amrgpu_set_listbc(int level, int grid, int n, const int* list5) {
    int rc = check_level(level);
    if (rc) return rc;
    Level& L = g.levels[level];
    if (level < 1 || grid < 0 || grid >= L.ngrids || n < 0 || (n > 0 && !list5))
        return fail(AMRGPU_ERR_ARG, "bad set_listbc arguments");
    if ((int)L.listbc.size() != L.ngrids) {
        L.listbc.assign(L.ngrids, {});
        L.listbc_set.assign(L.ngrids, 0);
    }
    auto& v = L.listbc[grid];
    v.resize(n);
    for (int e = 0; e < n; e++)
        for (int c = 0; c < 5; c++) v[e][c] = list5[5 * e + c];
    L.listbc_set[grid] = 1;
    // the fluxsv and upbnd programs come from the lists; captured level
    // steps name their tables
    g.regs_valid = false;
    g.layout_version++;
    for (auto& Lv : g.levels) free_graphs(Lv);
    return AMRGPU_OK;
}

int amrgpu_set_gauges(int n, const double* x, const double* y, const double* t_start,
                      const double* t_end, const double* min_dt) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    if (n < 0 || (n > 0 && (!x || !y || !t_start || !t_end || !min_dt)))
        return fail(AMRGPU_ERR_ARG, "bad gauge arrays");
    free_gauges();
    if (n == 0) return AMRGPU_OK;
    g.gauges.resize(n);
    for (int i = 0; i < n; i++) {
        GaugeDev& G = g.gauges[i];
        G = GaugeDev{};
        G.x = x[i];
        G.y = y[i];
        G.t_start = t_start[i];
        G.t_end = t_end[i];
        G.min_dt = min_dt[i];
        G.last_time = t_start[i];      // gauges_module.f90:131
        G.level = 0;
        G.grid = -1;
    }
    g.gcap = std::max(4096, 64 * n);
    HIPCHK(hipMalloc(&g.d_gauges, n * sizeof(GaugeDev)));
    HIPCHK(hipMalloc(&g.d_gmeta, (size_t)g.gcap * 2 * sizeof(int32_t)));
    HIPCHK(hipMalloc(&g.d_gvals, (size_t)g.gcap * (g.cfg.nvar + 1) * sizeof(double)));
    HIPCHK(hipMalloc(&g.d_gcount, 2 * sizeof(unsigned int)));
    HIPCHK(h2d(g.d_gauges, g.gauges.data(), n * sizeof(GaugeDev)));
    HIPCHK(launch_fill_u64(reinterpret_cast<unsigned long long*>(g.d_gcount), 0ull, 1, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    return AMRGPU_OK;
}

int amrgpu_set_gauge_sources(const int* level, const int* grid) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    const int n = (int)g.gauges.size();
    if (n == 0) return AMRGPU_OK;
    if (!level || !grid) return fail(AMRGPU_ERR_ARG, "bad gauge source arrays");
    // the device holds each gauge's last record time: read it back first
    HIPCHK(d2h(g.gauges.data(), g.d_gauges, n * sizeof(GaugeDev)));
    for (int i = 0; i < n; i++) {
        GaugeDev& G = g.gauges[i];
        const int l = level[i], k = grid[i];
        if (l < 1 || l >= MAXLV || k < 0 || k >= g.levels[l].ngrids) {
            G.level = 0;
            G.grid = -1;
            continue;
        }
        const Level& L = g.levels[l];
        G.level = l;
        G.grid = k;
        G.hx = L.hx;
        G.hy = L.hy;
        G.xlow = L.cxlo[k] - NGHOST * L.hx;     // advanc.f:229-230
        G.ylow = L.cylo[k] - NGHOST * L.hy;
    }
    HIPCHK(h2d(g.d_gauges, g.gauges.data(), n * sizeof(GaugeDev)));
    return AMRGPU_OK;
}

int amrgpu_get_gauge_records(int maxrec, int* nrec, int* gauge, int* level, double* vals) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    if (!nrec) return fail(AMRGPU_ERR_ARG, "nrec is NULL");
    *nrec = 0;
    if (g.gauges.empty()) return AMRGPU_OK;
    unsigned int cnt[2] = {0, 0};
    HIPCHK(d2h(cnt, g.d_gcount, sizeof cnt));
    if (cnt[1] > 0)
        return fail(AMRGPU_ERR_STATE, "%u gauge records were refused (record buffer of %d full, "
                    "or a gauge outside its source grid)", cnt[1], g.gcap);
    const int n = (int)cnt[0];
    if (n > maxrec) return fail(AMRGPU_ERR_ARG, "%d gauge records waiting, room for %d", n, maxrec);
    if (n > 0) {
        if (!gauge || !level || !vals) return fail(AMRGPU_ERR_ARG, "bad gauge record arrays");
        std::vector<int32_t> meta(2 * (size_t)n);
        HIPCHK(d2h(meta.data(), g.d_gmeta, meta.size() * sizeof(int32_t)));
        HIPCHK(d2h(vals, g.d_gvals, (size_t)n * (g.cfg.nvar + 1) * sizeof(double)));
        for (int r = 0; r < n; r++) {
            gauge[r] = meta[2 * r];
            level[r] = meta[2 * r + 1];
        }
    }
    HIPCHK(launch_fill_u64(reinterpret_cast<unsigned long long*>(g.d_gcount), 0ull, 1, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    *nrec = n;
    return AMRGPU_OK;
}

int amrgpu_set_level(int level, int ngrids, const int* ilo, const int* ihi, const int* jlo,
                     const int* jhi, double hx, double hy) {
    int rc = check_level(level);
    if (rc) return rc;
    if (level == 0) return fail(AMRGPU_ERR_ARG, "levels are numbered from 1");
    if (ngrids < 0 || (ngrids > 0 && (!ilo || !ihi || !jlo || !jhi)))
        return fail(AMRGPU_ERR_ARG, "bad grid arrays");
    if (!(hx > 0) || !(hy > 0)) return fail(AMRGPU_ERR_ARG, "mesh widths must be positive");
    return define_level(level, ngrids, ilo, ihi, jlo, jhi, hx, hy);
}

int amrgpu_put_grid(int level, int grid, int which, const double* q) {
    int rc = check_level(level);
    if (rc) return rc;
    Level& L = g.levels[level];
    if (grid < 0 || grid >= L.ngrids || (which != 0 && which != 1) || !q)
        return fail(AMRGPU_ERR_ARG, "bad put_grid arguments");
    return put_grid(L, grid, which, q);
}

int amrgpu_get_grid(int level, int grid, int which, double* q) {
    int rc = check_level(level);
    if (rc) return rc;
    Level& L = g.levels[level];
    if (grid < 0 || grid >= L.ngrids || (which != 0 && which != 1) || !q)
        return fail(AMRGPU_ERR_ARG, "bad get_grid arguments");
    return get_grid(L, grid, which, q);
}

int amrgpu_put_level(int level, int which, const double* alloc_base, const int64_t* offsets) {
    int rc = check_level(level);
    if (rc) return rc;
    if ((which != 0 && which != 1) || !alloc_base || !offsets)
        return fail(AMRGPU_ERR_ARG, "bad put_level arguments");
    return xfer_level(g.levels[level], true, which, const_cast<double*>(alloc_base), offsets);
}

int amrgpu_get_level(int level, int which, double* alloc_base, const int64_t* offsets) {
    int rc = check_level(level);
    if (rc) return rc;
    if ((which != 0 && which != 1) || !alloc_base || !offsets)
        return fail(AMRGPU_ERR_ARG, "bad get_level arguments");
    return xfer_level(g.levels[level], false, which, alloc_base, offsets);
}

int amrgpu_put_level_aux(int level, const double* alloc_base, const int64_t* offsets) {
    int rc = check_level(level);
    if (rc) return rc;
    if (!alloc_base || !offsets) return fail(AMRGPU_ERR_ARG, "bad put_level_aux arguments");
    if (g.auxk == 0) return fail(AMRGPU_ERR_STATE, "the run has no aux arrays (naux = 0)");
    Level& L = g.levels[level];
    if (L.ngrids == 0) return AMRGPU_OK;
    const int na = g.cfg.naux;
    std::vector<double> h((size_t)L.nelem * g.auxk, 0.0);
    for (int p = 0; p < L.ngrids; p++) {
        const PatchDesc& pd = L.patches[p];
        const double* src = alloc_base + offsets[p];
        double* dst = h.data() + pd.off * g.auxk;
        for (int64_t c = 0; c < pd.plane; c++)
            for (int m = 0; m < na; m++) dst[(int64_t)m * pd.plane + c] = src[c * na + m];
    }
    HIPCHK(h2d(L.d_aux, h.data(), h.size() * sizeof(double)));
    L.aux_set = true;
    return AMRGPU_OK;
}

int amrgpu_sweep_level(int level, double dt, double* cfl_level, double* dtlevnew) {
    int rc = check_level(level);
    if (rc) return rc;
    Level& L = g.levels[level];
    rc = sweep(L, dt, nullptr, nullptr);
    if (rc) return rc;
    if (cfl_level || dtlevnew) return level_cfl(L, dt, cfl_level, dtlevnew);
    return AMRGPU_OK;
}

int amrgpu_set_geometry(int mxnest, const int* iregsz, const int* jregsz, const int* intratx,
                        const int* intraty, const double* hxposs, const double* hyposs) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    if (mxnest < 1 || mxnest >= MAXLV || !iregsz || !jregsz || !intratx || !intraty || !hxposs ||
        !hyposs)
        return fail(AMRGPU_ERR_ARG, "bad geometry arguments");
    bool same = mxnest == g.mxnest;
    for (int l = 1; l <= mxnest; l++) {
        if (iregsz[l - 1] < 1 || jregsz[l - 1] < 1 || intratx[l - 1] < 1 || intraty[l - 1] < 1 ||
            !(hxposs[l - 1] > 0) || !(hyposs[l - 1] > 0))
            return fail(AMRGPU_ERR_ARG, "bad geometry of level %d", l);
        same = same && g.iregsz[l] == iregsz[l - 1] && g.jregsz[l] == jregsz[l - 1] &&
               g.ratx[l] == intratx[l - 1] && g.raty[l] == intraty[l - 1] &&
               g.hxposs[l] == hxposs[l - 1] && g.hyposs[l] == hyposs[l - 1];
    }
    // the same geometry again (a driver hands it over after every regrid):
    // nothing planned from it is stale
    if (same) return AMRGPU_OK;
    for (int l = 1; l <= mxnest; l++) {
        g.iregsz[l] = iregsz[l - 1];
        g.jregsz[l] = jregsz[l - 1];
        g.ratx[l] = intratx[l - 1];
        g.raty[l] = intraty[l - 1];
        g.hxposs[l] = hxposs[l - 1];
        g.hyposs[l] = hyposs[l - 1];
    }
    g.mxnest = mxnest;
    // fill programs and registers are planned from the geometry, and captured
    // level steps name their device tables: all of them are stale now
    for (auto& L : g.levels) {
        free_graphs(L);
        free_plan(L);
    }
    g.regs_valid = false;
    g.geom_gen++;
    g.layout_version++;
    return AMRGPU_OK;
}

int amrgpu_set_time(int level, double time, double possk) {
    int rc = check_level(level);
    if (rc) return rc;
    g.levels[level].time = time;
    g.levels[level].possk = possk;
    return AMRGPU_OK;
}

int amrgpu_bound_level(int level, double time) {
    int rc = check_level(level);
    if (rc) return rc;
    if (level < 1 || level > g.mxnest) return fail(AMRGPU_ERR_ARG, "level %d outside 1..mxnest", level);
    return bound_level(level, time);
}

int amrgpu_advance_level(int level, double* dtlevnew, double* cfl_level) {
    int rc = check_level(level);
    if (rc) return rc;
    if (level < 1 || level > g.mxnest) return fail(AMRGPU_ERR_ARG, "level %d outside 1..mxnest", level);
    return advance_level(level, dtlevnew, cfl_level);
}

int amrgpu_update_level(int level) {
    int rc = check_level(level);
    if (rc) return rc;
    if (level < 1 || level > g.mxnest) return fail(AMRGPU_ERR_ARG, "level %d outside 1..mxnest", level);
    return update_level(level);
}

int amrgpu_conck(int level, double* totmass) {
    int rc = check_level(level);
    if (rc) return rc;
    if (!totmass) return fail(AMRGPU_ERR_ARG, "totmass is NULL");
    Level& L = g.levels[level];
    *totmass = 0.0;
    if (L.ngrids == 0) return AMRGPU_OK;
    if ((rc = aux_ready(level, level))) return rc;
    hipError_t e = launch_conck((int)g.esize, L.d_patches, L.ngrids, L.buf[L.cur], L.d_partial, g.stream,
                                L.d_aux, L.d_aux ? g.cfg.mcapa : 0, g.auxk);
    if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "conck: %s", hipGetErrorString(e));
    double tot = 0.0;   // the sequential sum over the level (conck.f:36-60)
    HIPCHK(hipMemcpyAsync(&tot, L.d_partial, sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPCHK(hipStreamSynchronize(g.stream));
    *totmass = tot * L.hx * L.hy;                                   // conck.f:63
    return AMRGPU_OK;
}

int amrgpu_set_corners(int level, const double* xlo, const double* ylo) {
    int rc = check_level(level);
    if (rc) return rc;
    Level& L = g.levels[level];
    if ((L.ngrids > 0 && (!xlo || !ylo))) return fail(AMRGPU_ERR_ARG, "bad corner arrays");
    L.cxlo.assign(xlo, xlo + L.ngrids);
    L.cylo.assign(ylo, ylo + L.ngrids);
    return AMRGPU_OK;
}

int amrgpu_flag_level(int level, int flag_gradient, double tolsp, int flag_richardson,
                      double tol, int iorder, double start_time, int nregions,
                      const double* regions, int mbuff, double* flags_base,
                      const int64_t* offsets, int64_t* nflagged) {
    int rc = check_level(level);
    if (rc) return rc;
    if (level < 1 || level > g.mxnest) return fail(AMRGPU_ERR_ARG, "level %d outside 1..mxnest", level);
    if (nregions < 0 || (nregions > 0 && !regions) || mbuff < NGHOST || mbuff > 64 ||
        (flags_base && !offsets) || iorder < 1 || iorder > 8)
        return fail(AMRGPU_ERR_ARG, "bad flag_level arguments");
    if (flag_richardson && g.auxk > 0)
        return fail(AMRGPU_ERR_UNSUPPORTED, "Richardson estimation with aux arrays");
    if (flag_richardson && level == g.mxnest)
        return fail(AMRGPU_ERR_ARG, "Richardson estimation needs old values, which level %d "
                                    "(mxnest) does not keep", level);
    return flag_level(level, flag_gradient != 0, tolsp, flag_richardson != 0, tol, iorder, start_time,
                      nregions, regions, mbuff, flags_base, offsets, nflagged);
}

int amrgpu_set_launch_mode(int mode) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    if (mode != AMRGPU_LAUNCH_BATCHED && mode != AMRGPU_LAUNCH_PER_PATCH &&
        mode != AMRGPU_LAUNCH_PERSISTENT && mode != AMRGPU_LAUNCH_MARCH)
        return fail(AMRGPU_ERR_ARG, "unknown launch mode %d", mode);
    if (mode == g.launch_mode) return AMRGPU_OK;
    int rc0 = check_tiling(g.txmax, g.tymax, mode);
    if (rc0) return rc0;
    g.launch_mode = mode;
    for (auto& L : g.levels)
        if (L.ngrids) {
            int rc = plan_level(L);
            if (rc) return rc;
        }
    return AMRGPU_OK;
}

int amrgpu_set_graphs(int on) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    g.graphs = on != 0;
    if (!g.graphs)
        for (auto& L : g.levels) free_graphs(L);
    return AMRGPU_OK;
}

int amrgpu_set_timing(int on) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    quiesce();
    g.timing = on != 0;
    for (auto& u : g.tev_used) u = 0;
    return AMRGPU_OK;
}

int amrgpu_kernel_time(int kind, double* ms, int64_t* launches) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    if (kind < 0 || kind >= AMRGPU_KT_COUNT || !ms) return fail(AMRGPU_ERR_ARG, "bad kernel_time arguments");
    quiesce();
    double tot = 0.0;
    for (size_t k = 0; k < g.tev_used[kind]; k++) {
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, g.tev[kind][k].first, g.tev[kind][k].second));
        tot += t;
    }
    *ms = tot;
    if (launches) *launches = (int64_t)g.tev_used[kind];
    return AMRGPU_OK;
}

int amrgpu_get_registers(int level, int grid, double* svdflx_qc1d, int* lenbc) {
    int rc = check_level(level);
    if (rc) return rc;
    Level& L = g.levels[level];
    if (grid < 0 || grid >= L.ngrids || !svdflx_qc1d || !lenbc)
        return fail(AMRGPU_ERR_ARG, "bad get_registers arguments");
    rc = ensure_regs();
    if (rc) return rc;
    if (!L.d_reg) return fail(AMRGPU_ERR_STATE, "level %d has no coarse-fine registers", level);
    const PatchDesc& pd = L.patches[grid];
    *lenbc = pd.lenbc;
    HIPCHK(hipStreamSynchronize(g.stream));
    HIPCHK(d2h(svdflx_qc1d, L.d_reg + pd.reg_off, 2 * (size_t)g.cfg.nvar * pd.lenbc * sizeof(double)));
    return AMRGPU_OK;
}

static_assert(sizeof(amrgpu_fill_op) == sizeof(FillOp), "fill op layout");
static_assert(sizeof(amrgpu_fill_buf) == sizeof(FillBuf), "fill buffer layout");
static_assert(offsetof(amrgpu_fill_op, stage) == offsetof(FillOp, stage), "fill op layout");
static_assert(offsetof(amrgpu_fill_buf, off) == offsetof(FillBuf, off), "fill buffer layout");

static FillPlan g_inspect;

int amrgpu_regrid_level(int level, int ngrids, const int* ilo, const int* ihi, const int* jlo,
                        const int* jhi, double time) {
    int rc = check_level(level);
    if (rc) return rc;
    if (level < 2 || level >= FILL_OLDSLOT || level > g.mxnest)
        return fail(AMRGPU_ERR_ARG, "regrid: level %d outside 2..mxnest", level);
    if (ngrids < 0 || (ngrids > 0 && (!ilo || !ihi || !jlo || !jhi)))
        return fail(AMRGPU_ERR_ARG, "bad grid arrays");
    if (g.levels[level - 1].ngrids == 0)
        return fail(AMRGPU_ERR_STATE, "regrid: level %d has no grids to interpolate from", level - 1);
    if (g.auxk > 0)
        return fail(AMRGPU_ERR_UNSUPPORTED,
                    "regrid with aux arrays (filval.f90 then calls setaux and fixcapaq)");
    if (!(g.hxposs[level] > 0) || g.iregsz[level] < 1)
        return fail(AMRGPU_ERR_STATE, "regrid: the geometry of level %d is not set", level);
    Level& L = g.levels[level];
    // keep the old grids' new-time values and descriptors: define_level
    // frees the rest of the old level
    std::vector<std::array<int, 4>> old_boxes;
    for (const auto& pd : L.patches) old_boxes.push_back({pd.ilo, pd.ihi, pd.jlo, pd.jhi});
    void* old_buf = L.buf[L.cur];
    PatchDesc* old_d = L.d_patches;
    L.buf[L.cur] = nullptr;
    L.d_patches = nullptr;
    auto drop_old = [&]() {
        if (old_buf) hipFree(old_buf);
        if (old_d) hipFree(old_d);
        old_buf = nullptr;
        old_d = nullptr;
    };
    rc = define_level(level, ngrids, ilo, ihi, jlo, jhi, g.hxposs[level], g.hyposs[level]);
    if (rc) {
        drop_old();
        return rc;
    }
    L.time = time;
    if (ngrids == 0) {
        drop_old();
        return AMRGPU_OK;
    }
    HierGeom H = make_geom();
    LevelGeom& O = H.lev[FILL_OLDSLOT];
    O = H.lev[level];
    O.boxes = old_boxes;
    O.listbc.clear();
    FillPlan P;
    std::string err;
    if (!plan_filval(H, level, g.cfg.nvar, P, err)) {
        drop_old();
        return fail(AMRGPU_ERR_ARG, "level %d regrid: %s", level, err.c_str());
    }
    FillOp* d_ops = nullptr;
    FillBuf* d_bufs = nullptr;
    unsigned char* d_masks = nullptr;
    double* d_scratch = nullptr;
    auto drop_plan = [&]() {
        void* ps[] = {d_ops, d_bufs, d_masks, d_scratch};
        for (void* p : ps)
            if (p) hipFree(p);
    };
    hipError_t e = hipSuccess;
    if (!P.ops.empty()) e = hipMalloc(&d_ops, P.ops.size() * sizeof(FillOp));
    if (e == hipSuccess && !P.ops.empty()) e = h2d(d_ops, P.ops.data(), P.ops.size() * sizeof(FillOp));
    if (e == hipSuccess && !P.bufs.empty()) e = hipMalloc(&d_bufs, P.bufs.size() * sizeof(FillBuf));
    if (e == hipSuccess && !P.bufs.empty()) e = h2d(d_bufs, P.bufs.data(), P.bufs.size() * sizeof(FillBuf));
    if (e == hipSuccess && !P.masks.empty()) e = hipMalloc(&d_masks, P.masks.size());
    if (e == hipSuccess && !P.masks.empty()) e = h2d(d_masks, P.masks.data(), P.masks.size());
    if (e == hipSuccess && P.scratch_elems > 0)
        e = hipMalloc(&d_scratch, (size_t)P.scratch_elems * sizeof(double));
    FillArgs a{};
    a.bufs = d_bufs;
    a.scratch = d_scratch;
    a.masks = d_masks;
    for (int l = 1; l < MAXLV && l < FILL_MAXLV; l++) {
        Level& S = g.levels[l];
        a.lvl_new[l] = S.buf[S.cur];
        a.lvl_old[l] = S.buf[1 - S.cur];
        a.lvl_patches[l] = S.d_patches;
        a.lvl_reg[l] = S.d_reg;
        a.hx[l] = g.hxposs[l];
        a.hy[l] = g.hyposs[l];
        a.tmode[l] = TIME_NEW;
        a.alphai[l] = 1.0;
        a.alphac[l] = 0.0;
    }
    a.lvl_new[FILL_OLDSLOT] = old_buf;
    a.lvl_old[FILL_OLDSLOT] = old_buf;
    a.lvl_patches[FILL_OLDSLOT] = old_d;
    a.xlower = g.cfg.xlower;
    a.ylower = g.cfg.ylower;
    for (int k = 0; k < 4; k++) a.mthbc[k] = g.cfg.mthbc[k];
    a.nv = g.cfg.nvar;
    for (int s = 0; e == hipSuccess && s < P.nstages; s++) {
        a.ops = d_ops + P.stage_start[s];
        a.nops = P.stage_start[s + 1] - P.stage_start[s];
        e = launch_fill_stage((int)g.esize, a, g.stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
    drop_plan();
    drop_old();
    if (e != hipSuccess) return fail(AMRGPU_ERR_HIP, "level %d regrid: %s", level, hipGetErrorString(e));
    return AMRGPU_OK;
}

int amrgpu_bndry_lists(int n, const int* boxes, int iregsz, int jregsz, int* counts, int* items,
                       int64_t cap, int64_t* total) {
    if (n < 0 || (n > 0 && (!boxes || !counts)) || !total || cap < 0 || iregsz < 1 || jregsz < 1)
        return fail(AMRGPU_ERR_ARG, "bad amrgpu_bndry_lists arguments");
    std::vector<std::array<int, 4>> bx(n);
    for (int k = 0; k < n; k++)
        bx[k] = {boxes[4 * k], boxes[4 * k + 1], boxes[4 * k + 2], boxes[4 * k + 3]};
    BoxIndex index;
    index.build(bx, iregsz, jregsz);
    std::vector<int> cand;
    int64_t pos = 0;
    for (int k = 0; k < n; k++) {
        const int ilo = bx[k][0] - NGHOST, ihi = bx[k][1] + NGHOST;
        const int jlo = bx[k][2] - NGHOST, jhi = bx[k][3] + NGHOST;
        index.query(ilo, ihi, jlo, jhi, cand);
        int cnt = 0;
        // each grid found goes to the front of the reference's list, so the
        // list runs from the highest index down
        for (auto it = cand.rbegin(); it != cand.rend(); ++it) {
            const int s = *it;
            if (s == k) continue;
            const auto& b = bx[s];
            if (std::max(b[0], ilo) <= std::min(b[1], ihi) && std::max(b[2], jlo) <= std::min(b[3], jhi)) {
                if (items && pos < cap) items[pos] = s;
                pos++;
                cnt++;
            }
        }
        counts[k] = cnt;
    }
    *total = pos;
    if (pos > cap) return fail(AMRGPU_ERR_ARG, "amrgpu_bndry_lists: %lld entries, room for %lld",
                               (long long)pos, (long long)cap);
    return AMRGPU_OK;
}

int amrgpu_plan_fill(int nlev, const int* ngrids, const int* boxes, const int* iregsz,
                     const int* jregsz, const int* intratx, const int* intraty,
                     const double* hxposs, const double* hyposs, const double* domain,
                     const int* mthbc, int nvar, int level, int* nops, int* nbufs,
                     int64_t* nmask, int64_t* nscratch) {
    if (nlev < 1 || nlev >= FILL_MAXLV || level < 1 || level > nlev || !ngrids || !boxes ||
        !iregsz || !jregsz || !intratx || !intraty || !hxposs || !hyposs || !domain || !mthbc ||
        nvar < 1 || !nops || !nbufs || !nmask || !nscratch)
        return fail(AMRGPU_ERR_ARG, "bad plan_fill arguments");
    HierGeom H;
    H.nlev = nlev;
    int nb = 0;
    for (int l = 1; l <= nlev; l++) {
        LevelGeom& G = H.lev[l];
        for (int k = 0; k < ngrids[l - 1]; k++, nb++)
            G.boxes.push_back({boxes[4 * nb], boxes[4 * nb + 1], boxes[4 * nb + 2], boxes[4 * nb + 3]});
        G.iregsz = iregsz[l - 1];
        G.jregsz = jregsz[l - 1];
        G.ratx = intratx[l - 1];
        G.raty = intraty[l - 1];
        G.hx = hxposs[l - 1];
        G.hy = hyposs[l - 1];
    }
    H.xlower = domain[0];
    H.xupper = domain[1];
    H.ylower = domain[2];
    H.yupper = domain[3];
    H.xperdom = mthbc[0] == 2 && mthbc[1] == 2;
    H.yperdom = mthbc[2] == 2 && mthbc[3] == 2;
    for (int k = 0; k < 4; k++) H.mthbc[k] = mthbc[k];
    std::string err;
    if (!plan_fill_level(H, level, nvar, g_inspect, err))
        return fail(AMRGPU_ERR_ARG, "ghost fill: %s", err.c_str());
    *nops = (int)g_inspect.ops.size();
    *nbufs = (int)g_inspect.bufs.size();
    *nmask = (int64_t)g_inspect.masks.size();
    *nscratch = g_inspect.scratch_elems;
    return AMRGPU_OK;
}

static std::vector<SvEntry> g_inspect_sv;

int amrgpu_plan_listbc(int nlev, const int* ngrids, const int* boxes, const int* iregsz,
                       const int* jregsz, const int* intratx, const int* intraty,
                       const int* mthbc, int nvar, int level, int grid, int* list5,
                       int maxn, int* n) {
    if (nlev < 1 || nlev >= FILL_MAXLV || level < 1 || level >= nlev || !ngrids || !boxes ||
        !iregsz || !jregsz || !intratx || !intraty || !mthbc || nvar < 1 || !n || (maxn > 0 && !list5))
        return fail(AMRGPU_ERR_ARG, "bad plan_listbc arguments");
    HierGeom H;
    H.nlev = nlev;
    int nb = 0;
    for (int l = 1; l <= nlev; l++) {
        LevelGeom& G = H.lev[l];
        for (int k = 0; k < ngrids[l - 1]; k++, nb++)
            G.boxes.push_back({boxes[4 * nb], boxes[4 * nb + 1], boxes[4 * nb + 2], boxes[4 * nb + 3]});
        G.iregsz = iregsz[l - 1];
        G.jregsz = jregsz[l - 1];
        G.ratx = intratx[l - 1];
        G.raty = intraty[l - 1];
        G.hx = G.hy = 1.0;
    }
    H.xperdom = mthbc[0] == 2 && mthbc[1] == 2;
    H.yperdom = mthbc[2] == 2 && mthbc[3] == 2;
    if (grid < 0 || grid >= (int)H.lev[level].boxes.size()) return fail(AMRGPU_ERR_ARG, "grid out of range");
    std::vector<RegLevel> R(FILL_MAXLV);
    std::string err;
    if (!plan_registers(H, nvar, R.data(), err)) return fail(AMRGPU_ERR_ARG, "registers: %s", err.c_str());
    const RegLevel& Rl = R[level];
    *n = 0;
    if (Rl.sv_grid_start.empty()) return AMRGPU_OK;
    for (int e = Rl.sv_grid_start[grid]; e < Rl.sv_grid_start[grid + 1]; e++, (*n)++) {
        if (*n >= maxn) continue;
        const SvEntry& s = Rl.sv[e];
        // the reference's listbc(5, .) with i, j counted from the first ghost cell
        int* o = list5 + 5 * (*n);
        o[0] = s.I + NGHOST;
        o[1] = s.J + NGHOST;
        o[2] = s.side;
        o[3] = s.kid;
        o[4] = s.lkid;
    }
    return AMRGPU_OK;
}

int amrgpu_plan_fill_get(amrgpu_fill_op* ops, amrgpu_fill_buf* bufs, unsigned char* masks) {
    if (!ops || !bufs || !masks) return fail(AMRGPU_ERR_ARG, "bad plan_fill_get arguments");
    std::memcpy(ops, g_inspect.ops.data(), g_inspect.ops.size() * sizeof(FillOp));
    std::memcpy(bufs, g_inspect.bufs.data(), g_inspect.bufs.size() * sizeof(FillBuf));
    std::memcpy(masks, g_inspect.masks.data(), g_inspect.masks.size());
    return AMRGPU_OK;
}

int amrgpu_event_record(int slot) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    if (slot < 0 || slot >= NEVENTS) return fail(AMRGPU_ERR_ARG, "event slot %d", slot);
    HIPCHK(hipEventRecord(g.ev[slot], g.stream));
    return AMRGPU_OK;
}

int amrgpu_event_elapsed(int a, int b, float* ms) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    if (a < 0 || a >= NEVENTS || b < 0 || b >= NEVENTS || !ms)
        return fail(AMRGPU_ERR_ARG, "bad event arguments");
    HIPCHK(hipEventSynchronize(g.ev[b]));
    HIPCHK(hipEventElapsedTime(ms, g.ev[a], g.ev[b]));
    return AMRGPU_OK;
}

int amrgpu_synchronize(void) {
    if (!g.inited) return fail(AMRGPU_ERR_STATE, "amrgpu_init has not been called");
    HIPCHK(hipStreamSynchronize(g.stream));
    return AMRGPU_OK;
}

}  // extern "C"
static_assert(AMRGPU_SOLVER_PARAMS_MAX == SOLVER_PARAMS_MAX, "plug-in constants");
