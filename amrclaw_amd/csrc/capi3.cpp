// capi3.cpp -- host side of the 3-D patch step: the C ABI of
// include/amrgpu3.h.
//
// Owns the device copies of a batch of 3-D grids (solution in the compute
// precision, structure of arrays per grid, and the six flux arrays of each
// grid), the work lists of the four launches of a step and the library's 3-D
// stream.  Grid layout follows the reference's node(ndilo..ndkhi) boxes with
// nghost = 2 (src/3d/amr_module.f90).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/amrgpu.h"
#include "../../include/amrgpu3.h"
#include "capi3_internal.hpp"
#include "kernels3.hpp"

using namespace amrgpu;

namespace {

const double RINFINITY3 = (double)10.e32f;   // amr_module's rinfinity
constexpr int NKIND = 4;                      // x, y, z sweeps, update

struct Context3 {
    bool inited = false;
    amrgpu3_config cfg{};
    Method3 meth{};
    size_t esize = 8;
    hipStream_t stream = nullptr;
    int npatch = 0;
    int plane_block = 4;
    std::vector<Patch3> patches;
    Patch3* d_patches = nullptr;
    std::vector<Block3> blocks[NKIND];
    Block3* d_blocks[NKIND] = {};
    void* d_q = nullptr;
    void* d_face = nullptr;
    int64_t q_elems = 0, face_elems = 0;
    // the fused FP32 step (sweep3_tile.hpp): its tiles, the buffer it writes
    // (swapped with d_q after each step; both hold the ghost cells put_patch
    // wrote, which the step never changes) and its LDS size
    int step_mode = 0;                  // 0 phase kernels, 1 fused tiles
    bool fluxes_valid = false;          // the last step kept the flux arrays
    bool last_fused = false;            // the last step was the fused one
    std::vector<Tile3> tiles;
    Tile3* d_tiles = nullptr;
    void* d_q2 = nullptr;
    size_t tile_lds = 0;
    int xcd_start[TILE3_XCDS] = {}, xcd_count[TILE3_XCDS] = {};
    unsigned long long* d_cfl = nullptr;
    std::vector<unsigned long long> h_cfl;
    // flagging (flag2refine) and conck state
    double* d_flags = nullptr;          // FP64 flags of every grid, ghost border = goodpt
    int64_t* d_foff = nullptr;
    std::vector<int64_t> foff;
    unsigned long long* d_nflag = nullptr;
    double* d_partial = nullptr;        // conck total (sized one per update block)
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> tev[NKIND];
    size_t tev_used[NKIND] = {};
    void* h_pin = nullptr;              // pinned staging of small read-backs
};
constexpr size_t PIN3_BYTES = 1 << 16;

Context3 g3;
std::string g3_err;

int fail3(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g3_err = buf;
    return code;
}

#define HIPCHK3(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail3(AMRGPU_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// all transfers go through the 3-D stream (ordered after its kernels)
hipError_t h2d3(void* dst, const void* src, size_t n) {
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, g3.stream);
    return e != hipSuccess ? e : hipStreamSynchronize(g3.stream);
}
// device -> host, synchronous on the library stream.  Small read-backs (the
// per-level-step Courant numbers, counts, totals) go through a pinned
// staging block: a copy into pageable memory is a slower staged transfer,
// and these sit on the critical path of every level step.
hipError_t d2h3(void* dst, const void* src, size_t n) {
    if (n <= PIN3_BYTES) {
        if (!g3.h_pin) {
            hipError_t e = hipHostMalloc(&g3.h_pin, PIN3_BYTES, hipHostMallocDefault);
            if (e != hipSuccess) return e;
        }
        hipError_t e = hipMemcpyAsync(g3.h_pin, src, n, hipMemcpyDeviceToHost, g3.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g3.stream);
        if (e == hipSuccess) std::memcpy(dst, g3.h_pin, n);
        return e;
    }
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, g3.stream);
    return e != hipSuccess ? e : hipStreamSynchronize(g3.stream);
}

void free_batch() {
    if (g3.stream) hipStreamSynchronize(g3.stream);
    hipFree(g3.d_patches);
    hipFree(g3.d_q);
    hipFree(g3.d_face);
    hipFree(g3.d_q2);
    hipFree(g3.d_tiles);
    g3.d_q2 = nullptr;
    g3.d_tiles = nullptr;
    g3.tiles.clear();
    g3.tile_lds = 0;
    g3.fluxes_valid = false;
    g3.last_fused = false;
    hipFree(g3.d_cfl);
    hipFree(g3.d_flags);
    hipFree(g3.d_foff);
    hipFree(g3.d_nflag);
    hipFree(g3.d_partial);
    g3.d_flags = nullptr;
    g3.d_foff = nullptr;
    g3.d_nflag = nullptr;
    g3.d_partial = nullptr;
    g3.foff.clear();
    for (int k = 0; k < NKIND; k++) {
        hipFree(g3.d_blocks[k]);
        g3.d_blocks[k] = nullptr;
        g3.blocks[k].clear();
    }
    g3.d_patches = nullptr;
    g3.d_q = g3.d_face = nullptr;
    g3.d_cfl = nullptr;
    g3.patches.clear();
    g3.npatch = 0;
    g3.q_elems = g3.face_elems = 0;
}

void free_events() {
    for (int k = 0; k < NKIND; k++) {
        for (auto& pr : g3.tev[k]) {
            hipEventDestroy(pr.first);
            hipEventDestroy(pr.second);
        }
        g3.tev[k].clear();
        g3.tev_used[k] = 0;
    }
}

int need_init() {
    if (!g3.inited) return fail3(AMRGPU_ERR_STATE, "amrgpu3_init has not been called");
    return AMRGPU_OK;
}

int check_patch(int p) {
    int rc = need_init();
    if (rc) return rc;
    if (p < 0 || p >= g3.npatch) return fail3(AMRGPU_ERR_ARG, "patch %d out of range", p);
    return AMRGPU_OK;
}

// event pair for one launch of `kind` (timing on)
int tbegin(int kind, hipEvent_t* stop) {
    if (!g3.timing) {
        *stop = nullptr;
        return AMRGPU_OK;
    }
    auto& v = g3.tev[kind];
    if (g3.tev_used[kind] == v.size()) {
        hipEvent_t a, b;
        HIPCHK3(hipEventCreate(&a));
        HIPCHK3(hipEventCreate(&b));
        v.emplace_back(a, b);
    }
    auto& pr = v[g3.tev_used[kind]++];
    HIPCHK3(hipEventRecord(pr.first, g3.stream));
    *stop = pr.second;
    return AMRGPU_OK;
}
int tend(hipEvent_t stop) {
    if (stop) HIPCHK3(hipEventRecord(stop, g3.stream));
    return AMRGPU_OK;
}

// cost of a sweep work item: planes x rounds of 256 lanes per plane
int64_t block_cost(const Patch3& pt, int d, const Block3& b) {
    const int n = d == 0 ? pt.mx : (d == 1 ? pt.my : pt.mz);
    const int nq = d == 0 ? pt.my : pt.mx;
    const int lines_per = SWEEP3_LANES / (n + 3);
    const int rounds = (nq + 2 + lines_per - 1) / lines_per;
    return (int64_t)(b.pe - b.pb + 2) * rounds;
}

// the batch's work lists: per sweep, (patch, plane block) in decreasing
// cost, so that the longest workgroups start first; update, (patch, 256
// cells)
void plan_blocks_of(const std::vector<Patch3>& patches, int pbk, std::vector<Block3> out[NKIND]) {
    for (int k = 0; k < NKIND; k++) out[k].clear();
    for (int p = 0; p < (int)patches.size(); p++) {
        const Patch3& pt = patches[p];
        const int np[3] = {pt.mz, pt.mz, pt.my};
        for (int d = 0; d < 3; d++)
            for (int b = 1; b <= np[d]; b += pbk)
                out[d].push_back(Block3{p, b, std::min(b + pbk, np[d] + 1), 0});
        const int ncell = pt.mx * pt.my * pt.mz;
        for (int c = 0; c < ncell; c += 256)
            out[3].push_back(Block3{p, c, std::min(c + 256, ncell), 0});
    }
    for (int d = 0; d < 3; d++) {
        auto& v = out[d];
        std::stable_sort(v.begin(), v.end(), [&patches, d](const Block3& a, const Block3& b) {
            return block_cost(patches[a.patch], d, a) > block_cost(patches[b.patch], d, b);
        });
    }
}

void plan_blocks() { plan_blocks_of(g3.patches, g3.plane_block, g3.blocks); }

// n cells as k near-equal parts: (start, length)
void split_tile3(int n, int k, std::vector<std::pair<int, int>>& out) {
    out.clear();
    int s = 0;
    for (int i = 0; i < k; i++) {
        const int len = n / k + (i < n % k ? 1 : 0);
        out.emplace_back(s, len);
        s += len;
    }
}

// lane slots the fused kernel spends on a tile (its schedule: per sweep,
// rows of lines in three classes, four waves of chunks of 64 / (tr + 1)
// lines per round, 256 lanes per round)
int64_t tile3_slots(int tx, int ty, int tz) {
    const int t3[3][3] = {{tx, ty, tz}, {ty, tx, tz}, {tz, tx, ty}};
    int64_t slots = 0;
    for (const auto& t : t3) {
        const int tr = t[0], tq = t[1], tp = t[2];
        const int lpw = 64 / (tr + 1), K = (tq + 2 + lpw - 1) / lpw;
        for (int c3 = 0; c3 < 3; c3++) {
            const int nrows = (tp + 4 - c3) / 3, wpr = std::min(nrows, 4);
            slots += (int64_t)((nrows * K + wpr - 1) / wpr) * 256;
        }
    }
    return slots;
}
size_t tile3_bytes(int nv, int tx, int ty, int tz) {
    return (size_t)nv * ((size_t)(tx + 4) * (ty + 4) * (tz + 4) + (size_t)tx * ty * tz) *
           sizeof(float);
}

// the fused step's tiles, largest first, and the LDS the largest needs.  Each
// grid is cut into kx x ky x kz near-equal tiles of at most TILE3_MAX cells
// per direction whose LDS lets two workgroups
// share a CU; of those cuts, the one with the fewest lane slots.
void plan_tiles3(const std::vector<Patch3>& patches, int nv, std::vector<Tile3>& tiles,
                 size_t& lds, int xcd_start[TILE3_XCDS], int xcd_count[TILE3_XCDS]) {
    tiles.clear();
    lds = 0;
    std::vector<std::vector<Tile3>> per(patches.size());
    std::vector<int64_t> cost(patches.size(), 0);
    const size_t cap = 160 * 1024 / TILE3_WG_PER_CU;
    std::vector<std::pair<int, int>> sx, sy, sz;
    for (int p = 0; p < (int)patches.size(); p++) {
        const Patch3& pt = patches[p];
        const int n[3] = {pt.mx, pt.my, pt.mz};
        int kbest[3] = {0, 0, 0};
        int64_t best = -1;
        int klo[3], khi[3];
        for (int a = 0; a < 3; a++) {
            klo[a] = (n[a] + TILE3_MAX - 1) / TILE3_MAX;
            khi[a] = std::max(klo[a], (n[a] + 3) / 4);
        }
        for (int kx = klo[0]; kx <= khi[0]; kx++)
            for (int ky = klo[1]; ky <= khi[1]; ky++)
                for (int kz = klo[2]; kz <= khi[2]; kz++) {
                    const int k[3] = {kx, ky, kz};
                    // tile extents: n / k and n / k + 1 (the first n % k parts)
                    int ext[3][2], cnt[3][2];
                    for (int a = 0; a < 3; a++) {
                        ext[a][0] = n[a] / k[a];
                        ext[a][1] = ext[a][0] + 1;
                        cnt[a][1] = n[a] % k[a];
                        cnt[a][0] = k[a] - cnt[a][1];
                    }
                    const int mx_ = cnt[0][1] ? ext[0][1] : ext[0][0];
                    const int my_ = cnt[1][1] ? ext[1][1] : ext[1][0];
                    const int mz_ = cnt[2][1] ? ext[2][1] : ext[2][0];
                    if (tile3_bytes(nv, mx_, my_, mz_) > cap) continue;
                    int64_t slots = 0;
                    for (int i = 0; i < 2; i++)
                        for (int j = 0; j < 2; j++)
                            for (int l = 0; l < 2; l++) {
                                const int64_t c = (int64_t)cnt[0][i] * cnt[1][j] * cnt[2][l];
                                if (c) slots += c * tile3_slots(ext[0][i], ext[1][j], ext[2][l]);
                            }
                    if (best < 0 || slots < best) {
                        best = slots;
                        kbest[0] = kx;
                        kbest[1] = ky;
                        kbest[2] = kz;
                    }
                }
        if (best < 0)   // (never: tiles of at most 4 cells per direction always fit)
            for (int a = 0; a < 3; a++) kbest[a] = (n[a] + 3) / 4;
        split_tile3(pt.mx, kbest[0], sx);
        split_tile3(pt.my, kbest[1], sy);
        split_tile3(pt.mz, kbest[2], sz);
        for (auto& z : sz)
            for (auto& y : sy)
                for (auto& x : sx) {
                    Tile3 t{};
                    t.patch = p;
                    t.i0 = (int16_t)x.first;
                    t.j0 = (int16_t)y.first;
                    t.k0 = (int16_t)z.first;
                    t.tx = (int16_t)x.second;
                    t.ty = (int16_t)y.second;
                    t.tz = (int16_t)z.second;
                    per[p].push_back(t);
                    cost[p] += tile3_slots(t.tx, t.ty, t.tz);
                    lds = std::max(lds, tile3_bytes(nv, t.tx, t.ty, t.tz));
                }
    }
    // whole grids to one XCD each (workgroups go to the XCDs in turn), so
    // that neighbouring tiles share their halo rows in that XCD's L2: grids
    // in decreasing cost to the least loaded XCD, then each XCD's tiles in
    // one contiguous range, grid by grid
    std::vector<int> order(patches.size());
    for (int p = 0; p < (int)order.size(); p++) order[p] = p;
    std::stable_sort(order.begin(), order.end(), [&cost](int a, int b) { return cost[a] > cost[b]; });
    std::vector<std::vector<int>> bin(TILE3_XCDS);
    int64_t load[TILE3_XCDS] = {};
    for (int p : order) {
        int x = 0;
        for (int y = 1; y < TILE3_XCDS; y++)
            if (load[y] < load[x]) x = y;
        bin[x].push_back(p);
        load[x] += cost[p];
    }
    for (int x = 0; x < TILE3_XCDS; x++) {
        xcd_start[x] = (int)tiles.size();
        for (int p : bin[x]) tiles.insert(tiles.end(), per[p].begin(), per[p].end());
        xcd_count[x] = (int)tiles.size() - xcd_start[x];
    }
}

int define_batch(int npatch, const int* mx, const int* my, const int* mz) {
    free_batch();
    const int nv = g3.cfg.nvar;
    g3.patches.resize(npatch);
    int64_t qo = 0, fo = 0;
    for (int p = 0; p < npatch; p++) {
        Patch3& pt = g3.patches[p];
        pt.mx = mx[p];
        pt.my = my[p];
        pt.mz = mz[p];
        pt.pad = 0;
        pt.qoff = qo;
        pt.foff = fo;
        // align each grid's arrays to 64 elements
        qo += (q3_elems(pt, nv) + 63) / 64 * 64;
        fo += (faces3_elems(pt, nv) + 63) / 64 * 64;
    }
    g3.npatch = npatch;
    g3.q_elems = qo;
    g3.face_elems = fo;
    plan_blocks();
    HIPCHK3(hipMalloc(&g3.d_patches, std::max<size_t>(1, npatch) * sizeof(Patch3)));
    if (npatch) HIPCHK3(h2d3(g3.d_patches, g3.patches.data(), npatch * sizeof(Patch3)));
    HIPCHK3(hipMalloc(&g3.d_q, std::max<int64_t>(1, qo) * g3.esize));
    HIPCHK3(hipMalloc(&g3.d_face, std::max<int64_t>(1, fo) * g3.esize));
    HIPCHK3(hipMemsetAsync(g3.d_q, 0, std::max<int64_t>(1, qo) * g3.esize, g3.stream));
    HIPCHK3(hipMemsetAsync(g3.d_face, 0, std::max<int64_t>(1, fo) * g3.esize, g3.stream));
    if (g3.esize == 4) {
        // the fused step writes the second buffer
        plan_tiles3(g3.patches, nv, g3.tiles, g3.tile_lds, g3.xcd_start, g3.xcd_count);
        HIPCHK3(hipMalloc(&g3.d_q2, std::max<int64_t>(1, qo) * g3.esize));
        HIPCHK3(hipMemsetAsync(g3.d_q2, 0, std::max<int64_t>(1, qo) * g3.esize, g3.stream));
        HIPCHK3(hipMalloc(&g3.d_tiles, std::max<size_t>(1, g3.tiles.size()) * sizeof(Tile3)));
        if (!g3.tiles.empty())
            HIPCHK3(h2d3(g3.d_tiles, g3.tiles.data(), g3.tiles.size() * sizeof(Tile3)));
    }
    HIPCHK3(hipMalloc(&g3.d_cfl, std::max(1, npatch) * sizeof(unsigned long long)));
    for (int k = 0; k < NKIND; k++) {
        const auto& b = g3.blocks[k];
        HIPCHK3(hipMalloc(&g3.d_blocks[k], std::max<size_t>(1, b.size()) * sizeof(Block3)));
        if (!b.empty()) HIPCHK3(h2d3(g3.d_blocks[k], b.data(), b.size() * sizeof(Block3)));
    }
    g3.h_cfl.assign(npatch, 0);
    HIPCHK3(hipStreamSynchronize(g3.stream));
    return AMRGPU_OK;
}

// host AoS q(nvar, mx+4, my+4, mz+4) <-> device SoA [m][k][j][i]
int put_q(int p, const double* q) {
    const Patch3& pt = g3.patches[p];
    const int nv = g3.cfg.nvar;
    const int64_t cells = (int64_t)(pt.mx + 4) * (pt.my + 4) * (pt.mz + 4);
    if (g3.esize == 8) {
        std::vector<double> h(nv * cells);
        for (int64_t c = 0; c < cells; c++)
            for (int m = 0; m < nv; m++) h[m * cells + c] = q[c * nv + m];
        HIPCHK3(h2d3(static_cast<double*>(g3.d_q) + pt.qoff, h.data(), h.size() * 8));
    } else {
        std::vector<float> h(nv * cells);
        for (int64_t c = 0; c < cells; c++)
            for (int m = 0; m < nv; m++) h[m * cells + c] = (float)q[c * nv + m];
        HIPCHK3(h2d3(static_cast<float*>(g3.d_q) + pt.qoff, h.data(), h.size() * 4));
        if (g3.d_q2)   // the fused step's output buffer keeps the same ghost cells
            HIPCHK3(h2d3(static_cast<float*>(g3.d_q2) + pt.qoff, h.data(), h.size() * 4));
    }
    return AMRGPU_OK;
}

int get_q(int p, double* q) {
    const Patch3& pt = g3.patches[p];
    const int nv = g3.cfg.nvar;
    const int64_t cells = (int64_t)(pt.mx + 4) * (pt.my + 4) * (pt.mz + 4);
    if (g3.esize == 8) {
        std::vector<double> h(nv * cells);
        HIPCHK3(d2h3(h.data(), static_cast<double*>(g3.d_q) + pt.qoff, h.size() * 8));
        for (int64_t c = 0; c < cells; c++)
            for (int m = 0; m < nv; m++) q[c * nv + m] = h[m * cells + c];
    } else {
        std::vector<float> h(nv * cells);
        HIPCHK3(d2h3(h.data(), static_cast<float*>(g3.d_q) + pt.qoff, h.size() * 4));
        for (int64_t c = 0; c < cells; c++)
            for (int m = 0; m < nv; m++) q[c * nv + m] = (double)h[m * cells + c];
    }
    return AMRGPU_OK;
}

int step_io(const Step3IO& io, double dx, double dy, double dz, double dt, int es_override = 0) {
    const size_t es = es_override ? (size_t)es_override : g3.esize;
    if (!(dx > 0) || !(dy > 0) || !(dz > 0) || !(dt >= 0))
        return fail3(AMRGPU_ERR_ARG, "bad mesh widths or time step");
    HIPCHK3(hipMemsetAsync(io.d_cfl, 0, std::max(1, io.npatch) * sizeof(unsigned long long),
                           g3.stream));
    Sweep3Args a{};
    a.q = io.q;
    a.face = io.face;
    a.patches = io.d_patches;
    a.dtd[0] = dt / dx;
    a.dtd[1] = dt / dy;
    a.dtd[2] = dt / dz;
    a.meth = g3.meth;
    a.cfl_patch = io.d_cfl;
    const int solver = g3.cfg.solver;
    for (int d = 1; d <= 3; d++) {
        a.phase = d;
        a.blocks = io.d_blocks[d - 1];
        a.nblocks = io.nblocks[d - 1];
        if (a.nblocks == 0) continue;
        hipEvent_t stop;
        int rc = tbegin(d - 1, &stop);
        if (rc) return rc;
        HIPCHK3(es == 8 ? launch_sweep3_f64(solver, a, g3.stream)
                        : launch_sweep3_f32(solver, a, g3.stream));
        rc = tend(stop);
        if (rc) return rc;
    }
    const int nb = io.nblocks[3];
    if (nb > 0) {
        hipEvent_t stop;
        int rc = tbegin(3, &stop);
        if (rc) return rc;
        if (es == 8)
            HIPCHK3(launch_update3_f64(g3.cfg.nvar, static_cast<double*>(io.q),
                                       static_cast<const double*>(io.face), io.d_patches,
                                       io.d_blocks[3], nb, a.dtd[0], a.dtd[1], a.dtd[2],
                                       g3.stream));
        else
            HIPCHK3(launch_update3_f32(g3.cfg.nvar, static_cast<float*>(io.q),
                                       static_cast<const float*>(io.face), io.d_patches,
                                       io.d_blocks[3], nb, a.dtd[0], a.dtd[1], a.dtd[2],
                                       g3.stream));
        rc = tend(stop);
        if (rc) return rc;
    }
    return AMRGPU_OK;
}

// the fused FP32 step of the batch: d_q -> d_q2 tile by tile, then swap
int step_fused(double dx, double dy, double dz, double dt) {
    if (!(dx > 0) || !(dy > 0) || !(dz > 0) || !(dt >= 0))
        return fail3(AMRGPU_ERR_ARG, "bad mesh widths or time step");
    if (g3.tile_lds > 160 * 1024)
        return fail3(AMRGPU_ERR_UNSUPPORTED, "tile LDS %zu B exceeds a CU's", g3.tile_lds);
    HIPCHK3(hipMemsetAsync(g3.d_cfl, 0, std::max(1, g3.npatch) * sizeof(unsigned long long),
                           g3.stream));
    Tile3Args a{};
    a.qin = g3.d_q;
    a.qout = g3.d_q2;
    a.patches = g3.d_patches;
    a.tiles = g3.d_tiles;
    a.ntiles = (int)g3.tiles.size();
    for (int x = 0; x < TILE3_XCDS; x++) {
        a.xcd_start[x] = g3.xcd_start[x];
        a.xcd_count[x] = g3.xcd_count[x];
    }
    a.nv = g3.cfg.nvar;
    a.dtd[0] = dt / dx;
    a.dtd[1] = dt / dy;
    a.dtd[2] = dt / dz;
    a.meth = g3.meth;
    a.cfl_patch = g3.d_cfl;
    hipEvent_t stop;
    int rc = tbegin(0, &stop);
    if (rc) return rc;
    HIPCHK3(launch_sweep3_tile_f32(g3.cfg.solver, a, g3.tile_lds, g3.stream));
    rc = tend(stop);
    if (rc) return rc;
    std::swap(g3.d_q, g3.d_q2);
    g3.fluxes_valid = false;
    g3.last_fused = true;
    return AMRGPU_OK;
}

int step_batch(double dx, double dy, double dz, double dt, double* cflgrid, bool fused) {
    if (fused) {
        int rc = step_fused(dx, dy, dz, dt);
        if (rc) return rc;
        if (cflgrid) {
            HIPCHK3(d2h3(g3.h_cfl.data(), g3.d_cfl, g3.npatch * sizeof(unsigned long long)));
            for (int p = 0; p < g3.npatch; p++) {
                double v;
                std::memcpy(&v, &g3.h_cfl[p], sizeof v);
                cflgrid[p] = v;
            }
        }
        return AMRGPU_OK;
    }
    Step3IO io{};
    io.q = g3.d_q;
    io.face = g3.d_face;
    io.d_patches = g3.d_patches;
    for (int k = 0; k < NKIND; k++) {
        io.d_blocks[k] = g3.d_blocks[k];
        io.nblocks[k] = (int)g3.blocks[k].size();
    }
    io.npatch = g3.npatch;
    io.d_cfl = g3.d_cfl;
    int rc = step_io(io, dx, dy, dz, dt);
    if (rc) return rc;
    g3.fluxes_valid = true;
    g3.last_fused = false;
    if (cflgrid) {
        HIPCHK3(d2h3(g3.h_cfl.data(), g3.d_cfl, g3.npatch * sizeof(unsigned long long)));
        for (int p = 0; p < g3.npatch; p++) {
            double v;
            std::memcpy(&v, &g3.h_cfl[p], sizeof v);
            cflgrid[p] = v;
        }
    }
    return AMRGPU_OK;
}

int get_fluxes(int p, double* const out[6]) {
    if (!g3.fluxes_valid)
        return g3.last_fused
                   ? fail3(AMRGPU_ERR_STATE, "no flux arrays: the last step was the fused one, which "
                                             "keeps every face on chip (amrgpu3_set_step_mode(0))")
                   : fail3(AMRGPU_ERR_STATE, "no flux arrays: no step has run since the patches "
                                             "were set (amrgpu3_step_patches)");
    const Patch3& pt = g3.patches[p];
    const int nv = g3.cfg.nvar;
    const Faces3 F = faces3(pt, nv);
    const int64_t n = faces3_elems(pt, nv);
    std::vector<double> h(n);
    if (g3.esize == 8) {
        HIPCHK3(d2h3(h.data(), static_cast<double*>(g3.d_face) + pt.foff, n * 8));
    } else {
        std::vector<float> f(n);
        HIPCHK3(d2h3(f.data(), static_cast<float*>(g3.d_face) + pt.foff, n * 4));
        for (int64_t k = 0; k < n; k++) h[k] = f[k];
        // the FP32 step keeps fp = fm + (faddp - faddm), gp likewise
        // (sweep3.hpp, merged form): rebuild fp and gp as the update uses them
        for (int64_t k = 0; k < nv * F.sx; k++)
            h[F.fp - pt.foff + k] = (double)(f[F.fm - pt.foff + k] + f[F.fp - pt.foff + k]);
        for (int64_t k = 0; k < nv * F.sy; k++)
            h[F.gp - pt.foff + k] = (double)(f[F.gm - pt.foff + k] + f[F.gp - pt.foff + k]);
    }
    const int mit = pt.mx + 4, mjt = pt.my + 4;
    auto at = [&](int i, int j, int k) { return (((int64_t)(k + 1) * mjt + (j + 1)) * mit + (i + 1)) * nv; };
    const int64_t base[6] = {F.fm, F.fp, F.gm, F.gp, F.hm, F.hp};
    const int64_t cs[3] = {F.sx, F.sy, F.sz};
    for (int a = 0; a < 6; a++) {
        if (!out[a]) continue;
        const int dir = a / 2;
        const int ni = pt.mx + (dir == 0), nj = pt.my + (dir == 1), nk = pt.mz + (dir == 2);
        const double* src = h.data() + (base[a] - pt.foff);
        for (int k = 1; k <= nk; k++)
            for (int j = 1; j <= nj; j++)
                for (int i = 1; i <= ni; i++) {
                    const int64_t f = ((int64_t)(k - 1) * nj + (j - 1)) * ni + (i - 1);
                    for (int m = 0; m < nv; m++) out[a][at(i, j, k) + m] = src[m * cs[dir] + f];
                }
    }
    return AMRGPU_OK;
}

// flag and conck buffers, allocated on first use for the current batch
int ensure_flag_buffers() {
    if (g3.d_flags) return AMRGPU_OK;
    g3.foff.assign(g3.npatch + 1, 0);
    for (int p = 0; p < g3.npatch; p++) {
        const Patch3& pt = g3.patches[p];
        g3.foff[p + 1] = g3.foff[p] + (int64_t)(pt.mx + 4) * (pt.my + 4) * (pt.mz + 4);
    }
    const int64_t n = std::max<int64_t>(1, g3.foff[g3.npatch]);
    HIPCHK3(hipMalloc(&g3.d_flags, n * sizeof(double)));
    HIPCHK3(hipMemsetAsync(g3.d_flags, 0, n * sizeof(double), g3.stream));   // goodpt = 0
    HIPCHK3(hipMalloc(&g3.d_foff, (g3.npatch + 1) * sizeof(int64_t)));
    HIPCHK3(h2d3(g3.d_foff, g3.foff.data(), (g3.npatch + 1) * sizeof(int64_t)));
    HIPCHK3(hipMalloc(&g3.d_nflag, sizeof(unsigned long long)));
    HIPCHK3(hipMalloc(&g3.d_partial, std::max<size_t>(1, g3.blocks[3].size()) * sizeof(double)));
    return AMRGPU_OK;
}

}  // namespace

namespace amrgpu {
bool ctx3_inited() { return g3.inited; }
const amrgpu3_config& ctx3_cfg() { return g3.cfg; }
size_t ctx3_esize() { return g3.esize; }
hipStream_t ctx3_stream() { return g3.stream; }
int ctx3_plane_block() { return g3.plane_block; }
int ctx3_fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g3_err = buf;
    return code;
}
hipError_t ctx3_h2d(void* dst, const void* src, size_t n) { return h2d3(dst, src, n); }
hipError_t ctx3_d2h(void* dst, const void* src, size_t n) { return d2h3(dst, src, n); }
int ctx3_step(const Step3IO& io, double dx, double dy, double dz, double dt, int esize) {
    return step_io(io, dx, dy, dz, dt, esize);
}
void ctx3_plan_blocks(const std::vector<Patch3>& patches, int plane_block,
                      std::vector<Block3> out[4]) {
    plan_blocks_of(patches, plane_block, out);
}
}  // namespace amrgpu

extern "C" {

int amrgpu3_flag_patches(double tolsp, int64_t* nflagged) {
    int rc = need_init();
    if (rc) return rc;
    rc = ensure_flag_buffers();
    if (rc) return rc;
    HIPCHK3(hipMemsetAsync(g3.d_nflag, 0, sizeof(unsigned long long), g3.stream));
    HIPCHK3(launch_flag3((int)g3.esize, g3.cfg.nvar, g3.d_q, g3.d_flags, g3.d_foff, g3.d_patches,
                         g3.d_blocks[3], (int)g3.blocks[3].size(), tolsp, g3.d_nflag, g3.stream));
    unsigned long long n = 0;
    HIPCHK3(d2h3(&n, g3.d_nflag, sizeof n));
    if (nflagged) *nflagged = (int64_t)n;
    return AMRGPU_OK;
}

int amrgpu3_get_flags(int patch, double* amrflags) {
    int rc = check_patch(patch);
    if (rc) return rc;
    if (!amrflags) return fail3(AMRGPU_ERR_ARG, "null flags");
    if (!g3.d_flags) return fail3(AMRGPU_ERR_STATE, "amrgpu3_flag_patches has not been called");
    HIPCHK3(d2h3(amrflags, g3.d_flags + g3.foff[patch],
                 (g3.foff[patch + 1] - g3.foff[patch]) * sizeof(double)));
    return AMRGPU_OK;
}

int amrgpu3_conck(double hx, double hy, double hz, double* totmass) {
    int rc = need_init();
    if (rc) return rc;
    if (!totmass) return fail3(AMRGPU_ERR_ARG, "null totmass");
    rc = ensure_flag_buffers();
    if (rc) return rc;
    const int np = (int)g3.patches.size();
    double s = 0.0;   // the sequential sum over the batch (conck.f (3-D):47-60)
    if (np) {
        HIPCHK3(launch_conck3((int)g3.esize, g3.d_q, g3.d_patches, np, g3.d_partial, g3.stream));
        HIPCHK3(d2h3(&s, g3.d_partial, sizeof(double)));
    }
    *totmass = s * hx * hy * hz;
    return AMRGPU_OK;
}

int amrgpu3_last_error(char* buf, int len) {
    if (!buf || len <= 0) return AMRGPU_ERR_ARG;
    std::snprintf(buf, len, "%s", g3_err.c_str());
    return AMRGPU_OK;
}

int amrgpu3_init(const amrgpu3_config* cfg) {
    if (!cfg) return fail3(AMRGPU_ERR_ARG, "null config");
    const amrgpu3_config& c = *cfg;
    if (c.solver == AMRGPU_ADVECTION) {
        if (c.nvar != 1 || c.mwaves != 1)
            return fail3(AMRGPU_ERR_ARG, "3-D advection has nvar = mwaves = 1");
    } else if (c.solver == AMRGPU_EULER) {
        if (c.nvar != 5 || c.mwaves != 3)
            return fail3(AMRGPU_ERR_ARG, "3-D Euler has nvar = 5, mwaves = 3");
        if (!(c.gamma > 1.0)) return fail3(AMRGPU_ERR_ARG, "gamma must exceed 1");
    } else {
        return fail3(AMRGPU_ERR_UNSUPPORTED, "3-D solver %d not supported", c.solver);
    }
    if (c.naux != 0) return fail3(AMRGPU_ERR_UNSUPPORTED, "aux arrays are not supported");
    if (c.mcapa != 0) return fail3(AMRGPU_ERR_UNSUPPORTED, "capacity is not supported");
    if (c.nghost != 2) return fail3(AMRGPU_ERR_UNSUPPORTED, "nghost must be 2");
    if (c.precision != AMRGPU_FP64 && c.precision != AMRGPU_FP32)
        return fail3(AMRGPU_ERR_ARG, "bad precision");
    const int mt = c.method[2];
    int m3, m4;
    if (mt < 0) {
        m3 = -1;
        m4 = 0;
    } else {
        m3 = mt / 10;
        m4 = mt - 10 * m3;
    }
    if (mt >= 0 && (m3 > 2 || m4 > 2))
        return fail3(AMRGPU_ERR_ARG, "method(3) = %d is not 10 m3 + m4 with m3, m4 <= 2", mt);
    // flux3.f:390-405 reads the cqxx splits, which only m3 = 2 computes
    if (m4 == 2 && m3 == 1)
        return fail3(AMRGPU_ERR_UNSUPPORTED, "method(3) = %d: m4 = 2 needs m3 = 2", mt);
    for (int w = 0; w < c.mwaves; w++)
        if (c.mthlim[w] < 0 || c.mthlim[w] > 5)
            return fail3(AMRGPU_ERR_ARG, "limiter %d not supported", c.mthlim[w]);
    if (g3.inited) amrgpu3_finalize();
    g3.cfg = c;
    g3.esize = c.precision == AMRGPU_FP64 ? 8 : 4;
    g3.step_mode = g3.esize == 4 ? 1 : 0;   // FP32: the fused step by default
    Method3 M{};
    M.ph.ubar = c.ubar;
    M.ph.vbar = c.vbar;
    M.ph.wbar = c.wbar;
    M.ph.gamma = c.gamma;
    M.ph.efix = c.efix;
    M.order = c.method[1];
    M.m3 = m3;
    M.m4 = m4;
    M.use_fwaves = c.use_fwaves;
    for (int w = 0; w < NWMAX3; w++) M.mthlim[w] = w < c.mwaves ? c.mthlim[w] : 0;
    g3.meth = M;
    HIPCHK3(hipStreamCreateWithFlags(&g3.stream, hipStreamNonBlocking));
    g3.inited = true;
    return AMRGPU_OK;
}

int amrgpu3_finalize(void) {
    if (!g3.inited) return AMRGPU_OK;
    amr3_release();
    free_batch();
    free_events();
    if (g3.stream) hipStreamDestroy(g3.stream);
    g3.stream = nullptr;
    if (g3.h_pin) hipHostFree(g3.h_pin);
    g3.h_pin = nullptr;
    g3.inited = false;
    g3.timing = false;
    return AMRGPU_OK;
}

int amrgpu3_set_patches(int npatch, const int* mx, const int* my, const int* mz, int plane_block) {
    int rc = need_init();
    if (rc) return rc;
    if (npatch < 0 || (npatch > 0 && (!mx || !my || !mz)))
        return fail3(AMRGPU_ERR_ARG, "bad patch arrays");
    if (plane_block < 0) return fail3(AMRGPU_ERR_ARG, "plane_block must be >= 0");
    for (int p = 0; p < npatch; p++) {
        if (mx[p] < 1 || my[p] < 1 || mz[p] < 1)
            return fail3(AMRGPU_ERR_ARG, "patch %d: sizes must be positive", p);
        // one round must hold a whole line: n + 3 interfaces <= 256 lanes
        if (std::max(mx[p], std::max(my[p], mz[p])) + 3 > SWEEP3_LANES)
            return fail3(AMRGPU_ERR_UNSUPPORTED, "patch %d: a side exceeds %d cells", p,
                         SWEEP3_LANES - 3);
    }
    g3.plane_block = plane_block ? plane_block : 4;
    return define_batch(npatch, mx, my, mz);
}

int amrgpu3_put_patch(int patch, const double* q) {
    int rc = check_patch(patch);
    if (rc) return rc;
    if (!q) return fail3(AMRGPU_ERR_ARG, "null q");
    return put_q(patch, q);
}

int amrgpu3_get_patch(int patch, double* q) {
    int rc = check_patch(patch);
    if (rc) return rc;
    if (!q) return fail3(AMRGPU_ERR_ARG, "null q");
    return get_q(patch, q);
}

int amrgpu3_step_patches(double dx, double dy, double dz, double dt, double* cflgrid) {
    int rc = need_init();
    if (rc) return rc;
    return step_batch(dx, dy, dz, dt, cflgrid, g3.esize == 4 && g3.step_mode == 1);
}

int amrgpu3_set_step_mode(int mode) {
    int rc = need_init();
    if (rc) return rc;
    if (mode != 0 && mode != 1) return fail3(AMRGPU_ERR_ARG, "step mode must be 0 or 1");
    if (mode == 1 && g3.esize != 4)
        return fail3(AMRGPU_ERR_UNSUPPORTED,
                     "the fused step runs in FP32; FP64 keeps the reference's face order");
    g3.step_mode = mode;
    return AMRGPU_OK;
}

int amrgpu3_get_fluxes(int patch, double* fm, double* fp, double* gm, double* gp, double* hm,
                       double* hp) {
    int rc = check_patch(patch);
    if (rc) return rc;
    double* const out[6] = {fm, fp, gm, gp, hm, hp};
    return get_fluxes(patch, out);
}

int amrgpu3_stepgrid(int mx, int my, int mz, const double* q_in, double* q_out, double dx,
                     double dy, double dz, double dt, double* fm, double* fp, double* gm,
                     double* gp, double* hm, double* hp, double* cflgrid, double* dtnew) {
    int rc = need_init();
    if (rc) return rc;
    if (!q_in || !q_out) return fail3(AMRGPU_ERR_ARG, "null q");
    rc = amrgpu3_set_patches(1, &mx, &my, &mz, 0);
    if (rc) return rc;
    rc = put_q(0, q_in);
    if (rc) return rc;
    double cfl = 0.0;
    // flux outputs come from the phase kernels' arrays
    const bool want = fm && fp && gm && gp && hm && hp;
    rc = step_batch(dx, dy, dz, dt, &cfl, g3.esize == 4 && g3.step_mode == 1 && !want);
    if (rc) return rc;
    rc = get_q(0, q_out);
    if (rc) return rc;
    if (fm && fp && gm && gp && hm && hp) {
        double* const out[6] = {fm, fp, gm, gp, hm, hp};
        rc = get_fluxes(0, out);
        if (rc) return rc;
    }
    if (cflgrid) *cflgrid = cfl;
    if (dtnew) *dtnew = cfl > 0.0 ? dt * g3.cfg.cfl / cfl : RINFINITY3;
    return AMRGPU_OK;
}

int amrgpu3_set_timing(int on) {
    int rc = need_init();
    if (rc) return rc;
    HIPCHK3(hipStreamSynchronize(g3.stream));
    for (int k = 0; k < NKIND; k++) g3.tev_used[k] = 0;
    g3.timing = on != 0;
    return AMRGPU_OK;
}

int amrgpu3_kernel_time(int kind, double* ms, int64_t* launches) {
    int rc = need_init();
    if (rc) return rc;
    if (kind < 0 || kind >= NKIND || !ms) return fail3(AMRGPU_ERR_ARG, "bad kernel kind");
    HIPCHK3(hipStreamSynchronize(g3.stream));
    double t = 0.0;
    for (size_t k = 0; k < g3.tev_used[kind]; k++) {
        float x = 0.f;
        HIPCHK3(hipEventElapsedTime(&x, g3.tev[kind][k].first, g3.tev[kind][k].second));
        t += x;
    }
    *ms = t;
    if (launches) *launches = (int64_t)g3.tev_used[kind];
    return AMRGPU_OK;
}

int amrgpu3_synchronize(void) {
    int rc = need_init();
    if (rc) return rc;
    HIPCHK3(hipStreamSynchronize(g3.stream));
    return AMRGPU_OK;
}

}  // extern "C"
