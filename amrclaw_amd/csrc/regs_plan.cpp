// regs_plan.cpp -- host side of the coarse-fine flux registers.
//
// prepf.f:18-53 (register layout of every fine grid), prepc.f:11-157 with
// setuse.f:61-213 (the listbc segments of every coarse grid, including the
// periodic images), and from them the device programs of fluxsv.f (slot
// writes), upbnd.f (per coarse cell, in list order, with the double-update
// guard of upbnd.f:79-82), update.f's averaging (update.f:95-177) and qad.f
// (one work item per register slot).
#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "regs_plan.hpp"

namespace amrgpu {

namespace {

void setuse(std::vector<SvEntry>& out, int mkid, int ilo, int ihi, int jlo, int jhi, int iclo,
            int ichi, int jclo, int jchi) {
    const int ist = iclo - 1, iend = ichi + 1, jst = jclo - 1, jend = jchi + 1;
    // (i, j) below are the coarse grid's interior 1-based indices
    if (!(ist < ilo)) {
        int lkid = std::max(jlo, jclo) - jclo + 1;
        for (int j = std::max(jlo, jclo); j <= std::min(jhi, jchi); j++)
            out.push_back({ist - ilo + 1, j - jlo + 1, 3, mkid, lkid++, 0});
    }
    if (!(jend > jhi)) {
        int lkid = (jchi - jclo + 1) + std::max(ilo, iclo) - iclo + 1;
        for (int i = std::max(ilo, iclo); i <= std::min(ihi, ichi); i++)
            out.push_back({i - ilo + 1, jend - jlo + 1, 4, mkid, lkid++, 0});
    }
    if (!(iend > ihi)) {
        int lkid = (ichi - iclo + 1) + (jchi - jclo + 1) + std::max(jlo, jclo) - jclo + 1;
        for (int j = std::max(jlo, jclo); j <= std::min(jhi, jchi); j++)
            out.push_back({iend - ilo + 1, j - jlo + 1, 1, mkid, lkid++, 0});
    }
    if (!(jst < jlo)) {
        int lkid = 2 * (jchi - jclo + 1) + (ichi - iclo + 1) + std::max(ilo, iclo) - iclo + 1;
        for (int i = std::max(ilo, iclo); i <= std::min(ihi, ichi); i++)
            out.push_back({i - ilo + 1, jst - jlo + 1, 2, mkid, lkid++, 0});
    }
}

}  // namespace

bool plan_registers(const HierGeom& H, int nvar, RegLevel* R, std::string& err) {
    for (int l = 0; l < FILL_MAXLV; l++) R[l] = RegLevel();
    for (int l = 2; l <= H.nlev; l++) {
        const LevelGeom& F = H.lev[l];
        const LevelGeom& C = H.lev[l - 1];
        RegLevel& Rl = R[l];
        for (size_t k = 0; k < F.boxes.size(); k++) {
            const auto& b = F.boxes[k];
            int nx = b[1] - b[0] + 1, ny = b[3] - b[2] + 1;
            if (nx % C.ratx || ny % C.raty || b[0] % C.ratx || b[2] % C.raty) {
                err = "fine grid not aligned with the coarse cells";
                return false;
            }
            if (ny == 1) {
                err = "single-row fine grids are not supported";
                return false;
            }
            int lenbc = 2 * (nx / C.ratx + ny / C.raty);
            Rl.reg_off.push_back(Rl.reg_elems);
            Rl.lenbc.push_back(lenbc);
            // svdflx(nvar, lenbc), qc1d(nvar, lenbc), perimeter fluxes and,
            // with aux arrays, auxc1d(naux, lenbc) (prepf.f:33-48)
            Rl.reg_elems += 2 * (int64_t)nvar * lenbc + 2 * (int64_t)nvar * (nx + ny) +
                            (int64_t)H.naux * lenbc;
            Rl.reg_elems = (Rl.reg_elems + 31) / 32 * 32;
            const int nyc = ny / C.raty, nxc = nx / C.ratx;
            for (int s = 1; s <= nyc; s++) Rl.qad.push_back({(int)k, 1, s, s});
            for (int s = 1; s <= nxc; s++) Rl.qad.push_back({(int)k, 2, s, nyc + s});
            for (int s = 1; s <= nyc; s++) Rl.qad.push_back({(int)k, 3, s, nyc + nxc + s});
            for (int s = 1; s <= nxc; s++) Rl.qad.push_back({(int)k, 4, s, 2 * nyc + nxc + s});
        }
    }
    for (int l = 1; l < H.nlev; l++) {
        const LevelGeom& C = H.lev[l];
        const LevelGeom& F = H.lev[l + 1];
        if (F.boxes.empty()) continue;
        RegLevel& Rl = R[l];
        const int rx = C.ratx, ry = C.raty;
        const int imax = C.iregsz - 1, jmax = C.jregsz - 1;
        Rl.sv_grid_start.push_back(0);
        const bool given = !C.listbc.empty();
        if (given && C.listbc.size() != C.boxes.size()) {
            err = "listbc given for some grids of a level only";
            return false;
        }
        for (size_t m = 0; m < C.boxes.size(); m++) {
            const auto& cb = C.boxes[m];
            const int ilo = cb[0], ihi = cb[1], jlo = cb[2], jhi = cb[3];
            if (given) {
                // the caller's list, in its order: upbnd adds a coarse cell's
                // entries in list order (upbnd.f:59-113), so the order of the
                // kids prepc visited fixes the FP64 sums
                for (const auto& e : C.listbc[m]) {
                    const int I = e[0] - NGHOST, J = e[1] - NGHOST, side = e[2], kid = e[3];
                    if (I < 1 || I > ihi - ilo + 1 || J < 1 || J > jhi - jlo + 1 || side < 1 ||
                        side > 4 || kid < 0 || kid >= (int)F.boxes.size() || e[4] < 1 ||
                        e[4] > R[l + 1].lenbc[kid]) {
                        err = "listbc entry outside its grids";
                        return false;
                    }
                    Rl.sv.push_back({I, J, side, kid, e[4], 0});
                }
                Rl.sv_grid_start.push_back((int32_t)Rl.sv.size());
                continue;
            }
            for (size_t kid = 0; kid < F.boxes.size(); kid++) {
                const auto& fb = F.boxes[kid];
                int iclo = fb[0] / rx, jclo = fb[2] / ry, ichi = fb[1] / rx, jchi = fb[3] / ry;
                int iplo = std::max(ilo, iclo), jplo = std::max(jlo, jclo);
                int iphi = std::min(ihi, ichi), jphi = std::min(jhi, jchi);
                if (iplo <= iphi + 1 && jplo <= jphi + 1)
                    setuse(Rl.sv, (int)kid, ilo, ihi, jlo, jhi, iclo, ichi, jclo, jchi);
                if (H.xperdom && ilo == 0 && ichi == imax)
                    setuse(Rl.sv, (int)kid, ilo, ihi, jlo, jhi, iclo - C.iregsz, ichi - C.iregsz, jclo, jchi);
                if (H.xperdom && iclo == 0 && ihi == imax)
                    setuse(Rl.sv, (int)kid, ilo, ihi, jlo, jhi, iclo + C.iregsz, ichi + C.iregsz, jclo, jchi);
                if (H.yperdom && jlo == 0 && jchi == jmax)
                    setuse(Rl.sv, (int)kid, ilo, ihi, jlo, jhi, iclo, ichi, jclo - C.jregsz, jchi - C.jregsz);
                if (H.yperdom && jclo == 0 && jhi == jmax)
                    setuse(Rl.sv, (int)kid, ilo, ihi, jlo, jhi, iclo, ichi, jclo + C.jregsz, jchi + C.jregsz);
            }
            Rl.sv_grid_start.push_back((int32_t)Rl.sv.size());
        }
        // fluxsv: a slot reached twice keeps the last write in grid order
        Rl.sv_write.assign(Rl.sv.size(), 1);
        std::map<std::pair<int, int>, size_t> last;
        for (size_t e = 0; e < Rl.sv.size(); e++) {
            auto key = std::make_pair(Rl.sv[e].kid, Rl.sv[e].lkid);
            auto it = last.find(key);
            if (it != last.end()) Rl.sv_write[it->second] = 0;
            last[key] = e;
        }
        // upbnd programs: per coarse cell, its entries in list order
        for (size_t m = 0; m < C.boxes.size(); m++) {
            const int mitot = C.boxes[m][1] - C.boxes[m][0] + 1 + 2 * NGHOST;
            std::map<int, std::vector<UpbndEntry>> cells;   // key: cell index
            std::map<int, int> iused;
            std::vector<int> order;
            for (int e = Rl.sv_grid_start[m]; e < Rl.sv_grid_start[m + 1]; e++) {
                const SvEntry& s = Rl.sv[e];
                int ci = s.I + NGHOST - 1, cj = s.J + NGHOST - 1;     // 0-based with ghosts
                int key = cj * mitot + ci;
                int norm = 1 << (s.side - 1);
                int& u = iused[key];
                if ((u / norm) % 2 == 1) continue;
                u += norm;
                double sgnm = (s.side == 2 || s.side == 3) ? -1.0 : 1.0;
                if (!cells.count(key)) order.push_back(key);
                cells[key].push_back({s.kid, s.lkid, sgnm});
            }
            for (int key : order) {
                UpbndCell c;
                c.grid = (int)m;
                c.i = key % mitot;
                c.j = key / mitot;
                c.e0 = (int32_t)Rl.uents.size();
                for (const auto& en : cells[key]) Rl.uents.push_back(en);
                c.e1 = (int32_t)Rl.uents.size();
                Rl.ucells.push_back(c);
            }
        }
        // averaging blocks
        Rl.avg_start.push_back(0);
        for (size_t m = 0; m < C.boxes.size(); m++) {
            const auto& cb = C.boxes[m];
            for (size_t kid = 0; kid < F.boxes.size(); kid++) {
                const auto& fb = F.boxes[kid];
                int iclo = fb[0] / rx, jclo = fb[2] / ry, ichi = fb[1] / rx, jchi = fb[3] / ry;
                int iplo = std::max(cb[0], iclo), jplo = std::max(cb[2], jclo);
                int iphi = std::min(cb[1], ichi), jphi = std::min(cb[3], jchi);
                if (iplo > iphi || jplo > jphi) continue;
                AvgBlock a;
                a.grid = (int)m;
                a.kid = (int)kid;
                a.ci0 = iplo - cb[0] + NGHOST;
                a.cj0 = jplo - cb[2] + NGHOST;
                a.ni = iphi - iplo + 1;
                a.nj = jphi - jplo + 1;
                a.fi0 = iplo * rx - fb[0] + NGHOST;
                a.fj0 = jplo * ry - fb[2] + NGHOST;
                Rl.avg.push_back(a);
                Rl.avg_start.push_back(Rl.avg_start.back() + (int64_t)a.ni * a.nj);
            }
        }
    }
    return true;
}

}  // namespace amrgpu
