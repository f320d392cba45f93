// kernels.hpp -- host-side launchers of the device kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "flag.hpp"
#include "regs.hpp"
#include "sweep.hpp"

namespace amrgpu {
hipError_t launch_sweep_f64(int solver, const SweepArgs& a, size_t lds, hipStream_t st);
hipError_t launch_sweep_f32(int solver, const SweepArgs& a, size_t lds, hipStream_t st);
// wgs: the occupancy bound of the mixed sweep, 5, 6 or 7 workgroups per CU (0: 7)
hipError_t launch_sweep_mixed(int solver, const SweepArgs& a, size_t lds, hipStream_t st, int wgs);
hipError_t launch_finereg(int solver, int esize, const QadArgs& a, const Physics& ph, hipStream_t st);
hipError_t launch_update(int esize, const UpdateArgs& a, const int64_t* d_block_start,
                         int64_t navg_cells, hipStream_t st);
hipError_t launch_flag(int esize, const FlagArgs& a, int ngrids, hipStream_t st);
hipError_t launch_errf1(const ErrfArgs& a, int ngrids, hipStream_t st);
// conck.f: the level's sum of q(1) over the interiors, in the reference's
// sequential order, into out[0]
hipError_t launch_conck(int esize, const PatchDesc* patches, int npatch, const void* q,
                        double* out, hipStream_t st, const double* aux = nullptr, int mcapa = 0,
                        int auxk = 0);
}  // namespace amrgpu
