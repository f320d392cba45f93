// amrgpu_solver.hpp -- the Riemann-solver plug-in contract of libamrgpu
// (HIP C++, compiled by the application with hipcc for gfx950).
//
// The reference's solvers are Fortran externals with the interface
//   rpn2(ixy,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,auxl,auxr,wave,s,amdq,apdq)
//   rpt2(ixy,imp,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,aux1,aux2,aux3,
//        asdq,bmasdq,bpasdq)
// called once per 1-D slice (src/2d/flux2.f:125,202,219) and by qad.f:129.
// Here a solver is a class template Sol<T, Mth> of __device__ functions
// called once per interface, with the same argument meaning:
//
//   template <typename T, class Mth = amrgpu::ExactMath<T>> struct MySolver {
//     template <typename U> using Rebind = MySolver<U, amrgpu::ExactMath<U>>;
//     static constexpr int NV = meqn, NW = mwaves;
//     using Math = Mth;   // Mth::div, Mth::sqrt: FP64 exact, FP32 hardware
//     struct Cache { ... };  // what rpn2 and both rpt2 calls of an
//                            // interface share (e.g. Roe averages)
//     // components m of wave mw that are zero by construction (the FP32
//     // paths skip their products); return false when unsure
//     __device__ static constexpr bool wave_zero(int ixy, int mw, int m);
//     __device__ static Cache prep(const amrgpu::Physics&, int ixy,
//                                  const T* ql, const T* qr);
//     // ql = qr(:,i-1) and qr = ql(:,i) of the reference: the states left
//     // and right of the interface; aux pointers are NULL (no aux arrays)
//     __device__ static void rpn2(const amrgpu::Physics&, int ixy,
//         const T* ql, const T* qr, const T* auxl, const T* auxr,
//         const Cache&, T wave[NW][NV], T s[NW], T amdq[NV], T apdq[NV]);
//     // imp = 1: asdq is A-dq, imp = 2: A+dq; bm/bp = B-asdq / B+asdq
//     __device__ static void rpt2(const amrgpu::Physics&, int ixy, int imp,
//         const T* ql, const T* qr, const T* aux1, const T* aux2,
//         const T* aux3, const Cache&, const T asdq[NV], T bm[NV], T bp[NV]);
//     // the physical flux along ixy (the mixed path's flux form)
//     __device__ static void flux(const amrgpu::Physics&, int ixy,
//                                 const T* q, T f[NV]);
//   };
//   AMRGPU_DEFINE_SOLVER(my_solver_ops, MySolver, "my solver")
//
// The solver's constants come from Physics::user[] (amrgpu_set_solver_params,
// the /cparam/ block); ixy = 1 is an x-slice (normal momentum component 1),
// ixy = 2 a y-slice.  FP64 instantiations evaluate exactly what the solver
// writes; build the translation unit with -ffp-contract=off so that no
// multiply-add is fused (the library's FP64 units are built the same way).
//
// The application's translation unit then holds its own instantiations of
// the library's batched sweep (all three precisions, every occupancy bound,
// batched and persistent launches) and qad kernels for the solver; the
// exported function fills an amrgpu_solver_ops table, which the application
// passes to amrgpu_register_solver (include/amrgpu.h).
#pragma once
#include "amrgpu.h"
#include "../amrclaw_amd/csrc/qad.hpp"
#include "../amrclaw_amd/csrc/sweep_launch.hpp"

namespace amrgpu {

template <template <typename, class> class SolT>
struct SolverPlugin {
    using Sol64 = SolT<double, ExactMath<double>>;
    static int sweep_f64(const void* a, size_t lds, void* st) {
        return (int)sweep_solver<SolT, double, ExactMath<double>, double, false>(
            *static_cast<const SweepArgs*>(a), lds, static_cast<hipStream_t>(st), 0);
    }
    static int sweep_f32(const void* a, size_t lds, void* st) {
        return (int)sweep_solver<SolT, float, FastMathF32, float, false>(
            *static_cast<const SweepArgs*>(a), lds, static_cast<hipStream_t>(st), 0);
    }
    static int sweep_mixed(const void* a, size_t lds, void* st, int wgs) {
        return (int)sweep_solver<SolT, float, FastMathF32, double, false>(
            *static_cast<const SweepArgs*>(a), lds, static_cast<hipStream_t>(st), wgs);
    }
    static int qad(int esize, const void* a, const void* ph, void* st) {
        return (int)finereg_launch<Sol64>(esize, *static_cast<const QadArgs*>(a),
                                          *static_cast<const Physics*>(ph),
                                          static_cast<hipStream_t>(st));
    }
    static amrgpu_solver_ops ops(const char* name) {
        amrgpu_solver_ops o{};
        o.abi = AMRGPU_SOLVER_ABI;
        o.nvar = Sol64::NV;
        o.mwaves = Sol64::NW;
        o.sweep_args_bytes = (int)sizeof(SweepArgs);
        o.qad_args_bytes = (int)sizeof(QadArgs);
        o.physics_bytes = (int)sizeof(Physics);
        o.name = name;
        o.sweep_f64 = &sweep_f64;
        o.sweep_f32 = &sweep_f32;
        o.sweep_mixed = &sweep_mixed;
        o.qad = &qad;
        return o;
    }
};

}  // namespace amrgpu

// Defines `extern "C" int entry(amrgpu_solver_ops*)`, which fills the table
// of solver template SolverTemplate (returns 0; 1 for a NULL table).
#define AMRGPU_DEFINE_SOLVER(entry, SolverTemplate, label)                      \
    extern "C" int entry(amrgpu_solver_ops* ops) {                              \
        if (!ops) return 1;                                                      \
        *ops = amrgpu::SolverPlugin<SolverTemplate>::ops(label);                 \
        return 0;                                                                \
    }
