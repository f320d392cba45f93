// shallow_water.hip -- a Riemann solver defined outside the library: the
// 2-D shallow-water equations q = (h, hu, hv) with the Roe split (LeVeque
// 2002 sec. 15.3.3; transverse split sec. 21.6), written against the
// plug-in contract of include/amrgpu_solver.hpp and registered at run time
// (amrgpu_register_solver).  Gravity is the solver's constant user[0]
// (amrgpu_set_solver_params), the counterpart of its setprob's /cparam/.
// Built with -ffp-contract=off so that its FP64 arithmetic matches the
// oracle's restatement (oracle/orc_riemann.c, rpn2_shallow / rpt2_shallow)
// bit for bit.
#include "amrgpu_solver.hpp"

template <typename T, class Mth = amrgpu::ExactMath<T>>
struct ShallowWater {
    template <typename U> using Rebind = ShallowWater<U, amrgpu::ExactMath<U>>;
    static constexpr int NV = 3, NW = 3;
    using Math = Mth;
    struct Cache { T u, v, c; };
    // the shear wave carries neither depth nor normal momentum
    __device__ static constexpr bool wave_zero(int ixy, int mw, int m) {
        return mw == 1 && (m == 0 || m == (ixy == 1 ? 1 : 2));
    }
    // Roe averages: velocities weighted by sqrt(h), celerity of the mean depth
    __device__ static Cache prep(const amrgpu::Physics& ph, int ixy, const T* L, const T* R) {
        const int mu = ixy == 1 ? 1 : 2, mv = ixy == 1 ? 2 : 1;
        const T hsl = Mth::sqrt(L[0]), hsr = Mth::sqrt(R[0]);
        const T hsq = hsl + hsr;
        Cache c;
        c.u = Mth::div(Mth::div(L[mu], hsl) + Mth::div(R[mu], hsr), hsq);
        c.v = Mth::div(Mth::div(L[mv], hsl) + Mth::div(R[mv], hsr), hsq);
        c.c = Mth::sqrt(T(ph.user[0]) * T(0.5) * (L[0] + R[0]));
        return c;
    }
    __device__ static void rpn2(const amrgpu::Physics&, int ixy, const T* L, const T* R,
                                const T*, const T*, const Cache& c, T wave[NW][NV], T s[NW],
                                T amdq[NV], T apdq[NV]) {
        const int mu = ixy == 1 ? 1 : 2, mv = ixy == 1 ? 2 : 1;
        const T d1 = R[0] - L[0], d2 = R[mu] - L[mu], d3 = R[mv] - L[mv];
        const T a1 = Mth::div((c.u + c.c) * d1 - d2, T(2) * c.c);
        const T a2 = d3 - c.v * d1;
        const T a3 = Mth::div(d2 - (c.u - c.c) * d1, T(2) * c.c);
        wave[0][0] = a1; wave[0][mu] = a1 * (c.u - c.c); wave[0][mv] = a1 * c.v; s[0] = c.u - c.c;
        wave[1][0] = T(0); wave[1][mu] = T(0); wave[1][mv] = a2; s[1] = c.u;
        wave[2][0] = a3; wave[2][mu] = a3 * (c.u + c.c); wave[2][mv] = a3 * c.v; s[2] = c.u + c.c;
#pragma unroll
        for (int m = 0; m < NV; m++) {
            T am = T(0), ap = T(0);
#pragma unroll
            for (int mw = 0; mw < NW; mw++) {
                am = am + amrgpu::dmin1(s[mw], T(0)) * wave[mw][m];
                ap = ap + amrgpu::dmax1(s[mw], T(0)) * wave[mw][m];
            }
            amdq[m] = am;
            apdq[m] = ap;
        }
    }
    __device__ static void rpt2(const amrgpu::Physics&, int ixy, int, const T*, const T*,
                                const T*, const T*, const T*, const Cache& c, const T asdq[NV],
                                T bm[NV], T bp[NV]) {
        const int mu = ixy == 1 ? 1 : 2, mv = ixy == 1 ? 2 : 1;
        const T A1 = asdq[0], Amu = asdq[mu], Amv = asdq[mv];
        const T a1 = Mth::div((c.v + c.c) * A1 - Amv, T(2) * c.c);
        const T a2 = Amu - c.u * A1;
        const T a3 = Mth::div(Amv - (c.v - c.c) * A1, T(2) * c.c);
        T wb[3][3], sb[3];
        wb[0][0] = a1; wb[0][mu] = a1 * c.u; wb[0][mv] = a1 * (c.v - c.c); sb[0] = c.v - c.c;
        wb[1][0] = T(0); wb[1][mu] = a2; wb[1][mv] = T(0); sb[1] = c.v;
        wb[2][0] = a3; wb[2][mu] = a3 * c.u; wb[2][mv] = a3 * (c.v + c.c); sb[2] = c.v + c.c;
#pragma unroll
        for (int m = 0; m < NV; m++) {
            T xm = T(0), xp = T(0);
#pragma unroll
            for (int mw = 0; mw < 3; mw++) {
                xm = xm + amrgpu::dmin1(sb[mw], T(0)) * wb[mw][m];
                xp = xp + amrgpu::dmax1(sb[mw], T(0)) * wb[mw][m];
            }
            bm[m] = xm;
            bp[m] = xp;
        }
    }
    // f = (hu, hu^2/h + g h^2/2, hu hv/h) along ixy (the mixed path's flux form)
    __device__ static void flux(const amrgpu::Physics& ph, int ixy, const T* q, T f[NV]) {
        const int mu = ixy == 1 ? 1 : 2, mv = ixy == 1 ? 2 : 1;
        const T un = Mth::div(q[mu], q[0]);
        f[0] = q[mu];
        f[mu] = q[mu] * un + T(0.5) * T(ph.user[0]) * q[0] * q[0];
        f[mv] = q[mv] * un;
    }
};

AMRGPU_DEFINE_SOLVER(amrgpu_plugin_shallow_water, ShallowWater, "shallow water (Roe)")
