// riemann3.hpp -- device Riemann solvers behind the 3-D rpn3/rpt3/rptt3
// plugin contract.
//
// The reference calls the application's solvers once per 1-D slice
// (flux3.f:220 rpn3; flux3.f:302-321 rpt3 of A-dq and A+dq in the second and
// third transverse directions; flux3.f:326-377 rpt3 of the correction cqxx;
// flux3.f:411-422,538-545 rptt3).  Here they are __device__ functions called
// once per interface with the same argument meaning:
//   ixyz     1, 2, 3: direction of the slice; (mu, mv, mw) is the cyclic
//            triple of momentum components (normal, first and second
//            transverse)
//   icoor    2: split along mv, 3: split along mw (rpt3/rptt3)
//   L, R     the reference's qr(:,i-1) and ql(:,i): states left and right of
//            interface i
//   asdq     the fluctuation to split; bm/bp its down- and up-going parts
// rpt3 and rptt3 of the solvers here split a fluctuation the same way (the
// double-transverse solve of a fluctuation that was already split is one
// more split of that vector), so both are `split`.  Neither reads aux: imp
// and impt only select aux rows in the reference's convention.
//
// Algorithms (clawpack/riemann, not in the reference's tree): LeVeque,
// Finite Volume Methods for Hyperbolic Problems (2002) -- advection
// sec. 20.2, Euler Roe solver sec. 15.3 with the Harten-Hyman entropy fix
// sec. 15.3.5, 3-D transverse splitting sec. 21.6-21.7.  The Euler solver
// has three waves: the acoustic waves u -+ c and one wave at speed u that
// carries the entropy jump and both shear jumps.  FP64 instantiations
// evaluate every expression in the order of the CPU oracle (compiled without
// FMA contraction), so results agree bit for bit.
#pragma once
#include "riemann.hpp"

namespace amrgpu {

struct Physics3 {
    double ubar, vbar, wbar;   // advection velocities
    double gamma;              // Euler
    int32_t efix;              // Euler entropy fix
    int32_t pad;
};

// 0-based components of the normal momentum and the two transverse momenta
// of sweep ixyz (x: 1,2,3; y: 2,3,1; z: 3,1,2)
__device__ __forceinline__ void triple3(int ixyz, int& mu, int& mv, int& mw) {
    mu = ixyz;
    mv = ixyz == 3 ? 1 : ixyz + 1;
    mw = ixyz == 1 ? 3 : ixyz - 1;
}

// ----------------------------------------------------------------------------
template <typename T, class Mth = ExactMath<T>> struct Advection3 {
    static constexpr int NV = 1, NW = 1;
    using Math = Mth;
    struct Cache {};
    __device__ static Cache prep(const Physics3&, const T*, const T*) { return {}; }
    __device__ static T velocity(const Physics3& ph, int comp) {
        return comp == 1 ? T(ph.ubar) : (comp == 2 ? T(ph.vbar) : T(ph.wbar));
    }
    __device__ static void rpn3(const Physics3& ph, int ixyz, const T* L, const T* R, const Cache&,
                                T wave[NW][NV], T s[NW], T amdq[NV], T apdq[NV]) {
        const T sp = velocity(ph, ixyz);
        wave[0][0] = R[0] - L[0];
        s[0] = sp;
        amdq[0] = dmin1(sp, T(0)) * wave[0][0];
        apdq[0] = dmax1(sp, T(0)) * wave[0][0];
    }
    __device__ static void split(const Physics3& ph, int ixyz, int icoor, const T*, const T*,
                                 const Cache&, const T asdq[NV], T bm[NV], T bp[NV]) {
        int mu, mv, mw;
        triple3(ixyz, mu, mv, mw);
        const T sp = velocity(ph, icoor == 2 ? mv : mw);
        bm[0] = dmin1(sp, T(0)) * asdq[0];
        bp[0] = dmax1(sp, T(0)) * asdq[0];
    }
};

// ----------------------------------------------------------------------------
template <typename T, class Mth = ExactMath<T>> struct Euler3 {
    static constexpr int NV = 5, NW = 3;
    using Math = Mth;
    // Roe average of the interface (velocity components u[0..2] = x, y, z)
    struct Cache { T u[3]; T enth, a, g1a2, euv, u2; };

    __device__ static T pressure(T gamma1, const T* q) {
        return gamma1 * (q[4] - Mth::div(T(0.5) * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]), q[0]));
    }

    __device__ static Cache prep(const Physics3& ph, const T* L, const T* R) {
        const T gamma1 = T(ph.gamma) - T(1);
        Cache c;
        T rhsqrtl = Mth::sqrt(L[0]);
        T rhsqrtr = Mth::sqrt(R[0]);
        T pl = pressure(gamma1, L);
        T pr = pressure(gamma1, R);
        T rhsq2 = rhsqrtl + rhsqrtr;
#pragma unroll
        for (int m = 0; m < 3; m++)
            c.u[m] = Mth::div(Mth::div(L[m + 1], rhsqrtl) + Mth::div(R[m + 1], rhsqrtr), rhsq2);
        c.enth = Mth::div(Mth::div(L[4] + pl, rhsqrtl) + Mth::div(R[4] + pr, rhsqrtr), rhsq2);
        c.u2 = c.u[0] * c.u[0] + c.u[1] * c.u[1] + c.u[2] * c.u[2];
        T a2 = gamma1 * (c.enth - T(0.5) * c.u2);
        c.a = Mth::sqrt(a2);
        c.g1a2 = Mth::div(gamma1, a2);
        c.euv = c.enth - c.u2;
        return c;
    }

    // eigen-decomposition of d along momentum component md (0-based 1..3)
    // with the other two momenta o1, o2: waves wv[3][5], speeds sp[3]
    __device__ static void waves(const Cache& c, int md, int o1, int o2, const T* d, T wv[3][5],
                                 T sp[3]) {
        const T ud = c.u[md - 1], a = c.a, u1 = c.u[o1 - 1], u2 = c.u[o2 - 1];
        T a4 = c.g1a2 * (c.euv * d[0] + c.u[0] * d[1] + c.u[1] * d[2] + c.u[2] * d[3] - d[4]);
        T a2 = d[o1] - u1 * d[0];
        T a3 = d[o2] - u2 * d[0];
        T a5 = Mth::div(d[md] + (a - ud) * d[0] - a * a4, T(2) * a);
        T a1 = d[0] - a4 - a5;
        wv[0][0] = a1;
        wv[0][md] = a1 * (ud - a);
        wv[0][o1] = a1 * u1;
        wv[0][o2] = a1 * u2;
        wv[0][4] = a1 * (c.enth - ud * a);
        sp[0] = ud - a;
        wv[1][0] = a4;
        wv[1][md] = a4 * ud;
        wv[1][o1] = a4 * u1 + a2;
        wv[1][o2] = a4 * u2 + a3;
        wv[1][4] = a4 * T(0.5) * c.u2 + a2 * u1 + a3 * u2;
        sp[1] = ud;
        wv[2][0] = a5;
        wv[2][md] = a5 * (ud + a);
        wv[2][o1] = a5 * u1;
        wv[2][o2] = a5 * u2;
        wv[2][4] = a5 * (c.enth + ud * a);
        sp[2] = ud + a;
    }

    __device__ static void rpn3(const Physics3& ph, int ixyz, const T* L, const T* R,
                                const Cache& c, T wave[NW][NV], T s[NW], T amdq[NV],
                                T apdq[NV]) {
        int mu, mv, mw;
        triple3(ixyz, mu, mv, mw);
        T d[5];
#pragma unroll
        for (int m = 0; m < 5; m++) d[m] = R[m] - L[m];
        waves(c, mu, mv, mw, d, wave, s);
        T am[5] = {T(0), T(0), T(0), T(0), T(0)};
        if (!ph.efix) {
#pragma unroll
            for (int w = 0; w < 3; w++)
#pragma unroll
                for (int m = 0; m < 5; m++) am[m] = am[m] + dmin1(s[w], T(0)) * wave[w][m];
        } else {
            const T gamma = T(ph.gamma), gamma1 = T(ph.gamma) - T(1);
            T cim1 = Mth::sqrt(Mth::div(gamma * pressure(gamma1, L), L[0]));
            T s0 = Mth::div(L[mu], L[0]) - cim1;
            if (!(s0 >= T(0) && s[0] > T(0))) {
                T q1[5];
#pragma unroll
                for (int m = 0; m < 5; m++) q1[m] = L[m] + wave[0][m];
                T c1 = Mth::sqrt(Mth::div(gamma * pressure(gamma1, q1), q1[0]));
                T s1 = Mth::div(q1[mu], q1[0]) - c1;
                T sfract;
                if (s0 < T(0) && s1 > T(0))
                    sfract = Mth::div(s0 * (s1 - s[0]), s1 - s0);
                else if (s[0] < T(0))
                    sfract = s[0];
                else
                    sfract = T(0);
#pragma unroll
                for (int m = 0; m < 5; m++) am[m] = sfract * wave[0][m];
                if (s[1] < T(0)) {
#pragma unroll
                    for (int m = 0; m < 5; m++) am[m] = am[m] + s[1] * wave[1][m];
                    T ci = Mth::sqrt(Mth::div(gamma * pressure(gamma1, R), R[0]));
                    T s3 = Mth::div(R[mu], R[0]) + ci;
                    T q2[5];
#pragma unroll
                    for (int m = 0; m < 5; m++) q2[m] = R[m] - wave[2][m];
                    T c2 = Mth::sqrt(Mth::div(gamma * pressure(gamma1, q2), q2[0]));
                    T s2 = Mth::div(q2[mu], q2[0]) + c2;
                    T sf3 = T(0);
                    bool add3 = true;
                    if (s2 < T(0) && s3 > T(0))
                        sf3 = Mth::div(s2 * (s3 - s[2]), s3 - s2);
                    else if (s[2] < T(0))
                        sf3 = s[2];
                    else
                        add3 = false;
                    if (add3) {
#pragma unroll
                        for (int m = 0; m < 5; m++) am[m] = am[m] + sf3 * wave[2][m];
                    }
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 5; m++) {
            T df = T(0);
#pragma unroll
            for (int w = 0; w < 3; w++) df = df + s[w] * wave[w][m];
            amdq[m] = am[m];
            if (ph.efix) {
                apdq[m] = df - am[m];
            } else {
                T ap = T(0);
#pragma unroll
                for (int w = 0; w < 3; w++) ap = ap + dmax1(s[w], T(0)) * wave[w][m];
                apdq[m] = ap;
            }
        }
    }

    __device__ static void split(const Physics3&, int ixyz, int icoor, const T*, const T*,
                                 const Cache& c, const T asdq[NV], T bm[NV], T bp[NV]) {
        int mu, mv, mw;
        triple3(ixyz, mu, mv, mw);
        const int md = icoor == 2 ? mv : mw;
        const int o1 = icoor == 2 ? mw : mu, o2 = icoor == 2 ? mu : mv;
        T wv[3][5], sp[3];
        waves(c, md, o1, o2, asdq, wv, sp);
#pragma unroll
        for (int m = 0; m < 5; m++) {
            T a = T(0), b = T(0);
#pragma unroll
            for (int w = 0; w < 3; w++) {
                a = a + dmin1(sp[w], T(0)) * wv[w][m];
                b = b + dmax1(sp[w], T(0)) * wv[w][m];
            }
            bm[m] = a;
            bp[m] = b;
        }
    }
};

}  // namespace amrgpu
