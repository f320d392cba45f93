// riemann.hpp -- device Riemann solvers behind the rpn2/rpt2 plugin contract.
//
// The reference calls the application's solvers once per 1-D slice
// (flux2.f:125 rpn2; flux2.f:202,219 rpt2 with imp = 1, 2; qad.f:129 rpn2).
// Here a solver is a struct of __device__ functions called once per
// interface with the same argument meaning:
//   ixy      1 = x-slice, 2 = y-slice (selects the normal momentum mu)
//   imp      1: asdq is A-dq, 2: asdq is A+dq (flux2.f:202,219)
//   ql, qr   the reference's qr(:,i-1) and ql(:,i): the states left and
//            right of interface i (AMRClaw passes one slice for both,
//            flux2.f:125-126)
//   auxl/auxr the aux values (naux each) of the cells left and right of
//            the interface; aux1..3 those of the cell a transverse
//            fluctuation entered and of its neighbours below and above it
//            (x-slice; left and right of it for a y-slice), the entries
//            step2.f90:98-102 hands to rpt2 at index i-2+imp.  Solvers
//            that read aux declare NAUX (the components they read);
//            the others receive null pointers
//   outputs  wave[mw][m], s[mw], amdq[m], apdq[m]; bmasdq[m], bpasdq[m]
// A solver also defines Cache/prep(): quantities both rpn2 and the two rpt2
// calls at one interface derive from the same two states (Roe averages).
// Computing them once is exactly the recomputation the reference does, so
// results are unchanged.
//
// Algorithms (the reference links them from clawpack/riemann, which is not in
// its tree): LeVeque, Finite Volume Methods for Hyperbolic Problems (2002) --
// advection sec. 20.2, acoustics sec. 3.10/21.2, Euler Roe solver sec. 15.3
// with the Harten-Hyman entropy fix sec. 15.3.5, transverse splitting
// sec. 21.6-21.7.  FP64 instantiations evaluate every expression in the same
// order as the CPU oracle so that results agree bit for bit.
#pragma once
#include <hip/hip_runtime.h>

namespace amrgpu {

enum SolverId { SOLVER_ADVECTION = 1, SOLVER_ACOUSTICS = 2, SOLVER_EULER = 3, SOLVER_VCADV = 4 };

// the aux components a solver reads (Sol::NAUX when it declares one, else 0)
template <class S, class = void> struct sol_naux { static constexpr int value = 0; };
template <class S> struct sol_naux<S, decltype((void)S::NAUX, void())> {
    static constexpr int value = S::NAUX;
};

// number of constants a plug-in solver may read (Physics::user)
constexpr int SOLVER_PARAMS_MAX = 16;

// physics constants (the reference's common /cparam/)
struct Physics {
    double ubar, vbar;          // advection velocities
    double rho, bulk, cc, zz;   // acoustics: density, bulk modulus, c, Z
    double gamma;               // Euler
    int efix;                   // Euler entropy fix
    // the constants of a plug-in solver, in the order its application sets
    // them (amrgpu_set_solver_params): the /cparam/ block of a solver the
    // library does not build in
    double user[SOLVER_PARAMS_MAX];
};

template <typename T> __device__ __forceinline__ T dmin1(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T dmax1(T a, T b) { return a > b ? a : b; }

template <typename T> __device__ __forceinline__ T dsqrt(T x);
template <> __device__ __forceinline__ double dsqrt<double>(double x) { return __builtin_sqrt(x); }
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_sqrtf(x); }

// Arithmetic policies.  Exact: IEEE division and square root, the
// reference's FP64 arithmetic.  Fast (FP32 bulk path): hardware reciprocal
// and square root (about 1 ulp), which removes the refinement sequences of
// correctly rounded FP32 division from the solvers' inner loops.
template <typename T> struct ExactMath {
    using type = T;
    __device__ static __forceinline__ T div(T a, T b) { return a / b; }
    __device__ static __forceinline__ T sqrt(T x) { return dsqrt(x); }
    // the wave limiter's ratio r = dot / |wave|^2 (inlinelimiter.f:36-50)
    __device__ static __forceinline__ T ratio(T a, T b) { return a / b; }
};
struct FastMathF32 {
    using type = float;
    __device__ static __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    // The limiter ratio r = dot / |wave|^2.  A wave below about 1e-19 has a
    // subnormal |wave|^2, whose hardware reciprocal overflows to inf, and a
    // zero neighbour dot would then give 0 * inf = NaN.  Both operands are
    // scaled by 2^64 there (exact: powers of two), which makes the divisor
    // normal and leaves every other ratio bit for bit as before.
    __device__ static __forceinline__ float ratio(float a, float b) {
        const float sc = b < 1.17549435e-38f ? 0x1p64f : 1.0f;
        return (a * sc) * __builtin_amdgcn_rcpf(b * sc);
    }
};

// ----------------------------------------------------------------------------
template <typename T, class Mth = ExactMath<T>> struct Advection {
    // the same solver in another precision (FP64 fluxes of the mixed path)
    template <typename U> using Rebind = Advection<U, ExactMath<U>>;
    static constexpr int NV = 1, NW = 1;
    using Math = Mth;
    struct Cache {};
    // components of a wave that are zero by construction (none here)
    __device__ static constexpr bool wave_zero(int, int, int) { return false; }
    __device__ static Cache prep(const Physics&, int, const T*, const T*) { return {}; }
    __device__ static void rpn2(const Physics& ph, int ixy, const T* ql, const T* qr,
                                const T*, const T*, const Cache&, T wave[NW][NV],
                                T s[NW], T amdq[NV], T apdq[NV]) {
        wave[0][0] = qr[0] - ql[0];
        s[0] = (ixy == 1) ? T(ph.ubar) : T(ph.vbar);
        amdq[0] = dmin1(s[0], T(0)) * wave[0][0];
        apdq[0] = dmax1(s[0], T(0)) * wave[0][0];
    }
    // physical flux f(q) in direction ixy (the flux form of the mixed path)
    __device__ static void flux(const Physics& ph, int ixy, const T* q, T f[NV]) {
        f[0] = ((ixy == 1) ? T(ph.ubar) : T(ph.vbar)) * q[0];
    }
    __device__ static void rpt2(const Physics& ph, int ixy, int, const T*, const T*,
                                const T*, const T*, const T*, const Cache&,
                                const T asdq[NV], T bm[NV], T bp[NV]) {
        T stran = (ixy == 1) ? T(ph.vbar) : T(ph.ubar);
        bm[0] = dmin1(stran, T(0)) * asdq[0];
        bp[0] = dmax1(stran, T(0)) * asdq[0];
    }
};

// ----------------------------------------------------------------------------
// Variable-coefficient advection q_t + u q_x + v q_y = 0 with the edge
// velocities as aux fields (clawpack/riemann's rpn2_vc_advection /
// rpt2_vc_advection, the solvers of the reference's
// tests/advection_2d_annulus): aux(1) the x-velocity at a cell's left edge,
// aux(2) the y-velocity at its bottom edge.  The wave is the jump, moving at
// the velocity of the interface (the right cell's left / bottom edge); a
// transverse fluctuation splits by the velocities of the edges below and
// above the cell it entered (LeVeque 2002 sec. 9.5, 21.2).
template <typename T, class Mth = ExactMath<T>> struct VcAdvection {
    template <typename U> using Rebind = VcAdvection<U, ExactMath<U>>;
    static constexpr int NV = 1, NW = 1, NAUX = 2;
    using Math = Mth;
    struct Cache {};
    __device__ static constexpr bool wave_zero(int, int, int) { return false; }
    __device__ static Cache prep(const Physics&, int, const T*, const T*) { return {}; }
    __device__ static void rpn2(const Physics&, int ixy, const T* ql, const T* qr, const T*,
                                const T* auxr, const Cache&, T wave[NW][NV], T s[NW],
                                T amdq[NV], T apdq[NV]) {
        wave[0][0] = qr[0] - ql[0];
        s[0] = auxr[ixy - 1];
        amdq[0] = dmin1(auxr[ixy - 1], T(0)) * wave[0][0];
        apdq[0] = dmax1(auxr[ixy - 1], T(0)) * wave[0][0];
    }
    // no state-only flux (the velocities live on the edges): this solver runs
    // in the fluctuation form only (FP64 and FP32 storage)
    __device__ static void flux(const Physics&, int, const T*, T f[NV]) { f[0] = T(0); }
    __device__ static void rpt2(const Physics&, int ixy, int, const T*, const T*, const T*,
                                const T* aux2, const T* aux3, const Cache&, const T asdq[NV],
                                T bm[NV], T bp[NV]) {
        const int kv = 3 - ixy;
        bm[0] = dmin1(aux2[kv - 1], T(0)) * asdq[0];
        bp[0] = dmax1(aux3[kv - 1], T(0)) * asdq[0];
    }
};

// ----------------------------------------------------------------------------
template <typename T, class Mth = ExactMath<T>> struct Acoustics {
    // the same solver in another precision (FP64 fluxes of the mixed path)
    template <typename U> using Rebind = Acoustics<U, ExactMath<U>>;
    static constexpr int NV = 3, NW = 2;
    using Math = Mth;
    struct Cache {};
    // components of a wave that are zero by construction: the transverse
    // momentum of both waves
    __device__ static constexpr bool wave_zero(int ixy, int, int m) { return m == (ixy == 1 ? 2 : 1); }
    __device__ static Cache prep(const Physics&, int, const T*, const T*) { return {}; }
    __device__ static void rpn2(const Physics& ph, int ixy, const T* ql, const T* qr,
                                const T*, const T*, const Cache&, T wave[NW][NV],
                                T s[NW], T amdq[NV], T apdq[NV]) {
        const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
        const T zz = T(ph.zz), cc = T(ph.cc);
        T d1 = qr[0] - ql[0];
        T d2 = qr[mu] - ql[mu];
        T a1 = Mth::div(-d1 + zz * d2, T(2) * zz);
        T a2 = Mth::div(d1 + zz * d2, T(2) * zz);
        wave[0][0] = -a1 * zz; wave[0][mu] = a1; wave[0][mv] = T(0); s[0] = -cc;
        wave[1][0] = a2 * zz;  wave[1][mu] = a2; wave[1][mv] = T(0); s[1] = cc;
#pragma unroll
        for (int m = 0; m < NV; m++) {
            amdq[m] = s[0] * wave[0][m];
            apdq[m] = s[1] * wave[1][m];
        }
    }
    // f(q) = A q: (K u, p / rho, 0) in x, (K v, 0, p / rho) in y
    __device__ static void flux(const Physics& ph, int ixy, const T* q, T f[NV]) {
        const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
        f[0] = T(ph.bulk) * q[mu];
        f[mu] = Mth::div(q[0], T(ph.rho));
        f[mv] = T(0);
    }
    __device__ static void rpt2(const Physics& ph, int ixy, int, const T*, const T*,
                                const T*, const T*, const T*, const Cache&,
                                const T asdq[NV], T bm[NV], T bp[NV]) {
        const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
        const T zz = T(ph.zz), cc = T(ph.cc);
        T a1 = Mth::div(-asdq[0] + zz * asdq[mv], T(2) * zz);
        T a2 = Mth::div(asdq[0] + zz * asdq[mv], T(2) * zz);
        bm[0] = cc * a1 * zz; bm[mu] = T(0); bm[mv] = -cc * a1;
        bp[0] = cc * a2 * zz; bp[mu] = T(0); bp[mv] = cc * a2;
    }
};

// ----------------------------------------------------------------------------
template <typename T, class Mth = ExactMath<T>> struct EulerRoe {
    // the same solver in another precision (FP64 fluxes of the mixed path)
    template <typename U> using Rebind = EulerRoe<U, ExactMath<U>>;
    static constexpr int NV = 4, NW = 4;
    using Math = Mth;
    struct Cache { T u, v, enth, a, g1a2, euv, u2v2; };
    // components of a wave that are zero by construction: density and normal
    // momentum of the shear wave (wave 2)
    __device__ static constexpr bool wave_zero(int ixy, int mw, int m) {
        return mw == 1 && (m == 0 || m == (ixy == 1 ? 1 : 2));
    }
    // FP32 arithmetic leaves out the products with those zero components
    // (x + s*0); the FP64 instantiation keeps the reference's sequence
    static constexpr bool SKIPZ = sizeof(T) == 4;

    __device__ static Cache prep(const Physics& ph, int ixy, const T* L, const T* R) {
        const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
        const T gamma1 = T(ph.gamma) - T(1);
        Cache c;
        T rhsqrtl = Mth::sqrt(L[0]);
        T rhsqrtr = Mth::sqrt(R[0]);
        T pl = gamma1 * (L[3] - Mth::div(T(0.5) * (L[1] * L[1] + L[2] * L[2]), L[0]));
        T pr = gamma1 * (R[3] - Mth::div(T(0.5) * (R[1] * R[1] + R[2] * R[2]), R[0]));
        T rhsq2 = rhsqrtl + rhsqrtr;
        c.u = Mth::div(Mth::div(L[mu], rhsqrtl) + Mth::div(R[mu], rhsqrtr), rhsq2);
        c.v = Mth::div(Mth::div(L[mv], rhsqrtl) + Mth::div(R[mv], rhsqrtr), rhsq2);
        c.enth = Mth::div(Mth::div(L[3] + pl, rhsqrtl) + Mth::div(R[3] + pr, rhsqrtr), rhsq2);
        c.u2v2 = c.u * c.u + c.v * c.v;
        T a2 = gamma1 * (c.enth - T(0.5) * c.u2v2);
        c.a = Mth::sqrt(a2);
        c.g1a2 = Mth::div(gamma1, a2);
        c.euv = c.enth - c.u2v2;
        return c;
    }

    __device__ static T pressure(T gamma1, T rho, T mx, T my, T en) {
        return gamma1 * (en - Mth::div(T(0.5) * (mx * mx + my * my), rho));
    }

    // f(q) = (rho u, rho u^2 + p, rho u v, u (E + p)) along momentum mu (the
    // mixed path's FP64 fluxes): one division, by rho, shared by p and u
    __device__ static void flux(const Physics& ph, int ixy, const T* q, T f[NV]) {
        const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
        const T rinv = Mth::div(T(1), q[0]);
        const T p = (T(ph.gamma) - T(1)) * (q[3] - T(0.5) * (q[1] * q[1] + q[2] * q[2]) * rinv);
        const T u = q[mu] * rinv;
        f[0] = q[mu];
        f[mu] = q[mu] * u + p;
        f[mv] = q[mv] * u;
        f[3] = u * (q[3] + p);
    }

    __device__ static void rpn2(const Physics& ph, int ixy, const T* L, const T* R,
                                const T*, const T*, const Cache& c, T wave[NW][NV],
                                T s[NW], T amdq[NV], T apdq[NV]) {
        const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
        const T gamma = T(ph.gamma), gamma1 = T(ph.gamma) - T(1);
        const T u = c.u, v = c.v, enth = c.enth, a = c.a;
        T d1 = R[0] - L[0];
        T d2 = R[mu] - L[mu];
        T d3 = R[mv] - L[mv];
        T d4 = R[3] - L[3];
        T a3 = c.g1a2 * (c.euv * d1 + u * d2 + v * d3 - d4);
        T a2 = d3 - v * d1;
        T a4 = Mth::div(d2 + (a - u) * d1 - a * a3, T(2) * a);
        T a1 = d1 - a3 - a4;
        wave[0][0] = a1; wave[0][mu] = a1 * (u - a); wave[0][mv] = a1 * v;
        wave[0][3] = a1 * (enth - u * a); s[0] = u - a;
        wave[1][0] = T(0); wave[1][mu] = T(0); wave[1][mv] = a2;
        wave[1][3] = a2 * v; s[1] = u;
        wave[2][0] = a3; wave[2][mu] = a3 * u; wave[2][mv] = a3 * v;
        wave[2][3] = a3 * T(0.5) * c.u2v2; s[2] = u;
        wave[3][0] = a4; wave[3][mu] = a4 * (u + a); wave[3][mv] = a4 * v;
        wave[3][3] = a4 * (enth + u * a); s[3] = u + a;

        T am[4] = {T(0), T(0), T(0), T(0)};
        if (!ph.efix) {
#pragma unroll
            for (int mw = 0; mw < 4; mw++)
#pragma unroll
                for (int m = 0; m < 4; m++)
                    if (!(SKIPZ && wave_zero(ixy, mw, m)))
                        am[m] = am[m] + dmin1(s[mw], T(0)) * wave[mw][m];
        } else {
            T rhoim1 = L[0];
            T pim1 = pressure(gamma1, L[0], L[1], L[2], L[3]);
            T cim1 = Mth::sqrt(Mth::div(gamma * pim1, rhoim1));
            T s0 = Mth::div(L[mu], rhoim1) - cim1;
            if (!(s0 >= T(0) && s[0] > T(0))) {
                T rho1 = L[0] + wave[0][0];
                T rhou1 = L[mu] + wave[0][mu];
                T rhov1 = L[mv] + wave[0][mv];
                T en1 = L[3] + wave[0][3];
                T p1 = pressure(gamma1, rho1, rhou1, rhov1, en1);
                T c1 = Mth::sqrt(Mth::div(gamma * p1, rho1));
                T s1 = Mth::div(rhou1, rho1) - c1;
                T sfract;
                if (s0 < T(0) && s1 > T(0))
                    sfract = Mth::div(s0 * (s1 - s[0]), s1 - s0);
                else if (s[0] < T(0))
                    sfract = s[0];
                else
                    sfract = T(0);
#pragma unroll
                for (int m = 0; m < 4; m++) am[m] = sfract * wave[0][m];
                if (s[1] < T(0)) {
#pragma unroll
                    for (int m = 0; m < 4; m++)
                        am[m] = (SKIPZ && wave_zero(ixy, 1, m) ? am[m] : am[m] + s[1] * wave[1][m]) +
                                s[2] * wave[2][m];
                    T rhoi = R[0];
                    T pi_ = pressure(gamma1, R[0], R[1], R[2], R[3]);
                    T ci = Mth::sqrt(Mth::div(gamma * pi_, rhoi));
                    T s3 = Mth::div(R[mu], rhoi) + ci;
                    T rho2 = R[0] - wave[3][0];
                    T rhou2 = R[mu] - wave[3][mu];
                    T rhov2 = R[mv] - wave[3][mv];
                    T en2 = R[3] - wave[3][3];
                    T p2 = pressure(gamma1, rho2, rhou2, rhov2, en2);
                    T c2 = Mth::sqrt(Mth::div(gamma * p2, rho2));
                    T s2 = Mth::div(rhou2, rho2) + c2;
                    T sf4 = T(0);
                    bool add4 = true;
                    if (s2 < T(0) && s3 > T(0))
                        sf4 = Mth::div(s2 * (s3 - s[3]), s3 - s2);
                    else if (s[3] < T(0))
                        sf4 = s[3];
                    else
                        add4 = false;
                    if (add4) {
#pragma unroll
                        for (int m = 0; m < 4; m++) am[m] = am[m] + sf4 * wave[3][m];
                    }
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 4; m++) {
            amdq[m] = am[m];
            if (ph.efix) {
                T df = T(0);
#pragma unroll
                for (int mw = 0; mw < 4; mw++)
                    if (!(SKIPZ && wave_zero(ixy, mw, m))) df = df + s[mw] * wave[mw][m];
                apdq[m] = df - am[m];
            } else {
                T ap = T(0);
#pragma unroll
                for (int mw = 0; mw < 4; mw++)
                    if (!(SKIPZ && wave_zero(ixy, mw, m))) ap = ap + dmax1(s[mw], T(0)) * wave[mw][m];
                apdq[m] = ap;
            }
        }
    }

    __device__ static void rpt2(const Physics&, int ixy, int, const T*, const T*,
                                const T*, const T*, const T*, const Cache& c,
                                const T asdq[NV], T bm[NV], T bp[NV]) {
        const int mu = (ixy == 1) ? 1 : 2, mv = (ixy == 1) ? 2 : 1;
        const T u = c.u, v = c.v, enth = c.enth, a = c.a;
        T A1 = asdq[0], Amu = asdq[mu], Amv = asdq[mv], A4 = asdq[3];
        T a3 = c.g1a2 * (c.euv * A1 + u * Amu + v * Amv - A4);
        T a2 = Amu - u * A1;
        T a4 = Mth::div(Amv + (a - v) * A1 - a * a3, T(2) * a);
        T a1 = A1 - a3 - a4;
        T wb[4][4], sb[4];
        wb[0][0] = a1; wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a);
        wb[0][3] = a1 * (enth - v * a); sb[0] = v - a;
        wb[1][0] = a3; wb[1][mu] = a3 * u; wb[1][mv] = a3 * v;
        wb[1][3] = a3 * T(0.5) * c.u2v2; sb[1] = v;
        wb[2][0] = T(0); wb[2][mu] = a2; wb[2][mv] = T(0);
        wb[2][3] = a2 * u; sb[2] = v;
        wb[3][0] = a4; wb[3][mu] = a4 * u; wb[3][mv] = a4 * (v + a);
        wb[3][3] = a4 * (enth + v * a); sb[3] = v + a;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            T xm = T(0), xp = T(0);
#pragma unroll
            for (int mw = 0; mw < 4; mw++) {
                // wave 3 of the transverse split carries no density and no
                // transverse momentum
                if (SKIPZ && mw == 2 && (m == 0 || m == mv)) continue;
                xm = xm + dmin1(sb[mw], T(0)) * wb[mw][m];
                xp = xp + dmax1(sb[mw], T(0)) * wb[mw][m];
            }
            bm[m] = xm;
            bp[m] = xp;
        }
    }
};

}  // namespace amrgpu
