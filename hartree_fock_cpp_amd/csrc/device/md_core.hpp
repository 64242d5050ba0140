// md_core.hpp -- McMurchie-Davidson building blocks for gfx950 kernels.
//
// Compile-time angular shapes, the Boys function F_m(T), the 1-D Hermite
// expansion coefficients E^{ab}_t and the Hermite Coulomb integrals R_tuv.
// Everything here is templated on the scalar type (double for the FP64 form,
// float for the small-bound quartets of the mixed form) and on the angular
// momenta, so that loops unroll fully and all arrays live in registers.
//
// Mathematics follows the reference's McMurchie-Davidson formulation
// (KLYYCYNN/hartree_fock_cpp include/integrals.hpp:269-322 for E,
// 541-618 for R, 223-266 for Boys) but the evaluation order is our own:
// R is built once per primitive quartet for every t+u+v <= L (the reference
// rebuilds it per Hermite triple), and Boys uses a Taylor table plus downward
// recursion instead of the reference's ill-conditioned upward recursion.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#if defined(__HIP_DEVICE_COMPILE__)
#define HFX_HD __device__
#else
#define HFX_HD __host__ __device__
#endif

namespace hfx {

// ------------------------------------------------------------------ shapes
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// Cartesian component k of a shell of angular momentum l, in the reference's
// angmom_partition order (basis.hpp:52-64): for lx = 0..l, for ly = 0..l-lx,
// lz = l-lx-ly.  p: (0,0,1),(0,1,0),(1,0,0)  -> z, y, x.
__host__ __device__ constexpr int cart_x(int l, int k) {
    int idx = 0;
    for (int a = 0; a <= l; ++a)
        for (int b = 0; b <= l - a; ++b) {
            if (idx == k) return a;
            ++idx;
        }
    return -1;
}
__host__ __device__ constexpr int cart_y(int l, int k) {
    int idx = 0;
    for (int a = 0; a <= l; ++a)
        for (int b = 0; b <= l - a; ++b) {
            if (idx == k) return b;
            ++idx;
        }
    return -1;
}
__host__ __device__ constexpr int cart_z(int l, int k) { return l - cart_x(l, k) - cart_y(l, k); }

// Packed Hermite index of (t,u,v), t+u+v = s: all orders below s first, then
// t ascending, then u ascending (same packing as integrals.hpp:343-358):
// s(s+1)(s+2)/6 + sum_{a<t} (s-a+1) + u, in closed form.
__host__ __device__ constexpr int hidx(int t, int u, int v) {
    const int s = t + u + v;
    return s * (s + 1) * (s + 2) / 6 + t * (s + 1) - t * (t - 1) / 2 + u;
}

// ------------------------------------------------------------------ Boys
// F_m(T) = int_0^1 x^{2m} exp(-T x^2) dx.
// Table: F_n(T_i) at T_i = i*h, h = 1/80, i = 0..BOYS_NGRID-1, n = 0..BOYS_NCOL-1.
// For T < BOYS_TMAX: Taylor expansion of F_mmax about the nearest grid point
// (|dT| <= h/2 = 1/160; BOYS_NTAYLOR = 6 terms leave (h/2)^6 / 6! < 1e-16 of
// F_{m+6} <= F_m in FP64, 4 terms 7e-11 in FP32), then the stable downward
// recursion F_m = (2T F_{m+1} + e^{-T}) / (2m+1).  The fine grid keeps a
// lookup to six table values (three 16-byte loads per lane; eight terms on a
// 1/20 grid took four): the low-L family-A kernels are bound by the texture
// address unit's load processing, not by arithmetic.
// For T >= BOYS_TMAX: F_0 = sqrt(pi/T)/2 (erfc(sqrt(40)) < 1e-19) and the
// upward recursion, which is stable for T > m.
constexpr int BOYS_MMAX = 12;        // (ff|ff) needs L = 12
constexpr int BOYS_NTAYLOR = 6;
constexpr int BOYS_NCOL = BOYS_MMAX + BOYS_NTAYLOR + 1;   // 19
constexpr double BOYS_H = 0.0125;
constexpr double BOYS_INV_H = 80.0;
constexpr double BOYS_TMAX = 40.0;
constexpr int BOYS_NGRID = 3201;     // T_i = 0 .. 40
constexpr int BOYS_NTAYLOR_F32 = 4;

template <class Real> struct MathOps;
template <> struct MathOps<double> {
    static HFX_HD __forceinline__ double exp_(double x) { return exp(x); }
    static HFX_HD __forceinline__ double sqrt_(double x) { return sqrt(x); }
    static HFX_HD __forceinline__ double rcp(double x) { return 1.0 / x; }
};
template <> struct MathOps<float> {
#if defined(__HIP_DEVICE_COMPILE__)
    static __device__ __forceinline__ float exp_(float x) { return __expf(x); }
    static __device__ __forceinline__ float sqrt_(float x) { return __fsqrt_rn(x); }
    static __device__ __forceinline__ float rcp(float x) { return __frcp_rn(x); }
#else
    static inline float exp_(float x) { return expf(x); }
    static inline float sqrt_(float x) { return sqrtf(x); }
    static inline float rcp(float x) { return 1.0f / x; }
#endif
};

// Fills F[0..M] with F_m(T).  `tab` is the row-major table [BOYS_NGRID][BOYS_NCOL].
template <int M, class Real>
HFX_HD __forceinline__ void boys_fm(Real T, const Real* __restrict__ tab, Real F[M + 1]) {
    constexpr int NT = sizeof(Real) == 8 ? BOYS_NTAYLOR : BOYS_NTAYLOR_F32;
    // e^{-T} for either recursion, evaluated once before the branch: lanes of
    // one wavefront often take both branches, which then share one exp
    Real e = Real(0);
    if constexpr (M > 0) e = MathOps<Real>::exp_(-T);
    if (T < Real(BOYS_TMAX)) {
        int i = (int)(T * Real(BOYS_INV_H) + Real(0.5));
        Real d = Real(i) * Real(BOYS_H) - T;   // -(T - T_i)
        const Real* row = tab + i * BOYS_NCOL + M;
        Real f = row[NT - 1];
#pragma unroll
        for (int k = NT - 2; k >= 0; --k) f = row[k] + f * d * Real(1.0 / (k + 1));
        F[M] = f;
        if (M > 0) {
            Real T2 = T + T;
#pragma unroll
            for (int m = M - 1; m >= 0; --m) F[m] = (T2 * F[m + 1] + e) * Real(1.0 / (2 * m + 1));
        }
    } else {
        Real rT = MathOps<Real>::rcp(T);
        F[0] = Real(0.886226925452758013649083741671) * MathOps<Real>::sqrt_(rT);  // sqrt(pi)/2 / sqrt(T)
        if (M > 0) {
            Real h = Real(0.5) * rT;
#pragma unroll
            for (int m = 0; m < M; ++m) F[m + 1] = (Real(2 * m + 1) * F[m] - e) * h;
        }
    }
}

// ------------------------------------------------------------------ Hermite E
// 1-D expansion coefficients for one primitive pair along one axis:
//   E[a][b][t], a <= LA, b <= LB, t <= a+b, with E^{00}_0 = 1 (the Gaussian
//   product prefactor exp(-mu X_AB^2) is folded into the pair weight).
//   E^{a+1,b}_t = E^{ab}_{t-1}/(2p) + X_PA E^{ab}_t + (t+1) E^{ab}_{t+1}
//   E^{a,b+1}_t = E^{ab}_{t-1}/(2p) + X_PB E^{ab}_t + (t+1) E^{ab}_{t+1}
// (integrals.hpp:269-322 with X_PA = -(beta/p) X_AB and X_PB = (alpha/p) X_AB.)
template <int LA, int LB, class Real>
struct Herm1D {
    Real e[LA + 1][LB + 1][LA + LB + 1];
    HFX_HD __forceinline__ void build(Real i2p, Real PA, Real PB) {
#pragma unroll
        for (int a = 0; a <= LA; ++a)
#pragma unroll
            for (int b = 0; b <= LB; ++b)
#pragma unroll
                for (int t = 0; t <= LA + LB; ++t) e[a][b][t] = Real(0);
        e[0][0][0] = Real(1);
#pragma unroll
        for (int a = 1; a <= LA; ++a)
#pragma unroll
            for (int t = 0; t <= a; ++t) {
                Real v = PA * e[a - 1][0][t];
                if (t > 0) v += i2p * e[a - 1][0][t - 1];
                if (t + 1 <= a - 1) v += Real(t + 1) * e[a - 1][0][t + 1];
                e[a][0][t] = v;
            }
#pragma unroll
        for (int b = 1; b <= LB; ++b)
#pragma unroll
            for (int a = 0; a <= LA; ++a)
#pragma unroll
                for (int t = 0; t <= a + b; ++t) {
                    Real v = (t <= a + b - 1) ? PB * e[a][b - 1][t] : Real(0);
                    if (t > 0) v += i2p * e[a][b - 1][t - 1];
                    if (t + 1 <= a + b - 1) v += Real(t + 1) * e[a][b - 1][t + 1];
                    e[a][b][t] = v;
                }
    }
};

// ------------------------------------------------------------------ R_tuv
// R[hidx(t,u,v)] = R^{(0)}_{tuv} for t+u+v <= L, given a[n] = pref*(-2 rho)^n F_n
// for n = 0..L.  In-place downward sweep: at level n the orders are updated
// from high to low, so every read still sees level n+1.
template <int L, class Real>
HFX_HD __forceinline__ void hermite_R(const Real a[L + 1], Real X, Real Y, Real Z, Real R[nherm(L)]) {
    R[0] = a[L];
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
#pragma unroll
        for (int s = L - n; s >= 1; --s) {
#pragma unroll
            for (int t = s; t >= 0; --t) {
#pragma unroll
                for (int u = s - t; u >= 0; --u) {
                    const int v = s - t - u;
                    Real val;
                    if (t > 0) {
                        val = X * R[hidx(t - 1, u, v)];
                        if (t > 1) val += Real(t - 1) * R[hidx(t - 2, u, v)];
                    } else if (u > 0) {
                        val = Y * R[hidx(t, u - 1, v)];
                        if (u > 1) val += Real(u - 1) * R[hidx(t, u - 2, v)];
                    } else {
                        val = Z * R[hidx(t, u, v - 1)];
                        if (v > 1) val += Real(v - 1) * R[hidx(t, u, v - 2)];
                    }
                    R[hidx(t, u, v)] = val;
                }
            }
        }
        R[0] = a[n];
    }
}

}  // namespace hfx
