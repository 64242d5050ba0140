// oneint_device.hip -- one-electron integrals S, T, V on the device.
//
// Replaces the reference's host overlap_matrix / kinetic_matrix /
// attraction_matrix (include/integrals.hpp:99-118, 199-218, 516-536) with the
// same contracted quantities sum_ij c_i c_j N_i N_j (primitive integral):
//   S  = K_AB (pi/p)^{3/2} E^{lx,mx}_0 E^{ly,my}_0 E^{lz,mz}_0
//   T  = sum over axes of T1_x S1_y S1_z with the 1-D identity
//        T1(l,m) = b(2m+1) S1(l,m) - 2b^2 S1(l,m+2) - m(m-1)/2 S1(l,m-2)
//   V  = -sum_C Z_C (2 pi/p) K_AB sum_tuv E_t E_u E_v R_tuv(p, P - C)
// (McMurchie-Davidson; E^{00}_0 = 1 with K_AB = exp(-ab/p |AB|^2) applied
// once; R from the FP64 Boys table + downward recursion of the ERI kernels).
//
// One lane per function pair (m >= n in the higher-L-first order of its pair
// class (LA, LB)); the 1-D E rows of the lane's own components are built by
// the recursion in compile-time-bounded loops, so every array index is a
// constant and nothing spills.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../../include/hfx.h"
#include "devmem.hpp"
#include "eri_body.hpp"

namespace hfx_detail {
void set_error(const std::string& msg);
const double* device_boys_table();
}  // namespace hfx_detail

namespace hfx {
namespace {

constexpr double PI = 3.14159265358979323846;

// E^{l,b}_t over t for one axis: rows for b = m (all t) and the t = 0 values
// at b = m - 2, m, m + 2.  l <= LA, m <= LB at run time.
template <int LA, int LB>
struct ERow {
    static constexpr int NT = LA + LB + 3;
    double row[LA + LB + 1];   // E^{l,m}_t, t <= l + m
    double s_m2, s_0, s_p2;    // E^{l,m-2}_0, E^{l,m}_0, E^{l,m+2}_0
    __device__ __forceinline__ void build(int l, int m, double PA, double PB, double i2p) {
        double cur[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) cur[t] = t == 0 ? 1.0 : 0.0;
        // a-steps: E^{a,0}_t = i2p E^{a-1,0}_{t-1} + PA E^{a-1,0}_t + (t+1) E^{a-1,0}_{t+1}
#pragma unroll
        for (int a = 1; a <= LA; ++a) {
            if (a > l) break;
            double nxt[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                double v = PA * cur[t];
                if (t > 0) v += i2p * cur[t - 1];
                if (t + 1 < NT) v += (t + 1) * cur[t + 1];
                nxt[t] = v;
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) cur[t] = nxt[t];
        }
        s_m2 = m == 2 ? cur[0] : 0.0;   // E^{l,0}_0 when m - 2 = 0
        s_0 = cur[0];
        s_p2 = 0.0;
#pragma unroll
        for (int t = 0; t <= LA + LB; ++t) row[t] = m == 0 ? cur[t] : 0.0;
        // b-steps: E^{l,b}_t = i2p E^{l,b-1}_{t-1} + PB E^{l,b-1}_t + (t+1) E^{l,b-1}_{t+1}
#pragma unroll
        for (int b = 1; b <= LB + 2; ++b) {
            if (b > m + 2) break;
            double nxt[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                double v = PB * cur[t];
                if (t > 0) v += i2p * cur[t - 1];
                if (t + 1 < NT) v += (t + 1) * cur[t + 1];
                nxt[t] = v;
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) cur[t] = nxt[t];
            if (b == m - 2) s_m2 = cur[0];
            if (b == m) {
                s_0 = cur[0];
#pragma unroll
                for (int t = 0; t <= LA + LB; ++t) row[t] = cur[t];
            }
            if (b == m + 2) s_p2 = cur[0];
        }
    }
};

struct OneIntArgs {
    const int2* pairs;     // (m, n) with L(m) = LA, L(n) = LB
    int npair;
    int K;
    const double* xyz;     // 3K
    const int* lmn;        // 3K
    const int* poff;       // K+1
    const double* alpha;
    const double* coeff;
    const double* norm;
    int natom;
    const double* Zn;      // natom charges
    const double* Rn;      // 3 natom
    const double* boys;
    double* S;
    double* T;
    double* V;
};

template <int LA, int LB>
__global__ __launch_bounds__(128) void k_oneint(OneIntArgs g) {
    constexpr int L = LA + LB;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.npair) return;
    const int m = g.pairs[idx].x, n = g.pairs[idx].y;
    int lm[3], ln[3];
    double A[3], B[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        lm[x] = g.lmn[3 * m + x];
        ln[x] = g.lmn[3 * n + x];
        A[x] = g.xyz[3 * m + x];
        B[x] = g.xyz[3 * n + x];
    }
    const double AB2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
    double sS = 0.0, sT = 0.0, sV = 0.0;
    for (int i = g.poff[m]; i < g.poff[m + 1]; ++i) {
        for (int j = g.poff[n]; j < g.poff[n + 1]; ++j) {
            const double a = g.alpha[i], b = g.alpha[j], p = a + b, inv = 1.0 / p;
            const double w = g.coeff[i] * g.coeff[j] * g.norm[i] * g.norm[j] * exp(-a * b * inv * AB2);
            const double i2p = 0.5 * inv;
            double P[3];
            ERow<LA, LB> E[3];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                P[x] = (a * A[x] + b * B[x]) * inv;
                E[x].build(lm[x], ln[x], P[x] - A[x], P[x] - B[x], i2p);
            }
            // overlap and kinetic
            const double sq = sqrt(PI * inv);
            double S1[3], T1[3];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const int k = ln[x];
                S1[x] = E[x].s_0 * sq;
                T1[x] = b * (2 * k + 1) * S1[x] - 2.0 * b * b * E[x].s_p2 * sq - 0.5 * k * (k - 1) * E[x].s_m2 * sq;
            }
            sS += w * S1[0] * S1[1] * S1[2];
            sT += w * (T1[0] * S1[1] * S1[2] + S1[0] * T1[1] * S1[2] + S1[0] * S1[1] * T1[2]);
            // nuclear attraction
            double tot = 0.0;
            for (int c = 0; c < g.natom; ++c) {
                const double X = P[0] - g.Rn[3 * c], Y = P[1] - g.Rn[3 * c + 1], Z = P[2] - g.Rn[3 * c + 2];
                double F[L + 1], an[L + 1], R[nherm(L)];
                boys_fm<L, double>(p * (X * X + Y * Y + Z * Z), g.boys, F);
                double f = 1.0;
#pragma unroll
                for (int q = 0; q <= L; ++q) {
                    an[q] = f * F[q];
                    f *= -2.0 * p;
                }
                hermite_R<L, double>(an, X, Y, Z, R);
                double s = 0.0;
                static_for<0, L + 1>([&](auto ti) {
                    constexpr int t = decltype(ti)::value;
                    static_for<0, L + 1 - t>([&](auto ui) {
                        constexpr int u = decltype(ui)::value;
                        static_for<0, L + 1 - t - u>([&](auto vi) {
                            constexpr int v = decltype(vi)::value;
                            s += E[0].row[t] * E[1].row[u] * E[2].row[v] * R[hidx(t, u, v)];
                        });
                    });
                });
                tot -= g.Zn[c] * s;
            }
            sV += w * 2.0 * PI * inv * tot;
        }
    }
    const size_t K = (size_t)g.K;
    if (g.S) g.S[m * K + n] = g.S[n * K + m] = sS;
    if (g.T) g.T[m * K + n] = g.T[n * K + m] = sT;
    if (g.V) g.V[m * K + n] = g.V[n * K + m] = sV;
}

template <int LA, int LB>
hipError_t launch_oneint(const OneIntArgs& g, hipStream_t st) {
    if (g.npair <= 0) return hipSuccess;
    k_oneint<LA, LB><<<(g.npair + 127) / 128, 128, 0, st>>>(g);
    return hipGetLastError();
}

hipError_t launch_pair_class(int la, int lb, const OneIntArgs& g, hipStream_t st) {
    switch (la * 4 + lb) {
        case 0: return launch_oneint<0, 0>(g, st);
        case 4: return launch_oneint<1, 0>(g, st);
        case 5: return launch_oneint<1, 1>(g, st);
        case 8: return launch_oneint<2, 0>(g, st);
        case 9: return launch_oneint<2, 1>(g, st);
        case 10: return launch_oneint<2, 2>(g, st);
        case 12: return launch_oneint<3, 0>(g, st);
        case 13: return launch_oneint<3, 1>(g, st);
        case 14: return launch_oneint<3, 2>(g, st);
        case 15: return launch_oneint<3, 3>(g, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace
}  // namespace hfx

extern "C" int hfx_one_electron_device(int K, const double* xyz, const int* lmn, const int* poff, const double* alpha,
                                       const double* coeff, const double* norm, int natom, const double* Z,
                                       const double* Rn, double* S, double* T, double* V) {
    using namespace hfx;
    if (K <= 0 || !xyz || !lmn || !poff || !alpha || !coeff || !norm || natom < 0 || (natom > 0 && (!Z || !Rn))) {
        hfx_detail::set_error("hfx_one_electron_device: invalid argument");
        return HFX_ERR_INVALID;
    }
    const double* boys = hfx_detail::device_boys_table();
    if (!boys) return HFX_ERR_DEVICE;
    // function pairs by class, higher L first (the matrices are symmetric)
    std::vector<int> Lf(K);
    for (int f = 0; f < K; ++f) {
        Lf[f] = lmn[3 * f] + lmn[3 * f + 1] + lmn[3 * f + 2];
        if (Lf[f] > 3) {
            hfx_detail::set_error("hfx_one_electron_device: angular momentum above f");
            return HFX_ERR_UNSUPPORTED;
        }
    }
    std::vector<int2> pairs[16];
    for (int m = 0; m < K; ++m)
        for (int n = 0; n <= m; ++n) {
            int a = m, b = n;
            if (Lf[a] < Lf[b]) std::swap(a, b);
            pairs[Lf[a] * 4 + Lf[b]].push_back(make_int2(a, b));
        }
    const int np = poff[K];
    const size_t K2 = (size_t)K * K;
    std::vector<size_t> off(17, 0);
    for (int c = 0; c < 16; ++c) off[c + 1] = off[c] + pairs[c].size();
    hipStream_t st = nullptr;
    char* buf = nullptr;
    // one device block: pairs | xyz | lmn | poff | alpha | coeff | norm | Z | R | S | T | V
    const size_t b_pairs = off[16] * sizeof(int2), b_xyz = 3 * (size_t)K * 8, b_lmn = 3 * (size_t)K * 4,
                 b_poff = (size_t)(K + 1) * 4, b_prim = (size_t)np * 8, b_nuc = (size_t)natom * 8, b_mat = K2 * 8;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t o = 0;
    const size_t o_pairs = o; o += al(b_pairs);
    const size_t o_xyz = o; o += al(b_xyz);
    const size_t o_lmn = o; o += al(b_lmn);
    const size_t o_poff = o; o += al(b_poff);
    const size_t o_al = o; o += al(b_prim);
    const size_t o_co = o; o += al(b_prim);
    const size_t o_no = o; o += al(b_prim);
    const size_t o_Z = o; o += al(b_nuc);
    const size_t o_R = o; o += al(3 * b_nuc);
    const size_t o_S = o; o += al(b_mat);
    const size_t o_T = o; o += al(b_mat);
    const size_t o_V = o; o += al(b_mat);
    std::vector<int2> allp;
    allp.reserve(off[16]);
    for (int c = 0; c < 16; ++c) allp.insert(allp.end(), pairs[c].begin(), pairs[c].end());
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hfx_detail::dev_alloc((void**)&buf, o);
    auto h2d = [&](size_t at, const void* src, size_t n) {
        if (e == hipSuccess && n) e = hipMemcpyAsync(buf + at, src, n, hipMemcpyHostToDevice, st);
    };
    h2d(o_pairs, allp.data(), b_pairs);
    h2d(o_xyz, xyz, b_xyz);
    h2d(o_lmn, lmn, b_lmn);
    h2d(o_poff, poff, b_poff);
    h2d(o_al, alpha, b_prim);
    h2d(o_co, coeff, b_prim);
    h2d(o_no, norm, b_prim);
    h2d(o_Z, Z, b_nuc);
    h2d(o_R, Rn, 3 * b_nuc);
    OneIntArgs g;
    g.K = K;
    g.xyz = (const double*)(buf + o_xyz);
    g.lmn = (const int*)(buf + o_lmn);
    g.poff = (const int*)(buf + o_poff);
    g.alpha = (const double*)(buf + o_al);
    g.coeff = (const double*)(buf + o_co);
    g.norm = (const double*)(buf + o_no);
    g.natom = natom;
    g.Zn = (const double*)(buf + o_Z);
    g.Rn = (const double*)(buf + o_R);
    g.boys = boys;
    g.S = (double*)(buf + o_S);
    g.T = (double*)(buf + o_T);
    g.V = (double*)(buf + o_V);
    for (int la = 0; la < 4 && e == hipSuccess; ++la)
        for (int lb = 0; lb <= la && e == hipSuccess; ++lb) {
            const int c = la * 4 + lb;
            g.pairs = (const int2*)(buf + o_pairs) + off[c];
            g.npair = (int)pairs[c].size();
            e = launch_pair_class(la, lb, g, st);
        }
    auto d2h = [&](void* dst, size_t at) {
        if (e == hipSuccess && dst) e = hipMemcpyAsync(dst, buf + at, b_mat, hipMemcpyDeviceToHost, st);
    };
    d2h(S, o_S);
    d2h(T, o_T);
    d2h(V, o_V);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hfx_detail::dev_free(buf);
    if (st) (void)hipStreamDestroy(st);
    if (e != hipSuccess) {
        hfx_detail::set_error(std::string("hfx_one_electron_device: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? HFX_ERR_NOMEM : HFX_ERR_DEVICE;
    }
    return 0;
}
