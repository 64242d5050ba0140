// eri_body.hpp -- bodies of the shell-quartet ERI kernels.
//
// Two kernel families, both McMurchie-Davidson with per-primitive-quartet
// Boys + R recursion and a two-stage Hermite -> Cartesian contraction:
//
//  * family A ("one lane, one shell quartet"): every lane owns one shell
//    quartet of a low-L class; all primitive loops, E tables, R_tuv and the
//    output block live in registers.  Used for classes whose block and R fit
//    the FP64 register budget (see use_family_a()).
//
//  * family B ("one wavefront, one shell quartet"): the 64 lanes of a
//    wavefront cooperate on one shell quartet.  Per primitive quartet the
//    Boys values and the R_tuv levels are built lane-parallel in LDS, then
//    the ket contraction g[tuv][cd] and the bra contraction acc[ab][cd] are
//    spread over lanes.  The ket components are processed in chunks so the
//    LDS footprint stays under 64 KiB even for (ff|ff).
//
// Bodies are __host__ __device__ so the identical code also builds into a
// CPU emulation library (lane loop with nlanes = 1) used by the CPU tests.
#pragma once
#include <stdint.h>
#include <type_traits>
#include <utility>

#include "eri_types.hpp"
#include "md_core.hpp"

#ifndef HFX_DENSE_R
#define HFX_DENSE_R 1   // dense last R level in family B (L <= 7); 0 = packed only
#endif

namespace hfx {

#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }
#else
inline void atomic_add(double* p, double v) { *p += v; }
#endif

constexpr double TWO_PI_POW_2_5 = 34.98683665524972497;   // 2 pi^{5/2}

template <int Begin, int End, class F>
HFX_HD __forceinline__ void static_for(F&& f) {
    if constexpr (Begin < End) {
        f(std::integral_constant<int, Begin>{});
        static_for<Begin + 1, End>(f);
    }
}

// Family A is used while the output block, R_tuv and the E tables stay within
// about 96 FP64 registers' worth of values.  Measured: moving the classes with
// 64 < nout + nherm(L) <= 128 ((pp|pp), (ds|pp), (dp|ps), ...) to family B
// slows the benzene unique build 1.37x although family A runs them at one or
// two resident waves per SIMD.
__host__ __device__ constexpr bool use_family_a(int la, int lb, int lc, int ld) {
    int L = la + lb + lc + ld;
    int nout = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
    return L <= 4 && nout + nherm(L) <= 128;
}

// Family M (eri_mfma.hpp: both Hermite -> Cartesian contractions on the
// matrix cores) runs the classes whose ket Hermite and Cartesian sizes (at
// least 4 and 3: a ket of p or more) keep the MFMA tiles from being mostly
// padding: every such class beyond family A's reach, and the family-A class
// (pp|pp), which it runs 2x faster in the benzene / cc-pVDZ unique build
// (measured; (ds|ds) ran 1.5x slower there and stays in family A).  Moving the
// (xx|ps) classes from family B (6 ket components required before) took the
// benzene build 2.31 -> 2.19 ms and the (H2O)4 direct build 33.7 -> 31.8 ms.
__host__ __device__ constexpr bool use_family_m(int la, int lb, int lc, int ld) {
    if (nherm(lc + ld) < 4 || ncart(lc) * ncart(ld) < 3) return false;
    if (!use_family_a(la, lb, lc, ld)) return true;
    return la == 1 && lb == 1 && lc == 1 && ld == 1;
}

// The class runs the family-A kernel (else family B or M, which share the
// work-unit semantics of famb_range).
__host__ __device__ constexpr bool runs_family_a(int la, int lb, int lc, int ld) {
    return use_family_a(la, lb, lc, ld) && !use_family_m(la, lb, lc, ld);
}

// ------------------------------------------------------------------ epilogue
// A shell quartet block v[a][b][c][d] for the kernel-order shells
// (s1 s2 | s3 s4) is consumed in one of three ways: the Schwarz diagonal and
// the unique array take the final values (raw block times the four function
// scales); EPI_JK takes the raw block, the scales being folded into its
// densities and outputs (F = diag(fscale): J = F J'(F Dj F) F, the same for K).
//
// Degeneracy for EPI_JK: with f = 1/|stabiliser| of the shell quartet under
// the 8-fold permutation group, summing all 8 images over the full block is
//   Jacc_ab += 2f sum_cd v Dj_cd ; Jacc_cd += 2f sum_ab v Dj_ab
//   Kacc_ac += f sum_bd v D_bd ; Kacc_bc += f sum_ad v D_ad
//   Kacc_ad += f sum_bc v D_bc ; Kacc_bd += f sum_ac v D_ac
// and J = Jacc + Jacc^T, K = Kacc + Kacc^T afterwards.
struct QuartetCtx {
    int fn1, fn2, fn3, fn4;
    bool d12, d34, same;   // s1 == s2, s3 == s4, bra pair == ket pair
};

// Canonical filter: each unique function quartet is emitted exactly once.
HFX_HD __forceinline__ bool keep_unique(const QuartetCtx& q, int mu, int nu, int la, int si) {
    if (q.d12 && mu < nu) return false;
    if (q.d34 && la < si) return false;
    if (q.same && pair_of(mu, nu) < pair_of(la, si)) return false;
    return true;
}

HFX_HD __forceinline__ bool mask_pass(const EpiArgs& ep, int mu, int nu, int la, int si) {
    return ep.bounds[pair_of(mu, nu)] * ep.bounds[pair_of(la, si)] >= ep.threshold;
}

// ------------------------------------------------------------------ family A
template <int LA, int LB, int LC, int LD, class Real>
struct FamA {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
    static constexpr int NAB = NA * NB, NCD = NC * ND, NOUT = NAB * NCD;

    // acc[ab*NCD + cd] += sum_tuv Ebra_ab(tuv) sum_{tau nu phi} Eket_cd(tau nu phi) R(t+tau, u+nu, v+phi)
    static HFX_HD __forceinline__ void contract(Real* acc, const Herm1D<LA, LB, Real>& ex,
                                                const Herm1D<LA, LB, Real>& ey, const Herm1D<LA, LB, Real>& ez,
                                                const Herm1D<LC, LD, Real>& fx, const Herm1D<LC, LD, Real>& fy,
                                                const Herm1D<LC, LD, Real>& fz, const Real* R) {
        static_for<0, NC>([&](auto ci) {
            static_for<0, ND>([&](auto di) {
                constexpr int c = decltype(ci)::value, d = decltype(di)::value;
                constexpr int cx = cart_x(LC, c), cy = cart_y(LC, c), cz = cart_z(LC, c);
                constexpr int dx = cart_x(LD, d), dy = cart_y(LD, d), dz = cart_z(LD, d);
                Real g[nherm(LAB)];
                static_for<0, LAB + 1>([&](auto ti) {
                    static_for<0, LAB + 1 - decltype(ti)::value>([&](auto ui) {
                        static_for<0, LAB + 1 - decltype(ti)::value - decltype(ui)::value>([&](auto vi) {
                            constexpr int t = decltype(ti)::value, u = decltype(ui)::value, v = decltype(vi)::value;
                            Real s = Real(0);
                            static_for<0, cx + dx + 1>([&](auto tai) {
                                constexpr int ta = decltype(tai)::value;
                                Real sy = Real(0);
                                static_for<0, cy + dy + 1>([&](auto nui) {
                                    constexpr int nn = decltype(nui)::value;
                                    Real sz = Real(0);
                                    static_for<0, cz + dz + 1>([&](auto phi) {
                                        constexpr int ph = decltype(phi)::value;
                                        sz += fz.e[cz][dz][ph] * R[hidx(t + ta, u + nn, v + ph)];
                                    });
                                    sy += fy.e[cy][dy][nn] * sz;
                                });
                                s += fx.e[cx][dx][ta] * sy;
                            });
                            g[hidx(t, u, v)] = s;
                        });
                    });
                });
                static_for<0, NA>([&](auto ai) {
                    static_for<0, NB>([&](auto bi) {
                        constexpr int a = decltype(ai)::value, b = decltype(bi)::value;
                        constexpr int ax = cart_x(LA, a), ay = cart_y(LA, a), az = cart_z(LA, a);
                        constexpr int bx = cart_x(LB, b), by = cart_y(LB, b), bz = cart_z(LB, b);
                        Real s = Real(0);
                        static_for<0, ax + bx + 1>([&](auto ti) {
                            constexpr int t = decltype(ti)::value;
                            Real sy = Real(0);
                            static_for<0, ay + by + 1>([&](auto ui) {
                                constexpr int u = decltype(ui)::value;
                                Real sz = Real(0);
                                static_for<0, az + bz + 1>([&](auto vi) {
                                    constexpr int v = decltype(vi)::value;
                                    sz += ez.e[az][bz][v] * g[hidx(t, u, v)];
                                });
                                sy += ey.e[ay][by][u] * sz;
                            });
                            s += ex.e[ax][bx][t] * sy;
                        });
                        acc[(a * NB + b) * NCD + c * ND + d] += s;
                    });
                });
            });
        });
    }

    // Primitive loops of one shell quartet on a group of G lanes.  The slice's
    // bra rows are the bra primitive pairs slice, slice + nslice, ...  acc must
    // be zeroed and holds this lane's partial sum.
    //  * L <= 1 ((ss|ss), (ps|ss)): the lanes form a gi x gj grid (gi gj = G);
    //    lane sub takes rows sub / gj, + gi, ... and, for each, ket pairs
    //    sub % gj, + gj, ..., so a bra pair's data and E tables are loaded once
    //    per row (these kernels are bound by vector-load address processing);
    //    (ss|ss) takes two rows per ket load.  gi is the power of two with the
    //    fewest steps per lane (ties: the larger gi, fewer bra loads).
    //  * otherwise the flattened (row, ket) index f = sub, sub + G, ... (no idle
    //    lanes; measured faster for the classes with larger blocks).
    // One primitive quartet (bra pair b, ket pair k with their E tables, the
    // ket's odd Hermite orders negated) added into acc.
    static HFX_HD __forceinline__ void prim(const PrimPair<Real>& b, const Herm1D<LA, LB, Real>& ex,
                                            const Herm1D<LA, LB, Real>& ey, const Herm1D<LA, LB, Real>& ez,
                                            const PrimPair<Real>& k, const Herm1D<LC, LD, Real>& fx,
                                            const Herm1D<LC, LD, Real>& fy, const Herm1D<LC, LD, Real>& fz,
                                            const Real* tab, Real* acc) {
        Real s = b.p + k.p;
        Real inv = Real(1) / s;
        Real rho = b.p * k.p * inv;
        Real X = b.P[0] - k.P[0], Y = b.P[1] - k.P[1], Z = b.P[2] - k.P[2];
        Real T = rho * (X * X + Y * Y + Z * Z);
        Real F[L + 1];
        boys_fm<L, Real>(T, tab, F);
        Real pref = Real(TWO_PI_POW_2_5) * b.w * k.w * MathOps<Real>::sqrt_(inv);
        Real a[L + 1];
        Real m2r = Real(-2) * rho;
        Real fct = pref;
#pragma unroll
        for (int n = 0; n <= L; ++n) {
            a[n] = fct * F[n];
            fct *= m2r;
        }
        Real R[nherm(L)];
        hermite_R<L, Real>(a, X, Y, Z, R);
        contract(acc, ex, ey, ez, fx, fy, fz, R);
    }

    static HFX_HD __forceinline__ void ket_tables(const PrimPair<Real>& k, Herm1D<LC, LD, Real>& fx,
                                                  Herm1D<LC, LD, Real>& fy, Herm1D<LC, LD, Real>& fz) {
        fx.build(k.i2p, k.PA[0], k.PB[0]);
        fy.build(k.i2p, k.PA[1], k.PB[1]);
        fz.build(k.i2p, k.PA[2], k.PB[2]);
        // (-1)^{tau+nu+phi}: negate odd Hermite orders of the ket tables
        static_for<0, LC + 1>([&](auto ci) {
            static_for<0, LD + 1>([&](auto di) {
                static_for<0, LCD + 1>([&](auto ti) {
                    constexpr int c = decltype(ci)::value, d = decltype(di)::value, t = decltype(ti)::value;
                    if constexpr (t % 2 == 1) {
                        fx.e[c][d][t] = -fx.e[c][d][t];
                        fy.e[c][d][t] = -fy.e[c][d][t];
                        fz.e[c][d][t] = -fz.e[c][d][t];
                    }
                });
            });
        });
    }

    static HFX_HD __forceinline__ void quartet(const PairHdr& hb, const PairHdr& hk, const PrimPair<Real>* pp,
                                               const Real* tab, Real* acc, int sub, int G, int slice, int nslice) {
        const int nk = hk.nprim;
        const int rows = (hb.nprim - slice + nslice - 1) / nslice;
        if constexpr (L <= 1) {
            // lane grid gi x gj; for (ss|ss) a lane step takes RB = 2 of the
            // lane's rows against one ket pair, so each ket load serves two
            // primitive quartets (for (ps|ss) the second row's registers cost
            // more occupancy than the loads save: RB = 1, measured):
            // steps = ceil(ceil(rows / gi) / RB) ceil(nk / gj)
            constexpr int RB = L == 0 ? 2 : 1;
            int gi = 1, best = 0x7fffffff;
            for (int g = 1; g <= G; g <<= 1) {
                const int gj = G / g;
                const int c = (((rows + g - 1) / g + RB - 1) / RB) * ((nk + gj - 1) / gj);
                if (c <= best) {
                    best = c;
                    gi = g;
                }
            }
            const int gj = G / gi;
            for (int r = sub / gj; r < rows; r += RB * gi) {
                const bool two = RB == 2 && r + gi < rows;
                const PrimPair<Real> b0 = pp[hb.prim_off + slice + r * nslice];
                PrimPair<Real> b1 = b0;
                if constexpr (RB == 2) b1 = pp[hb.prim_off + slice + (two ? r + gi : r) * nslice];
                Herm1D<LA, LB, Real> ex0, ey0, ez0, ex1, ey1, ez1;
                ex0.build(b0.i2p, b0.PA[0], b0.PB[0]);
                ey0.build(b0.i2p, b0.PA[1], b0.PB[1]);
                ez0.build(b0.i2p, b0.PA[2], b0.PB[2]);
                if constexpr (RB == 2) {
                    ex1.build(b1.i2p, b1.PA[0], b1.PB[0]);
                    ey1.build(b1.i2p, b1.PA[1], b1.PB[1]);
                    ez1.build(b1.i2p, b1.PA[2], b1.PB[2]);
                }
                for (int j = sub % gj; j < nk; j += gj) {
                    const PrimPair<Real> k = pp[hk.prim_off + j];
                    Herm1D<LC, LD, Real> fx, fy, fz;
                    ket_tables(k, fx, fy, fz);
                    prim(b0, ex0, ey0, ez0, k, fx, fy, fz, tab, acc);
                    if constexpr (RB == 2)
                        if (two) prim(b1, ex1, ey1, ez1, k, fx, fy, fz, tab, acc);
                }
            }
        } else {
            // flattened (row, ket) index f = sub, sub + G, ...
            const int tot = rows * nk;
            int r = nk > 0 ? sub / nk : 0, j = sub - r * nk;
            int last_r = -1;
            PrimPair<Real> b;
            Herm1D<LA, LB, Real> ex, ey, ez;
            for (int f = sub; f < tot; f += G) {
                if (r != last_r) {
                    b = pp[hb.prim_off + slice + r * nslice];
                    ex.build(b.i2p, b.PA[0], b.PB[0]);
                    ey.build(b.i2p, b.PA[1], b.PB[1]);
                    ez.build(b.i2p, b.PA[2], b.PB[2]);
                    last_r = r;
                }
                const PrimPair<Real> k = pp[hk.prim_off + j];
                Herm1D<LC, LD, Real> fx, fy, fz;
                ket_tables(k, fx, fy, fz);
                prim(b, ex, ey, ez, k, fx, fy, fz, tab, acc);
                j += G;
                if (j >= nk) {   // G may exceed nk when rows are short
                    const int dr = j / nk;
                    r += dr;
                    j -= dr * nk;
                }
            }
        }
    }
};

#if defined(__HIP_DEVICE_COMPILE__)
// Value of lane (lane - SHIFT) of the lane's 16-lane row (DPP row_shr), 0 when
// that lane lies in another row.
template <int SHIFT>
__device__ __forceinline__ double dpp_row_shr(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x110 + SHIFT, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x110 + SHIFT, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int LANE>
__device__ __forceinline__ double read_lane(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, LANE);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), LANE);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Segmented inclusive scan over the group leaders of one wavefront (lanes 0,
// G, 2G, ... for G <= 16).  A run starts at every leader whose `head` flag is
// set (lane 0 must set it) and ends before the next such leader; keys may
// recur in later runs.  After scan(x) the last leader of each run (`tail`)
// holds the sum of x over the run's leaders.  In-row steps use DPP row_shr by
// G, 2G, 4G, 8G; the rows are then chained through lanes 15, 31, 47 (for G = 1;
// 16 r - G in general) with readlane.  Every lane of the wavefront must be
// active, and x must be 0 on lanes that are not leaders.
template <int G>
struct LeaderScan {
    static constexpr int NIN = G == 1 ? 4 : G == 2 ? 3 : G == 4 ? 2 : G == 8 ? 1 : 0;
    double m[NIN > 0 ? NIN : 1];   // 1.0 where the in-row source leader is in the run
    double c[3];                   // 1.0 where the run reaches back to lane 16 r - G
    bool tail;
    __device__ __forceinline__ void init(bool head, int lane) {
        const bool lead = lane % G == 0;
        const unsigned long long H = __ballot(lead && head);
        const int first = 63 - __clzll(H & (~0ull >> (63 - lane)));   // first leader of the lane's run
#pragma unroll
        for (int k = 0; k < NIN; ++k) m[k] = (lane & 15) >= (G << k) && lane - (G << k) >= first ? 1.0 : 0.0;
#pragma unroll
        for (int r = 1; r <= 3; ++r) c[r - 1] = (lane >> 4) == r && first <= 16 * r - G ? 1.0 : 0.0;
        tail = lead && (lane + G >= 64 || ((H >> (lane + G)) & 1ull) != 0);
    }
    __device__ __forceinline__ double scan(double x) const {
        if constexpr (NIN > 0) x = fma(dpp_row_shr<G>(x), m[0], x);
        if constexpr (NIN > 1) x = fma(dpp_row_shr<2 * G>(x), m[1], x);
        if constexpr (NIN > 2) x = fma(dpp_row_shr<4 * G>(x), m[2], x);
        if constexpr (NIN > 3) x = fma(dpp_row_shr<8 * G>(x), m[3], x);
        x = fma(read_lane<16 - G>(x), c[0], x);
        x = fma(read_lane<32 - G>(x), c[1], x);
        x = fma(read_lane<48 - G>(x), c[2], x);
        return x;
    }
};
#endif

// J/K epilogue of a register block acc[NOUT] (raw, unscaled) for the
// kernel-order shell quartet q, against densities that carry the function
// scales.  Every operand is loaded before the first atomic (the compiler
// cannot move a load above an atomic that may alias it, so interleaving them
// would serialise one memory latency per term), then the block is masked once,
// the J and K sums formed in registers, and the atomics issued back to back.
//
// WAVE = false: one lane, its own quartet (`lead` must be true).
// WAVE = true (device, family A with G <= 16 lanes per unit): every lane of
// the wavefront calls; `lead` marks the group leaders whose unit has work.
// The planner orders a J/K list by (bra pair, ket first shell, ket pair)
// inside each primitive-work class, so neighbouring units of a wavefront
// share J_ab (key: bra pair) and K_ac, K_bc (key: bra pair and ket shell c)
// targets: those sums are added along runs of equal key (LeaderScan) and only
// the last unit of a run issues their atomics; J_cd, K_ad and K_bd stay one
// atomic per unit.  Units past the list end pass active = false.
template <int NA, int NB, int NC, int ND, class Real, int G, bool WAVE>
HFX_HD __forceinline__ void epilogue_jk(const QuartetCtx& q, const EpiArgs& ep, const Real* acc, bool lead,
                                        bool active, int bra, int cshell, int lane) {
    constexpr int NCD = NC * ND, NAB = NA * NB;
    const int K = ep.K;
    const double f = (q.d12 ? 0.5 : 1.0) * (q.d34 ? 0.5 : 1.0) * (q.same ? 0.5 : 1.0);
    double v[NAB * NCD];
    double jab[NAB], jcd[NCD];
    double dac[NA * NC], dad[NA * ND], dbc[NB * NC], dbd[NB * ND];
#pragma unroll
    for (int x = 0; x < NAB; ++x) jab[x] = 0.0;
#pragma unroll
    for (int y = 0; y < NCD; ++y) jcd[y] = 0.0;
    auto load_k = [&](const double* D) {
#pragma unroll
        for (int a = 0; a < NA; ++a) {
#pragma unroll
            for (int c = 0; c < NC; ++c) dac[a * NC + c] = D[(q.fn1 + a) * K + q.fn3 + c];
#pragma unroll
            for (int d = 0; d < ND; ++d) dad[a * ND + d] = D[(q.fn1 + a) * K + q.fn4 + d];
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
#pragma unroll
            for (int c = 0; c < NC; ++c) dbc[b * NC + c] = D[(q.fn2 + b) * K + q.fn3 + c];
#pragma unroll
            for (int d = 0; d < ND; ++d) dbd[b * ND + d] = D[(q.fn2 + b) * K + q.fn4 + d];
        }
    };
    if (lead) {
        double bab[NAB], bcd[NCD], djab[NAB], djcd[NCD];
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                bab[a * NB + b] = ep.bounds[pair_of(q.fn1 + a, q.fn2 + b)];
                djab[a * NB + b] = ep.Dj[(q.fn1 + a) * K + q.fn2 + b];
            }
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                bcd[c * ND + d] = ep.bounds[pair_of(q.fn3 + c, q.fn4 + d)];
                djcd[c * ND + d] = ep.Dj[(q.fn3 + c) * K + q.fn4 + d];
            }
        load_k(ep.Da);
        // the raw block: the function scales sit in the densities and outputs
        // (enqueue_direct scales them), only the Schwarz mask applies here
#pragma unroll
        for (int x = 0; x < NAB; ++x)
#pragma unroll
            for (int y = 0; y < NCD; ++y)
                v[x * NCD + y] = bab[x] * bcd[y] >= ep.threshold ? (double)acc[x * NCD + y] : 0.0;
#pragma unroll
        for (int x = 0; x < NAB; ++x)
#pragma unroll
            for (int y = 0; y < NCD; ++y) {
                jab[x] += v[x * NCD + y] * djcd[y];
                jcd[y] += v[x * NCD + y] * djab[x];
            }
#pragma unroll
        for (int x = 0; x < NAB; ++x) jab[x] *= 2.0 * f;
#pragma unroll
        for (int y = 0; y < NCD; ++y) jcd[y] *= 2.0 * f;
    } else {
#pragma unroll
        for (int e = 0; e < NAB * NCD; ++e) v[e] = 0.0;
    }
    bool tail_ab = lead, tail_c = lead;
#if defined(__HIP_DEVICE_COMPILE__)
    LeaderScan<G> sab, sc;
    if constexpr (WAVE) {
        const int pbra = __shfl_up(bra, G, 64), pc = __shfl_up(cshell, G, 64);
        const bool head_ab = lane < G || pbra != bra;
        sab.init(head_ab, lane);
        sc.init(head_ab || pc != cshell, lane);
        tail_ab = active && sab.tail;
        tail_c = active && sc.tail;
#pragma unroll
        for (int x = 0; x < NAB; ++x) jab[x] = sab.scan(jab[x]);
    }
#endif
    // K sums of one spin: Kacc_ac += f sum_bd v D_bd, Kacc_bc += f sum_ad v D_ad,
    // Kacc_ad += f sum_bc v D_bc, Kacc_bd += f sum_ac v D_ac
    auto k_terms = [&](double* Kacc) {
        double kac[NA * NC], kbc[NB * NC], kad[NA * ND], kbd[NB * ND];
#pragma unroll
        for (int x = 0; x < NA * NC; ++x) kac[x] = 0.0;
#pragma unroll
        for (int x = 0; x < NB * NC; ++x) kbc[x] = 0.0;
#pragma unroll
        for (int x = 0; x < NA * ND; ++x) kad[x] = 0.0;
#pragma unroll
        for (int x = 0; x < NB * ND; ++x) kbd[x] = 0.0;
        if (lead) {
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int c = 0; c < NC; ++c)
#pragma unroll
                        for (int d = 0; d < ND; ++d) {
                            const double x = v[(a * NB + b) * NCD + c * ND + d];
                            kac[a * NC + c] += x * dbd[b * ND + d];
                            kbc[b * NC + c] += x * dad[a * ND + d];
                            kad[a * ND + d] += x * dbc[b * NC + c];
                            kbd[b * ND + d] += x * dac[a * NC + c];
                        }
#pragma unroll
            for (int x = 0; x < NA * NC; ++x) kac[x] *= f;
#pragma unroll
            for (int x = 0; x < NB * NC; ++x) kbc[x] *= f;
#pragma unroll
            for (int x = 0; x < NA * ND; ++x) kad[x] *= f;
#pragma unroll
            for (int x = 0; x < NB * ND; ++x) kbd[x] *= f;
        }
#if defined(__HIP_DEVICE_COMPILE__)
        if constexpr (WAVE) {
#pragma unroll
            for (int x = 0; x < NA * NC; ++x) kac[x] = sc.scan(kac[x]);
#pragma unroll
            for (int x = 0; x < NB * NC; ++x) kbc[x] = sc.scan(kbc[x]);
        }
#endif
        if (tail_c) {
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int c = 0; c < NC; ++c) atomic_add(&Kacc[(q.fn1 + a) * K + q.fn3 + c], kac[a * NC + c]);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int c = 0; c < NC; ++c) atomic_add(&Kacc[(q.fn2 + b) * K + q.fn3 + c], kbc[b * NC + c]);
        }
        if (lead) {
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int d = 0; d < ND; ++d) atomic_add(&Kacc[(q.fn1 + a) * K + q.fn4 + d], kad[a * ND + d]);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int d = 0; d < ND; ++d) atomic_add(&Kacc[(q.fn2 + b) * K + q.fn4 + d], kbd[b * ND + d]);
        }
    };
    k_terms(ep.Ka);
    if (tail_ab) {
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) atomic_add(&ep.J[(q.fn1 + a) * K + q.fn2 + b], jab[a * NB + b]);
    }
    if (lead) {
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int d = 0; d < ND; ++d) atomic_add(&ep.J[(q.fn3 + c) * K + q.fn4 + d], jcd[c * ND + d]);
    }
    if (ep.nspin == 2) {
        if (lead) load_k(ep.Db);
        k_terms(ep.Kb);
    }
}

// Family-A epilogue over a register block acc[NOUT] (raw, unscaled).
template <int NA, int NB, int NC, int ND, class Real, bool JK>
HFX_HD __forceinline__ void epilogue_regs(const QuartetCtx& q, const EpiArgs& ep, const Real* acc, bool partial) {
    constexpr int NCD = NC * ND;
    const double* fs = ep.fscale;
    if constexpr (!JK) {
    if (ep.mode == EPI_UNIQUE) {
#pragma unroll
        for (int e = 0; e < NA * NB * NCD; ++e) {
            const int a = e / (NB * NCD), b = (e / NCD) % NB, c = (e / ND) % NC, d = e % ND;
            const int mu = q.fn1 + a, nu = q.fn2 + b, la = q.fn3 + c, si = q.fn4 + d;
            if (!keep_unique(q, mu, nu, la, si)) continue;
            const int64_t IJ = pair_of(mu, nu), KL = pair_of(la, si);
            if (ep.bounds[IJ] * ep.bounds[KL] < ep.threshold) continue;
            const double v = (double)acc[e] * (fs[mu] * fs[nu] * fs[la] * fs[si]);
            if (partial)
                atomic_add(&ep.out[quartet_of(IJ, KL)], v);
            else
                ep.out[quartet_of(IJ, KL)] = v;
        }
        return;
    }
    if (ep.mode == EPI_DIAG) {
        // bra pair == ket pair; (mu nu|mu nu) sits at a == c, b == d
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int mu = q.fn1 + a, nu = q.fn2 + b;
                if (q.d12 && mu < nu) continue;
                const int e = (a * NB + b) * NCD + a * ND + b;
                const double v = (double)acc[e] * (fs[mu] * fs[nu] * fs[mu] * fs[nu]);
                if (partial)
                    atomic_add(&ep.out[pair_of(mu, nu)], v);
                else
                    ep.out[pair_of(mu, nu)] = v;
            }
        return;
    }
    } else {
    epilogue_jk<NA, NB, NC, ND, Real, 1, false>(q, ep, acc, true, true, 0, 0, 0);
    }
}

// True when the density-difference bound of a unit is below ep.dthr (the J/K
// terms of (ab|cd) read the density blocks cd, ab, bd, bc, ad and ac).
HFX_HD __forceinline__ bool dens_skip(const EpiArgs& ep, uint2 e, const PairHdr& hb, const PairHdr& hk) {
    if (ep.dpmax == nullptr) return false;
    const int n = ep.nshell;
    const double* m = ep.dpmax;
    const int a = hb.s1, b = hb.s2, c = hk.s1, d = hk.s2;
    const double x = fmax(fmax(fmax(m[c * n + d], m[a * n + b]), fmax(m[b * n + d], m[b * n + c])),
                          fmax(m[a * n + d], m[a * n + c]));
    return ep.qpair[unit_bra(e)] * ep.qpair[unit_ket(e)] * x < ep.dthr;
}

HFX_HD __forceinline__ QuartetCtx make_ctx(const PairHdr& hb, const PairHdr& hk) {
    QuartetCtx q;
    q.fn1 = hb.fn1; q.fn2 = hb.fn2; q.fn3 = hk.fn1; q.fn4 = hk.fn2;
    q.d12 = hb.diag != 0; q.d34 = hk.diag != 0; q.same = hb.canon == hk.canon;
    return q;
}

// One work unit per group of G lanes (G = 1, 4, 16 or 64, a power of two
// dividing the wavefront).  `sub` is the lane's index inside its group; every
// lane of the group must call this (inactive groups pass active = false) so
// the cross-lane reduction sees the whole group.  Lane 0 of the group runs the
// epilogue.  A unit that is slice s of S covers the bra primitive pairs s,
// s + S, ... against all ket pairs, and its epilogue adds a partial block
// (J/K and the unique array are linear in it).
template <int LA, int LB, int LC, int LD, class Real, int G, bool JK>
HFX_HD __forceinline__ void family_a_group(uint2 e, bool active, int sub, const PairHdr* hdr, const PrimPair<Real>* pp,
                                           const Real* tab, const EpiArgs& ep) {
    using F = FamA<LA, LB, LC, LD, Real>;
    const PairHdr hb = hdr[unit_bra(e)], hk = hdr[unit_ket(e)];
    const int slice = unit_slice(e), nslice = unit_nslice(e);
    Real acc[F::NOUT];
#pragma unroll
    for (int i = 0; i < F::NOUT; ++i) acc[i] = Real(0);
    const bool work = active && slice < hb.nprim && !dens_skip(ep, e, hb, hk);
    if (work) F::quartet(hb, hk, pp, tab, acc, sub, G, slice, nslice);
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (G > 1) {
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
            for (int i = 0; i < F::NOUT; ++i) acc[i] += __shfl_xor(acc[i], off, G);
    }
#endif
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (JK && G < 64) {
        epilogue_jk<F::NA, F::NB, F::NC, F::ND, Real, G, true>(make_ctx(hb, hk), ep, acc, work && sub == 0, active,
                                                               active ? (int)unit_bra(e) : -1, active ? hk.s1 : -1,
                                                               (int)__lane_id());
        return;
    }
#endif
    if (work && sub == 0)
        epilogue_regs<F::NA, F::NB, F::NC, F::ND, Real, JK>(make_ctx(hb, hk), ep, acc, nslice > 1);
}

template <int LA, int LB, int LC, int LD, class Real>
HFX_HD __forceinline__ void family_a_one(uint2 e, const PairHdr* hdr, const PrimPair<Real>* pp, const Real* tab,
                                         const EpiArgs& ep) {
    if (ep.mode == EPI_JK)
        family_a_group<LA, LB, LC, LD, Real, 1, true>(e, true, 0, hdr, pp, tab, ep);
    else
        family_a_group<LA, LB, LC, LD, Real, 1, false>(e, true, 0, hdr, pp, tab, ep);
}

#if !defined(__HIP_DEVICE_COMPILE__)
// CPU emulation of a G-lane group: the G partial sums of the unit are formed
// one lane at a time and added, exactly the partition the device uses.
template <int LA, int LB, int LC, int LD, class Real>
inline void family_a_group_emu(uint2 e, int G, const PairHdr* hdr, const PrimPair<Real>* pp, const Real* tab,
                               const EpiArgs& ep) {
    using F = FamA<LA, LB, LC, LD, Real>;
    const PairHdr hb = hdr[unit_bra(e)], hk = hdr[unit_ket(e)];
    const int slice = unit_slice(e), nslice = unit_nslice(e);
    if (slice >= hb.nprim || dens_skip(ep, e, hb, hk)) return;
    Real tot[F::NOUT] = {};
    for (int sub = 0; sub < G; ++sub) {
        Real acc[F::NOUT] = {};
        F::quartet(hb, hk, pp, tab, acc, sub, G, slice, nslice);
        for (int i = 0; i < F::NOUT; ++i) tot[i] += acc[i];
    }
    if (ep.mode == EPI_JK)
        epilogue_regs<F::NA, F::NB, F::NC, F::ND, Real, true>(make_ctx(hb, hk), ep, tot, nslice > 1);
    else
        epilogue_regs<F::NA, F::NB, F::NC, F::ND, Real, false>(make_ctx(hb, hk), ep, tot, nslice > 1);
}
#endif

// ------------------------------------------------------------------ family B
// Ket primitive pairs per step of a family-B class of total angular momentum L.
__host__ __device__ constexpr int famb_kb(int L) { return (HFX_DENSE_R && L <= 7) ? 2 : (L <= 8 ? 4 : 1); }

// Range of primitive pairs a family-B work unit covers.  Slice s of S with
// S <= nb (bra primitive pairs): bra pairs s, s + S, ... against all ket
// pairs.  S > nb (the planner then makes S a multiple of nb): bra pair s % nb
// against ket chunk s / nb of S / nb chunks, each chunk a whole number of
// steps (kb pairs).
struct FamBRange {
    int i0, istep, j0, j1;
};
__host__ __device__ inline FamBRange famb_range(int slice, int nslice, int nb, int nk, int kb) {
    FamBRange r{slice, nslice, 0, nk};
    if (nslice > nb) {
        const int nkc = nslice / nb, kc = slice / nb;
        const int per = (nk + nkc - 1) / nkc;
        const int chunk = (per + kb - 1) / kb * kb;
        r.i0 = slice % nb;
        r.istep = nb;
        r.j0 = kc * chunk;
        r.j1 = r.j0 + chunk < nk ? r.j0 + chunk : nk;
        if (kc >= nkc) r.i0 = nb;   // no work
    }
    return r;
}

// Shared-memory footprint is sized per class; the ket component pairs (cd)
// are processed in chunks of CDC columns.
template <int LA, int LB, int LC, int LD>
struct FamBShape {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
    static constexpr int NAB = NA * NB, NCD = NC * ND;
    static constexpr int NHAB = nherm(LAB), NR = nherm(L);
    // L <= 7: R lives in one dense cube per ket primitive pair, R3[(t D + u) D + v]
    // with D = L + 1, updated in place level by level (orders high to low), so
    // the ket contraction walks it with constant strides; two ket primitive
    // pairs per step keep the LDS block small (resident wavefronts per CU are
    // bounded by LDS).  Larger L: packed ping-pong levels, four (L = 8) or one
    // pair per step.
    static constexpr bool DENSE = HFX_DENSE_R && L <= 7;
    static constexpr int D = L + 1, ND3 = DENSE ? D * D * D : 1, NRP = DENSE ? 1 : NR;
    static constexpr int KB = famb_kb(L);   // ket primitive pairs per step
    // epilogue staging (doubles): per-function scales of the four shells, the
    // bra / ket function-pair bounds, the J densities over ab and cd, and per
    // spin the four exchange density blocks D_ac, D_ad, D_bc, D_bd
    static constexpr int O_FS = 0, O_BAB = NA + NB + NC + ND, O_BCD = O_BAB + NAB, O_DJAB = O_BCD + NCD,
                         O_DJCD = O_DJAB + NAB, O_DK = O_DJCD + NCD, NDK = (NA + NB) * (NC + ND),
                         NEPI = O_DK + 2 * NDK;
    static constexpr int BUDGET = 5600 - (DENSE ? KB * ND3 : (KB - 1) * 2 * NR) - NEPI;   // doubles for g + acc
    static constexpr int CDC0 = BUDGET / (NHAB + NAB);
    static constexpr int CDC = CDC0 < NCD ? CDC0 : NCD;
    static constexpr int NCHUNK = (NCD + CDC - 1) / CDC;
    static constexpr int EB = 3 * (LA + 1) * (LB + 1) * (LAB + 1);
    static constexpr int EK = 3 * (LC + 1) * (LD + 1) * (LCD + 1);
};

// Hermite triple tables for the runtime-indexed loops of family B.
struct HermTab {
    unsigned char t[455], u[455], v[455];
};

// JK = false (write epilogues) stages only the scales and bounds, so those
// kernels hold a smaller LDS block (more resident wavefronts per CU).
template <int LA, int LB, int LC, int LD, class Real, bool JK = true>
struct FamBSmem {
    using S = FamBShape<LA, LB, LC, LD>;
    Real R[S::KB][2][S::NRP];          // packed ping-pong levels (not S::DENSE)
    Real g[S::NHAB * S::CDC];
    Real acc[S::NAB * S::CDC];
    Real eb[S::EB];
    Real ek[S::KB][S::EK];
    Real a[S::KB][S::L + 1];
    Real rx[S::KB][3];
    // class-constant decode tables, filled once per workgroup
    unsigned char ht[S::NR][3];        // packed Hermite index -> (t, u, v)
    unsigned char abc[S::NAB][6];      // bra component pair -> (ax, ay, az, bx, by, bz)
    unsigned char cdc[S::NCD][6];      // ket component pair -> (cx, cy, cz, dx, dy, dz)
    double epi[JK ? S::NEPI : S::O_DJAB];   // epilogue staging, per quartet (offsets in FamBShape)
    Real R3[S::KB][S::ND3];            // dense R cube (S::DENSE)
    // R recursion sources per packed Hermite index w (class constants):
    // packed levels: idx1 | idx2 << 10 | coef << 20 | axis << 26 (axis 3: the base a[n]);
    // dense cube: dense index | axis << 16 | coef << 18
    int rsrc[S::NR];
#ifdef HFX_LDS_PAD
    double pad[HFX_LDS_PAD];
#endif
};

// Stages what the epilogue of one quartet reads: function scales, pair bounds
// (not for EPI_DIAG, which is producing them) and, for J/K, the density
// blocks, so the epilogue reads LDS instead of HBM / L2.
template <int LA, int LB, int LC, int LD, class Real, bool JK, class SM>
HFX_HD __forceinline__ void famb_stage_epi(SM& sm, const QuartetCtx& q, const EpiArgs& ep, int lane, int nlanes) {
    using S = FamBShape<LA, LB, LC, LD>;
    double* E = sm.epi;
    const int K = ep.K;
    for (int w = lane; w < S::O_BAB; w += nlanes) {
        const int fn = w < S::NA ? q.fn1 + w
                     : w < S::NA + S::NB ? q.fn2 + (w - S::NA)
                     : w < S::NA + S::NB + S::NC ? q.fn3 + (w - S::NA - S::NB)
                     : q.fn4 + (w - S::NA - S::NB - S::NC);
        E[S::O_FS + w] = ep.fscale[fn];
    }
    if (ep.mode == EPI_DIAG) return;
    for (int w = lane; w < S::NAB; w += nlanes)
        E[S::O_BAB + w] = ep.bounds[pair_of(q.fn1 + w / S::NB, q.fn2 + w % S::NB)];
    for (int w = lane; w < S::NCD; w += nlanes)
        E[S::O_BCD + w] = ep.bounds[pair_of(q.fn3 + w / S::ND, q.fn4 + w % S::ND)];
    if constexpr (JK) {
        for (int w = lane; w < S::NAB; w += nlanes)
            E[S::O_DJAB + w] = ep.Dj[(q.fn1 + w / S::NB) * K + q.fn2 + w % S::NB];
        for (int w = lane; w < S::NCD; w += nlanes)
            E[S::O_DJCD + w] = ep.Dj[(q.fn3 + w / S::ND) * K + q.fn4 + w % S::ND];
        for (int spin = 0; spin < ep.nspin; ++spin) {
            const double* D = spin == 0 ? ep.Da : ep.Db;
            double* dk = E + S::O_DK + spin * S::NDK;
            // [D_ac: NA x NC][D_ad: NA x ND][D_bc: NB x NC][D_bd: NB x ND]
            for (int w = lane; w < S::NDK; w += nlanes) {
                int row, col;
                if (w < S::NA * S::NC) {
                    row = q.fn1 + w / S::NC; col = q.fn3 + w % S::NC;
                } else if (w < S::NA * (S::NC + S::ND)) {
                    const int x = w - S::NA * S::NC;
                    row = q.fn1 + x / S::ND; col = q.fn4 + x % S::ND;
                } else if (w < S::NA * (S::NC + S::ND) + S::NB * S::NC) {
                    const int x = w - S::NA * (S::NC + S::ND);
                    row = q.fn2 + x / S::NC; col = q.fn3 + x % S::NC;
                } else {
                    const int x = w - S::NA * (S::NC + S::ND) - S::NB * S::NC;
                    row = q.fn2 + x / S::ND; col = q.fn4 + x % S::ND;
                }
                dk[w] = D[row * K + col];
            }
        }
    }
}

// Fills the decode tables of a family-B workgroup (lanes stride the entries).
template <int LA, int LB, int LC, int LD, class Real, class SM>
HFX_HD __forceinline__ void famb_tables(SM& sm, const HermTab& ht, int lane, int nlanes) {
    using S = FamBShape<LA, LB, LC, LD>;
    for (int w = lane; w < S::NR; w += nlanes) {
        sm.ht[w][0] = ht.t[w];
        sm.ht[w][1] = ht.u[w];
        sm.ht[w][2] = ht.v[w];
    }
    for (int w = lane; w < S::NAB; w += nlanes) {
        const int a = w / S::NB, b = w % S::NB;
        sm.abc[w][0] = (unsigned char)cart_x(LA, a);
        sm.abc[w][1] = (unsigned char)cart_y(LA, a);
        sm.abc[w][2] = (unsigned char)cart_z(LA, a);
        sm.abc[w][3] = (unsigned char)cart_x(LB, b);
        sm.abc[w][4] = (unsigned char)cart_y(LB, b);
        sm.abc[w][5] = (unsigned char)cart_z(LB, b);
    }
    for (int w = lane; w < S::NCD; w += nlanes) {
        const int c = w / S::ND, d = w % S::ND;
        sm.cdc[w][0] = (unsigned char)cart_x(LC, c);
        sm.cdc[w][1] = (unsigned char)cart_y(LC, c);
        sm.cdc[w][2] = (unsigned char)cart_z(LC, c);
        sm.cdc[w][3] = (unsigned char)cart_x(LD, d);
        sm.cdc[w][4] = (unsigned char)cart_y(LD, d);
        sm.cdc[w][5] = (unsigned char)cart_z(LD, d);
    }
    // R_{tuv} = X R_{t-1,u,v} + (t-1) R_{t-2,u,v} (first non-zero index), base a[n] at (0,0,0)
    for (int w = lane; w < S::NR; w += nlanes) {
        const int t = ht.t[w], u = ht.u[w], v = ht.v[w];
        int i1 = 0, i2 = 0, c = 0, ax = 3;
        if (t > 0) {
            ax = 0; i1 = hidx(t - 1, u, v); c = t - 1; i2 = t > 1 ? hidx(t - 2, u, v) : i1;
        } else if (u > 0) {
            ax = 1; i1 = hidx(t, u - 1, v); c = u - 1; i2 = u > 1 ? hidx(t, u - 2, v) : i1;
        } else if (v > 0) {
            ax = 2; i1 = hidx(t, u, v - 1); c = v - 1; i2 = v > 1 ? hidx(t, u, v - 2) : i1;
        }
        if constexpr (S::DENSE)
            sm.rsrc[w] = ((t * S::D + u) * S::D + v) | (ax << 16) | (c << 18);
        else
            sm.rsrc[w] = i1 | (i2 << 10) | (c << 20) | (ax << 26);
    }
}

// 1-D E tables (x, y, z) of a primitive pair into LDS, odd Hermite orders
// negated for the ket.  All indices are compile-time constants so the pair
// stays in registers; one lane per group runs it (`lane == 0`).
template <int LA, int LB, class Real>
HFX_HD __forceinline__ void build_e_lanes(Real* dst, const PrimPair<Real>& pp, bool sign, int lane, int nlanes) {
    if (lane != 0) return;
    (void)nlanes;
    static_for<0, 3>([&](auto xi) {
        constexpr int x = decltype(xi)::value;
        Herm1D<LA, LB, Real> h;
        h.build(pp.i2p, pp.PA[x], pp.PB[x]);
        constexpr int S = (LA + 1) * (LB + 1) * (LA + LB + 1);
#pragma unroll
        for (int a = 0; a <= LA; ++a)
#pragma unroll
            for (int b = 0; b <= LB; ++b)
#pragma unroll
                for (int t = 0; t <= LA + LB; ++t)
                    dst[x * S + (a * (LB + 1) + b) * (LA + LB + 1) + t] = (sign && (t & 1)) ? -h.e[a][b][t] : h.e[a][b][t];
    });
}

// Epilogue of one chunk of ket component pairs cd0 .. cd0 + ncd - 1 of a
// family-B / family-M work unit: the raw block values sit in
// sm.acc[ab * CDC + cd - cd0]; sm.epi holds the staged scales, bounds and
// densities (famb_stage_epi).  Shared by the VALU and the MFMA kernels.
template <int LA, int LB, int LC, int LD, int CDC, bool JK, class Real, class SM, class Sync>
HFX_HD void famb_epilogue(SM& sm, const QuartetCtx& q, const EpiArgs& ep, int cd0_in, int ncd_in, int nslice,
                          double fdeg, int lane, int nlanes, Sync sync) {
    using S = FamBShape<LA, LB, LC, LD>;
    // one chunk holds every ket component pair: compile-time bounds let the
    // compiler unroll the sums below and drop their range checks
    constexpr bool ONE = S::NCD <= CDC;
    const int cd0 = ONE ? 0 : cd0_in, ncd = ONE ? S::NCD : ncd_in;
    const int K = ep.K;
    const double* E = sm.epi;
    auto fs4 = [&](int ab, int cd) -> double {
        return E[S::O_FS + ab / S::NB] * E[S::O_FS + S::NA + ab % S::NB] *
               E[S::O_FS + S::NA + S::NB + cd / S::ND] * E[S::O_FS + S::NA + S::NB + S::NC + cd % S::ND];
    };
    if (!JK && ep.mode == EPI_UNIQUE) {
        for (int w = lane; w < S::NAB * ncd; w += nlanes) {
            const int ab = w / ncd, cdl = w % ncd, cd = cd0 + cdl;
            const int mu = q.fn1 + ab / S::NB, nu = q.fn2 + ab % S::NB;
            const int la = q.fn3 + cd / S::ND, si = q.fn4 + cd % S::ND;
            if (!keep_unique(q, mu, nu, la, si)) continue;
            if (E[S::O_BAB + ab] * E[S::O_BCD + cd] < ep.threshold) continue;
            const double v = (double)sm.acc[ab * CDC + cdl] * fs4(ab, cd);
            const int64_t o = quartet_of(pair_of(mu, nu), pair_of(la, si));
            if (nslice > 1)
                atomic_add(&ep.out[o], v);
            else
                ep.out[o] = v;
        }
    } else if (!JK && ep.mode == EPI_DIAG) {
        for (int w = lane; w < S::NAB; w += nlanes) {
            const int cd = w;   // (ab|ab): same component pair on both sides
            if (cd < cd0 || cd >= cd0 + ncd) continue;
            const int mu = q.fn1 + w / S::NB, nu = q.fn2 + w % S::NB;
            if (q.d12 && mu < nu) continue;
            const double v = (double)sm.acc[w * CDC + (cd - cd0)] * fs4(w, cd);
            if (nslice > 1)
                atomic_add(&ep.out[pair_of(mu, nu)], v);
            else
                ep.out[pair_of(mu, nu)] = v;
        }
    } else if constexpr (JK) {
        // masked raw values in place (the function scales sit in the densities
        // and outputs of a direct build), then every J/K sum reads LDS only
        for (int w = lane; w < S::NAB * ncd; w += nlanes) {
            const int ab = w / ncd, cdl = w % ncd, cd = cd0 + cdl;
            Real& a = sm.acc[ab * CDC + cdl];
            if (E[S::O_BAB + ab] * E[S::O_BCD + cd] < ep.threshold) a = Real(0);
        }
        sync();
        auto v = [&](int ab, int cdl) -> double { return (double)sm.acc[ab * CDC + cdl]; };
        // J_ab over this chunk's cd
        for (int ab = lane; ab < S::NAB; ab += nlanes) {
            double s = 0.0;
            for (int cdl = 0; cdl < ncd; ++cdl) s += v(ab, cdl) * E[S::O_DJCD + cd0 + cdl];
            atomic_add(&ep.J[(q.fn1 + ab / S::NB) * K + q.fn2 + ab % S::NB], 2.0 * fdeg * s);
        }
        for (int cdl = lane; cdl < ncd; cdl += nlanes) {
            const int cd = cd0 + cdl;
            double s = 0.0;
            for (int ab = 0; ab < S::NAB; ++ab) s += v(ab, cdl) * E[S::O_DJAB + ab];
            atomic_add(&ep.J[(q.fn3 + cd / S::ND) * K + q.fn4 + cd % S::ND], 2.0 * fdeg * s);
        }
        for (int spin = 0; spin < ep.nspin; ++spin) {
            const double* dk = E + S::O_DK + spin * S::NDK;
            const double* Dac = dk;
            const double* Dad = Dac + S::NA * S::NC;
            const double* Dbc = Dad + S::NA * S::ND;
            const double* Dbd = Dbc + S::NB * S::NC;
            double* Kacc = spin == 0 ? ep.Ka : ep.Kb;
            // Kacc_ac += f sum_bd v D_bd ; Kacc_bc += f sum_ad v D_ad : targets (x, c)
            for (int w = lane; w < (S::NA + S::NB) * S::NC; w += nlanes) {
                const bool isA = w < S::NA * S::NC;
                const int x = isA ? w / S::NC : (w - S::NA * S::NC) / S::NC;
                const int c = w % S::NC;
                double s = 0.0;
                bool any = false;
                for (int d = 0; d < S::ND; ++d) {
                    const int cd = c * S::ND + d;
                    if (cd < cd0 || cd >= cd0 + ncd) continue;
                    any = true;
                    const int cdl = cd - cd0;
                    if (isA) {
                        for (int bb = 0; bb < S::NB; ++bb) s += v(x * S::NB + bb, cdl) * Dbd[bb * S::ND + d];
                    } else {
                        for (int a = 0; a < S::NA; ++a) s += v(a * S::NB + x, cdl) * Dad[a * S::ND + d];
                    }
                }
                if (any) atomic_add(&Kacc[((isA ? q.fn1 : q.fn2) + x) * K + q.fn3 + c], fdeg * s);
            }
            // Kacc_ad += f sum_bc v D_bc ; Kacc_bd += f sum_ac v D_ac : targets (x, d)
            for (int w = lane; w < (S::NA + S::NB) * S::ND; w += nlanes) {
                const bool isA = w < S::NA * S::ND;
                const int x = isA ? w / S::ND : (w - S::NA * S::ND) / S::ND;
                const int d = w % S::ND;
                double s = 0.0;
                bool any = false;
                for (int c = 0; c < S::NC; ++c) {
                    const int cd = c * S::ND + d;
                    if (cd < cd0 || cd >= cd0 + ncd) continue;
                    any = true;
                    const int cdl = cd - cd0;
                    if (isA) {
                        for (int bb = 0; bb < S::NB; ++bb) s += v(x * S::NB + bb, cdl) * Dbc[bb * S::NC + c];
                    } else {
                        for (int a = 0; a < S::NA; ++a) s += v(a * S::NB + x, cdl) * Dac[a * S::NC + c];
                    }
                }
                if (any) atomic_add(&Kacc[((isA ? q.fn1 : q.fn2) + x) * K + q.fn4 + d], fdeg * s);
            }
        }
    }
}

// One wavefront (or one CPU thread with nlanes = 1) computes one work unit: a
// shell quartet or, for slice s of S, the part of it from bra primitive pairs
// s, s + S, ... (the epilogue then adds its partial block).  chunk >= 0
// restricts the work to that chunk of ket components (chunk < 0: all chunks),
// so the chunks of a high-L quartet can run on different wavefronts.
// `sync()` must be a wavefront/workgroup barrier on the device.
template <int LA, int LB, int LC, int LD, class Real, bool JK, class Sync>
HFX_HD void family_b_one(uint2 e, int chunk, const PairHdr* hdr, const PrimPair<Real>* pp,
                         const Real* tab, const EpiArgs& ep, FamBSmem<LA, LB, LC, LD, Real, JK>& sm, const HermTab& ht,
                         int lane, int nlanes, Sync sync) {
    using S = FamBShape<LA, LB, LC, LD>;
    constexpr int SB = (LA + 1) * (LB + 1) * (S::LAB + 1);
    constexpr int SK = (LC + 1) * (LD + 1) * (S::LCD + 1);
    const PairHdr hb = hdr[unit_bra(e)], hk = hdr[unit_ket(e)];
    const int slice = unit_slice(e), nslice = unit_nslice(e);
    const FamBRange rg = famb_range(slice, nslice, hb.nprim, hk.nprim, S::KB);
    // uniform over the workgroup: an empty slice, or screened by the density difference
    if (rg.i0 >= hb.nprim || rg.j0 >= rg.j1 || dens_skip(ep, e, hb, hk)) return;
    const QuartetCtx q = make_ctx(hb, hk);
    const double* fs = ep.fscale;
    const int K = ep.K;
    const double fdeg = (q.d12 ? 0.5 : 1.0) * (q.d34 ? 0.5 : 1.0) * (q.same ? 0.5 : 1.0);
    famb_tables<LA, LB, LC, LD, Real>(sm, ht, lane, nlanes);
    famb_stage_epi<LA, LB, LC, LD, Real, JK>(sm, q, ep, lane, nlanes);

    const int ch_begin = chunk < 0 ? 0 : chunk, ch_end = chunk < 0 ? S::NCHUNK : chunk + 1;
    for (int ch = ch_begin; ch < ch_end; ++ch) {
        const int cd0 = ch * S::CDC;
        // a compile-time count when the class has one chunk, so the item
        // decodes below divide by a constant
        const int ncd = S::NCHUNK == 1 ? S::NCD : ((S::NCD - cd0) < S::CDC ? (S::NCD - cd0) : S::CDC);
        if (ch > ch_begin) sync();   // the previous chunk's epilogue has read sm.acc
        for (int w = lane; w < S::NAB * S::CDC; w += nlanes) sm.acc[w] = Real(0);
        for (int i = rg.i0; i < hb.nprim; i += rg.istep) {
            const PrimPair<Real> b = pp[hb.prim_off + i];
            sync();
            build_e_lanes<LA, LB, Real>(sm.eb, b, false, lane, nlanes);
            for (int w = lane; w < S::NHAB * S::CDC; w += nlanes) sm.g[w] = Real(0);
            // KB ket primitive pairs per step: on the device the wavefront splits
            // into KB groups of grp lanes, one pair each; with one lane the
            // groups run one after another.
            const int grp = nlanes >= S::KB ? nlanes / S::KB : 1;
            const int kb0 = nlanes >= S::KB ? lane / grp : 0;
            const int kbstep = nlanes >= S::KB ? S::KB : 1;
            const int gl = nlanes >= S::KB ? lane % grp : 0;
            for (int j0 = rg.j0; j0 < rg.j1; j0 += S::KB) {
                const int nkb = (rg.j1 - j0) < S::KB ? (rg.j1 - j0) : S::KB;
                sync();
                for (int kb = kb0; kb < nkb; kb += kbstep) {
                    const PrimPair<Real> k = pp[hk.prim_off + j0 + kb];
                    build_e_lanes<LC, LD, Real>(sm.ek[kb], k, true, gl, grp);
                    if (gl == 1 || grp == 1) {
                        Real s = b.p + k.p;
                        Real inv = Real(1) / s;
                        Real rho = b.p * k.p * inv;
                        Real X = b.P[0] - k.P[0], Y = b.P[1] - k.P[1], Z = b.P[2] - k.P[2];
                        Real T = rho * (X * X + Y * Y + Z * Z);
                        Real F[S::L + 1];
                        boys_fm<S::L, Real>(T, tab, F);
                        Real f = Real(TWO_PI_POW_2_5) * b.w * k.w * MathOps<Real>::sqrt_(inv);
                        Real m2r = Real(-2) * rho;
                        for (int n = 0; n <= S::L; ++n) {
                            sm.a[kb][n] = f * F[n];
                            f *= m2r;
                        }
                        sm.rx[kb][0] = X;
                        sm.rx[kb][1] = Y;
                        sm.rx[kb][2] = Z;
                    }
                }
                sync();
                // R levels n = L .. 0 : level n holds orders <= L - n, ping-pong buffers
                int cur = 0;
                if constexpr (S::DENSE) {
                    // in place: level n overwrites orders s = L-n .. 1 from high to
                    // low (reading orders s-1, s-2 that still hold level n+1), then
                    // the base R_000 = a[n]
                    for (int kb = kb0; kb < nkb; kb += kbstep)
                        if (gl == 0) sm.R3[kb][0] = sm.a[kb][S::L];
                    constexpr int st[3] = {S::D * S::D, S::D, 1};
                    for (int n = S::L - 1; n >= 0; --n) {
                        for (int so = S::L - n; so >= 0; --so) {
                            sync();
                            const int w0 = so == 0 ? 0 : nherm(so - 1), w1 = nherm(so);
                            for (int kb = kb0; kb < nkb; kb += kbstep) {
                                Real* R = sm.R3[kb];
                                for (int w = w0 + gl; w < w1; w += grp) {
                                    const int src = sm.rsrc[w], dd = src & 0xFFFF, ax = (src >> 16) & 3;
                                    if (ax == 3) {
                                        R[0] = sm.a[kb][n];
                                    } else {
                                        const int sd = st[ax];
                                        const int c = src >> 18;
                                        Real val = sm.rx[kb][ax] * R[dd - sd];
                                        if (c > 0) val += Real(c) * R[dd - 2 * sd];
                                        R[dd] = val;
                                    }
                                }
                            }
                        }
                    }
                } else {
                // packed levels n = L .. 0 : level n holds orders <= L - n, ping-pong buffers
                for (int kb = kb0; kb < nkb; kb += kbstep)
                    if (gl == 0) sm.R[kb][0][0] = sm.a[kb][S::L];
                for (int n = S::L - 1; n >= 0; --n) {
                    sync();
                    const int cnt = nherm(S::L - n);
                    for (int kb = kb0; kb < nkb; kb += kbstep) {
                        const Real* Rp = sm.R[kb][cur];
                        Real* Rn = sm.R[kb][cur ^ 1];
                        for (int w = gl; w < cnt; w += grp) {
                            const int src = sm.rsrc[w], ax = src >> 26;
                            Rn[w] = ax == 3 ? sm.a[kb][n]
                                            : sm.rx[kb][ax] * Rp[src & 1023] + Real((src >> 20) & 63) * Rp[(src >> 10) & 1023];
                        }
                    }
                    cur ^= 1;
                }
                }
                sync();
                // stage 1: g[tuv][cd] += sum_kb sum Eket_cd(tau nu phi) R_kb(t+tau, u+nu, v+phi)
                for (int w = lane; w < S::NHAB * ncd; w += nlanes) {
                    const int h = w / ncd, cdl = w % ncd, cd = cd0 + cdl;
                    const int t = sm.ht[h][0], u = sm.ht[h][1], v = sm.ht[h][2];
                    const int cx = sm.cdc[cd][0], cy = sm.cdc[cd][1], cz = sm.cdc[cd][2];
                    const int dx = sm.cdc[cd][3], dy = sm.cdc[cd][4], dz = sm.cdc[cd][5];
                    const int ox = (cx * (LD + 1) + dx) * (S::LCD + 1);
                    const int oy = SK + (cy * (LD + 1) + dy) * (S::LCD + 1);
                    const int oz = 2 * SK + (cz * (LD + 1) + dz) * (S::LCD + 1);
                    Real s = Real(0);
                    if constexpr (S::DENSE) {
                        const int b0 = (t * S::D + u) * S::D + v;
                        const int nx = cx + dx, ny = cy + dy, nz = cz + dz;
                        for (int kb = 0; kb < nkb; ++kb) {
                            const Real* Ex = sm.ek[kb] + ox;
                            const Real* Ey = sm.ek[kb] + oy;
                            const Real* Ez = sm.ek[kb] + oz;
                            const Real* Rt = sm.R3[kb] + b0;
                            for (int ta = 0; ta <= nx; ++ta, Rt += S::D * S::D) {
                                Real sy = Real(0);
                                const Real* Ru = Rt;
                                for (int nn = 0; nn <= ny; ++nn, Ru += S::D) {
                                    Real sz = Real(0);
                                    for (int ph = 0; ph <= nz; ++ph) sz += Ez[ph] * Ru[ph];
                                    sy += Ey[nn] * sz;
                                }
                                s += Ex[ta] * sy;
                            }
                        }
                    } else {
                        for (int kb = 0; kb < nkb; ++kb) {
                            const Real* R = sm.R[kb][cur];
                            const Real* Ex = sm.ek[kb] + ox;
                            const Real* Ey = sm.ek[kb] + oy;
                            const Real* Ez = sm.ek[kb] + oz;
                            for (int ta = 0; ta <= cx + dx; ++ta) {
                                Real sy = Real(0);
                                for (int nn = 0; nn <= cy + dy; ++nn) {
                                    Real sz = Real(0);
                                    for (int ph = 0; ph <= cz + dz; ++ph) sz += Ez[ph] * R[hidx(t + ta, u + nn, v + ph)];
                                    sy += Ey[nn] * sz;
                                }
                                s += Ex[ta] * sy;
                            }
                        }
                    }
                    sm.g[h * S::CDC + cdl] += s;
                }
            }
            sync();
            // stage 2: acc[ab][cd] += sum_tuv Ebra_ab(tuv) g[tuv][cd]
            for (int w = lane; w < S::NAB * ncd; w += nlanes) {
                const int ab = w / ncd, cdl = w % ncd;
                const int ax = sm.abc[ab][0], ay = sm.abc[ab][1], az = sm.abc[ab][2];
                const int bx = sm.abc[ab][3], by = sm.abc[ab][4], bz = sm.abc[ab][5];
                const Real* Ex = sm.eb + (ax * (LB + 1) + bx) * (S::LAB + 1);
                const Real* Ey = sm.eb + SB + (ay * (LB + 1) + by) * (S::LAB + 1);
                const Real* Ez = sm.eb + 2 * SB + (az * (LB + 1) + bz) * (S::LAB + 1);
                Real s = Real(0);
                for (int t = 0; t <= ax + bx; ++t) {
                    Real sy = Real(0);
                    for (int u = 0; u <= ay + by; ++u) {
                        Real sz = Real(0);
                        for (int v = 0; v <= az + bz; ++v) sz += Ez[v] * sm.g[hidx(t, u, v) * S::CDC + cdl];
                        sy += Ey[u] * sz;
                    }
                    s += Ex[t] * sy;
                }
                sm.acc[ab * S::CDC + cdl] += s;
            }
        }
        sync();
        // ---------------- epilogue for this chunk (values read from sm.acc)
        famb_epilogue<LA, LB, LC, LD, S::CDC, JK, Real>(sm, q, ep, cd0, ncd, nslice, fdeg, lane, nlanes, sync);
    }
    sync();
}

}  // namespace hfx
