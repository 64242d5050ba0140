// eri_mfma.hpp -- family M: shell-quartet ERIs with the two Hermite ->
// Cartesian contractions on the matrix cores (gfx950 MFMA).
//
// For a shell quartet (ab|cd) with bra primitive pairs p and ket primitive
// pairs q (McMurchie-Davidson; the reference's 6-deep Hermite loop,
// KLYYCYNN/hartree_fock_cpp cuda/include/kernels.cuh:51-77 and
// include/integrals.hpp:624-699):
//
//   (ab|cd) = sum_p sum_tuv Eb_p[ab][tuv] G_p[tuv][cd]
//   G_p[tuv][cd] = sum_q sum_{tau nu phi} (-1)^{tau+nu+phi} R^{pq}_{t+tau,u+nu,v+phi} Ek_q[cd][tau nu phi]
//
// Both sums are dense small GEMMs once the primitive pairs are concatenated
// into the contracted dimension:
//   stage 1: G[(p,tuv)][cd]  = A1[(p,tuv)][(q,tau nu phi)] . B1[(q,tau nu phi)][cd]
//            A1 = R values gathered from per-(p,q) Hermite-Coulomb arrays in
//            LDS, B1 = the ket pairs' Hermite -> Cartesian matrices with the
//            sign (-1)^{tau+nu+phi} folded in (precomputed per primitive
//            pair, BasisPlan::etkd);
//   stage 2: acc[ab][cd]     = A2[ab][(p,tuv)] . G[(p,tuv)][cd]
//            A2 = the bra pairs' matrices, B2 = G straight from the stage-1
//            accumulators: in the v_mfma_f64_16x16x4f64 layout a lane's
//            accumulator register r of a 16x16 tile is the B operand of
//            k-block r of that tile, so no data moves between the stages.
// MFMA lane layout (16x16x4): lane l supplies A[l%16][l/16] and B[l/16][l%16]
// and holds C[l/16 + 4r][l%16], r = 0..3.
//
// FP32 form (the mixed-precision path's small-bound quartets): E matrices,
// R recursion, both GEMMs and their accumulation in FP32 (v_mfma_f32_16x16x4f32,
// whose accumulator layout, lane l holding C[4(l/16) + r][l%16], is the
// transpose of the FP64 one across lane groups, so G passes from stage 1 to
// stage 2 through LDS), while the Boys function, the prefactors and P - Q are
// evaluated in FP64 from the FP64 primitive pairs and rounded once (the FP64
// islands: the Boys function at any T and n, and the E matrices, which are
// computed in FP64 on the host).
//
// One wavefront per work unit (a shell quartet or a primitive slice of one,
// famb_range), bra primitive pairs in groups of PG (stage-1 rows PG x NHAB),
// ket primitive pairs in passes of QC (their R arrays in LDS), ket
// components in chunks of CDM columns; the R_tuv of all (p,q) of a pass are
// built lane-parallel in LDS by the downward recursion (integrals.hpp:541-618
// restated once per primitive quartet); the epilogues are family B's.
#pragma once
#include "eri_body.hpp"

namespace hfx {

template <class Real> struct MfmaOps;
#if defined(__HIP_DEVICE_COMPILE__)
typedef double hfx_f64x4 __attribute__((ext_vector_type(4)));
typedef float hfx_f32x4 __attribute__((ext_vector_type(4)));
template <> struct MfmaOps<double> {
    using V4 = hfx_f64x4;
    static __device__ __forceinline__ V4 mma(double a, double b, V4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct MfmaOps<float> {
    using V4 = hfx_f32x4;
    static __device__ __forceinline__ V4 mma(float a, float b, V4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
#endif

// PGV: bra primitive pairs per group; 0 selects the class default PGD.  The
// ket-component chunking (NT, CDM, NCHUNK) and the pass capacity NCUBE are
// those of the default in every variant, so all variants share one LDS layout.
template <int LA, int LB, int LC, int LD, int PGV = 0>
struct FamMShape {
    using Bs = FamBShape<LA, LB, LC, LD>;
    static constexpr int NAB = Bs::NAB, NCD = Bs::NCD;
    static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
    static constexpr int NHAB = nherm(LAB), NHCD = nherm(LCD), NR = nherm(L);
    // default: bra primitive pairs per group with stage-1 rows PGD x NHAB within 64 where possible
    static constexpr int PG0 = 64 / NHAB;
    static constexpr int PGD = PG0 < 1 ? 1 : PG0 > 4 ? 4 : PG0;
    static constexpr int PG = PGV > 0 ? PGV : PGD;
    static constexpr int MT1 = (PG * NHAB + 15) / 16;
    static constexpr int KT1 = (NHCD + 3) / 4;
    static constexpr int MT2 = (NAB + 15) / 16;
    static constexpr int KT2 = (PG * NHAB + 3) / 4;
    // ket component chunk: accumulators (MT1 + MT2) x NT tiles of 4 values per lane, at most 64
    static constexpr int MT1D = (PGD * NHAB + 15) / 16;
    static constexpr int NTMAX = (NCD + 15) / 16;
    static constexpr int NTFIT = 16 / (MT1D + MT2) < 1 ? 1 : 16 / (MT1D + MT2);
    static constexpr int NT = NTMAX < NTFIT ? NTMAX : NTFIT;
    static constexpr int CDM = 16 * NT;
    static constexpr int NCHUNK = (NCD + CDM - 1) / CDM;
    // ket primitive pairs per pass: the pass's R arrays within ~1024 values
    // (LDS per wavefront bounds the resident wavefronts per CU; measured:
    // staging the ket and bra matrices in LDS as well, one resident wavefront
    // per SIMD, ran (dp|pp) 1.8x slower than reading them from L2)
    static constexpr int QC0 = 1024 / (PGD * NR);
    static constexpr int QCD = QC0 < 1 ? 1 : QC0 > 64 ? 64 : QC0;
    static constexpr int NCUBE = QCD * PGD;
    static constexpr int QC = NCUBE / PG;
    // stage-1 gather offsets held per lane in registers when few
    static constexpr bool OFFREG = MT1 * KT1 <= 12;
};


template <int LA, int LB, int LC, int LD, class Real, bool JK>
struct FamMSmem {
    using S = FamMShape<LA, LB, LC, LD>;
    using Bs = FamBShape<LA, LB, LC, LD>;
    Real cube[S::NCUBE][S::NR];        // R^{(n)}_tuv of each (p, q) of a pass, packed Hermite index, in place
    Real a[S::NCUBE][S::L + 1];        // pref (-2 rho)^n F_n(T)
    Real x[S::NCUBE][3];               // P - Q
    Real acc[Bs::NAB * S::CDM];        // a chunk's block values for the epilogue
    Real gt[sizeof(Real) == 4 ? 16 * S::MT1 * S::CDM : 1];   // FP32: G in [(p,tuv)][cd] layout for stage 2
    double epi[JK ? Bs::NEPI : Bs::O_DJAB];
    int rsrc[S::NR];                   // recursion sources: i1 | i2 << 10 | coef << 20 | axis << 26 (axis 3: base)
    unsigned short lut[S::NHAB * S::NHCD];   // hidx(t+tau, u+nu, v+phi)
};

#if defined(__HIP_DEVICE_COMPILE__)
// Class-constant tables of a family-M workgroup (one wavefront).
template <int LA, int LB, int LC, int LD, class SM>
__device__ __forceinline__ void famm_tables(SM& sm, const HermTab& ht, int lane) {
    using S = FamMShape<LA, LB, LC, LD>;
    for (int w = lane; w < S::NR; w += 64) {
        const int t = ht.t[w], u = ht.u[w], v = ht.v[w];
        int i1 = 0, i2 = 0, c = 0, ax = 3;
        if (t > 0) {
            ax = 0; i1 = hidx(t - 1, u, v); c = t - 1; i2 = t > 1 ? hidx(t - 2, u, v) : i1;
        } else if (u > 0) {
            ax = 1; i1 = hidx(t, u - 1, v); c = u - 1; i2 = u > 1 ? hidx(t, u - 2, v) : i1;
        } else if (v > 0) {
            ax = 2; i1 = hidx(t, u, v - 1); c = v - 1; i2 = v > 1 ? hidx(t, u, v - 2) : i1;
        }
        sm.rsrc[w] = i1 | (i2 << 10) | (c << 20) | (ax << 26);
    }
    for (int w = lane; w < S::NHAB * S::NHCD; w += 64) {
        const int h = w / S::NHCD, k = w % S::NHCD;
        const int t = ht.t[h] + ht.t[k], u = ht.u[h] + ht.u[k], v = ht.v[h] + ht.v[k];
        sm.lut[w] = (unsigned short)hidx(t, u, v);
    }
}

// Offset into the R arrays of a pass (relative to the pass's first ket pair)
// of the stage-1 A value that lane `lane` supplies in tile m, k-block kb:
// row (p, tuv) = 16 m + lane % 16, column tau nu phi = 4 kb + lane / 16;
// -1 outside the matrix.
template <class S, class SM>
__device__ __forceinline__ int famm_off(const SM& sm, int m, int kb, int lane) {
    const int row = 16 * m + (lane & 15), tnp = 4 * kb + (lane >> 4);
    const int pi = row / S::NHAB, tuv = row - pi * S::NHAB;
    return (pi < S::PG && tnp < S::NHCD) ? pi * S::NR + sm.lut[tuv * S::NHCD + tnp] : -1;
}

// One work unit on one wavefront (lane = 0..63, all active).  etab / etabk:
// the E tables in bra and ket form (BasisPlan::etd / etkd, or the FP32
// copies) in the precision of the class kernel.
// pp / tab: the FP64 primitive pairs and Boys table (both precisions
// evaluate the Boys function and prefactors in FP64).
template <int LA, int LB, int LC, int LD, class Real, bool JK, int PGV>
__device__ void family_m_body(const PairHdr& hb, const PairHdr& hk, const FamBRange& rg, int np, int nslice,
                              const QuartetCtx& q, const PrimPair<double>* __restrict__ pp,
                              const Real* __restrict__ etab, const Real* __restrict__ etabk,
                              const double* __restrict__ tab, const EpiArgs& ep,
                              FamMSmem<LA, LB, LC, LD, Real, JK>& sm, int lane) {
    using S = FamMShape<LA, LB, LC, LD, PGV>;
    using V4 = typename MfmaOps<Real>::V4;
    auto sync = []() { __syncthreads(); };
    const double fdeg = (q.d12 ? 0.5 : 1.0) * (q.d34 ? 0.5 : 1.0) * (q.same ? 0.5 : 1.0);
    const int g = lane >> 4, c16 = lane & 15;                        // MFMA lane coordinates
    const Real* eb = etab + hb.eoff;
    const Real* ek = etabk + hk.eoff;
    int off[S::OFFREG ? S::MT1 : 1][S::OFFREG ? S::KT1 : 1];
    if constexpr (S::OFFREG) {
#pragma unroll
        for (int m = 0; m < S::MT1; ++m)
#pragma unroll
            for (int kb = 0; kb < S::KT1; ++kb) off[m][kb] = famm_off<S>(sm, m, kb, lane);
    }
    for (int ch = 0; ch < S::NCHUNK; ++ch) {
        const int cd0 = ch * S::CDM;
        const int ncd = S::NCD - cd0 < S::CDM ? S::NCD - cd0 : S::CDM;
        V4 acc[S::MT2][S::NT];
#pragma unroll
        for (int m = 0; m < S::MT2; ++m)
#pragma unroll
            for (int n = 0; n < S::NT; ++n) acc[m][n] = V4{0, 0, 0, 0};
        for (int pg0 = 0; pg0 < np; pg0 += S::PG) {
            const int npg = np - pg0 < S::PG ? np - pg0 : S::PG;
            V4 G[S::MT1][S::NT];
#pragma unroll
            for (int m = 0; m < S::MT1; ++m)
#pragma unroll
                for (int n = 0; n < S::NT; ++n) G[m][n] = V4{0, 0, 0, 0};
            for (int q0 = rg.j0; q0 < rg.j1; q0 += S::QC) {
                const int nqc = rg.j1 - q0 < S::QC ? rg.j1 - q0 : S::QC;
                sync();   // the previous pass has read its R arrays
                // Boys function and prefactors of each (p, q): one lane each
                for (int c = lane; c < nqc * S::PG; c += 64) {
                    const int pi = c % S::PG, qi = c / S::PG;
                    if (pi >= npg) {   // absent bra pair: zero arrays, so stage 1 reads them unmasked
#pragma unroll
                        for (int n = 0; n <= S::L; ++n) sm.a[c][n] = Real(0);
                        sm.x[c][0] = sm.x[c][1] = sm.x[c][2] = Real(0);
                        sm.cube[c][0] = Real(0);
                        continue;
                    }
                    const PrimPair<double> b = pp[hb.prim_off + rg.i0 + (pg0 + pi) * rg.istep];
                    const PrimPair<double> k = pp[hk.prim_off + q0 + qi];
                    const double s = b.p + k.p;
                    const double inv = 1.0 / s;
                    const double rho = b.p * k.p * inv;
                    const double X = b.P[0] - k.P[0], Y = b.P[1] - k.P[1], Z = b.P[2] - k.P[2];
                    const double T = rho * (X * X + Y * Y + Z * Z);
                    double F[S::L + 1];
                    boys_fm<S::L, double>(T, tab, F);
                    double f = TWO_PI_POW_2_5 * b.w * k.w * sqrt(inv);
                    const double m2r = -2.0 * rho;
#pragma unroll
                    for (int n = 0; n <= S::L; ++n) {
                        sm.a[c][n] = (Real)(f * F[n]);
                        f *= m2r;
                    }
                    sm.x[c][0] = (Real)X;
                    sm.x[c][1] = (Real)Y;
                    sm.x[c][2] = (Real)Z;
                    sm.cube[c][0] = sm.a[c][S::L];
                }
                sync();
                // R recursion, all arrays of the pass at once: level n rewrites
                // orders L-n .. 1 from high to low (each reads orders one and two
                // below, which still hold level n+1), then the base a[n]
                const int ncube = nqc * S::PG;
                static_for<0, S::L>([&](auto ni) {
                    constexpr int n = S::L - 1 - decltype(ni)::value;
                    static_for<0, S::L - n>([&](auto si) {
                        constexpr int so = S::L - n - decltype(si)::value;
                        constexpr int w0 = nherm(so - 1), cnt = nherm(so) - w0;   // order so's entries
                        for (int idx = lane; idx < ncube * cnt; idx += 64) {
                            const int c = idx / cnt, w = w0 + (idx - c * cnt);
                            const int src = sm.rsrc[w];
                            Real* R = sm.cube[c];
                            const int ax = src >> 26;
                            R[w] = sm.x[c][ax] * R[src & 1023] + Real((src >> 20) & 63) * R[(src >> 10) & 1023];
                        }
                        sync();
                    });
                    for (int c = lane; c < ncube; c += 64) sm.cube[c][0] = sm.a[c][n];
                    sync();
                });
                // stage 1: G[(p,tuv)][cd] += sum_{tau nu phi} A1 B1, per ket pair of the pass.
                // Every load reads a valid address (indices clamped) and is masked
                // afterwards, so the loads of a k-block issue together.
                for (int qi = 0; qi < nqc; ++qi) {
                    const Real* ekq = ek + (long long)(q0 + qi) * (S::NCD * S::NHCD);   // [tau nu phi][cd]
                    const Real* cq = sm.cube[qi * S::PG];
                    int gl = g, cl = c16;
                    asm volatile("" : "+v"(gl), "+v"(cl));
                    static_for<0, S::KT1>([&](auto kbi) {
                        constexpr int kb = decltype(kbi)::value;
                        const int tnp = 4 * kb + gl;   // this lane's tau nu phi
                        const bool tin = tnp < S::NHCD;
                        const Real* er = ekq + (tin ? tnp : 0) * S::NCD;
                        Real bf[S::NT];
#pragma unroll
                        for (int n = 0; n < S::NT; ++n) {
                            const int cd = cd0 + 16 * n + cl;
                            const Real v = er[cd < S::NCD ? cd : S::NCD - 1];
                            bf[n] = (tin && cd < S::NCD) ? v : Real(0);
                        }
#pragma unroll
                        for (int m = 0; m < S::MT1; ++m) {
                            int o;
                            if constexpr (S::OFFREG)
                                o = off[m][kb];
                            else
                                o = famm_off<S>(sm, m, kb, lane);
                            const Real v = cq[o >= 0 ? o : 0];
                            const Real av = o >= 0 ? v : Real(0);
#pragma unroll
                            for (int n = 0; n < S::NT; ++n) G[m][n] = MfmaOps<Real>::mma(av, bf[n], G[m][n]);
                        }
                    });
                }
            }
            // stage 2: acc[ab][cd] += sum_{(p,tuv)} Eb[ab][(p,tuv)] G[(p,tuv)][cd]
            if constexpr (sizeof(Real) == 4) {   // FP32 layout: G through LDS, [(p,tuv)][cd]
#pragma unroll
                for (int m = 0; m < S::MT1; ++m)
#pragma unroll
                    for (int n = 0; n < S::NT; ++n)
#pragma unroll
                        for (int r = 0; r < 4; ++r) sm.gt[(16 * m + 4 * g + r) * S::CDM + 16 * n + c16] = G[m][n][r];
                sync();
            }
            {
                const long long pstr = (long long)rg.istep * (S::NAB * S::NHAB);
                const Real* ebg = eb + (long long)(rg.i0 + pg0 * rg.istep) * (S::NAB * S::NHAB);   // [tuv][ab]
                // lane coordinates made opaque here: the per-k-block offsets below are
                // recomputed in this loop instead of hoisted (and spilled) as invariants
                int gl = g, cl = c16;
                asm volatile("" : "+v"(gl), "+v"(cl));
                static_for<0, S::KT2>([&](auto kki) {
                    constexpr int kk = decltype(kki)::value;
                    const int krow = 4 * kk + gl;   // this lane's (p, tuv) of the A operand
                    const int pi = krow / S::NHAB, tuv = krow - pi * S::NHAB;
                    const bool kin = pi < npg;
                    const Real* src = ebg + (kin ? pi : 0) * pstr + tuv * S::NAB;
#pragma unroll
                    for (int m = 0; m < S::MT2; ++m) {
                        const int ab = 16 * m + cl;
                        const Real v = src[ab < S::NAB ? ab : S::NAB - 1];
                        const Real av = (kin && ab < S::NAB) ? v : Real(0);
#pragma unroll
                        for (int n = 0; n < S::NT; ++n) {
                            Real bv;
                            if constexpr (sizeof(Real) == 4)
                                bv = sm.gt[krow * S::CDM + 16 * n + c16];
                            else
                                bv = G[kk / 4][n][kk % 4];
                            acc[m][n] = MfmaOps<Real>::mma(av, bv, acc[m][n]);
                        }
                    }
                });
            }
        }
        // the chunk's block into LDS (lane holds acc[ab = 16m + g + 4r][cd = 16n + c16]
        // in FP64, acc[ab = 16m + 4g + r][cd] in FP32)
#pragma unroll
        for (int m = 0; m < S::MT2; ++m)
#pragma unroll
            for (int n = 0; n < S::NT; ++n)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int ab = sizeof(Real) == 4 ? 16 * m + 4 * g + r : 16 * m + g + 4 * r, cdl = 16 * n + c16;
                    if (ab < S::NAB && cdl < ncd) sm.acc[ab * S::CDM + cdl] = acc[m][n][r];
                }
        sync();
        famb_epilogue<LA, LB, LC, LD, S::CDM, JK, Real>(sm, q, ep, cd0, ncd, nslice, fdeg, lane, 64, sync);
        sync();
    }
}

// One work unit on one wavefront.  A unit with a single bra primitive pair
// (common for d and f shells of one primitive) takes the PG = 1 body: with the
// class's default group it would run its stage-1 tiles, stage-2 k-blocks, Boys
// and R recursions on rows of absent bra pairs.
template <int LA, int LB, int LC, int LD, class Real, bool JK>
__device__ void family_m_one(uint2 e, const PairHdr* __restrict__ hdr, const PrimPair<double>* __restrict__ pp,
                             const Real* __restrict__ etab, const Real* __restrict__ etabk,
                             const double* __restrict__ tab, const EpiArgs& ep,
                             FamMSmem<LA, LB, LC, LD, Real, JK>& sm, int lane) {
    using S = FamMShape<LA, LB, LC, LD>;
    const PairHdr hb = hdr[unit_bra(e)], hk = hdr[unit_ket(e)];
    const int slice = unit_slice(e), nslice = unit_nslice(e);
    const FamBRange rg = famb_range(slice, nslice, hb.nprim, hk.nprim, famb_kb(S::L));
    if (rg.i0 >= hb.nprim || rg.j0 >= rg.j1 || dens_skip(ep, e, hb, hk)) return;   // uniform
    const QuartetCtx q = make_ctx(hb, hk);
    famb_stage_epi<LA, LB, LC, LD, Real, JK>(sm, q, ep, lane, 64);
    const int np = (hb.nprim - rg.i0 + rg.istep - 1) / rg.istep;   // bra primitive pairs of this unit
    if constexpr (S::PGD > 1) {
        if (np == 1) {
            family_m_body<LA, LB, LC, LD, Real, JK, 1>(hb, hk, rg, np, nslice, q, pp, etab, etabk, tab, ep, sm, lane);
            return;
        }
    }
    family_m_body<LA, LB, LC, LD, Real, JK, 0>(hb, hk, rg, np, nslice, q, pp, etab, etabk, tab, ep, sm, lane);
}
#endif

}  // namespace hfx
