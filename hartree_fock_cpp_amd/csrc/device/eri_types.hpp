// eri_types.hpp -- device-side data layout shared by the ERI kernels and the
// host-side planner (hfx_api.hip).
//
// HBM layout (all built once per basis, resident for the life of the basis):
//   shell pairs  : PairHdr[npair]           one per canonical shell pair A >= B,
//                                            stored in kernel order (larger L first)
//   prim pairs   : PrimPair<double>[nprimpair] and PrimPair<float>[...] (AoS, 96/48 B),
//                                            contiguous per shell pair, primitive
//                                            pairs with exp(-mu AB^2) < e^-60 dropped
//   fn scale     : double[K]                per-Cartesian-component normalisation
//                                            ratio (norm / radial norm)
//   bounds       : double[K(K+1)/2]         function-pair Schwarz bounds
//   quartet lists: uint2[n] per (class, precision): (bra pair id, ket pair id)
#pragma once
#include <stdint.h>

namespace hfx {

template <class Real>
struct alignas(16) PrimPair {
    Real p;       // alpha + beta
    Real i2p;     // 1 / (2p)
    Real w;       // c_a c_b exp(-mu |AB|^2) / p  (radial norms folded into c)
    Real P[3];    // Gaussian product centre
    Real PA[3];   // P - A  (A = first shell in kernel order)
    Real PB[3];   // P - B
};

struct PairHdr {
    int s1, s2;        // shells in kernel order (L(s1) >= L(s2))
    int fn1, fn2;      // first basis function of s1, s2
    int prim_off;      // offset into the PrimPair arrays
    int nprim;         // number of kept primitive pairs
    int diag;          // 1 when s1 == s2
    int canon;         // canonical pair index A(A+1)/2 + B
    long long eoff;    // first entry of the pair's Hermite -> Cartesian matrices (family-M E tables)
};

enum EpiMode : int {
    EPI_DIAG = 0,     // Schwarz: out[pair(mu,nu)] (+)= (mu nu|mu nu) (primitive slices add);
                      // the caller then takes bounds = sqrt(max(out, 0)) (launch_sqrt_clamp)
    EPI_UNIQUE = 1,   // write the 8-fold unique array at eri_index(mu,nu,la,si)
    EPI_JK = 2,       // contract with densities into J / K accumulators
};

struct EpiArgs {
    int mode;
    int K;
    int nspin;                 // 1 = RHF (K from Dj), 2 = UHF (Ka from Da, Kb from Db)
    double threshold;          // function-level Schwarz threshold
    const double* bounds;      // function-pair bounds (P entries)
    const double* fscale;      // per-function component scale
    double* out;               // EPI_DIAG: bounds out; EPI_UNIQUE: unique ERI array
    const double* Dj;          // density for J (Pa + Pb), K x K
    const double* Da;          // density for K (alpha or total)
    const double* Db;          // beta density (UHF)
    double* J;                 // J accumulator (symmetrised afterwards)
    double* Ka;                // K accumulators
    double* Kb;
    // density-difference screening of incremental J/K builds: a unit is
    // skipped when qpair[bra] * qpair[ket] * (max |dP| over the six shell
    // blocks its J/K terms read) < dthr.  dpmax == nullptr: no screening.
    const double* qpair;       // shell-pair Schwarz maxima by pair id
    const double* dpmax;       // nshell x nshell block maxima of |dP|
    int nshell;
    double dthr;
};

// A work unit of a launch list: a shell quartet (bra, ket pair ids < 2^24)
// or one primitive slice of it, with the slice index in the top 8 bits of .x
// and the slice count in the top 8 bits of .y (0 reads as 1: a whole
// quartet).  Slices add partial blocks in the epilogue through atomics, so
// EPI_UNIQUE then needs a zeroed output.
constexpr unsigned UNIT_ID_MASK = 0xFFFFFFu;
__host__ __device__ inline unsigned unit_bra(uint2 e) { return e.x & UNIT_ID_MASK; }
__host__ __device__ inline unsigned unit_ket(uint2 e) { return e.y & UNIT_ID_MASK; }
__host__ __device__ inline int unit_slice(uint2 e) { return (int)(e.x >> 24); }
__host__ __device__ inline int unit_nslice(uint2 e) { return (e.y >> 24) ? (int)(e.y >> 24) : 1; }

__host__ __device__ inline int64_t pair_of(int a, int b) {
    return a >= b ? (int64_t)a * (a + 1) / 2 + b : (int64_t)b * (b + 1) / 2 + a;
}
__host__ __device__ inline int64_t quartet_of(int64_t IJ, int64_t KL) {
    return IJ >= KL ? IJ * (IJ + 1) / 2 + KL : KL * (KL + 1) / 2 + IJ;
}

}  // namespace hfx
