// eri_launch_impl.hpp -- __global__ ERI kernels and per-class launchers.
//
// Included by the eri_part*.hip translation units; each instantiates a range
// of the 55 classes so the build parallelises.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <type_traits>
#include <utility>

#include "eri_body.hpp"
#include "eri_classes.hpp"
#include "eri_mfma.hpp"
#include "hfx_launch.hpp"

namespace hfx {

constexpr int FAM_A_BLOCK = 64;

// G lanes per work unit; FAM_A_BLOCK / G units per workgroup.
template <int LA, int LB, int LC, int LD, class Real, int G, bool JK>
__global__ __launch_bounds__(FAM_A_BLOCK) void k_fam_a(const uint2* __restrict__ list, int n,
                                                        const PairHdr* __restrict__ hdr,
                                                        const PrimPair<Real>* __restrict__ pp,
                                                        const Real* __restrict__ tab, EpiArgs ep) {
    const int u = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / G);
    const bool active = u < n;
    family_a_group<LA, LB, LC, LD, Real, G, JK>(list[active ? u : n - 1], active, (int)(threadIdx.x % G), hdr, pp,
                                                tab, ep);
}

// Minimum resident wavefronts per SIMD the register allocator must leave room
// for (1: no constraint).  Unconstrained, family-B kernels take 128-165 VGPRs
// (3-4 waves); for the classes below a few bytes of spill buy a fourth or
// fifth wave and pay, elsewhere the spill traffic costs more than the
// occupancy gains.  Chosen per class from a sweep of 1 / 4 / 5 / 6 on the
// benzene / cc-pVDZ unique build (write) and the (H2O)4 / cc-pVTZ and
// n-octane / cc-pVDZ direct builds (J/K); a class keeps 1 unless it gains 2 %.
template <int LA, int LB, int LC, int LD, bool JK>
constexpr int famb_waves_per_eu() {
    constexpr int c = ((LA * 4 + LB) * 4 + LC) * 4 + LD;
    constexpr auto id = [](int a, int b, int cc, int d) { return ((a * 4 + b) * 4 + cc) * 4 + d; };
    if constexpr (!JK) return (c == id(2, 1, 2, 0) || c == id(2, 2, 1, 0)) ? 5 : 1;
    if (c == id(2, 2, 1, 0) || c == id(3, 0, 2, 0)) return 5;
    if (c == id(2, 1, 1, 1) || c == id(2, 1, 2, 0) || c == id(2, 2, 1, 1) || c == id(3, 0, 3, 0) ||
        c == id(3, 1, 1, 0) || c == id(3, 1, 3, 0) || c == id(3, 2, 0, 0) || c == id(3, 2, 1, 0) || c == id(3, 3, 1, 0))
        return 4;
    return 1;
}

// One wavefront per (work unit, ket-component chunk), grid-strided.
template <int LA, int LB, int LC, int LD, class Real, bool JK>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(famb_waves_per_eu<LA, LB, LC, LD, JK>()))) void k_fam_b(const uint2* __restrict__ list, int n, const PairHdr* __restrict__ hdr,
                                               const PrimPair<Real>* __restrict__ pp, const Real* __restrict__ tab,
                                               const HermTab* __restrict__ ht, EpiArgs ep) {
    __shared__ FamBSmem<LA, LB, LC, LD, Real, JK> sm;
    auto sync = []() { __syncthreads(); };
    constexpr int NCH = FamBShape<LA, LB, LC, LD>::NCHUNK;
    const long long total = (long long)n * NCH;
    for (long long u = blockIdx.x; u < total; u += gridDim.x)
        family_b_one<LA, LB, LC, LD, Real, JK>(list[u / NCH], (int)(u % NCH), hdr, pp, tab, ep, sm, *ht,
                                               (int)threadIdx.x, 64, sync);
}

// One wavefront per work unit, grid-strided; class tables built once per workgroup.
template <int LA, int LB, int LC, int LD, class Real, bool JK>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_fam_m(
    const uint2* __restrict__ list, int n, const PairHdr* __restrict__ hdr, const PrimPair<double>* __restrict__ ppd,
    const Real* __restrict__ etab, const Real* __restrict__ etabk, const double* __restrict__ tabd,
    const HermTab* __restrict__ ht, EpiArgs ep) {
#if defined(__HIP_DEVICE_COMPILE__)
    __shared__ FamMSmem<LA, LB, LC, LD, Real, JK> sm;
    famm_tables<LA, LB, LC, LD>(sm, *ht, (int)threadIdx.x);
    __syncthreads();
    for (long long u = blockIdx.x; u < n; u += gridDim.x)
        family_m_one<LA, LB, LC, LD, Real, JK>(list[u], hdr, ppd, etab, etabk, tabd, ep, sm, (int)threadIdx.x);
#endif
}

// Family M (MFMA contractions) for the classes that use_family_m admits, in
// both precisions; HFX_FAMM=0 in the environment selects family B instead
// (measurement).
inline bool famm_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("HFX_FAMM");
        return !(e && e[0] == '0');
    }();
    return on;
}

template <int QC, class Real, bool JK>
hipError_t launch_class_mode(const ClassLaunch& L, hipStream_t st) {
    using C = QClass<QC>;
    const PrimPair<Real>* pp = static_cast<const PrimPair<Real>*>(L.prim);
    const Real* tab = static_cast<const Real*>(L.boys);
    if constexpr (runs_family_a(C::LA, C::LB, C::LC, C::LD)) {
        const long long lanes = (long long)L.n * L.G;
        const int grid = (int)((lanes + FAM_A_BLOCK - 1) / FAM_A_BLOCK);
        switch (L.G) {
            case 1: k_fam_a<C::LA, C::LB, C::LC, C::LD, Real, 1, JK><<<grid, FAM_A_BLOCK, 0, st>>>(L.list, L.n, L.hdr, pp, tab, L.ep); break;
            case 4: k_fam_a<C::LA, C::LB, C::LC, C::LD, Real, 4, JK><<<grid, FAM_A_BLOCK, 0, st>>>(L.list, L.n, L.hdr, pp, tab, L.ep); break;
            case 16: k_fam_a<C::LA, C::LB, C::LC, C::LD, Real, 16, JK><<<grid, FAM_A_BLOCK, 0, st>>>(L.list, L.n, L.hdr, pp, tab, L.ep); break;
            case 64: k_fam_a<C::LA, C::LB, C::LC, C::LD, Real, 64, JK><<<grid, FAM_A_BLOCK, 0, st>>>(L.list, L.n, L.hdr, pp, tab, L.ep); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        if constexpr (use_family_m(C::LA, C::LB, C::LC, C::LD)) {
            if (famm_enabled()) {
                const int grid = L.n < 65536 ? L.n : 65536;
                k_fam_m<C::LA, C::LB, C::LC, C::LD, Real, JK><<<grid, 64, 0, st>>>(
                    L.list, L.n, L.hdr, static_cast<const PrimPair<double>*>(L.prim64),
                    static_cast<const Real*>(L.etab), static_cast<const Real*>(L.etabk),
                    static_cast<const double*>(L.boys64), L.herm, L.ep);
                return hipGetLastError();
            }
        }
        const long long units = (long long)L.n * FamBShape<C::LA, C::LB, C::LC, C::LD>::NCHUNK;
        const int grid = units < 65536 ? (int)units : 65536;
        k_fam_b<C::LA, C::LB, C::LC, C::LD, Real, JK><<<grid, 64, 0, st>>>(L.list, L.n, L.hdr, pp, tab, L.herm, L.ep);
    }
    return hipGetLastError();
}

// The write epilogues (Schwarz diagonal, unique array) and the J/K epilogue
// are separate instantiations so the write kernels keep a small register file.
template <int QC, class Real>
hipError_t launch_class(const ClassLaunch& L, hipStream_t st) {
    if (L.n <= 0) return hipSuccess;
    return L.ep.mode == EPI_JK ? launch_class_mode<QC, Real, true>(L, st) : launch_class_mode<QC, Real, false>(L, st);
}

template <int Begin, int... I>
hipError_t launch_range(std::integer_sequence<int, I...>, int qc, bool fp32, const ClassLaunch& L, hipStream_t st) {
    static constexpr LaunchFn t64[] = {&launch_class<Begin + I, double>...};
    static constexpr LaunchFn t32[] = {&launch_class<Begin + I, float>...};
    return fp32 ? t32[qc - Begin](L, st) : t64[qc - Begin](L, st);
}

}  // namespace hfx

#define HFX_DEFINE_PART(NAME, BEGIN, END)                                                               \
    namespace hfx {                                                                                     \
    hipError_t NAME(int qc, bool fp32, const ClassLaunch& L, hipStream_t st) {                         \
        return launch_range<BEGIN>(std::make_integer_sequence<int, (END) - (BEGIN)>{}, qc, fp32, L, st); \
    }                                                                                                   \
    }
