// hfx_launch.hpp -- host-side launch interface of hfx_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "eri_body.hpp"

namespace hfx {

struct ClassLaunch {
    const uint2* list;     // device (bra pair, ket pair) ids of one class
    int n;
    const PairHdr* hdr;    // device pair headers (indexed by canonical pair id)
    const void* prim;      // device PrimPair<Real> array
    const void* boys;      // device Boys table (Real)
    const HermTab* herm;   // device Hermite (t,u,v) decode table
    const void* etab;      // device Hermite -> Cartesian matrices of the primitive pairs (Real), family M
    const void* etabk;     // the same, ket form (odd Hermite orders negated)
    const void* prim64;    // FP64 primitive pairs and Boys table (family M evaluates Boys in FP64 in both precisions)
    const void* boys64;
    int G;                 // family A lanes per quartet (1, 4, 16, 64)
    EpiArgs ep;
};

using LaunchFn = hipError_t (*)(const ClassLaunch&, hipStream_t);

hipError_t launch_eri_class(int qc, bool fp32, const ClassLaunch& L, hipStream_t st);
bool class_uses_family_a(int qc);
// one-pass J/K from the stored unique array (jk_stored.hip); Jacc / Kacc must be zeroed
hipError_t launch_jk_rows(const double* eri, const int2* pairs, int K, int nspin, const double* Dj, const double* Da,
                          const double* Db, double* J, double* Ka, double* Kb, hipStream_t st);
hipError_t launch_symmetrize(double* M, int K, hipStream_t st);
// nmat <= 3; fs != nullptr also scales entry (i, j) by fs[i] fs[j]
hipError_t launch_symmetrize_n(double* M0, double* M1, double* M2, int nmat, int K, hipStream_t st,
                               const double* fs = nullptr);
// O_m = diag(fs) I_m diag(fs) for nmat <= 3 K x K matrices
hipError_t launch_scale_n(double* O0, double* O1, double* O2, const double* I0, const double* I1, const double* I2,
                          int nmat, const double* fs, int K, hipStream_t st);
hipError_t launch_zero(double* p, size_t n, hipStream_t st);
hipError_t launch_sqrt_clamp(double* p, size_t n, hipStream_t st);   // p = sqrt(max(p, 0))
hipError_t launch_axpy(double* y, const double* x, size_t n, hipStream_t st);   // y += x
hipError_t launch_add(double* z, const double* x, const double* y, size_t n, hipStream_t st);   // z = x + y
// incremental J/K builds
hipError_t launch_delta(double* dp, const double* p, double* prev, size_t n, hipStream_t st);   // dp = p - prev; prev = p
hipError_t launch_accum(double* out, double* acc, size_t n, hipStream_t st);                    // out += acc; acc = out
hipError_t launch_copy(double* dst, const double* src, size_t n, hipStream_t st);
hipError_t launch_shell_dpmax(const int2* shells, int nshell, int K, const double* a, const double* b, const double* c,
                              double* out, hipStream_t st);

}  // namespace hfx
