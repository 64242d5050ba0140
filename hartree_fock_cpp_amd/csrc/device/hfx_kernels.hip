// hfx_kernels.hip -- __global__ entry points for gfx950 and their launch table.
//
//  * class dispatch to the ERI kernels instantiated in eri_part0..4.hip
//    (k_fam_a: one lane per shell quartet; k_fam_b: one wavefront per shell quartet)
//  * k_symmetrize, k_zero, k_axpy : small helpers of the Fock build
//
// Every ERI kernel reads a class-homogeneous list of (bra pair, ket pair)
// ids, so one launch covers one angular class and one precision.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "eri_body.hpp"
#include "eri_classes.hpp"
#include "hfx_launch.hpp"

namespace hfx {

hipError_t launch_eri_part0(int, bool, const ClassLaunch&, hipStream_t);
hipError_t launch_eri_part1(int, bool, const ClassLaunch&, hipStream_t);
hipError_t launch_eri_part2(int, bool, const ClassLaunch&, hipStream_t);
hipError_t launch_eri_part3(int, bool, const ClassLaunch&, hipStream_t);
hipError_t launch_eri_part4(int, bool, const ClassLaunch&, hipStream_t);
hipError_t launch_eri_part5(int, bool, const ClassLaunch&, hipStream_t);
hipError_t launch_eri_part6(int, bool, const ClassLaunch&, hipStream_t);
hipError_t launch_eri_part7(int, bool, const ClassLaunch&, hipStream_t);

// Class ranges instantiated by eri_part0..7.hip.
hipError_t launch_eri_class(int qc, bool fp32, const ClassLaunch& L, hipStream_t st) {
    if (qc < 0 || qc >= NQCLASS) return hipErrorInvalidValue;
    if (qc < 4) return launch_eri_part0(qc, fp32, L, st);
    if (qc < 7) return launch_eri_part1(qc, fp32, L, st);
    if (qc < 10) return launch_eri_part2(qc, fp32, L, st);
    if (qc < 13) return launch_eri_part3(qc, fp32, L, st);
    if (qc < 21) return launch_eri_part4(qc, fp32, L, st);
    if (qc < 30) return launch_eri_part5(qc, fp32, L, st);
    if (qc < 42) return launch_eri_part6(qc, fp32, L, st);
    return launch_eri_part7(qc, fp32, L, st);
}

bool class_uses_family_a(int qc) {
    int pb = qc_bra(qc), pk = qc_ket(qc);
    return runs_family_a(pc_la(pb), pc_lb(pb), pc_la(pk), pc_lb(pk));
}

// ------------------------------------------------------------------ finish
// J <- Jacc + Jacc^T ; Kx <- Kacc + Kacc^T (in place, one thread per i <= j pair)
__global__ void k_symmetrize(double* M, int K) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np = (int64_t)K * (K + 1) / 2;
    if (p >= np) return;
    int i = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while ((int64_t)i * (i + 1) / 2 > p) --i;
    while ((int64_t)(i + 1) * (i + 2) / 2 <= p) ++i;
    const int j = (int)(p - (int64_t)i * (i + 1) / 2);
    const double s = M[i * K + j] + M[j * K + i];
    M[i * K + j] = s;
    M[j * K + i] = s;
}

// The same for up to three matrices in one launch (blockIdx.y picks one);
// fs != nullptr also scales entry (i, j) by fs[i] fs[j].
__global__ void k_symmetrize3(double* M0, double* M1, double* M2, int K, const double* fs) {
    double* M = blockIdx.y == 0 ? M0 : blockIdx.y == 1 ? M1 : M2;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np = (int64_t)K * (K + 1) / 2;
    if (p >= np) return;
    int i = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while ((int64_t)i * (i + 1) / 2 > p) --i;
    while ((int64_t)(i + 1) * (i + 2) / 2 <= p) ++i;
    const int j = (int)(p - (int64_t)i * (i + 1) / 2);
    double s = M[i * K + j] + M[j * K + i];
    if (fs) s *= fs[i] * fs[j];
    M[i * K + j] = s;
    M[j * K + i] = s;
}

hipError_t launch_symmetrize_n(double* M0, double* M1, double* M2, int nmat, int K, hipStream_t st,
                               const double* fs) {
    const int64_t np = (int64_t)K * (K + 1) / 2;
    const int block = 256;
    k_symmetrize3<<<dim3((unsigned)((np + block - 1) / block), (unsigned)nmat), block, 0, st>>>(M0, M1, M2, K, fs);
    return hipGetLastError();
}

// out_m[i K + j] = fs[i] fs[j] in_m[i K + j] for up to three K x K matrices
// (blockIdx.y picks one).
__global__ void k_scale3(double* O0, double* O1, double* O2, const double* I0, const double* I1, const double* I2,
                         const double* fs, int K) {
    double* O = blockIdx.y == 0 ? O0 : blockIdx.y == 1 ? O1 : O2;
    const double* I = blockIdx.y == 0 ? I0 : blockIdx.y == 1 ? I1 : I2;
    const int64_t n = (int64_t)K * K;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(x / K), j = (int)(x - (int64_t)i * K);
        O[x] = fs[i] * fs[j] * I[x];
    }
}

hipError_t launch_scale_n(double* O0, double* O1, double* O2, const double* I0, const double* I1, const double* I2,
                          int nmat, const double* fs, int K, hipStream_t st) {
    const int64_t n = (int64_t)K * K;
    const int block = 256;
    const int64_t nb = (n + block - 1) / block;
    k_scale3<<<dim3((unsigned)(nb < 4096 ? nb : 4096), (unsigned)nmat), block, 0, st>>>(O0, O1, O2, I0, I1, I2, fs, K);
    return hipGetLastError();
}

hipError_t launch_symmetrize(double* M, int K, hipStream_t st) {
    const int64_t np = (int64_t)K * (K + 1) / 2;
    const int block = 256;
    k_symmetrize<<<(unsigned)((np + block - 1) / block), block, 0, st>>>(M, K);
    return hipGetLastError();
}

__global__ void k_zero(double* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

__global__ void k_sqrt_clamp(double* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = p[i] > 0.0 ? sqrt(p[i]) : 0.0;
}

hipError_t launch_sqrt_clamp(double* p, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    k_sqrt_clamp<<<(unsigned)(blocks < 1024 ? blocks : 1024), 256, 0, st>>>(p, n);
    return hipGetLastError();
}

__global__ void k_axpy(double* y, const double* x, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] += x[i];
}

hipError_t launch_axpy(double* y, const double* x, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    k_axpy<<<(unsigned)blocks, 256, 0, st>>>(y, x, n);
    return hipGetLastError();
}

// z = x + y (a kernel, not a memcpy + add, so the Fock build can be captured
// into a graph with kernel nodes only)
__global__ void k_add(double* z, const double* x, const double* y, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        z[i] = x[i] + y[i];
}

hipError_t launch_add(double* z, const double* x, const double* y, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    k_add<<<(unsigned)blocks, 256, 0, st>>>(z, x, y, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------ incremental builds
// dp = p - prev, then prev = p
__global__ void k_delta(double* dp, const double* p, double* prev, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double v = p[i];
        dp[i] = v - prev[i];
        prev[i] = v;
    }
}

// out += acc, then acc = out
__global__ void k_accum(double* out, double* acc, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double v = out[i] + acc[i];
        out[i] = v;
        acc[i] = v;
    }
}

__global__ void k_copy(double* dst, const double* src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

// out[A * nshell + B] = max over the (A, B) block of |a|, |b|, |c| (b, c may
// be null); shells[A] = (first function, function count).  One thread per
// shell pair.
__global__ void k_shell_dpmax(const int2* shells, int nshell, int K, const double* a, const double* b, const double* c,
                              double* out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nshell * nshell) return;
    const int A = (int)(t / nshell), B = (int)(t % nshell);
    const int2 sa = shells[A], sb = shells[B];
    double m = 0.0;
    for (int i = sa.x; i < sa.x + sa.y; ++i)
        for (int j = sb.x; j < sb.x + sb.y; ++j) {
            const size_t o = (size_t)i * K + j;
            m = fmax(m, fabs(a[o]));
            if (b) m = fmax(m, fabs(b[o]));
            if (c) m = fmax(m, fabs(c[o]));
        }
    out[t] = m;
}

namespace {
unsigned grid_1d(size_t n) {
    size_t blocks = (n + 255) / 256;
    return (unsigned)(blocks > 65536 ? 65536 : blocks == 0 ? 1 : blocks);
}
}  // namespace

hipError_t launch_delta(double* dp, const double* p, double* prev, size_t n, hipStream_t st) {
    k_delta<<<grid_1d(n), 256, 0, st>>>(dp, p, prev, n);
    return hipGetLastError();
}

hipError_t launch_accum(double* out, double* acc, size_t n, hipStream_t st) {
    k_accum<<<grid_1d(n), 256, 0, st>>>(out, acc, n);
    return hipGetLastError();
}

hipError_t launch_copy(double* dst, const double* src, size_t n, hipStream_t st) {
    k_copy<<<grid_1d(n), 256, 0, st>>>(dst, src, n);
    return hipGetLastError();
}

hipError_t launch_shell_dpmax(const int2* shells, int nshell, int K, const double* a, const double* b, const double* c,
                              double* out, hipStream_t st) {
    const int64_t n = (int64_t)nshell * nshell;
    k_shell_dpmax<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(shells, nshell, K, a, b, c, out);
    return hipGetLastError();
}

hipError_t launch_zero(double* p, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    k_zero<<<(unsigned)blocks, 256, 0, st>>>(p, n);
    return hipGetLastError();
}

}  // namespace hfx
