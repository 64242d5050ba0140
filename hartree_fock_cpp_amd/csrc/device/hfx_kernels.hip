// hfx_kernels.hip -- __global__ entry points for gfx950 and their launch table.
//
//  * class dispatch to the ERI kernels instantiated in eri_part0..4.hip
//    (k_fam_a: one lane per shell quartet; k_fam_b: one wavefront per shell quartet)
//  * k_jk_unique            : J/K from the stored 8-fold unique ERI array
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
    return use_family_a(pc_la(pb), pc_lb(pb), pc_la(pk), pc_lb(pk));
}

// ------------------------------------------------------------------ stored-ERI J/K
// One workgroup per function-pair row IJ = pair(i, j) of the unique array
// (rows longest-first).  Row elements (ij|kl), KL <= IJ, are read once,
// coalesced.  With f the function-level degeneracy factor:
//   Jacc_ij += 2f v Dj_kl   (block reduction, one atomic per row)
//   Kacc_ik += f v D_jl, Kacc_il += f v D_jk  (row i, LDS accumulator)
//   Kacc_jk += f v D_il, Kacc_jl += f v D_ik  (row j, LDS accumulator)
// Pairs are ordered k-major, so the lanes of a wave hold a few runs of equal
// k: the k-keyed terms are summed over each run with a segmented wave scan
// (DPP lane shifts, no LDS traffic) and added by the run's last lane, one LDS
// atomic per run instead of 64 atomics on one address; the l-keyed terms hit
// distinct addresses.  Each thread keeps JK_UNROLL elements in flight (their
// pair ids, integrals and Dj_kl gathers issue together); reading (k, l) from
// the pair table measured faster than computing it from KL with sqrtf.
// The column terms Jacc_kl += 2f v Dj_ij are k_j_columns' (a per-element
// atomic here would cost one global atomic per stored integral).
constexpr int JK_BLOCK = 256;
constexpr int JK_UNROLL = 4;
constexpr int JK_MAXK = 900;   // 2 spins x 4 rows x K doubles must fit 64 KiB of LDS

// value of lane (lane - shift) under DPP control CTRL, 0 where that lane lies
// outside the row or the destination row is masked off by RM
template <int CTRL, int RM>
__device__ inline int dpp_int(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, RM, 0xf, false);
}
template <int CTRL, int RM>
__device__ inline double dpp_dbl(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = dpp_int<CTRL, RM>((int)b);
    const int hi = dpp_int<CTRL, RM>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL, int RM>
__device__ inline void run_step(double& a, double& b, int key) {
    const int ok = dpp_int<CTRL, RM>(key);
    const double oa = dpp_dbl<CTRL, RM>(a), ob = dpp_dbl<CTRL, RM>(b);
    if (ok == key) {
        a += oa;
        b += ob;
    }
}
// Inclusive scan of (a, b) over runs of equal key (key >= 1, nondecreasing
// over the wave's lanes; a shifted-in 0 never matches): row_shr 1, 2, 4, 8
// within each 16-lane row, then row_bcast 15 (rows 1, 3) and row_bcast 31
// (rows 2, 3).  Afterwards the last lane of a run holds the run sums.  All
// 64 lanes must be active.
__device__ inline void run_scan2(double& a, double& b, int key) {
    run_step<0x111, 0xf>(a, b, key);
    run_step<0x112, 0xf>(a, b, key);
    run_step<0x114, 0xf>(a, b, key);
    run_step<0x118, 0xf>(a, b, key);
    run_step<0x142, 0xa>(a, b, key);
    run_step<0x143, 0xc>(a, b, key);
}

__global__ __launch_bounds__(JK_BLOCK) void k_jk_unique(const double* __restrict__ eri, const int2* __restrict__ pairs,
                                                         int K, int npair, int nspin, const double* __restrict__ Dj,
                                                         const double* __restrict__ Da, const double* __restrict__ Db,
                                                         double* J, double* Ka, double* Kb) {
    extern __shared__ double lds[];
    // layout: [spin][ Krow_i[K], Krow_j[K], D_i[K], D_j[K] ], then red[JK_BLOCK/64]
    const int IJ = npair - 1 - (int)blockIdx.x;
    const int2 ij = pairs[IJ];
    const int i = ij.x, j = ij.y;
    const int tid = threadIdx.x, lane = tid & 63;
    for (int s = 0; s < nspin; ++s) {
        double* base = lds + s * 4 * K;
        const double* D = s == 0 ? Da : Db;
        for (int c = tid; c < K; c += JK_BLOCK) {
            base[c] = 0.0;
            base[K + c] = 0.0;
            base[2 * K + c] = D[i * K + c];
            base[3 * K + c] = D[j * K + c];
        }
    }
    __syncthreads();
    const double fij = (i == j) ? 0.5 : 1.0;
    const size_t rowbase = (size_t)IJ * (IJ + 1) / 2;
    double srow = 0.0;
    // the trip count is uniform over the block: every lane takes part in the scans
    for (int KL0 = 0; KL0 <= IJ; KL0 += JK_BLOCK * JK_UNROLL) {
        double v[JK_UNROLL], dkl[JK_UNROLL];
        int kk[JK_UNROLL], ll[JK_UNROLL];
#pragma unroll
        for (int u = 0; u < JK_UNROLL; ++u) {
            const int KL = KL0 + u * JK_BLOCK + tid;
            const bool in = KL <= IJ;
            const int2 kl = pairs[in ? KL : 0];
            const int k = kl.x, l = kl.y;
            kk[u] = in ? k : 0x3fffffff;   // lanes past the row form a run of their own
            ll[u] = l;
            v[u] = in ? eri[rowbase + KL] : 0.0;
            dkl[u] = in ? Dj[k * K + l] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < JK_UNROLL; ++u) {
            const int KL = KL0 + u * JK_BLOCK + tid;
            const int k = kk[u], l = ll[u];
            const double fv = fij * (k == l ? 0.5 : 1.0) * (KL == IJ ? 0.5 : 1.0) * v[u];
            srow += 2.0 * fv * dkl[u];
            const bool last = lane == 63 || l == k || KL == IJ;
            for (int s = 0; s < nspin; ++s) {
                double* base = lds + s * 4 * K;
                const double* Di = base + 2 * K;
                const double* Dj_ = base + 3 * K;
                double a = 0.0, b = 0.0;
                if (fv != 0.0) {
                    a = fv * Dj_[l];                        // Kacc_ik += f v D_jl
                    b = fv * Di[l];                         // Kacc_jk += f v D_il
                    atomicAdd(&base[l], fv * Dj_[k]);       // Kacc_il += f v D_jk
                    atomicAdd(&base[K + l], fv * Di[k]);    // Kacc_jl += f v D_ik
                }
                run_scan2(a, b, k + 1);
                if (last && k < K) {
                    if (a != 0.0) atomicAdd(&base[k], a);
                    if (b != 0.0) atomicAdd(&base[K + k], b);
                }
            }
        }
    }
    // block reduction of the row sum
    for (int off = 32; off > 0; off >>= 1) srow += __shfl_down(srow, off, 64);
    double* red = lds + nspin * 4 * K;
    if (lane == 0) red[tid >> 6] = srow;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < JK_BLOCK / 64; ++w) s += red[w];
        unsafeAtomicAdd(&J[i * K + j], s);
    }
    for (int s = 0; s < nspin; ++s) {
        double* base = lds + s * 4 * K;
        double* Kx = s == 0 ? Ka : Kb;
        for (int c = tid; c < K; c += JK_BLOCK) {
            if (base[c] != 0.0) unsafeAtomicAdd(&Kx[i * K + c], base[c]);
            if (base[K + c] != 0.0) unsafeAtomicAdd(&Kx[j * K + c], base[K + c]);
        }
    }
}

// Column terms of J from the unique array: Jacc_kl += sum over rows IJ >= KL
// of 2 f v(IJ, KL) Dj_ij.  A workgroup takes a block of JC_COLS columns (one
// per thread) against a chunk of `rows` rows, reads each row's segment of
// the block (coalesced) and adds its column sums with one atomic per column:
// (P / rows) x P atomics instead of one per stored integral.  rows shrinks
// with P (16 .. 256) so small matrices still spread over many workgroups.
constexpr int JC_COLS = 256;

__global__ __launch_bounds__(JC_COLS) void k_j_columns(const double* __restrict__ eri, const int2* __restrict__ pairs,
                                                       int K, int npair, int rows, const double* __restrict__ Dj,
                                                       double* J) {
    const int KL = (int)blockIdx.x * JC_COLS + (int)threadIdx.x;
    const int col0 = (int)blockIdx.x * JC_COLS;
    const int r_lo = max((int)blockIdx.y * rows, col0);
    const int r_hi = min((int)(blockIdx.y + 1) * rows, npair);
    if (r_lo >= r_hi) return;   // uniform: the chunk lies above the block's diagonal
    double acc = 0.0;
    for (int IJ = r_lo; IJ < r_hi; ++IJ) {
        const int2 ij = pairs[IJ];
        const double rowf = 2.0 * (ij.x == ij.y ? 0.5 : 1.0) * Dj[ij.x * K + ij.y];
        if (KL <= IJ) {
            const double v = eri[(size_t)IJ * (IJ + 1) / 2 + KL];
            acc += (KL == IJ ? 0.5 : 1.0) * rowf * v;
        }
    }
    if (KL < npair && acc != 0.0) {
        const int2 kl = pairs[KL];
        unsafeAtomicAdd(&J[kl.x * K + kl.y], (kl.x == kl.y ? 0.5 : 1.0) * acc);
    }
}

hipError_t launch_jk_unique(const double* eri, const int2* pairs, int K, int nspin, const double* Dj, const double* Da,
                            const double* Db, double* J, double* Ka, double* Kb, hipStream_t st) {
    if (K > JK_MAXK) return hipErrorInvalidValue;
    const int npair = K * (K + 1) / 2;
    const size_t lds = sizeof(double) * (size_t)(nspin * 4 * K + JK_BLOCK / 64);
    k_jk_unique<<<npair, JK_BLOCK, lds, st>>>(eri, pairs, K, npair, nspin, Dj, Da, Db, J, Ka, Kb);
    if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    const int rows = std::min(256, std::max(16, npair / 64));
    const dim3 grid((unsigned)((npair + JC_COLS - 1) / JC_COLS), (unsigned)((npair + rows - 1) / rows));
    k_j_columns<<<grid, JC_COLS, 0, st>>>(eri, pairs, K, npair, rows, Dj, J);
    return hipGetLastError();
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

// The same for up to three matrices in one launch (blockIdx.y picks one).
__global__ void k_symmetrize3(double* M0, double* M1, double* M2, int K) {
    double* M = blockIdx.y == 0 ? M0 : blockIdx.y == 1 ? M1 : M2;
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

hipError_t launch_symmetrize_n(double* M0, double* M1, double* M2, int nmat, int K, hipStream_t st) {
    const int64_t np = (int64_t)K * (K + 1) / 2;
    const int block = 256;
    k_symmetrize3<<<dim3((unsigned)((np + block - 1) / block), (unsigned)nmat), block, 0, st>>>(M0, M1, M2, K);
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
