// scf_device.hip -- device-resident SCF iteration (see scf_device.hpp).
//
// Kernels: a tiled FP64 GEMM for the K x K products (X^T F X, F P S, X C',
// C_occ C_occ^T), Fock assembly, antisymmetrisation of F P S, and fused
// reductions (energy, Pulay B-matrix row, density RMSD) that accumulate into a
// small scalar block read back once per stage.  The loop mirrors the host
// driver in host/scf.cpp step for step.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../host/linalg.hpp"
#include "devmem.hpp"
#include "scf_device.hpp"

namespace hfxd {

namespace {

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

void hchk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("device SCF: ") + what + ": " + hipGetErrorString(e));
}
void fchk(int rc) {
    if (rc != 0) throw std::runtime_error(std::string("GPU engine: ") + hfx_last_error());
}

constexpr int MAXH = 8;            // DIIS history length (include/scf.hpp:213)
constexpr int GT = 64, GKT = 16;   // GEMM tile: 64 x 64 outputs, 16-deep k slices
constexpr int RB = 256;            // reduction block

// C[M x N] = alpha op(A) op(B) + beta C, row-major; op(A) is M x Kd, op(B) Kd x N.
// 256 threads, each owning a 4 x 4 block of the 64 x 64 tile.
__global__ __launch_bounds__(256) void k_dgemm(int M, int N, int Kd, double alpha, const double* __restrict__ A,
                                               int lda, int ta, const double* __restrict__ B, int ldb, int tb,
                                               double beta, double* __restrict__ C, int ldc) {
    __shared__ double As[GKT][GT + 1];
    __shared__ double Bs[GKT][GT + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < Kd; k0 += GKT) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int idx = threadIdx.x + r * 256;   // 0 .. 1023 over the 64 x 16 tile
            int mm, kk;
            if (ta) {
                mm = idx % GT;
                kk = idx / GT;
            } else {
                kk = idx % GKT;
                mm = idx / GKT;
            }
            const int gm = m0 + mm, gk = k0 + kk;
            double v = 0.0;
            if (gm < M && gk < Kd) v = ta ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            As[kk][mm] = v;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int idx = threadIdx.x + r * 256;
            int nn, kk;
            if (tb) {
                kk = idx % GKT;
                nn = idx / GKT;
            } else {
                nn = idx % GT;
                kk = idx / GT;
            }
            const int gn = n0 + nn, gk = k0 + kk;
            double v = 0.0;
            if (gn < N && gk < Kd) v = tb ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GKT; ++kk) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gm = m0 + ty * 4 + i, gn = n0 + tx * 4 + j;
            if (gm < M && gn < N) {
                double* c = C + (size_t)gm * ldc + gn;
                *c = alpha * acc[i][j] + (beta != 0.0 ? beta * *c : 0.0);
            }
        }
}

// F = H + J - fk Kx
__global__ void k_assemble(size_t n, const double* __restrict__ H, const double* __restrict__ J,
                           const double* __restrict__ Kx, double fk, double* __restrict__ F) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        F[i] = H[i] + (J[i] - fk * Kx[i]);
}

// R = M - M^T
__global__ void k_antisym(int K, const double* __restrict__ M, double* __restrict__ R) {
    const size_t n = (size_t)K * K;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(x / K), j = (int)(x % K);
        R[x] = M[x] - M[(size_t)j * K + i];
    }
}

__device__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < RB / 64; ++k) s += red[k];
    return s;   // valid in thread 0
}

// out[i] += <R_new, R_i> for the nslot history slots (each slot n doubles).
__global__ __launch_bounds__(RB) void k_dots(size_t n, int nslot, const double* __restrict__ Rnew,
                                             const double* __restrict__ Rhist, double* out) {
    __shared__ double red[RB / 64];
    double acc[MAXH];
#pragma unroll
    for (int s = 0; s < MAXH; ++s) acc[s] = 0.0;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x) {
        const double r = Rnew[x];
#pragma unroll
        for (int s = 0; s < MAXH; ++s)
            if (s < nslot) acc[s] += r * Rhist[(size_t)s * n + x];
    }
#pragma unroll
    for (int s = 0; s < MAXH; ++s) {
        if (s >= nslot) break;
        const double t = block_sum(acc[s], red);
        if (threadIdx.x == 0) atomicAdd(&out[s], t);
    }
}

// out += 1/2 sum_x P[x] (H[x mod K2] + F[x]) over nspin stacked matrices
__global__ __launch_bounds__(RB) void k_energy(size_t K2, int nspin, const double* __restrict__ P,
                                               const double* __restrict__ H, const double* __restrict__ F, double* out) {
    __shared__ double red[RB / 64];
    double acc = 0.0;
    const size_t n = K2 * nspin;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x)
        acc += P[x] * (H[x % K2] + F[x]);
    const double t = block_sum(0.5 * acc, red);
    if (threadIdx.x == 0) atomicAdd(out, t);
}

// out[s] += sum (A_s - B_s)^2 per spin
__global__ __launch_bounds__(RB) void k_sqdiff(size_t K2, const double* __restrict__ A, const double* __restrict__ B,
                                               double* out) {
    __shared__ double red[RB / 64];
    const int s = blockIdx.y;
    double acc = 0.0;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < K2; x += (size_t)gridDim.x * blockDim.x) {
        const double d = A[(size_t)s * K2 + x] - B[(size_t)s * K2 + x];
        acc += d * d;
    }
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) atomicAdd(&out[s], t);
}

// out = sum_i c[i] hist[i]
__global__ void k_lincomb(size_t n, int nslot, const double* __restrict__ c, const double* __restrict__ hist,
                          double* __restrict__ out) {
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int s = 0; s < nslot; ++s) v += c[s] * hist[(size_t)s * n + x];
        out[x] = v;
    }
}

// P = d Pn + (1 - d) P
__global__ void k_damp(size_t n, double d, const double* __restrict__ Pn, double* __restrict__ P) {
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x)
        P[x] = d * Pn[x] + (1.0 - d) * P[x];
}

// Copies between device memory and pinned host memory go through this kernel
// (the GPU reads / writes the pinned block over the host link): an SDMA copy
// of the same few kilobytes costs ~0.1 ms of latency.
__global__ void k_copy(size_t n, const double* __restrict__ src, double* __restrict__ dst) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

__global__ void k_fill(size_t n, double v, double* p) {
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x) p[x] = v;
}

unsigned grid_for(size_t n, unsigned block = 256, unsigned cap = 1024) {
    size_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g > cap ? cap : g);
}

}  // namespace

// scalar block layout
enum : int { SC_DOT = 0, SC_E = MAXH, SC_RMSD = MAXH + 1, SC_N = MAXH + 3 };

struct DevScf::Impl {
    int K = 0, nspin = 1;
    size_t K2 = 0;
    std::vector<int> nocc;
    double occ = 2.0;
    hipStream_t st = nullptr;
    double *H = nullptr, *S = nullptr, *X = nullptr;
    double *P = nullptr, *Pn = nullptr, *C = nullptr;          // nspin x K2 each
    double *J = nullptr, *Ka = nullptr, *Kb = nullptr;
    double *hF = nullptr, *hR = nullptr;                       // MAXH x nspin x K2
    double *Fuse = nullptr, *Fc = nullptr, *Rc = nullptr;      // nspin x K2
    double *T1 = nullptr, *T2 = nullptr;                       // K2
    double *sc = nullptr, *coef = nullptr;                     // SC_N, MAXH
    double *h_sc = nullptr, *h_mat = nullptr, *h_coef = nullptr;   // pinned host
    double *h_sc_dev = nullptr, *h_mat_dev = nullptr, *h_coef_dev = nullptr;   // the same, as the device sees them

    void copy(const double* src, double* dst, size_t n) {
        k_copy<<<grid_for(n), 256, 0, st>>>(n, src, dst);
        hchk(hipGetLastError(), "copy launch");
    }

    void gemm(int M, int N, int Kd, double alpha, const double* A, int lda, bool ta, const double* B, int ldb,
              bool tb, double* Cm, int ldc) {
        dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
        k_dgemm<<<grid, 256, 0, st>>>(M, N, Kd, alpha, A, lda, ta ? 1 : 0, B, ldb, tb ? 1 : 0, 0.0, Cm, ldc);
        hchk(hipGetLastError(), "gemm launch");
    }
    void zero_scalars() { k_fill<<<1, 64, 0, st>>>(SC_N, 0.0, sc); }
    void read_scalars() {
        copy(sc, h_sc_dev, SC_N);
        hchk(hipStreamSynchronize(st), "sync");
    }

    // Fock matrices of density D into F (nspin x K2) and error vectors into R;
    // leaves E and <R, R_i> for the given slots in the scalar block.
    void fock_error(hfx_fock* f, const double* D, double* F, double* R, const double* Rhist, int nslot) {
        fchk(hfx_fock_build_device(f, D, nspin == 2 ? D + K2 : nullptr, J, Ka, nspin == 2 ? Kb : nullptr, st));
        for (int s = 0; s < nspin; ++s) {
            k_assemble<<<grid_for(K2), 256, 0, st>>>(K2, H, J, s == 0 ? Ka : Kb, nspin == 1 ? 0.5 : 1.0,
                                                     F + (size_t)s * K2);
            gemm(K, K, K, 1.0, F + (size_t)s * K2, K, false, D + (size_t)s * K2, K, false, T1, K);
            gemm(K, K, K, 1.0, T1, K, false, S, K, false, T2, K);
            k_antisym<<<grid_for(K2), 256, 0, st>>>(K, T2, R + (size_t)s * K2);
        }
        zero_scalars();
        k_energy<<<grid_for(K2 * nspin, RB, 256), RB, 0, st>>>(K2, nspin, D, H, F, sc + SC_E);
        if (nslot > 0)
            k_dots<<<grid_for(K2 * nspin, RB, 256), RB, 0, st>>>(K2 * nspin, nslot, R, Rhist, sc + SC_DOT);
        hchk(hipGetLastError(), "fock/error launch");
        read_scalars();
    }
};

DevScf::DevScf(int K, int nspin, const int* nocc, const double* H, const double* S, const double* X)
    : p_(new Impl()) {
    Impl& d = *p_;
    d.K = K;
    d.nspin = nspin;
    d.K2 = (size_t)K * K;
    d.nocc.assign(nocc, nocc + nspin);
    d.occ = nspin == 1 ? 2.0 : 1.0;
    e_.assign(nspin, std::vector<double>(K, 0.0));
    const size_t K2 = d.K2, ns = (size_t)nspin;
    auto alloc = [&](double** p, size_t n) { hchk(hfx_detail::dev_alloc((void**)p, n * sizeof(double)), "device allocation"); };
    try {
        hchk(hipStreamCreateWithFlags(&d.st, hipStreamNonBlocking), "stream");
        alloc(&d.H, K2);
        alloc(&d.S, K2);
        alloc(&d.X, K2);
        alloc(&d.P, ns * K2);
        alloc(&d.Pn, ns * K2);
        alloc(&d.C, ns * K2);
        alloc(&d.J, K2);
        alloc(&d.Ka, K2);
        alloc(&d.Kb, K2);
        alloc(&d.hF, MAXH * ns * K2);
        alloc(&d.hR, MAXH * ns * K2);
        alloc(&d.Fuse, ns * K2);
        alloc(&d.Fc, ns * K2);
        alloc(&d.Rc, ns * K2);
        alloc(&d.T1, K2);
        alloc(&d.T2, K2);
        alloc(&d.sc, SC_N);
        alloc(&d.coef, MAXH);
        hchk(hfx_detail::host_alloc((void**)&d.h_sc, SC_N * sizeof(double)), "pinned allocation");
        hchk(hfx_detail::host_alloc((void**)&d.h_mat, K2 * sizeof(double)), "pinned allocation");
        hchk(hfx_detail::host_alloc((void**)&d.h_coef, MAXH * sizeof(double)), "pinned allocation");
        hchk(hipHostGetDevicePointer((void**)&d.h_sc_dev, d.h_sc, 0), "pinned device pointer");
        hchk(hipHostGetDevicePointer((void**)&d.h_mat_dev, d.h_mat, 0), "pinned device pointer");
        hchk(hipHostGetDevicePointer((void**)&d.h_coef_dev, d.h_coef, 0), "pinned device pointer");
        hchk(hipMemcpyAsync(d.H, H, K2 * sizeof(double), hipMemcpyHostToDevice, d.st), "H2D");
        hchk(hipMemcpyAsync(d.S, S, K2 * sizeof(double), hipMemcpyHostToDevice, d.st), "H2D");
        hchk(hipMemcpyAsync(d.X, X, K2 * sizeof(double), hipMemcpyHostToDevice, d.st), "H2D");
        k_fill<<<grid_for(ns * K2), 256, 0, d.st>>>(ns * K2, 0.0, d.P);
        hchk(hipStreamSynchronize(d.st), "sync");
    } catch (...) {
        release();
        throw;
    }
}

DevScf::~DevScf() { release(); }

void DevScf::release() {
    if (!p_) return;
    Impl& d = *p_;
    if (d.st) (void)hipStreamSynchronize(d.st);
    for (double* q : {d.H, d.S, d.X, d.P, d.Pn, d.C, d.J, d.Ka, d.Kb, d.hF, d.hR, d.Fuse, d.Fc, d.Rc, d.T1, d.T2, d.sc,
                      d.coef})
        hfx_detail::dev_free(q);
    for (double* q : {d.h_sc, d.h_mat, d.h_coef})
        hfx_detail::host_free(q);
    if (d.st) (void)hipStreamDestroy(d.st);
    delete p_;
    p_ = nullptr;
}

void DevScf::set_density(int spin, const double* P) {
    Impl& d = *p_;
    hchk(hipMemcpyAsync(d.P + (size_t)spin * d.K2, P, d.K2 * sizeof(double), hipMemcpyHostToDevice, d.st), "H2D");
    hchk(hipStreamSynchronize(d.st), "sync");
}

void DevScf::get_density(int spin, double* P) {
    Impl& d = *p_;
    hchk(hipMemcpyAsync(P, d.P + (size_t)spin * d.K2, d.K2 * sizeof(double), hipMemcpyDeviceToHost, d.st), "D2H");
    hchk(hipStreamSynchronize(d.st), "sync");
}

void DevScf::get_coefficients(int spin, double* Cm) {
    Impl& d = *p_;
    hchk(hipMemcpyAsync(Cm, d.C + (size_t)spin * d.K2, d.K2 * sizeof(double), hipMemcpyDeviceToHost, d.st), "D2H");
    hchk(hipStreamSynchronize(d.st), "sync");
}

bool DevScf::run(int max_it, int start_diis, bool use_damping, double eps, double damping, double e_conv,
                 bool verbose, int& n_it, double& E_ele, DevScfTimes& t) {
    if (!f_) throw std::runtime_error("device SCF: no Fock builder");
    Impl& d = *p_;
    const int K = d.K, ns = d.nspin;
    const size_t K2 = d.K2, nK2 = (size_t)ns * K2;
    double B[MAXH][MAXH] = {};
    int nh = 0, head = 0;
    double E_prev = 0.0;
    bool have_prev = false;
    std::vector<double> w(K), Cp(K2);
    // F' of the last diagonalisation per spin; all_orbitals() solves it again
    // with every eigenvector for C (the iterations form the occupied ones)
    std::vector<std::vector<double>> last_Fp(ns);
    auto all_orbitals = [&]() {
        for (int s = 0; s < ns; ++s) {
            if (last_Fp[s].empty()) continue;
            hfxh::sym_eig(K, last_Fp[s].data(), w.data(), Cp.data());
            e_[s] = w;
            for (size_t x = 0; x < K2; ++x) d.h_mat[x] = Cp[x];
            d.copy(d.h_mat_dev, d.T1, K2);
            d.gemm(K, K, K, 1.0, d.X, K, false, d.T1, K, false, d.C + (size_t)s * K2, K);
            hchk(hipStreamSynchronize(d.st), "sync");   // h_mat is reused by the next spin
        }
    };
    for (int it = 0; it < max_it; ++it) {
        ++n_it;
        auto t0 = clk::now();
        // Fock + error into history slot `head`; dots against every valid slot
        const int nvalid = nh < MAXH ? nh + 1 : MAXH;
        double* F = d.hF + (size_t)head * nK2;
        double* R = d.hR + (size_t)head * nK2;
        d.fock_error(f_, d.P, F, R, d.hR, nvalid);
        nh = nvalid;
        for (int j = 0; j < nh; ++j) B[head][j] = B[j][head] = d.h_sc[SC_DOT + j];
        const double E_it = d.h_sc[SC_E];
        const double err = std::sqrt(std::max(0.0, B[head][head]));
        auto t1 = clk::now();
        t.fock += secs(t0, t1);
        // DIIS extrapolation (augmented Pulay system, include/linalg.hpp:405-511)
        const double* Fuse = F;
        if (it + 1 >= start_diis) {
            const int n = nh;
            std::vector<double> A((size_t)(n + 1) * (n + 1), 0.0), rhs(n + 1, 0.0), c(n + 1, 0.0);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) A[(size_t)i * (n + 1) + j] = B[i][j];
            for (int i = 0; i < n; ++i) {
                A[(size_t)i * (n + 1) + n] = -1.0;
                A[(size_t)n * (n + 1) + i] = 1.0;
            }
            rhs[n] = 1.0;
            hfxh::solve_fullpiv(n + 1, A.data(), rhs.data(), c.data());
            for (int i = 0; i < n; ++i) d.h_coef[i] = c[i];
            d.copy(d.h_coef_dev, d.coef, (size_t)n);
            k_lincomb<<<grid_for(nK2), 256, 0, d.st>>>(nK2, n, d.coef, d.hF, d.Fuse);
            Fuse = d.Fuse;
        }
        // F' = X^T F X, host eigen-solve, C = X C', P_new = occ C_occ C_occ^T
        for (int s = 0; s < ns; ++s) {
            d.gemm(K, K, K, 1.0, d.X, K, true, Fuse + (size_t)s * K2, K, false, d.T1, K);
            d.gemm(K, K, K, 1.0, d.T1, K, false, d.X, K, false, d.T2, K);
            d.copy(d.T2, d.h_mat_dev, K2);
            hchk(hipStreamSynchronize(d.st), "sync");
            last_Fp[s].assign(d.h_mat, d.h_mat + K2);
            // occupied orbitals only (all orbital energies); all_orbitals()
            // forms the rest once the phase ends
            hfxh::sym_eig_lowest(K, d.h_mat, d.nocc[s], w.data(), Cp.data());
            e_[s] = w;
            for (size_t x = 0; x < K2; ++x) d.h_mat[x] = Cp[x];
            d.copy(d.h_mat_dev, d.T1, K2);
            double* Cs = d.C + (size_t)s * K2;
            d.gemm(K, K, K, 1.0, d.X, K, false, d.T1, K, false, Cs, K);
            d.gemm(K, K, d.nocc[s], d.occ, Cs, K, false, Cs, K, true, d.Pn + (size_t)s * K2, K);
        }
        d.zero_scalars();
        k_sqdiff<<<dim3(grid_for(K2, RB, 256), ns), RB, 0, d.st>>>(K2, d.Pn, d.P, d.sc + SC_RMSD);
        d.read_scalars();
        double dmax = 0.0;
        for (int s = 0; s < ns; ++s) dmax = std::max(dmax, std::sqrt(d.h_sc[SC_RMSD + s] / (double)K2));
        head = (head + 1) % MAXH;
        t.diag += secs(t1, clk::now());
        const bool econv = e_conv <= 0.0 || (have_prev && std::fabs(E_it - E_prev) < e_conv);
        E_prev = E_it;
        have_prev = true;
        if (verbose) std::printf("  iter %3d  E_ele %.12f  dP %.3e  |R| %.3e  (device)\n", n_it, E_it, dmax, err);
        if (dmax < eps && err < eps && econv) {
            auto tc = clk::now();
            d.fock_error(f_, d.Pn, d.Fc, d.Rc, d.Rc, 1);   // <Rc, Rc> lands in slot 0
            t.fock += secs(tc, clk::now());
            const double chk = std::sqrt(std::max(0.0, d.h_sc[SC_DOT]));
            if (chk < eps) {
                hchk(hipMemcpyAsync(d.P, d.Pn, nK2 * sizeof(double), hipMemcpyDeviceToDevice, d.st), "D2D");
                all_orbitals();
                E_ele = d.h_sc[SC_E];
                return true;
            }
            if (verbose) std::printf("  final check failed, error: %.3e\n", chk);
        }
        if (!use_damping || it + 1 >= start_diis)
            hchk(hipMemcpyAsync(d.P, d.Pn, nK2 * sizeof(double), hipMemcpyDeviceToDevice, d.st), "D2D");
        else
            k_damp<<<grid_for(nK2), 256, 0, d.st>>>(nK2, damping, d.Pn, d.P);
        if (it == max_it - 1) {
            d.fock_error(f_, d.P, d.Fc, d.Rc, d.Rc, 1);
            E_ele = d.h_sc[SC_E];
            if (verbose) std::printf("  last iteration reached\n");
        }
    }
    all_orbitals();
    return false;
}

}  // namespace hfxd
