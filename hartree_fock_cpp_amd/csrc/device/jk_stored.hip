// jk_stored.hip -- J/K from the stored 8-fold unique ERI array in one pass.
//
// Replaces the reference's host loops coulomb_matrix / exchange_matrix /
// repulsion_matrix (include/linalg.hpp:100-129, 215-296), which visit every
// (ij|kl) of the K^4 index space; here every stored integral is read from HBM
// exactly once and contributes its 8 symmetry images.  With f the weight of
// the integral among its images (jkr_weight) and Dj the density of J, D the
// density of K (per spin):
//   Jacc_ij += 2f v Dj_kl   Jacc_kl += 2f v Dj_ij
//   Kacc_ik += f v D_jl     Kacc_il += f v D_jk
//   Kacc_jk += f v D_il     Kacc_jl += f v D_ik
// and J = Jacc + Jacc^T, K = Kacc + Kacc^T afterwards (k_symmetrize3).
//
// Work decomposition (jk_rows.hpp): a workgroup takes one row block (up to R
// rows IJ = (i, j0..j0+R-1) of one i) and one segment of its columns; blocks
// of larger i come first.  Each lane owns one column KL = (k, l) of a
// JKR_CHUNK-column step and walks the block's rows, so one row's integrals
// arrive as one coalesced load per wave.  Per column, the terms keyed by
// (i, .) or by the column alone are summed over the rows in registers
// (Jacc_kl, Kacc_ik, Kacc_il); per row, Jacc_ij is a register sum over the
// columns, Kacc_jl (a distinct l per lane within a run of equal k) goes to an
// LDS row by LDS atomics, and Kacc_jk (equal over runs of lanes, since pairs
// are k-major) is summed over each run inside 16-lane rows by a segmented DPP
// scan whose lane masks depend only on the columns (built once per step);
// the last lane of each run piece adds the piece by an LDS atomic.  The
// densities of the block's rows sit in LDS, pre-scaled by their row weight,
// at a compile-time row stride KS so that row addresses are immediate offsets.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "hfx_launch.hpp"
#include "jk_rows.hpp"

namespace hfx {

constexpr int JKR_THREADS = 256;
constexpr int JKR_CHUNK = JKR_THREADS;   // columns per workgroup step (one per lane)
constexpr int JKR_WAVES = JKR_THREADS / 64;
constexpr size_t JKR_LDS_MAX = 80 * 1024;

// value of lane (lane - shift) under DPP control CTRL, 0 where that lane lies
// outside its 16-lane row
template <int CTRL>
__device__ inline int jkr_dpp_int(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ inline double jkr_dpp_dbl(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = jkr_dpp_int<CTRL>((int)b);
    const int hi = jkr_dpp_int<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Inclusive scan over runs of equal key inside each 16-lane row (key >= 1,
// nondecreasing over the lanes; a shifted-in 0 never matches): row_shr 1, 2,
// 4, 8.  The masks (1.0 where the source lane holds the same key) depend only
// on the keys and are built once.  Afterwards the last lane of a run inside a
// row (`flush`) holds the sum of the run's lanes in that row.  All 64 lanes
// must be active.
struct RowScan {
    double m[4];
    bool flush;
    __device__ inline void init(int key, int lane) {
        m[0] = jkr_dpp_int<0x111>(key) == key ? 1.0 : 0.0;
        m[1] = jkr_dpp_int<0x112>(key) == key ? 1.0 : 0.0;
        m[2] = jkr_dpp_int<0x114>(key) == key ? 1.0 : 0.0;
        m[3] = jkr_dpp_int<0x118>(key) == key ? 1.0 : 0.0;
        const int next = __shfl_down(key, 1, 64);
        flush = (lane & 15) == 15 || next != key;
    }
    __device__ inline double scan(double x) const {
        x = fma(jkr_dpp_dbl<0x111>(x), m[0], x);
        x = fma(jkr_dpp_dbl<0x112>(x), m[1], x);
        x = fma(jkr_dpp_dbl<0x114>(x), m[2], x);
        x = fma(jkr_dpp_dbl<0x118>(x), m[3], x);
        return x;
    }
};

// LDS layout (doubles), per spin s at s * jkr_spin_stride(R, KS):
//   Drow[R][KS] = w_r D_s[j0 + r][.], Kr[R][KS], Ki[KS]
// then djr[R] = 2 w_r Dj[i][j0 + r] and red[JKR_WAVES][R].
__host__ __device__ constexpr int jkr_spin_stride(int R, int KS) { return (2 * R + 1) * KS; }

template <int R, int NS, int KS>
__global__ __launch_bounds__(JKR_THREADS) void k_jk_rows(const double* __restrict__ eri, const int2* __restrict__ pairs,
                                                          int K, long long nblk, int seg, const double* __restrict__ Dj,
                                                          const double* __restrict__ Da, const double* __restrict__ Db,
                                                          double* J, double* Ka, double* Kb) {
    extern __shared__ double lds[];
    constexpr int SK = jkr_spin_stride(R, KS);
    double* djr = lds + NS * SK;
    double* red = djr + R;
    int i, j0;
    jkr_block(nblk - 1 - (long long)blockIdx.x, K, R, i, j0);
    const int nrow = min(R, i + 1 - j0);
    const long long IJ0 = (long long)i * (i + 1) / 2 + j0;
    const int last = (int)(IJ0 + nrow - 1);   // the block's columns are 0 .. last
    // this workgroup's segment of them: [cbeg, cend)
    const long long cb = (long long)blockIdx.y * seg;
    if (cb > last) return;   // uniform: the block is shorter than this segment
    const int cbeg = (int)cb;
    const int cend = cbeg + seg < last + 1 ? cbeg + seg : last + 1;
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int s = 0; s < NS; ++s) {
        const double* D = s == 0 ? Da : Db;
        double* Drow = lds + s * SK;
        for (int x = tid; x < R * KS; x += JKR_THREADS) {
            const int r = x / KS, c = x - r * KS;
            const int j = j0 + r;
            Drow[x] = (r < nrow && c < K) ? (i == j ? 0.5 : 1.0) * D[(size_t)j * K + c] : 0.0;
            Drow[R * KS + x] = 0.0;
        }
        for (int c = tid; c < KS; c += JKR_THREADS) Drow[2 * R * KS + c] = 0.0;
    }
    if (tid < R) {
        const int j = j0 + tid;
        djr[tid] = tid < nrow ? 2.0 * (i == j ? 0.5 : 1.0) * Dj[(size_t)i * K + j] : 0.0;
    }
    __syncthreads();
    double jr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) jr[r] = 0.0;
    // row r of the block starts at element rowbase(IJ0 + r) of the unique array
    const double* rows = eri + (size_t)(IJ0 * (IJ0 + 1) / 2);
    // the trip count is uniform over the workgroup: every lane takes part in the scans
    for (int c0 = cbeg; c0 < cend; c0 += JKR_CHUNK) {
        const int KL = c0 + tid;
        const bool valid = KL < cend;
        const int2 p = pairs[valid ? KL : 0];
        const int k = p.x, l = p.y;
        const double fc = k == l ? 0.5 : 1.0;
        // loads are unconditional (an invalid lane reads pair 0) so that
        // none of them sits in a branch of its own waiting for its data
        const double djkl = Dj[(size_t)k * K + l];
        const double dkl = valid ? 2.0 * fc * djkl : 0.0;
        double dil[NS], dik[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double* D = s == 0 ? Da : Db;
            const double xl = D[(size_t)i * K + l], xk = D[(size_t)i * K + k];
            dil[s] = valid ? fc * xl : 0.0;
            dik[s] = valid ? fc * xk : 0.0;
        }
        // the block's integrals of this step: all rows' loads issued together,
        // each from an in-bounds address (a row past the block repeats its last
        // row, a column past a row's end reads the row's diagonal element)
        double vrow[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int rr = r < nrow ? r : nrow - 1;
            const int IJ = (int)IJ0 + rr;
            // rowbase(IJ0 + rr) - rowbase(IJ0) = rr IJ0 + rr (rr + 1) / 2
            const double* row = rows + ((long long)rr * IJ0 + rr * (rr + 1) / 2);
            vrow[r] = row[(valid && KL <= IJ) ? KL : IJ];
        }
        const int key = valid ? k + 1 : K + 1 + lane;   // an invalid lane is a run of its own
        RowScan sc;
        sc.init(key, lane);
        const bool flush = sc.flush && valid;
        double jc = 0.0, kik[NS], kil[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) kik[s] = kil[s] = 0.0;
        // some row's diagonal element (KL == IJ) lies in this step
        const bool diag_here = c0 + JKR_CHUNK > (int)IJ0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r < nrow) {   // uniform
                const int IJ = (int)IJ0 + r;
                double v = (valid && KL <= IJ) ? vrow[r] : 0.0;
                if (diag_here && KL == IJ) v *= 0.5;
                const double wr = (i == j0 + r) ? 0.5 : 1.0;
                jr[r] = fma(v, dkl, jr[r]);
                jc = fma(v, djr[r], jc);
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const double* Dr = lds + s * SK + r * KS;
                    double* Kr = lds + s * SK + (R + r) * KS;
                    kik[s] = fma(v, Dr[l], kik[s]);
                    kil[s] = fma(v, Dr[k], kil[s]);
                    const double pk = sc.scan(v * dil[s]);
                    if (flush) atomicAdd(&Kr[k], wr * pk);
                    atomicAdd(&Kr[l], (wr * dik[s]) * v);
                }
            }
        }
        // column sums of this step (consecutive lanes: consecutive addresses within a run)
        if (valid && jc != 0.0) unsafeAtomicAdd(&J[(size_t)k * K + l], fc * jc);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double* Ki = lds + s * SK + 2 * R * KS;
            const double q = sc.scan(fc * kik[s]);
            if (flush) atomicAdd(&Ki[k], q);
            atomicAdd(&Ki[l], fc * kil[s]);
        }
    }
    // row sums Jacc_ij over the workgroup
#pragma unroll
    for (int r = 0; r < R; ++r) {
        double x = jr[r];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) red[wave * R + r] = x;
    }
    __syncthreads();
    if (tid < nrow) {
        double x = 0.0;
        for (int w = 0; w < JKR_WAVES; ++w) x += red[w * R + tid];
        const int j = j0 + tid;
        if (x != 0.0) unsafeAtomicAdd(&J[(size_t)i * K + j], (i == j ? 0.5 : 1.0) * x);
    }
    // the accumulators hold entries up to the segment's largest k only
    const int kmax = pairs[cend - 1].x + 1;
    for (int s = 0; s < NS; ++s) {
        const double* Kr = lds + s * SK + R * KS;
        const double* Ki = Kr + R * KS;
        double* Kx = s == 0 ? Ka : Kb;
        for (int r = 0; r < nrow; ++r)
            for (int c = tid; c < kmax; c += JKR_THREADS)
                if (Kr[r * KS + c] != 0.0) unsafeAtomicAdd(&Kx[(size_t)(j0 + r) * K + c], Kr[r * KS + c]);
        for (int c = tid; c < kmax; c += JKR_THREADS)
            if (Ki[c] != 0.0) unsafeAtomicAdd(&Kx[(size_t)i * K + c], Ki[c]);
    }
}

namespace {

constexpr size_t jkr_lds_bytes(int R, int NS, int KS) {
    return sizeof(double) * ((size_t)NS * jkr_spin_stride(R, KS) + (size_t)R * (1 + JKR_WAVES));
}

// Columns per workgroup segment: long rows are cut so that the workgroups of
// one launch carry similar work, about Q / (1024 R) columns, between one and
// eight steps (measured: benzene / cc-pVDZ 1536-2048, O2 / cc-pVTZ 256-512,
// n-octane / cc-pVDZ 2048 columns are the fastest of 256 .. 4096).
// HFX_JK_SEG in the environment (steps per segment) overrides it.
int jkr_segment(int K, int R) {
    static const int env = [] {
        const char* e = std::getenv("HFX_JK_SEG");
        return e ? std::atoi(e) : 0;
    }();
    if (env > 0) return env * JKR_CHUNK;
    const double P = (double)K * (K + 1) / 2;
    const double per = P * (P + 1) / 2 / R / 1024.0;
    int steps = (int)(per / JKR_CHUNK + 0.5);
    steps = steps < 1 ? 1 : steps > 8 ? 8 : steps;
    return steps * JKR_CHUNK;
}

template <int R, int NS, int KS>
hipError_t jkr_launch(const double* eri, const int2* pairs, int K, const double* Dj, const double* Da,
                      const double* Db, double* J, double* Ka, double* Kb, hipStream_t st) {
    static_assert(jkr_lds_bytes(R, NS, KS) <= JKR_LDS_MAX, "LDS of the row block");
    const int seg = jkr_segment(K, R);
    constexpr size_t lds = jkr_lds_bytes(R, NS, KS);
    if constexpr (lds > 64 * 1024) {
        static const hipError_t attr =
            hipFuncSetAttribute((const void*)k_jk_rows<R, NS, KS>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (attr != hipSuccess) return attr;
    }
    const long long nblk = jkr_blocks_before(K, R);
    const long long P = (long long)K * (K + 1) / 2;
    const dim3 grid((unsigned)nblk, (unsigned)((P + seg - 1) / seg));
    k_jk_rows<R, NS, KS><<<grid, JKR_THREADS, lds, st>>>(eri, pairs, K, nblk, seg, Dj, Da, Db,
                                                                             J, Ka, Kb);
    return hipGetLastError();
}

// (rows per block, LDS row stride) per K range: the most rows whose LDS fits
// (more rows per block: fewer column atomics per integral).
template <int NS>
hipError_t jkr_dispatch(const double* eri, const int2* pairs, int K, const double* Dj, const double* Da,
                        const double* Db, double* J, double* Ka, double* Kb, hipStream_t st) {
    if (K <= 128) return jkr_launch<NS == 1 ? 16 : 8, NS, 128>(eri, pairs, K, Dj, Da, Db, J, Ka, Kb, st);
    if (K <= 256) return jkr_launch<NS == 1 ? 16 : 8, NS, 256>(eri, pairs, K, Dj, Da, Db, J, Ka, Kb, st);
    if (K <= 512) return jkr_launch<NS == 1 ? 8 : 4, NS, 512>(eri, pairs, K, Dj, Da, Db, J, Ka, Kb, st);
    if (K <= 1024) return jkr_launch<NS == 1 ? 4 : 1, NS, 1024>(eri, pairs, K, Dj, Da, Db, J, Ka, Kb, st);
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_jk_rows(const double* eri, const int2* pairs, int K, int nspin, const double* Dj, const double* Da,
                          const double* Db, double* J, double* Ka, double* Kb, hipStream_t st) {
    if (K <= 0) return hipSuccess;
    return nspin == 2 ? jkr_dispatch<2>(eri, pairs, K, Dj, Da, Db, J, Ka, Kb, st)
                      : jkr_dispatch<1>(eri, pairs, K, Dj, Da, Db, J, Ka, Kb, st);
}

}  // namespace hfx
