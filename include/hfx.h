/*
 * hfx.h -- C ABI of the MI355X (gfx950) ERI + Fock engine.
 *
 * Drop-in boundary for the GPU path of KLYYCYNN/hartree_fock_cpp.  Plain C
 * types only: host pointers unless a name ends in _device, sizes as size_t,
 * status codes as int (0 = success, negative = error class; the message is
 * available from hfx_last_error()).  All functions are synchronous unless a
 * stream argument is taken.  Thread compatibility: distinct hfx_basis /
 * hfx_fock objects may be used from distinct threads concurrently.  Planning
 * a large basis (above 2^18 shell-pair pairs: the first unique build or
 * hfx_fock_create at a new threshold) runs on up to 16 short-lived host
 * threads of the library's own, joined before the call returns; so does the
 * host copy of hfx_eri_unique above 32 MB (up to 8 threads).
 *
 * Conventions reproduced from the reference (all cited file:line are in
 * KLYYCYNN/hartree_fock_cpp):
 *   - basis functions are Cartesian, in basis_construction order
 *     (include/basis.hpp:157-262): whole shells, components in
 *     angmom_partition order (basis.hpp:52-64), L <= 3 (cuda/include/recursion.cuh:5);
 *   - primitives carry (alpha, coeff, norm) with norm = cart_norm per
 *     component (basis.hpp:28-49); the library never renormalises;
 *   - the unique ERI array has Q = P(P+1)/2 entries, P = K(K+1)/2, value of
 *     (ij|kl) at eri_index(i,j,k,l) (include/indexing.hpp:33-52); entries whose
 *     Schwarz product b_IJ*b_KL < threshold are exactly 0.0
 *     (include/integrals.hpp:771-776, cuda/include/kernels.cuh:104-110);
 *   - matrices are K x K row-major doubles (include/indexing.hpp:4-7).
 */
#ifndef HFX_H
#define HFX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HFX_OK 0
#define HFX_ERR_INVALID (-1)      /* bad argument / unsupported basis layout */
#define HFX_ERR_DEVICE (-2)       /* HIP runtime error */
#define HFX_ERR_NOMEM (-3)        /* device allocation failed or over budget */
#define HFX_ERR_UNSUPPORTED (-4)  /* angular momentum or size beyond limits */

#define HFX_MODE_STORED 0   /* unique ERIs resident in HBM, J/K streamed from them */
#define HFX_MODE_DIRECT 1   /* ERIs recomputed and contracted every build */
#define HFX_MODE_AUTO 2     /* stored if it fits the HBM budget, else direct */

#define HFX_PREC_FP64 0     /* oracle-matched drop-in form */
#define HFX_PREC_MIXED 1    /* FP32 for quartets with Schwarz bound < split, FP64 otherwise */

typedef struct hfx_basis hfx_basis;
typedef struct hfx_fock hfx_fock;

/* Last error message of the calling thread ("" when none). */
const char* hfx_last_error(void);

/* Number of visible HIP devices (0 when none). */
int hfx_device_count(void);

/* Replaces convert_basis_to_gpu (include/gpu_types.hpp:21-52) and the
 * per-call H2D copy of ERI_GPU (cuda/src/eri.cu:40-62): builds the device
 * resident shell, shell-pair and primitive-pair tables once.
 * Function mu has centre xyz[3mu..3mu+2], exponents lmn[3mu..3mu+2] and
 * primitives poff[mu] .. poff[mu+1]-1 of alpha/coeff/norm (the fields of
 * primitive_gaussian, include/types.hpp:16-20). */
int hfx_basis_create(int K, const double* xyz, const int* lmn, const int* poff, const double* alpha,
                     const double* coeff, const double* norm, hfx_basis** out);
void hfx_basis_destroy(hfx_basis* b);

/* Shell / shell-pair / kept primitive-pair counts of a basis. */
int hfx_basis_info(const hfx_basis* b, int* nshell, int* npair, int* nprimpair);

/* Replaces schwarz_bounds (include/integrals.hpp:732-752), which the
 * reference runs on the host inside ERI_GPU (cuda/src/eri.cu:32):
 * bounds[pair_index(i,j)] = sqrt(max((ij|ij), 0)), K(K+1)/2 entries. */
int hfx_schwarz_bounds(hfx_basis* b, double* bounds);

/* Replaces ERI_GPU (cuda/src/eri.cu:20-89, declared include/scf.hpp:41-44):
 * the full unique ERI array (Q entries, host memory) in the eri_index layout,
 * FP64, function-level Schwarz mask. */
int hfx_eri_unique(hfx_basis* b, double threshold, double* out, size_t Q);

/* Same, into caller-owned device memory (Q doubles) on `stream`
 * (a hipStream_t, NULL = default stream); no host round trip. */
int hfx_eri_unique_device(hfx_basis* b, double threshold, double* d_out, size_t Q, void* stream);

/* Replaces repulsion_matrix (include/linalg.hpp:100-129) for a
 * device-resident unique array: G = J - 1/2 K with
 * G_ij = sum_kl P_kl [(ij|kl) - 1/2 (ik|jl)]. P, G: host K x K. */
int hfx_repulsion_matrix(hfx_basis* b, const double* d_eri, const double* P, double* G);

/* One-electron matrices on the GPU (K x K row-major host outputs; any may be
 * NULL): S, T and V = -sum_C Z_C <a|1/|r-C||b> for the flat basis of
 * hfx_basis_create and natom point charges Z at Rn (3*natom bohr).  Replaces
 * the reference's host overlap_matrix / kinetic_matrix / attraction_matrix
 * (include/integrals.hpp:99-118, 199-218, 516-536).  L <= 3. */
int hfx_one_electron_device(int K, const double* xyz, const int* lmn, const int* poff, const double* alpha,
                            const double* coeff, const double* norm, int natom, const double* Z, const double* Rn,
                            double* S, double* T, double* V);

/* Replaces coulomb_matrix (include/linalg.hpp:215-257) and exchange_matrix
 * (include/linalg.hpp:260-296) for a device-resident unique array, in one
 * pass: J from Pa + Pb, Ka from Pa, Kb from Pb (host K x K). */
int hfx_coulomb_exchange(hfx_basis* b, const double* d_eri, const double* Pa, const double* Pb, double* J,
                         double* Ka, double* Kb);

/* No counterpart in the reference: an allocation-free variant of
 * hfx_coulomb_exchange for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Densities in, J / Ka / Kb out, all caller-owned device K x K
 * matrices; enqueued on `stream` (a hipStream_t, NULL = the basis' stream),
 * asynchronous.  RHF: d_Pb == NULL (J and Ka from Pa).  UHF: J from Pa + Pb,
 * which is formed in a fixed device scratch array of HFX_NOALLOC_MAX_K^2
 * doubles (2 MiB per GPU, held until the process ends); a UHF call with
 * K > HFX_NOALLOC_MAX_K returns HFX_ERR_UNSUPPORTED.  The array is shared by
 * every call on one GPU: UHF calls that overlap on one GPU (different streams
 * or threads) race on it.  No hipMalloc / hipFree or allocating library call
 * is made. */
#define HFX_NOALLOC_MAX_K 512
int hfx_coulomb_exchange_noalloc(hfx_basis* b, const double* d_eri, const double* d_Pa, const double* d_Pb,
                                 double* d_J, double* d_Ka, double* d_Kb, void* stream);

typedef struct {
    int mode;              /* HFX_MODE_* */
    int precision;         /* HFX_PREC_* */
    double threshold;      /* function-level Schwarz threshold (reference screening_threshold) */
    double mixed_split;    /* quartet bound below which the mixed form uses FP32 (default 1e-5) */
    size_t hbm_budget;     /* bytes this job may hold on the device; 0 = unlimited */
    int incremental;       /* direct mode: 1 = build G(P) as G(P_prev) + G(P - P_prev) with the second
                              term screened by |P - P_prev| (full build every 8th call); 0 = off.
                              No reference counterpart (the reference rebuilds in full every iteration). */
    double incremental_tau;/* skip a quartet of the difference term when Q_ab Q_cd max|dP| < tau;
                              0 = 1e-3 x threshold */
} hfx_fock_opts;

/* Fock-build context for one SCF job: owns its stored ERIs (stored mode) or
 * quartet lists (direct mode), device densities and accumulators.  Mixed
 * precision: direct builds evaluate and accumulate the small-bound quartets in
 * FP32 every build; stored mode evaluates them in FP32 once and keeps them in
 * the FP64 unique array.  AUTO picks stored when the unique array fits
 * hbm_budget (0: always), for both precisions.  Destroying the context waits
 * for its last build. */
int hfx_fock_create(hfx_basis* b, const hfx_fock_opts* opts, hfx_fock** out);
void hfx_fock_destroy(hfx_fock* f);

/* Selected mode (HFX_MODE_STORED or HFX_MODE_DIRECT) and device bytes held. */
int hfx_fock_info(const hfx_fock* f, int* mode, size_t* device_bytes, size_t* n_quartets);

/* One Fock build.  RHF: Pb == NULL, J from Pa and Ka from Pa (the caller forms
 * G = J - 1/2 Ka, as repulsion_matrix does).  UHF: J from Pa + Pb, Ka from Pa,
 * Kb from Pb (coulomb_matrix / exchange_matrix).  All host K x K; J, Ka, Kb
 * may be NULL when not wanted. */
int hfx_fock_build(hfx_fock* f, const double* Pa, const double* Pb, double* J, double* Ka, double* Kb);

/* Same on caller-owned device matrices (K x K), on `stream`. */
int hfx_fock_build_device(hfx_fock* f, const double* d_Pa, const double* d_Pb, double* d_J, double* d_Ka,
                          double* d_Kb, void* stream);

/* Per-class launch statistics of the last build: for each of the 55 classes
 * (ss|ss)..(ff|ff) (index qc = pb(pb+1)/2 + pk over pair classes
 * ss,ps,pp,ds,dp,dd,fs,fp,fd,ff), the number of FP64 and FP32 shell quartets. */
int hfx_fock_class_counts(const hfx_fock* f, long long* n_fp64, long long* n_fp32);

/* Work-unit autotuner of a direct context (no reference counterpart): after
 * an untimed first build, the next three builds run in full with the slice
 * target T, T/4 and 4T (T = 2048 wavefronts, or HFX_SLICE_TARGET), timed with
 * HIP events; the fastest plan serves the later builds.  Off by default
 * (hfx_set_autotune(1) or HFX_AUTOTUNE=1 turns it on for contexts created
 * afterwards): for an SCF of ~30 builds the tuning builds cost more than they
 * save; it pays for long sequences of builds on one context.  Results do not
 * depend on the plan beyond summation order.  done: 1 once chosen; target:
 * the current plan's; targets / ms: the candidates and their measured build
 * times (0 while not measured). */
int hfx_set_autotune(int on);
int hfx_fock_autotune_info(const hfx_fock* f, int* done, long long* target, long long* targets, float* ms);

/* Measurement hooks (no reference counterpart).  With class timing on, every
 * ERI class launch is bracketed by HIP events on its stream; hfx_class_times
 * returns the last build's per-launch milliseconds, ms[2*qc + p] with p = 0
 * for FP64 and 1 for FP32 (0 for classes not launched).  on = 1: the lists
 * of one build still run on up to four streams at once, so a class's time
 * includes the device it shares with the classes beside it; on = 2: every
 * list runs on the caller's stream, one class at a time. */
int hfx_set_class_timing(int on);

/* Graph replay of the Fock build (no reference counterpart): each hfx_fock
 * captures its build's kernel launches into a HIP graph for the last
 * (densities, outputs, stream) it was called with and replays it while they
 * stay the same (direct builds; stored builds are a few launches).  On by
 * default (HFX_FOCK_GRAPH=0 in the environment turns it off); builds with
 * class timing on, and builds from the server's worker threads, run eagerly. */
int hfx_set_fock_graphs(int on);
int hfx_class_times(float* ms);

/* Screened shell-quartet counts and primitive-quartet work (sum over quartets
 * of kept bra x ket primitive pairs) per class, nq[2*qc + p], nprimq[2*qc + p]. */
int hfx_plan_stats(hfx_basis* b, double threshold, int precision, double mixed_split, long long* nq,
                   long long* nprimq);

#ifdef __cplusplus
}
#endif

#endif /* HFX_H */
