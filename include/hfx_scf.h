/*
 * hfx_scf.h -- C ABI of the host SCF drivers built on hfx.h.
 *
 * Counterparts in KLYYCYNN/hartree_fock_cpp:
 *   hfx_scf_run (RHF)  -> rhf_diis  include/scf.hpp:205-419
 *   hfx_scf_run (UHF)  -> uhf_diis  include/scf.hpp:423-682
 *   hfx_basis_size     -> basis_construction include/basis.hpp:157-262 (size only)
 *   hfx_one_electron   -> overlap_matrix / kinetic_matrix / attraction_matrix
 *                         include/integrals.hpp:99-118, 199-218, 516-536
 * Options carry the reference's JSON keys (include/run.hpp:28-51) plus the
 * engine's own (eri_mode, precision, mixed_split, energy_convergence,
 * hbm_budget, polish).  Errors: non-zero return, message from hfx_last_error()
 * (include/hfx.h) -- the reference throws std::runtime_error instead.
 */
#ifndef HFX_SCF_H
#define HFX_SCF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int charge;
    int n_alpha;               /* UHF when n_alpha >= 0 (spin_occupancy.alpha), RHF when < 0 */
    int n_beta;
    int max_iterations;        /* reference default 10000 */
    double epsilon;            /* 1e-6: RMSD(P) and ||FPS - SPF|| */
    double damping;            /* 0.3 for iterations < 4 */
    double screening_threshold;/* 1e-8 */
    int eri_mode;              /* HFX_MODE_* */
    int precision;             /* HFX_PREC_* */
    double mixed_split;        /* 1e-5 */
    double energy_convergence; /* 0 = off */
    size_t hbm_budget;         /* 0 = unlimited */
    int polish;                /* FP64 polish after a mixed-precision solve */
    int verbose;
    int device_iteration;      /* 1: SCF iteration resident on the GPU (Fock assembly, DIIS, transforms,
                                  densities; host eigen-solve), 0: on the host, -1: auto (K >= 96) */
    int incremental_fock;      /* direct builds from density differences (hfx_fock_opts.incremental):
                                  1 on, 0 off, -1 auto (on) */
} hfx_scf_opts;

typedef struct {
    double E_tot, E_ele, E_nuc;
    double E_before_polish;    /* mixed form: energy of the mixed-precision solution */
    int iterations, polish_iterations, converged;
    int K, n_electrons, mode;
    long long n_quartets;
    double t_setup, t_eri, t_fock, t_diag, t_total;   /* seconds */
} hfx_scf_result;

void hfx_scf_default_opts(hfx_scf_opts* o);

/* Number of basis functions for a molecule (symbols: natom NUL-separated). */
int hfx_basis_size(int natom, const char* symbols, const char* basis_set, const char* basis_path, int* K);

/* Flat basis arrays (sizes from hfx_basis_size and *nprim): xyz 3K, lmn 3K, poff K+1, alpha/coeff/norm nprim. */
int hfx_basis_build(int natom, const char* symbols, const double* xyz_atoms, const char* basis_set,
                    const char* basis_path, int K, double* xyz, int* lmn, int* poff, int* nprim, double* alpha,
                    double* coeff, double* norm, int max_prim);

/* S, T, V (K x K each, any may be NULL). */
int hfx_one_electron(int natom, const char* symbols, const double* xyz_atoms, const char* basis_set,
                     const char* basis_path, double* S, double* T, double* V);

/* The host eigensolver used by the SCF (Householder + implicit QL; no
 * reference counterpart beyond Eigen's SelfAdjointEigenSolver at
 * include/linalg.hpp:138,162): w ascending, V row-major, column k = vector k. */
int hfx_host_sym_eig(int n, const double* A, double* w, double* V);

/* All eigenvalues (ascending) and the eigenvectors of the m lowest (columns
 * 0..m-1 of V, the rest zero): what an SCF iteration needs (the occupied
 * orbitals).  No reference counterpart; inverse iteration on the tridiagonal
 * form, Gram-Schmidt inside eigenvalue clusters. */
int hfx_host_sym_eig_lowest(int n, const double* A, int m, double* w, double* V);

/* Runs one SCF.  symbols: natom NUL-separated element symbols; xyz_atoms:
 * 3*natom bohr.  P0a/P0b: optional initial densities (K x K; P0b for UHF).
 * Outputs (each optional): Pa/Pb densities, Ca/Cb coefficients (K x K,
 * column k = orbital k), ea/eb orbital energies (K). */
int hfx_scf_run(int natom, const char* symbols, const double* xyz_atoms, const char* basis_set,
                const char* basis_path, const hfx_scf_opts* opts, const double* P0a, const double* P0b,
                hfx_scf_result* res, double* Pa, double* Pb, double* Ca, double* Cb, double* ea, double* eb);

#ifdef __cplusplus
}
#endif

#endif /* HFX_SCF_H */
