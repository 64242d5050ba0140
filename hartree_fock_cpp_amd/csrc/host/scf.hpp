// scf.hpp -- RHF / UHF DIIS self-consistent field drivers.
//
// Same algorithm and stopping rules as the reference's rhf_diis / uhf_diis
// (KLYYCYNN/hartree_fock_cpp include/scf.hpp:205-682):
//   * P0 = 0 (or the caller's density), H = T + V, X = S^{-1/2} (symmetric);
//   * each iteration: Fock build, R = F P S - S P F, history of the last 8
//     (F, R), DIIS extrapolation from iteration 4 (augmented Pulay system,
//     include/linalg.hpp:405-511), F' = X^T F X, C = X C', P from the
//     occupied columns, damping P = d P_new + (1 - d) P for iterations < 4;
//   * converged when RMSD(P_new, P) < eps and ||R||_F < eps and a final
//     check with a Fock matrix rebuilt from P_new passes (scf.hpp:107-185);
//   * E = 1/2 sum P (H + F) + E_nuc.
// The Fock build (J and K) runs on the GPU through the C ABI of hfx.h
// (stored or direct, FP64 or mixed).  In the mixed form a converged solution
// is polished by continuing the iteration with FP64 Fock builds until the
// same criteria hold again.
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "basis.hpp"

namespace hfxh {

struct ScfOptions {
    int charge = 0;
    int max_it = 10000;
    double epsilon = 1e-6;
    double damping = 0.3;
    double screening_threshold = 1e-8;
    int eri_mode = 2;              // HFX_MODE_AUTO
    int precision = 0;             // HFX_PREC_FP64
    double mixed_split = 1e-5;
    double energy_convergence = 0.0;   // 0: not used (the reference has no energy criterion)
    size_t hbm_budget = 0;
    bool polish = true;
    bool verbose = true;
    int device_iteration = -1;     // 1 GPU-resident iteration, 0 host, -1 auto (K >= 96)
    int incremental_fock = -1;     // direct builds from density differences: 1 on, 0 off, -1 auto (on)
};

struct ScfTimings {
    double setup = 0.0;   // basis, one-electron integrals, S^{-1/2}
    double eri = 0.0;     // device tables, Schwarz bounds, stored ERIs or quartet plan
    double fock = 0.0;    // sum of Fock builds (incl. the final checks)
    double diag = 0.0;    // sum of DIIS + eigen-solves + density updates
    double total = 0.0;
};

struct RhfResult {
    std::vector<double> density, coefficient, mo_energy;
    double E_tot = 0.0, E_ele = 0.0, E_nuc = 0.0;
    double E_before_polish = 0.0;
    int Ne = 0, K = 0, iterations = 0, polish_iterations = 0;
    bool converged = false;
    int mode = 0;
    size_t n_quartets = 0;
    ScfTimings t;
    std::string start, end;
    BasisInfo info;
};

struct UhfResult {
    std::vector<double> Pa, Pb, Ca, Cb, Ea, Eb;
    double E_tot = 0.0, E_ele = 0.0, E_nuc = 0.0;
    double E_before_polish = 0.0;
    int Na = 0, Nb = 0, K = 0, iterations = 0, polish_iterations = 0;
    bool converged = false;
    int mode = 0;
    size_t n_quartets = 0;
    ScfTimings t;
    std::string start, end;
    BasisInfo info;
};

double nuclear_repulsion(const Config& config);
std::string now_string();

RhfResult rhf_diis(const Config& config, const std::string& basis_set, const std::string& param_path,
                   const ScfOptions& opt, const std::vector<double>* P0 = nullptr);

UhfResult uhf_diis(const Config& config, const std::string& basis_set, const std::string& param_path,
                   std::pair<int, int> spin_occ, const ScfOptions& opt, const std::vector<double>* Pa0 = nullptr,
                   const std::vector<double>* Pb0 = nullptr);

}  // namespace hfxh
