// scf_device.hpp -- device-resident SCF iteration (host-visible interface).
//
// The per-iteration state of the reference's DIIS loop (include/scf.hpp:
// 307-384; include/linalg.hpp:132-511) kept in HBM: densities, Fock
// matrices, the last 8 (F, R) pairs of the DIIS history, X = S^{-1/2}, H, S.
// Each iteration runs on one HIP stream:
//   Fock build (J/K kernels) -> F = H + J - K/2 (or H + J - K_s)
//   -> R = F P S - S P F (two GEMMs and an antisymmetrisation)
//   -> E = 1/2 sum P (H + F), B-matrix row <R_new, R_i>  (one small readback)
//   -> F_diis = sum c_i F_i with the host-solved Pulay coefficients
//   -> F' = X^T F_diis X  (two GEMMs; F' read back for the eigen-solve)
//   -> C = X C', P_new = occ C_occ C_occ^T, RMSD(P_new, P)  (C' written back)
// Only F' and C' (K x K each) cross the host link per iteration; the small
// Pulay system (at most 9 x 9) and the symmetric eigen-solve stay on the host.
#pragma once
#include <vector>

#include "../../../include/hfx.h"

namespace hfxd {

struct DevScfTimes {
    double fock = 0.0;   // Fock builds incl. assembly and error vectors
    double diag = 0.0;   // DIIS, transforms, eigen-solves, densities
};

class DevScf {
public:
    // H, S, X: K x K row-major host matrices (copied); nocc per spin; occ = 2 (RHF) or 1 (UHF).
    DevScf(int K, int nspin, const int* nocc, const double* H, const double* S, const double* X);
    ~DevScf();
    DevScf(const DevScf&) = delete;
    DevScf& operator=(const DevScf&) = delete;

    void set_fock(hfx_fock* f) { f_ = f; }
    void set_density(int spin, const double* P);   // host -> device
    void get_density(int spin, double* P);         // device -> host
    void get_coefficients(int spin, double* C);
    const std::vector<double>& orbital_energies(int spin) const { return e_[spin]; }

    // One SCF phase with the reference's rules (same as the host loop):
    // DIIS from iteration start_diis, damping before it when use_damping,
    // convergence on RMSD(P) < eps and ||R|| < eps (and |dE| < e_conv when
    // e_conv > 0) confirmed by a final check.  Returns true when converged;
    // n_it is incremented per iteration and E_ele is the final-check energy.
    bool run(int max_it, int start_diis, bool use_damping, double eps, double damping, double e_conv, bool verbose,
             int& n_it, double& E_ele, DevScfTimes& t);

private:
    void release();
    struct Impl;
    Impl* p_;
    hfx_fock* f_ = nullptr;
    std::vector<std::vector<double>> e_;
};

// Minimum basis size for which the device-resident iteration pays for its
// launches and read-backs (auto selection).
constexpr int DEVSCF_MIN_K = 96;

}  // namespace hfxd
