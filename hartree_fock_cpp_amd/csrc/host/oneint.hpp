// oneint.hpp -- one-electron integrals S, T, V on the host (FP64).
//
// Same quantities as the reference's overlap_matrix / kinetic_matrix /
// attraction_matrix (include/integrals.hpp:99-118, 199-218, 516-536):
// contracted values sum_ij c_i c_j N_i N_j (primitive integral) over
// Cartesian functions given in the flat basis layout of include/hfx.h.
// Evaluation is McMurchie-Davidson: 1-D overlaps from E^{ab}_0 sqrt(pi/p),
// kinetic from the 1-D identity T = b(2m+1)S(l,m) - 2b^2 S(l,m+2)
// - m(m-1)/2 S(l,m-2), nuclear attraction from -Z (2 pi/p) sum E_tuv R_tuv
// with an accurate Boys function (Taylor table + downward recursion).
#pragma once
#include <vector>

namespace hfxh {

struct FlatBasis {
    int K = 0;
    std::vector<double> xyz;      // 3K
    std::vector<int> lmn;         // 3K
    std::vector<int> poff;        // K+1
    std::vector<double> alpha, coeff, norm;
};

void overlap_matrix(const FlatBasis& b, double* S);
void kinetic_matrix(const FlatBasis& b, double* T);
void attraction_matrix(const FlatBasis& b, int natom, const int* Z, const double* R, double* V);

// F_m(T) for m = 0..mmax (mmax <= 16), FP64.
void boys_host(int mmax, double T, double* F);

}  // namespace hfxh
