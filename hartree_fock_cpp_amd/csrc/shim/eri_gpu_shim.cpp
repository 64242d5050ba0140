// eri_gpu_shim.cpp -- the reference-side binding: defines the reference's own
// symbol
//     std::vector<double> ERI_GPU(const std::vector<basis_function>&, double)
// (declared at include/scf.hpp:41-44 of KLYYCYNN/hartree_fock_cpp, defined by
// cuda/src/eri.cu:20-89 there) on top of this engine's C ABI, plus GPU
// versions of the host Fock builds of include/linalg.hpp:100-129 and 215-296.
//
// It is compiled inside the reference tree (it needs the reference's
// include/types.hpp for basis_function); replacing cuda/src/eri.cu by this file
// and linking libhfx.so is the whole integration.  Failures throw
// std::runtime_error, which the reference's app mains already catch.
#include <stdexcept>
#include <string>
#include <vector>

#include "hfx.h"
#include "types.hpp"

namespace {

struct FlatBasis {
    std::vector<double> xyz, alpha, coeff, norm;
    std::vector<int> lmn, poff;
};

FlatBasis flatten(const std::vector<basis_function>& basis) {
    FlatBasis f;
    f.poff.push_back(0);
    for (const auto& b : basis) {
        f.xyz.insert(f.xyz.end(), {b.center.x, b.center.y, b.center.z});
        f.lmn.insert(f.lmn.end(), {b.lx, b.ly, b.lz});
        for (const auto& p : b.prims) {
            f.alpha.push_back(p.alpha);
            f.coeff.push_back(p.coeff);
            f.norm.push_back(p.norm);
        }
        f.poff.push_back((int)f.alpha.size());
    }
    return f;
}

void check(int rc) {
    if (rc != HFX_OK) throw std::runtime_error(std::string("hfx: ") + hfx_last_error());
}

struct BasisHandle {
    hfx_basis* h = nullptr;
    explicit BasisHandle(const std::vector<basis_function>& basis) {
        FlatBasis f = flatten(basis);
        check(hfx_basis_create((int)basis.size(), f.xyz.data(), f.lmn.data(), f.poff.data(), f.alpha.data(),
                               f.coeff.data(), f.norm.data(), &h));
    }
    ~BasisHandle() { hfx_basis_destroy(h); }
};

}  // namespace

std::vector<double> ERI_GPU(const std::vector<basis_function>& basis, double threshold) {
    BasisHandle b(basis);
    const size_t K = basis.size();
    const size_t P = K * (K + 1) / 2;
    std::vector<double> eri(P * (P + 1) / 2);
    check(hfx_eri_unique(b.h, threshold, eri.data(), eri.size()));
    return eri;
}

// G = J - K/2 on the GPU for a density P, with ERIs recomputed (direct) or
// stored in HBM according to the budget -- the per-iteration replacement of
// repulsion_matrix(P, ERI) when the ERIs never leave the device.
std::vector<double> repulsion_matrix_gpu(const std::vector<basis_function>& basis, const std::vector<double>& P,
                                         double threshold) {
    BasisHandle b(basis);
    hfx_fock_opts o{HFX_MODE_AUTO, HFX_PREC_FP64, threshold, 1e-5, 0};
    hfx_fock* f = nullptr;
    check(hfx_fock_create(b.h, &o, &f));
    std::vector<double> J(P.size()), Kx(P.size()), G(P.size());
    const int rc = hfx_fock_build(f, P.data(), nullptr, J.data(), Kx.data(), nullptr);
    hfx_fock_destroy(f);
    check(rc);
    for (size_t i = 0; i < G.size(); ++i) G[i] = J[i] - 0.5 * Kx[i];
    return G;
}
