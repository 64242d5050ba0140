// basis.hpp -- basis construction from per-element CSV files, in the
// reference's conventions (KLYYCYNN/hartree_fock_cpp include/basis.hpp):
//   cart_norm         basis.hpp:28-49   per primitive and Cartesian component
//   angmom_partition  basis.hpp:52-64   (lx, ly, lz) for lx = 0..l, ly = 0..l-lx
//   basis_construction basis.hpp:157-262 atoms in input order, shells in
//                     shell_id order (which must be 0, 1, 2, ...)
// and the element table of include/atoms.hpp:12-16 (H..Kr).
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "oneint.hpp"

namespace hfxh {

struct Vec3 {
    double x, y, z;
};

using Config = std::vector<std::pair<std::string, Vec3>>;

// Per-function description written to basis.csv (include/types.hpp:33-48).
struct BasisInfo {
    std::vector<int> index, shells, lx, ly, lz;
    std::vector<double> cx, cy, cz;
    std::vector<std::string> types, atoms;
};

// Atomic number of an element symbol; throws std::runtime_error for an
// unknown symbol (the reference silently returns 0, include/atoms.hpp:35-39).
int atomic_number(const std::string& sym);
const std::string& element_symbol(int Z);

double cart_norm(double alpha, int lx, int ly, int lz);

// Throws std::runtime_error with the reference's messages for a missing file,
// shells out of order or a wrong primitive count.
FlatBasis basis_construction(const Config& config, const std::string& basis_set, const std::string& param_path,
                             BasisInfo* info = nullptr);

// Rebuilds functions from a saved basis.csv (include/basis.hpp:298-371).
FlatBasis basis_recon(const std::string& basis_csv, const std::string& param_path, const std::string& basis_set);

}  // namespace hfxh
