// ref_harness.cpp -- extern "C" wrappers around the reference's own CPU integral
// code, compiled in place from the reference checkout by oracle/Makefile into
// oracle/_ref/libref.so.  Test infrastructure only: it produces golden vectors
// (tests/golden/make_golden.py) and the "reference" CPU baseline timed by
// bench.py.  No reference source is copied; this file only includes the
// reference headers (include/integrals.hpp and its closure: basis.hpp,
// atoms.hpp, types.hpp, indexing.hpp, external/rapidcsv/rapidcsv.h).
#include <cassert>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "integrals.hpp"

namespace {

std::vector<basis_function> g_basis;   // last basis built by ref_build_basis
basis_info g_info;

std::vector<basis_function> from_flat(int K, const double* xyz, const int* lmn, const int* poff,
                                      const double* alpha, const double* coeff, const double* norm) {
    std::vector<basis_function> b(K);
    for (int mu = 0; mu < K; ++mu) {
        b[mu].center = {xyz[3 * mu], xyz[3 * mu + 1], xyz[3 * mu + 2]};
        b[mu].lx = lmn[3 * mu];
        b[mu].ly = lmn[3 * mu + 1];
        b[mu].lz = lmn[3 * mu + 2];
        for (int i = poff[mu]; i < poff[mu + 1]; ++i) b[mu].prims.push_back({alpha[i], coeff[i], norm[i]});
    }
    return b;
}

}  // namespace

#define REF_BASIS_ARGS int K, const double *xyz, const int *lmn, const int *poff, \
                       const double *alpha, const double *coeff, const double *norm

extern "C" {

// basis.hpp:157-262.  symbols: natom NUL-separated element symbols.
// Returns K (and the number of primitives in *nprim_out), or -1 on error.
int ref_build_basis(int natom, const char* symbols, const double* coords, const char* basis_set,
                    const char* param_path, int* nprim_out) {
    try {
        std::vector<std::pair<std::string, vector3>> config;
        const char* s = symbols;
        for (int a = 0; a < natom; ++a) {
            config.push_back({std::string(s), {coords[3 * a], coords[3 * a + 1], coords[3 * a + 2]}});
            s += std::strlen(s) + 1;
        }
        auto res = basis_construction(config, basis_set, param_path);
        g_basis = res.first;
        g_info = res.second;
        int np = 0;
        for (auto& f : g_basis) np += (int)f.prims.size();
        *nprim_out = np;
        return (int)g_basis.size();
    } catch (...) {
        return -1;
    }
}

// Copy the last built basis into flat arrays (sizes from ref_build_basis);
// shell_id receives basis_info.shells.
void ref_basis_fill(double* xyz, int* lmn, int* poff, double* alpha, double* coeff, double* norm,
                    int* shell_id) {
    int p = 0;
    for (size_t mu = 0; mu < g_basis.size(); ++mu) {
        const auto& f = g_basis[mu];
        xyz[3 * mu] = f.center.x;
        xyz[3 * mu + 1] = f.center.y;
        xyz[3 * mu + 2] = f.center.z;
        lmn[3 * mu] = f.lx;
        lmn[3 * mu + 1] = f.ly;
        lmn[3 * mu + 2] = f.lz;
        poff[mu] = p;
        for (auto& g : f.prims) {
            alpha[p] = g.alpha;
            coeff[p] = g.coeff;
            norm[p] = g.norm;
            ++p;
        }
        shell_id[mu] = g_info.shells[mu];
    }
    poff[g_basis.size()] = p;
}

double ref_boys(int m, double T) { return boys(m, T); }                       // integrals.hpp:223
double ref_cart_norm(double a, int lx, int ly, int lz) { return cart_norm(a, lx, ly, lz); }

int ref_hermite_E(int i, int j, double a, double b, double A, double B, double* out) {
    auto v = hermite_E_array(i, j, a, b, A, B);                                // integrals.hpp:269
    for (size_t t = 0; t < v.size(); ++t) out[t] = v[t];
    return (int)v.size();
}

double ref_coulomb_R(int t, int u, int v, int n, double p, double q, const double* P, const double* Q) {
    return coulomb_R(t, u, v, n, p, q, {P[0], P[1], P[2]}, {Q[0], Q[1], Q[2]});  // integrals.hpp:541
}

void ref_overlap_matrix(REF_BASIS_ARGS, double* out) {
    auto b = from_flat(K, xyz, lmn, poff, alpha, coeff, norm);
    auto S = overlap_matrix(b);
    std::memcpy(out, S.data(), sizeof(double) * S.size());
}

void ref_kinetic_matrix(REF_BASIS_ARGS, double* out) {
    auto b = from_flat(K, xyz, lmn, poff, alpha, coeff, norm);
    auto T = kinetic_matrix(b);
    std::memcpy(out, T.data(), sizeof(double) * T.size());
}

void ref_attraction_matrix(REF_BASIS_ARGS, int natom, const int* Z, const double* R, double* out) {
    auto b = from_flat(K, xyz, lmn, poff, alpha, coeff, norm);
    std::vector<atom> mol(natom);
    for (int a = 0; a < natom; ++a) mol[a] = {Z[a], {R[3 * a], R[3 * a + 1], R[3 * a + 2]}};
    auto V = attraction_matrix(b, mol);
    std::memcpy(out, V.data(), sizeof(double) * V.size());
}

void ref_schwarz_bounds(REF_BASIS_ARGS, double* out) {
    auto b = from_flat(K, xyz, lmn, poff, alpha, coeff, norm);
    auto v = schwarz_bounds(b);                                                // integrals.hpp:732
    std::memcpy(out, v.data(), sizeof(double) * v.size());
}

// integrals.hpp:754-787 evaluated at selected unique quartet ids (NULL = all).
void ref_eri_quartets(REF_BASIS_ARGS, const double* bounds, double threshold, const uint64_t* qids,
                      size_t n, double* out) {
    auto b = from_flat(K, xyz, lmn, poff, alpha, coeff, norm);
    for (size_t s = 0; s < n; ++s) {
        size_t q = qids ? (size_t)qids[s] : s;
        auto pr = quartet_to_pairs(q);
        auto p1 = pair_to_indices(pr.first);
        auto p2 = pair_to_indices(pr.second);
        if (bounds[pr.first] * bounds[pr.second] < threshold) {
            out[s] = 0.0;
            continue;
        }
        out[s] = basis_repulsion(b[p1.first], b[p1.second], b[p2.first], b[p2.second]);
    }
}

double ref_basis_repulsion(REF_BASIS_ARGS, int f1, int f2, int f3, int f4) {
    auto b = from_flat(K, xyz, lmn, poff, alpha, coeff, norm);
    return basis_repulsion(b[f1], b[f2], b[f3], b[f4]);
}

}  // extern "C"
