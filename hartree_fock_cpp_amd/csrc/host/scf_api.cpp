// scf_api.cpp -- extern "C" entry points of include/hfx_scf.h.
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../../include/hfx.h"
#include "../../../include/hfx_scf.h"
#include "linalg.hpp"
#include "scf.hpp"

namespace hfx_detail {
// hfx_api.hip owns the thread-local error string; the SCF layer reports
// through the same channel.
void set_error(const std::string& msg);
}

namespace {

hfxh::Config make_config(int natom, const char* symbols, const double* xyz) {
    hfxh::Config c;
    const char* s = symbols;
    for (int a = 0; a < natom; ++a) {
        c.push_back({std::string(s), {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]}});
        s += std::strlen(s) + 1;
    }
    return c;
}

hfxh::ScfOptions to_opts(const hfx_scf_opts* o) {
    hfxh::ScfOptions r;
    r.charge = o->charge;
    r.max_it = o->max_iterations;
    r.epsilon = o->epsilon;
    r.damping = o->damping;
    r.screening_threshold = o->screening_threshold;
    r.eri_mode = o->eri_mode;
    r.precision = o->precision;
    r.mixed_split = o->mixed_split;
    r.energy_convergence = o->energy_convergence;
    r.hbm_budget = o->hbm_budget;
    r.polish = o->polish != 0;
    r.verbose = o->verbose != 0;
    r.device_iteration = o->device_iteration;
    r.incremental_fock = o->incremental_fock;
    return r;
}

void copy(double* dst, const std::vector<double>& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(double));
}

template <class R>
void fill_result(hfx_scf_result* res, const R& r, int Ne) {
    res->E_tot = r.E_tot;
    res->E_ele = r.E_ele;
    res->E_nuc = r.E_nuc;
    res->E_before_polish = r.E_before_polish;
    res->iterations = r.iterations;
    res->polish_iterations = r.polish_iterations;
    res->converged = r.converged;
    res->K = r.K;
    res->n_electrons = Ne;
    res->mode = r.mode;
    res->n_quartets = (long long)r.n_quartets;
    res->t_setup = r.t.setup;
    res->t_eri = r.t.eri;
    res->t_fock = r.t.fock;
    res->t_diag = r.t.diag;
    res->t_total = r.t.total;
}

}  // namespace

extern "C" {

void hfx_scf_default_opts(hfx_scf_opts* o) {
    std::memset(o, 0, sizeof(*o));
    o->n_alpha = -1;
    o->n_beta = -1;
    o->max_iterations = 10000;
    o->epsilon = 1e-6;
    o->damping = 0.3;
    o->screening_threshold = 1e-8;
    o->eri_mode = HFX_MODE_AUTO;
    o->precision = HFX_PREC_FP64;
    o->mixed_split = 1e-5;
    o->polish = 1;
    o->verbose = 0;
    o->device_iteration = -1;
    o->incremental_fock = -1;
}

int hfx_basis_size(int natom, const char* symbols, const char* basis_set, const char* basis_path, int* K) {
    try {
        std::vector<double> xyz(3 * natom, 0.0);
        auto fb = hfxh::basis_construction(make_config(natom, symbols, xyz.data()), basis_set, basis_path);
        *K = fb.K;
        return 0;
    } catch (const std::exception& e) {
        hfx_detail::set_error(e.what());
        return HFX_ERR_INVALID;
    }
}

int hfx_basis_build(int natom, const char* symbols, const double* xyz_atoms, const char* basis_set,
                    const char* basis_path, int K, double* xyz, int* lmn, int* poff, int* nprim, double* alpha,
                    double* coeff, double* norm, int max_prim) {
    try {
        auto fb = hfxh::basis_construction(make_config(natom, symbols, xyz_atoms), basis_set, basis_path);
        if (fb.K != K) throw std::runtime_error("K does not match the basis");
        *nprim = (int)fb.alpha.size();
        if ((int)fb.alpha.size() > max_prim) throw std::runtime_error("primitive buffer too small");
        std::memcpy(xyz, fb.xyz.data(), fb.xyz.size() * sizeof(double));
        std::memcpy(lmn, fb.lmn.data(), fb.lmn.size() * sizeof(int));
        std::memcpy(poff, fb.poff.data(), fb.poff.size() * sizeof(int));
        std::memcpy(alpha, fb.alpha.data(), fb.alpha.size() * sizeof(double));
        std::memcpy(coeff, fb.coeff.data(), fb.coeff.size() * sizeof(double));
        std::memcpy(norm, fb.norm.data(), fb.norm.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        hfx_detail::set_error(e.what());
        return HFX_ERR_INVALID;
    }
}

int hfx_one_electron(int natom, const char* symbols, const double* xyz_atoms, const char* basis_set,
                     const char* basis_path, double* S, double* T, double* V) {
    try {
        auto cfg = make_config(natom, symbols, xyz_atoms);
        auto fb = hfxh::basis_construction(cfg, basis_set, basis_path);
        if (S) hfxh::overlap_matrix(fb, S);
        if (T) hfxh::kinetic_matrix(fb, T);
        if (V) {
            std::vector<int> Z;
            for (auto& a : cfg) Z.push_back(hfxh::atomic_number(a.first));
            hfxh::attraction_matrix(fb, natom, Z.data(), xyz_atoms, V);
        }
        return 0;
    } catch (const std::exception& e) {
        hfx_detail::set_error(e.what());
        return HFX_ERR_INVALID;
    }
}

int hfx_host_sym_eig(int n, const double* A, double* w, double* V) {
    if (n <= 0 || !A || !w || !V) {
        hfx_detail::set_error("invalid argument");
        return HFX_ERR_INVALID;
    }
    hfxh::sym_eig(n, A, w, V);
    return 0;
}

int hfx_host_sym_eig_lowest(int n, const double* A, int m, double* w, double* V) {
    if (n <= 0 || m < 0 || m > n || !A || !w || !V) {
        hfx_detail::set_error("invalid argument");
        return HFX_ERR_INVALID;
    }
    hfxh::sym_eig_lowest(n, A, m, w, V);
    return 0;
}

int hfx_scf_run(int natom, const char* symbols, const double* xyz_atoms, const char* basis_set,
                const char* basis_path, const hfx_scf_opts* opts, const double* P0a, const double* P0b,
                hfx_scf_result* res, double* Pa, double* Pb, double* Ca, double* Cb, double* ea, double* eb) {
    if (!symbols || !xyz_atoms || !basis_set || !basis_path || !opts || !res) {
        hfx_detail::set_error("null argument");
        return HFX_ERR_INVALID;
    }
    try {
        auto cfg = make_config(natom, symbols, xyz_atoms);
        auto o = to_opts(opts);
        int K = 0;
        if (opts->n_alpha < 0) {
            std::vector<double> p0;
            const std::vector<double>* pp = nullptr;
            if (P0a) {
                auto fb = hfxh::basis_construction(cfg, basis_set, basis_path);
                p0.assign(P0a, P0a + (size_t)fb.K * fb.K);
                pp = &p0;
            }
            auto r = hfxh::rhf_diis(cfg, basis_set, basis_path, o, pp);
            fill_result(res, r, r.Ne);
            copy(Pa, r.density);
            copy(Ca, r.coefficient);
            copy(ea, r.mo_energy);
            (void)K;
        } else {
            std::vector<double> a0, b0;
            const std::vector<double>*pa = nullptr, *pb = nullptr;
            if (P0a && P0b) {
                auto fb = hfxh::basis_construction(cfg, basis_set, basis_path);
                a0.assign(P0a, P0a + (size_t)fb.K * fb.K);
                b0.assign(P0b, P0b + (size_t)fb.K * fb.K);
                pa = &a0;
                pb = &b0;
            }
            auto r = hfxh::uhf_diis(cfg, basis_set, basis_path, {opts->n_alpha, opts->n_beta}, o, pa, pb);
            fill_result(res, r, r.Na + r.Nb);
            copy(Pa, r.Pa);
            copy(Pb, r.Pb);
            copy(Ca, r.Ca);
            copy(Cb, r.Cb);
            copy(ea, r.Ea);
            copy(eb, r.Eb);
        }
        return 0;
    } catch (const std::exception& e) {
        hfx_detail::set_error(e.what());
        return HFX_ERR_INVALID;
    }
}

}  // extern "C"
