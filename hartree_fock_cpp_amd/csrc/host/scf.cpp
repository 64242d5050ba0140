// scf.cpp -- see scf.hpp.
#include "scf.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <deque>
#include <stdexcept>

#include "../../../include/hfx.h"
#include "../device/scf_device.hpp"
#include "linalg.hpp"

namespace hfxh {

namespace {

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

void check(int rc) {
    if (rc != 0) throw std::runtime_error(std::string("GPU engine: ") + hfx_last_error());
}

// Owns the device tables of one job and its Fock-build context.
struct Engine {
    hfx_basis* b = nullptr;
    hfx_fock* f = nullptr;
    ~Engine() {
        if (f) hfx_fock_destroy(f);
        if (b) hfx_basis_destroy(b);
    }
    void make_fock(const ScfOptions& o, int precision) {
        if (f) hfx_fock_destroy(f);
        f = nullptr;
        hfx_fock_opts fo;
        fo.mode = o.eri_mode;
        fo.precision = precision;
        fo.threshold = o.screening_threshold;
        fo.mixed_split = o.mixed_split;
        fo.hbm_budget = o.hbm_budget;
        fo.incremental = o.incremental_fock != 0;
        // skipped difference terms three orders below both the integral screen
        // and the convergence threshold, so they cannot stall convergence
        fo.incremental_tau = 1e-3 * std::min(o.screening_threshold, o.epsilon);
        check(hfx_fock_create(b, &fo, &f));
    }
};

// ||A||_F over a vector
double vnorm(const std::vector<double>& v) {
    double s = 0.0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}

double rmsd(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += (a[i] - b[i]) * (a[i] - b[i]);
    return std::sqrt(s / (double)a.size());
}

// Generic closed/open-shell SCF state: nspin = 1 (RHF, P = 2 C_occ C_occ^T,
// F = H + J - K/2) or 2 (UHF, P_s = C_occ C_occ^T, F_s = H + J - K_s).
struct ScfCore {
    int K = 0, nspin = 1;
    std::vector<int> nocc;
    std::vector<double> H, S, X;
    Engine* eng = nullptr;
    std::vector<std::vector<double>> P, F, C, e;
    ScfTimings* t = nullptr;

    void fock(const std::vector<std::vector<double>>& Pin, std::vector<std::vector<double>>& Fout) {
        auto t0 = clk::now();
        const size_t K2 = (size_t)K * K;
        std::vector<double> J(K2), Ka(K2), Kb(nspin == 2 ? K2 : 0);
        if (nspin == 1) {
            check(hfx_fock_build(eng->f, Pin[0].data(), nullptr, J.data(), Ka.data(), nullptr));
            Fout.assign(1, std::vector<double>(K2));
            for (size_t i = 0; i < K2; ++i) Fout[0][i] = H[i] + (J[i] - 0.5 * Ka[i]);
        } else {
            check(hfx_fock_build(eng->f, Pin[0].data(), Pin[1].data(), J.data(), Ka.data(), Kb.data()));
            Fout.assign(2, std::vector<double>(K2));
            for (size_t i = 0; i < K2; ++i) {
                Fout[0][i] = H[i] + J[i] - Ka[i];
                Fout[1][i] = H[i] + J[i] - Kb[i];
            }
        }
        t->fock += secs(t0, clk::now());
    }

    // R = F P S - S P F  (= M - M^T with M = F P S for symmetric F, P, S)
    std::vector<double> error(const std::vector<double>& Fm, const std::vector<double>& Pm) {
        const size_t K2 = (size_t)K * K;
        std::vector<double> FP(K2), M(K2), R(K2);
        gemm(K, Fm.data(), false, Pm.data(), false, FP.data());
        gemm(K, FP.data(), false, S.data(), false, M.data());
        for (int i = 0; i < K; ++i)
            for (int j = 0; j < K; ++j) R[(size_t)i * K + j] = M[(size_t)i * K + j] - M[(size_t)j * K + i];
        return R;
    }

    double energy(const std::vector<std::vector<double>>& Pm, const std::vector<std::vector<double>>& Fm) {
        double E = 0.0;
        for (int s = 0; s < nspin; ++s) {
            double es = 0.0;
            for (size_t i = 0; i < H.size(); ++i) es += Pm[s][i] * (H[i] + Fm[s][i]);
            E += 0.5 * es;
        }
        return E;
    }

    // Eigen-solve of X^T F X; returns C = X C' and orbital energies; density.
    // all_vectors false: only the n_occ occupied columns of C are formed (all
    // orbital energies still are); the SCF iterations need no more.
    void diag(const std::vector<double>& Fm, std::vector<double>& Cout, std::vector<double>& eout,
              std::vector<double>& Pout, int n_occ, double occ, bool all_vectors = true) {
        const size_t K2 = (size_t)K * K;
        std::vector<double> T1(K2), Fp(K2), Cp(K2);
        gemm(K, X.data(), true, Fm.data(), false, T1.data());
        gemm(K, T1.data(), false, X.data(), false, Fp.data());
        eout.assign(K, 0.0);
        Cout.assign(K2, 0.0);
        if (all_vectors) {
            sym_eig(K, Fp.data(), eout.data(), Cp.data());
            gemm(K, X.data(), false, Cp.data(), false, Cout.data());
        } else {
            sym_eig_lowest(K, Fp.data(), n_occ, eout.data(), Cp.data());
            for (int i = 0; i < K; ++i) {
                double* __restrict__ c = &Cout[(size_t)i * K];
                for (int k = 0; k < K; ++k) {
                    const double x = X[(size_t)i * K + k];
                    const double* __restrict__ cp = &Cp[(size_t)k * K];
                    for (int j = 0; j < n_occ; ++j) c[j] += x * cp[j];
                }
            }
        }
        Pout.assign(K2, 0.0);
        for (int i = 0; i < K; ++i)
            for (int j = 0; j <= i; ++j) {
                double s = 0.0;
                for (int k = 0; k < n_occ; ++k) s += Cout[(size_t)i * K + k] * Cout[(size_t)j * K + k];
                Pout[(size_t)i * K + j] = occ * s;
                Pout[(size_t)j * K + i] = occ * s;
            }
    }

    // DIIS extrapolation over the stored (F, R) history (augmented Pulay
    // system, include/linalg.hpp:405-511).
    static std::vector<double> diis(const std::deque<std::vector<double>>& Fh, const std::deque<std::vector<double>>& Rh) {
        const int n = (int)Fh.size();
        std::vector<double> B((size_t)(n + 1) * (n + 1), 0.0), rhs(n + 1, 0.0), c(n + 1, 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double d = 0.0;
                for (size_t k = 0; k < Rh[i].size(); ++k) d += Rh[i][k] * Rh[j][k];
                B[(size_t)i * (n + 1) + j] = d;
            }
        for (int i = 0; i < n; ++i) {
            B[(size_t)i * (n + 1) + n] = -1.0;
            B[(size_t)n * (n + 1) + i] = 1.0;
        }
        rhs[n] = 1.0;
        solve_fullpiv(n + 1, B.data(), rhs.data(), c.data());
        std::vector<double> out(Fh[0].size(), 0.0);
        for (int i = 0; i < n; ++i)
            for (size_t k = 0; k < out.size(); ++k) out[k] += c[i] * Fh[i][k];
        return out;
    }

    // Fock matrices from P, concatenated error vector; returns ||R||.
    double final_check(const std::vector<std::vector<double>>& Pm, std::vector<std::vector<double>>& Fm, double& E) {
        fock(Pm, Fm);
        double n2 = 0.0;
        for (int s = 0; s < nspin; ++s) {
            const double r = vnorm(error(Fm[s], Pm[s]));
            n2 += r * r;
        }
        E = energy(Pm, Fm);
        return std::sqrt(n2);
    }

    // One SCF phase.  Returns true on convergence; P/C/e hold the accepted
    // state and E_ele the final-check energy.
    bool run(const ScfOptions& o, int max_it, int start_diis, bool use_damping, int& n_it, double& E_ele) {
        const int max_hist = 8;
        std::deque<std::vector<double>> Fh, Rh;
        const size_t K2 = (size_t)K * K;
        double E_prev = 0.0;
        bool have_prev = false;
        for (int i = 0; i < max_it; ++i) {
            ++n_it;
            std::vector<std::vector<double>> Fm;
            fock(P, Fm);
            auto t0 = clk::now();
            std::vector<double> Fcat, Rcat;
            for (int s = 0; s < nspin; ++s) {
                auto R = error(Fm[s], P[s]);
                Fcat.insert(Fcat.end(), Fm[s].begin(), Fm[s].end());
                Rcat.insert(Rcat.end(), R.begin(), R.end());
            }
            const double E_it = energy(P, Fm);
            if ((int)Fh.size() == max_hist) {
                Fh.pop_front();
                Rh.pop_front();
            }
            Fh.push_back(Fcat);
            Rh.push_back(Rcat);
            std::vector<double> Fuse = Fcat;
            if (i + 1 >= start_diis) Fuse = diis(Fh, Rh);
            std::vector<std::vector<double>> Pn(nspin);
            C.assign(nspin, {});
            e.assign(nspin, {});
            last_F.assign(nspin, {});
            double dmax = 0.0;
            for (int s = 0; s < nspin; ++s) {
                last_F[s].assign(Fuse.begin() + s * K2, Fuse.begin() + (s + 1) * K2);
                diag(last_F[s], C[s], e[s], Pn[s], nocc[s], nspin == 1 ? 2.0 : 1.0, false);
                dmax = std::max(dmax, rmsd(Pn[s], P[s]));
            }
            const double err = vnorm(Rcat);
            t->diag += secs(t0, clk::now());
            bool econv = o.energy_convergence <= 0.0 || (have_prev && std::fabs(E_it - E_prev) < o.energy_convergence);
            E_prev = E_it;
            have_prev = true;
            if (o.verbose) std::printf("  iter %3d  E_ele %.12f  dP %.3e  |R| %.3e\n", n_it, E_it, dmax, err);
            if (dmax < o.epsilon && err < o.epsilon && econv) {
                std::vector<std::vector<double>> Fc;
                double Ec = 0.0;
                const double chk = final_check(Pn, Fc, Ec);
                if (chk < o.epsilon) {
                    P = Pn;
                    E_ele = Ec;
                    all_orbitals();
                    return true;
                }
                if (o.verbose) std::printf("  final check failed, error: %.3e\n", chk);
            }
            if (!use_damping || i + 1 >= start_diis) {
                P = Pn;
            } else {
                for (int s = 0; s < nspin; ++s)
                    for (size_t k = 0; k < K2; ++k) P[s][k] = o.damping * Pn[s][k] + (1.0 - o.damping) * P[s][k];
            }
            if (i == max_it - 1) {
                std::vector<std::vector<double>> Fc;
                final_check(P, Fc, E_ele);
                if (o.verbose) std::printf("  last iteration reached\n");
            }
        }
        all_orbitals();
        return false;
    }

    // Extrapolated Fock matrices of the last diagonalisation; the iterations
    // form only the occupied orbitals, the result gets all of them (the same
    // eigen-problem solved again with every vector).
    std::vector<std::vector<double>> last_F;
    void all_orbitals() {
        std::vector<double> Pd;
        for (int s = 0; s < (int)last_F.size(); ++s) diag(last_F[s], C[s], e[s], Pd, 0, 0.0, true);
    }
};

// Runs the SCF phases (the solve, then for the mixed form the FP64 polish)
// on the host loop above or, when selected, on the GPU-resident iteration
// (device/scf_device.hip); P, C and e of `core` hold the result either way.
bool solve(ScfCore& core, Engine& eng, const ScfOptions& opt, int& n_it, int& n_polish, double& E_ele,
           double& E_first) {
    const bool device = opt.device_iteration > 0 || (opt.device_iteration < 0 && core.K >= hfxd::DEVSCF_MIN_K);
    const bool polish = opt.precision == HFX_PREC_MIXED && opt.polish;
    if (!device) {
        bool ok = core.run(opt, opt.max_it, 4, true, n_it, E_ele);
        E_first = E_ele;
        if (ok && polish) {
            eng.make_fock(opt, HFX_PREC_FP64);
            ok = core.run(opt, opt.max_it, 1, false, n_polish, E_ele);
        }
        return ok;
    }
    hfxd::DevScf dev(core.K, core.nspin, core.nocc.data(), core.H.data(), core.S.data(), core.X.data());
    for (int s = 0; s < core.nspin; ++s) dev.set_density(s, core.P[s].data());
    dev.set_fock(eng.f);
    hfxd::DevScfTimes dt;
    bool ok = dev.run(opt.max_it, 4, true, opt.epsilon, opt.damping, opt.energy_convergence, opt.verbose, n_it, E_ele,
                      dt);
    E_first = E_ele;
    if (ok && polish) {
        eng.make_fock(opt, HFX_PREC_FP64);
        dev.set_fock(eng.f);
        ok = dev.run(opt.max_it, 1, false, opt.epsilon, opt.damping, opt.energy_convergence, opt.verbose, n_polish,
                     E_ele, dt);
    }
    const size_t K2 = (size_t)core.K * core.K;
    core.C.assign(core.nspin, std::vector<double>(K2));
    core.e.assign(core.nspin, {});
    for (int s = 0; s < core.nspin; ++s) {
        dev.get_density(s, core.P[s].data());
        dev.get_coefficients(s, core.C[s].data());
        core.e[s] = dev.orbital_energies(s);
    }
    core.t->fock += dt.fock;
    core.t->diag += dt.diag;
    return ok;
}

struct Prepared {
    FlatBasis fb;
    BasisInfo info;
    int Ne = 0;
    double E_nuc = 0.0;
    std::vector<double> H, S, X;
};

Prepared prepare(const Config& config, const std::string& basis_set, const std::string& param_path, int charge) {
    Prepared p;
    p.Ne = -charge;
    std::vector<int> Z;
    std::vector<double> R;
    for (const auto& a : config) {
        Z.push_back(atomic_number(a.first));
        p.Ne += Z.back();
        R.insert(R.end(), {a.second.x, a.second.y, a.second.z});
    }
    p.fb = basis_construction(config, basis_set, param_path, &p.info);
    const int K = p.fb.K;
    const size_t K2 = (size_t)K * K;
    std::vector<double> T(K2), V(K2), Zd(Z.begin(), Z.end());
    p.S.assign(K2, 0.0);
    // S, T, V on the GPU (device/oneint_device.hip)
    check(hfx_one_electron_device(K, p.fb.xyz.data(), p.fb.lmn.data(), p.fb.poff.data(), p.fb.alpha.data(),
                                  p.fb.coeff.data(), p.fb.norm.data(), (int)Z.size(), Zd.data(), R.data(),
                                  p.S.data(), T.data(), V.data()));
    p.H.resize(K2);
    for (size_t i = 0; i < K2; ++i) p.H[i] = T[i] + V[i];
    p.E_nuc = nuclear_repulsion(config);
    // X = U s^{-1/2} U^T
    std::vector<double> w(K), U(K2), Us(K2);
    sym_eig(K, p.S.data(), w.data(), U.data());
    for (int i = 0; i < K; ++i)
        for (int k = 0; k < K; ++k) Us[(size_t)i * K + k] = U[(size_t)i * K + k] / std::sqrt(w[k]);
    p.X.assign(K2, 0.0);
    gemm(K, Us.data(), false, U.data(), true, p.X.data());
    return p;
}

void make_engine(Engine& eng, const FlatBasis& fb, const ScfOptions& o, int precision) {
    check(hfx_basis_create(fb.K, fb.xyz.data(), fb.lmn.data(), fb.poff.data(), fb.alpha.data(), fb.coeff.data(),
                           fb.norm.data(), &eng.b));
    eng.make_fock(o, precision);
}

}  // namespace

double nuclear_repulsion(const Config& config) {
    double E = 0.0;
    for (size_t i = 0; i + 1 < config.size(); ++i)
        for (size_t j = i + 1; j < config.size(); ++j) {
            const Vec3& a = config[i].second;
            const Vec3& b = config[j].second;
            const double d = std::sqrt((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z));
            E += atomic_number(config[i].first) * atomic_number(config[j].first) / d;
        }
    return E;
}

std::string now_string() {
    std::time_t t = std::time(nullptr);
    char buf[64];
    std::strftime(buf, sizeof(buf), "%H:%M:%S %d/%m/%y", std::localtime(&t));
    return buf;
}

RhfResult rhf_diis(const Config& config, const std::string& basis_set, const std::string& param_path,
                   const ScfOptions& opt, const std::vector<double>* P0) {
    RhfResult out;
    out.start = now_string();
    auto t0 = clk::now();
    Prepared p = prepare(config, basis_set, param_path, opt.charge);
    const int K = p.fb.K;
    if (p.Ne % 2) throw std::runtime_error("RHF needs an even number of electrons");
    auto t1 = clk::now();
    out.t.setup = secs(t0, t1);
    Engine eng;
    make_engine(eng, p.fb, opt, opt.precision);
    out.t.eri = secs(t1, clk::now());
    hfx_fock_info(eng.f, &out.mode, nullptr, &out.n_quartets);

    ScfCore core;
    core.K = K;
    core.nspin = 1;
    core.nocc = {p.Ne / 2};
    core.H = p.H;
    core.S = p.S;
    core.X = p.X;
    core.eng = &eng;
    core.t = &out.t;
    core.P.assign(1, std::vector<double>((size_t)K * K, 0.0));
    if (P0) {
        if (P0->size() != (size_t)K * K) throw std::runtime_error("initial density matrix dimension inconsistent with basis size");
        core.P[0] = *P0;
    }
    int n_it = 0, n2 = 0;
    double E_ele = 0.0, E_first = 0.0;
    const bool ok = solve(core, eng, opt, n_it, n2, E_ele, E_first);
    out.iterations = n_it;
    out.polish_iterations = n2;
    out.E_before_polish = E_first + p.E_nuc;
    out.converged = ok;
    out.E_ele = E_ele;
    out.E_nuc = p.E_nuc;
    out.E_tot = E_ele + p.E_nuc;
    out.density = core.P[0];
    out.coefficient = core.C.empty() ? std::vector<double>() : core.C[0];
    out.mo_energy = core.e.empty() ? std::vector<double>() : core.e[0];
    out.Ne = p.Ne;
    out.K = K;
    out.info = p.info;
    out.t.total = secs(t0, clk::now());
    out.end = now_string();
    return out;
}

UhfResult uhf_diis(const Config& config, const std::string& basis_set, const std::string& param_path,
                   std::pair<int, int> spin_occ, const ScfOptions& opt, const std::vector<double>* Pa0,
                   const std::vector<double>* Pb0) {
    UhfResult out;
    out.start = now_string();
    auto t0 = clk::now();
    Prepared p = prepare(config, basis_set, param_path, opt.charge);
    if (p.Ne != spin_occ.first + spin_occ.second) throw std::runtime_error("inconsistency in electron number");
    const int K = p.fb.K;
    auto t1 = clk::now();
    out.t.setup = secs(t0, t1);
    Engine eng;
    make_engine(eng, p.fb, opt, opt.precision);
    out.t.eri = secs(t1, clk::now());
    hfx_fock_info(eng.f, &out.mode, nullptr, &out.n_quartets);

    ScfCore core;
    core.K = K;
    core.nspin = 2;
    core.nocc = {spin_occ.first, spin_occ.second};
    core.H = p.H;
    core.S = p.S;
    core.X = p.X;
    core.eng = &eng;
    core.t = &out.t;
    core.P.assign(2, std::vector<double>((size_t)K * K, 0.0));
    if (Pa0 && Pb0) {
        if (Pa0->size() != (size_t)K * K || Pb0->size() != (size_t)K * K)
            throw std::runtime_error("initial density matrix dimension inconsistent with basis size");
        core.P[0] = *Pa0;
        core.P[1] = *Pb0;
    }
    int n_it = 0, n2 = 0;
    double E_ele = 0.0, E_first = 0.0;
    const bool ok = solve(core, eng, opt, n_it, n2, E_ele, E_first);
    out.iterations = n_it;
    out.polish_iterations = n2;
    out.E_before_polish = E_first + p.E_nuc;
    out.converged = ok;
    out.E_ele = E_ele;
    out.E_nuc = p.E_nuc;
    out.E_tot = E_ele + p.E_nuc;
    out.Pa = core.P[0];
    out.Pb = core.P[1];
    if (!core.C.empty()) {
        out.Ca = core.C[0];
        out.Cb = core.C[1];
        out.Ea = core.e[0];
        out.Eb = core.e[1];
    }
    out.Na = spin_occ.first;
    out.Nb = spin_occ.second;
    out.K = K;
    out.info = p.info;
    out.t.total = secs(t0, clk::now());
    out.end = now_string();
    return out;
}

}  // namespace hfxh
