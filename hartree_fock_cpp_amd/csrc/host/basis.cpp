// basis.cpp -- see basis.hpp.
#include "basis.hpp"

#include <cmath>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "csv.hpp"

namespace hfxh {

namespace {

const std::vector<std::string> kElements = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar",
    "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};
const char* kLsym[] = {"S", "P", "D", "F", "G", "H"};

int dfact(int n) {
    int r = 1;
    for (int k = n; k > 0; k -= 2) r *= k;
    return r;
}

struct ShellRec {
    int L;
    std::string Lsym;
    std::vector<double> exp, coef;
};

std::vector<ShellRec> read_shells(const std::string& file) {
    if (!std::filesystem::exists(file) || !std::filesystem::is_regular_file(file))
        throw std::runtime_error("Critical Error: File does not exist or is invalid at path: " + file);
    CsvTable t = read_csv(file);
    const auto& sid = t.column("shell_id");
    const auto& L = t.column("L");
    const auto& pid = t.column("primitive_id");
    const auto& ex = t.column("exponent");
    const auto& co = t.column("coefficient");
    std::vector<ShellRec> out;
    std::vector<int> maxpid;
    int cur = -1;
    for (size_t r = 0; r < sid.size(); ++r) {
        const int s = std::stoi(sid[r]);
        if (s != cur) {
            if (s != (int)out.size()) throw std::runtime_error("building shells in the wrong order");
            ShellRec rec;
            rec.Lsym = L[r];
            rec.L = -1;
            for (int k = 0; k < 6; ++k)
                if (L[r] == kLsym[k]) rec.L = k;
            if (rec.L < 0) throw std::runtime_error("unknown angular momentum symbol " + L[r]);
            out.push_back(rec);
            maxpid.push_back(-1);
            cur = s;
        }
        out.back().exp.push_back(std::stod(ex[r]));
        out.back().coef.push_back(std::stod(co[r]));
        maxpid.back() = std::max(maxpid.back(), std::stoi(pid[r]));
    }
    for (size_t s = 0; s < out.size(); ++s)
        if (maxpid[s] + 1 != (int)out[s].exp.size()) throw std::runtime_error("wrong primitive number");
    return out;
}

}  // namespace

int atomic_number(const std::string& sym) {
    for (size_t i = 0; i < kElements.size(); ++i)
        if (kElements[i] == sym) return (int)i + 1;
    throw std::runtime_error("unknown element symbol: " + sym);
}

const std::string& element_symbol(int Z) {
    if (Z < 1 || Z > (int)kElements.size()) throw std::runtime_error("atomic number out of range");
    return kElements[Z - 1];
}

double cart_norm(double alpha, int lx, int ly, int lz) {
    const int L = lx + ly + lz;
    const double num = std::pow(4.0 * alpha, L);
    const double den = (double)dfact(2 * lx - 1) * (double)dfact(2 * ly - 1) * (double)dfact(2 * lz - 1);
    return std::pow(2.0 * alpha / M_PI, 0.75) * std::sqrt(num / den);
}

FlatBasis basis_construction(const Config& config, const std::string& basis_set, const std::string& param_path,
                             BasisInfo* info) {
    FlatBasis b;
    b.poff.push_back(0);
    for (const auto& at : config) {
        atomic_number(at.first);
        const auto shells = read_shells(param_path + "/" + basis_set + "/" + at.first + ".csv");
        for (size_t s = 0; s < shells.size(); ++s) {
            const ShellRec& sh = shells[s];
            for (int lx = 0; lx <= sh.L; ++lx)
                for (int ly = 0; ly <= sh.L - lx; ++ly) {
                    const int lz = sh.L - lx - ly;
                    b.xyz.insert(b.xyz.end(), {at.second.x, at.second.y, at.second.z});
                    b.lmn.insert(b.lmn.end(), {lx, ly, lz});
                    for (size_t k = 0; k < sh.exp.size(); ++k) {
                        b.alpha.push_back(sh.exp[k]);
                        b.coeff.push_back(sh.coef[k]);
                        b.norm.push_back(cart_norm(sh.exp[k], lx, ly, lz));
                    }
                    b.poff.push_back((int)b.alpha.size());
                    if (info) {
                        info->index.push_back((int)info->index.size());
                        info->shells.push_back((int)s);
                        info->lx.push_back(lx);
                        info->ly.push_back(ly);
                        info->lz.push_back(lz);
                        info->cx.push_back(at.second.x);
                        info->cy.push_back(at.second.y);
                        info->cz.push_back(at.second.z);
                        info->types.push_back(sh.Lsym);
                        info->atoms.push_back(at.first);
                    }
                }
        }
    }
    b.K = (int)b.poff.size() - 1;
    return b;
}

FlatBasis basis_recon(const std::string& basis_csv, const std::string& param_path, const std::string& basis_set) {
    CsvTable t = read_csv(basis_csv);
    const auto& atom = t.column("atom");
    const auto& sid = t.column("shell_id");
    const auto& lx = t.column("lx");
    const auto& ly = t.column("ly");
    const auto& lz = t.column("lz");
    const auto& x = t.column("x");
    const auto& y = t.column("y");
    const auto& z = t.column("z");
    FlatBasis b;
    b.poff.push_back(0);
    for (size_t r = 0; r < atom.size(); ++r) {
        const auto shells = read_shells(param_path + "/" + basis_set + "/" + atom[r] + ".csv");
        const int s = std::stoi(sid[r]);
        if (s < 0 || s >= (int)shells.size()) throw std::runtime_error("shell mismatch encountered in basis reconstruction");
        const int l[3] = {std::stoi(lx[r]), std::stoi(ly[r]), std::stoi(lz[r])};
        b.xyz.insert(b.xyz.end(), {std::stod(x[r]), std::stod(y[r]), std::stod(z[r])});
        b.lmn.insert(b.lmn.end(), {l[0], l[1], l[2]});
        for (size_t k = 0; k < shells[s].exp.size(); ++k) {
            b.alpha.push_back(shells[s].exp[k]);
            b.coeff.push_back(shells[s].coef[k]);
            b.norm.push_back(cart_norm(shells[s].exp[k], l[0], l[1], l[2]));
        }
        b.poff.push_back((int)b.alpha.size());
    }
    b.K = (int)b.poff.size() - 1;
    return b;
}

}  // namespace hfxh
