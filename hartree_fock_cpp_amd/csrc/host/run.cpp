// run.cpp -- see run.hpp.
#include "run.hpp"

#include <chrono>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <stdexcept>

#include "../../../include/hfx.h"
#include "csv.hpp"
#include "json.hpp"

namespace hfxh {

namespace {

namespace fs = std::filesystem;
using clk = std::chrono::steady_clock;

constexpr double PI = 3.14159265358979323846;

void require_key(const Json& input, const std::string& key) {
    if (!input.contains(key)) throw std::runtime_error("Missing required JSON field: " + key);
}

double deg2rad(double theta) { return theta * PI / 180.00; }

Vec3 polar3d(double r, double theta, double phi) {
    return {r * std::sin(theta) * std::cos(phi), r * std::sin(theta) * std::sin(phi), r * std::cos(theta)};
}

// the fields shared by read_scf_options and read_blscan_options
void read_common(const Json& input, RunOptions& opt) {
    require_key(input, "basis_path");
    require_key(input, "save_path");
    require_key(input, "basis");
    require_key(input, "atoms");
    opt.basis = input.at("basis").as_string();
    opt.basis_path = input.at("basis_path").as_string();
    opt.save_path = input.at("save_path").as_string();
    opt.folder_name = input.value("folder_name", "test");
    opt.charge = input.value("charge", 0);
    opt.eri_device = input.value("eri_hardware", "CPU");
    opt.max_it = input.value("max_iterations", 10000);
    opt.epsilon = input.value("epsilon", 1e-6);
    opt.damping = input.value("damping", 0.3);
    opt.screening_threshold = input.value("screening_threshold", 1e-8);
    opt.eri_mode = input.value("eri_mode", "auto");
    opt.precision = input.value("precision", "fp64");
    opt.mixed_split = input.value("mixed_split", 1e-5);
    if (opt.eri_mode != "auto" && opt.eri_mode != "stored" && opt.eri_mode != "direct")
        throw std::runtime_error("Unknown eri_mode: " + opt.eri_mode + " (auto, stored or direct)");
    if (opt.precision != "fp64" && opt.precision != "mixed")
        throw std::runtime_error("Unknown precision: " + opt.precision + " (fp64 or mixed)");
}

std::vector<std::string> symbols_of(const Config& c) {
    std::vector<std::string> s;
    for (const auto& a : c) s.push_back(a.first);
    return s;
}

void create_folder(const std::string& folder) {
    std::error_code ec;
    if (!fs::create_directories(folder, ec)) {
        if (ec)
            std::cerr << "Error creating directory: " << ec.message() << std::endl;
        else
            std::cout << "Folder already exists: " << folder << std::endl;
    }
}

void save_basis(const BasisInfo& info, const std::string& folder) {
    write_csv(folder + "/basis.csv", {"index", "atom", "shell_id", "angmom", "lx", "ly", "lz", "x", "y", "z"},
              {info.index, info.atoms, info.shells, info.types, info.lx, info.ly, info.lz, info.cx, info.cy, info.cz});
}

void save_geometry(const Config& config, const std::string& folder) {
    std::vector<std::string> at;
    std::vector<double> x, y, z;
    for (const auto& a : config) {
        at.push_back(a.first);
        x.push_back(a.second.x);
        y.push_back(a.second.y);
        z.push_back(a.second.z);
    }
    write_csv(folder + "/geometry.csv", {"atom", "x", "y", "z"}, {at, x, y, z});
}

void write_json(const std::string& path, const Json& j, bool trailing_newline) {
    std::ofstream f(path);
    if (!f.is_open()) throw std::runtime_error("Could not open JSON file: " + path);
    f << json_dump(j, 4);
    if (trailing_newline) f << "\n";
}

std::string hardware_of(const RunOptions& opt) { return opt.eri_device; }

void note_hardware(const RunOptions& opt) {
    if (opt.eri_device != "GPU")
        std::cout << "eri_hardware \"" << opt.eri_device
                  << "\": this build computes ERIs and Fock builds on the GPU engine" << std::endl;
}

}  // namespace

std::vector<double> linspace(double start, double stop, int n) {
    const double step = (stop - start) / (n - 1);
    std::vector<double> v;
    for (int i = 0; i < n; ++i) v.push_back(start + i * step);
    return v;
}

std::vector<double> scan_distances(const std::vector<std::array<double, 3>>& ranges) {
    std::vector<double> d;
    for (const auto& r : ranges) {
        std::vector<double> seg = linspace(r[0], r[1], (int)r[2]);
        if (!d.empty() && !seg.empty()) seg.erase(seg.begin());   // shared end point
        d.insert(d.end(), seg.begin(), seg.end());
    }
    return d;
}

std::string format_element_counts(const std::vector<std::string>& elements) {
    std::vector<std::pair<std::string, int>> runs;   // consecutive runs, in order of appearance
    for (const auto& e : elements) {
        if (!runs.empty() && runs.back().first == e)
            ++runs.back().second;
        else
            runs.push_back({e, 1});
    }
    std::string out;
    for (size_t i = 0; i < runs.size(); ++i) {
        if (runs[i].second > 1) out += std::to_string(runs[i].second);
        out += runs[i].first;
        if (i + 1 < runs.size()) out += ", ";
    }
    return out;
}

// (alpha, beta) occupancies of the isolated ground-state atoms H..Ar used by
// the dissociation runs (include/run.hpp:468-503).
std::pair<int, int> atom_spin_occupancy(const std::string& atom) {
    static const std::map<std::string, std::pair<int, int>> occ = {
        {"H", {1, 0}},  {"He", {1, 1}}, {"Li", {2, 1}}, {"Be", {2, 2}}, {"B", {3, 2}},  {"C", {4, 2}},
        {"N", {4, 3}},  {"O", {5, 3}},  {"F", {5, 4}},  {"Ne", {5, 5}}, {"Na", {6, 5}}, {"Mg", {6, 6}},
        {"Al", {7, 6}}, {"Si", {8, 6}}, {"P", {8, 7}},  {"S", {9, 7}},  {"Cl", {9, 8}}, {"Ar", {9, 9}}};
    auto it = occ.find(atom);
    if (it == occ.end()) throw std::runtime_error("Spin occupancy not implemented for atom: " + atom);
    return it->second;
}

ScfOptions engine_options(const RunOptions& opt) {
    ScfOptions o;
    o.charge = opt.charge;
    o.max_it = opt.max_it;
    o.epsilon = opt.epsilon;
    o.damping = opt.damping;
    o.screening_threshold = opt.screening_threshold;
    o.eri_mode = opt.eri_mode == "stored" ? HFX_MODE_STORED : opt.eri_mode == "direct" ? HFX_MODE_DIRECT : HFX_MODE_AUTO;
    o.precision = opt.precision == "mixed" ? HFX_PREC_MIXED : HFX_PREC_FP64;
    o.mixed_split = opt.mixed_split;
    o.verbose = true;
    return o;
}

RunOptions read_scf_options(const std::string& json_file) {
    const Json input = json_read_file(json_file);
    RunOptions opt;
    read_common(input, opt);
    opt.method = input.value("method", "RHF");
    if (opt.method == "UHF") {
        require_key(input, "spin_occupancy");
        opt.n_alpha = input.at("spin_occupancy").at("alpha").as_int();
        opt.n_beta = input.at("spin_occupancy").at("beta").as_int();
    }
    for (const Json& atom : input.at("atoms").arr) {
        const std::string symbol = atom.at("symbol").as_string();
        const std::string coord_type = atom.value("coordinate_type", "cartesian");
        const Json& r = atom.at("position");
        Vec3 pos;
        if (coord_type == "cartesian") {
            pos = {r.at(0).as_double(), r.at(1).as_double(), r.at(2).as_double()};
        } else if (coord_type == "polar") {
            pos = polar3d(r.at(0).as_double(), deg2rad(r.at(1).as_double()), deg2rad(r.at(2).as_double()));
        } else {
            throw std::runtime_error("Unknown coordinate_type: " + coord_type);
        }
        opt.config.push_back({symbol, pos});
    }
    return opt;
}

BlScanOptions read_blscan_options(const std::string& json_file) {
    const Json input = json_read_file(json_file);
    BlScanOptions opt;
    read_common(input, opt.base);
    require_key(input, "central_atom");
    require_key(input, "distances");
    opt.central_atom = input.at("central_atom").as_string();
    for (const Json& atom : input.at("atoms").arr) {
        const std::string symbol = atom.at("symbol").as_string();
        require_key(atom, "angular_position");
        const Json& r = atom.at("angular_position");
        if (!r.is_array() || r.size() != 2) throw std::runtime_error("angular_position must be [theta, phi]");
        opt.satellites.push_back({symbol, {r.at(0).as_double(), r.at(1).as_double()}});
    }
    for (const Json& d : input.at("distances").arr) {
        if (!d.is_array() || d.size() != 3)
            throw std::runtime_error("each distance range must be [start, stop, n_points]");
        const double start = d.at(0).as_double(), stop = d.at(1).as_double();
        const int n = d.at(2).as_int();
        if (n < 2) throw std::runtime_error("distance range n_points must be at least 2");
        if (stop <= start) throw std::runtime_error("distance range stop must be larger than start");
        opt.distances.push_back({start, stop, (double)n});
    }
    return opt;
}

void save_rhf_data(const std::string& path, const std::string& folder_name, const RhfResult& r, const Config& config,
                   const std::string& basis_set, int charge, const std::string& hardware) {
    const std::string folder = path + "/" + folder_name;
    create_folder(folder);
    Json m = Json::object();
    m["method"] = "RHF";
    m["basis"] = basis_set;
    m["n_basis"] = r.K;
    m["unit"] = "Bohr";
    m["n_iterations"] = r.iterations;
    m["converged"] = r.converged;
    m["ERI_hardware"] = hardware;
    m["system"] = format_element_counts(symbols_of(config));
    m["charge_e"] = charge;
    m["electron_number"] = r.Ne;
    m["total_energy"] = r.E_tot;
    m["electronic_energy"] = r.E_ele;
    m["nuclear_energy"] = r.E_nuc;
    m["start_time"] = r.start;
    m["finish_time"] = r.end;
    m["duration_s"] = r.t.total;
    write_json(folder + "/metadata.json", m, false);
    write_csv(folder + "/density.csv", {"density_matrix"}, {r.density});
    write_csv(folder + "/coefficient.csv", {"coefficient_matrix"}, {r.coefficient});
    write_csv(folder + "/orbital_energies.csv", {"mo_energies"}, {r.mo_energy});
    save_basis(r.info, folder);
    save_geometry(config, folder);
    std::cout << "Data successfully saved." << std::endl;
}

void save_uhf_data(const std::string& path, const std::string& folder_name, const UhfResult& r, const Config& config,
                   const std::string& basis_set, int charge, const std::string& hardware) {
    const std::string folder = path + "/" + folder_name;
    create_folder(folder);
    Json m = Json::object();
    m["method"] = "UHF";
    m["basis"] = basis_set;
    m["n_basis"] = r.K;
    m["unit"] = "Bohr";
    m["n_iterations"] = r.iterations;
    m["converged"] = r.converged;
    m["ERI_hardware"] = hardware;
    m["system"] = format_element_counts(symbols_of(config));
    m["charge_e"] = charge;
    m["alpha_occupancy"] = r.Na;
    m["beta_occupancy"] = r.Nb;
    m["electronic_energy"] = r.E_ele;
    m["nuclear_energy"] = r.E_nuc;
    m["total_energy"] = r.E_tot;
    m["start_time"] = r.start;
    m["finish_time"] = r.end;
    m["duration_s"] = r.t.total;
    write_json(folder + "/metadata.json", m, false);
    write_csv(folder + "/density.csv", {"density_alpha", "density_beta"}, {r.Pa, r.Pb});
    write_csv(folder + "/coefficient.csv", {"coefficient_alpha", "coefficient_beta"}, {r.Ca, r.Cb});
    save_basis(r.info, folder);
    save_geometry(config, folder);
    write_csv(folder + "/orbital_energies.csv", {"e_a", "e_b"}, {r.Ea, r.Eb});
    std::cout << "Data successfully saved." << std::endl;
}

void run_scf_from_options(const RunOptions& opt) {
    note_hardware(opt);
    const ScfOptions o = engine_options(opt);
    if (opt.method == "RHF") {
        RhfResult r = rhf_diis(opt.config, opt.basis, opt.basis_path, o);
        save_rhf_data(opt.save_path, opt.folder_name, r, opt.config, opt.basis, opt.charge, hardware_of(opt));
    } else if (opt.method == "UHF") {
        UhfResult r = uhf_diis(opt.config, opt.basis, opt.basis_path, {opt.n_alpha, opt.n_beta}, o);
        save_uhf_data(opt.save_path, opt.folder_name, r, opt.config, opt.basis, opt.charge, hardware_of(opt));
    } else {
        throw std::runtime_error("Unknown SCF method: " + opt.method);
    }
}

void run_bl_scan(const BlScanOptions& opt) {
    const std::string start = now_string();
    const auto t0 = clk::now();
    note_hardware(opt.base);
    const std::vector<double> d = scan_distances(opt.distances);
    const int n = (int)d.size();
    std::vector<double> E(n), homo(n), lumo(n), its(n);
    int K = -1;
    std::vector<double> prev;
    ScfOptions o = engine_options(opt.base);
    for (int i = 0; i < n; ++i) {
        Config c;
        c.push_back({opt.central_atom, {0.0, 0.0, 0.0}});
        for (const auto& s : opt.satellites)
            c.push_back({s.first, polar3d(d[i], deg2rad(s.second.first), deg2rad(s.second.second))});
        // warm start from the previous point's density, as the reference does
        RhfResult r = rhf_diis(c, opt.base.basis, opt.base.basis_path, o, prev.empty() ? nullptr : &prev);
        prev = r.density;
        const int ih = r.Ne / 2 - 1, il = r.Ne / 2;
        if (ih < 0 || il >= (int)r.mo_energy.size()) throw std::runtime_error("Invalid RHF HOMO/LUMO index");
        E[i] = r.E_tot;
        homo[i] = r.mo_energy[ih];
        lumo[i] = r.mo_energy[il];
        its[i] = (double)r.iterations;
        K = r.K;
        std::cout << i + 1 << "/" << n << " data points collect" << std::endl;
    }
    std::vector<std::string> elements = {opt.central_atom};
    for (const auto& s : opt.satellites) elements.push_back(s.first);
    const double duration = std::chrono::duration<double>(clk::now() - t0).count();

    const fs::path out_dir = fs::path(opt.base.save_path) / opt.base.folder_name;
    fs::create_directories(out_dir);
    write_csv((out_dir / "scan_data.csv").string(),
              {"distance", "total_energy", "homo_energy", "lumo_energy", "iterations"}, {d, E, homo, lumo, its});
    Json info = Json::object();
    info["system"] = format_element_counts(elements);
    info["method"] = "RHF";
    info["basis"] = opt.base.basis;
    info["basis_size"] = K;
    info["n_points"] = n;
    info["min_distance"] = d.front();
    info["max_distance"] = d.back();
    info["start"] = start;
    info["end"] = now_string();
    info["duration_seconds"] = duration;
    Json files = Json::object();
    files["scan_data"] = "scan_data.csv";
    info["files"] = files;
    write_json((out_dir / "scan_info.json").string(), info, true);
    std::cout << "Bond-length scan saved to: " << out_dir.string() << std::endl;
}

void run_dissociation(const RunOptions& opt) {
    const std::string start = now_string();
    const auto t0 = clk::now();
    if (opt.method != "RHF") throw std::runtime_error("dissociation energy only support RHF molecules");
    note_hardware(opt);
    const ScfOptions o = engine_options(opt);
    RhfResult mol = rhf_diis(opt.config, opt.basis, opt.basis_path, o);

    // distinct elements in sorted order with their counts (std::map order, as the reference)
    std::map<std::string, int> counts;
    for (const auto& a : opt.config) ++counts[a.first];
    std::vector<std::string> atoms;
    std::vector<int> numbers;
    std::vector<double> atom_E;
    double total_atoms = 0.0;
    ScfOptions oa = o;
    oa.charge = 0;
    for (const auto& kv : counts) {
        Config single = {{kv.first, {0.0, 0.0, 0.0}}};
        UhfResult a = uhf_diis(single, opt.basis, opt.basis_path, atom_spin_occupancy(kv.first), oa);
        atoms.push_back(kv.first);
        numbers.push_back(kv.second);
        atom_E.push_back(a.E_tot);
        total_atoms += kv.second * a.E_tot;
    }
    const double duration = std::chrono::duration<double>(clk::now() - t0).count();
    const double Ed = total_atoms - mol.E_tot;

    const fs::path out_dir = fs::path(opt.save_path) / opt.folder_name;
    fs::create_directories(out_dir);
    const std::string folder = out_dir.string();
    save_geometry(opt.config, folder);
    save_basis(mol.info, folder);
    write_csv((out_dir / "density.csv").string(), {"density"}, {mol.density});
    write_csv((out_dir / "coefficients.csv").string(), {"coefficient"}, {mol.coefficient});
    write_csv((out_dir / "orbital_energies.csv").string(), {"orbital_energy"}, {mol.mo_energy});
    Json info = Json::object();
    info["method"] = "RHF molecule + UHF atoms";
    info["basis"] = opt.basis;
    info["basis_size"] = mol.K;
    info["charge"] = opt.charge;
    info["n_electrons"] = mol.Ne;
    info["hardware"] = hardware_of(opt);
    info["molecule_energy"] = mol.E_tot;
    info["total_atom_energy"] = mol.E_tot + Ed;
    info["dissociation_energy_hartree"] = Ed;
    info["dissociation_energy_ev"] = Ed * 27.211386245988;
    info["dissociation_energy_kj_mol"] = Ed * 2625.499638;
    info["atoms"] = Json::array(atoms);
    info["atom_numbers"] = Json::array(numbers);
    info["atom_energies"] = Json::array(atom_E);
    info["start"] = start;
    info["end"] = now_string();
    info["duration_seconds"] = duration;
    Json files = Json::object();
    files["geometry"] = "geometry.csv";
    files["basis"] = "basis.csv";
    files["density"] = "density.csv";
    files["coefficients"] = "coefficients.csv";
    files["orbital_energies"] = "orbital_energies.csv";
    info["files"] = files;
    write_json((out_dir / "dissociation_info.json").string(), info, true);
    std::cout << "Dissociation data saved to: " << folder << std::endl;
}

}  // namespace hfxh
