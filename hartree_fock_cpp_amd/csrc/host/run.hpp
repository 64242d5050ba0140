// run.hpp -- the apps' drivers: JSON options, single runs, bond-length scans
// and dissociation energies, with the reference's result folders.
//
//   read_scf_options / run_scf_from_options   include/run.hpp:25-172
//   read_blscan_options / run_bl_scan         include/run.hpp:174-423
//   run_dissociation                          include/run.hpp:425-577
//   save_rhf_data / save_uhf_data / save_blscan / save_disstn_data
//                                             include/save.hpp:12-461
//
// Same JSON keys, defaults and error messages; the ERIs and Fock builds run
// on the GPU engine whatever "eri_hardware" says (the reference's "CPU" path
// is its host ERI tensor, which this build does not ship), and the keys
// "eri_mode" (auto | stored | direct), "precision" (fp64 | mixed) and
// "mixed_split" select the engine's forms.
#pragma once
#include <array>
#include <string>
#include <utility>
#include <vector>

#include "basis.hpp"
#include "scf.hpp"

namespace hfxh {

struct RunOptions {
    std::string method = "RHF";
    std::string basis, basis_path, save_path;
    std::string folder_name = "test";
    int charge = 0;
    int n_alpha = -1, n_beta = -1;
    std::string eri_device = "CPU";
    int max_it = 10000;
    double epsilon = 1e-6;
    double damping = 0.3;
    double screening_threshold = 1e-8;
    std::string eri_mode = "auto";
    std::string precision = "fp64";
    double mixed_split = 1e-5;
    Config config;
};

struct BlScanOptions {
    RunOptions base;   // method is always RHF; config unused
    std::string central_atom;
    std::vector<std::pair<std::string, std::pair<double, double>>> satellites;   // symbol, {theta, phi} in degrees
    std::vector<std::array<double, 3>> distances;                                // {start, stop, n_points}
};

RunOptions read_scf_options(const std::string& json_file);
BlScanOptions read_blscan_options(const std::string& json_file);

void run_scf_from_options(const RunOptions& opt);
void run_bl_scan(const BlScanOptions& opt);
void run_dissociation(const RunOptions& opt);

// helpers shared with the tests
ScfOptions engine_options(const RunOptions& opt);
std::string format_element_counts(const std::vector<std::string>& elements);
std::vector<double> linspace(double start, double stop, int n);
std::pair<int, int> atom_spin_occupancy(const std::string& atom);
std::vector<double> scan_distances(const std::vector<std::array<double, 3>>& ranges);

void save_rhf_data(const std::string& path, const std::string& folder_name, const RhfResult& r, const Config& config,
                   const std::string& basis_set, int charge, const std::string& hardware);
void save_uhf_data(const std::string& path, const std::string& folder_name, const UhfResult& r, const Config& config,
                   const std::string& basis_set, int charge, const std::string& hardware);

}  // namespace hfxh
