"""Helpers for the app tests: write a JSON input, run an app binary."""
import json
import os
import subprocess

from tests.conftest import ROOT

LIB = os.path.join(ROOT, "hartree_fock_cpp_amd", "_lib")
BASIS = os.path.join(ROOT, "hartree_fock_cpp_amd", "basis")

# the reference's result-file schemas (include/save.hpp)
RHF_METADATA_KEYS = ["method", "basis", "n_basis", "unit", "n_iterations", "converged", "ERI_hardware", "system",
                     "charge_e", "electron_number", "total_energy", "electronic_energy", "nuclear_energy",
                     "start_time", "finish_time", "duration_s"]
UHF_METADATA_KEYS = ["method", "basis", "n_basis", "unit", "n_iterations", "converged", "ERI_hardware", "system",
                     "charge_e", "alpha_occupancy", "beta_occupancy", "electronic_energy", "nuclear_energy",
                     "total_energy", "start_time", "finish_time", "duration_s"]
SCAN_INFO_KEYS = ["system", "method", "basis", "basis_size", "n_points", "min_distance", "max_distance", "start",
                  "end", "duration_seconds", "files"]
DISSOCIATION_KEYS = ["method", "basis", "basis_size", "charge", "n_electrons", "hardware", "molecule_energy",
                     "total_atom_energy", "dissociation_energy_hartree", "dissociation_energy_ev",
                     "dissociation_energy_kj_mol", "atoms", "atom_numbers", "atom_energies", "start", "end",
                     "duration_seconds", "files"]
BASIS_CSV_HEADER = "index,atom,shell_id,angmom,lx,ly,lz,x,y,z"


def run_app(name, tmp_path, spec, args=None, timeout=300):
    """Writes spec (dict or raw str) to tmp_path/input.json and runs the app."""
    inp = os.path.join(str(tmp_path), "input.json")
    with open(inp, "w") as f:
        f.write(spec if isinstance(spec, str) else json.dumps(spec, indent=4))
    cmd = [os.path.join(LIB, name)] + ([inp] if args is None else args)
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)


def base_spec(tmp_path, **kw):
    spec = {"method": "RHF", "basis": "STO-3G", "basis_path": BASIS, "save_path": str(tmp_path),
            "folder_name": "out", "charge": 0, "eri_hardware": "GPU", "max_iterations": 10000,
            "epsilon": 1e-6, "damping": 0.3, "screening_threshold": 1e-8}
    spec.update(kw)
    return spec


def read_csv(path):
    with open(path) as f:
        lines = f.read().splitlines()
    return lines[0], [l.split(",") for l in lines[1:]]
