"""The native apps end to end on the GPU: single_run (RHF, UHF), bl_scan and
dissociation write the reference's result folders (include/save.hpp) with the
same energies as the engine's SCF drivers.

Energies: 1e-8 Ha against the Python-bound drivers at the same options
(epsilon 1e-6, the apps' default; a warm-started scan point converges to the
same state within that); doubles in the CSV and JSON files round-trip."""
import json
import math
import os

import numpy as np
import pytest

from tests.apps_util import (BASIS_CSV_HEADER, DISSOCIATION_KEYS, RHF_METADATA_KEYS, SCAN_INFO_KEYS,
                             UHF_METADATA_KEYS, base_spec, read_csv, run_app)

pytestmark = pytest.mark.gpu

H2O = [{"symbol": "O", "coordinate_type": "cartesian", "position": [0.0, 0.0, 0.0]},
       {"symbol": "H", "coordinate_type": "polar", "position": [1.811, 52.25, 0.0]},
       {"symbol": "H", "coordinate_type": "polar", "position": [1.811, 52.25, 180.0]}]


def _h2o_config():
    t = math.radians(52.25)
    return [("O", (0.0, 0.0, 0.0)), ("H", (1.811 * math.sin(t), 0.0, 1.811 * math.cos(t))),
            ("H", (1.811 * math.sin(t) * math.cos(math.pi), 1.811 * math.sin(t) * math.sin(math.pi),
                   1.811 * math.cos(t)))]


def test_single_run_rhf(gpu, tmp_path):
    from hartree_fock_cpp_amd.scf import rhf_diis
    r = run_app("single_run", tmp_path, base_spec(tmp_path, atoms=H2O))
    print(r.stdout[-500:], r.stderr[-500:])
    assert r.returncode == 0, r.stderr
    out = tmp_path / "out"
    with open(out / "metadata.json") as f:
        text = f.read()
    m = json.loads(text)
    assert list(m) == RHF_METADATA_KEYS
    assert m["method"] == "RHF" and m["basis"] == "STO-3G" and m["n_basis"] == 7 and m["unit"] == "Bohr"
    assert m["system"] == "O, 2H" and m["electron_number"] == 10 and m["converged"] is True
    assert m["ERI_hardware"] == "GPU"
    assert text.startswith('{\n    "method": "RHF",\n')
    ref = rhf_diis(_h2o_config(), "STO-3G")
    assert abs(m["total_energy"] - ref["E_tot"]) < 1e-8
    assert abs(m["total_energy"] - (m["electronic_energy"] + m["nuclear_energy"])) < 1e-12
    hdr, rows = read_csv(out / "density.csv")
    assert hdr == "density_matrix" and len(rows) == 49
    P = np.array([float(x[0]) for x in rows]).reshape(7, 7)
    assert np.abs(P - ref["density"]).max() < 1e-6
    assert np.abs(P - P.T).max() == 0.0
    hdr, rows = read_csv(out / "coefficient.csv")
    assert hdr == "coefficient_matrix" and len(rows) == 49
    hdr, rows = read_csv(out / "orbital_energies.csv")
    assert hdr == "mo_energies" and len(rows) == 7
    e = np.array([float(x[0]) for x in rows])
    assert np.all(np.diff(e) >= 0) and np.abs(e - ref["mo_energy"]).max() < 1e-6
    hdr, rows = read_csv(out / "basis.csv")
    assert hdr == BASIS_CSV_HEADER and len(rows) == 7
    assert [r_[1] for r_ in rows] == ["O"] * 5 + ["H", "H"]
    assert [r_[3] for r_ in rows] == ["S", "S", "P", "P", "P", "S", "S"]
    hdr, rows = read_csv(out / "geometry.csv")
    assert hdr == "atom,x,y,z" and [r_[0] for r_ in rows] == ["O", "H", "H"]
    xyz = np.array([[float(v) for v in r_[1:]] for r_ in rows])
    assert np.abs(xyz - np.array([p for _, p in _h2o_config()])).max() < 1e-15


def test_single_run_uhf(gpu, tmp_path):
    from hartree_fock_cpp_amd.scf import uhf_diis
    atoms = [{"symbol": "O", "position": [-1.14, 0.0, 0.0]}, {"symbol": "O", "position": [1.14, 0.0, 0.0]}]
    spec = base_spec(tmp_path, method="UHF", spin_occupancy={"alpha": 9, "beta": 7}, atoms=atoms)
    r = run_app("single_run", tmp_path, spec)
    assert r.returncode == 0, r.stderr
    out = tmp_path / "out"
    m = json.load(open(out / "metadata.json"))
    assert list(m) == UHF_METADATA_KEYS
    assert m["alpha_occupancy"] == 9 and m["beta_occupancy"] == 7 and m["system"] == "2O"
    ref = uhf_diis([("O", (-1.14, 0.0, 0.0)), ("O", (1.14, 0.0, 0.0))], "STO-3G", (9, 7))
    assert abs(m["total_energy"] - ref["E_tot"]) < 1e-8
    hdr, rows = read_csv(out / "density.csv")
    assert hdr == "density_alpha,density_beta" and len(rows) == 100
    hdr, rows = read_csv(out / "coefficient.csv")
    assert hdr == "coefficient_alpha,coefficient_beta"
    hdr, rows = read_csv(out / "orbital_energies.csv")
    assert hdr == "e_a,e_b" and len(rows) == 10


def test_bl_scan(gpu, tmp_path):
    from hartree_fock_cpp_amd.scf import rhf_diis
    spec = base_spec(tmp_path, atoms=[{"symbol": "H", "angular_position": [0.0, 0.0]}], central_atom="H",
                     distances=[[1.0, 1.4, 3], [1.4, 3.0, 3]])
    r = run_app("bl_scan", tmp_path, spec)
    assert r.returncode == 0, r.stderr
    assert "5/5 data points collect" in r.stdout and "Bond-length scan completed successfully." in r.stdout
    out = tmp_path / "out"
    info = json.load(open(out / "scan_info.json"))
    assert list(info) == SCAN_INFO_KEYS
    assert info["system"] == "2H" and info["n_points"] == 5 and info["basis_size"] == 2
    assert info["min_distance"] == 1.0 and info["max_distance"] == 3.0
    assert info["files"] == {"scan_data": "scan_data.csv"}
    hdr, rows = read_csv(out / "scan_data.csv")
    assert hdr == "distance,total_energy,homo_energy,lumo_energy,iterations"
    d = [float(x[0]) for x in rows]
    assert np.allclose(d, [1.0, 1.2, 1.4, 2.2, 3.0], atol=1e-15)
    # the scan warm-starts each point from the previous density; the driver run
    # the same way follows the same iterations (HOMO / LUMO are eigenvalues of
    # the last DIIS-extrapolated Fock matrix, as in the reference, so they are
    # compared on the same path)
    prev = None
    for row in rows:
        R = float(row[0])
        ref = rhf_diis([("H", (0.0, 0.0, 0.0)), ("H", (0.0, 0.0, R))], "STO-3G", P0=prev)
        prev = ref["density"]
        print(R, row, ref["E_tot"], ref["mo_energy"], ref["iterations"])
        assert abs(float(row[1]) - ref["E_tot"]) < 1e-8
        assert abs(float(row[2]) - ref["mo_energy"][0]) < 1e-8
        assert abs(float(row[3]) - ref["mo_energy"][1]) < 1e-8
        assert int(float(row[4])) == ref["iterations"]


def test_dissociation(gpu, tmp_path):
    from hartree_fock_cpp_amd.scf import rhf_diis, uhf_diis
    r = run_app("dissociation", tmp_path, base_spec(tmp_path, atoms=H2O))
    assert r.returncode == 0, r.stderr
    assert "Dissociation calculation completed." in r.stdout
    out = tmp_path / "out"
    info = json.load(open(out / "dissociation_info.json"))
    assert list(info) == DISSOCIATION_KEYS
    assert info["atoms"] == ["H", "O"] and info["atom_numbers"] == [2, 1]
    mol = rhf_diis(_h2o_config(), "STO-3G")
    eH = uhf_diis([("H", (0.0, 0.0, 0.0))], "STO-3G", (1, 0))["E_tot"]
    eO = uhf_diis([("O", (0.0, 0.0, 0.0))], "STO-3G", (5, 3))["E_tot"]
    assert abs(info["molecule_energy"] - mol["E_tot"]) < 1e-8
    assert abs(info["atom_energies"][0] - eH) < 1e-8 and abs(info["atom_energies"][1] - eO) < 1e-8
    Ed = 2 * eH + eO - mol["E_tot"]
    assert abs(info["dissociation_energy_hartree"] - Ed) < 1e-8
    assert abs(info["dissociation_energy_ev"] - Ed * 27.211386245988) < 1e-6
    assert info["total_atom_energy"] == pytest.approx(2 * eH + eO, abs=1e-8)
    for name in ("geometry.csv", "basis.csv", "density.csv", "coefficients.csv", "orbital_energies.csv"):
        assert (out / name).exists()
    assert read_csv(out / "density.csv")[0] == "density"
    assert read_csv(out / "coefficients.csv")[0] == "coefficient"
    assert read_csv(out / "orbital_energies.csv")[0] == "orbital_energy"
