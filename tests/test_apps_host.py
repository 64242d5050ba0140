"""The apps' option handling on the CPU: usage lines, the reference's error
messages for missing or malformed JSON fields (include/run.hpp:15-300), and
exit status 1 with "Error: ..." on stderr (apps/*.cpp).  These paths stop
before any GPU call."""
import os

import pytest

from tests.apps_util import LIB, base_spec, run_app

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(LIB, "single_run")), reason="apps not built")

H2O = [{"symbol": "O", "coordinate_type": "cartesian", "position": [0.0, 0.0, 0.0]},
       {"symbol": "H", "coordinate_type": "polar", "position": [1.811, 52.25, 0.0]},
       {"symbol": "H", "coordinate_type": "polar", "position": [1.811, 52.25, 180.0]}]


@pytest.mark.parametrize("app,usage", [("single_run", "input.json"), ("bl_scan", "<input.json>"),
                                       ("dissociation", "<input.json>")])
def test_usage(tmp_path, app, usage):
    r = run_app(app, tmp_path, "{}", args=[])
    assert r.returncode == 1 and "Usage:" in r.stderr and usage in r.stderr


@pytest.mark.parametrize("missing", ["basis_path", "save_path", "basis", "atoms"])
def test_missing_required_field(tmp_path, missing):
    spec = base_spec(tmp_path, atoms=H2O)
    del spec[missing]
    r = run_app("single_run", tmp_path, spec)
    assert r.returncode == 1
    assert r.stderr.strip() == f"Error: Missing required JSON field: {missing}"


def test_unknown_coordinate_type(tmp_path):
    atoms = [dict(H2O[0], coordinate_type="spherical")]
    r = run_app("single_run", tmp_path, base_spec(tmp_path, atoms=atoms))
    assert r.returncode == 1 and "Error: Unknown coordinate_type: spherical" in r.stderr


def test_uhf_needs_spin_occupancy(tmp_path):
    r = run_app("single_run", tmp_path, base_spec(tmp_path, method="UHF", atoms=H2O))
    assert r.returncode == 1 and "Missing required JSON field: spin_occupancy" in r.stderr


def test_unknown_method(tmp_path):
    r = run_app("single_run", tmp_path, base_spec(tmp_path, method="ROHF", atoms=H2O))
    assert r.returncode == 1 and "Error: Unknown SCF method: ROHF" in r.stderr


def test_dissociation_rejects_uhf(tmp_path):
    spec = base_spec(tmp_path, method="UHF", spin_occupancy={"alpha": 5, "beta": 5}, atoms=H2O)
    r = run_app("dissociation", tmp_path, spec)
    assert r.returncode == 1 and "dissociation energy only support" in r.stderr


def test_malformed_json(tmp_path):
    r = run_app("single_run", tmp_path, '{"basis": "STO-3G",, }')
    assert r.returncode == 1 and r.stderr.startswith("Error: JSON parse error")


def test_missing_file(tmp_path):
    r = run_app("single_run", tmp_path, "{}", args=[os.path.join(str(tmp_path), "nope.json")])
    assert r.returncode == 1 and "Could not open JSON file" in r.stderr


def test_engine_form_keys_are_checked(tmp_path):
    r = run_app("single_run", tmp_path, base_spec(tmp_path, atoms=H2O, precision="fp16"))
    assert r.returncode == 1 and "Unknown precision: fp16" in r.stderr
    r = run_app("single_run", tmp_path, base_spec(tmp_path, atoms=H2O, eri_mode="cached"))
    assert r.returncode == 1 and "Unknown eri_mode: cached" in r.stderr


SCAN_ATOMS = [{"symbol": "H", "angular_position": [0.0, 0.0]}]


@pytest.mark.parametrize("distances,msg", [
    ([[1.0, 2.0, 1]], "distance range n_points must be at least 2"),
    ([[2.0, 1.0, 5]], "distance range stop must be larger than start"),
    ([[1.0, 2.0]], "each distance range must be [start, stop, n_points]"),
])
def test_scan_distance_errors(tmp_path, distances, msg):
    spec = base_spec(tmp_path, atoms=SCAN_ATOMS, central_atom="H", distances=distances)
    r = run_app("bl_scan", tmp_path, spec)
    assert r.returncode == 1 and msg in r.stderr


def test_scan_angular_position_shape(tmp_path):
    spec = base_spec(tmp_path, atoms=[{"symbol": "H", "angular_position": [0.0]}], central_atom="H",
                     distances=[[1.0, 2.0, 3]])
    r = run_app("bl_scan", tmp_path, spec)
    assert r.returncode == 1 and "angular_position must be [theta, phi]" in r.stderr


@pytest.mark.parametrize("missing", ["central_atom", "distances"])
def test_scan_required_fields(tmp_path, missing):
    spec = base_spec(tmp_path, atoms=SCAN_ATOMS, central_atom="H", distances=[[1.0, 2.0, 3]])
    del spec[missing]
    r = run_app("bl_scan", tmp_path, spec)
    assert r.returncode == 1 and f"Missing required JSON field: {missing}" in r.stderr
