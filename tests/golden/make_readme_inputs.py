"""Writes readme_molecules.json: the reference's example inputs for the eight
molecules of its published timing table (README.md:137-154: RHF/cc-pVTZ SCF
wall time, i9-14900K CPU path vs RTX 5090 "eri_hardware":"GPU" path), with
the published seconds beside them.  Run in the container that holds the
reference checkout:  python tests/golden/make_readme_inputs.py
Geometry is copied as data (symbol, coordinate_type, position) from
examples/single_run/<molecule>.json."""
import json
import os

REF = "/root/reference/examples/single_run"
HERE = os.path.dirname(os.path.abspath(__file__))

# README.md:145-152: molecule -> (K, CPU s, GPU s)
PUBLISHED = {"CH4": (95, 352, 134), "Cl2": (78, 2301, 251), "MgO": (74, 1185, 161), "CO2": (105, 2018, 520),
             "NH3": (80, 207, 76), "NaF": (74, 1180, 159), "NaCl": (78, 2525, 268), "SO2": (109, 3883, 609)}


def main():
    out = {}
    for name, (K, cpu_s, gpu_s) in PUBLISHED.items():
        with open(os.path.join(REF, name + ".json")) as f:
            d = json.load(f)
        out[name] = {"basis": d["basis"], "charge": d.get("charge", 0), "method": d.get("method", "RHF"),
                     "atoms": [{"symbol": a["symbol"], "coordinate_type": a.get("coordinate_type", "cartesian"),
                                "position": a["position"]} for a in d["atoms"]],
                     "K": K, "published_cpu_s": cpu_s, "published_gpu_s": gpu_s,
                     "source": f"examples/single_run/{name}.json; README.md timing table"}
    with open(os.path.join(HERE, "readme_molecules.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
