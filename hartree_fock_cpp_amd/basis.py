"""Basis construction in the reference's conventions (Python mirror of the
native host code in csrc/host/basis.cpp).

Follows KLYYCYNN/hartree_fock_cpp include/basis.hpp:
  cart_norm            basis.hpp:28-49   N = (2a/pi)^{3/4} sqrt((4a)^L / ((2lx-1)!!(2ly-1)!!(2lz-1)!!))
  angmom_partition     basis.hpp:52-64   components (lx, ly, lz) for lx = 0..l, ly = 0..l-lx
  basis_construction   basis.hpp:157-262 atoms in input order, shells in CSV shell_id order
The element table follows include/atoms.hpp:12-16.  Basis data: the CSV files
under hartree_fock_cpp_amd/basis/<set>/<El>.csv (columns shell_id, L,
primitive_id, exponent, coefficient).
"""
import csv
import math
import os

import numpy as np

BASIS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")

ELEMENTS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar", "K",
            "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"]
LSYM = ["S", "P", "D", "F", "G", "H"]


def atomic_number(sym):
    if sym not in ELEMENTS:
        raise ValueError(f"unknown element symbol {sym!r}")
    return ELEMENTS.index(sym) + 1


def double_factorial(n):
    r = 1
    while n > 0:
        r *= n
        n -= 2
    return r


def cart_norm(alpha, lx, ly, lz):
    L = lx + ly + lz
    num = (4.0 * alpha) ** L
    den = float(double_factorial(2 * lx - 1) * double_factorial(2 * ly - 1) * double_factorial(2 * lz - 1))
    return (2.0 * alpha / math.pi) ** 0.75 * math.sqrt(num / den)


def angmom_partition(l):
    return [(a, b, l - a - b) for a in range(l + 1) for b in range(l - a + 1)]


def read_shells(basis_set, symbol, basis_dir=None):
    path = os.path.join(basis_dir or BASIS_DIR, basis_set, symbol + ".csv")
    if not os.path.isfile(path):
        raise FileNotFoundError(f"Critical Error: File does not exist or is invalid at path: {path}")
    shells = []
    with open(path) as f:
        rows = list(csv.DictReader(f))
    cur = None
    for r in rows:
        sid = int(r["shell_id"])
        if cur is None or sid != cur["id"]:
            if sid != len(shells):
                raise ValueError("building shells in the wrong order")
            cur = {"id": sid, "L": LSYM.index(r["L"]), "exp": [], "coef": [], "pid": []}
            shells.append(cur)
        cur["exp"].append(float(r["exponent"]))
        cur["coef"].append(float(r["coefficient"]))
        cur["pid"].append(int(r["primitive_id"]))
    for s in shells:
        if max(s["pid"]) + 1 != len(s["exp"]):
            raise ValueError("wrong primitive number")
    return shells


def basis_construction(atoms, basis_set, basis_dir=None):
    """atoms: [(symbol, (x, y, z)), ...] in bohr.  Returns a dict of flat
    arrays (xyz, lmn, poff, alpha, coeff, norm), per-function shell ids and
    atom symbols, and the nuclear data (Zs, R)."""
    xyz, lmn, poff, al, co, nm, shell_id, fatom = [], [], [0], [], [], [], [], []
    for sym, pos in atoms:
        atomic_number(sym)
        for s in read_shells(basis_set, sym, basis_dir):
            for (lx, ly, lz) in angmom_partition(s["L"]):
                xyz.extend(float(c) for c in pos)
                lmn.extend((lx, ly, lz))
                for a, c in zip(s["exp"], s["coef"]):
                    al.append(a)
                    co.append(c)
                    nm.append(cart_norm(a, lx, ly, lz))
                poff.append(len(al))
                shell_id.append(s["id"])
                fatom.append(sym)
    return {
        "xyz": np.array(xyz, float), "lmn": np.array(lmn, np.int32), "poff": np.array(poff, np.int32),
        "alpha": np.array(al), "coeff": np.array(co), "norm": np.array(nm),
        "shell": np.array(shell_id, np.int32), "atoms": fatom, "K": len(poff) - 1,
        "Zs": np.array([atomic_number(s) for s, _ in atoms], np.int32),
        "R": np.array([p for _, p in atoms], float),
    }
