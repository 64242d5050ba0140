"""Writes the golden fixtures under tests/golden/ from the reference's own CPU code.

Run in the build container only (it needs the reference checkout):
    make -C oracle && python tests/golden/make_golden.py

The reference path is compiled from its sources by oracle/Makefile into
oracle/_ref/libref.so (include/integrals.hpp and its header closure, wrapped by
oracle/ref_harness.cpp).  Fixtures hold inputs and the reference's outputs
only: basis arrays as built by basis_construction (include/basis.hpp:157-262),
S/T/V (integrals.hpp:99-536), Schwarz bounds (integrals.hpp:732-752), unique
ERIs (integrals.hpp:754-787, full or at seeded sampled quartet ids), the
reference Boys function (integrals.hpp:223-266) on a grid, mpmath Boys values
on the same grid, and rows of the reference's STO-3G energy curves
(results/STO-3G/*_vr.csv).
"""
import ctypes as C
import csv
import json
import os
import sys

import numpy as np

REF = os.environ.get("HFX_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")

Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10, "Na": 11, "Mg": 12,
     "Al": 13, "Si": 14, "P": 15, "S": 16, "Cl": 17, "Ar": 18}


def lib():
    L = C.CDLL(LIB)
    L.ref_boys.restype = C.c_double
    L.ref_basis_repulsion.restype = C.c_double
    return L


def build_basis(L, atoms, bset):
    syms = b"".join(s.encode() + b"\0" for s, _ in atoms)
    xyz = np.ascontiguousarray(np.array([x for _, x in atoms], float).ravel())
    npr = C.c_int()
    K = L.ref_build_basis(len(atoms), syms, xyz.ctypes.data_as(C.c_void_p), bset.encode(),
                          os.path.join(REF, "basis").encode(), C.byref(npr))
    assert K > 0, "reference basis construction failed"
    b = dict(xyz=np.zeros(3 * K), lmn=np.zeros(3 * K, np.int32), poff=np.zeros(K + 1, np.int32),
             alpha=np.zeros(npr.value), coeff=np.zeros(npr.value), norm=np.zeros(npr.value),
             shell=np.zeros(K, np.int32))
    L.ref_basis_fill(*[b[k].ctypes.data_as(C.c_void_p) for k in ["xyz", "lmn", "poff", "alpha", "coeff", "norm", "shell"]])
    return K, b


def bargs(K, b):
    return [C.c_int(K)] + [b[k].ctypes.data_as(C.c_void_p) for k in ["xyz", "lmn", "poff", "alpha", "coeff", "norm"]]


def one_electron(L, K, b, atoms):
    S = np.zeros(K * K); T = np.zeros(K * K); V = np.zeros(K * K)
    L.ref_overlap_matrix(*bargs(K, b), S.ctypes.data_as(C.c_void_p))
    L.ref_kinetic_matrix(*bargs(K, b), T.ctypes.data_as(C.c_void_p))
    Zs = np.array([Z[s] for s, _ in atoms], np.int32)
    R = np.ascontiguousarray(np.array([x for _, x in atoms], float).ravel())
    L.ref_attraction_matrix(*bargs(K, b), C.c_int(len(atoms)), Zs.ctypes.data_as(C.c_void_p),
                            R.ctypes.data_as(C.c_void_p), V.ctypes.data_as(C.c_void_p))
    return S.reshape(K, K), T.reshape(K, K), V.reshape(K, K)


def bounds_of(L, K, b):
    P = K * (K + 1) // 2
    bd = np.zeros(P)
    L.ref_schwarz_bounds(*bargs(K, b), bd.ctypes.data_as(C.c_void_p))
    return bd


def eri_at(L, K, b, bd, thr, qids):
    out = np.zeros(len(qids))
    q = np.ascontiguousarray(qids, dtype=np.uint64)
    L.ref_eri_quartets(*bargs(K, b), bd.ctypes.data_as(C.c_void_p), C.c_double(thr), q.ctypes.data_as(C.c_void_p),
                       C.c_size_t(len(q)), out.ctypes.data_as(C.c_void_p))
    return out


def sample_qids(K, b, per_class, seed, thr_bd=None):
    """Seeded sample of unique quartet ids, stratified by the angular class of
    the function quartet (sorted la>=lb, lc>=ld, bra>=ket): candidates are
    drawn uniformly from the Q unique ids in batches and each class keeps the
    first per_class distinct ids it receives."""
    rng = np.random.default_rng(seed)
    lsum = b["lmn"].reshape(K, 3).sum(1).astype(np.int64)
    P = K * (K + 1) // 2
    Q = P * (P + 1) // 2
    pi = np.repeat(np.arange(K), np.arange(1, K + 1))
    pj = np.concatenate([np.arange(i + 1) for i in range(K)])
    lo_of_pair = np.minimum(lsum[pi], lsum[pj])
    hi_of_pair = np.maximum(lsum[pi], lsum[pj])
    pcode = hi_of_pair * 4 + lo_of_pair      # (la, lb) with la >= lb, ordered like the tuple
    idx = {}
    for tries in range(400):
        cand = rng.integers(0, Q, size=200000, dtype=np.int64)
        IJ = ((np.sqrt(8.0 * cand + 1.0) - 1.0) // 2).astype(np.int64)
        IJ -= (IJ * (IJ + 1) // 2 > cand)
        IJ += ((IJ + 1) * (IJ + 2) // 2 <= cand)
        KL = cand - IJ * (IJ + 1) // 2
        cb, ck = pcode[IJ], pcode[KL]
        code = np.maximum(cb, ck) * 16 + np.minimum(cb, ck)
        for c in np.unique(code):
            have = idx.setdefault(int(c), [])
            if len(have) >= per_class:
                continue
            seen = set(have)
            for q in cand[code == c]:
                q = int(q)
                if q not in seen:
                    have.append(q)
                    seen.add(q)
                    if len(have) >= per_class:
                        break
        if tries >= 20 and all(len(v) >= per_class for v in idx.values()):
            break
    out = sorted(q for v in idx.values() for q in v)
    return np.array(out, np.uint64)


def save(name, **kw):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **kw)
    print("wrote", path, os.path.getsize(path), "bytes")


def molecule_fixture(L, name, atoms, bset, thr=1e-8, full=True, per_class=0, seed=0):
    K, b = build_basis(L, atoms, bset)
    S, T, V = one_electron(L, K, b, atoms)
    bd = bounds_of(L, K, b)
    P = K * (K + 1) // 2
    Q = P * (P + 1) // 2
    kw = dict(b, K=K, S=S, T=T, V=V, bounds=bd, threshold=thr,
              Zs=np.array([Z[s] for s, _ in atoms], np.int32), R=np.array([x for _, x in atoms], float),
              symbols=np.array([s for s, _ in atoms]), basis_set=bset)
    if full:
        kw["eri"] = eri_at(L, K, b, bd, thr, np.arange(Q, dtype=np.uint64))
    else:
        q = sample_qids(K, b, per_class, seed)
        kw["qids"] = q
        kw["eri_sample"] = eri_at(L, K, b, bd, thr, q)
        kw["sample_seed"] = seed
    save(name, **kw)


def boys_fixture(L):
    import mpmath as mp
    mp.mp.dps = 40
    Ts = [0.0, 1e-12, 1e-9, 1e-6, 5e-5, 1e-4, 2e-4, 1e-3, 5e-3, 0.01, 0.02475, 0.05, 0.1, 0.2, 0.3, 0.5, 0.77, 1.0,
          1.5, 2.0, 3.0, 5.0, 7.5, 10.0, 12.34, 15.0, 20.0, 25.0, 30.0, 35.0, 39.99, 40.0, 40.01, 45.0, 50.0,
          75.0, 100.0, 250.0, 1000.0]
    ms = list(range(0, 17))
    exact = []
    ref = []
    for m in ms:
        row, rrow = [], []
        for T in Ts:
            if T == 0.0:
                v = mp.mpf(1) / (2 * m + 1)
            else:
                v = mp.gammainc(m + mp.mpf(1) / 2, 0, T) / (2 * mp.power(T, m + mp.mpf(1) / 2))
            row.append(mp.nstr(v, 30))
            rrow.append(repr(L.ref_boys(C.c_int(m), C.c_double(T))))
        exact.append(row)
        ref.append(rrow)
    path = os.path.join(HERE, "boys_mpmath.json")
    with open(path, "w") as f:
        json.dump({"T": Ts, "m": ms, "exact": exact, "reference": ref,
                   "note": "exact = mpmath gammainc(m+1/2,0,T)/(2 T^(m+1/2)) at 40 digits; "
                           "reference = boys() of include/integrals.hpp:223-266 compiled from the reference"}, f)
    print("wrote", path)


def vr_fixture():
    rows = []
    picks = {"H2_vr.csv": ("H", "H", 0), "HeHp_vr.csv": ("He", "H", 1), "LiH_vr_1.csv": ("Li", "H", 0),
             "HF_vr.csv": ("H", "F", 0)}
    for fn, (a, bb, ch) in picks.items():
        with open(os.path.join(REF, "results", "STO-3G", fn)) as f:
            data = list(csv.reader(f))[1:]
        n = len(data)
        for k in np.linspace(0, n - 1, 9).astype(int):
            rows.append([fn, a, bb, ch, data[k][0], data[k][1]])
        # the point nearest the minimum too
        emin = min(range(n), key=lambda k: float(data[k][1]))
        rows.append([fn, a, bb, ch, data[emin][0], data[emin][1]])
    path = os.path.join(HERE, "sto3g_vr_rows.csv")
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["source", "atom_a", "atom_b", "charge", "R", "E"])
        w.writerows(rows)
    print("wrote", path)


H2O_C1 = [("O", (0.0, 0.0, 0.0)), ("H", (1.430, 1.107, 0.0)), ("H", (-1.430, 1.107, 0.0))]
O2 = [("O", (-1.14, 0.0, 0.0)), ("O", (1.14, 0.0, 0.0))]


def benzene():
    import math
    at = []
    for k in range(6):
        c, s = math.cos(math.radians(60 * k)), math.sin(math.radians(60 * k))
        at.append(("C", (2.637 * c, 2.637 * s, 0.0)))
    for k in range(6):
        c, s = math.cos(math.radians(60 * k)), math.sin(math.radians(60 * k))
        at.append(("H", (4.684 * c, 4.684 * s, 0.0)))
    return at


def eri_unique_threaded(fn, Q, nthreads=8, chunk=1 << 20):
    """The full unique array through fn(qids) -> values, on host threads
    (ctypes releases the GIL inside the C call)."""
    from concurrent.futures import ThreadPoolExecutor
    out = np.zeros(Q)

    def work(lo):
        hi = min(Q, lo + chunk)
        out[lo:hi] = fn(np.arange(lo, hi, dtype=np.uint64))

    with ThreadPoolExecutor(nthreads) as ex:
        list(ex.map(work, range(0, Q, chunk)))
    return out


def benzene_energy_fixture(L):
    """Benzene / cc-pVDZ RHF total energy (threshold 1e-8, epsilon 1e-10) from
    the numpy SCF oracle on (a) the reference's own unique ERIs (libref, the
    reference's Boys function) and (b) the oracle's exact-Boys unique ERIs,
    with the reference's S, T, V."""
    import time
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import scf_oracle
    from tests import oracle_py as orc
    from hartree_fock_cpp_amd.workloads import nuclear_repulsion
    atoms = benzene()
    K, b = build_basis(L, atoms, "cc-pVDZ")
    S, T, V = one_electron(L, K, b, atoms)
    bd = bounds_of(L, K, b)
    thr = 1e-8
    P = K * (K + 1) // 2
    Q = P * (P + 1) // 2
    d = dict(b, bounds=bd)
    res = {"molecule": "benzene", "basis_set": "cc-pVDZ", "K": K, "threshold": thr, "epsilon": 1e-10,
           "geometry": [[s, list(x)] for s, x in atoms]}
    E_nuc = nuclear_repulsion(atoms)
    ne = 42
    t0 = time.time()
    eri_ref = eri_unique_threaded(lambda q: eri_at(L, K, b, bd, thr, q), Q)
    print("reference unique array", time.time() - t0, "s", flush=True)
    E, Pm, Cm, e, it, ok = scf_oracle.rhf(S, T + V, E_nuc, ne, lambda Pd: orc.repulsion_matrix(Pd, eri_ref), eps=1e-10)
    assert ok
    res["E_reference_eris"] = float(E)
    res["iterations_reference_eris"] = it
    del eri_ref
    bd_x = orc.schwarz_bounds_exact(d)
    t0 = time.time()
    eri_x = eri_unique_threaded(lambda q: orc.eri_quartets(d, thr, q, bounds=bd_x, exact=True), Q)
    print("exact-Boys unique array", time.time() - t0, "s", flush=True)
    E2, Pm2, Cm2, e2, it2, ok2 = scf_oracle.rhf(S, T + V, E_nuc, ne, lambda Pd: orc.repulsion_matrix(Pd, eri_x), eps=1e-10)
    assert ok2
    res["E_exact_eris"] = float(E2)
    res["iterations_exact_eris"] = it2
    res["note"] = ("E_reference_eris: numpy SCF oracle on the reference's own unique ERIs and S/T/V "
                   "(include/integrals.hpp compiled in place); E_exact_eris: the same on the oracle's "
                   "exact-Boys unique ERIs and bounds")
    path = os.path.join(HERE, "benzene_ccpvdz_energy.json")
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    print("wrote", path, res)


if __name__ == "__main__":
    L = lib()
    which = set(sys.argv[1:]) or {"boys", "vr", "h2o_sto3g", "h2o_ccpvdz", "h2_sto3g", "o2_ccpvtz", "benzene",
                                  "octane", "water4"}
    if "boys" in which:
        boys_fixture(L)
    if "vr" in which:
        vr_fixture()
    if "h2_sto3g" in which:
        molecule_fixture(L, "h2_sto3g.npz", [("H", (0.0, 0.0, 0.0)), ("H", (0.0, 0.0, 1.4))], "STO-3G")
    if "h2o_sto3g" in which:
        molecule_fixture(L, "h2o_sto3g.npz", H2O_C1, "STO-3G")
    if "h2o_ccpvdz" in which:
        molecule_fixture(L, "h2o_ccpvdz.npz", H2O_C1, "cc-pVDZ")
    if "o2_ccpvtz" in which:
        molecule_fixture(L, "o2_ccpvtz.npz", O2, "cc-pVTZ", full=False, per_class=200, seed=7)
    if "benzene" in which:
        molecule_fixture(L, "benzene_ccpvdz.npz", benzene(), "cc-pVDZ", full=False, per_class=200, seed=11)
    if "octane" in which or "water4" in which:
        sys.path.insert(0, ROOT)
        from hartree_fock_cpp_amd import workloads as W
        if "octane" in which:   # n-octane / cc-pVDZ, jitter seed 1, Schwarz 1e-12
            molecule_fixture(L, "octane_ccpvdz.npz", W.n_octane(seed=1), "cc-pVDZ", thr=1e-12, full=False,
                             per_class=200, seed=1)
        if "water4" in which:   # (H2O)4 / cc-pVTZ, jitter seed 4, Schwarz 1e-12
            molecule_fixture(L, "water4_ccpvtz.npz", W.water_tetramer(seed=4), "cc-pVTZ", thr=1e-12, full=False,
                             per_class=200, seed=4)
    if "benzene_energy" in which:
        benzene_energy_fixture(L)
