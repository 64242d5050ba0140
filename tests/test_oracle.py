"""Pins the oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against something independent of its own
formula: the compiled reference's golden vectors, mpmath, closed forms and
textbook values, brute-force differentiation, and invariants of the tensor.
"""
import json
import math
import os

import numpy as np
import pytest

from tests import oracle_py as orc
from tests.conftest import GOLDEN, golden


def _boys_golden():
    with open(os.path.join(GOLDEN, "boys_mpmath.json")) as f:
        return json.load(f)


def test_boys_reference_restatement_matches_compiled_reference():
    g = _boys_golden()
    for mi, m in enumerate(g["m"]):
        for ti, T in enumerate(g["T"]):
            ref = float(g["reference"][mi][ti])
            got = orc.boys_ref(m, T)
            assert got == ref or abs(got - ref) <= 1e-15 * abs(ref), (m, T, got, ref)


def test_boys_exact_matches_mpmath():
    g = _boys_golden()
    for mi, m in enumerate(g["m"]):
        for ti, T in enumerate(g["T"]):
            ex = float(g["exact"][mi][ti])
            got = orc.boys_exact(m, T)
            assert abs(got - ex) <= 2e-15 * abs(ex), (m, T, got, ex)


def test_boys_reference_is_ill_conditioned_where_expected():
    # The reference's upward recursion loses all digits at high m / small T
    # (integrals.hpp:256-263); the exact form does not.
    g = _boys_golden()
    m = 12
    T = 1e-3
    ex = float(g["exact"][g["m"].index(m)][g["T"].index(T)])
    assert abs(orc.boys_ref(m, T) - ex) > 1.0
    assert abs(orc.boys_exact(m, T) - ex) < 1e-16


@pytest.mark.parametrize("name", ["h2_sto3g.npz", "h2o_sto3g.npz", "h2o_ccpvdz.npz"])
def test_oracle_eri_matches_golden(name):
    d = golden(name)
    got = orc.eri_unique(d, float(d["threshold"]))
    assert np.array_equal(got, d["eri"]) or np.abs(got - d["eri"]).max() < 1e-15


@pytest.mark.parametrize("name", ["o2_ccpvtz.npz", "benzene_ccpvdz.npz", "octane_ccpvdz.npz", "water4_ccpvtz.npz"])
def test_oracle_eri_sampled_matches_golden(name):
    d = golden(name)
    got = orc.eri_quartets(d, float(d["threshold"]), d["qids"])
    assert np.abs(got - d["eri_sample"]).max() < 1e-14


@pytest.mark.parametrize("name", ["h2o_sto3g.npz", "h2o_ccpvdz.npz"])
def test_oracle_bounds_and_one_electron_match_golden(name):
    d = golden(name)
    assert np.abs(orc.schwarz_bounds(d) - d["bounds"]).max() < 1e-15
    S, T, V = orc.one_electron(d)
    for got, ref in ((S, d["S"]), (T, d["T"]), (V, d["V"])):
        assert np.abs(got - ref).max() < 1e-13 * max(1.0, np.abs(ref).max())


def test_szabo_ostlund_h2_sto3g():
    # Szabo & Ostlund, Modern Quantum Chemistry, H2 / STO-3G (zeta = 1.24), R = 1.4 bohr
    d = golden("h2_sto3g.npz")
    S, T, V = orc.one_electron(d)
    assert abs(S[0, 1] - 0.6593) < 1e-4
    eri = orc.eri_unique(d, 0.0, exact=True)
    g = lambda i, j, k, l: eri[orc.eri_index(i, j, k, l)]
    assert abs(g(0, 0, 0, 0) - 0.7746) < 1e-4
    assert abs(g(0, 0, 1, 1) - 0.5697) < 1e-4
    assert abs(g(1, 0, 1, 0) - 0.2970) < 1e-4
    assert abs(g(1, 0, 0, 0) - 0.4441) < 1e-4


def _ssss_closed_form(d, i, j, k, l):
    """(ss|ss) = sum c c c c N N N N 2 pi^{5/2}/(pq sqrt(p+q)) K_AB K_CD F0(T)."""
    xyz = d["xyz"].reshape(-1, 3)
    poff = d["poff"]
    tot = 0.0
    for a in range(poff[i], poff[i + 1]):
        for b in range(poff[j], poff[j + 1]):
            for c in range(poff[k], poff[k + 1]):
                for e in range(poff[l], poff[l + 1]):
                    al, be, ga, de = d["alpha"][a], d["alpha"][b], d["alpha"][c], d["alpha"][e]
                    p, q = al + be, ga + de
                    A, B, Cc, D = xyz[i], xyz[j], xyz[k], xyz[l]
                    P = (al * A + be * B) / p
                    Q = (ga * Cc + de * D) / q
                    Kab = math.exp(-al * be / p * np.dot(A - B, A - B))
                    Kcd = math.exp(-ga * de / q * np.dot(Cc - D, Cc - D))
                    T = p * q / (p + q) * np.dot(P - Q, P - Q)
                    F0 = 1.0 if T == 0 else 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
                    w = 1.0
                    for x in (a, b, c, e):
                        w *= d["coeff"][x] * d["norm"][x]
                    tot += w * 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kcd * F0
    return tot


def test_ssss_closed_form():
    d = golden("h2o_ccpvdz.npz")
    L = d["lmn"].reshape(-1, 3).sum(1)
    s = [i for i in range(len(L)) if L[i] == 0]
    rng = np.random.default_rng(3)
    for _ in range(12):
        i, j, k, l = rng.choice(s, 4)
        got = orc.basis_repulsion(d, int(i), int(j), int(k), int(l), exact=True)
        assert abs(got - _ssss_closed_form(d, i, j, k, l)) < 1e-13


def test_eight_fold_symmetry_and_schwarz():
    d = golden("o2_ccpvtz.npz")
    K = int(d["K"])
    rng = np.random.default_rng(5)
    for _ in range(6):
        i, j, k, l = (int(x) for x in rng.integers(0, K, 4))
        v = orc.basis_repulsion(d, i, j, k, l, exact=True)
        for perm in ((j, i, k, l), (i, j, l, k), (k, l, i, j), (l, k, j, i)):
            assert abs(orc.basis_repulsion(d, *perm, exact=True) - v) < 1e-13
        bij = orc.basis_repulsion(d, i, j, i, j, exact=True)
        bkl = orc.basis_repulsion(d, k, l, k, l, exact=True)
        assert v * v <= bij * bkl * (1 + 1e-12) + 1e-24


def test_supermatrix_positive_semidefinite():
    d = golden("h2o_ccpvdz.npz")
    K = int(d["K"])
    P = K * (K + 1) // 2
    eri = orc.eri_unique(d, 0.0, exact=True)
    M = np.zeros((P, P))
    iu = np.tril_indices(P)
    M[iu] = eri
    M = M + M.T - np.diag(np.diag(M))
    assert np.linalg.eigvalsh(M).min() > -1e-10


def test_hermite_E_closed_forms():
    import ctypes as C
    a, b, A, B = 0.8, 1.3, 0.2, -0.7
    p = a + b
    mu = a * b / p
    X = A - B
    P = (a * A + b * B) / p
    E00 = math.exp(-mu * X * X)
    out = np.zeros(9)
    orc._lib.orc_hermite_E(0, 0, C.c_double(a), C.c_double(b), C.c_double(A), C.c_double(B), orc._p(out))
    assert abs(out[0] - E00) < 1e-15
    orc._lib.orc_hermite_E(1, 1, C.c_double(a), C.c_double(b), C.c_double(A), C.c_double(B), orc._p(out))
    # E^{11}_0 = (X_PA X_PB + 1/(2p)) E00, E^{11}_1 = (X_PA + X_PB)/(2p) E00, E^{11}_2 = E00/(4p^2)
    assert abs(out[0] - ((P - A) * (P - B) + 0.5 / p) * E00) < 1e-15
    assert abs(out[1] - ((P - A) + (P - B)) / (2 * p) * E00) < 1e-15
    assert abs(out[2] - E00 / (4 * p * p)) < 1e-15


def test_coulomb_R_brute_force_derivatives():
    # R_tuv = d^t/dX^t d^u/dY^u d^v/dZ^v F0(rho (X^2+Y^2+Z^2))  (n = 0)
    import ctypes as C

    import mpmath as mp
    mp.mp.dps = 30
    p, q = 1.7, 0.9
    rho = p * q / (p + q)
    Pv = np.array([0.3, -0.2, 0.5])
    Qv = np.array([-0.4, 0.1, 0.25])
    X0 = Pv - Qv

    def F0(x, y, z):
        T = rho * (x * x + y * y + z * z)
        return mp.sqrt(mp.pi / T) / 2 * mp.erf(mp.sqrt(T))

    with orc.boys_mode(1):
        for t, u, v in [(0, 0, 0), (1, 0, 0), (0, 2, 0), (1, 1, 1), (2, 0, 2), (3, 1, 0)]:
            got = orc._lib.orc_coulomb_R(t, u, v, 0, C.c_double(p), C.c_double(q), orc._p(Pv), orc._p(Qv))
            ref = float(mp.diff(F0, (mp.mpf(X0[0]), mp.mpf(X0[1]), mp.mpf(X0[2])), (t, u, v)))
            assert abs(got - ref) < 1e-12 * max(1.0, abs(ref)), (t, u, v, got, ref)


@pytest.mark.parametrize("name", ["h2o_sto3g.npz", "h2o_ccpvdz.npz"])
def test_jk_oracle_matches_einsum(name):
    d = golden(name)
    K = int(d["K"])
    eri = d["eri"]
    full = orc.expand(eri, K)
    rng = np.random.default_rng(1)
    A = rng.standard_normal((K, K))
    P = A @ A.T / K
    Jb = np.einsum("ijkl,kl->ij", full, P)
    Kb = np.einsum("ikjl,kl->ij", full, P)
    assert np.abs(orc.repulsion_matrix(P, eri) - (Jb - 0.5 * Kb)).max() < 1e-12
    assert np.abs(orc.coulomb_matrix(P, P, eri) - 2 * Jb).max() < 1e-12
    assert np.abs(orc.exchange_matrix(P, eri) - Kb).max() < 1e-12
