"""Numpy restatement of the reference's SCF drivers (test infrastructure only).

Follows KLYYCYNN/hartree_fock_cpp include/scf.hpp:
  rhf_diis  scf.hpp:205-419    uhf_diis  scf.hpp:423-682
  Hcore scf.hpp:11-25, energy scf.hpp:28-39, update_damping scf.hpp:81-104,
  rhf/uhf_final_check scf.hpp:107-185
and the numerics of include/linalg.hpp (Xmatrix 132-147, EigProblem 150-170,
update_Pmatrix 174-192, update_density_uhf 195-212, error_matrix 355-379,
DIIS 382-511).  The Fock build is a callable (the oracle's C restatement of
repulsion_matrix / coulomb_matrix / exchange_matrix, or einsum on small
tensors), so the loop runs on the reference's golden integrals.

Eigen-solver and LU come from numpy (the reference uses Eigen, which is not
available here; parity at that boundary is pinned through converged energies,
which do not depend on eigenvector signs or the DIIS path).
"""
from collections import deque

import numpy as np


def _xmatrix(S):
    w, U = np.linalg.eigh(S)
    return U @ np.diag(1.0 / np.sqrt(w)) @ U.T


def _eig(F, X):
    Fp = X.T @ F @ X
    e, Cp = np.linalg.eigh(Fp)
    return X @ Cp, e


def _diis(Fh, Rh):
    n = len(Fh)
    B = np.zeros((n + 1, n + 1))
    for i in range(n):
        for j in range(n):
            B[i, j] = np.dot(Rh[i], Rh[j])
    B[:n, n] = -1.0
    B[n, :n] = 1.0
    rhs = np.zeros(n + 1)
    rhs[n] = 1.0
    c = np.linalg.lstsq(B, rhs, rcond=None)[0]
    return sum(c[i] * Fh[i] for i in range(n))


def rhf(S, H, E_nuc, n_electrons, G_of, max_it=200, eps=1e-10, damping=0.3, P0=None):
    """G_of(P) -> G = J - K/2.  Returns (E_tot, P, C, e, n_it, converged)."""
    K = S.shape[0]
    X = _xmatrix(S)
    P = np.zeros((K, K)) if P0 is None else P0.copy()
    nocc = n_electrons // 2
    Fh, Rh = deque(maxlen=8), deque(maxlen=8)
    C = e = None
    for it in range(1, max_it + 1):
        F = H + G_of(P)
        R = F @ P @ S - S @ P @ F
        Fh.append(F.ravel().copy())
        Rh.append(R.ravel().copy())
        if it >= 4:
            F = _diis(list(Fh), list(Rh)).reshape(K, K)
        C, e = _eig(F, X)
        Pn = 2.0 * C[:, :nocc] @ C[:, :nocc].T
        dP = np.sqrt(np.mean((Pn - P) ** 2))
        if dP < eps and np.linalg.norm(R) < eps:
            Fc = H + G_of(Pn)
            Rc = Fc @ Pn @ S - S @ Pn @ Fc
            if np.linalg.norm(Rc) < eps:
                E = 0.5 * np.sum(Pn * (H + Fc)) + E_nuc
                return E, Pn, C, e, it, True
        P = Pn if it >= 4 else damping * Pn + (1 - damping) * P
    Fc = H + G_of(P)
    return 0.5 * np.sum(P * (H + Fc)) + E_nuc, P, C, e, max_it, False


def uhf(S, H, E_nuc, na, nb, JK_of, max_it=300, eps=1e-10, damping=0.3):
    """JK_of(Pa, Pb) -> (J, Ka, Kb).  Returns (E_tot, Pa, Pb, n_it, converged)."""
    K = S.shape[0]
    X = _xmatrix(S)
    Pa = np.zeros((K, K))
    Pb = np.zeros((K, K))
    Fh, Rh = deque(maxlen=8), deque(maxlen=8)
    for it in range(1, max_it + 1):
        J, Ka, Kb = JK_of(Pa, Pb)
        Fa, Fb = H + J - Ka, H + J - Kb
        Ra = Fa @ Pa @ S - S @ Pa @ Fa
        Rb = Fb @ Pb @ S - S @ Pb @ Fb
        Fh.append(np.concatenate([Fa.ravel(), Fb.ravel()]))
        Rh.append(np.concatenate([Ra.ravel(), Rb.ravel()]))
        if it >= 4:
            Fd = _diis(list(Fh), list(Rh))
            Fa, Fb = Fd[:K * K].reshape(K, K), Fd[K * K:].reshape(K, K)
        Ca, _ = _eig(Fa, X)
        Cb, _ = _eig(Fb, X)
        Pan = Ca[:, :na] @ Ca[:, :na].T
        Pbn = Cb[:, :nb] @ Cb[:, :nb].T
        da = np.sqrt(np.mean((Pan - Pa) ** 2))
        db = np.sqrt(np.mean((Pbn - Pb) ** 2))
        err = np.sqrt(np.linalg.norm(Ra) ** 2 + np.linalg.norm(Rb) ** 2)
        if da < eps and db < eps and err < eps:
            J, Ka, Kb = JK_of(Pan, Pbn)
            Fa, Fb = H + J - Ka, H + J - Kb
            Ra = Fa @ Pan @ S - S @ Pan @ Fa
            Rb = Fb @ Pbn @ S - S @ Pbn @ Fb
            if np.sqrt(np.linalg.norm(Ra) ** 2 + np.linalg.norm(Rb) ** 2) < eps:
                E = 0.5 * np.sum(Pan * (H + Fa)) + 0.5 * np.sum(Pbn * (H + Fb)) + E_nuc
                return E, Pan, Pbn, it, True
        if it >= 4:
            Pa, Pb = Pan, Pbn
        else:
            Pa = damping * Pan + (1 - damping) * Pa
            Pb = damping * Pbn + (1 - damping) * Pb
    return None, Pa, Pb, max_it, False
