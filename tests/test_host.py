"""CPU tests of the native host code (basis construction, one-electron
integrals, eigensolver) through the C ABI of include/hfx_scf.h."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import golden

hfx = pytest.importorskip("hartree_fock_cpp_amd")
from hartree_fock_cpp_amd import _native as N  # noqa: E402
from hartree_fock_cpp_amd import scf as S  # noqa: E402
from hartree_fock_cpp_amd.basis import basis_construction  # noqa: E402


def _config(d):
    return [(str(s), tuple(float(x) for x in r)) for s, r in zip(d["symbols"], d["R"])]


@pytest.mark.parametrize("name", ["h2o_sto3g.npz", "h2o_ccpvdz.npz", "o2_ccpvtz.npz", "benzene_ccpvdz.npz"])
def test_native_basis_matches_reference(name):
    d = golden(name)
    cfg = _config(d)
    K = S.basis_size(cfg, str(d["basis_set"]))
    assert K == int(d["K"])
    npr = len(d["alpha"])
    xyz, lmn, poff = np.zeros(3 * K), np.zeros(3 * K, np.int32), np.zeros(K + 1, np.int32)
    al, co, no = np.zeros(npr), np.zeros(npr), np.zeros(npr)
    n = C.c_int()
    x = np.ascontiguousarray(np.array([p for _, p in cfg], float).ravel())
    N.check(N.lib.hfx_basis_build(len(cfg), S._syms(cfg), N.ptr(x), str(d["basis_set"]).encode(),
                                  S.BASIS_DIR.encode(), K, N.ptr(xyz), N.ptr(lmn), N.ptr(poff), C.byref(n),
                                  N.ptr(al), N.ptr(co), N.ptr(no), npr))
    for got, key in ((xyz, "xyz"), (lmn, "lmn"), (poff, "poff"), (al, "alpha"), (co, "coeff"), (no, "norm")):
        assert np.array_equal(got, d[key]), key


def test_python_basis_matches_reference():
    d = golden("o2_ccpvtz.npz")
    b = basis_construction(_config(d), "cc-pVTZ")
    for key in ("xyz", "lmn", "poff", "alpha", "coeff", "norm"):
        assert np.array_equal(b[key], d[key]), key


@pytest.mark.parametrize("name", ["h2_sto3g.npz", "h2o_sto3g.npz", "h2o_ccpvdz.npz", "o2_ccpvtz.npz"])
def test_one_electron_matches_reference(name):
    d = golden(name)
    Sm, Tm, Vm = S.one_electron(_config(d), str(d["basis_set"]))
    for got, key in ((Sm, "S"), (Tm, "T"), (Vm, "V")):
        ref = d[key]
        assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), key


def test_errors_are_reported_not_fatal():
    with pytest.raises(N.HfxError, match="unknown element"):
        S.basis_size([("Xx", (0.0, 0.0, 0.0))], "STO-3G")
    with pytest.raises(N.HfxError, match="does not exist"):
        S.basis_size([("H", (0.0, 0.0, 0.0))], "no-such-basis")


@pytest.mark.parametrize("n", [1, 2, 7, 40, 150])
def test_host_eigensolver(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = A + A.T
    A[0, :] *= 1e-3   # some spread of scales
    A[:, 0] *= 1e-3
    w, V = np.zeros(n), np.zeros((n, n))
    N.check(N.lib.hfx_host_sym_eig(n, N.ptr(np.ascontiguousarray(A)), N.ptr(w), N.ptr(V)))
    wr = np.linalg.eigvalsh(A)
    assert np.abs(w - wr).max() < 1e-12 * max(1, np.abs(wr).max())
    assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12
    assert np.abs(A @ V - V * w).max() < 1e-11 * max(1, np.abs(wr).max())


def test_host_eigensolver_degenerate():
    n = 12
    Q, _ = np.linalg.qr(np.random.default_rng(0).standard_normal((n, n)))
    d = np.array([1.0] * 4 + [2.0] * 4 + [3.0] * 4)
    A = Q @ np.diag(d) @ Q.T
    w, V = np.zeros(n), np.zeros((n, n))
    N.check(N.lib.hfx_host_sym_eig(n, N.ptr(np.ascontiguousarray(A)), N.ptr(w), N.ptr(V)))
    assert np.abs(w - d).max() < 1e-13
    assert np.abs(A @ V - V * w).max() < 1e-12


def _lowest(A, m):
    n = A.shape[0]
    w, V = np.zeros(n), np.zeros((n, n))
    N.check(N.lib.hfx_host_sym_eig_lowest(n, N.ptr(np.ascontiguousarray(A)), m, N.ptr(w), N.ptr(V)))
    return w, V


@pytest.mark.parametrize("n,m", [(1, 1), (3, 1), (40, 7), (150, 25), (150, 150)])
def test_host_eigensolver_lowest(n, m):
    """All eigenvalues and the m lowest eigenvectors (what an SCF iteration
    uses): eigenvalues as numpy's LAPACK, residual and orthonormality of the
    m vectors, zero columns beyond m."""
    rng = np.random.default_rng(n + m)
    A = rng.standard_normal((n, n))
    A = A + A.T
    w, V = _lowest(A, m)
    wr = np.linalg.eigvalsh(A)
    s = max(1, np.abs(wr).max())
    Vm = V[:, :m]
    assert np.abs(w - wr).max() < 1e-12 * s
    assert np.abs(Vm.T @ Vm - np.eye(m)).max() < 1e-12
    assert np.abs(A @ Vm - Vm * w[:m]).max() < 1e-11 * s
    assert np.all(V[:, m:] == 0.0)


@pytest.mark.parametrize("kind", ["repeated", "block", "identity"])
def test_host_eigensolver_lowest_clusters(kind):
    """Degenerate and clustered spectra (orbital shells of symmetric
    molecules): inverse iteration with Gram-Schmidt inside clusters still
    returns an orthonormal eigenbasis of every degenerate subspace it cuts."""
    rng = np.random.default_rng(3)
    if kind == "repeated":
        Q, _ = np.linalg.qr(rng.standard_normal((20, 20)))
        A = Q @ np.diag(np.r_[np.ones(5), 2 * np.ones(5), np.arange(3, 13)]) @ Q.T
    elif kind == "block":
        B = rng.standard_normal((10, 10))
        A = np.kron(np.eye(4), B + B.T)
    else:
        A = np.eye(12)
    n = A.shape[0]
    for m in (1, 3, n // 2):
        w, V = _lowest(A, m)
        Vm = V[:, :m]
        assert np.abs(w - np.linalg.eigvalsh(A)).max() < 1e-12 * max(1, np.abs(w).max())
        assert np.abs(Vm.T @ Vm - np.eye(m)).max() < 1e-12
        assert np.abs(A @ Vm - Vm * w[:m]).max() < 1e-12 * max(1, np.abs(w).max())


@pytest.mark.parametrize("kind", ["tridiagonal", "nearly_split", "split", "graded", "zero_diagonal"])
def test_host_eigenvalues_root_free_qr(kind):
    """The eigenvalue sweep (root-free QR on the tridiagonal form) on inputs
    that stress its deflation: already-tridiagonal matrices, off-diagonals at
    1e-9 of the diagonal, exact zeros (split blocks), magnitudes spread over
    six decades, and a zero diagonal (deflation next to exact zeros).
    Eigenvalues as numpy's LAPACK eigvalsh."""
    rng = np.random.default_rng(11)
    n = 60
    d = rng.standard_normal(n)
    e = rng.standard_normal(n - 1)
    if kind == "nearly_split":
        e *= 1e-9
    elif kind == "split":
        e[::4] = 0.0
    elif kind == "graded":
        d *= np.logspace(0, 6, n)
        e *= np.logspace(0, 6, n - 1) * 0.1
    elif kind == "zero_diagonal":
        d[:] = 0.0
    A = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    if kind == "tridiagonal":
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        A = Q @ A @ Q.T
        A = 0.5 * (A + A.T)
    w, V = _lowest(A, 5)
    wr = np.linalg.eigvalsh(A)
    s = max(1.0, np.abs(wr).max())
    assert np.abs(w - wr).max() < 1e-12 * s
    Vm = V[:, :5]
    assert np.abs(A @ Vm - Vm * w[:5]).max() < 1e-11 * s
