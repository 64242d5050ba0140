"""ctypes view of the oracle (oracle/liboracle.so, the C restatement of the
reference's CPU path).  Test infrastructure: used by tests/, smoke() and the
cpu_baseline leg of bench.py only, always as the checker."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_lib = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
_lib.orc_boys_ref.restype = C.c_double
_lib.orc_boys_exact.restype = C.c_double
_lib.orc_basis_repulsion.restype = C.c_double
_lib.orc_coulomb_R.restype = C.c_double
_lib.orc_eri_index.restype = C.c_size_t
_vp = C.c_void_p


def _p(a):
    return a.ctypes.data_as(_vp)


def _bargs(d):
    K = int(len(d["poff"]) - 1)
    arrs = [np.ascontiguousarray(d["xyz"], np.float64), np.ascontiguousarray(d["lmn"], np.int32),
            np.ascontiguousarray(d["poff"], np.int32), np.ascontiguousarray(d["alpha"], np.float64),
            np.ascontiguousarray(d["coeff"], np.float64), np.ascontiguousarray(d["norm"], np.float64)]
    return K, arrs, [C.c_int(K)] + [_p(a) for a in arrs]


class boys_mode:
    """Context manager: 0 = reference Boys (integrals.hpp:223-266), 1 = exact Boys."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        _lib.orc_set_boys_mode(C.c_int(self.mode))

    def __exit__(self, *a):
        _lib.orc_set_boys_mode(C.c_int(0))


def boys_ref(m, T):
    return _lib.orc_boys_ref(C.c_int(m), C.c_double(T))


def boys_exact(m, T):
    return _lib.orc_boys_exact(C.c_int(m), C.c_double(T))


def schwarz_bounds(d):
    K, keep, a = _bargs(d)
    out = np.zeros(K * (K + 1) // 2)
    _lib.orc_schwarz_bounds(*a, _p(out))
    return out


def eri_quartets(d, threshold, qids, bounds=None, exact=False):
    K, keep, a = _bargs(d)
    bd = np.ascontiguousarray(bounds if bounds is not None else d["bounds"], np.float64)
    q = np.ascontiguousarray(qids, np.uint64)
    out = np.zeros(len(q))
    with boys_mode(1 if exact else 0):
        _lib.orc_eri_quartets(*a, _p(bd), C.c_double(threshold), _p(q), C.c_size_t(len(q)), _p(out))
    return out


def eri_unique(d, threshold, bounds=None, exact=False):
    K = int(len(d["poff"]) - 1)
    P = K * (K + 1) // 2
    Q = P * (P + 1) // 2
    K_, keep, a = _bargs(d)
    bd = np.ascontiguousarray(bounds if bounds is not None else d["bounds"], np.float64)
    out = np.zeros(Q)
    with boys_mode(1 if exact else 0):
        _lib.orc_eri_quartets(*a, _p(bd), C.c_double(threshold), None, C.c_size_t(Q), _p(out))
    return out


def basis_repulsion(d, i, j, k, l, exact=False):
    K, keep, a = _bargs(d)
    with boys_mode(1 if exact else 0):
        return _lib.orc_basis_repulsion(*a, C.c_int(i), C.c_int(j), C.c_int(k), C.c_int(l))


def one_electron(d):
    K, keep, a = _bargs(d)
    S = np.zeros(K * K)
    T = np.zeros(K * K)
    V = np.zeros(K * K)
    _lib.orc_overlap_matrix(*a, _p(S))
    _lib.orc_kinetic_matrix(*a, _p(T))
    Zs = np.ascontiguousarray(d["Zs"], np.int32)
    R = np.ascontiguousarray(np.asarray(d["R"], float).ravel())
    _lib.orc_attraction_matrix(*a, C.c_int(len(Zs)), _p(Zs), _p(R), _p(V))
    return S.reshape(K, K), T.reshape(K, K), V.reshape(K, K)


def repulsion_matrix(P, eri):
    K = P.shape[0]
    P = np.ascontiguousarray(P, np.float64)
    G = np.zeros((K, K))
    _lib.orc_repulsion_matrix(C.c_int(K), _p(P), _p(np.ascontiguousarray(eri)), _p(G))
    return G


def coulomb_matrix(P1, P2, eri):
    K = P1.shape[0]
    J = np.zeros((K, K))
    _lib.orc_coulomb_matrix(C.c_int(K), _p(np.ascontiguousarray(P1, np.float64)),
                            _p(np.ascontiguousarray(P2, np.float64)), _p(np.ascontiguousarray(eri)), _p(J))
    return J


def exchange_matrix(P, eri):
    K = P.shape[0]
    Kx = np.zeros((K, K))
    _lib.orc_exchange_matrix(C.c_int(K), _p(np.ascontiguousarray(P, np.float64)), _p(np.ascontiguousarray(eri)),
                             _p(Kx))
    return Kx


def eri_index(i, j, k, l):
    return _lib.orc_eri_index(C.c_int(i), C.c_int(j), C.c_int(k), C.c_int(l))


def expand(eri, K):
    """Full K^4 tensor from the unique array (numpy, small K only)."""
    P = K * (K + 1) // 2
    pi = np.zeros((K, K), np.int64)
    for i in range(K):
        for j in range(K):
            a, b = max(i, j), min(i, j)
            pi[i, j] = a * (a + 1) // 2 + b
    IJ = pi[:, :, None, None]
    KL = pi[None, None, :, :]
    hi = np.maximum(IJ, KL)
    lo = np.minimum(IJ, KL)
    return eri[hi * (hi + 1) // 2 + lo]


def schwarz_bounds_exact(d):
    with boys_mode(1):
        return schwarz_bounds(d)
