"""One-electron integrals on the GPU (hfx_one_electron_device) against the
reference's own S, T, V (golden fixtures from the compiled reference) and the
host implementation, for every pair class up to (ff| (O2/cc-pVTZ).

Tolerance: 1e-12 absolute (both sides FP64; the reference's Boys function
enters V, so V is compared to 1e-10 where the golden fixture holds it)."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def _device(d):
    from hartree_fock_cpp_amd import _native as N
    K = int(len(d["poff"]) - 1)
    arr = [np.ascontiguousarray(d["xyz"], np.float64), np.ascontiguousarray(d["lmn"], np.int32),
           np.ascontiguousarray(d["poff"], np.int32), np.ascontiguousarray(d["alpha"], np.float64),
           np.ascontiguousarray(d["coeff"], np.float64), np.ascontiguousarray(d["norm"], np.float64)]
    Z = np.ascontiguousarray(d["Zs"], np.float64)
    R = np.ascontiguousarray(d["R"], np.float64).ravel()
    S, T, V = (np.zeros((K, K)) for _ in range(3))
    N.check(N.lib.hfx_one_electron_device(C.c_int(K), *[N.ptr(a) for a in arr], C.c_int(len(Z)), N.ptr(Z), N.ptr(R),
                                          N.ptr(S), N.ptr(T), N.ptr(V)))
    return S, T, V


@pytest.mark.parametrize("name", ["h2o_sto3g.npz", "h2o_ccpvdz.npz", "benzene_ccpvdz.npz", "o2_ccpvtz.npz"])
def test_one_electron_device_matches_reference(gpu, name):
    d = golden(name)
    S, T, V = _device(d)
    for key, M in (("S", S), ("T", T), ("V", V)):
        if key in d:
            err = np.abs(M - d[key]).max()
            print(name, key, err)
            assert err < (1e-10 if key == "V" else 1e-12), (key, err)
    assert np.abs(S - S.T).max() == 0.0


def test_one_electron_device_matches_host(gpu):
    from hartree_fock_cpp_amd.scf import one_electron
    from hartree_fock_cpp_amd import workloads as W
    from hartree_fock_cpp_amd.basis import basis_construction
    atoms = W.water_tetramer()
    d = basis_construction(atoms, "cc-pVTZ")
    S, T, V = _device(d)
    Sh, Th, Vh = one_electron(atoms, "cc-pVTZ")
    print(np.abs(S - Sh).max(), np.abs(T - Th).max(), np.abs(V - Vh).max())
    assert np.abs(S - Sh).max() < 1e-13 and np.abs(T - Th).max() < 1e-12 and np.abs(V - Vh).max() < 1e-11
