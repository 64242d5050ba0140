"""GPU parity: J/K (stored and direct, RHF and UHF, FP64 and mixed) against the
oracle's restatement of repulsion_matrix / coulomb_matrix / exchange_matrix
(include/linalg.hpp:100-296) on the golden unique ERIs."""
import numpy as np
import pytest

from tests import oracle_py as orc
from tests.conftest import golden

pytestmark = pytest.mark.gpu


def _dens(K, seed, n_occ=None):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((K, K))
    return (A @ A.T) / K


@pytest.mark.parametrize("name", ["h2o_sto3g.npz", "h2o_ccpvdz.npz"])
@pytest.mark.parametrize("mode", ["stored", "direct"])
def test_rhf_jk(gpu, name, mode):
    d = golden(name)
    K = int(d["K"])
    thr = float(d["threshold"])
    b = gpu.Basis.from_arrays(d)
    eri = orc.eri_unique(d, thr, exact=True)
    P = _dens(K, 1)
    fb = gpu.FockBuilder(b, mode=mode, threshold=thr)
    assert fb.info()["mode"] == mode
    J, Kx = fb.build(P)
    G = J - 0.5 * Kx
    Gref = orc.repulsion_matrix(P, eri)
    Jref = orc.coulomb_matrix(P, np.zeros_like(P), eri)
    scale = max(1.0, np.abs(Gref).max())
    print(f"{name} {mode}: max|dG|={np.abs(G - Gref).max():.2e} max|dJ|={np.abs(J - Jref).max():.2e}")
    assert np.abs(J - Jref).max() < 1e-11 * scale
    assert np.abs(G - Gref).max() < 1e-11 * scale


@pytest.mark.parametrize("mode", ["stored", "direct"])
def test_uhf_jk(gpu, mode):
    d = golden("h2o_ccpvdz.npz")
    K = int(d["K"])
    thr = float(d["threshold"])
    b = gpu.Basis.from_arrays(d)
    eri = orc.eri_unique(d, thr, exact=True)
    Pa, Pb = _dens(K, 2), _dens(K, 3)
    fb = gpu.FockBuilder(b, mode=mode, threshold=thr)
    J, Ka, Kb = fb.build(Pa, Pb)
    Jref = orc.coulomb_matrix(Pa, Pb, eri)
    Karef = orc.exchange_matrix(Pa, eri)
    Kbref = orc.exchange_matrix(Pb, eri)
    for got, ref in ((J, Jref), (Ka, Karef), (Kb, Kbref)):
        assert np.abs(got - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


def test_mixed_direct_close_to_fp64(gpu):
    d = golden("h2o_ccpvdz.npz")
    K = int(d["K"])
    b = gpu.Basis.from_arrays(d)
    P = _dens(K, 4)
    f64 = gpu.FockBuilder(b, mode="direct", precision="fp64", threshold=1e-12)
    mix = gpu.FockBuilder(b, mode="direct", precision="mixed", threshold=1e-12, mixed_split=1e-3)
    n64, n32 = mix.class_counts()
    assert n32.sum() > 0
    J0, K0 = f64.build(P)
    J1, K1 = mix.build(P)
    dJ, dK = np.abs(J1 - J0).max(), np.abs(K1 - K0).max()
    print(f"mixed vs fp64: fp32 quartets {n32.sum()} of {n32.sum() + n64.sum()}, max|dJ|={dJ:.2e} max|dK|={dK:.2e}")
    assert dJ < 1e-6 and dK < 1e-6


def test_mixed_stored_close_to_fp64(gpu):
    """Stored mode of the mixed form: the unique array holds FP32-evaluated
    small-bound quartets; J/K within the mixed form's error of FP64 and of the
    direct mixed build."""
    d = golden("h2o_ccpvdz.npz")
    K = int(d["K"])
    b = gpu.Basis.from_arrays(d)
    P = _dens(K, 4)
    f64 = gpu.FockBuilder(b, mode="stored", precision="fp64", threshold=1e-12)
    mst = gpu.FockBuilder(b, mode="stored", precision="mixed", threshold=1e-12, mixed_split=1e-3)
    mdi = gpu.FockBuilder(b, mode="direct", precision="mixed", threshold=1e-12, mixed_split=1e-3)
    assert mst.info()["mode"] == "stored"
    n64, n32 = mst.class_counts()
    assert n32.sum() > 0
    J0, K0 = f64.build(P)
    J1, K1 = mst.build(P)
    J2, K2 = mdi.build(P)
    print(f"stored mixed vs fp64: max|dJ|={np.abs(J1 - J0).max():.2e} max|dK|={np.abs(K1 - K0).max():.2e}; "
          f"vs direct mixed: {np.abs(J1 - J2).max():.2e} {np.abs(K1 - K2).max():.2e}")
    assert np.abs(J1 - J0).max() < 1e-6 and np.abs(K1 - K0).max() < 1e-6
    assert np.abs(J1 - J2).max() < 1e-6 and np.abs(K1 - K2).max() < 1e-6
    assert np.abs(J1 - J0).max() > 0.0   # the FP32 quartets were really evaluated in FP32


@pytest.mark.parametrize("uhf", [False, True])
def test_incremental_matches_full_builds(gpu, uhf):
    """Incremental direct builds (G(P_prev) + G(P - P_prev), the difference
    screened by |dP|) follow a sequence of SCF-like densities: every build
    matches a full build of the same density, across the periodic full
    rebuild (12 builds), and against the oracle's J/K on the golden ERIs.
    Screening is on: most late differences are far below the bounds."""
    d = golden("h2o_ccpvdz.npz")
    K = int(d["K"])
    thr = 1e-12
    b = gpu.Basis.from_arrays(d)
    eri = orc.eri_unique(d, float(d["threshold"]), exact=True)
    full = gpu.FockBuilder(b, mode="direct", threshold=thr)
    inc = gpu.FockBuilder(b, mode="direct", threshold=thr, incremental=True)
    rng = np.random.default_rng(11)
    Pa, Pb = _dens(K, 12), _dens(K, 13)
    worst = 0.0
    for it in range(12):
        step = 10.0 ** (-1 - it)   # shrinking updates, as an SCF converges
        Da, Db = rng.standard_normal((K, K)), rng.standard_normal((K, K))
        Pa = Pa + step * (Da + Da.T)
        Pb = Pb + step * (Db + Db.T)
        if uhf:
            r_full, r_inc = full.build(Pa, Pb), inc.build(Pa, Pb)
        else:
            r_full, r_inc = full.build(Pa), inc.build(Pa)
        for x, y in zip(r_full, r_inc):
            worst = max(worst, np.abs(x - y).max())
            assert np.abs(x - y).max() < 1e-10, it
    J = r_inc[0]
    Jref = orc.coulomb_matrix(Pa, Pb if uhf else np.zeros_like(Pa), eri)
    print(f"incremental vs full: max diff {worst:.2e}; J vs oracle {np.abs(J - Jref).max():.2e}")
    assert np.abs(J - Jref).max() < 1e-9


def test_direct_autotuner_keeps_results(gpu):
    """The work-unit autotuner of direct builds tries three slice targets on
    builds 2-4 and keeps the fastest; every build's J/K equals the first
    (summation order only) and the chosen target is one of the candidates."""
    from hartree_fock_cpp_amd import _native as N
    d = golden("h2o_ccpvdz.npz")
    K = int(d["K"])
    b = gpu.Basis.from_arrays(d)
    N.check(N.lib.hfx_set_autotune(1))
    try:
        fb = gpu.FockBuilder(b, mode="direct", threshold=1e-12)
    finally:
        N.check(N.lib.hfx_set_autotune(0))
    P = _dens(K, 21)
    J0, K0 = fb.build(P)
    for _ in range(5):
        J, Kx = fb.build(P)
        assert np.abs(J - J0).max() < 1e-12 and np.abs(Kx - K0).max() < 1e-12
    info = fb.autotune_info()
    print("autotune:", info)
    assert info["done"] and info["target"] in info["targets"]
    assert all(m > 0 for m in info["ms"])
