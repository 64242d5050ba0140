"""GPU parity of the small-plan launch shapes, against the oracle's exact-Boys
bounds and integrals.

Plans of few primitive quartets (the serving stream's molecules) run with
shapes the large benchmark molecules do not use: the Schwarz diagonal is cut
into primitive slices (the epilogue leaves (mu nu|mu nu), the slices add into a
zeroed array, the square roots follow), and unique-array plans of at most 2^25
primitive quartets slice their family-A quartets as well.  NaCl and HCl in
cc-pVDZ have s shells of 12 primitives, so single quartets carry 20,736
primitive quartets and are cut into many slices.
"""
import numpy as np
import pytest

from tests import oracle_py as orc

pytestmark = pytest.mark.gpu

ABS = 1e-12
CASES = [("NaCl", "cc-pVDZ", 1.3), ("HCl", "cc-pVDZ", 2.1), ("NaCl", "6-31G*", 0.9)]


@pytest.mark.parametrize("name,basis_set,scale", CASES)
def test_small_plan_bounds_and_eris(gpu, name, basis_set, scale):
    from hartree_fock_cpp_amd.basis import basis_construction
    from hartree_fock_cpp_amd.serve import EQUILIBRIUM, scan_molecule
    d = basis_construction(scan_molecule(name, scale * EQUILIBRIUM[name]), basis_set)
    b = gpu.Basis.from_arrays(d)
    K = b.K
    bd = gpu.schwarz_bounds(b)
    ex = orc.schwarz_bounds_exact(d)
    db = np.abs(bd - ex).max()
    thr = 1e-8
    eri = gpu.ERI_GPU(b, thr)
    b.close()
    # a seeded sample of all unique quartets plus every quartet of the most
    # contracted functions (the longest quartets of the plan)
    Q = eri.size
    rng = np.random.default_rng(11)
    q = rng.choice(Q, size=min(Q, 1500), replace=False)
    nprim = np.diff(np.asarray(d["poff"]))
    heavy = np.nonzero(nprim >= nprim.max() - 2)[0]
    hq = []
    for a in heavy:
        for bb in heavy[heavy <= a]:
            IJ = a * (a + 1) // 2 + bb
            for c in heavy:
                for e in heavy[heavy <= c]:
                    KL = c * (c + 1) // 2 + e
                    if KL <= IJ:
                        hq.append(IJ * (IJ + 1) // 2 + KL)
    q = np.unique(np.concatenate([q, np.asarray(hq, np.int64)[:1500]]))
    exact = orc.eri_quartets(d, thr, q, bounds=bd, exact=True)
    de = np.abs(eri[q] - exact).max()
    print(f"{name}/{basis_set}: K={K} max nprim={nprim.max()} max|d bounds|={db:.2e} "
          f"n={q.size} ({len(hq)} heavy) max|d eri|={de:.2e}", flush=True)
    assert db < ABS
    assert de < ABS
    # masked quartets are exact zeros
    IJ = ((np.sqrt(8.0 * q + 1.0) - 1.0) // 2).astype(np.int64)
    IJ -= (IJ * (IJ + 1) // 2 > q)
    IJ += ((IJ + 1) * (IJ + 2) // 2 <= q)
    KL = q - IJ * (IJ + 1) // 2
    masked = bd[IJ] * bd[KL] < thr
    assert np.all(eri[q][masked] == 0.0)
