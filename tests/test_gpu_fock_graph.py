"""Graph replay of the Fock build (hfx_set_fock_graphs): every replay, with
other device work and synchronisations between replays and densities that
change between calls, equals the eager build of the same density.

Tolerance 1e-12 (J/K atomics reorder FP64 sums between runs)."""
import numpy as np
import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def _dens(K, seed):
    A = np.random.default_rng(seed).standard_normal((K, K))
    return A @ A.T / K


@pytest.mark.parametrize("mode", ["stored", "direct"])
@pytest.mark.parametrize("uhf", [False, True])
def test_graph_replays_match_eager(gpu, mode, uhf):
    import torch
    import hartree_fock_cpp_amd as hfx
    from hartree_fock_cpp_amd import _native as N
    d = golden("h2o_ccpvdz.npz")
    b = hfx.Basis.from_arrays(d)
    K = b.K
    dens = [(_dens(K, s), _dens(K, s + 10) if uhf else None) for s in range(3)]
    N.check(N.lib.hfx_set_fock_graphs(0))
    try:
        fb = hfx.FockBuilder(b, mode=mode, threshold=1e-10)
        eager = [fb.build(Pa, Pb) for Pa, Pb in dens]
        fb.close()
        N.check(N.lib.hfx_set_fock_graphs(1))
        fb = hfx.FockBuilder(b, mode=mode, threshold=1e-10)
        x = torch.ones(1 << 20, device="cuda", dtype=torch.float64)
        for rep in range(3):
            for (Pa, Pb), ref in zip(dens, eager):
                got = fb.build(Pa, Pb)
                x.mul_(1.0000001).add_(1.0)            # unrelated eager work between replays
                torch.cuda.synchronize()
                for g, r in zip(got, ref):
                    err = np.abs(g - r).max()
                    assert err < 1e-12, (mode, uhf, rep, err)
        fb.close()
    finally:
        N.check(N.lib.hfx_set_fock_graphs(1))
