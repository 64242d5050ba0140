"""The allocation-free J/K entry point (hfx_coulomb_exchange_noalloc) against
the allocating one it varies (hfx_coulomb_exchange, which replaces the
reference's coulomb_matrix / exchange_matrix, include/linalg.hpp:215-296),
through the C ABI on device buffers: RHF and UHF, two calls in a row with the
second on a smaller basis (the shared scratch keeps the first call's larger
contents), bit-level agreement up to summation order (1e-12 relative), and
the error for a UHF basis beyond the scratch maximum."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def _dens(K, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((K, K))
    return (A @ A.T) / K


def _case(gpu, name, uhf, seed):
    import torch

    from hartree_fock_cpp_amd import _native as N
    d = golden(name)
    thr = float(d["threshold"])
    b = gpu.Basis.from_arrays(d)
    K, Q = b.K, b.n_quartets
    print(f"noalloc: {name} K={K} {'UHF' if uhf else 'RHF'}", flush=True)
    eri = torch.empty(Q, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream()
    N.check(N.lib.hfx_eri_unique_device(b.handle(), C.c_double(thr), C.c_void_p(eri.data_ptr()), C.c_size_t(Q),
                                        C.c_void_p(st.cuda_stream)))
    Pa, Pb = _dens(K, seed), _dens(K, seed + 1)
    # the allocating entry point, host matrices
    J0, Ka0, Kb0 = (np.zeros((K, K)) for _ in range(3))
    torch.cuda.synchronize()
    N.check(N.lib.hfx_coulomb_exchange(b.handle(), C.c_void_p(eri.data_ptr()), N.ptr(Pa), N.ptr(Pb) if uhf else None,
                                       N.ptr(J0), N.ptr(Ka0), N.ptr(Kb0) if uhf else None))
    # the allocation-free one, device matrices (outputs pre-filled with garbage)
    dPa = torch.from_numpy(Pa).cuda()
    dPb = torch.from_numpy(Pb).cuda()
    dJ, dKa, dKb = (torch.full((K, K), 7.0, dtype=torch.float64, device="cuda") for _ in range(3))
    p = lambda t: C.c_void_p(t.data_ptr())
    N.check(N.lib.hfx_coulomb_exchange_noalloc(b.handle(), p(eri), p(dPa), p(dPb) if uhf else None, p(dJ), p(dKa),
                                               p(dKb) if uhf else None, C.c_void_p(st.cuda_stream)))
    torch.cuda.synchronize()
    got = [dJ.cpu().numpy(), dKa.cpu().numpy()] + ([dKb.cpu().numpy()] if uhf else [])
    ref = [J0, Ka0] + ([Kb0] if uhf else [])
    for lab, x, y in zip(("J", "Ka", "Kb"), got, ref):
        err = np.abs(x - y).max()
        print(f"  {lab}: max|noalloc - alloc| = {err:.2e}", flush=True)
        assert err <= 1e-12 * max(1.0, np.abs(y).max())
    b.close()


@pytest.mark.parametrize("uhf", [False, True])
def test_noalloc_matches_allocating_entry_point_twice(gpu, uhf):
    _case(gpu, "h2o_ccpvdz.npz", uhf, 3)    # K = 25
    _case(gpu, "h2o_sto3g.npz", uhf, 5)     # K = 7, after the larger call


def test_noalloc_rejects_uhf_beyond_scratch(gpu):
    import torch

    from hartree_fock_cpp_amd import _native as N
    from hartree_fock_cpp_amd import workloads as W
    from hartree_fock_cpp_amd.basis import basis_construction
    # K = 530 > 512: rejected before any device work (no unique array needed)
    atoms = W.water_tetramer() + [(s, (x + 20.0, y, z)) for s, (x, y, z) in W.water_tetramer()]
    b = gpu.Basis.from_arrays(basis_construction(atoms, "cc-pVTZ"))
    assert b.K > 512
    x = torch.zeros(1, dtype=torch.float64, device="cuda")
    p = C.c_void_p(x.data_ptr())
    rc = N.lib.hfx_coulomb_exchange_noalloc(b.handle(), p, p, p, p, p, p, None)
    msg = N.lib.hfx_last_error().decode()
    print("rejected:", rc, msg)
    assert rc == -4 and "exceeds the scratch maximum" in msg
    b.close()
