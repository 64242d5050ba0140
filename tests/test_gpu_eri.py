"""GPU parity: unique ERIs and Schwarz bounds from the native gfx950 library
against the reference's golden vectors and the oracle.

Parity contract (FP64 form):
  * every quartet whose reference value is not Boys-noise (|golden - exact| <
    1e-13, "eligible") matches the golden value to 1e-12 absolute;
  * every quartet matches the exact-Boys oracle to 1e-12 absolute;
  * the zero pattern (function-level Schwarz mask) equals the reference's,
    except for integrals that vanish by symmetry (|value| < 1e-15).
"""
import numpy as np
import pytest

from tests import oracle_py as orc
from tests.conftest import golden

pytestmark = pytest.mark.gpu

ABS = 1e-12


def _check_values(got, gold, exact, label):
    elig = np.abs(gold - exact) < 1e-13
    d_gold = np.abs(got - gold)[elig]
    d_exact = np.abs(got - exact)
    print(f"{label}: n={got.size} eligible={int(elig.sum())} max|d golden|={d_gold.max():.2e} "
          f"max|d exact|={d_exact.max():.2e}", flush=True)
    assert d_gold.max() < ABS
    assert d_exact.max() < ABS
    # mask parity: a zero in one and a non-zero in the other only for symmetry zeros
    flip = (got != 0) != (gold != 0)
    assert np.all(np.abs(exact[flip]) < 1e-15)


@pytest.mark.parametrize("name", ["h2_sto3g.npz", "h2o_sto3g.npz", "h2o_ccpvdz.npz"])
def test_schwarz_bounds(gpu, name):
    d = golden(name)
    b = gpu.Basis.from_arrays(d)
    got = gpu.schwarz_bounds(b)
    exact = orc.schwarz_bounds_exact(d)
    assert np.abs(got - exact).max() < 1e-12
    assert np.abs(got - d["bounds"]).max() < 1e-10


@pytest.mark.parametrize("name", ["h2_sto3g.npz", "h2o_sto3g.npz", "h2o_ccpvdz.npz"])
def test_eri_unique_full(gpu, name):
    d = golden(name)
    b = gpu.Basis.from_arrays(d)
    thr = float(d["threshold"])
    got = gpu.ERI_GPU(b, thr)
    exact = orc.eri_unique(d, thr, exact=True)
    _check_values(got, d["eri"], exact, name)


@pytest.mark.parametrize("name", ["o2_ccpvtz.npz", "benzene_ccpvdz.npz"])
def test_eri_unique_sampled(gpu, name):
    d = golden(name)
    b = gpu.Basis.from_arrays(d)
    thr = float(d["threshold"])
    got = gpu.ERI_GPU(b, thr)
    q = d["qids"].astype(np.int64)
    exact = orc.eri_quartets(d, thr, q, exact=True)
    _check_values(got[q], d["eri_sample"], exact, name)
    # every entry finite (the zero pattern over the whole array, and its
    # non-zeros, are checked on the device in test_gpu_configs)
    assert np.all(np.isfinite(got))


@pytest.mark.parametrize("mode", [1, 2])
def test_class_timing_keeps_values(gpu, mode):
    """Class timing (1: events on each list's own stream, lists still on four
    streams; 2: lists serialized on the launch stream) reports a positive time
    for every launched class and leaves the unique array unchanged."""
    from hartree_fock_cpp_amd import _native as N
    d = golden("h2o_ccpvdz.npz")
    thr = float(d["threshold"])
    b = gpu.Basis.from_arrays(d)
    ref = gpu.ERI_GPU(b, thr)
    N.check(N.lib.hfx_set_class_timing(mode))
    try:
        got = gpu.ERI_GPU(b, thr)
        ms = np.zeros(110, np.float32)
        N.check(N.lib.hfx_class_times(N.ptr(ms)))
    finally:
        N.check(N.lib.hfx_set_class_timing(0))
    nq = np.zeros(110, np.int64)
    npq = np.zeros(110, np.int64)
    N.check(N.lib.hfx_plan_stats(b.handle(), N.C.c_double(thr), 0, N.C.c_double(0.0), N.ptr(nq), N.ptr(npq)))
    print(f"mode {mode}: {int((ms > 0).sum())} timed lists, {int((nq > 0).sum())} launched, sum {ms.sum():.3f} ms")
    assert np.all(ms[nq > 0] > 0) and np.all(ms[nq == 0] == 0)
    assert np.abs(got - ref).max() < 1e-14


def test_eri_unique_host_copy_matches_device_array(gpu):
    """The host-buffer entry point (the drop-in's calling form) returns the
    device-resident array when the copy runs in pinned chunks: benzene /
    cc-pVDZ, Q = 26,357,430 (211 MB, seven chunks, the last partial).  The two
    calls are two builds: quartets cut into three or more primitive slices add
    their partial blocks by atomics in an order that varies between builds,
    so values agree to rounding (1e-15 relative) and the zero pattern exactly;
    a misplaced chunk would differ at O(1)."""
    import ctypes as C

    import torch

    from hartree_fock_cpp_amd import _native as N
    from hartree_fock_cpp_amd import workloads as W
    from hartree_fock_cpp_amd.basis import basis_construction

    b = gpu.Basis.from_arrays(basis_construction(W.benzene(), "cc-pVDZ"))
    Q = b.n_quartets
    print(f"benzene/cc-pVDZ Q={Q}", flush=True)
    dev = torch.empty(Q, dtype=torch.float64, device="cuda")
    st = torch.cuda.Stream()
    N.check(N.lib.hfx_eri_unique_device(b.handle(), C.c_double(1e-8), C.c_void_p(dev.data_ptr()), C.c_size_t(Q),
                                        C.c_void_p(st.cuda_stream)))
    torch.cuda.synchronize()
    ref = dev.cpu().numpy()
    del dev
    got = gpu.ERI_GPU(b, 1e-8)
    assert got.shape == ref.shape
    assert np.array_equal(got == 0, ref == 0)
    assert np.abs(got - ref).max() <= 1e-15 * np.abs(ref).max()
    assert np.count_nonzero(got) > Q // 2
