"""GPU parity on the large configurations: n-octane / cc-pVDZ (K = 210, jitter
seed 1) and (H2O)4 / cc-pVTZ (K = 260, jitter seed 4, f functions), both at
Schwarz threshold 1e-12, plus the sampled molecules of test_gpu_eri
(O2 / cc-pVTZ, benzene / cc-pVDZ) for the whole-array layout checks.

  * unique ERIs (the array ERI_GPU returns, cuda/src/eri.cu:20-89): a
    stratified sample of 200 function quartets per angular class, computed by
    the reference's own basis_repulsion (include/integrals.hpp:701-729,
    compiled in place), must match to 1e-12 absolute where the reference's
    Boys function is not noise, and the exact-Boys oracle to 1e-12 everywhere;
  * the whole Q-long array, on the device: exactly 0.0 wherever the
    function-level Schwarz product b_IJ b_KL is below the threshold
    (integrals.hpp:771-776), and elsewhere zero only where the oracle says
    the integral vanishes by symmetry, so a write to a wrong eri_index slot
    (indexing.hpp:42-52) cannot hide outside the sample;
  * J / K of the direct build (ERIs contracted in the kernel epilogue, no
    unique array) equal J / K streamed from the stored unique array to 1e-11
    relative: every stored integral enters, at its eri_index position;
  * mixed versus FP64 direct SCF at split 1e-5: |dE| <= 1 uHa after the FP64
    polish, with max|dF| and max|dD| printed and bounded.
"""
import ctypes as C

import numpy as np
import pytest

from tests import oracle_py as orc
from tests.conftest import golden

pytestmark = pytest.mark.gpu

ABS = 1e-12
CONFIGS = ["octane_ccpvdz.npz", "water4_ccpvtz.npz"]


def _config(d):
    return [(str(s), tuple(float(x) for x in r)) for s, r in zip(d["symbols"], d["R"])]


def _device_unique(gpu, b, thr):
    """The unique array in HBM (torch float64 tensor) from hfx_eri_unique_device."""
    import torch

    from hartree_fock_cpp_amd import _native as N
    Q = b.n_quartets
    out = torch.empty(Q, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream()
    N.check(N.lib.hfx_eri_unique_device(b.handle(), C.c_double(thr), C.c_void_p(out.data_ptr()), C.c_size_t(Q),
                                        C.c_void_p(st.cuda_stream)))
    torch.cuda.synchronize()
    return out


def _mask_checks(eri, bounds, thr, label):
    """Walks the device array in chunks.  Returns (n_masked, n_written_where
    masked, ids of entries that are exactly zero where the mask keeps them)."""
    import torch
    bd = torch.from_numpy(bounds).cuda()
    Q = eri.numel()
    n_masked = n_bad = 0
    zero_kept = []
    step = 1 << 25
    for lo in range(0, Q, step):
        q = torch.arange(lo, min(Q, lo + step), dtype=torch.int64, device="cuda")
        IJ = torch.floor((torch.sqrt(8.0 * q.double() + 1.0) - 1.0) / 2.0).long()
        IJ = IJ - (IJ * (IJ + 1) // 2 > q).long()
        IJ = IJ + ((IJ + 1) * (IJ + 2) // 2 <= q).long()
        KL = q - IJ * (IJ + 1) // 2
        keep = bd[IJ] * bd[KL] >= thr
        v = eri[lo:lo + q.numel()]
        n_bad += int(((~keep) & (v != 0)).sum())
        n_masked += int((~keep).sum())
        zero_kept.append(q[keep & (v == 0)].cpu().numpy())
    zero_kept = np.concatenate(zero_kept)
    print(f"{label}: Q={Q} masked={n_masked} written-where-masked={n_bad} zero-where-kept={len(zero_kept)}",
          flush=True)
    return n_masked, n_bad, zero_kept


@pytest.mark.parametrize("name", CONFIGS)
def test_sampled_eris_match_reference(gpu, name):
    d = golden(name)
    thr = float(d["threshold"])
    b = gpu.Basis.from_arrays(d)
    print(f"{name}: K={b.K} Q={b.n_quartets} threshold={thr:g}", flush=True)
    eri = _device_unique(gpu, b, thr)
    q = d["qids"].astype(np.int64)
    import torch
    got = eri[torch.from_numpy(q).cuda()].cpu().numpy()
    gold = d["eri_sample"]
    exact = orc.eri_quartets(d, thr, q, exact=True)
    elig = np.abs(gold - exact) < 1e-13
    d_gold = np.abs(got - gold)[elig].max()
    d_exact = np.abs(got - exact).max()
    print(f"{name}: n={len(q)} eligible={int(elig.sum())} max|d golden|={d_gold:.2e} max|d exact|={d_exact:.2e}",
          flush=True)
    assert elig.sum() > 0.9 * len(q)
    assert d_gold < ABS and d_exact < ABS
    flip = (got != 0) != (gold != 0)
    assert np.all(np.abs(exact[flip]) < 1e-15)


@pytest.mark.parametrize("name", ["o2_ccpvtz.npz", "benzene_ccpvdz.npz"] + CONFIGS)
def test_whole_array_zero_pattern(gpu, name):
    d = golden(name)
    thr = float(d["threshold"])
    b = gpu.Basis.from_arrays(d)
    eri = _device_unique(gpu, b, thr)
    bounds = gpu.schwarz_bounds(b)
    # the library's bounds against the reference's (Boys noise only)
    assert np.abs(bounds - d["bounds"]).max() < 1e-10
    n_masked, n_bad, zero_kept = _mask_checks(eri, bounds, thr, name)
    assert n_bad == 0
    # a kept entry may be exactly zero only where the integral vanishes by
    # symmetry (point group, or the local symmetry of one- and two-centre
    # integrals, which the jittered molecules keep): the oracle's exact-Boys
    # value there is zero too (seeded sample of up to 400 such entries)
    if name in CONFIGS:   # jittered: no point-group zeros, only the few local ones
        assert len(zero_kept) < 0.01 * (eri.numel() - n_masked)
    if len(zero_kept):
        pick = np.random.default_rng(5).choice(len(zero_kept), min(400, len(zero_kept)), replace=False)
        ex = orc.eri_quartets(d, thr, zero_kept[pick], bounds=bounds, exact=True)
        print(f"{name}: max |oracle| at zero-kept entries {np.abs(ex).max():.1e}", flush=True)
        assert np.abs(ex).max() < 1e-15


def _dens(K, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((K, K))
    return (A @ A.T) / K


@pytest.mark.parametrize("name", CONFIGS)
@pytest.mark.parametrize("uhf", [False, True])
def test_direct_jk_matches_stored_jk(gpu, name, uhf):
    d = golden(name)
    thr = float(d["threshold"])
    K = int(d["K"])
    b = gpu.Basis.from_arrays(d)
    Pa, Pb = _dens(K, 31), _dens(K, 32)
    st = gpu.FockBuilder(b, mode="stored", threshold=thr)
    di = gpu.FockBuilder(b, mode="direct", threshold=thr)
    assert st.info()["mode"] == "stored" and di.info()["mode"] == "direct"
    r_st = st.build(Pa, Pb) if uhf else st.build(Pa)
    r_di = di.build(Pa, Pb) if uhf else di.build(Pa)
    st.close()
    di.close()
    for lab, x, y in zip(("J", "Ka", "Kb"), r_st, r_di):
        scale = max(1.0, np.abs(x).max())
        err = np.abs(x - y).max()
        print(f"{name} {'UHF' if uhf else 'RHF'} {lab}: max|direct - stored|={err:.2e} (scale {scale:.1f})", flush=True)
        assert err < 1e-11 * scale


@pytest.mark.parametrize("name", CONFIGS)
def test_mixed_scf_within_1uha_of_fp64(gpu, name):
    from hartree_fock_cpp_amd.scf import rhf_diis
    d = golden(name)
    cfg = _config(d)
    bset = str(d["basis_set"])
    thr, split = 1e-12, 1e-5
    r64 = rhf_diis(cfg, bset, eri_mode="direct", precision="fp64", screening_threshold=thr)
    rmx = rhf_diis(cfg, bset, eri_mode="direct", precision="mixed", screening_threshold=thr, mixed_split=split,
                   polish=True)
    assert r64["converged"] and rmx["converged"]
    b = gpu.Basis.from_arrays(d)
    P = r64["density"]
    G = {}
    for prec in ("fp64", "mixed"):
        fb = gpu.FockBuilder(b, mode="direct", precision=prec, threshold=thr, mixed_split=split)
        J, Kx = fb.build(P)
        G[prec] = J - 0.5 * Kx
        fb.close()
    dE = rmx["E_tot"] - r64["E_tot"]
    dE0 = rmx["E_before_polish"] - r64["E_tot"]
    dF = float(np.abs(G["mixed"] - G["fp64"]).max())
    dD = float(np.abs(rmx["density"] - P).max())
    print(f"{name}: E64={r64['E_tot']:.10f} dE={dE:.2e} (before polish {dE0:.2e}) max|dF|={dF:.2e} "
          f"max|dD|={dD:.2e} it {r64['iterations']}/{rmx['iterations']}+{rmx['polish_iterations']}", flush=True)
    assert abs(dE) <= 1e-6
    assert abs(dE0) <= 1e-6
    assert dF < 1e-6 and dD < 1e-4
