"""CPU checks of the product's kernel bodies (csrc/device/eri_body.hpp and the
planner in plan.hpp), compiled for the host by tests/emu/cpu_emu.cpp and run
one lane at a time.  The same source is what the gfx950 kernels execute; the
GPU tests (test_gpu_*.py) repeat these checks on the device."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from tests import oracle_py as orc
from tests.conftest import GOLDEN, ROOT, golden

# HFX_EMU_LIBRARY: a sanitizer build of the same library (tests/sanitize)
EMU_PATH = os.environ.get("HFX_EMU_LIBRARY") or os.path.join(ROOT, "tests", "emu", "libhfx_emu.so")
if not os.path.exists(EMU_PATH):
    pytest.skip("CPU emulation library not built (run __graft_entry__.build())", allow_module_level=True)
emu = C.CDLL(EMU_PATH)
emu.emu_last_error.restype = C.c_char_p
_vp = C.c_void_p


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def _bargs(d):
    K = int(len(d["poff"]) - 1)
    arrs = [np.ascontiguousarray(d["xyz"], np.float64), np.ascontiguousarray(d["lmn"], np.int32),
            np.ascontiguousarray(d["poff"], np.int32), np.ascontiguousarray(d["alpha"], np.float64),
            np.ascontiguousarray(d["coeff"], np.float64), np.ascontiguousarray(d["norm"], np.float64)]
    return K, arrs, [C.c_int(K)] + [_p(a) for a in arrs]


def emu_unique(d, thr):
    K, keep, a = _bargs(d)
    P = K * (K + 1) // 2
    out = np.zeros(P * (P + 1) // 2)
    assert emu.emu_eri_unique(*a, C.c_double(thr), _p(out)) == 0, emu.emu_last_error()
    return out


def emu_fock(d, thr, Pa, Pb=None, mixed=False, split=1e-5):
    K, keep, a = _bargs(d)
    J = np.zeros((K, K))
    Ka = np.zeros((K, K))
    Kb = np.zeros((K, K)) if Pb is not None else None
    Pa = np.ascontiguousarray(Pa)
    Pb = np.ascontiguousarray(Pb) if Pb is not None else None
    assert emu.emu_fock_direct(*a, C.c_double(thr), C.c_int(int(mixed)), C.c_double(split), _p(Pa), _p(Pb),
                               _p(J), _p(Ka), _p(Kb)) == 0
    return (J, Ka) if Pb is None else (J, Ka, Kb)


def test_boys_table_path_matches_mpmath():
    with open(os.path.join(GOLDEN, "boys_mpmath.json")) as f:
        g = json.load(f)
    F64 = np.zeros(13)
    F32 = np.zeros(13, np.float32)
    worst64 = worst32 = 0.0
    for ti, T in enumerate(g["T"]):
        emu.emu_boys(C.c_int(12), C.c_double(T), _p(F64), _p(F32))
        for m in range(13):
            ex = float(g["exact"][m][ti])
            worst64 = max(worst64, abs(F64[m] - ex) / ex)
            worst32 = max(worst32, abs(float(F32[m]) - ex) / ex)
    assert worst64 < 5e-15, worst64
    assert worst32 < 2e-6, worst32


@pytest.mark.parametrize("name", ["h2_sto3g.npz", "h2o_sto3g.npz", "h2o_ccpvdz.npz"])
def test_emu_unique_matches_reference_and_exact(name):
    d = golden(name)
    thr = float(d["threshold"])
    got = emu_unique(d, thr)
    exact = orc.eri_unique(d, thr, exact=True)
    gold = d["eri"]
    elig = np.abs(gold - exact) < 1e-13
    assert np.abs(got - gold)[elig].max() < 1e-12
    assert np.abs(got - exact).max() < 1e-12


def test_emu_schwarz_bounds():
    d = golden("h2o_ccpvdz.npz")
    K, keep, a = _bargs(d)
    out = np.zeros(K * (K + 1) // 2)
    assert emu.emu_schwarz(*a, _p(out)) == 0
    assert np.abs(out - orc.schwarz_bounds_exact(d)).max() < 1e-13


@pytest.mark.slow
def test_emu_unique_o2_ccpvtz_all_classes():
    # cc-pVTZ O2 has every class up to (ff|ff): exercises both kernel families
    d = golden("o2_ccpvtz.npz")
    thr = float(d["threshold"])
    got = emu_unique(d, thr)
    q = d["qids"].astype(np.int64)
    exact = orc.eri_quartets(d, thr, q, exact=True)
    gold = d["eri_sample"]
    elig = np.abs(gold - exact) < 1e-13
    assert elig.sum() > 0.5 * len(q)
    assert np.abs(got[q] - gold)[elig].max() < 1e-12
    assert np.abs(got[q] - exact).max() < 1e-12


@pytest.mark.parametrize("name", ["h2o_sto3g.npz", "h2o_ccpvdz.npz"])
def test_emu_direct_jk_rhf(name):
    d = golden(name)
    K = int(d["K"])
    thr = float(d["threshold"])
    rng = np.random.default_rng(2)
    A = rng.standard_normal((K, K))
    P = A @ A.T / K
    J, Kx = emu_fock(d, thr, P)
    eri = orc.eri_unique(d, thr, exact=True)
    Gref = orc.repulsion_matrix(P, eri)
    assert np.abs((J - 0.5 * Kx) - Gref).max() < 1e-11 * max(1, np.abs(Gref).max())


def test_emu_direct_jk_uhf():
    d = golden("h2o_ccpvdz.npz")
    K = int(d["K"])
    thr = float(d["threshold"])
    rng = np.random.default_rng(4)
    A, B = rng.standard_normal((K, K)), rng.standard_normal((K, K))
    Pa, Pb = A @ A.T / K, B @ B.T / K
    J, Ka, Kb = emu_fock(d, thr, Pa, Pb)
    eri = orc.eri_unique(d, thr, exact=True)
    assert np.abs(J - orc.coulomb_matrix(Pa, Pb, eri)).max() < 1e-11 * K
    assert np.abs(Ka - orc.exchange_matrix(Pa, eri)).max() < 1e-11 * K
    assert np.abs(Kb - orc.exchange_matrix(Pb, eri)).max() < 1e-11 * K


def test_emu_mixed_close_to_fp64():
    d = golden("h2o_ccpvdz.npz")
    K = int(d["K"])
    rng = np.random.default_rng(6)
    A = rng.standard_normal((K, K))
    P = A @ A.T / K
    J0, K0 = emu_fock(d, 1e-12, P)
    J1, K1 = emu_fock(d, 1e-12, P, mixed=True, split=1e-3)
    assert np.abs(J1 - J0).max() < 1e-6
    assert np.abs(K1 - K0).max() < 1e-6
    n64 = np.zeros(55, np.int64)
    n32 = np.zeros(55, np.int64)
    K_, keep, a = _bargs(d)
    emu.emu_class_counts(*a, C.c_double(1e-12), C.c_int(1), C.c_double(1e-3), _p(n64), _p(n32))
    assert n32.sum() > 0 and n64.sum() > 0


@pytest.mark.parametrize("S", [1, 3, 7, -2, -5])
def test_emu_primitive_slices_are_exact(S):
    """Slicing a quartet's primitive loop and adding the partial blocks (unique
    array and J/K) reproduces the unsliced result to rounding.  Negative S:
    family-B quartets cut into |S| ket-primitive chunks per bra pair."""
    d = golden("h2o_ccpvdz.npz")
    K = int(d["K"])
    thr = float(d["threshold"])
    rng = np.random.default_rng(5)
    A = rng.standard_normal((K, K))
    P = A @ A.T / K
    emu.emu_set_slices(C.c_int(1))
    try:
        u1 = emu_unique(d, thr)
        J1, K1 = emu_fock(d, thr, P)
        emu.emu_set_slices(C.c_int(S))
        uS = emu_unique(d, thr)
        JS, KS = emu_fock(d, thr, P)
    finally:
        emu.emu_set_slices(C.c_int(0))
    assert np.abs(uS - u1).max() < 1e-14
    assert np.abs(JS - J1).max() < 1e-12 and np.abs(KS - K1).max() < 1e-12
    exact = orc.eri_unique(d, thr, exact=True)
    assert np.abs(uS - exact).max() < 1e-12


@pytest.mark.parametrize("threads,mixed", [(2, 0), (7, 0), (16, 1)])
def test_plan_threads_agree(threads, mixed):
    """The quartet planner cut over several threads (row ranges of equal
    area, per-list counting sorts in parallel) yields the same lists, in the
    same order, with the same lane-group buckets as on one thread: benzene /
    cc-pVDZ (1.1 M quartets), FP64 and the mixed split."""
    from hartree_fock_cpp_amd import workloads as W
    from hartree_fock_cpp_amd.basis import basis_construction
    d = basis_construction(W.benzene(), "cc-pVDZ")
    K, keep, a = _bargs(d)
    nq = C.c_longlong(0)
    bad = emu.emu_plan_threads_agree(*a, C.c_double(1e-8), C.c_int(mixed), C.c_double(1e-5), C.c_int(threads),
                                     C.byref(nq))
    assert bad == 0, (bad, emu.emu_last_error())
    assert nq.value > 1_000_000


@pytest.mark.parametrize("threads,mixed", [(1, 0), (16, 1)])
def test_plan_jk_order(threads, mixed):
    """The J/K plan order: every class list holds the same quartets and
    lane-group buckets as the default plan, ordered by primitive work
    (descending), then bra pair, ket first shell and ket pair (ascending,
    no repeats) -- the order the J/K epilogue's run merging relies on:
    benzene / cc-pVDZ at 1e-8, FP64 and the mixed split."""
    from hartree_fock_cpp_amd import workloads as W
    from hartree_fock_cpp_amd.basis import basis_construction
    d = basis_construction(W.benzene(), "cc-pVDZ")
    K, keep, a = _bargs(d)
    nq = C.c_longlong(0)
    bad = emu.emu_plan_jk_order_check(*a, C.c_double(1e-8), C.c_int(mixed), C.c_double(1e-5), C.c_int(threads),
                                      C.byref(nq))
    assert bad == 0, (bad, emu.emu_last_error())
    assert nq.value > 1_000_000


@pytest.mark.parametrize("threads", [1, 4, 16])
def test_parallel_tasks_propagates_exceptions(threads):
    """The planner's thread helper runs every task once when none throws, and
    hands a task's exception back to the calling thread (not std::terminate)
    when one does."""
    ran = C.c_int(0)
    assert emu.emu_parallel_tasks_check(C.c_int(200), C.c_int(threads), C.c_int(-1), C.byref(ran)) == 0
    assert ran.value == 200
    assert emu.emu_parallel_tasks_check(C.c_int(200), C.c_int(threads), C.c_int(57), C.byref(ran)) == 1
    assert ran.value < 200
