"""Python mirror of the reference's GPU ERI boundary and Fock build.

Names follow the reference (KLYYCYNN/hartree_fock_cpp):
  ERI_GPU(basis, threshold)            cuda/src/eri.cu:20-89 (declared include/scf.hpp:41-44)
  schwarz_bounds(basis)                include/integrals.hpp:732-752
  repulsion_matrix(P, eri)             include/linalg.hpp:100-129   (G = J - K/2)
  coulomb_matrix(Pa, Pb, eri)          include/linalg.hpp:215-257
  exchange_matrix(P, eri)              include/linalg.hpp:260-296
Arrays are numpy; everything is computed by the native gfx950 library.
"""
import ctypes as C

import numpy as np

from . import _native as N


class Basis:
    """Flat Cartesian basis in the reference's basis_function layout
    (include/types.hpp:16-29): per function a centre, (lx, ly, lz) and its
    primitives (alpha, coeff, norm) at poff[mu]..poff[mu+1]-1."""

    def __init__(self, xyz, lmn, poff, alpha, coeff, norm):
        self.xyz = N.as_f64(xyz).reshape(-1)
        self.lmn = N.as_i32(lmn).reshape(-1)
        self.poff = N.as_i32(poff).reshape(-1)
        self.alpha = N.as_f64(alpha)
        self.coeff = N.as_f64(coeff)
        self.norm = N.as_f64(norm)
        self.K = len(self.poff) - 1
        if self.xyz.size != 3 * self.K or self.lmn.size != 3 * self.K:
            raise ValueError("xyz and lmn must hold 3 entries per function")
        self._h = None

    @classmethod
    def from_arrays(cls, d):
        return cls(d["xyz"], d["lmn"], d["poff"], d["alpha"], d["coeff"], d["norm"])

    @property
    def n_pairs(self):
        return self.K * (self.K + 1) // 2

    @property
    def n_quartets(self):
        P = self.n_pairs
        return P * (P + 1) // 2

    def handle(self):
        """Device-resident tables (created on first use)."""
        if self._h is None:
            h = C.c_void_p()
            N.check(N.lib.hfx_basis_create(C.c_int(self.K), N.ptr(self.xyz), N.ptr(self.lmn), N.ptr(self.poff),
                                           N.ptr(self.alpha), N.ptr(self.coeff), N.ptr(self.norm), C.byref(h)))
            self._h = h
        return self._h

    def info(self):
        ns, npair, npp = C.c_int(), C.c_int(), C.c_int()
        N.check(N.lib.hfx_basis_info(self.handle(), C.byref(ns), C.byref(npair), C.byref(npp)))
        return {"n_shells": ns.value, "n_shell_pairs": npair.value, "n_prim_pairs": npp.value}

    def close(self):
        if self._h is not None:
            N.lib.hfx_basis_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def schwarz_bounds(basis):
    out = np.zeros(basis.n_pairs)
    N.check(N.lib.hfx_schwarz_bounds(basis.handle(), N.ptr(out)))
    return out


def ERI_GPU(basis, threshold):
    """Unique ERI array (eri_index layout), FP64, function-level Schwarz mask."""
    out = np.zeros(basis.n_quartets)
    N.check(N.lib.hfx_eri_unique(basis.handle(), C.c_double(threshold), N.ptr(out), C.c_size_t(out.size)))
    return out


class FockBuilder:
    """One SCF job's Fock-build context (stored or direct, FP64 or mixed)."""

    def __init__(self, basis, mode="auto", precision="fp64", threshold=1e-8, mixed_split=1e-5, hbm_budget=0,
                 incremental=False):
        self.basis = basis
        o = N.FockOpts()
        o.mode = {"stored": N.HFX_MODE_STORED, "direct": N.HFX_MODE_DIRECT, "auto": N.HFX_MODE_AUTO}[mode]
        o.precision = {"fp64": N.HFX_PREC_FP64, "mixed": N.HFX_PREC_MIXED}[precision]
        o.threshold = threshold
        o.mixed_split = mixed_split
        o.hbm_budget = hbm_budget
        o.incremental = int(bool(incremental))
        self.opts = o
        h = C.c_void_p()
        N.check(N.lib.hfx_fock_create(basis.handle(), C.byref(o), C.byref(h)))
        self._h = h

    def info(self):
        m, b, n = C.c_int(), C.c_size_t(), C.c_size_t()
        N.check(N.lib.hfx_fock_info(self._h, C.byref(m), C.byref(b), C.byref(n)))
        return {"mode": "stored" if m.value == 0 else "direct", "device_bytes": b.value, "n_quartets": n.value}

    def autotune_info(self):
        """dict(done, target, targets, ms) of the direct-build work-unit autotuner."""
        done, tgt = C.c_int(), C.c_longlong()
        targets = np.zeros(3, np.int64)
        ms = np.zeros(3, np.float32)
        N.check(N.lib.hfx_fock_autotune_info(self._h, C.byref(done), C.byref(tgt), N.ptr(targets), N.ptr(ms)))
        return {"done": bool(done.value), "target": tgt.value, "targets": targets.tolist(), "ms": ms.tolist()}

    def class_counts(self):
        n64 = np.zeros(55, np.int64)
        n32 = np.zeros(55, np.int64)
        N.check(N.lib.hfx_fock_class_counts(self._h, N.ptr(n64), N.ptr(n32)))
        return n64, n32

    def build(self, Pa, Pb=None):
        """RHF (Pb None): returns (J, K).  UHF: returns (J, Ka, Kb)."""
        K = self.basis.K
        Pa = N.as_f64(Pa)
        J = np.zeros((K, K))
        Ka = np.zeros((K, K))
        if Pb is None:
            N.check(N.lib.hfx_fock_build(self._h, N.ptr(Pa), None, N.ptr(J), N.ptr(Ka), None))
            return J, Ka
        Pb = N.as_f64(Pb)
        Kb = np.zeros((K, K))
        N.check(N.lib.hfx_fock_build(self._h, N.ptr(Pa), N.ptr(Pb), N.ptr(J), N.ptr(Ka), N.ptr(Kb)))
        return J, Ka, Kb

    def close(self):
        if getattr(self, "_h", None) is not None:
            N.lib.hfx_fock_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def repulsion_matrix(basis, P, threshold=1e-8):
    """G = J - K/2 from the stored unique ERIs (linalg.hpp:100-129)."""
    fb = FockBuilder(basis, mode="stored", threshold=threshold)
    J, K = fb.build(P)
    fb.close()
    return J - 0.5 * K
