"""Python mirror of the reference's SCF entry points, running the native
drivers (csrc/host/scf.cpp) with the GPU Fock build.

  rhf_diis(config, basis_set, ...)   include/scf.hpp:205-419
  uhf_diis(config, basis_set, (na, nb), ...)   include/scf.hpp:423-682
  one_electron(config, basis_set)    S, T, V (include/integrals.hpp:99-536)
config: [(symbol, (x, y, z)), ...] in bohr, as the reference's
std::vector<std::pair<std::string, vector3>>.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .basis import BASIS_DIR

MODES = {"stored": 0, "direct": 1, "auto": 2}
PRECISIONS = {"fp64": 0, "mixed": 1}


class ScfOpts(C.Structure):
    _fields_ = [("charge", C.c_int), ("n_alpha", C.c_int), ("n_beta", C.c_int), ("max_iterations", C.c_int),
                ("epsilon", C.c_double), ("damping", C.c_double), ("screening_threshold", C.c_double),
                ("eri_mode", C.c_int), ("precision", C.c_int), ("mixed_split", C.c_double),
                ("energy_convergence", C.c_double), ("hbm_budget", C.c_size_t), ("polish", C.c_int),
                ("verbose", C.c_int), ("device_iteration", C.c_int), ("incremental_fock", C.c_int)]


class ScfResult(C.Structure):
    _fields_ = [("E_tot", C.c_double), ("E_ele", C.c_double), ("E_nuc", C.c_double),
                ("E_before_polish", C.c_double), ("iterations", C.c_int), ("polish_iterations", C.c_int),
                ("converged", C.c_int), ("K", C.c_int), ("n_electrons", C.c_int), ("mode", C.c_int),
                ("n_quartets", C.c_longlong), ("t_setup", C.c_double), ("t_eri", C.c_double),
                ("t_fock", C.c_double), ("t_diag", C.c_double), ("t_total", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


N.lib.hfx_scf_default_opts.restype = None


def _syms(config):
    return b"".join(s.encode() + b"\0" for s, _ in config)


def _xyz(config):
    return np.ascontiguousarray(np.array([p for _, p in config], float).ravel())


def basis_size(config, basis_set, basis_path=BASIS_DIR):
    K = C.c_int()
    N.check(N.lib.hfx_basis_size(len(config), _syms(config), basis_set.encode(), basis_path.encode(), C.byref(K)))
    return K.value


def one_electron(config, basis_set, basis_path=BASIS_DIR):
    K = basis_size(config, basis_set, basis_path)
    S, T, V = (np.zeros((K, K)) for _ in range(3))
    x = _xyz(config)
    N.check(N.lib.hfx_one_electron(len(config), _syms(config), N.ptr(x), basis_set.encode(), basis_path.encode(),
                                   N.ptr(S), N.ptr(T), N.ptr(V)))
    return S, T, V


DEVICE_ITERATION = {"auto": -1, "on": 1, "off": 0, True: 1, False: 0, -1: -1, 1: 1, 0: 0}


def _opts(charge, max_iterations, epsilon, damping, screening_threshold, eri_mode, precision, mixed_split,
          energy_convergence, hbm_budget, polish, verbose, spin=None, device_iteration="auto", incremental_fock="auto"):
    o = ScfOpts()
    N.lib.hfx_scf_default_opts(C.byref(o))
    o.charge = charge
    o.max_iterations = max_iterations
    o.epsilon = epsilon
    o.damping = damping
    o.screening_threshold = screening_threshold
    o.eri_mode = MODES[eri_mode]
    o.precision = PRECISIONS[precision]
    o.mixed_split = mixed_split
    o.energy_convergence = energy_convergence
    o.hbm_budget = hbm_budget
    o.polish = int(polish)
    o.verbose = int(verbose)
    o.device_iteration = DEVICE_ITERATION[device_iteration]
    o.incremental_fock = DEVICE_ITERATION[incremental_fock]
    if spin is not None:
        o.n_alpha, o.n_beta = spin
    return o


def rhf_diis(config, basis_set, basis_path=BASIS_DIR, charge=0, P0=None, max_iterations=10000, epsilon=1e-6,
             damping=0.3, screening_threshold=1e-8, eri_mode="auto", precision="fp64", mixed_split=1e-5,
             energy_convergence=0.0, hbm_budget=0, polish=True, verbose=False, device_iteration="auto",
             incremental_fock="auto"):
    """Returns a dict with energies, timings, density, coefficient, mo_energy.
    device_iteration: "auto" (GPU-resident iteration for K >= 96), "on" or "off".
    incremental_fock: direct builds from density differences, "auto" (on), "on" or "off"."""
    K = basis_size(config, basis_set, basis_path)
    o = _opts(charge, max_iterations, epsilon, damping, screening_threshold, eri_mode, precision, mixed_split,
              energy_convergence, hbm_budget, polish, verbose, device_iteration=device_iteration,
              incremental_fock=incremental_fock)
    res = ScfResult()
    P, Cm, e = np.zeros((K, K)), np.zeros((K, K)), np.zeros(K)
    p0 = N.as_f64(P0) if P0 is not None else None
    x = _xyz(config)
    N.check(N.lib.hfx_scf_run(len(config), _syms(config), N.ptr(x), basis_set.encode(), basis_path.encode(),
                              C.byref(o), N.ptr(p0), None, C.byref(res), N.ptr(P), None, N.ptr(Cm), None, N.ptr(e),
                              None))
    out = res.as_dict()
    out.update(density=P, coefficient=Cm, mo_energy=e)
    return out


def uhf_diis(config, basis_set, spin_occ, basis_path=BASIS_DIR, charge=0, max_iterations=10000, epsilon=1e-6,
             damping=0.3, screening_threshold=1e-8, eri_mode="auto", precision="fp64", mixed_split=1e-5,
             energy_convergence=0.0, hbm_budget=0, polish=True, verbose=False, device_iteration="auto",
             incremental_fock="auto"):
    K = basis_size(config, basis_set, basis_path)
    o = _opts(charge, max_iterations, epsilon, damping, screening_threshold, eri_mode, precision, mixed_split,
              energy_convergence, hbm_budget, polish, verbose, spin=spin_occ, device_iteration=device_iteration,
              incremental_fock=incremental_fock)
    res = ScfResult()
    Pa, Pb, Ca, Cb = (np.zeros((K, K)) for _ in range(4))
    ea, eb = np.zeros(K), np.zeros(K)
    x = _xyz(config)
    N.check(N.lib.hfx_scf_run(len(config), _syms(config), N.ptr(x), basis_set.encode(), basis_path.encode(),
                              C.byref(o), None, None, C.byref(res), N.ptr(Pa), N.ptr(Pb), N.ptr(Ca), N.ptr(Cb),
                              N.ptr(ea), N.ptr(eb)))
    out = res.as_dict()
    out.update(Pa=Pa, Pb=Pb, Ca=Ca, Cb=Cb, Ea=ea, Eb=eb)
    return out
