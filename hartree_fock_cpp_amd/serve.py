"""Continuous-batching SCF server (include/hfx_serve.h) and the bond-length
scan stream it is measured on.

The reference solves scan points one after another (run_bl_scan,
include/run.hpp:300-423: one rhf_diis per bond length, each with a fresh
ERI_GPU).  Server admits jobs as in-flight slots and HBM budget free up and
runs each admitted job's whole SCF on its own worker thread and HIP stream, so
the device work of jobs in flight overlaps.

  srv = Server(max_inflight=8, hbm_budget=8 << 30)
  jid = srv.submit(config, "cc-pVDZ", precision="mixed")
  r = srv.result(jid)           # dict: E_tot, iterations, mode, t_submit/t_admit/t_done, ...
  srv.close()

scan_stream(n, rate, seed) draws the synthetic stream: HCl, N2, NaCl, H2O, NH3,
CH4 with every bond length uniform in [0.8, 2.5] x equilibrium, basis uniform
over {STO-3G, 6-31G*, cc-pVDZ}, Poisson arrivals at `rate` jobs/s.
"""
import ctypes as C
import math
import time

import numpy as np

from . import _native as N
from .basis import BASIS_DIR
from .scf import MODES, PRECISIONS, ScfOpts, ScfResult, _opts, _syms, _xyz


class ServerOpts(C.Structure):
    _fields_ = [("max_inflight", C.c_int), ("hbm_budget", C.c_size_t), ("allow_skip_ahead", C.c_int)]


class JobDesc(C.Structure):
    _fields_ = [("natom", C.c_int), ("symbols", C.c_char_p), ("xyz", C.c_void_p), ("basis_set", C.c_char_p),
                ("basis_path", C.c_char_p), ("scf", ScfOpts), ("P0", C.c_void_p)]


class JobResult(C.Structure):
    _fields_ = [("status", C.c_int), ("done", C.c_int), ("scf", ScfResult), ("K", C.c_int),
                ("device_bytes", C.c_size_t), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("error", C.c_char * 256)]


N.lib.hfx_server_destroy.restype = None
N.lib.hfx_server_now.restype = C.c_double
N.lib.hfx_server_now.argtypes = [C.c_void_p]
N.lib.hfx_server_result.argtypes = [C.c_void_p, C.c_longlong, C.c_int, C.c_void_p]
N.lib.hfx_server_density.argtypes = [C.c_void_p, C.c_longlong, C.c_void_p]

MODE_NAMES = {v: k for k, v in MODES.items()}


class Server:
    """Continuous-batching SCF server over the native library."""

    def __init__(self, max_inflight=8, hbm_budget=0, allow_skip_ahead=True):
        o = ServerOpts(max_inflight, hbm_budget, int(allow_skip_ahead))
        h = C.c_void_p()
        N.check(N.lib.hfx_server_create(C.byref(o), C.byref(h)))
        self._h = h
        self._K = {}

    def submit(self, config, basis_set, basis_path=BASIS_DIR, precision="fp64", eri_mode="auto", spin=None,
               charge=0, max_iterations=10000, epsilon=1e-6, damping=0.3, screening_threshold=1e-8,
               mixed_split=1e-5, energy_convergence=0.0, polish=True, P0=None):
        """Queues one SCF job (RHF, or UHF when spin=(n_alpha, n_beta)); returns its id."""
        o = _opts(charge, max_iterations, epsilon, damping, screening_threshold, eri_mode, precision, mixed_split,
                  energy_convergence, 0, polish, False, spin=spin)
        x = _xyz(config)
        p0 = N.as_f64(P0) if P0 is not None else None
        d = JobDesc(len(config), _syms(config), x.ctypes.data, basis_set.encode(), basis_path.encode(), o,
                    p0.ctypes.data if p0 is not None else None)
        jid = C.c_longlong()
        N.check(N.lib.hfx_server_submit(self._h, C.byref(d), C.byref(jid)))
        return jid.value

    def result(self, jid, wait=True):
        """Result dict of a job, or None while it is pending (wait=False)."""
        r = JobResult()
        rc = N.lib.hfx_server_result(self._h, jid, int(wait), C.byref(r))
        N.check(min(rc, 0))
        if rc == 0:
            return None
        out = r.scf.as_dict()
        out.update(status=r.status, error=r.error.decode(), K=r.K, device_bytes=r.device_bytes,
                   t_submit=r.t_submit, t_admit=r.t_admit, t_done=r.t_done, mode_name=MODE_NAMES.get(r.scf.mode))
        self._K[jid] = r.K
        return out

    def density(self, jid):
        K = self._K.get(jid) or self.result(jid)["K"]
        P = np.zeros((K, K))
        N.check(N.lib.hfx_server_density(self._h, jid, P.ctypes.data))
        return P

    def drain(self):
        N.check(N.lib.hfx_server_drain(self._h))

    def now(self):
        return N.lib.hfx_server_now(self._h)

    def stats(self):
        pi, pb, nd = C.c_int(), C.c_size_t(), C.c_longlong()
        N.check(N.lib.hfx_server_stats(self._h, C.byref(pi), C.byref(pb), C.byref(nd)))
        return {"peak_inflight": pi.value, "peak_bytes": pb.value, "n_done": nd.value}

    def close(self):
        if self._h:
            N.lib.hfx_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --- the synthetic scan stream ------------------------------------------------
# equilibrium bond lengths (bohr) and angles of the six scan molecules
EQUILIBRIUM = {"HCl": 2.409, "N2": 2.068, "NaCl": 4.46, "H2O": 1.811, "NH3": 1.913, "CH4": 2.052}
BASES = ("STO-3G", "6-31G*", "cc-pVDZ")


def scan_molecule(name, r):
    """Geometry (bohr) of a scan molecule at bond length r (all X-H bonds equal)."""
    if name in ("HCl", "N2", "NaCl"):
        a, b = {"HCl": ("Cl", "H"), "N2": ("N", "N"), "NaCl": ("Na", "Cl")}[name]
        return [(a, (0.0, 0.0, 0.0)), (b, (0.0, 0.0, r))]
    if name == "H2O":
        h = math.radians(104.5) / 2
        return [("O", (0.0, 0.0, 0.0)), ("H", (r * math.sin(h), r * math.cos(h), 0.0)),
                ("H", (-r * math.sin(h), r * math.cos(h), 0.0))]
    if name == "NH3":   # C3v, H-N-H 106.7 deg
        hnh = math.radians(106.7)
        # angle theta between the C3 axis and each N-H bond: cos(hnh) = 1 - 1.5 sin^2(theta)
        st = math.sqrt((1.0 - math.cos(hnh)) / 1.5)
        ct = math.sqrt(1.0 - st * st)
        at = [("N", (0.0, 0.0, 0.0))]
        for k in range(3):
            ph = 2 * math.pi * k / 3
            at.append(("H", (r * st * math.cos(ph), r * st * math.sin(ph), -r * ct)))
        return at
    if name == "CH4":
        s = r / math.sqrt(3.0)
        return [("C", (0.0, 0.0, 0.0))] + [("H", (s * x, s * y, s * z))
                                           for x, y, z in ((1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1))]
    raise ValueError(f"unknown scan molecule {name}")


def scan_stream(n, rate, seed=2026, precisions=("fp64", "mixed")):
    """n jobs: dicts {t_arrive, name, scale, config, basis, precision}; Poisson arrivals at `rate`/s."""
    rng = np.random.default_rng(seed)
    names = sorted(EQUILIBRIUM)
    t = 0.0
    jobs = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        name = names[rng.integers(len(names))]
        scale = float(rng.uniform(0.8, 2.5))
        jobs.append({"t_arrive": t, "name": name, "scale": scale,
                     "config": scan_molecule(name, scale * EQUILIBRIUM[name]),
                     "basis": BASES[rng.integers(len(BASES))],
                     "precision": precisions[rng.integers(len(precisions))]})
    return jobs


def run_stream(server, jobs, **scf_kw):
    """Submits each job at its arrival time (relative to now) and waits for all.

    Returns (results, summary) with latency = t_done - t_submit per job,
    p50/p99 latency, median service time (t_done - t_admit) and solves/s over
    the wall time from the first arrival to the last completion.
    """
    t0 = time.perf_counter()
    ids = []
    for j in jobs:
        dt = j["t_arrive"] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids.append(server.submit(j["config"], j["basis"], precision=j["precision"], **scf_kw))
    server.drain()
    wall = time.perf_counter() - t0
    res = [server.result(i) for i in ids]
    lat = np.array([r["t_done"] - r["t_submit"] for r in res])
    svc = np.array([r["t_done"] - r["t_admit"] for r in res])
    summary = {"n": len(res), "wall_s": wall, "solves_per_s": len(res) / wall,
               "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
               "median_service_s": float(np.median(svc)), "failed": sum(r["status"] != 0 for r in res),
               "unconverged": sum(not r["converged"] for r in res)}
    summary.update(server.stats())
    return res, summary
