#!/usr/bin/env python3
"""Benchmark: unique-ERI build throughput of the FP64 drop-in on one MI355X.

Workload (the configuration the headline metric is quoted on): benzene
RHF/cc-pVDZ (K = 120, Q = 26,357,430 unique function quartets), stored FP64
ERIs with the reference's screening threshold 1e-8.  One step = one complete
unique-ERI build into HBM (all 55 class kernels, function-level Schwarz mask,
eri_index layout) with the basis tables and the quartet plan resident, i.e.
the ERI phase of the reference's "eri_hardware":"GPU" path.

value = screened unique quartets (b_IJ b_KL >= threshold) per second over the
timed steps (the library's own launch schedule: class lists spread over four
streams).  The roofline object covers the dominant class kernel, timed with
HIP events on its launch stream while the same build is replayed, after the
timed region, with the class lists serialized on one stream.  The CPU
baseline times the reference's own basis_repulsion (oracle/_ref, compiled
from the reference sources) on one host core over a bounded random sample of
the same quartets.

Beside the headline line's own fields, the same JSON line carries
  scf         SCF wall times: benzene/cc-pVDZ stored FP64; n-octane/cc-pVDZ and
              (H2O)4/cc-pVTZ direct, FP64 vs mixed with dE, max|dF|, max|dD|
  readme_scf  RHF/cc-pVTZ SCF wall time of the eight molecules of the
              reference's published timing table, beside the published seconds
  serving     the scan-serving stream at 0.7 x measured saturation: p50 / p99
              latency, median solve time and solves/s
  stored_jk   J/K from the stored unique array (device buffers) per build, with
              its HBM roofline
  pcie_inclusive  the drop-in as the reference calls it: hfx_eri_unique into a
              host array (build + device-to-host copy of the whole array)
(--no-extras skips the first three).

Usage: python bench.py [--steps K] [--warmup W] [--cpu-seconds S] [--no-cpu] [--no-extras]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def count_screened(bounds, thr):
    import numpy as np
    b = np.sort(bounds)
    with np.errstate(divide="ignore"):
        need = np.where(b > 0, thr / b, np.inf)
    n_all = int((len(b) - np.searchsorted(b, need, side="left")).sum())
    n_diag = int((b * b >= thr).sum())
    return (n_all + n_diag) // 2


def cpu_baseline(arrs, bounds, thr, budget_s, seed=0):
    """Reference basis_repulsion (oracle/_ref/libref.so) on one core, or the
    oracle's restatement when the reference build is absent."""
    import numpy as np
    K = len(arrs["poff"]) - 1
    P = K * (K + 1) // 2
    Q = P * (P + 1) // 2
    rng = np.random.default_rng(seed)
    cand = rng.integers(0, Q, size=400000, dtype=np.int64)
    IJ = ((np.sqrt(8.0 * cand + 1.0) - 1.0) // 2).astype(np.int64)
    IJ -= (IJ * (IJ + 1) // 2 > cand)
    IJ += ((IJ + 1) * (IJ + 2) // 2 <= cand)
    KL = cand - IJ * (IJ + 1) // 2
    keep = cand[bounds[IJ] * bounds[KL] >= thr]
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    kind = "reference" if os.path.exists(ref_path) else "port"
    if kind == "reference":
        lib = C.CDLL(ref_path)
        fn = lib.ref_eri_quartets
    else:
        lib = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
        fn = lib.orc_eri_quartets
    keepalive = [np.ascontiguousarray(arrs[k], t) for k, t in
                 (("xyz", np.float64), ("lmn", np.int32), ("poff", np.int32), ("alpha", np.float64),
                  ("coeff", np.float64), ("norm", np.float64))]
    args = [C.c_int(K)] + [a.ctypes.data_as(C.c_void_p) for a in keepalive]
    bd = np.ascontiguousarray(bounds)

    def run(q):
        q = np.ascontiguousarray(q, np.uint64)
        out = np.zeros(len(q))
        t = time.perf_counter()
        fn(*args, bd.ctypes.data_as(C.c_void_p), C.c_double(thr), q.ctypes.data_as(C.c_void_p), C.c_size_t(len(q)),
           out.ctypes.data_as(C.c_void_p))
        return time.perf_counter() - t

    t_cal = run(keep[:64])
    n = int(min(len(keep) - 64, max(64, budget_s / max(t_cal / 64, 1e-9))))
    t = run(keep[64:64 + n])
    out = {"value": n / t, "unit": "quartets/s", "cores": 1, "kind": kind,
           "sample": f"{n} unique quartets drawn uniformly (seed {seed}) from the screened set of the same "
                     f"workload, {t:.1f} s of single-core basis_repulsion (include/integrals.hpp:701-729)"}
    # the same sample split over every host core (ctypes releases the GIL in the call)
    import threading
    # the host cores that go with one GPU (OMP_NUM_THREADS where it is set), at most 16
    cores = max(1, min(16, os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16)))
    parts = np.array_split(keep[64:64 + n], cores)
    t0 = time.perf_counter()
    th = [threading.Thread(target=run, args=(q,)) for q in parts]
    for x in th:
        x.start()
    for x in th:
        x.join()
    t_all = time.perf_counter() - t0
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    out["all_cores"] = {"value": n / t_all, "cores": cores, "cpu_model": cpu_model,
                        "sample": f"the same {n} quartets in {cores} threads"}
    return out


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def polar_config(atoms):
    """[(symbol, (x, y, z))] from the reference's JSON atom records (include/run.hpp:98-123)."""
    import math
    out = []
    for a in atoms:
        r = a["position"]
        if a.get("coordinate_type", "cartesian") == "polar":
            th, ph = r[1] * math.pi / 180.0, r[2] * math.pi / 180.0
            out.append((a["symbol"], (r[0] * math.sin(th) * math.cos(ph), r[0] * math.sin(th) * math.sin(ph),
                                      r[0] * math.cos(th))))
        else:
            out.append((a["symbol"], tuple(float(x) for x in r)))
    return out


def direct_scf_pair(name, atoms, basis_set, thr=1e-12, split=1e-5):
    """FP64 vs mixed direct SCF: wall times, dE (before / after the FP64
    polish), max|dF| of one Fock build at the FP64 density, max|dD|."""
    import numpy as np
    import hartree_fock_cpp_amd as hfx
    from hartree_fock_cpp_amd.basis import basis_construction
    from hartree_fock_cpp_amd.scf import rhf_diis
    t = time.perf_counter()
    r64 = rhf_diis(atoms, basis_set, eri_mode="direct", precision="fp64", screening_threshold=thr)
    w64 = time.perf_counter() - t
    t = time.perf_counter()
    rmx = rhf_diis(atoms, basis_set, eri_mode="direct", precision="mixed", screening_threshold=thr,
                   mixed_split=split, polish=True)
    wmx = time.perf_counter() - t
    b = hfx.Basis.from_arrays(basis_construction(atoms, basis_set))
    P = r64["density"]
    G = {}
    for prec in ("fp64", "mixed"):
        fb = hfx.FockBuilder(b, mode="direct", precision=prec, threshold=thr, mixed_split=split)
        J, K = fb.build(P)
        G[prec] = J - 0.5 * K
        fb.close()
    b.close()
    out = {"workload": f"{name} RHF direct SCF, Schwarz {thr:g}, FP32/FP64 split {split:g}", "K": r64["K"],
           "fp64_wall_s": w64, "mixed_wall_s": wmx, "mixed_over_fp64_speedup": w64 / wmx,
           "fp64_iterations": r64["iterations"], "mixed_iterations": rmx["iterations"],
           "mixed_polish_iterations": rmx["polish_iterations"],
           "fp64_fock_s": r64["t_fock"], "mixed_fock_s": rmx["t_fock"],
           "E_fp64": r64["E_tot"], "dE_before_polish": rmx["E_before_polish"] - r64["E_tot"],
           "dE": rmx["E_tot"] - r64["E_tot"], "max_dF": float(np.abs(G["mixed"] - G["fp64"]).max()),
           "max_dD": float(np.abs(rmx["density"] - P).max()), "converged": bool(r64["converged"] and rmx["converged"])}
    log(f"  {name}: fp64 {w64:.2f} s, mixed {wmx:.2f} s, dE {out['dE']:.2e}")
    return out


def stored_jk(hfx, N, basis, thr, steps):
    """J/K from the stored unique array (the per-iteration work of a stored SCF)
    on device buffers, timed with HIP events on the launch stream; HBM roofline
    of the one pass over the array (k_jk_rows) plus the K x K matrices."""
    import torch
    K, Q = basis.K, basis.n_quartets
    fb = hfx.FockBuilder(basis, mode="stored", threshold=thr)
    g = torch.Generator().manual_seed(0)
    A = torch.randn(K, K, generator=g, dtype=torch.float64)
    P = ((A @ A.T) / K).cuda()
    J = torch.empty(K, K, dtype=torch.float64, device="cuda")
    Kx = torch.empty_like(J)
    # a stream of its own: a null handle would send the build to the basis'
    # stream, away from the events
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    sh = C.c_void_p(st.cuda_stream)
    ptr = lambda t: C.c_void_p(t.data_ptr())
    def build():
        N.check(N.lib.hfx_fock_build_device(fb._h, ptr(P), None, ptr(J), ptr(Kx), None, sh))
    for _ in range(3):
        build()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(steps):
        build()
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    fb.close()
    # algorithmic bytes of one RHF build: one pass over the Q-long FP64 unique
    # array plus the K x K matrices (one density in; J and K out)
    hbm = 8 * Q + 8 * K * K * (1 + 2)
    gbs = hbm / (ms * 1e-3) / 1e9
    log(f"  stored J/K: {ms:.3f} ms per build, {gbs:.0f} GB/s of algorithmic bytes")
    traffic = None   # measured HBM bytes per build (FETCH_SIZE / WRITE_SIZE runs, profiles/)
    tpath = os.path.join(ROOT, "profiles", "traffic_benzene_stored_jk.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_build")
    return {"workload": "benzene RHF/cc-pVDZ J/K from the stored unique array (FP64)", "ms_per_build": ms,
            "hbm_bytes": hbm, "achieved_GBps": gbs, "peak_GBps": PEAK_HBM_GBS, "frac": gbs / PEAK_HBM_GBS,
            "traffic": traffic, "timing": f"HIP events on the launch stream over {steps} builds"}


def direct_build(name, atoms, basis_set, steps, thr=1e-12, split=1e-5):
    """Full (non-incremental) direct J/K builds, FP64 and mixed, on device
    buffers at a fixed density: per-build time from HIP events on the launch
    stream, model GFLOP per build (screened primitive quartets x per-class
    flops), and the dominant class kernel from a serialized replay."""
    import numpy as np
    import torch
    import hartree_fock_cpp_amd as hfx
    from hartree_fock_cpp_amd import _native as N
    from hartree_fock_cpp_amd import workloads as W
    from hartree_fock_cpp_amd.basis import basis_construction
    b = hfx.Basis.from_arrays(basis_construction(atoms, basis_set))
    K = b.K
    g = torch.Generator().manual_seed(0)
    A = torch.randn(K, K, generator=g, dtype=torch.float64)
    P = ((A @ A.T) / K).cuda()
    J = torch.empty(K, K, dtype=torch.float64, device="cuda")
    Kx = torch.empty_like(J)
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    sh = C.c_void_p(st.cuda_stream)
    ptr = lambda t: C.c_void_p(t.data_ptr())
    out = {"workload": f"{name} RHF direct J/K build (full, not incremental), Schwarz {thr:g}", "K": K}
    for prec in ("fp64", "mixed"):
        pcode = N.HFX_PREC_MIXED if prec == "mixed" else N.HFX_PREC_FP64
        nq = np.zeros(110, np.int64)
        nprimq = np.zeros(110, np.int64)
        N.check(N.lib.hfx_plan_stats(b.handle(), C.c_double(thr), pcode, C.c_double(split), N.ptr(nq),
                                     N.ptr(nprimq)))
        fl = np.array([W.model_flops(*W.qclass_l(i // 2)) for i in range(110)], float) * nprimq
        fb = hfx.FockBuilder(b, mode="direct", precision=prec, threshold=thr, mixed_split=split)
        build = lambda: N.check(N.lib.hfx_fock_build_device(fb._h, ptr(P), None, ptr(J), ptr(Kx), None, sh))
        for _ in range(2):
            build()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(steps):
            build()
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        N.check(N.lib.hfx_set_class_timing(2))
        cls = np.zeros(110)
        buf = np.zeros(110, np.float32)
        for _ in range(steps):
            build()
            st.synchronize()
            N.check(N.lib.hfx_class_times(N.ptr(buf)))
            cls += buf
        N.check(N.lib.hfx_set_class_timing(0))
        torch.cuda.synchronize()
        fb.close()
        cls /= steps
        dom = int(np.argmax(cls))
        r = {"ms_per_build": ms, "model_gflop": fl.sum() / 1e9, "tflops": fl.sum() / (ms * 1e-3) / 1e12,
             "frac_fp64_peak": fl.sum() / (ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
             "shell_quartets": int(nq.sum()), "fp32_shell_quartets": int(nq[1::2].sum()),
             "serialized_class_ms": float(cls.sum()),
             "top_classes_ms": {"%s %s fp%s" % (W.kernel_family(i // 2, i % 2 == 1), W.qclass_name(i // 2),
                                                 "32" if i % 2 else "64"): round(float(cls[i]), 4)
                                for i in np.argsort(-cls)[:12] if cls[i] > 0},
             "dominant": {"kernel": "%s %s fp%s" % (W.kernel_family(dom // 2, dom % 2 == 1), W.qclass_name(dom // 2),
                                                    "32" if dom % 2 else "64"),
                          "launch_ms": float(cls[dom]), "model_gflop": fl[dom] / 1e9,
                          "tflops": fl[dom] / (cls[dom] * 1e-3) / 1e12 if cls[dom] > 0 else None}}
        out[prec] = r
        log(f"  {name} direct build {prec}: {ms:.2f} ms, {r['tflops']:.2f} model TFLOP/s, "
            f"dominant {r['dominant']['kernel']} {cls[dom]:.3f} ms")
    out["mixed_over_fp64"] = out["fp64"]["ms_per_build"] / out["mixed"]["ms_per_build"]
    b.close()
    return out


def pcie_inclusive(hfx, N, basis, arrs, thr, q_screened, reps=5):
    """The drop-in as the reference calls it: hfx_eri_unique into a pageable
    host array (the unique build plus the device-to-host copy of the whole
    Q-long array), wall-clock per call.  Reported beside the headline, never
    as its value (the headline keeps the array resident in HBM).  "cold" is
    what one ERI_GPU call through the shim costs: a fresh basis handle (device
    tables, Schwarz bounds, quartet plan), the build, the copy, the teardown."""
    import numpy as np
    Q = basis.n_quartets
    out = np.empty(Q)
    call = lambda: N.check(N.lib.hfx_eri_unique(basis.handle(), C.c_double(thr), N.ptr(out), C.c_size_t(Q)))
    call()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        call()
        ts.append(time.perf_counter() - t)
    ms = float(np.median(ts)) * 1e3

    def cold():
        b = hfx.Basis.from_arrays(arrs)
        N.check(N.lib.hfx_eri_unique(b.handle(), C.c_double(thr), N.ptr(out), C.c_size_t(Q)))
        b.close()
    cold()
    tc = []
    for _ in range(reps):
        t = time.perf_counter()
        cold()
        tc.append(time.perf_counter() - t)
    ms_cold = float(np.median(tc)) * 1e3
    log(f"  PCIe-inclusive unique build: {ms:.2f} ms per call, cold {ms_cold:.2f} ms "
        f"({8 * Q / 1e6:.0f} MB to the host)")
    return {"ms_per_call": ms, "value": q_screened / (ms * 1e-3), "unit": "quartets/s",
            "bytes_to_host": 8 * Q, "timing": f"median wall clock of {reps} calls into a pageable numpy array",
            "cold_ms_per_call": ms_cold, "cold_value": q_screened / (ms_cold * 1e-3),
            "cold_timing": f"median of {reps} calls, each: basis handle from the basis arrays, Schwarz bounds, "
                           "quartet plan, build, copy to the host array, handle destroyed"}


def stored_scf(name, atoms, basis_set, thr=1e-8):
    from hartree_fock_cpp_amd.scf import rhf_diis
    t = time.perf_counter()
    r = rhf_diis(atoms, basis_set, eri_mode="auto", precision="fp64", screening_threshold=thr)
    w = time.perf_counter() - t
    log(f"  {name}: {w:.3f} s, E {r['E_tot']:.10f}, {r['iterations']} iterations")
    return {"wall_s": w, "E_tot": r["E_tot"], "iterations": r["iterations"], "K": r["K"],
            "eri_s": r["t_eri"], "fock_s": r["t_fock"], "diag_s": r["t_diag"], "mode": r["mode"],
            "converged": bool(r["converged"])}


def uhf_scf(name, atoms, basis_set, spin):
    from hartree_fock_cpp_amd.scf import uhf_diis
    t = time.perf_counter()
    r = uhf_diis(atoms, basis_set, spin)
    w = time.perf_counter() - t
    log(f"  {name}: {w:.3f} s, E {r['E_tot']:.10f}, {r['iterations']} iterations")
    return {"wall_s": w, "E_tot": r["E_tot"], "iterations": r["iterations"], "K": r["K"], "eri_s": r["t_eri"],
            "fock_s": r["t_fock"], "diag_s": r["t_diag"], "mode": r["mode"], "converged": bool(r["converged"])}


def readme_scf():
    """The reference's published table: RHF/cc-pVTZ SCF wall time per molecule."""
    with open(os.path.join(ROOT, "tests", "golden", "readme_molecules.json")) as f:
        mols = json.load(f)
    out = {}
    for name, m in mols.items():
        r = stored_scf(name, polar_config(m["atoms"]), m["basis"])
        r.update(published_gpu_s=m["published_gpu_s"], published_cpu_s=m["published_cpu_s"],
                 speedup_vs_published_gpu=m["published_gpu_s"] / r["wall_s"])
        out[name] = r
    return out


def serving(n_cal=200, n_jobs=500, load=0.7, seed=2026):
    """Scan-serving stream: saturation from a burst of n_cal jobs, then n_jobs
    Poisson arrivals at load x saturation."""
    from hartree_fock_cpp_amd.serve import Server, run_stream, scan_stream
    kw = dict(max_iterations=300)
    burst = scan_stream(n_cal, rate=1e6, seed=seed + 1)
    with Server(max_inflight=16, hbm_budget=32 << 30) as s:
        _, cal = run_stream(s, burst, **kw)
    sat = cal["solves_per_s"]
    rate = load * sat
    log(f"  serving: saturation {sat:.1f} solves/s, offered {rate:.1f}/s")
    jobs = scan_stream(n_jobs, rate=rate, seed=seed)
    with Server(max_inflight=16, hbm_budget=32 << 30) as s:
        res, summ = run_stream(s, jobs, **kw)
    import numpy as np
    svc = np.array([r["t_done"] - r["t_admit"] for r in res])
    summ.update(saturation_solves_per_s=sat, offered_rate=rate, load=load, seed=seed,
                median_solve_s=float(np.median(svc)),
                p99_over_median_solve=summ["p99_latency_s"] / float(np.median(svc)),
                fp64_jobs=sum(j["precision"] == "fp64" for j in jobs),
                mixed_jobs=sum(j["precision"] == "mixed" for j in jobs))
    log(f"  serving: p50 {summ['p50_latency_s']*1e3:.1f} ms, p99 {summ['p99_latency_s']*1e3:.1f} ms, "
        f"{summ['solves_per_s']:.1f} solves/s")
    return summ


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the SCF wall-time, mixed-vs-FP64 and serving sections")
    ap.add_argument("--direct-build", choices=["water4", "octane"], default=None,
                    help="only time full direct J/K builds (FP64 and mixed) of (H2O)4/cc-pVTZ or "
                         "n-octane/cc-pVDZ and print their JSON object")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times (the reference kernel "
                         "for scaling HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe << 20
        src = torch.ones(n // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n, "launches": 5}), flush=True)
        return
    if args.direct_build:
        from hartree_fock_cpp_amd import workloads as W
        mol = {"water4": ("(H2O)4/cc-pVTZ", W.water_tetramer, "cc-pVTZ"),
               "octane": ("n-octane/cc-pVDZ", W.n_octane, "cc-pVDZ")}[args.direct_build]
        print(json.dumps(direct_build(mol[0], mol[1](), mol[2], args.steps)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X: the engine runs on a single GPU; sharding over several GPUs "
              "(work-unit lists split, J/K all-reduced) is designed but not built.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    import hartree_fock_cpp_amd as hfx
    from hartree_fock_cpp_amd import _native as N
    from hartree_fock_cpp_amd import workloads as W
    from hartree_fock_cpp_amd.basis import basis_construction

    thr = 1e-8
    atoms = W.benzene()
    arrs = basis_construction(atoms, "cc-pVDZ")
    basis = hfx.Basis.from_arrays(arrs)
    bounds = hfx.schwarz_bounds(basis)
    Q = basis.n_quartets
    q_screened = count_screened(bounds, thr)

    nq = np.zeros(110, np.int64)
    nprimq = np.zeros(110, np.int64)
    N.check(N.lib.hfx_plan_stats(basis.handle(), C.c_double(thr), 0, C.c_double(0.0), N.ptr(nq), N.ptr(nprimq)))
    flops_cls = np.array([W.model_flops(*W.qclass_l(i // 2)) for i in range(110)], float) * nprimq
    out = torch.empty(Q, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    sh = C.c_void_p(stream.cuda_stream)

    def step():
        N.check(N.lib.hfx_eri_unique_device(basis.handle(), C.c_double(thr), C.c_void_p(out.data_ptr()),
                                            C.c_size_t(Q), sh))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # Roofline pass, after the timed region: the same build replayed with every
    # class list on the launch stream (one class kernel on the device at a
    # time) and HIP events around each list on that stream.  In the timed
    # builds the lists overlap on four streams, so a kernel's span there
    # measures the device it shares, not the kernel.
    N.check(N.lib.hfx_set_class_timing(2))
    cls_ms = np.zeros(110)
    buf = np.zeros(110, np.float32)
    for _ in range(args.steps):
        step()
        N.check(N.lib.hfx_class_times(N.ptr(buf)))
        cls_ms += buf
    N.check(N.lib.hfx_set_class_timing(0))
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    dom = int(np.argmax(cls_ms))
    dom_ms = cls_ms[dom] / args.steps
    dom_tflops = flops_cls[dom] / (dom_ms * 1e-3) / 1e12
    fam_dom = W.kernel_family(dom // 2, dom % 2 == 1)
    build_tflops = flops_cls.sum() / (ms_per_step * 1e-3) / 1e12

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_benzene_eri.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        if tj.get("kernel_class") == W.qclass_name(dom // 2):
            traffic = tj.get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(arrs, bounds, thr, args.cpu_seconds)

    line = {
        "metric": "ERI quartets/s (unique-ERI build, FP64 drop-in)",
        "value": q_screened * args.steps / elapsed,
        "unit": "quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "benzene RHF/cc-pVDZ unique-ERI build (stored FP64, eri_index layout)",
                   "K": basis.K, "Q_total": Q, "Q_screened": q_screened, "threshold": thr,
                   "shell_quartets": int(nq.sum()), "primitive_quartets": int(nprimq.sum()),
                   "model_gflop_per_build": flops_cls.sum() / 1e9},
        "roofline": {"bound": "mfma",
                     "pipe": "FP64 compute roofline; the kernel's FP64 work issues on the %s "
                             "(FP64 VALU and FP64 MFMA share the 78.6 TFLOP/s peak)"
                             % ("matrix cores" if fam_dom == "k_fam_m" else "VALU"),
                     "kernel": "%s %s fp%s" % (fam_dom, W.qclass_name(dom // 2),
                                               "32" if dom % 2 else "64"),
                     "achieved": dom_tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": dom_tflops / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "launch_ms": dom_ms, "launch_model_gflop": flops_cls[dom] / 1e9,
                     "build_tflops": build_tflops,
                     "timing": f"HIP events on the launch stream around each class list, {args.steps} builds "
                               "replayed after the timed region with the class lists serialized on one stream"},
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = line["value"] / cpu["value"]
    line["stored_jk"] = stored_jk(hfx, N, basis, thr, args.steps)
    line["pcie_inclusive"] = pcie_inclusive(hfx, N, basis, arrs, thr, q_screened)
    if not args.no_extras:
        log("extras: SCF wall times")
        scf = {"benzene_ccpvdz_stored_fp64": stored_scf("benzene/cc-pVDZ", atoms, "cc-pVDZ")}
        if cpu:
            scf["benzene_ccpvdz_stored_fp64"]["cpu_eri_phase_estimate_s"] = q_screened / cpu["value"]
        log("extras: mixed vs FP64 direct SCF")
        scf["octane_ccpvdz_direct"] = direct_scf_pair("n-octane/cc-pVDZ", W.n_octane(), "cc-pVDZ")
        scf["water4_ccpvtz_direct"] = direct_scf_pair("(H2O)4/cc-pVTZ", W.water_tetramer(), "cc-pVTZ")
        line["direct_build"] = {"water4_ccpvtz": direct_build("(H2O)4/cc-pVTZ", W.water_tetramer(), "cc-pVTZ",
                                                              min(args.steps, 5)),
                                "octane_ccpvdz": direct_build("n-octane/cc-pVDZ", W.n_octane(), "cc-pVDZ",
                                                              min(args.steps, 5))}
        scf["o2_ccpvtz_uhf_triplet"] = uhf_scf("O2/cc-pVTZ UHF triplet", W.o2(), "cc-pVTZ", (9, 7))
        line["scf"] = scf
        log("extras: published-table molecules")
        line["readme_scf"] = readme_scf()
        log("extras: serving stream")
        line["serving"] = serving()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
