"""Synthetic molecular inputs of the benchmark configurations (bohr) and the
algorithmic flop model used for roofline reporting.

Geometries:
  h2o()             O (0,0,0), H (+-1.430, 1.107, 0)                       (H2O, 7 fns in STO-3G)
  benzene()         D6h ring, C-C 2.637, C-H 2.047 (H at 4.684 from centre)  (K = 120 in cc-pVDZ)
  n_octane(seed)    all-anti C8H18, C-C 2.90, C-H 2.06, tetrahedral, N(0, 0.02) jitter (K = 210)
  water_tetramer(seed)  cyclic (H2O)4, O-O 5.3, O-H 1.81, HOH 104.5, N(0, 0.05) jitter (K = 260 cc-pVTZ)
  o2()              O2, R = 2.28 (UHF triplet, 9 alpha / 7 beta)
"""
import math

import numpy as np

TETRA = math.radians(109.4712206)


def h2o():
    return [("O", (0.0, 0.0, 0.0)), ("H", (1.430, 1.107, 0.0)), ("H", (-1.430, 1.107, 0.0))]


def o2():
    return [("O", (-1.14, 0.0, 0.0)), ("O", (1.14, 0.0, 0.0))]


def benzene():
    at = []
    for r, sym in ((2.637, "C"), (4.684, "H")):
        for k in range(6):
            a = math.radians(60 * k)
            at.append((sym, (r * math.cos(a), r * math.sin(a), 0.0)))
    return at


def _unit(v):
    return v / np.linalg.norm(v)


def n_octane(seed=1, cc=2.90, ch=2.06, jitter=0.02):
    n = 8
    half = TETRA / 2
    C = [np.array([k * cc * math.sin(half), (0.5 if k % 2 else -0.5) * cc * math.cos(half), 0.0]) for k in range(n)]
    z = np.array([0.0, 0.0, 1.0])
    H = []
    for k in range(n):
        nb = [C[j] for j in (k - 1, k + 1) if 0 <= j < n]
        us = [_unit(p - C[k]) for p in nb]
        if len(us) == 2:   # CH2: H-C-H in the plane bisecting the C-C-C angle
            bis = _unit(-(us[0] + us[1]))
            for s in (1, -1):
                H.append(C[k] + ch * (math.cos(TETRA / 2) * bis + s * math.sin(TETRA / 2) * z))
        else:              # CH3: three H at the tetrahedral angle to the C-C bond
            u = us[0]
            e1 = _unit(np.cross(u, z)) if abs(u[2]) < 0.9 else _unit(np.cross(u, [1.0, 0.0, 0.0]))
            e2 = np.cross(u, e1)
            for phi in (0.0, 2 * math.pi / 3, 4 * math.pi / 3):
                d = math.cos(TETRA) * u + math.sin(TETRA) * (math.cos(phi) * e1 + math.sin(phi) * e2)
                H.append(C[k] + ch * d)
    rng = np.random.default_rng(seed)
    atoms = [("C", c) for c in C] + [("H", h) for h in H]
    return [(s, tuple(float(x) for x in (p + rng.normal(0.0, jitter, 3)))) for s, p in atoms]


def water_tetramer(seed=4, oo=5.3, oh=1.81, hoh=104.5, jitter=0.05):
    half = oo / 2
    O = [np.array(v) for v in ((half, half, 0.0), (-half, half, 0.0), (-half, -half, 0.0), (half, -half, 0.0))]
    z = np.array([0.0, 0.0, 1.0])
    atoms = []
    for k in range(4):
        u = _unit(O[(k + 1) % 4] - O[k])            # H-bond donor along the ring edge
        s = 1.0 if k % 2 == 0 else -1.0              # free H alternately above / below
        w = _unit(s * z)
        a = math.radians(hoh)
        h1 = O[k] + oh * u
        h2 = O[k] + oh * (math.cos(a) * u + math.sin(a) * w)
        atoms += [("O", O[k]), ("H", h1), ("H", h2)]
    rng = np.random.default_rng(seed)
    return [(sym, tuple(float(x) for x in (p + rng.normal(0.0, jitter, 3)))) for sym, p in atoms]


def nuclear_repulsion(atoms):
    from .basis import atomic_number
    e = 0.0
    for i in range(len(atoms)):
        for j in range(i + 1, len(atoms)):
            r = math.dist(atoms[i][1], atoms[j][1])
            e += atomic_number(atoms[i][0]) * atomic_number(atoms[j][0]) / r
    return e


# ---------------------------------------------------------------- flop model
PAIR_L = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2), (3, 0), (3, 1), (3, 2), (3, 3)]


def qclass_l(qc):
    pb = 0
    while (pb + 1) * (pb + 2) // 2 <= qc:
        pb += 1
    pk = qc - pb * (pb + 1) // 2
    return PAIR_L[pb] + PAIR_L[pk]


def qclass_name(qc):
    return "(%s%s|%s%s)" % tuple("spdf"[l] for l in qclass_l(qc))


def _cart(l):
    return [(a, b, l - a - b) for a in range(l + 1) for b in range(l - a + 1)]


def _H(n):
    return (n + 1) * (n + 2) * (n + 3) // 6


def model_flops(a, b, c, d):
    """Algorithmic flops of one primitive shell quartet (ab|cd): prefactors,
    Boys, all R^{(m)}_tuv, then the ket and bra Hermite->Cartesian steps."""
    lab, lcd = a + b, c + d
    L = lab + lcd
    f = 40 + 30 + 3 * (L + 1) + 3 * math.comb(L + 4, 4)
    f += 2 * _H(lab) * sum(math.prod(x + y + 1 for x, y in zip(cc, dd)) for cc in _cart(c) for dd in _cart(d))
    f += 2 * len(_cart(c)) * len(_cart(d)) * sum(
        math.prod(x + y + 1 for x, y in zip(aa, bb)) for aa in _cart(a) for bb in _cart(b))
    return f


def kernel_family(qc, fp32=False):
    """Name of the kernel template that runs class qc (mirrors use_family_a,
    use_family_m and runs_family_a in csrc/device/eri_body.hpp)."""
    a, b, c, d = qclass_l(qc)
    L = a + b + c + d
    nout = len(_cart(a)) * len(_cart(b)) * len(_cart(c)) * len(_cart(d))
    fam_a = L <= 4 and nout + math.comb(L + 3, 3) <= 128
    ncd = len(_cart(c)) * len(_cart(d))
    fam_m = math.comb(c + d + 3, 3) >= 4 and ncd >= 3 and (not fam_a or (a, b, c, d) == (1, 1, 1, 1))
    if fam_m:
        return "k_fam_m"
    return "k_fam_a" if fam_a else "k_fam_b"
