"""Deterministic synthetic inputs of the benchmark configurations (species sorted heavy-first,
coordinates in Angstrom, reference padded layout).

C1  benzene, C-C 1.39 / C-H 1.09 A hexagon, N(0, 0.02 A) noise per coordinate
C2  linear alkanes CnH2n+2, n ~ U{2..30}, zig-zag C-C 1.54 A, tetrahedral C-H 1.09 A, N(0, 0.03 A)
C3/C4  water clusters (H2O)_N at 1 g/cm^3 in a sphere, O-O >= 2.6 A by random sequential
       addition, random orientations; O-H 0.9572 A, H-O-H 104.52 deg
"""
import math

import numpy as np

CC, CH = 1.54, 1.09
TET = math.radians(109.4712206)


def _pad(mols):
    K = max(len(s) for s, _ in mols)
    sp = np.zeros((len(mols), K), dtype=np.int64)
    xyz = np.zeros((len(mols), K, 3))
    for i, (s, x) in enumerate(mols):
        sp[i, :len(s)] = s
        xyz[i, :len(s)] = x
    return sp, xyz


def benzene(nmol=16, seed=0, noise=0.02):
    rng = np.random.default_rng(seed)
    ang = np.arange(6) * math.pi / 3
    ring = np.stack([np.cos(ang), np.sin(ang), np.zeros(6)], 1)
    base = np.concatenate([1.39 * ring, (1.39 + 1.09) * ring])
    sp = np.array([6] * 6 + [1] * 6)
    mols = [(sp, base + rng.normal(0.0, noise, base.shape)) for _ in range(nmol)]
    return _pad(mols)


def _carbon(k):
    return np.array([CC * math.sin(TET / 2) * k, CC * math.cos(TET / 2) * (k % 2), 0.0])


def _two_h(c, u1, u2):
    """Two tetrahedral H on carbon c whose other two bonds point along unit vectors u1, u2."""
    b = -(u1 + u2)
    b /= np.linalg.norm(b)
    n = np.cross(u1, u2)
    n /= np.linalg.norm(n)
    h = TET / 2
    return [c + CH * (math.cos(h) * b + math.sin(h) * n), c + CH * (math.cos(h) * b - math.sin(h) * n)]


def alkane(n):
    """CnH2n+2 zig-zag in the xy plane; CH2 hydrogens at +-z, one in-plane H per terminal CH3."""
    C = [_carbon(k) for k in range(n)]
    H = []
    unit = lambda v: v / np.linalg.norm(v)
    for k in range(n):
        nb = []
        if k > 0:
            nb.append(unit(C[k - 1] - C[k]))
        if k < n - 1:
            nb.append(unit(C[k + 1] - C[k]))
        if len(nb) == 1:  # terminal: in-plane H where the chain would continue
            virt = _carbon(k - 1) if k == 0 else _carbon(k + 1)
            u = unit(virt - C[k])
            H.append(C[k] + CH * u)
            nb.append(u)
        H.extend(_two_h(C[k], nb[0], nb[1]))
    sp = np.array([6] * n + [1] * len(H))
    return sp, np.array(C + H)


def alkanes(nmol=1024, seed=0, noise=0.03, nmin=2, nmax=30):
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=nmol)
    mols = []
    for n in ns:
        s, x = alkane(int(n))
        mols.append((s, x + rng.normal(0.0, noise, x.shape)))
    return _pad(mols)


def water_cluster(nw, seed=0, min_oo=2.6):
    """One (H2O)_nw cluster at 1 g/cm^3: sphere volume nw * 29.915 A^3."""
    rng = np.random.default_rng(seed)
    vol = nw * 18.015 / 6.02214076e23 * 1e24
    R = (3.0 * vol / (4.0 * math.pi)) ** (1.0 / 3.0)
    O = np.zeros((0, 3))
    tries = 0
    while len(O) < nw:
        tries += 1
        if tries > 2000000:
            raise RuntimeError("random sequential addition did not converge")
        p = rng.uniform(-R, R, 3)
        if p @ p > R * R:
            continue
        if len(O) and np.min(np.sum((O - p) ** 2, 1)) < min_oo ** 2:
            continue
        O = np.vstack([O, p])
    half = math.radians(104.52) / 2
    loc = 0.9572 * np.array([[math.sin(half), math.cos(half), 0.0], [-math.sin(half), math.cos(half), 0.0]])
    q = rng.normal(size=(nw, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    H = []
    for i in range(nw):
        a, b, c, d = q[i]
        Rm = np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                       [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                       [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])
        H.extend(O[i] + loc @ Rm.T)
    sp = np.array([8] * nw + [1] * (2 * nw))
    return sp, np.vstack([O, np.array(H)])


def water_clusters(nw=256, nmol=8, seed=0):
    return _pad([water_cluster(nw, seed + k) for k in range(nmol)])


CONFIGS = {
    "C1": dict(method="PM3", make=lambda seed=0: benzene(16, seed)),
    "C2": dict(method="AM1", make=lambda seed=0: alkanes(1024, seed)),
    "C3": dict(method="PM3", make=lambda seed=0: water_clusters(256, 8, seed)),
    "C4": dict(method="PM3", make=lambda seed=0: water_clusters(1024, 1, seed)),
}


def _rotate(x, rng):
    q = rng.normal(size=4)
    a, b, c, d = q / np.linalg.norm(q)
    Rm = np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                   [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                   [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])
    return x @ Rm.T


def serving_stream(rate=2000.0, window=10.0, seed=0, water_min=4, water_max=256, alkane_max=30, pool=48,
                   noise=0.02):
    """Open-loop Poisson request stream: exponential inter-arrival times of mean 1/rate over
    `window` seconds; half AM1 alkane single points (n ~ U{2..alkane_max}), half PM3 water clusters
    (N ~ U{water_min..water_max}).  Water clusters come from a pool of `pool` random-sequential-
    addition clusters (sizes drawn once), each request re-oriented at random and perturbed by
    N(0, noise); returns [(t_arrival, species, coords, method)] sorted by arrival."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * window * 1.2) + 16))
    t = t[t < window]
    sizes = rng.integers(water_min, water_max + 1, size=pool)
    wpool = [water_cluster(int(n), seed=1000 + k) for k, n in enumerate(sizes)]
    apool = {n: alkane(n) for n in range(2, alkane_max + 1)}
    out = []
    for ta in t:
        if rng.random() < 0.5:
            s, x = apool[int(rng.integers(2, alkane_max + 1))]
            out.append((float(ta), s, x + rng.normal(0.0, 0.03, x.shape), "AM1"))
        else:
            s, x = wpool[int(rng.integers(0, pool))]
            out.append((float(ta), s, _rotate(x, rng) + rng.normal(0.0, noise, x.shape), "PM3"))
    return out
