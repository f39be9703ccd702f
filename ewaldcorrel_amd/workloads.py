"""Synthetic SPME workloads (numpy default_rng(seed); max_m = 1.0 as in the
reference's ewald.py:122 / test.py:86; Ewald kernel with eta = beta).

  point_charges_1k   1,000 alternating +-1 charges, 30 A cube, 32^3, order 4, beta 0.35
  water_98k          32,768 rigid 3-site waters (98,304 atoms), 100 A cube, 128^3,
                     order 4, beta 0.32
  gaussian_1m        1,048,576 N(0, 0.5^2) charges shifted to net zero,
                     200 x 200 x 400 A, 192 x 192 x 384, order 6, beta 0.30
  uneven_frames      256 frames, N log-uniform in [500, 20000], density 0.1/A^3,
                     K = smallest 2^a 3^b 5^c >= L/1A clamped to [32, 96], order 4
  serving_stream     requests arriving as a Poisson process (2,000/s): N log-uniform
                     in [1000, 500000], triclinic boxes (angles uniform in 70-110
                     deg) at density 0.1/A^3, K_d = smallest 2^a 3^b 5^c >= |L_d|/1A,
                     order 4 or 6 and FP32 or FP64 with probability 1/2 each
"""
from __future__ import annotations

from dataclasses import dataclass, field
from math import cos, radians, sin

import numpy as np


@dataclass
class System:
    name: str
    L6: np.ndarray
    K: tuple
    order: int
    beta: float
    w: np.ndarray
    x: np.ndarray
    max_m: float = 1.0
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.w)


def smooth_235(n: int) -> int:
    """Smallest integer >= n whose only prime factors are 2, 3, 5."""
    m = max(1, int(n))
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def point_charges_1k(seed: int = 1) -> System:
    rng = np.random.default_rng(seed)
    n = 1000
    x = rng.uniform(0, 30, (n, 3))
    w = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return System("point_charges_1k", np.array([30., 30, 30, 0, 0, 0]), (32, 32, 32), 4,
                  0.35, w, x)


def water(nmol: int, box: float, K, seed: int = 2, name: str = "water") -> System:
    """Rigid TIP3P-like waters: O-H 0.9572 A, HOH 104.52 deg, uniform random
    orientation (unit quaternion), centres uniform in the box; order O,H,H."""
    rng = np.random.default_rng(seed)
    centre = rng.uniform(0, box, (nmol, 3))
    q = rng.normal(size=(nmol, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    a, b, c, d = q.T
    R = np.stack([
        np.stack([a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)], -1),
        np.stack([2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)], -1),
        np.stack([2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d], -1),
    ], 1)
    r, th = 0.9572, radians(104.52)
    body = np.array([[0, 0, 0], [r * sin(th / 2), 0, r * cos(th / 2)],
                     [-r * sin(th / 2), 0, r * cos(th / 2)]])
    x = centre[:, None, :] + np.einsum("mij,aj->mai", R, body)
    w = np.tile([-0.834, 0.417, 0.417], nmol)
    return System(name, np.array([box, box, box, 0, 0, 0]), tuple(K), 4, 0.32, w,
                  x.reshape(-1, 3))


def water_98k(seed: int = 2) -> System:
    return water(32768, 100.0, (128, 128, 128), seed, "water_98k")


def gaussian_charges(n: int, box, K, order: int = 6, beta: float = 0.30, seed: int = 3,
                     name: str = "gaussian") -> System:
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.5, n)
    w -= w.mean()
    x = rng.uniform(0, 1, (n, 3)) * np.asarray(box)
    return System(name, np.array([box[0], box[1], box[2], 0, 0, 0]), tuple(K), order, beta, w,
                  x)


def gaussian_1m(seed: int = 3) -> System:
    return gaussian_charges(1 << 20, (200., 200., 400.), (192, 192, 384), 6, 0.30, seed,
                            "gaussian_1m")


def uneven_frames(nframes: int = 256, seed: int = 4, nmin: int = 500, nmax: int = 20000):
    rng = np.random.default_rng(seed)
    out = []
    for f in range(nframes):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        n += n % 2
        Lb = (n / 0.1) ** (1 / 3)
        k = min(max(smooth_235(int(np.ceil(Lb))), 32), 96)
        x = rng.uniform(0, Lb, (n, 3))
        w = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
        out.append(System(f"frame{f}", np.array([Lb, Lb, Lb, 0, 0, 0]), (k, k, k), 4, 0.32, w,
                          x))
    return out


def triclinic_L6(a, b, c, alpha, beta_, gamma):
    """Lower-triangular box from lengths and angles (degrees)."""
    al, be, ga = radians(alpha), radians(beta_), radians(gamma)
    bx, by = b * cos(ga), b * sin(ga)
    cx = c * cos(be)
    cy = c * (cos(al) - cos(be) * cos(ga)) / sin(ga)
    cz2 = c * c - cx * cx - cy * cy
    if cz2 <= 0:
        raise ValueError("invalid cell angles")
    return np.array([a, by, np.sqrt(cz2), bx, cx, cy])


def serving_request(rng, nmin: int = 1000, nmax: int = 500000, name: str = "request"):
    """One request of the serving stream: N log-uniform in [nmin, nmax]
    (even, +-1 alternating charges), a triclinic box with all three vectors
    of length a and angles uniform in [70, 110] deg (resampled until valid)
    holding N atoms at 0.1 / A^3, K_d = smallest 2^a 3^b 5^c >= a / 1 A,
    order 4 or 6, FP32 or FP64; eta = 0.30, max_m = 1.0."""
    n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
    n += n % 2
    while True:
        al, be, ga = rng.uniform(70.0, 110.0, 3)
        ca, cb, cg = cos(radians(al)), cos(radians(be)), cos(radians(ga))
        vf = 1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg
        if vf > 1e-3:
            break
    a = (n / 0.1 / np.sqrt(vf)) ** (1 / 3)
    L6 = triclinic_L6(a, a, a, al, be, ga)
    k = smooth_235(int(np.ceil(a)))
    s = rng.uniform(0, 1, (n, 3))
    Lm = np.array([[L6[0], 0., 0.], [L6[3], L6[1], 0.], [L6[4], L6[5], L6[2]]])
    x = s @ Lm  # rows of Lm are the box vectors: r = L^T s
    w = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    order = 4 if rng.uniform() < 0.5 else 6
    precision = "fp32" if rng.uniform() < 0.5 else "fp64"
    return System(name, L6, (k, k, k), order, 0.30, w, x,
                  meta={"precision": precision, "angles": (al, be, ga)})


def serving_stream(count: int, seed: int = 5, rate: float = 2000.0, nmin: int = 1000,
                   nmax: int = 500000):
    """`count` requests of the serving stream and their Poisson arrival
    times (seconds from the first, exponential gaps of mean 1/rate)."""
    rng = np.random.default_rng(seed)
    reqs = [serving_request(rng, nmin, nmax, f"request{i}") for i in range(count)]
    gaps = rng.exponential(1.0 / rate, count)
    gaps[0] = 0.0
    return reqs, np.cumsum(gaps)
