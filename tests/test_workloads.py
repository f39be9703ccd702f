"""The synthetic workload generators (CPU only): the serving stream's
requests have valid triclinic boxes at 0.1 atoms per cubic Angstrom with
K_d the smallest 2^a 3^b 5^c not below the box vector length, the uneven
frames follow their stated ranges, and the arrival times are a Poisson
process of the requested rate."""
import numpy as np

from ewaldcorrel_amd import workloads as wl


def is_smooth(k):
    for p in (2, 3, 5):
        while k % p == 0:
            k //= p
    return k == 1


def test_smooth_235():
    assert [wl.smooth_235(n) for n in (1, 7, 13, 31, 97, 101)] == [1, 8, 15, 32, 100, 108]
    for n in range(1, 400):
        k = wl.smooth_235(n)
        assert k >= n and is_smooth(k)
        assert all(not is_smooth(m) for m in range(n, k))


def test_serving_requests_follow_their_definition():
    rng = np.random.default_rng(0)
    for _ in range(40):
        sy = wl.serving_request(rng, 1000, 500000)
        assert 1000 <= sy.n <= 500002 and sy.n % 2 == 0
        L = np.array([[sy.L6[0], 0, 0], [sy.L6[3], sy.L6[1], 0], [sy.L6[4], sy.L6[5], sy.L6[2]]])
        lengths = np.linalg.norm(L, axis=1)
        assert np.allclose(lengths, lengths[0], rtol=1e-12)
        vol = abs(np.linalg.det(L))
        assert abs(sy.n / vol - 0.1) < 1e-9
        al, be, ga = sy.meta["angles"]
        assert all(70.0 <= a <= 110.0 for a in (al, be, ga))
        k = sy.K[0]
        assert sy.K == (k, k, k) and is_smooth(k) and k >= lengths[0] > k / 1.25
        assert sy.order in (4, 6) and sy.meta["precision"] in ("fp32", "fp64")
        s = sy.x @ np.linalg.inv(L)  # fractional coordinates in [0, 1)
        assert s.min() >= -1e-12 and s.max() < 1 + 1e-12
        assert abs(sy.w.sum()) == 0.0


def test_serving_stream_arrivals_are_poisson():
    reqs, t = wl.serving_stream(2000, seed=5, rate=2000.0, nmin=100, nmax=200)
    gaps = np.diff(t)
    assert t[0] == 0.0 and np.all(gaps >= 0)
    assert abs(gaps.mean() - 1 / 2000.0) < 0.1 / 2000.0
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1  # exponential: cv = 1
    assert len(reqs) == 2000


def test_uneven_frames_follow_their_definition():
    frames = wl.uneven_frames(64, seed=4)
    for f in frames:
        assert 500 <= f.n <= 20002 and f.n % 2 == 0
        Lb = f.L6[0]
        assert abs(f.n / Lb ** 3 - 0.1) < 1e-9
        k = f.K[0]
        assert f.K == (k, k, k) and 32 <= k <= 96 and is_smooth(k)
        assert f.order == 4
