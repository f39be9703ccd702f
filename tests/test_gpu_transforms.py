"""The two transform paths of the engine: y-z slab transforms (a whole
x plane in LDS per workgroup) and separate z and y line passes.  Both
implement the same r2c/c2r 3-D transform, so an evaluation, the split entry
points and get_S must agree between them and with the oracle; the slab is
offered exactly when the plane fits in the 160 KB of LDS and K1, K2 factor
into 2, 3 and 5."""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from conftest import rel, rel_rms  # noqa: E402
from oracle.sfac_oracle import LofS, OracleSfac, ewald_rad  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ew():
    from ewaldcorrel_amd import ewald
    return ewald


def make_sys(L6, K, order, N, seed, beta=0.4):
    rng = np.random.default_rng(seed)
    L6 = np.asarray(L6, dtype=float)
    x = rng.uniform(-0.2, 1.2, (N, 3)) @ LofS(L6)
    w = rng.normal(size=N)
    w -= w.mean()
    return dict(L6=L6, K=tuple(K), order=order, beta=beta, w=w, x=x)


GRIDS = [
    # name, L6, K, order, N
    ("odd_K1_triclinic", [20., 18., 22., 3., -2., 1.5], (24, 25, 30), 6, 700),
    ("radix3_5", [30., 31., 33., 0., 0., 0.], (45, 48, 50), 4, 2000),
    ("long_z", [20., 20., 60., 0., 0., 0.], (16, 16, 384), 4, 500),
    ("odd_K1_K2", [12., 11., 13., 0., 0.5, 0.], (20, 15, 27), 4, 300),
    ("tiny_K", [5., 6., 7., 0., 0., 0.], (2, 3, 1), 4, 10),
    # 147 x-pass partials: their sum (finalized inside the 1024-thread slab
    # workgroup) spans several wavefronts
    ("many_x_blocks", [40., 44., 46., 2., 1., -3.], (48, 96, 96), 4, 3000),
    # 96 x 97 = 9312 plane elements: the slab variant for large planes
    ("large_slab_variant", [15., 40., 60., 1., -2., 3.], (16, 96, 192), 4, 1500),
    # 128 x 128 planes: FP32 runs the register-line slab kernels
    ("planes_128", [30., 34., 32., 2., 0., -1.], (24, 128, 128), 4, 3000),
]


def evaluate(s, sy):
    return s.evaluate(sy["w"], sy["x"])


@pytest.mark.parametrize("precision,fixed", [("fp64", False), ("fp64", True), ("fp32", True)])
@pytest.mark.parametrize("grid", GRIDS, ids=[g[0] for g in GRIDS])
def test_slab_equals_line_passes(grid, precision, fixed):
    name, L6, K, order, N = grid
    sy = make_sys(L6, K, order, N, 5)
    s = ew().Sfac(sy["L6"], np.array(K), order, precision=precision, fixed=fixed)
    s.set_A(ew().ewald_kernel(sy["beta"]), 1.0)
    info = s.info()
    say(name, precision, fixed, "slab", info["slab"], "available", info["slab_available"])
    assert info["slab_available"] == 1 and info["slab"] >= 1
    a = evaluate(s, sy)
    s.set_transform(1)
    assert s.info()["slab"] == 0
    b = evaluate(s, sy)
    s.set_transform(0)
    assert s.info()["slab"] == info["slab"]
    c = evaluate(s, sy)
    tol = 1e-13 if precision == "fp64" else 1e-6
    say("  E", abs(a[0] - b[0]) / abs(b[0]), "vir", rel(a[1], b[1]), "F", rel_rms(a[2], b[2]),
        "phi", rel_rms(a[3], b[3]))
    assert abs(a[0] - b[0]) <= tol * abs(b[0])
    assert rel(a[1], b[1]) <= 10 * tol
    assert rel_rms(a[2], b[2]) <= 10 * tol
    assert rel_rms(a[3], b[3]) <= 10 * tol
    # switching back runs the slab path again (the captured graph is rebuilt);
    # with deterministic spreading the result is bitwise the same
    if fixed:
        for u, v in zip(a, c):
            assert np.array_equal(np.asarray(u), np.asarray(v))


def test_fp32_largest_slab_variant():
    """FP32 planes of more than 512 x 20 elements run the 512-thread variant
    with 32 elements per thread (128 x 97 = 12,416 here; FP64 planes this
    large do not fit in LDS): slab and line passes agree and meet the
    oracle."""
    sy = make_sys([10., 40., 60., 0.5, -1., 2.], (8, 128, 192), 4, 2500, 7)
    s = ew().Sfac(sy["L6"], np.array(sy["K"]), 4, precision="fp32")
    s.set_A(ew().ewald_kernel(sy["beta"]), 1.0)
    assert s.info()["slab"] >= 1
    a = evaluate(s, sy)
    s.set_transform(1)
    b = evaluate(s, sy)
    o = OracleSfac(sy["L6"], sy["K"], 4)
    o.set_A(ewald_rad(sy["beta"]), 1.0)
    E0, v0, dx0, phi0 = o.evaluate(sy["w"], sy["x"])
    say("largest slab fp32: E", abs(a[0] - b[0]) / abs(b[0]), "oracle", abs(a[0] - E0) / abs(E0),
        "F", rel_rms(a[2], dx0))
    assert abs(a[0] - b[0]) <= 1e-6 * abs(b[0]) and rel_rms(a[2], b[2]) <= 1e-5
    assert abs(a[0] - E0) <= 1e-6 * abs(E0) and rel(a[1], v0) <= 1e-5
    assert rel_rms(a[2], dx0) <= 1e-5 and rel_rms(a[3], phi0) <= 1e-5


@pytest.mark.parametrize("mode", [0, 1], ids=["slab", "lines"])
def test_split_api_and_get_S_against_oracle(mode):
    name, L6, K, order, N = GRIDS[0]
    sy = make_sys(L6, K, order, N, 11)
    s = ew().Sfac(sy["L6"], np.array(K), order)
    s.set_transform(mode)
    s.set_A(ew().ewald_kernel(sy["beta"]), 1.0)
    o = OracleSfac(sy["L6"], K, order)
    o.set_A(ewald_rad(sy["beta"]), 1.0)
    s.sfac(sy["w"], sy["x"])
    o.sfac(sy["w"], sy["x"])
    S, Sref = s.get_S(), o.get_S()
    errS = np.abs(S - Sref).max() / np.abs(Sref).max()
    vir = np.zeros(6)
    E = s.de1(vir)
    Eref, vref = o.de1()
    n = N
    dx = np.zeros((n, 3))
    s.de2(n, sy["w"], sy["x"], dx)
    dref = o.de2(sy["w"], sy["x"])
    say("mode", mode, "S", errS, "E", abs(E - Eref) / abs(Eref), "F", rel_rms(dx, dref))
    assert errS < 1e-12
    assert abs(E - Eref) <= 1e-10 * abs(Eref)
    assert rel(vir, vref) <= 1e-10
    assert rel_rms(dx, dref) <= 1e-9


@pytest.mark.parametrize("precision,K,available", [
    ("fp64", (128, 128, 128), 1),
    ("fp32", (128, 128, 128), 1),
    ("fp32", (160, 160, 160), 1),   # 160 x 81 x 8 B = 104 KB
    ("fp64", (160, 160, 160), 0),   # 160 x 81 x 16 B = 207 KB > 160 KB
    ("fp64", (12, 13, 14), 0),      # radix 13 and 7
    ("fp64", (14, 12, 10), 1),      # only K1, K2 matter: x runs its own pass
])
def test_slab_availability(precision, K, available):
    L6 = np.array([30., 30., 30., 0., 0., 0.])
    s = ew().Sfac(L6, np.array(K), 4, precision=precision)
    info = s.info()
    say(precision, K, info)
    assert info["slab_available"] == available
    assert (info["slab"] >= 1) == (available == 1)
