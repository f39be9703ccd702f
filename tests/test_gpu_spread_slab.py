"""Spreading fused with the forward y-z slab transform (one workgroup per x
plane holding the plane's fixed-point sums in LDS) against the separate
spreading tiles followed by the slab transform: the same integer
contributions in the same sums, then the same transform compiled for
another workgroup shape (its rounding may differ where the compiler fuses
other multiply-adds), so the evaluation and get_S agree to the precision's
rounding; the fused path repeats bitwise, and meets the oracle at the usual
tolerances.  It applies with fixed-point sums when a plane of sums fits in
LDS and the grid has at least `order` cells per axis."""
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


@pytest.fixture(autouse=True)
def no_tuned_shapes():
    """The tests choose their paths: no launch shapes remembered from other
    tests (process-wide records) apply to their grids."""
    ew().tune_clear()
    yield


def make_sys(L6, K, order, N, seed, beta=0.4):
    rng = np.random.default_rng(seed)
    L6 = np.asarray(L6, dtype=float)
    x = rng.uniform(-0.2, 1.2, (N, 3)) @ LofS(L6)
    w = rng.normal(size=N)
    w -= w.mean()
    return dict(L6=L6, K=tuple(K), order=order, beta=beta, w=w, x=x)


GRIDS = [
    # name, L6, K, order, N
    ("cube_128_order4", [40., 40., 40., 0., 0., 0.], (128, 128, 128), 4, 6000),
    ("triclinic_odd_K1", [20., 18., 22., 3., -2., 1.5], (24, 25, 30), 6, 700),
    ("order8", [25., 26., 24., 0., 1., 0.], (32, 30, 36), 8, 900),
    ("order10_generic", [22., 20., 24., 1., 0., -1.], (30, 32, 30), 10, 600),
    ("long_z", [20., 20., 60., 0., 0., 0.], (16, 16, 384), 4, 500),
    ("K_equals_order", [8., 9., 10., 0., 0., 0.], (4, 4, 4), 4, 40),
]


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("grid", GRIDS, ids=[g[0] for g in GRIDS])
def test_fused_spreading_equals_tiles(grid, precision):
    name, L6, K, order, N = grid
    sy = make_sys(L6, K, order, N, 3)
    s = ew().Sfac(sy["L6"], np.array(K), order, precision=precision, fixed=True)
    s.set_A(ew().ewald_kernel(sy["beta"]), 1.0)
    info = s.info()
    say(name, precision, "slab", info["slab"])
    assert info["slab"] == 2
    a = s.evaluate(sy["w"], sy["x"])
    a2 = s.evaluate(sy["w"], sy["x"])  # graph replay
    for u, v in zip(a, a2):
        assert np.array_equal(np.asarray(u), np.asarray(v))
    s.sfac(sy["w"], sy["x"])
    Sa = s.get_S().copy()
    s.set_transform(2)
    assert s.info()["slab"] == 1
    b = s.evaluate(sy["w"], sy["x"])
    s.sfac(sy["w"], sy["x"])
    Sb = s.get_S().copy()
    tol = 1e-6 if precision == "fp32" else 1e-13
    eS = np.abs(Sa - Sb).max() / np.abs(Sb).max()
    say("  fused vs tiles: E", abs(a[0] - b[0]) / abs(b[0]), "vir", rel(a[1], b[1]),
        "F", rel_rms(a[2], b[2]), "phi", rel_rms(a[3], b[3]), "S", eS)
    assert abs(a[0] - b[0]) <= tol * abs(b[0])
    assert rel(a[1], b[1]) <= 10 * tol
    assert rel_rms(a[2], b[2]) <= 10 * tol
    assert rel_rms(a[3], b[3]) <= 10 * tol
    # S-hat carries 1 / |b|^2 (up to ~1e4 at the Nyquist points of order 10),
    # which scales the rounding of the spline weights: the fused kernel forms
    # its one x weight alone, the tiles every weight, rounded in other places
    assert eS <= (5e-5 if precision == "fp32" else 1e-12)
    # and against the oracle
    o = OracleSfac(sy["L6"], K, order)
    o.set_A(ewald_rad(sy["beta"]), 1.0)
    E0, v0, dx0, phi0 = o.evaluate(sy["w"], sy["x"])
    tE, tF = (1e-6, 1e-5) if precision == "fp32" else (1e-10, 1e-9)
    say("  oracle E", abs(a[0] - E0) / abs(E0), "vir", rel(a[1], v0), "F", rel_rms(a[2], dx0))
    assert abs(a[0] - E0) <= tE * abs(E0)
    assert rel(a[1], v0) <= 10 * tE
    assert rel_rms(a[2], dx0) <= tF
    assert rel_rms(a[3], phi0) <= tF


def test_fused_spreading_in_a_batch():
    """Frames of one batch through the fused launch (one workgroup per frame
    and plane, planes handed to the XCDs in runs): every frame bitwise as it
    evaluates alone, with a frame on the tile path (transform 2) and a frame
    of another grid in the same batch, over a capture and two replays."""
    dev = torch.device("cuda:0")
    systems = [make_sys([30., 31., 29., 0., 0., 0.], (48, 48, 48), 4, 2000, 20 + f)
               for f in range(3)]
    systems.append(make_sys([21., 19., 23., 2.5, -1.5, 1.0], (25, 27, 32), 4, 1500, 30))
    handles, bufs = [], []
    for f, sy in enumerate(systems):
        h = ew().Sfac(sy["L6"], np.array(sy["K"]), 4, precision="fp32")
        h.set_A(ew().ewald_kernel(sy["beta"]), 1.0)
        if f == 2:
            h.set_transform(2)
        assert h.info()["slab"] == (1 if f == 2 else 2)
        handles.append(h)
        n = len(sy["w"])
        bufs.append(dict(w=torch.tensor(sy["w"], device=dev), x=torch.tensor(sy["x"], device=dev),
                         out=torch.zeros(8, dtype=torch.float64, device=dev),
                         dx=torch.zeros((n, 3), dtype=torch.float64, device=dev),
                         phi=torch.zeros(n, dtype=torch.float64, device=dev)))
    alone = [h.evaluate(sy["w"], sy["x"]) for h, sy in zip(handles, systems)]
    batch = ew().SfacBatch(handles)
    for f, (sy, b) in enumerate(zip(systems, bufs)):
        batch.set(f, len(sy["w"]), b["w"].data_ptr(), b["x"].data_ptr(), b["out"].data_ptr(),
                  b["dx"].data_ptr(), b["phi"].data_ptr())
    torch.cuda.synchronize()
    for rep in range(3):
        batch.eval(None)
        batch.sync()
        torch.cuda.synchronize()
        for f, (a, b) in enumerate(zip(alone, bufs)):
            o = b["out"].cpu().numpy()
            assert o[0] == a[0] and np.array_equal(o[1:7], a[1]), (rep, f)
            assert np.array_equal(b["dx"].cpu().numpy(), a[2]), (rep, f)
            assert np.array_equal(b["phi"].cpu().numpy(), a[3]), (rep, f)
    batch.close()


@pytest.mark.parametrize("precision,fixed,K,order,expect", [
    ("fp32", True, (128, 128, 128), 6, 2),   # 128 x 133 x 8 B of sums
    ("fp64", True, (128, 128, 128), 4, 2),   # FP64 slab (133 KB) inside the sums (134 KB)
    ("fp64", False, (32, 32, 32), 4, 1),     # FP64 sums: no fixed point, tiles
    ("fp32", True, (160, 160, 160), 4, 1),   # 160 x 163 x 8 B of sums > LDS
    ("fp32", True, (3, 8, 8), 4, 1),         # fewer x planes than the order
    ("fp32", True, (16, 96, 192), 4, 1),     # 97 x 96 plane elements > 1024 x 9
])
def test_fused_spreading_availability(precision, fixed, K, order, expect):
    L6 = np.array([30., 30., 30., 0., 0., 0.])
    s = ew().Sfac(L6, np.array(K), order, precision=precision, fixed=fixed)
    info = s.info()
    say(precision, fixed, K, order, "slab", info["slab"])
    assert info["slab"] == expect


def test_batch_mixing_128_planes_with_other_grids():
    """A 128 x 128 x 128 frame (alone: the register-line slab transforms and
    their fused spreading) batched with a 48^3 frame: the launch runs the
    LDS-staged slab kernels for both, so each frame matches its evaluation
    alone to FP32 rounding (the transforms round in another order), and the
    oracle."""
    dev = torch.device("cuda:0")
    systems = [make_sys([40., 40., 40., 0., 0., 0.], (128, 128, 128), 4, 6000, 41),
               make_sys([30., 31., 29., 0., 0., 0.], (48, 48, 48), 4, 2000, 42)]
    handles, bufs = [], []
    for sy in systems:
        h = ew().Sfac(sy["L6"], np.array(sy["K"]), 4, precision="fp32")
        h.set_A(ew().ewald_kernel(sy["beta"]), 1.0)
        assert h.info()["slab"] == 2
        handles.append(h)
        n = len(sy["w"])
        bufs.append(dict(w=torch.tensor(sy["w"], device=dev), x=torch.tensor(sy["x"], device=dev),
                         out=torch.zeros(8, dtype=torch.float64, device=dev),
                         dx=torch.zeros((n, 3), dtype=torch.float64, device=dev),
                         phi=torch.zeros(n, dtype=torch.float64, device=dev)))
    alone = [h.evaluate(sy["w"], sy["x"]) for h, sy in zip(handles, systems)]
    batch = ew().SfacBatch(handles)
    for f, (sy, b) in enumerate(zip(systems, bufs)):
        batch.set(f, len(sy["w"]), b["w"].data_ptr(), b["x"].data_ptr(), b["out"].data_ptr(),
                  b["dx"].data_ptr(), b["phi"].data_ptr())
    torch.cuda.synchronize()
    for rep in range(2):
        batch.eval(None)
        batch.sync()
        torch.cuda.synchronize()
        for f, (a, b) in enumerate(zip(alone, bufs)):
            o = b["out"].cpu().numpy()
            dx = b["dx"].cpu().numpy()
            say("mixed batch", f, rep, abs(o[0] - a[0]) / abs(a[0]), rel_rms(dx, a[2]))
            assert abs(o[0] - a[0]) <= 1e-6 * abs(a[0]), (rep, f)
            assert rel(o[1:7], a[1]) <= 1e-5
            assert rel_rms(dx, a[2]) <= 1e-5
            assert rel_rms(b["phi"].cpu().numpy(), a[3]) <= 1e-5
    for f, sy in enumerate(systems):
        o = OracleSfac(sy["L6"], sy["K"], 4)
        o.set_A(ewald_rad(sy["beta"]), 1.0)
        E0, v0, dx0, phi0 = o.evaluate(sy["w"], sy["x"])
        E = bufs[f]["out"][0].item()
        assert abs(E - E0) <= 1e-6 * abs(E0)
        assert rel_rms(bufs[f]["dx"].cpu().numpy(), dx0) <= 1e-5
    batch.close()
