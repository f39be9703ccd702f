"""Fused evaluation, captured-graph replay, influence-table modes and
batched frames.  Results must not depend on how a frame is launched: the
split entry points (sfac/de1/de2/potl), the fused evaluation, its graph
replay and a batch of uneven frames give the same numbers; in the FP32 and
fixed-point modes (deterministic int64 spreading, fixed-order reductions)
they are bitwise identical."""
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


def split_api(s, w, x):
    n = len(w)
    s.sfac(w, x)
    vir = np.zeros(6)
    E = s.de1(vir)
    dx = np.zeros((n, 3))
    s.de2(n, w, x, dx)
    phi = np.zeros(n)
    s.potl(n, x, phi)
    return E, vir, dx, phi


@pytest.mark.parametrize("precision,fixed", [("fp64", False), ("fp64", True), ("fp32", True)])
def test_split_entry_points_equal_fused_evaluation(precision, fixed):
    sy = make_sys([20., 18., 22., 3., -2., 1.5], (24, 25, 30), 4, 700, 1)
    s = ew().Sfac(sy["L6"], np.array(sy["K"]), 4, precision=precision, fixed=fixed)
    s.set_A(ew().ewald_kernel(sy["beta"]), 1.0)
    a = split_api(s, sy["w"], sy["x"])
    b = s.evaluate(sy["w"], sy["x"])
    tol = 1e-12 if precision == "fp64" else 1e-6
    say(precision, fixed, abs(a[0] - b[0]) / abs(a[0]), rel(a[1], b[1]), rel_rms(a[2], b[2]))
    assert abs(a[0] - b[0]) <= tol * abs(a[0])
    assert rel(a[1], b[1]) <= tol * 10
    assert rel_rms(a[2], b[2]) <= tol * 10
    assert rel_rms(a[3], b[3]) <= tol * 10


def test_graph_replay_tracks_inputs():
    """The captured graph is replayed with the same device pointers while
    the charges and positions in them change; every replay matches a
    kernel-by-kernel evaluation of the same inputs."""
    sy = make_sys([16., 17., 18., 0.5, 0.2, -0.3], (20, 24, 30), 6, 900, 2)
    dev = torch.device("cuda:0")
    n = len(sy["w"])
    s = ew().Sfac(sy["L6"], np.array(sy["K"]), 6, precision="fp32")
    s.set_A(ew().ewald_kernel(sy["beta"]), 1.0)
    ref = ew().Sfac(sy["L6"], np.array(sy["K"]), 6, precision="fp32")
    ref.set_A(ew().ewald_kernel(sy["beta"]), 1.0)
    ref.set_graph(False)
    w = torch.tensor(sy["w"], device=dev)
    x = torch.tensor(sy["x"], device=dev)
    out = torch.zeros(8, dtype=torch.float64, device=dev)
    dx = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    phi = torch.zeros(n, dtype=torch.float64, device=dev)
    rng = np.random.default_rng(3)
    for it in range(4):
        torch.cuda.synchronize()
        s.eval_dev(n, w.data_ptr(), x.data_ptr(), out.data_ptr(), dx.data_ptr(), phi.data_ptr())
        s.sync()
        E, vir, dxr, phr = ref.evaluate(w.cpu().numpy(), x.cpu().numpy())
        o = out.cpu().numpy()
        say("replay", it, o[0], E)
        assert o[0] == E and np.array_equal(o[1:7], vir)
        assert np.array_equal(dx.cpu().numpy(), dxr)
        assert np.array_equal(phi.cpu().numpy(), phr)
        # new inputs, same pointers (program's own eager work between replays)
        w.mul_(torch.tensor(rng.uniform(0.5, 1.5, n), device=dev))
        x.add_(torch.tensor(rng.normal(0, 0.3, (n, 3)), device=dev))
    assert s.info()["graph"] == 1


def test_table_modes():
    """The closed-form Ewald influence is used when it reproduces the full
    table (fine grid), and the table is read when aliased images matter
    (coarse grid of the reference's own test); both agree with the oracle."""
    from ewaldcorrel_amd.workloads import water
    wsy = water(2048, 40.0, (48, 48, 48), seed=5)
    s = ew().Sfac(wsy.L6, np.array(wsy.K), 4)
    s.set_A(ew().ewald_kernel(wsy.beta), 1.0)
    assert s.info()["table_mode"] == 1
    E, vir, dx, phi = s.evaluate(wsy.w, wsy.x)
    o = OracleSfac(wsy.L6, wsy.K, 4)
    o.set_A(ewald_rad(wsy.beta), 1.0)
    E0, v0, dx0, phi0 = o.evaluate(wsy.w, wsy.x)
    say("analytic", abs(E - E0) / abs(E0), rel(vir, v0), rel_rms(dx, dx0))
    assert abs(E - E0) <= 1e-10 * abs(E0) and rel(vir, v0) <= 1e-10
    assert rel_rms(dx, dx0) <= 1e-9 and rel_rms(phi, phi0) <= 1e-9

    sy = make_sys([9., 10., 11., 2., -0.1, 1.0], (12, 13, 14), 4, 30, 6, beta=0.5)
    s = ew().Sfac(sy["L6"], np.array(sy["K"]), 4)
    s.set_A(ew().ewald_kernel(0.5), 1.0)
    assert s.info()["table_mode"] == 0


def _batch_systems():
    from ewaldcorrel_amd.workloads import uneven_frames
    frames = uneven_frames(nframes=12, seed=7, nmin=300, nmax=6000)
    # a triclinic frame and an odd grid in the mix
    tri = make_sys([21., 19., 23., 2.5, -1.5, 1.0], (25, 27, 32), 4, 2500, 8, beta=0.32)
    return frames, tri


def _check_batch(systems, precision, fixed, oracle_frames, exact=None):
    """exact (default: fixed): the batch must reproduce each frame alone
    bitwise; otherwise within 1e-12 (FP64) or 1e-6 (FP32) relative."""
    exact = fixed if exact is None else exact
    dev = torch.device("cuda:0")
    handles, bufs = [], []
    for sy in systems:
        h = ew().Sfac(sy["L6"], np.array(sy["K"]), sy["order"], precision=precision, fixed=fixed)
        h.set_A(ew().ewald_kernel(sy["beta"]), 1.0)
        handles.append(h)
        n = len(sy["w"])
        bufs.append(dict(w=torch.tensor(sy["w"], device=dev), x=torch.tensor(sy["x"], device=dev),
                         out=torch.zeros(8, dtype=torch.float64, device=dev),
                         dx=torch.zeros((n, 3), dtype=torch.float64, device=dev),
                         phi=torch.zeros(n, dtype=torch.float64, device=dev)))
    # each frame alone first
    alone = [h.evaluate(sy["w"], sy["x"]) for h, sy in zip(handles, systems)]
    batch = ew().SfacBatch(handles)
    for f, (sy, b) in enumerate(zip(systems, bufs)):
        batch.set(f, len(sy["w"]), b["w"].data_ptr(), b["x"].data_ptr(), b["out"].data_ptr(),
                  b["dx"].data_ptr(), b["phi"].data_ptr())
    torch.cuda.synchronize()
    for rep in range(3):  # graph capture, then replays
        batch.eval(None)
        batch.sync()
        torch.cuda.synchronize()
        for f, (a, b) in enumerate(zip(alone, bufs)):
            o = b["out"].cpu().numpy()
            if exact:
                assert o[0] == a[0] and np.array_equal(o[1:7], a[1]), (rep, f)
                assert np.array_equal(b["dx"].cpu().numpy(), a[2]), (rep, f)
                assert np.array_equal(b["phi"].cpu().numpy(), a[3]), (rep, f)
            else:
                tol = 1e-12 if precision == "fp64" else 1e-6
                assert abs(o[0] - a[0]) <= tol * abs(a[0]), (rep, f)
                assert rel(o[1:7], a[1]) <= 10 * tol
                assert rel_rms(b["dx"].cpu().numpy(), a[2]) <= 10 * tol
                assert rel_rms(b["phi"].cpu().numpy(), a[3]) <= 10 * tol
    # and the batch agrees with the oracle
    for f in oracle_frames:
        sy = systems[f]
        o = OracleSfac(sy["L6"], sy["K"], sy["order"])
        o.set_A(ewald_rad(sy["beta"]), 1.0)
        E0, v0, dx0, phi0 = o.evaluate(sy["w"], sy["x"])
        E = bufs[f]["out"][0].item()
        dx = bufs[f]["dx"].cpu().numpy()
        tol = 1e-10 if precision == "fp64" else 1e-6
        say(precision, f, sy["K"], abs(E - E0) / abs(E0), rel_rms(dx, dx0))
        assert abs(E - E0) <= tol * abs(E0)
        assert rel_rms(dx, dx0) <= (1e-9 if precision == "fp64" else 1e-5)
    batch.close()


@pytest.mark.parametrize("precision,fixed", [("fp32", True), ("fp64", True), ("fp64", False)])
def test_batch_equals_frames_alone(precision, fixed):
    frames, tri = _batch_systems()
    systems = [dict(L6=f.L6, K=f.K, order=4, beta=f.beta, w=f.w, x=f.x) for f in frames]
    systems.append(dict(tri, order=4))
    _check_batch(systems, precision, fixed, (0, len(systems) - 1))


@pytest.mark.parametrize("precision,fixed", [("fp32", True), ("fp64", True)])
def test_batch_mixing_plain_and_generic_radix_frames(precision, fixed):
    """K0 = 14 = 2 x 7 needs the generic x transform, so the whole batch's x
    launches run it, with its second set of lines in LDS; the frame with the
    plain K0 = 32 (64 lines of 33) needs twice the LDS it uses alone.  The
    generic transform rounds differently from the radix-2..8 one the plain
    frames run alone, so they agree within rounding, not bitwise."""
    plain = dict(make_sys([20., 16., 14., 0., 0., 0.], (32, 24, 20), 4, 1500, 11), order=4)
    gen = dict(make_sys([12., 15., 18., 1., -0.5, 0.5], (14, 18, 21), 4, 900, 12), order=4)
    plain2 = dict(make_sys([18., 18., 18., 0., 0., 0.], (32, 32, 32), 4, 2000, 13), order=4)
    _check_batch([plain, gen, plain2], precision, fixed, (0, 1, 2), exact=False)


@pytest.mark.parametrize("precision,fixed", [("fp32", True)])
def test_batch_of_many_planes_equals_frames_alone(precision, fixed):
    """A batch of 672 x planes over frames of two grid shapes (more
    workgroups per launch than the device holds at once): bitwise the
    frames alone and within the FP32 bars of the oracle."""
    systems = []
    for f in range(8):
        systems.append(dict(make_sys([22., 18., 16., 0.5, -0.5, 0.3], (64, 48, 40), 4, 1200,
                                     20 + f), order=4))
    for f in range(2):
        systems.append(dict(make_sys([26., 21., 19., 0., 1., 0.], (80, 64, 50), 4, 1800,
                                     40 + f), order=4))
    _check_batch(systems, precision, fixed, (0, 8, 9))


def test_batch_rejects_mixed_precision():
    a = ew().Sfac(np.array([10., 10, 10, 0, 0, 0]), np.array([16, 16, 16]), 4, precision="fp32")
    b = ew().Sfac(np.array([10., 10, 10, 0, 0, 0]), np.array([16, 16, 16]), 4)
    a.set_A(ew().ewald_kernel(0.4), 1.0)
    b.set_A(ew().ewald_kernel(0.4), 1.0)
    with pytest.raises(RuntimeError, match="precision"):
        ew().SfacBatch([a, b])


def test_fp32_influence_option():
    """The FP32 engine evaluates /B, the influence and the energy/virial sums
    in FP64 by default; fp32_influence evaluates the closed form and the
    thread sums in FP32 (its own kernel variant).  Both meet the FP32 bars
    against the FP64 oracle, the default is the closer of the two, and a
    batch refuses to mix them."""
    from ewaldcorrel_amd.workloads import water
    wsy = water(2048, 40.0, (48, 48, 48), seed=9)
    o = OracleSfac(wsy.L6, wsy.K, 4)
    o.set_A(ewald_rad(wsy.beta), 1.0)
    E0, v0, dx0, phi0 = o.evaluate(wsy.w, wsy.x)
    errs = {}
    hs = {}
    for opt in (False, True):
        s = ew().Sfac(wsy.L6, np.array(wsy.K), 4, precision="fp32", fp32_influence=opt)
        s.set_A(ew().ewald_kernel(wsy.beta), 1.0)
        assert s.info()["table_mode"] == 1
        assert bool(s.info()["flags"] & ew().SFAC_FP32_INFLUENCE) == opt
        E, vir, dx, phi = s.evaluate(wsy.w, wsy.x)
        errs[opt] = (abs(E - E0) / abs(E0), rel(vir, v0), rel_rms(dx, dx0))
        say("fp32_influence", opt, errs[opt])
        assert errs[opt][0] <= 1e-6 and errs[opt][1] <= 1e-5 and errs[opt][2] <= 1e-5
        hs[opt] = s
    with pytest.raises(RuntimeError):
        ew().SfacBatch([hs[False], hs[True]])
    with pytest.raises(ValueError):
        ew().Sfac(wsy.L6, np.array(wsy.K), 4, precision="fp64", fp32_influence=True)
