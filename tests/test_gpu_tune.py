"""Autotuner of launch shapes (transform path, FFT radix order per axis,
spreading and gather tiles).  Tuning must not change results beyond the
FFT's rounding: FP64 within 1e-12 of the untuned evaluation (and the oracle's
1e-10 / 1e-9 bars), FP32 within 1e-6 in energy and 1e-5 in forces.  Tuned
shapes are remembered for later handles of the same grid, order and flags."""
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


def system(L6, K, order, n, seed, beta=0.35):
    rng = np.random.default_rng(seed)
    L6 = np.asarray(L6, dtype=float)
    x = rng.uniform(0, 1, (n, 3)) @ LofS(L6)
    w = rng.normal(size=n)
    w -= w.mean()
    return L6, tuple(K), order, beta, w, x


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotune_keeps_results_and_is_remembered(precision):
    ew().tune_clear()
    L6, K, order, beta, w, x = system([30., 28., 32., 2., -1., 1.5], (32, 30, 36), 4, 3000, 7)
    s = ew().Sfac(L6, np.array(K), order, precision=precision)
    s.set_A(ew().ewald_kernel(beta), 1.0)
    a = s.evaluate(w, x)
    rep = s.autotune(n=3000, reps=5)
    say(rep)
    assert rep.startswith("start") and "end" in rep
    b = s.evaluate(w, x)
    tol_e, tol_f = (1e-12, 1e-11) if precision == "fp64" else (1e-6, 1e-5)
    assert abs(a[0] - b[0]) <= tol_e * abs(a[0])
    assert rel_rms(b[2], a[2]) <= tol_f and rel_rms(b[3], a[3]) <= tol_f
    o = OracleSfac(L6, K, order)
    o.set_A(ewald_rad(beta), 1.0)
    E0, v0, dx0, phi0 = o.evaluate(w, x)
    if precision == "fp64":
        assert abs(b[0] - E0) <= 1e-10 * abs(E0) and rel(b[1], v0) <= 1e-10
        assert rel_rms(b[2], dx0) <= 1e-9 and rel_rms(b[3], phi0) <= 1e-9
    else:
        assert abs(b[0] - E0) <= 1e-6 * abs(E0) and rel_rms(b[2], dx0) <= 1e-5
    text = ew().tune_export()
    say(text)
    flag = 1 if precision == "fp32" else 0
    assert f"32x30x36/4/{flag | (2 if precision == 'fp32' else 0)}" in text
    # a new handle of the same key starts from the tuned shapes: same results
    s2 = ew().Sfac(L6, np.array(K), order, precision=precision)
    s2.set_A(ew().ewald_kernel(beta), 1.0)
    c = s2.evaluate(w, x)
    if precision == "fp32":
        assert c[0] == b[0] and np.array_equal(c[2], b[2])
    else:
        assert abs(c[0] - b[0]) <= 1e-13 * abs(b[0])
    ew().tune_clear()


def test_batch_autotune_on_uneven_frames():
    ew().tune_clear()
    dev = torch.device("cuda:0")
    frames = [system([20. + 4 * i, 22. + 3 * i, 21. + 5 * i, 0, 0, 0],
                     (20 + 4 * i, 24 + 2 * i, 30 - 2 * i), 4, 500 + 300 * i, 20 + i)
              for i in range(4)]
    frames.append(frames[1])  # two frames of one class share a shape
    hs, bufs = [], []
    for L6, K, order, beta, w, x in frames:
        s = ew().Sfac(L6, np.array(K), order, precision="fp32")
        s.set_A(ew().ewald_kernel(beta), 1.0)
        hs.append(s)
        n = len(w)
        bufs.append(dict(w=torch.tensor(w, device=dev), x=torch.tensor(x, device=dev),
                         out=torch.zeros(8, dtype=torch.float64, device=dev),
                         dx=torch.zeros((n, 3), dtype=torch.float64, device=dev),
                         phi=torch.zeros(n, dtype=torch.float64, device=dev), n=n))
    b = ew().SfacBatch(hs)
    for f, d in enumerate(bufs):
        b.set(f, d["n"], d["w"].data_ptr(), d["x"].data_ptr(), d["out"].data_ptr(),
              d["dx"].data_ptr(), d["phi"].data_ptr())
    b.eval()
    b.sync()
    before = [(d["out"][0].item(), d["dx"].cpu().numpy().copy()) for d in bufs]
    rep = b.autotune(reps=5)
    say(rep)
    for _ in range(3):  # graph replays after tuning
        b.eval()
    b.sync()
    for f, d in enumerate(bufs):
        E, dx = d["out"][0].item(), d["dx"].cpu().numpy()
        assert abs(E - before[f][0]) <= 1e-6 * abs(before[f][0]), f
        assert rel_rms(dx, before[f][1]) <= 1e-5, f
    b.close()
    ew().tune_clear()


def test_autotune_on_caller_device_atoms_larger_than_the_initial_buffers():
    """Tuning on the caller's device arrays with more atoms than a new
    handle's staging buffers hold: the buffers are sized for n before the
    tuner's evaluations write forces and potentials into them."""
    ew().tune_clear()
    L6, K, order, beta, w, x = system([40., 40., 40., 0., 0., 0.], (40, 40, 40), 6, 20000, 8)
    s = ew().Sfac(L6, np.array(K), order, precision="fp32")
    s.set_A(ew().ewald_kernel(beta), 1.0)
    a = s.evaluate(w, x)
    dev = torch.device("cuda:0")
    wd = torch.tensor(w, device=dev)
    xd = torch.tensor(x, device=dev)
    rep = s.autotune(len(w), wd.data_ptr(), xd.data_ptr(), reps=3)
    assert rep.startswith("start")
    b = s.evaluate(w, x)
    assert abs(a[0] - b[0]) <= 1e-6 * abs(a[0])
    assert rel_rms(b[2], a[2]) <= 1e-5 and rel_rms(b[3], a[3]) <= 1e-5
    ew().tune_clear()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("slab", [1, 0])
def test_radix16_plans_against_the_oracle(precision, slab):
    """Radix-16 stages (with 2, 3 and 4 after them) on every axis, through
    the y-z slab or the line passes, imported as remembered shapes."""
    ew().tune_clear()
    L6, K, order, beta, w, x = system([31., 29., 33., 1.5, -2., 1.], (32, 48, 64), 4, 2000, 11)
    s = ew().Sfac(L6, np.array(K), order, precision=precision)
    s.set_A(ew().ewald_kernel(beta), 1.0)
    s.autotune(n=2000, reps=1)
    line = [ln for ln in ew().tune_export().splitlines() if ln.startswith("32x48x64/4/")][0]
    f = line.split()
    f[1], f[2], f[3] = "16,2", "16,3", "4,16"
    f[-1] = str(slab)
    ew().tune_clear()
    assert ew().tune_import(" ".join(f)) == 1
    s2 = ew().Sfac(L6, np.array(K), order, precision=precision)
    s2.set_A(ew().ewald_kernel(beta), 1.0)
    E, vir, dx, phi = s2.evaluate(w, x)
    say("radix-16 shapes", " ".join(f), "E", E)
    o = OracleSfac(L6, K, order)
    o.set_A(ewald_rad(beta), 1.0)
    E0, v0, dx0, phi0 = o.evaluate(w, x)
    if precision == "fp64":
        assert abs(E - E0) <= 1e-10 * abs(E0) and rel(vir, v0) <= 1e-10
        assert rel_rms(dx, dx0) <= 1e-9 and rel_rms(phi, phi0) <= 1e-9
    else:
        assert abs(E - E0) <= 1e-6 * abs(E0) and rel(vir, v0) <= 1e-5
        assert rel_rms(dx, dx0) <= 1e-5 and rel_rms(phi, phi0) <= 1e-5
    ew().tune_clear()
