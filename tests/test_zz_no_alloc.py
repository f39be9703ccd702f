"""The allocation-free evaluation (sfac_eval_dev_noalloc) against
sfac_eval_dev on the same handle configuration and inputs: two calls in a
row, the second with fewer atoms (the static scratch keeps the first call's
contents), bitwise equal in the deterministic modes (FP32, FP64 fixed point)
and within 1e-12 for FP64 LDS-atomic spreading; and a call whose atoms need
more than the scratch is refused before any device work."""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle.sfac_oracle import LofS  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ew():
    from ewaldcorrel_amd import ewald
    return ewald


def inputs(seed, n, L6, dev):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (n, 3)) @ LofS(L6)
    w = rng.normal(size=n)
    w -= w.mean()
    return dict(n=n, w=torch.tensor(w, device=dev), x=torch.tensor(x, device=dev))


def run(s, inp, noalloc, dev):
    n = inp["n"]
    out = torch.zeros(8, dtype=torch.float64, device=dev)
    dx = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    phi = torch.zeros(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    f = s.eval_dev_noalloc if noalloc else s.eval_dev
    f(n, inp["w"].data_ptr(), inp["x"].data_ptr(), out.data_ptr(), dx.data_ptr(),
      phi.data_ptr())
    s.sync()
    return out.cpu().numpy(), dx.cpu().numpy(), phi.cpu().numpy()


@pytest.mark.parametrize("precision,fixed,order,K", [
    ("fp32", True, 4, (32, 30, 36)),
    ("fp64", True, 6, (30, 32, 27)),
    ("fp64", False, 4, (32, 30, 36)),
])
def test_noalloc_equals_eval_dev(precision, fixed, order, K):
    dev = torch.device("cuda:0")
    L6 = np.array([25., 24., 27., 2., -1.5, 1.])
    say("noalloc", precision, fixed, order, K)
    s_ref = ew().Sfac(L6, np.array(K), order, precision=precision, fixed=fixed)
    s_na = ew().Sfac(L6, np.array(K), order, precision=precision, fixed=fixed)
    for s in (s_ref, s_na):
        s.set_A(ew().ewald_kernel(0.35), 1.0)
    tol = 0.0 if (precision == "fp32" or fixed) else 1e-12
    for call, n in enumerate((9000, 2500)):  # the second call smaller
        inp = inputs(40 + call, n, L6, dev)
        a = run(s_ref, inp, False, dev)
        b = run(s_na, inp, True, dev)
        say("call", call, "n", n, "E", a[0][0], b[0][0])
        for u, v in zip(a, b):
            if tol == 0.0:
                assert np.array_equal(u, v)
            else:
                assert np.max(np.abs(u - v)) <= tol * np.max(np.abs(u))
    # the same handle through both entry points, alternately
    inp = inputs(50, 4000, L6, dev)
    c = run(s_na, inp, False, dev)
    d = run(s_na, inp, True, dev)
    for u, v in zip(c, d):
        if tol == 0.0:
            assert np.array_equal(u, v)
        else:
            assert np.max(np.abs(u - v)) <= tol * np.max(np.abs(u))


def test_noalloc_refuses_more_atoms_than_its_scratch():
    """The size check runs on the host before any launch, so the device
    buffers given here (a few atoms) are never read."""
    dev = torch.device("cuda:0")
    L6 = np.array([25., 24., 27., 0., 0., 0.])
    s = ew().Sfac(L6, np.array((32, 32, 32)), 4, precision="fp32")
    s.set_A(ew().ewald_kernel(0.35), 1.0)
    inp = inputs(3, 16, L6, dev)
    out = torch.zeros(8, dtype=torch.float64, device=dev)
    with pytest.raises(RuntimeError, match="scratch"):
        s.eval_dev_noalloc(50_000_000, inp["w"].data_ptr(), inp["x"].data_ptr(),
                           out.data_ptr(), 0, 0)
    # the handle still works afterwards
    a = run(s, inp, True, dev)
    b = run(s, inp, False, dev)
    assert np.array_equal(a[0], b[0])
