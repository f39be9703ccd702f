"""Real-space Ewald part and corrections on the device, completing the
reciprocal engine to the total Coulomb energy (notes.tex:66-79): parity
with the oracle's brute-force sum, the rock-salt Madelung constant from
engine reciprocal + device real space, and independence of the total from
the splitting parameter."""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from conftest import rel, rel_rms  # noqa: E402
from oracle.sfac_oracle import LofS, real_space  # noqa: E402

pytestmark = pytest.mark.gpu

NACL = 1.747564594633182


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ew():
    from ewaldcorrel_amd import ewald
    return ewald


def system(n, L6, seed, neutral):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.3, 1.3, (n, 3)) @ LofS(L6)
    w = rng.normal(size=n)
    if neutral:
        w -= w.mean()
    return w, x


def test_real_space_matches_oracle():
    L6 = np.array([6., 6.5, 7., 0.5, -0.4, 0.3])
    w, x = system(150, L6, 1, neutral=False)
    eta, rc = 0.8, 8.0  # several images per axis
    s = ew().Sfac(L6, np.array([8, 8, 8]), 4)
    g = s.real_space(w, x, eta, rc)
    r = real_space(L6, w, x, eta, rc)
    say("E", g["E"], r["E"], "vir", rel(g["vir"], r["vir"]), "dx", rel_rms(g["dx"], r["dx"]),
        "phi", rel_rms(g["phi"], r["phi"]))
    assert abs(g["E"] - r["E"]) <= 1e-12 * abs(r["E"])
    assert rel(g["vir"], r["vir"]) <= 1e-12
    assert rel_rms(g["dx"], r["dx"]) <= 1e-12
    assert rel_rms(g["phi"], r["phi"]) <= 1e-12
    for k in ("E_self", "E_bg", "vir_bg"):
        assert abs(g[k] - r[k]) <= 1e-13 * abs(r[k])


def test_device_pointer_form_equals_host_form():
    L6 = np.array([9., 9., 9., 0., 0., 0.])
    w, x = system(300, L6, 2, neutral=True)
    s = ew().Sfac(L6, np.array([8, 8, 8]), 4)
    g = s.real_space(w, x, 1.0, 4.5)
    wd = torch.tensor(w, device="cuda")
    xd = torch.tensor(x, device="cuda")
    out = torch.zeros(10, dtype=torch.float64, device="cuda")
    dx = torch.zeros((300, 3), dtype=torch.float64, device="cuda")
    phi = torch.zeros(300, dtype=torch.float64, device="cuda")
    s.real_space_dev(300, wd.data_ptr(), xd.data_ptr(), 1.0, 4.5, out.data_ptr(), dx.data_ptr(),
                     phi.data_ptr())
    s.sync()
    o = out.cpu().numpy()
    assert o[0] == g["E"] and np.array_equal(o[1:7], g["vir"])
    assert np.array_equal(dx.cpu().numpy(), g["dx"])
    assert np.array_equal(phi.cpu().numpy(), g["phi"])


def rock_salt(m):
    idx = np.array([(i, j, k) for i in range(2 * m) for j in range(2 * m)
                    for k in range(2 * m)], float)
    return idx + 0.25, (-1.0) ** idx.sum(axis=1), 2.0 * m


@pytest.mark.parametrize("eta,rc", [(1.5, 4.5), (1.2, 5.5)])
def test_madelung_rock_salt_on_device(eta, rc):
    x, q, L = rock_salt(2)
    L6 = np.array([L, L, L, 0, 0, 0])
    s = ew().Sfac(L6, np.array([48, 48, 48]), 12)
    s.set_A(ew().ewald_kernel(eta), 3.5)
    s.sfac(q, x)
    E = s.en()
    r = s.real_space(q, x, eta, rc)
    M = -(E + r["E"] + r["E_self"] + r["E_bg"]) / (len(q) / 2)
    say("eta", eta, "Madelung", M, M - NACL)
    assert abs(M - NACL) < 1e-10


def test_total_energy_force_independent_of_eta():
    L6 = np.array([10., 11., 12., 1., -0.5, 0.7])
    w, x = system(200, L6, 3, neutral=False)
    res = []
    for eta, rc, K, mm in [(0.9, 6.0, 40, 2.2), (1.1, 5.0, 48, 2.6)]:
        s = ew().Sfac(L6, np.array([K, K, K]), 10)
        s.set_A(ew().ewald_kernel(eta), mm)
        E, vir, dx, phi = s.evaluate(w, x)
        r = s.real_space(w, x, eta, rc)
        res.append((E + r["E"] + r["E_self"] + r["E_bg"], dx + r["dx"], phi + r["phi"]))
    (E1, F1, P1), (E2, F2, P2) = res
    say("dE", abs(E1 - E2) / abs(E1), "dF", rel_rms(F1, F2), "dphi", rel_rms(P1, P2))
    assert abs(E1 - E2) <= 1e-8 * abs(E1)
    assert rel_rms(F1, F2) <= 1e-6
    assert rel_rms(P1, P2) <= 1e-7
