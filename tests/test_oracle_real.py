"""Pins of the oracle's real-space Ewald terms (notes.tex:66-79,
test.py:83,126-132) by what the mathematics fixes: Madelung constants of
rock salt and caesium chloride from reciprocal + real + self energies, the
independence of the total energy from the splitting parameter eta
(including the neutralising-background term of a charged cell), the
gradient and the virial by finite differences, and phi = dE/dw."""
import numpy as np
import pytest

from oracle.sfac_oracle import LofS, OracleSfac, ewald_rad, real_space

NACL = 1.747564594633182   # rock salt, per nearest-neighbour distance
CSCL = 1.762674773070987   # caesium chloride, per nearest-neighbour distance


def total(L6, w, x, eta, rc, K, order=12, max_m=3.5):
    r = real_space(L6, w, x, eta, rc)
    o = OracleSfac(L6, K, order)
    o.set_A(ewald_rad(eta), max_m)
    o.sfac(w, x)
    return o.en() + r["E"] + r["E_self"] + r["E_bg"]


def rock_salt(m):
    idx = np.array([(i, j, k) for i in range(2 * m) for j in range(2 * m)
                    for k in range(2 * m)], float)
    q = (-1.0) ** idx.sum(axis=1)
    return idx + 0.25, q, 2.0 * m


@pytest.mark.parametrize("eta,rc", [(1.5, 4.5), (1.2, 5.5)])
def test_madelung_rock_salt(eta, rc):
    x, q, L = rock_salt(2)
    L6 = np.array([L, L, L, 0, 0, 0])
    E = total(L6, q, x, eta, rc, (48, 48, 48))
    M = -E / (len(q) / 2)
    assert abs(M - NACL) < 1e-11


def test_madelung_caesium_chloride():
    m = 2
    a = 2.0 / np.sqrt(3.0)  # nearest-neighbour distance 1
    cells = np.array([(i, j, k) for i in range(m) for j in range(m) for k in range(m)], float)
    x = np.concatenate([cells * a, (cells + 0.5) * a]) + 0.1
    q = np.concatenate([np.ones(len(cells)), -np.ones(len(cells))])
    L = m * a
    L6 = np.array([L, L, L, 0, 0, 0])
    E = total(L6, q, x, 1.5, 4.5, (32, 32, 32))
    M = -E / (len(q) / 2)
    assert abs(M - CSCL) < 1e-10


def random_system(n, L6, seed, neutral):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (n, 3)) @ LofS(L6)
    w = rng.normal(size=n)
    if neutral:
        w -= w.mean()
    return w, x


@pytest.mark.parametrize("neutral", [True, False], ids=["neutral", "charged"])
def test_total_energy_independent_of_eta(neutral):
    L6 = np.array([6., 6.5, 7., 0.5, -0.4, 0.3])
    w, x = random_system(20, L6, 3, neutral)
    E1 = total(L6, w, x, 1.2, 6.0, (48, 48, 48), max_m=3.0)
    E2 = total(L6, w, x, 1.5, 5.0, (48, 48, 48), max_m=3.5)
    assert abs(E1 - E2) <= 1e-9 * abs(E1)


def test_real_space_gradient_phi_virial_by_fd():
    L6 = np.array([6., 6.5, 7., 0.5, -0.4, 0.3])
    w, x = random_system(8, L6, 4, False)
    eta, rc = 0.9, 5.0
    r = real_space(L6, w, x, eta, rc)
    h = 1e-6
    # gradient
    for a in range(3):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, d] += h
            xm[a, d] -= h
            fd = (real_space(L6, w, xp, eta, rc)["E"] - real_space(L6, w, xm, eta, rc)["E"]) / (2 * h)
            assert abs(fd - r["dx"][a, d]) < 1e-6 * max(1.0, abs(fd))
    # phi = dE/dw of all three terms (exact: E is quadratic in w)
    def Etot(ww):
        s = real_space(L6, ww, x, eta, rc)
        return s["E"] + s["E_self"] + s["E_bg"]
    for a in range(3):
        wp, wm = w.copy(), w.copy()
        wp[a] += 0.5
        wm[a] -= 0.5
        assert abs((Etot(wp) - Etot(wm)) - r["phi"][a]) < 1e-9 * max(1.0, abs(r["phi"][a]))
    # virial of the pairs: -(1/V) (dE/dL)^T L at fixed fractional coordinates
    # (the convention test.py:107-116 checks for de1)
    s = x @ np.linalg.inv(LofS(L6))

    def EL(L6b):
        return real_space(L6b, w, s @ LofS(L6b), eta, rc)["E"]
    g = np.zeros(6)
    for c in range(6):
        Lp, Lm = L6.copy(), L6.copy()
        Lp[c] += h
        Lm[c] -= h
        g[c] = (EL(Lp) - EL(Lm)) / (2 * h)
    Pi = LofS(g).T
    V = L6[0] * L6[1] * L6[2]
    M = (Pi @ LofS(L6)).T / (-V)
    vfd = np.array([M[0, 0], M[1, 1], M[2, 2], M[1, 0], M[2, 0], M[2, 1]])
    assert np.abs(vfd - r["vir"]).max() < 1e-6 * np.abs(r["vir"]).max()
