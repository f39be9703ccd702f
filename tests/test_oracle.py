"""Pins the FP64 oracle against the reference's golden log (test.log.ref),
closed forms, independent brute-force restatements and exact invariants.
CPU only."""
import itertools
import os
import sys
from math import erfc, pi

import numpy as np
import pytest

from conftest import ROOT, case_T, rel
from oracle.sfac_oracle import (Cell, LofS, OracleSfac, SofL, bspl_deval, bspl_dft,
                                bspl_eval, bspline_coeffs, ewald_rad)


# ---- golden log ------------------------------------------------------------------
@pytest.fixture(scope="module")
def oracle_T(golden):
    L, K, order, crds, atoms, q = case_T(golden)
    s = OracleSfac(L, K, order)
    s.set_A(ewald_rad(golden["case"]["eta"]), golden["case"]["max_m"])
    return s, L, crds, atoms, q


def test_golden_S_lines(golden, oracle_T):
    """test.log.ref:2-3 (Re S-hat(0,0,kz), 8 s.f.) and :6 (line errors)."""
    s, L, crds, atoms, q = oracle_T
    s.sfac(q, atoms)
    M = s.get_S()
    assert np.allclose(M[0, 0].real, golden["spl_S00"], atol=6e-9, rtol=0)
    mx = np.arange(M.shape[0]); mx[M.shape[0] // 2 + 1:] -= M.shape[0]
    my = np.arange(M.shape[1]); my[M.shape[1] // 2 + 1:] -= M.shape[1]
    mz = np.arange(M.shape[2])
    sx = sum(qq * np.exp(-2j * pi * x[0] * mx) for qq, x in zip(q, crds))
    sy = sum(qq * np.exp(-2j * pi * x[1] * my) for qq, x in zip(q, crds))
    sz = sum(qq * np.exp(-2j * pi * x[2] * mz) for qq, x in zip(q, crds))
    assert np.allclose(sz.real, golden["ref_S00"], atol=6e-9)
    errs = [np.abs(M[:, 0, 0] - sx)[:M.shape[0] // 4].max(),
            np.abs(M[0, :, 0] - sy)[:M.shape[1] // 4].max(),
            np.abs(M[0, 0, :] - sz)[:M.shape[2] // 2].max()]
    # 16-17 significant digits in the log; a few ulp of |S| ~ 4 apart
    assert np.allclose(errs, golden["line_errors"], rtol=0, atol=4e-15)


def test_golden_energy_forces_virial(golden, oracle_T):
    """test.log.ref:7-12,17-18: E (15 s.f.), dx (8 dp), virial (9 s.f.)."""
    s, L, crds, atoms, q = oracle_T
    s.sfac(q, atoms)
    e0 = s.en()
    e1, vir = s.de1()
    assert abs(e0 - golden["E"]) <= 1e-14 * golden["E"]
    assert abs(e1 - golden["E1"]) <= 1e-14 * golden["E1"]
    assert abs(e1 - e0) < 1e-14
    dx = s.de2(q, atoms)
    assert np.allclose(dx.ravel(), golden["spl_dx"], atol=6e-9, rtol=0)
    assert np.allclose(vir, golden["spl_vir"], rtol=6e-9, atol=0)
    # the FD columns of the log agree with the analytic ones to the stated bounds
    assert np.abs(dx.ravel() - np.array(golden["ref_dx"])).max() < 1e-7
    assert np.abs(vir - np.array(golden["ref_vir"])).max() < 1e-8


# ---- B-splines --------------------------------------------------------------------
def _M_recursive(n, x):
    """Cox-de Boor recursion M_n(x) = x/(n-1) M_{n-1}(x) + (n-x)/(n-1) M_{n-1}(x-1),
    M_1 = indicator of [0,1) -- independent of the coefficient derivation."""
    x = np.asarray(x, dtype=float)
    if n == 1:
        return ((x >= 0) & (x < 1)).astype(float)
    return (x * _M_recursive(n - 1, x) + (n - x) * _M_recursive(n - 1, x - 1)) / (n - 1)


@pytest.mark.parametrize("n", [2, 4, 6, 8, 10, 12])
def test_bspline_closed_form(n):
    M = bspline_coeffs(n)
    x = np.linspace(0, 1, 41, endpoint=False)
    for j in range(n):
        assert np.allclose(bspl_eval(M, x, j), _M_recursive(n, x + j), atol=1e-14)
        b, db = bspl_deval(M, x, j)
        # M_n' (y) = M_{n-1}(y) - M_{n-1}(y-1)
        ref = _M_recursive(n - 1, x + j) - _M_recursive(n - 1, x + j - 1)
        assert np.allclose(db, ref, atol=1e-13)
        assert np.allclose(b, bspl_eval(M, x, j), atol=0)
    # partition of unity
    tot = sum(bspl_eval(M, x, j) for j in range(n))
    assert np.allclose(tot, 1.0, atol=1e-14)


def test_bspline_nodes():
    """Node values M_4(j) = 0,1,4,1 /6 and M_6(j) = 0,1,26,66,26,1 /120 (the
    reference's test.cpp tabulates the same function)."""
    assert np.allclose(bspline_coeffs(4)[:, -1], np.array([0, 1, 4, 1]) / 6, atol=1e-16)
    assert np.allclose(bspline_coeffs(6)[:, -1], np.array([0, 1, 26, 66, 26, 1]) / 120,
                       atol=1e-16)


@pytest.mark.parametrize("K,n", [(12, 4), (13, 4), (14, 6), (32, 4), (7, 8)])
def test_bspl_dft_direct(K, n):
    """1/b(m) with b(m) = sum_j M_n(j) exp(-2 pi i (j - n/2) m/K), a complex
    DFT whose imaginary part vanishes by symmetry."""
    M = bspline_coeffs(n)
    m = np.arange(K)
    b = sum(_M_recursive(n, j) * np.exp(-2j * pi * (j - n // 2) * m / K) for j in range(n + 1))
    assert np.abs(b.imag).max() < 1e-14
    assert np.allclose(bspl_dft(K, n, M), 1.0 / b.real, rtol=1e-13)


# ---- geometry -----------------------------------------------------------------------
def test_cell_inverse_and_wrap():
    L6 = np.array([9., 10., 11., 2., -0.1, 1.0])
    c = Cell(L6)
    assert np.allclose(c.L @ c.iL, np.eye(3), atol=1e-15)
    assert np.isclose(c.iV, 1 / np.linalg.det(c.L))
    rng = np.random.default_rng(0)
    x = rng.uniform(-50, 50, (200, 3))
    s = c.scale(x)
    assert (s >= 0).all() and (s <= 1).all()
    # r = L^T s modulo lattice vectors
    d = x - s @ c.L
    n = d @ np.linalg.inv(c.L)
    assert np.allclose(n, np.round(n), atol=1e-9)
    assert np.allclose(SofL(LofS(L6)), L6)


# ---- spreading: scatter form vs the reference's gather form ---------------------------------
def _spread_gather_form(s, w, x):
    """Per-grid-point gather form of sfac.cpp:113-124,141-168 (pure loops)."""
    K0, K1, K2 = s.K
    n = s.order
    c, frac = s.cells(x)
    cells = {}
    for a in range(len(w)):
        cells.setdefault((c[a, 0], c[a, 1], c[a, 2]), []).append((frac[a], w[a]))
    Q = np.zeros((K0, K1, K2))
    for i1, j1, k1 in itertools.product(range(K0), range(K1), range(K2)):
        acc = 0.0
        for i, j, k in itertools.product(range(n), repeat=3):
            key = ((i1 - n // 2 + i) % K0, (j1 - n // 2 + j) % K1, (k1 - n // 2 + k) % K2)
            for f, ww in cells.get(key, ()):
                acc += (bspl_eval(s.M, f[0], i) * bspl_eval(s.M, f[1], j)
                        * bspl_eval(s.M, f[2], k) * ww)
        Q[i1, j1, k1] = acc
    return Q


@pytest.mark.parametrize("K,order", [((5, 7, 6), 4), ((6, 5, 9), 6), ((4, 4, 4), 2)])
def test_spread_scatter_equals_gather(K, order):
    rng = np.random.default_rng(1)
    L6 = np.array([7., 8., 9., 1., -0.5, 0.7])
    s = OracleSfac(L6, K, order)
    x = rng.uniform(-10, 20, (9, 3))
    w = rng.normal(size=9)
    assert np.allclose(s.spread(w, x), _spread_gather_form(s, w, x), atol=1e-14)


def test_structure_factor_origin_and_dft():
    """S-hat(0) = sum w exactly; r2c equals the brute-force DFT of the grid on a
    prime-sized grid."""
    rng = np.random.default_rng(2)
    s = OracleSfac([7., 8., 9., 1., -0.5, 0.7], (5, 7, 13), 4)
    x = rng.uniform(0, 9, (11, 3))
    w = rng.normal(size=11)
    s.sfac(w, x)
    assert abs(s.get_S()[0, 0, 0] - w.sum()) < 1e-13
    Q = s.spread(w, x)
    K0, K1, K2 = s.K
    g = np.stack(np.meshgrid(range(K0), range(K1), range(K2), indexing="ij"), -1).reshape(-1, 3)
    for kx, ky, kz in [(1, 2, 3), (4, 6, 6), (0, 3, 1)]:
        ph = np.exp(-2j * pi * (g[:, 0] * kx / K0 + g[:, 1] * ky / K1 + g[:, 2] * kz / K2))
        assert abs(np.fft.rfftn(Q)[kx, ky, kz] - (Q.ravel() * ph).sum()) < 1e-12


# ---- influence table ---------------------------------------------------------------------
def test_set_A_brute_force():
    """Box scan of lattice points with a direct |m|^2 <= R^2 test and the
    half-complex fold rule, vs the reference-ordered ball enumeration."""
    L6 = np.array([9., 10., 11., 2., -0.1, 1.0])
    K = (6, 5, 7)
    max_m = 0.93  # not on a lattice sphere
    s = OracleSfac(L6, K, 4)
    rad = ewald_rad(0.6)
    s.set_A(rad, max_m)
    iL = s.cell.iL
    ldim = s.ldim
    A = np.zeros((K[0], K[1], ldim))
    dA = np.zeros((K[0], K[1], ldim, 6))
    I = 12
    for i, j, k in itertools.product(range(-I, I + 1), repeat=3):
        m = np.array([i * iL[0, 0], i * iL[1, 0] + j * iL[1, 1],
                      i * iL[2, 0] + j * iL[2, 1] + k * iL[2, 2]])
        m2 = m @ m
        if m2 > max_m ** 2:
            continue
        f, fp = rad(np.array([m2]))
        a, b, c = i % K[0], j % K[1], k % K[2]
        if c >= ldim:
            a, b, c = (-i) % K[0], (-j) % K[1], K[2] - c
        A[a, b, c] += 0.5 * f[0]
        mm = [m[0] * m[0], m[1] * m[1], m[2] * m[2], m[1] * m[0], m[2] * m[0], m[2] * m[1]]
        dA[a, b, c] += fp[0] * np.array(mm)
    assert np.allclose(s.A, A, rtol=1e-13, atol=0)
    assert np.allclose(s.dA, dA, rtol=1e-12, atol=1e-300)
    assert abs(s.A).max() > 0


# ---- derivatives ---------------------------------------------------------------------------
def _rand_system(seed, N=12, L6=(9., 10., 11., 2., -0.1, 1.0), K=(12, 13, 14), order=4):
    rng = np.random.default_rng(seed)
    L6 = np.array(L6)
    s = OracleSfac(L6, K, order)
    s.set_A(ewald_rad(0.5), 1.0)
    crds = rng.uniform(0, 1, (N, 3))
    x = crds @ LofS(L6)
    w = rng.normal(size=N)
    return s, L6, crds, x, w


@pytest.mark.parametrize("order", [4, 6])
def test_forces_by_central_differences(order):
    s, L6, crds, x, w = _rand_system(3, order=order)
    s.sfac(w, x)
    e0 = s.en()
    e1, vir = s.de1()
    assert abs(e0 - e1) < 1e-14 * max(1, abs(e0))
    dx = s.de2(w, x)
    h = 1e-6
    fd = np.zeros_like(x)
    for a in range(len(w)):
        for d in range(3):
            xp = x.copy(); xp[a, d] += h
            xm = x.copy(); xm[a, d] -= h
            s.sfac(w, xp); ep = s.en()
            s.sfac(w, xm); em = s.en()
            fd[a, d] = (ep - em) / (2 * h)
    assert np.abs(dx - fd).max() < 1e-7 * max(1.0, np.abs(dx).max())


def test_virial_by_central_differences():
    """FD of E(L) at fixed fractional coordinates (test.py:107-116)."""
    s, L6, crds, x, w = _rand_system(4)
    s.sfac(w, x)
    _, vir = s.de1()
    V = L6[0] * L6[1] * L6[2]
    rad = ewald_rad(0.5)
    h = 1e-7
    g = np.zeros(6)
    for c in range(6):
        for sgn in (1, -1):
            Lp = L6.copy(); Lp[c] += sgn * h
            s.set_L(Lp); s.set_A(rad, 1.0)
            s.sfac(w, crds @ LofS(Lp))
            g[c] += sgn * s.en()
        g[c] /= 2 * h
    s.set_L(L6)
    Pi = LofS(g).T
    Pi = SofL((Pi @ LofS(L6)).T) / (-V)
    assert np.abs(vir - Pi).max() < 1e-8 * max(1.0, np.abs(vir).max() * 1e4)


def test_potential_is_exact_charge_derivative():
    """E is quadratic in w, so phi = dE/dw exactly and E = 1/2 w.phi."""
    s, L6, crds, x, w = _rand_system(5)
    s.sfac(w, x)
    E, _ = s.de1()
    phi = s.potl(x)
    assert abs(E - 0.5 * w @ phi) < 1e-13 * abs(E)
    for a in (0, 5):
        h = 0.37
        wp = w.copy(); wp[a] += h
        wm = w.copy(); wm[a] -= h
        s.sfac(wp, x); ep = s.en()
        s.sfac(wm, x); em = s.en()
        assert abs((ep - em) / (2 * h) - phi[a]) < 1e-12 * max(1, abs(phi[a]))


def test_state_machine_errors():
    s, L6, crds, x, w = _rand_system(6)
    with pytest.raises(RuntimeError):
        s.en()
    with pytest.raises(RuntimeError):
        s.de2(w, x)
    s.sfac(w, x)
    with pytest.raises(RuntimeError):
        s.potl(x)
    s.de1()
    with pytest.raises(RuntimeError):
        s.de1()
    with pytest.raises(ValueError):
        OracleSfac(L6, (8, 8, 8), 3)
    with pytest.raises(ValueError):
        OracleSfac(L6, (8, 0, 8), 4)


def test_ewald_beta_independence():
    """Physics sanity: recip (oracle) + real-space + self energy of a neutral
    system does not depend on the splitting parameter (notes.tex:73-79)."""
    rng = np.random.default_rng(7)
    Lb = 10.0
    L6 = np.array([Lb, Lb, Lb, 0, 0, 0])
    N = 6
    x = rng.uniform(0, Lb, (N, 3))
    q = np.array([1., -1., 1., -1., 0.5, -0.5])
    tot = []
    for beta in (0.30, 0.36):
        s = OracleSfac(L6, (40, 40, 40), 8)
        s.set_A(ewald_rad(beta), 2.4)
        s.sfac(q, x)
        er = s.en()
        rc = 14.0
        nimg = 2
        ereal = 0.0
        for i in range(N):
            for j in range(N):
                for t in itertools.product(range(-nimg, nimg + 1), repeat=3):
                    if i == j and t == (0, 0, 0):
                        continue
                    r = np.linalg.norm(x[j] - x[i] + Lb * np.array(t))
                    if r < rc:
                        ereal += 0.5 * q[i] * q[j] * erfc(beta * r) / r
        eself = -beta / np.sqrt(pi) * (q * q).sum()
        tot.append(er + ereal + eself)
    assert abs(tot[0] - tot[1]) < 2e-6 * abs(tot[0])


def test_plot_radial_against_reference_fixture():
    """oracle.plot_radial reproduces the reference's plot.py output stored in
    tests/golden/radial_plot.json (written by make_radial_golden.py)."""
    import json
    from oracle.sfac_oracle import plot_radial
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_radial_golden import make_input
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "radial_plot.json")))
    f = make_input(tuple(g["shape"]), g["seed"]).real
    assert f.sum() == g["f_real_sum"] and f.ravel()[:8].tolist() == g["f_real_first"]
    rows = plot_radial(f, g["L"])
    assert len(rows) == len(g["rows"])
    for (k, m, c), (k0, m0, c0) in zip(rows, g["rows"]):
        assert k == k0 and c == c0 and abs(m - m0) <= 1e-15 * abs(m0)
