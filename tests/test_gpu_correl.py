"""Structure-factor accumulation on the device (the reference's callers of
get_S: run.py sums |S|^2 over frames, gmx_ewald.py sums species cross terms
S_i conj(S_j), plot.py averages over |k| shells).  Checked against the same
sums formed on the host from the oracle's S-hat."""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle.sfac_oracle import Cell, LofS, OracleSfac  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ew():
    from ewaldcorrel_amd import ewald
    return ewald


L6 = np.array([20., 22., 24., 1., -2., 1.5])
K = (24, 25, 30)
H = K[0] * K[1] * (K[2] // 2 + 1)


def frame(seed, n=400):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (n, 3)) @ LofS(L6)
    w = rng.normal(size=n)
    return w, x


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_accum_S2_over_frames(precision, tol):
    s = ew().Sfac(L6, np.array(K), 4, precision=precision)
    o = OracleSfac(L6, K, 4)
    acc = torch.zeros(H, dtype=torch.float64, device="cuda")
    ref = np.zeros(H)
    for seed in range(3):
        w, x = frame(seed)
        s.sfac(w, x)
        s.accum_S2(acc.data_ptr())
        o.sfac(w, x)
        ref += np.abs(o.get_S().reshape(-1)) ** 2
    s.sync()
    got = acc.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say(precision, "accum |S|^2 rel err", err)
    assert err < tol


def test_accum_cross_species():
    """gmx_ewald.py:99-104: C += Q_i conj(Q_j) for two species."""
    a = ew().Sfac(L6, np.array(K), 4)
    b = ew().Sfac(L6, np.array(K), 4)
    oa, ob = OracleSfac(L6, K, 4), OracleSfac(L6, K, 4)
    acc = torch.zeros(2 * H, dtype=torch.float64, device="cuda")
    ref = np.zeros(H, dtype=complex)
    for seed in range(2):
        wa, xa = frame(10 + seed, 300)
        wb, xb = frame(20 + seed, 200)
        a.sfac(wa, xa)
        b.sfac(wb, xb)
        a.accum_cross(b, acc.data_ptr())
        oa.sfac(wa, xa)
        ob.sfac(wb, xb)
        ref += oa.get_S().reshape(-1) * np.conj(ob.get_S().reshape(-1))
    a.sync()
    got = acc.cpu().numpy().reshape(H, 2)
    got = got[:, 0] + 1j * got[:, 1]
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say("cross rel err", err)
    assert err < 1e-12


def test_radial_average():
    """plot.py: shell sums by |k| = 2 pi |m| with dk bins (the stored
    half-complex points once each; K/2 taken as the positive index)."""
    s = ew().Sfac(L6, np.array(K), 4)
    rng = np.random.default_rng(5)
    val = rng.uniform(0, 1, H)
    nb, dk = 40, 0.05
    vd = torch.tensor(val, device="cuda")
    sm = torch.zeros(nb, dtype=torch.float64, device="cuda")
    ct = torch.zeros(nb, dtype=torch.float64, device="cuda")
    s.radial(vd.data_ptr(), nb, dk, sm.data_ptr(), ct.data_ptr())
    s.radial(vd.data_ptr(), nb, dk, sm.data_ptr(), ct.data_ptr())  # accumulates
    s.sync()
    iL = Cell(L6).iL
    K0, K1, K2 = K
    nx = np.arange(K0)
    nx = np.where(nx <= K0 // 2, nx, nx - K0)
    ny = np.arange(K1)
    ny = np.where(ny <= K1 // 2, ny, ny - K1)
    nz = np.arange(K2 // 2 + 1)
    NX, NY, NZ = np.meshgrid(nx, ny, nz, indexing="ij")
    mx = NX * iL[0, 0]
    my = NX * iL[1, 0] + NY * iL[1, 1]
    mz = NX * iL[2, 0] + NY * iL[2, 1] + NZ * iL[2, 2]
    k = 2 * np.pi * np.sqrt(mx ** 2 + my ** 2 + mz ** 2).reshape(-1)
    b = np.floor(k / dk).astype(int)
    keep = b < nb
    rs = 2 * np.bincount(b[keep], weights=val[keep], minlength=nb)
    rc = 2 * np.bincount(b[keep], minlength=nb)
    gs, gc = sm.cpu().numpy(), ct.cpu().numpy()
    say("radial", np.abs(gs - rs).max(), np.abs(gc - rc).max())
    assert np.array_equal(gc, rc)
    assert np.abs(gs - rs).max() <= 1e-12 * np.abs(rs).max()


def test_accumulation_requires_sfac(capfd):
    s = ew().Sfac(L6, np.array(K), 4)
    acc = torch.zeros(H, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError):
        s.accum_S2(acc.data_ptr())
    assert "You must call sfac" in capfd.readouterr().out


def test_accumulate_sk_over_xyz_frames(tmp_path):
    """run.py: |S|^2 of neutron-weighted frames read from .xyz, summed on
    the device, equals the host sum over the oracle's S-hat."""
    from ewaldcorrel_amd.traj import accumulate_sk, neutron_weights, read_xyz
    rng = np.random.default_rng(9)
    names = ["O", "H", "H", "C"] * 50
    text = ""
    frames = []
    for f in range(3):
        x = rng.uniform(0, 20, (len(names), 3))
        frames.append(x)
        text += f"{len(names)}\nframe {f}\n" + "".join(
            f"{nm} {a:.10f} {b:.10f} {c:.10f}\n" for nm, (a, b, c) in zip(names, x))
    p = tmp_path / "t.xyz"
    p.write_text(text)
    L6c = np.array([20., 20., 20., 0., 0., 0.])
    s = ew().Sfac(L6c, np.array([24, 24, 24]), 4)
    acc, nfr = accumulate_sk(s, read_xyz(p))
    o = OracleSfac(L6c, (24, 24, 24), 4)
    ref = 0
    for nm_x in read_xyz(p):
        o.sfac(neutron_weights(nm_x[0]), nm_x[1])
        ref = ref + np.abs(o.get_S()) ** 2
    err = np.abs(acc - ref).max() / np.abs(ref).max()
    say("xyz frames", nfr, "rel err", err)
    assert nfr == 3 and err < 1e-12


def test_radial_plot_matches_reference_plot_py():
    """The device |k| shells in plot.py's convention (sfac_radial_plot via
    Sfac.plot_spectrum) reproduce the reference's plot.py output stored in
    tests/golden/radial_plot.json, on the array it was made from."""
    import json
    import os
    from conftest import ROOT
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_radial_golden import make_input
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "radial_plot.json")))
    shape = tuple(g["shape"])
    f = make_input(shape, g["seed"]).real
    L = g["L"]
    s = ew().Sfac(np.array([L, L, L, 0., 0., 0.]), np.array([shape[0], shape[1], 2 * (shape[2] - 1)]),
                  4)
    rows = s.plot_spectrum(f)
    say("plot rows", len(rows), "fixture", len(g["rows"]))
    assert len(rows) == len(g["rows"])
    for (k, m, c), (k0, m0, c0) in zip(rows, g["rows"]):
        assert k == k0 and c == c0 and abs(m - m0) <= 1e-15 * abs(m0)


def test_trr_frames_through_device_accumulation():
    """gmx_ewald.py:65-107 on the committed .trr trajectory
    (tests/golden/water16.trr): the handle is made from the first frame's
    box (SofL), each frame's oxygens and hydrogens are spread as two
    species, and |S_O|^2, |S_H|^2 and S_O conj(S_H) are summed over the
    frames on the device; against the oracle's sums of the same frames."""
    import os
    from conftest import ROOT
    from ewaldcorrel_amd.traj import read_trr
    from oracle.sfac_oracle import SofL
    frames = list(read_trr(os.path.join(ROOT, "tests", "golden", "water16.trr")))
    L6w = SofL(frames[0]["box"])
    Kw = (20, 21, 22)
    Hw = Kw[0] * Kw[1] * (Kw[2] // 2 + 1)
    hO = ew().Sfac(L6w, np.array(Kw), 4)
    hH = ew().Sfac(L6w, np.array(Kw), 4)
    oO, oH = OracleSfac(L6w, Kw, 4), OracleSfac(L6w, Kw, 4)
    aO = torch.zeros(Hw, dtype=torch.float64, device="cuda")
    aH = torch.zeros(Hw, dtype=torch.float64, device="cuda")
    aOH = torch.zeros(2 * Hw, dtype=torch.float64, device="cuda")
    rO = rH = rOH = 0
    for fr in frames:
        xO = fr["x"][0::3]
        xH = np.concatenate([fr["x"][1::3], fr["x"][2::3]])
        wO, wH = np.full(len(xO), -0.834), np.full(len(xH), 0.417)
        hO.sfac(wO, xO)
        hH.sfac(wH, xH)
        torch.cuda.synchronize()
        hO.accum_S2(aO.data_ptr())
        hH.accum_S2(aH.data_ptr())
        hO.accum_cross(hH, aOH.data_ptr())
        oO.sfac(wO, xO)
        oH.sfac(wH, xH)
        SO, SH = oO.get_S().ravel(), oH.get_S().ravel()
        rO = rO + np.abs(SO) ** 2
        rH = rH + np.abs(SH) ** 2
        rOH = rOH + SO * np.conj(SH)
    hO.sync()
    hH.sync()
    gOH = aOH.cpu().numpy()
    gOH = gOH[0::2] + 1j * gOH[1::2]
    eO = np.abs(aO.cpu().numpy() - rO).max() / np.abs(rO).max()
    eH = np.abs(aH.cpu().numpy() - rH).max() / np.abs(rH).max()
    eOH = np.abs(gOH - rOH).max() / np.abs(rOH).max()
    say("trr frames", len(frames), eO, eH, eOH)
    assert len(frames) == 3 and max(eO, eH, eOH) <= 1e-12
