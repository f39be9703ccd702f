"""Interleaving the entry points of one handle, of two handles and of a
batch with the handles it holds.

* The fused evaluation is replayed from a captured graph that reads the
  handle's device descriptor; the split entry points (potl, de2, sfac)
  upload their own descriptors in between, so every replay must upload the
  evaluation's descriptor again.  Each sequence is compared bitwise with a
  handle that launches kernel by kernel (no graph).
* Cross terms read the other handle's grid on this handle's stream; the
  other handle's next spread must wait for that read.
* A batch keeps a snapshot of each frame's descriptor; a new box or table on
  a member handle (set_L, set_A) must reach the batch's next evaluation.
"""
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


def system(seed, n=800, L6=(20., 18., 22., 3., -2., 1.5)):
    rng = np.random.default_rng(seed)
    L6 = np.asarray(L6, dtype=float)
    x = rng.uniform(0, 1, (n, 3)) @ LofS(L6)
    w = rng.normal(size=n)
    w -= w.mean()
    return L6, w, x


def pair(L6, K, precision, fixed, beta=0.4):
    out = []
    for graph in (True, False):
        s = ew().Sfac(L6, np.array(K), 4, precision=precision, fixed=fixed)
        s.set_A(ew().ewald_kernel(beta), 1.0)
        s.set_graph(graph)
        out.append(s)
    return out


@pytest.mark.parametrize("precision,fixed", [("fp32", True), ("fp64", True)])
def test_evaluate_potl_evaluate(precision, fixed):
    """sfac_eval, potl (other positions), sfac_eval: the second evaluation
    is a graph replay and must see its own atoms again."""
    L6, w, x = system(11)
    _, w2, x2 = system(12, n=300)
    K = (24, 25, 30)
    g, ref = pair(L6, K, precision, fixed)
    res = []
    for s in (g, ref):
        a = s.evaluate(w, x)
        phi = np.zeros(len(w2))
        s.potl(len(w2), x2, phi)
        b = s.evaluate(w, x)
        c = s.evaluate(w, x)  # a replay with nothing in between
        res.append((a, phi, b, c))
    say(precision, "E", res[0][2][0], res[1][2][0])
    (ga, gp, gb, gc), (ra, rp, rb, rc) = res
    assert np.array_equal(gp, rp)
    for u, v in ((ga, ra), (gb, rb), (gc, rc), (gb, ra)):
        assert u[0] == v[0]
        for k in (1, 2, 3):
            assert np.array_equal(u[k], v[k])


@pytest.mark.parametrize("precision,fixed", [("fp32", True), ("fp64", True)])
def test_eval_dev_de2_dev_eval_dev(precision, fixed):
    """eval_dev, de2_dev into another force buffer, eval_dev: the replayed
    evaluation writes its forces into its own buffer, not de2's."""
    L6, w, x = system(13)
    n = len(w)
    K = (24, 25, 30)
    dev = torch.device("cuda:0")
    wd = torch.tensor(w, device=dev)
    xd = torch.tensor(x, device=dev)
    results = []
    for s in pair(L6, K, precision, fixed):
        out = torch.zeros(8, dtype=torch.float64, device=dev)
        dx = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        phi = torch.zeros(n, dtype=torch.float64, device=dev)
        dx2 = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        s.eval_dev(n, wd.data_ptr(), xd.data_ptr(), out.data_ptr(), dx.data_ptr(),
                   phi.data_ptr())
        s.sync()
        first = (out.cpu().numpy().copy(), dx.cpu().numpy().copy(), phi.cpu().numpy().copy())
        # de2 into another buffer with half the charges (state 1 after eval_dev)
        wd2 = (wd * 0.5).contiguous()
        torch.cuda.synchronize()
        s.de2_dev(n, wd2.data_ptr(), xd.data_ptr(), dx2.data_ptr())
        s.sync()
        de2 = dx2.cpu().numpy().copy()
        dx.zero_()
        phi.zero_()
        out.zero_()
        torch.cuda.synchronize()
        s.eval_dev(n, wd.data_ptr(), xd.data_ptr(), out.data_ptr(), dx.data_ptr(),
                   phi.data_ptr())
        s.sync()
        second = (out.cpu().numpy(), dx.cpu().numpy(), phi.cpu().numpy())
        results.append((first, de2, second, dx2.cpu().numpy()))
    (gf, gd, gs, gdx2), (rf, rd, rs, rdx2) = results
    for k in range(3):
        assert np.array_equal(gf[k], rf[k])
        assert np.array_equal(gs[k], rs[k])
        assert np.array_equal(gs[k], gf[k])
    assert np.array_equal(gd, rd)
    # the replay did not touch de2's buffer: it still holds half the forces
    assert np.array_equal(gdx2, gd)
    assert rel_rms(gd, 0.5 * gf[1]) <= (1e-6 if precision == "fp32" else 1e-12)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_cross_terms_other_handle_first(precision):
    """Per frame: sfac on b first, then a, then the cross term on a's
    stream; b's spread of the next frame must not overwrite b's grid while
    the cross kernel reads it.  Checked against the oracle's S_a conj(S_b)
    summed over frames."""
    L6 = np.array([30., 31., 29., 0., 0., 0.])
    K = (40, 40, 40)
    a = ew().Sfac(L6, np.array(K), 4, precision=precision)
    b = ew().Sfac(L6, np.array(K), 4, precision=precision)
    oa = OracleSfac(L6, K, 4)
    ob = OracleSfac(L6, K, 4)
    H = K[0] * K[1] * (K[2] // 2 + 1)
    dev = torch.device("cuda:0")
    acc = torch.zeros(2 * H, dtype=torch.float64, device=dev)
    ref = np.zeros(H, dtype=complex)
    rng = np.random.default_rng(21)
    n = 6000
    for fr in range(5):
        xa = rng.uniform(0, 1, (n, 3)) @ LofS(L6)
        xb = rng.uniform(0, 1, (n, 3)) @ LofS(L6)
        wa = rng.normal(size=n)
        wb = rng.normal(size=n)
        b.sfac(wb, xb)
        a.sfac(wa, xa)
        torch.cuda.synchronize()
        a.accum_cross(b, acc.data_ptr())
        oa.sfac(wa, xa)
        ob.sfac(wb, xb)
        ref += (oa.get_S() * np.conj(ob.get_S())).ravel()
    a.sync()
    b.sync()
    got = acc.cpu().numpy()
    got = got[0::2] + 1j * got[1::2]
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    say(precision, "cross err", err)
    assert err <= (1e-12 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("precision,fixed", [("fp32", True), ("fp64", True)])
def test_batch_sees_new_box_and_table(precision, fixed):
    """A batch of two frames; set_L (which also rebuilds the table in the
    Python binding) on frame 0 between two batch evaluations: the second
    evaluation equals frame 0 evaluated alone in the new box, bitwise."""
    L6a, wa, xa = system(31, n=1500)
    L6b, wb, xb = system(32, n=900, L6=(19., 20., 21., 0., 0., 0.))
    K = (24, 25, 30)
    fa = ew().Sfac(L6a, np.array(K), 4, precision=precision, fixed=fixed)
    fb = ew().Sfac(L6b, np.array(K), 4, precision=precision, fixed=fixed)
    for f in (fa, fb):
        f.set_A(ew().ewald_kernel(0.4), 1.0)
    dev = torch.device("cuda:0")
    bufs = []
    for w, x in ((wa, xa), (wb, xb)):
        n = len(w)
        bufs.append(dict(n=n, w=torch.tensor(w, device=dev), x=torch.tensor(x, device=dev),
                         out=torch.zeros(8, dtype=torch.float64, device=dev),
                         dx=torch.zeros((n, 3), dtype=torch.float64, device=dev),
                         phi=torch.zeros(n, dtype=torch.float64, device=dev)))
    bt = ew().SfacBatch([fa, fb])
    for i, bf in enumerate(bufs):
        bt.set(i, bf["n"], bf["w"].data_ptr(), bf["x"].data_ptr(), bf["out"].data_ptr(),
               bf["dx"].data_ptr(), bf["phi"].data_ptr())
    torch.cuda.synchronize()
    bt.eval()
    bt.sync()
    E_old = bufs[0]["out"].cpu().numpy()[0]
    # new box for frame 0 (atoms keep their Cartesian positions)
    L6n = L6a * np.array([1.05, 0.97, 1.02, 1.0, 1.0, 1.0])
    fa.set_L(L6n)
    for _ in range(2):  # the first evaluation re-captures, the second replays
        torch.cuda.synchronize()
        bt.eval()
        bt.sync()
    got = [bufs[0][k].cpu().numpy() for k in ("out", "dx", "phi")]
    alone = ew().Sfac(L6n, np.array(K), 4, precision=precision, fixed=fixed)
    alone.set_A(ew().ewald_kernel(0.4), 1.0)
    E, vir, dx, phi = alone.evaluate(wa, xa)
    say(precision, "E old", E_old, "new", got[0][0], "alone", E)
    assert got[0][0] != E_old
    assert got[0][0] == E and np.array_equal(got[0][1:7], vir)
    assert np.array_equal(got[1], dx) and np.array_equal(got[2], phi)
    # frame 1 is unchanged by frame 0's box
    E1, vir1, dx1, phi1 = fb.evaluate(wb, xb)
    assert bufs[1]["out"].cpu().numpy()[0] == E1
    bt.close()
