"""The uneven-frame batch and the serving stream at their configured sizes.

* Uneven frames: 256 systems of 500-20,000 atoms (log-uniform, seed 4),
  cubic boxes at 0.1 atoms/A^3, grids 32^3-60^3 (2^a 3^b 5^c), order 4, in
  one FP32 batch (one launch per stage), autotuned.  Every frame's result
  equals its lone evaluation bitwise; the largest, the smallest and a 60^3
  frame match the FP64 oracle (E <= 1e-6, forces <= 1e-5 relative RMS).
* Serving stream: requests drawn from the stream's own generator (1k-500k
  atoms log-uniform, triclinic boxes with angles in 70-110 deg, order 4 or
  6, FP32 or FP64), among them a triclinic order-6 FP64 request of at least
  200k atoms and several FP32 ones, submitted with Poisson gaps (2,000/s)
  to a server with three slots and batches of four.  Each result equals the
  same request served alone (bitwise in FP32, 1e-12 for FP64 LDS atomics)
  and matches the FP64 oracle (FP64: E, virial <= 1e-10, forces, potentials
  <= 1e-9; FP32: E <= 1e-6, forces <= 1e-5).
"""
import sys
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from conftest import rel, rel_rms  # noqa: E402
from oracle.sfac_oracle import OracleSfac, ewald_rad  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ew():
    from ewaldcorrel_amd import ewald
    return ewald


def oracle_eval(sy):
    o = OracleSfac(sy.L6, sy.K, sy.order)
    o.set_A(ewald_rad(sy.beta), 1.0)
    return o.evaluate(sy.w, sy.x)


def test_uneven_frames_batch_at_size():
    from ewaldcorrel_amd import workloads as wl
    frames = wl.uneven_frames(256, seed=4)
    ns = np.array([f.n for f in frames])
    ks = np.array([f.K[0] for f in frames])
    say("uneven frames", len(frames), "atoms", ns.sum(), "n", ns.min(), ns.max(),
        "grids", sorted(set(ks.tolist())))
    dev = torch.device("cuda:0")
    hs, bufs = [], []
    for sy in frames:
        s = ew().Sfac(sy.L6, np.array(sy.K), 4, precision="fp32")
        s.set_A(ew().ewald_kernel(sy.beta), 1.0)
        hs.append(s)
        n = sy.n
        bufs.append(dict(n=n, w=torch.tensor(sy.w, device=dev), x=torch.tensor(sy.x, device=dev),
                         out=torch.zeros(8, dtype=torch.float64, device=dev),
                         dx=torch.zeros((n, 3), dtype=torch.float64, device=dev),
                         phi=torch.zeros(n, dtype=torch.float64, device=dev)))
    b = ew().SfacBatch(hs)
    for f, bf in enumerate(bufs):
        b.set(f, bf["n"], bf["w"].data_ptr(), bf["x"].data_ptr(), bf["out"].data_ptr(),
              bf["dx"].data_ptr(), bf["phi"].data_ptr())
    t0 = time.perf_counter()
    rep = b.autotune(reps=3)
    say("autotune", round(time.perf_counter() - t0, 2), "s;", rep.strip().splitlines()[-1])
    torch.cuda.synchronize()
    for _ in range(2):  # the second evaluation replays the captured graph
        b.eval()
    b.sync()
    got = [tuple(bf[k].cpu().numpy().copy() for k in ("out", "dx", "phi")) for bf in bufs]
    # each frame alone, on its own handle (same launch shapes after tuning)
    for f, (sy, s) in enumerate(zip(frames, hs)):
        E, vir, dx, phi = s.evaluate(sy.w, sy.x)
        out, gdx, gphi = got[f]
        assert out[0] == E and np.array_equal(out[1:7], vir), f
        assert np.array_equal(gdx, dx) and np.array_equal(gphi, phi), f
    # the oracle on the largest, the smallest and a 60^3 frame
    picks = {int(np.argmax(ns)), int(np.argmin(ns)), int(np.nonzero(ks == 60)[0][0])}
    for f in sorted(picks):
        sy = frames[f]
        E0, v0, dx0, phi0 = oracle_eval(sy)
        out, gdx, gphi = got[f]
        eE = abs(out[0] - E0) / abs(E0)
        say("frame", f, "n", sy.n, "K", sy.K, "errors", eE, rel(out[1:7], v0),
            rel_rms(gdx, dx0), rel_rms(gphi, phi0))
        assert eE <= 1e-6 and rel(out[1:7], v0) <= 1e-5
        assert rel_rms(gdx, dx0) <= 1e-5 and rel_rms(gphi, phi0) <= 1e-5
    b.close()


def stream_requests(seed=5):
    """Requests of the serving generator: the first triclinic order-6 FP64
    one of at least 200k atoms, four FP32 ones and two more FP64 ones of at
    most 100k atoms, in draw order, with their Poisson arrival gaps."""
    from ewaldcorrel_amd import workloads as wl
    rng = np.random.default_rng(seed)
    big, f32, f64 = [], [], []
    for i in range(2000):
        sy = wl.serving_request(rng, 1000, 500000, f"req{i}")
        p = sy.meta["precision"]
        if not big and p == "fp64" and sy.order == 6 and sy.n >= 200000:
            big.append(sy)
        elif p == "fp32" and len(f32) < 4 and sy.n <= 100000:
            f32.append(sy)
        elif p == "fp64" and len(f64) < 2 and sy.n <= 100000:
            f64.append(sy)
        if big and len(f32) == 4 and len(f64) == 2:
            break
    reqs = sorted(big + f32 + f64, key=lambda s: int(s.name[3:]))
    gaps = rng.exponential(1.0 / 2000.0, len(reqs))
    return reqs, gaps


def served(srv, sy):
    return srv.submit(sy.L6, sy.K, sy.order, sy.w, sy.x, sy.beta, 1.0,
                      precision=sy.meta["precision"])


def test_serving_stream_at_size():
    reqs, gaps = stream_requests()
    say("requests", [(s.name, s.n, s.K, s.order, s.meta["precision"],
                      tuple(round(a, 1) for a in s.meta["angles"])) for s in reqs])
    assert any(s.n >= 200000 and s.order == 6 and s.meta["precision"] == "fp64" for s in reqs)
    assert sum(s.meta["precision"] == "fp32" for s in reqs) >= 4
    alone_srv = ew().SfacServer(slots=1, max_batch=1)
    alone = []
    for sy in reqs:
        r = alone_srv.wait(served(alone_srv, sy))
        alone.append(tuple(np.copy(v) if isinstance(v, np.ndarray) else v for v in r))
    alone_srv.close()
    srv = ew().SfacServer(slots=3, max_batch=4)
    qs = []
    for sy, g in zip(reqs, gaps):  # Poisson arrivals while earlier requests run
        time.sleep(g)
        qs.append(served(srv, sy))
    mix = [srv.wait(q) for q in qs]
    st = srv.stats()
    say("stats", st)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    srv.close()
    for sy, a, m in zip(reqs, alone, mix):
        if sy.meta["precision"] == "fp32":
            assert a[0] == m[0] and np.array_equal(a[1], m[1]), sy.name
            assert np.array_equal(a[2], m[2]) and np.array_equal(a[3], m[3]), sy.name
        else:
            assert abs(a[0] - m[0]) <= 1e-12 * abs(a[0]) and rel(m[1], a[1]) <= 1e-12
            assert rel_rms(m[2], a[2]) <= 1e-12 and rel_rms(m[3], a[3]) <= 1e-12
        E0, v0, dx0, phi0 = oracle_eval(sy)
        eE = abs(m[0] - E0) / abs(E0)
        say(sy.name, sy.n, sy.K, sy.order, sy.meta["precision"], "errors", eE, rel(m[1], v0),
            rel_rms(m[2], dx0), rel_rms(m[3], phi0))
        if sy.meta["precision"] == "fp64":
            assert eE <= 1e-10 and rel(m[1], v0) <= 1e-10
            assert rel_rms(m[2], dx0) <= 1e-9 and rel_rms(m[3], phi0) <= 1e-9
        else:
            assert eE <= 1e-6 and rel(m[1], v0) <= 1e-5
            assert rel_rms(m[2], dx0) <= 1e-5 and rel_rms(m[3], phi0) <= 1e-5
