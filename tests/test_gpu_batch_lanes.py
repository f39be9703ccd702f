"""A batch evaluated as groups of frames on their own streams (forked from
and joined into the evaluation's stream): every frame bitwise as it
evaluates alone and as in one group, over graph captures and replays, after
new atoms for a frame, without graphs, and with frame 0 profiling (then one
group); the tuner's choice between one and two groups; argument checks."""
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
    ew().tune_clear()
    yield


def make_sys(L6, K, order, N, seed, beta=0.4):
    rng = np.random.default_rng(seed)
    L6 = np.asarray(L6, dtype=float)
    x = rng.uniform(-0.2, 1.2, (N, 3)) @ LofS(L6)
    w = rng.normal(size=N)
    w -= w.mean()
    return dict(L6=L6, K=tuple(K), order=order, beta=beta, w=w, x=x)


SYSTEMS = [
    ([30., 31., 29., 0., 0., 0.], (48, 48, 48), 2000, 51),
    ([30., 31., 29., 0., 0., 0.], (48, 48, 48), 1700, 52),
    ([21., 19., 23., 2.5, -1.5, 1.0], (25, 27, 32), 1500, 53),
    ([30., 31., 29., 0., 0., 0.], (48, 48, 48), 2300, 54),
    ([18., 20., 22., 0., 0., 0.], (20, 24, 30), 900, 55),
]


def make_batch(precision, dev):
    systems = [make_sys(L6, K, 4, N, seed) for L6, K, N, seed in SYSTEMS]
    handles, bufs = [], []
    for sy in systems:
        h = ew().Sfac(sy["L6"], np.array(sy["K"]), 4, precision=precision, fixed=True)
        h.set_A(ew().ewald_kernel(sy["beta"]), 1.0)
        handles.append(h)
        bufs.append(new_bufs(sy, dev))
    return systems, handles, bufs


def new_bufs(sy, dev):
    n = len(sy["w"])
    return dict(w=torch.tensor(sy["w"], device=dev), x=torch.tensor(sy["x"], device=dev),
                out=torch.zeros(8, dtype=torch.float64, device=dev),
                dx=torch.zeros((n, 3), dtype=torch.float64, device=dev),
                phi=torch.zeros(n, dtype=torch.float64, device=dev))


def bind(batch, f, sy, b):
    batch.set(f, len(sy["w"]), b["w"].data_ptr(), b["x"].data_ptr(), b["out"].data_ptr(),
              b["dx"].data_ptr(), b["phi"].data_ptr())


def clear(bufs):
    for b in bufs:
        b["out"].zero_()
        b["dx"].zero_()
        b["phi"].zero_()
    torch.cuda.synchronize()


def check_alone(bufs, alone, what):
    for f, (a, b) in enumerate(zip(alone, bufs)):
        o = b["out"].cpu().numpy()
        assert o[0] == a[0] and np.array_equal(o[1:7], a[1]), (what, f)
        assert np.array_equal(b["dx"].cpu().numpy(), a[2]), (what, f)
        assert np.array_equal(b["phi"].cpu().numpy(), a[3]), (what, f)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_lanes_equal_one_group_bitwise(precision):
    dev = torch.device("cuda:0")
    systems, handles, bufs = make_batch(precision, dev)
    alone = [h.evaluate(sy["w"], sy["x"]) for h, sy in zip(handles, systems)]
    batch = ew().SfacBatch(handles)
    assert batch.lanes() == 1
    for f, (sy, b) in enumerate(zip(systems, bufs)):
        bind(batch, f, sy, b)
    for lanes in (1, 2, 3, 4, 2):
        batch.set_lanes(lanes)
        assert batch.lanes() == lanes
        for rep in range(3):  # capture, then replays
            clear(bufs)
            batch.eval(None)
            batch.sync()
            check_alone(bufs, alone, ("lanes", lanes, rep))
        say(precision, "lanes", lanes, "bitwise as alone")
    # on the caller's stream (group 0 there, the others forked and joined)
    s = torch.cuda.Stream()
    clear(bufs)
    batch.eval(s.cuda_stream)
    s.synchronize()
    check_alone(bufs, alone, "caller stream")
    # launch by launch
    batch.set_graph(False)
    clear(bufs)
    batch.eval(None)
    batch.sync()
    check_alone(bufs, alone, "no graph")
    batch.set_graph(True)
    # frame 0 profiling: one group, same results
    handles[0].profile(True)
    clear(bufs)
    batch.eval(None)
    batch.sync()
    prof = handles[0].profile_read()
    handles[0].profile(False)
    check_alone(bufs, alone, "profiling")
    assert prof, "profiled launches of the whole batch"
    batch.close()


def test_lanes_new_atoms_after_grouping():
    """New atoms (another count, other buffers) for frames of the second and
    last group after set_lanes: each frame as it evaluates alone with them."""
    dev = torch.device("cuda:0")
    systems, handles, bufs = make_batch("fp32", dev)
    batch = ew().SfacBatch(handles)
    for f, (sy, b) in enumerate(zip(systems, bufs)):
        bind(batch, f, sy, b)
    batch.set_lanes(2)
    batch.eval(None)
    batch.sync()
    for f, N, seed in ((3, 1100, 61), (4, 1300, 62)):
        L6, K, _, _ = SYSTEMS[f]
        systems[f] = make_sys(L6, K, 4, N, seed)
        bufs[f] = new_bufs(systems[f], dev)
        bind(batch, f, systems[f], bufs[f])
    alone = [h.evaluate(sy["w"], sy["x"]) for h, sy in zip(handles, systems)]
    for rep in range(2):
        clear(bufs)
        batch.eval(None)
        batch.sync()
        check_alone(bufs, alone, ("new atoms", rep))
    batch.close()


def test_lanes_arguments():
    dev = torch.device("cuda:0")
    systems, handles, bufs = make_batch("fp32", dev)
    batch = ew().SfacBatch(handles[:1])
    bind(batch, 0, systems[0], bufs[0])
    batch.set_lanes(3)  # more groups than frames: one group per frame
    batch.eval(None)
    batch.sync()
    for k in (0, 5, -1):
        with pytest.raises(RuntimeError):
            batch.set_lanes(k)
    assert batch.lanes() == 3
    batch.close()


def test_autotune_tries_groups_and_keeps_results():
    """The tuner reports its trial of two groups; after it the frames meet
    the oracle and repeat bitwise, whichever it kept."""
    dev = torch.device("cuda:0")
    systems, handles, bufs = make_batch("fp32", dev)
    batch = ew().SfacBatch(handles)
    for f, (sy, b) in enumerate(zip(systems, bufs)):
        bind(batch, f, sy, b)
    rep = batch.autotune(reps=3)
    say(rep.strip().splitlines()[-2:])
    assert "batch in 2 groups on their own streams" in rep
    lanes = batch.lanes()
    assert lanes in (1, 2)
    clear(bufs)
    batch.eval(None)
    batch.sync()
    first = [(b["out"].cpu().numpy(), b["dx"].cpu().numpy()) for b in bufs]
    clear(bufs)
    batch.eval(None)
    batch.sync()
    for (o, dx), b in zip(first, bufs):
        assert np.array_equal(o, b["out"].cpu().numpy())
        assert np.array_equal(dx, b["dx"].cpu().numpy())
    for f, sy in enumerate(systems):
        o = OracleSfac(sy["L6"], sy["K"], 4)
        o.set_A(ewald_rad(sy["beta"]), 1.0)
        E0, v0, dx0, phi0 = o.evaluate(sy["w"], sy["x"])
        out = bufs[f]["out"].cpu().numpy()
        assert abs(out[0] - E0) <= 1e-6 * abs(E0), f
        assert rel(out[1:7], v0) <= 1e-5, f
        assert rel_rms(bufs[f]["dx"].cpu().numpy(), dx0) <= 1e-5, f
    batch.close()


def test_lanes_follow_a_box_change():
    """set_L (and the table it rebuilds) on frames of the first and the last
    group between two grouped evaluations: each group takes its frames'
    new descriptors, so every frame matches its evaluation alone in the new
    box bitwise."""
    dev = torch.device("cuda:0")
    systems, handles, bufs = make_batch("fp32", dev)
    batch = ew().SfacBatch(handles)
    for f, (sy, b) in enumerate(zip(systems, bufs)):
        bind(batch, f, sy, b)
    batch.set_lanes(2)
    batch.eval(None)
    batch.sync()
    for f, scale in ((0, 1.03), (4, 0.97)):
        L6 = systems[f]["L6"] * scale
        handles[f].set_L(L6)
        systems[f]["L6"] = L6
    alone = [h.evaluate(sy["w"], sy["x"]) for h, sy in zip(handles, systems)]
    for rep in range(2):
        clear(bufs)
        batch.eval(None)
        batch.sync()
        check_alone(bufs, alone, ("box change", rep))
    # and the oracle in the new box for the changed frames
    for f in (0, 4):
        sy = systems[f]
        o = OracleSfac(sy["L6"], sy["K"], 4)
        o.set_A(ewald_rad(sy["beta"]), 1.0)
        E0, v0, dx0, phi0 = o.evaluate(sy["w"], sy["x"])
        E = bufs[f]["out"][0].item()
        say("box change frame", f, abs(E - E0) / abs(E0))
        assert abs(E - E0) <= 1e-6 * abs(E0)
        assert rel_rms(bufs[f]["dx"].cpu().numpy(), dx0) <= 1e-5
    batch.close()
