"""Spreading and gather tiles split into z chunks (each workgroup owns a range
of z cells of its columns) against whole-column tiles.

The FP32 engine's spreading sums int32 contributions into int64 cells, so the
grid, and with it every output, is bitwise the same whatever the tiling; the
gather sums each atom's stencil from the same staged values in the same
order.  FP64 with LDS float atomics agrees within 1e-12 and matches the
oracle within its bars.  Chunk sizes that do not divide K2, chunks whose
stencils wrap around z = 0, and atoms sitting exactly on cell faces are
covered.  Shapes are forced through the tuned-shape record (tune_import),
which handles made afterwards start from."""
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


def system(L6, K, order, n, seed):
    rng = np.random.default_rng(seed)
    L6 = np.asarray(L6, dtype=float)
    s = rng.uniform(0, 1, (n, 3))
    # some atoms exactly on grid planes in z (fraction 0: the gather's ceil cell)
    s[: n // 10, 2] = rng.integers(0, K[2], n // 10) / K[2]
    x = s @ LofS(L6)
    w = rng.normal(size=n)
    w -= w.mean()
    return L6, tuple(K), order, w, x


def run(sysd, precision, record=None):
    L6, K, order, w, x = sysd
    ew().tune_clear()
    if record:
        flags = 3 if precision == "fp32" else 0
        assert ew().tune_import(f"{K[0]}x{K[1]}x{K[2]}/{order}/{flags} - - - {record}\n") == 1
    s = ew().Sfac(L6, np.array(K), order, precision=precision)
    if record:  # the record applied: the handle runs these tiles
        tiles = s.launch_shapes().split()[4:7]
        want = [t if t.count("x") == 2 else t + "x0" for t in record.split()]
        assert tiles[:2] == want[:2], (tiles, record)
    s.set_A(ew().ewald_kernel(0.35), 1.0)
    r = s.evaluate(w, x)
    ew().tune_clear()
    return r


CASES = [
    ("k96_o4", [24., 20., 60., 1., -2., 3.], (24, 20, 96), 4, 4000, 1),
    ("k60_o6", [18., 16., 40., 0., 0., 0.], (16, 16, 60), 6, 2500, 2),
    ("k50_o8", [15., 15., 30., 2., 0., -1.], (15, 15, 50), 8, 1500, 3),
]
# spreading tile TXxTYxTZ, gather tile GTXxGTYxGTZ (z = 0: whole columns)
RECORDS = ["8x8x32 8x8x32 0", "4x8x40 8x4x24 0", "8x8x16 8x8x0 0", "8x4x0 4x4x28 0"]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fp32_z_chunks_are_bitwise_whole_columns(case):
    name, L6, K, order, n, seed = case
    sysd = system(L6, K, order, n, seed)
    ref = run(sysd, "fp32", "8x8x0 8x8x0 0")
    for rec in RECORDS:
        if order >= 8 and rec.startswith("8x8x16"):
            continue  # wider stencils need chunks of at least `order` cells
        got = run(sysd, "fp32", rec)
        same = (got[0] == ref[0] and np.array_equal(got[1], ref[1]) and
                np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3]))
        say(f"{name} fp32 {rec}: E {got[0]:.15g} bitwise {same}")
        assert same, rec


@pytest.mark.parametrize("case", CASES[:2], ids=[c[0] for c in CASES[:2]])
def test_fp64_z_chunks_and_gather_tiles_match_and_meet_the_oracle(case):
    """FP64: chunked spreading, and gather tiles (whole columns and chunked)
    in place of the per-atom gather."""
    name, L6, K, order, n, seed = case
    sysd = system(L6, K, order, n, seed)
    ref = run(sysd, "fp64", "8x8x0 0x0x0 0")
    L6, K, order, w, x = sysd
    o = OracleSfac(L6, K, order)
    o.set_A(ewald_rad(0.35), 1.0)
    E0, v0, dx0, phi0 = o.evaluate(w, x)
    for rec in ("8x4x24 0x0x0 0", "8x8x0 4x4x0 0", "8x4x24 8x4x32 0"):
        got = run(sysd, "fp64", rec)
        eE = abs(got[0] - ref[0]) / abs(ref[0])
        say(f"{name} fp64 {rec}: relE {eE:.1e} F {rel_rms(got[2], ref[2]):.1e} "
            f"phi {rel_rms(got[3], ref[3]):.1e}")
        assert eE <= 1e-12 and rel(got[1], ref[1]) <= 1e-12
        assert rel_rms(got[2], ref[2]) <= 1e-12 and rel_rms(got[3], ref[3]) <= 1e-12
        assert abs(got[0] - E0) <= 1e-10 * abs(E0) and rel(got[1], v0) <= 1e-10
        assert rel_rms(got[2], dx0) <= 1e-9 and rel_rms(got[3], phi0) <= 1e-9


def test_default_shapes_chunk_long_columns():
    """Columns too long for a whole-column gather tile get chunked gather
    tiles by default (spreading keeps whole columns), and the result matches
    the oracle."""
    L6, K, order = np.array([20., 20., 200., 0., 0., 0.]), (16, 16, 384), 6
    rng = np.random.default_rng(9)
    n = 3000
    x = rng.uniform(0, 1, (n, 3)) @ LofS(L6)
    w = rng.normal(size=n)
    w -= w.mean()
    ew().tune_clear()
    o = OracleSfac(L6, K, order)
    o.set_A(ewald_rad(0.35), 1.0)
    E0, v0, dx0, phi0 = o.evaluate(w, x)
    for prec in ("fp32", "fp64"):
        s = ew().Sfac(L6, np.array(K), order, precision=prec)
        shapes = s.launch_shapes()
        say(prec, shapes)
        assert shapes.split()[4] == "4x4x0"  # whole-column spreading tiles
        if prec == "fp32":
            assert shapes.split()[5] == "8x8x77"  # five gather chunks
        s.set_A(ew().ewald_kernel(0.35), 1.0)
        E, vir, dx, phi = s.evaluate(w, x)
        eE, eF = abs(E - E0) / abs(E0), rel_rms(dx, dx0)
        say(f"long columns {prec}: relE {eE:.1e} F {eF:.1e} phi {rel_rms(phi, phi0):.1e}")
        if prec == "fp64":
            assert eE <= 1e-10 and rel(vir, v0) <= 1e-10 and eF <= 1e-9
        else:
            assert eE <= 1e-6 and eF <= 1e-5
