"""Influence table of the closed-form Ewald kernel built on the device.

set_A (sfac.cpp:213-292) enumerates the reciprocal-lattice ball |m| <= max_m
and adds each point to its aliased grid point, folding the k >= ldim half
onto the conjugate.  The device builder visits, for every grid point, the
residue classes of that point inside the same ball; these tests compare its
A and dA with the oracle's loop-for-loop restatement of the reference on
grids with heavy aliasing (tiny K), triclinic boxes, odd K and a small
cutoff, check the choice between the on-the-fly closed form and the table,
and that the Python-callback path (the reference's own route) builds the
same table."""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle.sfac_oracle import OracleSfac, ewald_rad  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ew():
    from ewaldcorrel_amd import ewald
    return ewald


TABLES = [
    # name, L6, K, eta, max_m, closed form expected
    ("golden_case", [9., 10., 11., 2., -0.1, 1.0], (12, 13, 14), 0.5, 1.0, 0),
    ("tiny_K_aliases", [5., 6., 7., 0., 0., 0.], (2, 3, 1), 0.7, 1.0, 0),
    ("odd_triclinic", [11., 12., 13., 0.5, 0.3, -0.8], (15, 21, 27), 0.45, 1.0, 0),
    ("small_cutoff", [20., 22., 24., 1., 2., -1.], (24, 25, 30), 2.0, 0.3, None),
    ("water_2048", [40., 40., 40., 0., 0., 0.], (48, 48, 48), 0.32, 1.0, 1),
    # skewed box: alias images weigh near the 1e-13 threshold; only the table
    # itself is asserted
    ("fine_triclinic", [30., 28., 26., 4., -3., 2.], (64, 60, 54), 0.35, 1.0, None),
]


@pytest.mark.parametrize("case", TABLES, ids=[t[0] for t in TABLES])
def test_device_table_equals_reference_fold(case):
    name, L6, K, eta, max_m, closed = case
    L6 = np.array(L6)
    s = ew().Sfac(L6, np.array(K), 4)
    s.set_A(ew().ewald_kernel(eta), max_m)
    A, dA = s.table()
    o = OracleSfac(L6, K, 4)
    o.set_A(ewald_rad(eta), max_m)
    Aref = o.A.reshape(-1)
    dref = o.dA.reshape(-1, 6)
    eA = np.abs(A - Aref).max() / np.abs(Aref).max()
    eD = np.abs(dA - dref).max() / np.abs(dref).max()
    mode = s.info()["table_mode"]
    say(name, "A err", eA, "dA err", eD, "table_mode", mode)
    assert eA < 1e-13
    assert eD < 1e-13
    if closed is not None:
        assert mode == closed


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_water_98k_uses_the_closed_form(precision):
    """At the benchmark configuration (100 A cube, 128^3, beta 0.32,
    max_m 1) the aliased images weigh ~exp(-44) and the kernels evaluate the
    closed form instead of reading the 60 MB table."""
    L6 = np.array([100., 100., 100., 0., 0., 0.])
    s = ew().Sfac(L6, np.array([128, 128, 128]), 4, precision=precision)
    s.set_A(ew().ewald_kernel(0.32), 1.0)
    assert s.info()["table_mode"] == 1


def test_callback_route_builds_the_same_table():
    """The reference's route (one rad_fn call per lattice point through
    ctypes) and the device builder give the same A and dA."""
    L6 = np.array([9., 10., 11., 2., -0.1, 1.0])
    K = np.array([12, 13, 14])
    eta = 0.5

    def f(r2, ptr, info):  # not recognised as the native kernel
        return ew().ewald_kernel(eta)(r2, ptr, info)

    a = ew().Sfac(L6, K, 4)
    a.set_A(f, 1.0)
    b = ew().Sfac(L6, K, 4)
    b.set_A(ew().ewald_kernel(eta), 1.0)
    A1, d1 = a.table()
    A2, d2 = b.table()
    say("callback vs device", np.abs(A1 - A2).max(), np.abs(d1 - d2).max())
    assert np.abs(A1 - A2).max() <= 1e-13 * np.abs(A1).max()
    assert np.abs(d1 - d2).max() <= 1e-13 * np.abs(d1).max()


def test_set_L_rebuilds_for_the_new_box():
    """Python set_L re-runs set_A with max_m = 1 (ewald.py:119-122): the
    table follows the box."""
    K = (16, 18, 20)
    L1 = np.array([20., 21., 22., 0., 0., 0.])
    L2 = np.array([22., 20., 21., 1.5, -1., 0.5])
    s = ew().Sfac(L1, np.array(K), 4)
    s.set_A(ew().ewald_kernel(0.6), 1.0)
    s.set_L(L2)
    A, dA = s.table()
    o = OracleSfac(L2, K, 4)
    o.set_A(ewald_rad(0.6), 1.0)
    eA = np.abs(A - o.A.reshape(-1)).max() / np.abs(o.A).max()
    say("after set_L", eA)
    assert eA < 1e-13


def test_table_requires_set_A(capfd):
    s = ew().Sfac(np.array([10., 10., 10., 0., 0., 0.]), np.array([8, 8, 8]), 4)
    with pytest.raises(RuntimeError):
        s.table()
