"""The spreading sort (x-plane binning, per-plane column counting sort) on
inputs that stress it: planes more crowded than the sort stages in LDS
(sorted straight into place), every atom in one column, most planes empty,
one atom.  Each case
against the FP64 oracle, and in the deterministic modes bitwise invariant
under a permutation of the atoms (the order within a column is the order
of the sort's atomics, which no result may depend on)."""
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


L6 = np.array([24., 22., 26., 1.5, -1.0, 0.8])
K = (30, 28, 32)


def crowded(n, seed):
    """Most atoms in two x planes (far beyond a uniform share), the rest
    spread uniformly."""
    rng = np.random.default_rng(seed)
    s = rng.uniform(0, 1, (n, 3))
    m = rng.uniform(0, 1, n) < 0.8
    s[m, 0] = rng.uniform(0.40, 0.40 + 1.9 / K[0], m.sum())
    w = rng.normal(size=n)
    w -= w.mean()
    return w, s @ LofS(L6)


def one_column(n, seed):
    rng = np.random.default_rng(seed)
    s = rng.uniform(0, 1, (n, 3))
    s[:, 0] = 0.31 + rng.uniform(0, 0.9 / K[0], n)
    s[:, 1] = 0.77 + rng.uniform(0, 0.9 / K[1], n)
    w = rng.normal(size=n)
    w -= w.mean()
    return w, s @ LofS(L6)


CASES = {
    "crowded_planes": lambda: crowded(6000, 1),
    "one_column": lambda: one_column(700, 2),
    "one_atom": lambda: (np.array([0.7]), np.array([[3.1, 7.2, 11.9]])),
}


@pytest.fixture(scope="module")
def oracle():
    o = OracleSfac(L6, K, 4)
    o.set_A(ewald_rad(0.4), 1.0)
    return o


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("precision,fixed", [("fp64", True), ("fp32", True), ("fp64", False)])
def test_sort_stress(case, precision, fixed, oracle):
    w, x = CASES[case]()
    say("sort case", case, precision, fixed, "n", len(w))
    E0, v0, dx0, phi0 = oracle.evaluate(w, x)
    s = ew().Sfac(L6, np.array(K), 4, precision=precision, fixed=fixed)
    s.set_A(ew().ewald_kernel(0.4), 1.0)
    E, vir, dx, phi = s.evaluate(w, x)
    eE = abs(E - E0) / max(abs(E0), 1e-300)
    say(" errors", eE, rel(vir, v0), rel_rms(dx, dx0), rel_rms(phi, phi0))
    if precision == "fp64":
        assert eE <= 1e-10 and rel(vir, v0) <= 1e-10
        assert rel_rms(dx, dx0) <= 1e-9 and rel_rms(phi, phi0) <= 1e-9
    else:
        assert eE <= 1e-6 and rel(vir, v0) <= 1e-5 and rel_rms(phi, phi0) <= 1e-5
        # a lone atom's force is only the grid's discretization residue
        # (~1e-4 of a Coulomb force at the box size): no relative FP32 bar
        if len(w) > 1:
            assert rel_rms(dx, dx0) <= 1e-5
    if fixed and len(w) > 1:
        # the same atoms in another order: bitwise the same results per atom
        perm = np.random.default_rng(7).permutation(len(w))
        E2, vir2, dx2, phi2 = s.evaluate(w[perm], x[perm])
        assert E2 == E and np.array_equal(vir2, vir)
        assert np.array_equal(dx2, dx[perm]) and np.array_equal(phi2, phi[perm])
