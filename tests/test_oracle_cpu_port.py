"""The C++ CPU restatement of the reference path (oracle/sfac_cpu.cpp, the
CPU baseline of bench.py) against the numpy oracle and the reference's
golden log: energy and virial to 1e-12, forces and potentials to 1e-11
(relative RMS), at one and at several OpenMP threads."""
import numpy as np
import pytest

from conftest import case_T, rel, rel_rms
from oracle.cpu_port import build, run_port
from oracle.sfac_oracle import LofS, OracleSfac, ewald_rad


@pytest.fixture(scope="module", autouse=True)
def built():
    build()


def system(L6, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.3, 1.3, (n, 3)) @ LofS(np.asarray(L6, float))
    w = rng.normal(size=n)
    return w - w.mean(), x


@pytest.mark.parametrize("L6,K,order,n", [
    ([20., 18., 22., 3., -2., 1.5], (24, 25, 30), 4, 300),
    ([15., 16., 14., 0., 0., 0.], (13, 18, 20), 6, 200),
    ([12., 13., 11., 1., 0.5, -0.7], (16, 12, 15), 8, 150),
])
@pytest.mark.parametrize("threads", [1, 4])
def test_port_matches_oracle(L6, K, order, n, threads):
    w, x = system(L6, n, 3)
    o = OracleSfac(np.array(L6), K, order)
    o.set_A(ewald_rad(0.45), 1.0)
    E0, v0, dx0, phi0 = o.evaluate(w, x)
    rep, (E, vir, dx, phi) = run_port(L6, K, order, 0.45, 1.0, w, x, reps=1, threads=threads)
    assert rep["threads"] == threads
    assert abs(E - E0) <= 1e-12 * abs(E0)
    assert rel(vir, v0) <= 1e-12
    assert rel_rms(dx, dx0) <= 1e-11 and rel_rms(phi, phi0) <= 1e-11


def test_port_golden_log(golden):
    """The reference's own 4-atom case (test.py:64-116, test.log.ref:7-12)."""
    L, K, order, crds, atoms, q = case_T(golden)
    rep, (E, vir, dx, phi) = run_port(L, K, order, 0.5, 1.0, q, atoms)
    g = golden
    assert abs(E - g["E"]) <= 1e-14 * g["E"]
    assert np.allclose(dx.ravel(), g["spl_dx"], atol=6e-9, rtol=0)
    assert np.allclose(vir, g["spl_vir"], rtol=6e-9, atol=0)
