"""The large orthorhombic configuration: 1,048,576 Gaussian charges in a
200 x 200 x 400 A box, 192 x 192 x 384 grid, order 6, FP32 with int64
fixed-point spreading (and FP64 beside it).

At full size the oracle is too slow for a test, so parity there rests on
size-independent properties; the same geometry at 1/8 of the atoms and half
the box and grid is compared with the oracle directly.
  * FP32 against the FP64 engine: E <= 1e-6, forces <= 1e-5 relative RMS
  * E = 1/2 w . phi (the energy is quadratic in the charges)
  * E(2w) = 4 E(w): doubling every charge doubles every spread value, and
    the transforms are linear; bitwise in FP32 (the fixed-point scale
    halves, so the integer grid is the same), within 1e-13 in FP64 (whose
    LDS atomics may add in another order)
  * moving every atom by one grid step along each axis shifts the grid
    cyclically: E, virial unchanged and forces unchanged (FP64, 1e-11)
  * two FP32 runs are bitwise identical (fixed-point spreading)
"""
import sys

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


def handle(sy, precision):
    s = ew().Sfac(sy.L6, np.array(sy.K), sy.order, precision=precision)
    s.set_A(ew().ewald_kernel(sy.beta), sy.max_m)
    return s


@pytest.fixture(scope="module")
def big():
    from ewaldcorrel_amd.workloads import gaussian_1m
    return gaussian_1m()


def test_gaussian_1m_fp32_against_fp64_and_invariants(big):
    sy = big
    s64 = handle(sy, "fp64")
    E64, v64, dx64, phi64 = s64.evaluate(sy.w, sy.x)
    s32 = handle(sy, "fp32")
    E32, v32, dx32, phi32 = s32.evaluate(sy.w, sy.x)
    eE = abs(E32 - E64) / abs(E64)
    eF = rel_rms(dx32, dx64)
    say(f"gaussian_1m: E64 {E64:.12g} fp32 relE {eE:.2e} vir {rel(v32, v64):.2e} "
        f"F {eF:.2e} phi {rel_rms(phi32, phi64):.2e}")
    assert eE <= 1e-6 and eF <= 1e-5
    # E = 1/2 w . phi
    assert abs(0.5 * np.dot(sy.w, phi64) - E64) <= 1e-12 * abs(E64)
    assert abs(0.5 * np.dot(sy.w, phi32) - E32) <= 1e-6 * abs(E32)
    # E(2w) = 4 E(w), exactly
    E64b = s64.evaluate(2 * sy.w, sy.x)[0]
    E32b = s32.evaluate(2 * sy.w, sy.x)[0]
    say(f"E(2w)/4E(w) - 1: fp64 {E64b / (4 * E64) - 1:.1e} fp32 {E32b / (4 * E32) - 1:.1e}")
    assert E32b == 4 * E32  # fixed point: bitwise
    assert abs(E64b - 4 * E64) <= 1e-13 * abs(E64)  # FP64 LDS atomics: order may differ
    # bitwise determinism of the fixed-point path
    a = s32.evaluate(sy.w, sy.x)
    assert a[0] == E32 and np.array_equal(a[2], dx32) and np.array_equal(a[3], phi32)
    # one grid step along every axis (orthorhombic box)
    step = np.array([sy.L6[0] / sy.K[0], sy.L6[1] / sy.K[1], sy.L6[2] / sy.K[2]])
    E2, v2, dx2, phi2 = s64.evaluate(sy.w, sy.x + step)
    say(f"shifted: relE {abs(E2 - E64) / abs(E64):.1e} vir {rel(v2, v64):.1e} "
        f"F {rel_rms(dx2, dx64):.1e}")
    assert abs(E2 - E64) <= 1e-11 * abs(E64) and rel(v2, v64) <= 1e-11
    assert rel_rms(dx2, dx64) <= 1e-10


def test_gaussian_eighth_against_the_oracle():
    """The same geometry at 1/8 of the atoms, half the box and grid."""
    from ewaldcorrel_amd.workloads import gaussian_charges
    sy = gaussian_charges(1 << 17, (100., 100., 200.), (96, 96, 192), 6, 0.30, seed=13)
    o = OracleSfac(sy.L6, sy.K, sy.order)
    o.set_A(ewald_rad(sy.beta), sy.max_m)
    E0, v0, dx0, phi0 = o.evaluate(sy.w, sy.x)
    for prec in ("fp64", "fp32"):
        E, vir, dx, phi = handle(sy, prec).evaluate(sy.w, sy.x)
        eE, ev, eF, eP = abs(E - E0) / abs(E0), rel(vir, v0), rel_rms(dx, dx0), rel_rms(phi, phi0)
        say(f"gaussian_131k {prec}: relE {eE:.2e} vir {ev:.2e} F {eF:.2e} phi {eP:.2e}")
        if prec == "fp64":
            assert eE <= 1e-10 and ev <= 1e-10 and eF <= 1e-9 and eP <= 1e-9
        else:
            assert eE <= 1e-6 and eF <= 1e-5
