"""Parity of the HIP engine (through libsfac.so and the ewald.py drop-in)
with the FP64 oracle and the reference's own tests.

Tolerances: FP64 path -- energy and virial <= 1e-10 relative, forces and
potentials <= 1e-9 relative RMS.  FP32 path -- energy <= 1e-6 relative,
forces <= 1e-5 relative RMS (virial and potentials stated beside them).
"""
import sys

import numpy as np
import pytest

# imported at collection: the first import on a fresh machine can take
# minutes and must not count against a test's time limit
torch = pytest.importorskip("torch")

from conftest import case_T, rel, rel_rms  # noqa: E402
from oracle.sfac_oracle import LofS, OracleSfac, SofL, ewald_rad

pytestmark = pytest.mark.gpu


def _ewald():
    from ewaldcorrel_amd import ewald
    return ewald


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_gpu(sysdef, precision="fp64", fixed=False, eta=None):
    ew = _ewald()
    s = ew.Sfac(sysdef["L6"], np.array(sysdef["K"]), sysdef["order"], precision=precision,
                fixed=fixed)
    s.set_A(ew.ewald_kernel(eta or sysdef["beta"]), sysdef.get("max_m", 1.0))
    E, vir, dx, phi = s.evaluate(sysdef["w"], sysdef["x"])
    return s, E, vir, dx, phi


def run_oracle(sysdef, eta=None):
    o = OracleSfac(sysdef["L6"], sysdef["K"], sysdef["order"])
    o.set_A(ewald_rad(eta or sysdef["beta"]), sysdef.get("max_m", 1.0))
    o.sfac(sysdef["w"], sysdef["x"])
    S = o.get_S()
    E, vir = o.de1()
    dx = o.de2(sysdef["w"], sysdef["x"])
    phi = o.potl(sysdef["x"])
    return S, E, vir, dx, phi


def make_sys(L6, K, order, N, seed, beta=0.5, spread_out=1.0, max_m=1.0):
    rng = np.random.default_rng(seed)
    L6 = np.asarray(L6, dtype=float)
    crds = rng.uniform(-0.5 * spread_out, 1 + 0.5 * spread_out, (N, 3))
    x = crds @ LofS(L6)
    w = rng.normal(size=N)
    w -= w.mean()
    return dict(L6=L6, K=tuple(K), order=order, beta=beta, w=w, x=x, max_m=max_m)


def check_fp64(name, got, ref):
    S_ref, E_ref, vir_ref, dx_ref, phi_ref = ref
    _, E, vir, dx, phi = got
    e_E = abs(E - E_ref) / abs(E_ref)
    e_v = rel(vir, vir_ref)
    e_f = rel_rms(dx, dx_ref)
    e_p = rel_rms(phi, phi_ref)
    say(f"{name}: E {E:.15g} relE {e_E:.2e} vir {e_v:.2e} F {e_f:.2e} phi {e_p:.2e}")
    assert e_E <= 1e-10
    assert e_v <= 1e-10
    assert e_f <= 1e-9
    assert e_p <= 1e-9


# ---- the reference's own tests (test.py), with assertions ------------------------------
def test_reference_test_py(golden):
    """test.py:test() -- line test, en == de1, de2 vs central FD, virial vs FD
    of E(L) at fixed fractional coordinates; plus the stored golden values."""
    ew = _ewald()
    L, K, order, crds, atoms, q = case_T(golden)
    N = len(q)
    s = ew.Sfac(L, K.astype(int), order)
    s.set_A(ew.ewald_f, 1.0)
    s.sfac(q, atoms)
    M = s.get_S()
    from math import pi
    mx = np.arange(M.shape[0]); mx[M.shape[0] // 2 + 1:] -= M.shape[0]
    my = np.arange(M.shape[1]); my[M.shape[1] // 2 + 1:] -= M.shape[1]
    mz = np.arange(M.shape[2])
    sx = sum(qq * np.exp(-2j * pi * x[0] * mx) for qq, x in zip(q, crds))
    sy = sum(qq * np.exp(-2j * pi * x[1] * my) for qq, x in zip(q, crds))
    sz = sum(qq * np.exp(-2j * pi * x[2] * mz) for qq, x in zip(q, crds))
    errs = [np.abs(M[:, 0, 0] - sx)[:M.shape[0] // 4].max(),
            np.abs(M[0, :, 0] - sy)[:M.shape[1] // 4].max(),
            np.abs(M[0, 0, :] - sz)[:M.shape[2] // 2].max()]
    say("line errors", errs)
    assert errs[0] < 0.01 and errs[1] < 0.01 and errs[2] < 0.1
    assert np.allclose(errs, golden["line_errors"], rtol=0, atol=1e-13)
    assert np.allclose(M[0, 0].real, golden["spl_S00"], atol=6e-9)

    def num_diff(f, x, h=1e-7):
        ih = 1. / (2.0 * h)
        h = 0.5 / ih
        shp = x.shape
        x = np.reshape(x, -1)
        df = np.zeros(len(x))
        for i in range(len(x)):
            x0 = x[i]
            x[i] = x0 + h
            df[i] = f(np.reshape(x, shp))
            x[i] = x0 - h
            df[i] -= f(np.reshape(x, shp))
            x[i] = x0
        return np.reshape(df * ih, shp)

    def Ex(x):
        s.sfac(q, x.reshape((N, 3)))
        return s.en()

    dx0 = num_diff(Ex, atoms.copy())
    s.sfac(q, atoms)
    en0 = s.en()
    vir = np.zeros(6)
    en1 = s.de1(vir)
    say("E", en0, en1)
    assert abs(en1 - en0) < 1e-14
    assert abs(en0 - golden["E"]) < 1e-13 * golden["E"]
    dx = np.zeros((N, 3))
    s.de2(N, q, atoms, dx)
    assert np.abs(dx - dx0).max() < 1e-7
    assert np.allclose(dx.ravel(), golden["spl_dx"], atol=6e-9)

    V = L[0] * L[1] * L[2]

    def E(Lx):
        s.set_L(Lx)
        return s.en()

    s.sfac(q, atoms)
    Pi = LofS(num_diff(E, L.copy())).transpose()
    Pi = SofL(np.dot(Pi, LofS(L)).transpose()) / (-V)
    say("vir", vir, Pi)
    assert np.abs(vir - Pi).max() < 1e-8
    assert np.allclose(vir, golden["spl_vir"], rtol=6e-9)


def test_reference_test_potl(golden):
    """test.py:test_potl() -- prints en0, en1, 1/2 q.pot, which must agree."""
    ew = _ewald()
    q = (np.arange(4) % 2) * 0.5 + 0.5
    L, K, order, crds, atoms, _ = case_T(golden)
    s = ew.Sfac(L, K.astype(int), 4)
    s.set_A(ew.ewald_f, 1.0)
    s.sfac(q, atoms)
    en0 = s.en()
    vir = np.zeros(6)
    en1 = s.de1(vir)
    pot = np.zeros(4)
    s.potl(4, atoms, pot)
    say(pot, en0, en1, 0.5 * np.dot(q, pot))
    assert abs(en0 - en1) < 1e-14
    assert abs(0.5 * np.dot(q, pot) - en0) < 1e-12 * abs(en0)
    o = OracleSfac(L, K, 4)
    o.set_A(ewald_rad(0.5), 1.0)
    o.sfac(q, atoms)
    o.de1()
    assert rel(pot, o.potl(atoms)) < 1e-12


def test_python_callback_table_equals_native():
    """set_A through a Python rad_fn callback (the reference path) and the
    native closed-form build give the same energies."""
    ew = _ewald()
    sysd = make_sys([9., 10., 11., 2., -0.1, 1.0], (12, 13, 14), 4, 20, 11)
    s = ew.Sfac(sysd["L6"], np.array(sysd["K"]), 4)
    s.set_A(ew.ewald_f, 1.0)
    E1, v1, d1, p1 = s.evaluate(sysd["w"], sysd["x"])

    def my_f(r2, ptr, info):  # same function, not recognised -> callback
        return ew.ewald_f(r2, ptr, info)

    s.set_A(my_f, 1.0)
    E2, v2, d2, p2 = s.evaluate(sysd["w"], sysd["x"])
    assert abs(E1 - E2) <= 1e-14 * abs(E1)
    assert rel(v1, v2) < 1e-13


# ---- oracle parity, FP64 ------------------------------------------------------------------
CASES = [
    # name, L6, K, order, N, seed, beta
    ("triclinic_prime13", [9., 10., 11., 2., -0.1, 1.0], (12, 13, 14), 4, 50, 1, 0.5),
    ("odd_grid", [11., 12., 13., 0.5, 0.3, -0.8], (15, 21, 27), 4, 80, 2, 0.45),
    ("order6_triclinic", [20., 18., 22., 3., -2., 1.5], (24, 25, 30), 6, 300, 3, 0.4),
    ("order8", [16., 16., 16., 0., 0., 0.], (20, 20, 20), 8, 200, 4, 0.45),
    ("order2", [10., 10., 12., 0., 0., 0.], (16, 16, 18), 2, 60, 5, 0.5),
    ("order10_generic", [14., 15., 16., 1., 1., 1.], (18, 20, 22), 10, 100, 6, 0.45),
    ("primes_7_11_13", [10., 11., 12., -1., 0.5, 0.2], (7, 11, 13), 4, 40, 7, 0.6),
    ("radix3_5", [30., 31., 33., 0., 0., 0.], (45, 48, 50), 4, 2000, 8, 0.35),
    ("tiny_K", [5., 6., 7., 0., 0., 0.], (2, 3, 1), 4, 10, 9, 0.7),
    ("long_z", [20., 20., 60., 0., 0., 0.], (16, 16, 384), 4, 500, 10, 0.4),
]


@pytest.mark.parametrize("fixed", [False, True], ids=["atomics", "fixedpoint"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_parity_fp64(case, fixed):
    name, L6, K, order, N, seed, beta = case
    sysd = make_sys(L6, K, order, N, seed, beta=beta, spread_out=1.0)
    say("case", name, "fixed", fixed)
    ref = run_oracle(sysd)
    got = run_gpu(sysd, "fp64", fixed)
    check_fp64(name, got, ref)


@pytest.mark.parametrize("case", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_get_S_after_sfac(case):
    name, L6, K, order, N, seed, beta = case
    sysd = make_sys(L6, K, order, N, seed, beta=beta)
    ew = _ewald()
    s = ew.Sfac(sysd["L6"], np.array(K), order)
    s.sfac(sysd["w"], sysd["x"])
    S = s.get_S()
    o = OracleSfac(sysd["L6"], K, order)
    o.sfac(sysd["w"], sysd["x"])
    Sref = o.get_S()
    err = np.abs(S - Sref).max() / np.abs(Sref).max()
    say(name, "get_S rel err", err)
    assert err < 1e-12


def test_point_charges_1k_fp64():
    from ewaldcorrel_amd.workloads import point_charges_1k
    sy = point_charges_1k()
    sysd = dict(L6=sy.L6, K=sy.K, order=sy.order, beta=sy.beta, w=sy.w, x=sy.x)
    check_fp64("point_charges_1k", run_gpu(sysd), run_oracle(sysd))


def test_water_98k_fp64_and_fp32():
    """The full water configuration (98,304 atoms, 128^3, order 4)."""
    from ewaldcorrel_amd.workloads import water_98k
    sy = water_98k()
    sysd = dict(L6=sy.L6, K=sy.K, order=sy.order, beta=sy.beta, w=sy.w, x=sy.x)
    ref = run_oracle(sysd)
    check_fp64("water_98k", run_gpu(sysd), ref)
    _, E, vir, dx, phi = run_gpu(sysd, "fp32")
    eE = abs(E - ref[1]) / abs(ref[1])
    eF = rel_rms(dx, ref[3])
    say(f"water_98k fp32: relE {eE:.2e} vir {rel(vir, ref[2]):.2e} F {eF:.2e} "
        f"phi {rel_rms(phi, ref[4]):.2e}")
    assert eE <= 1e-6
    assert eF <= 1e-5


# ---- FP32 --------------------------------------------------------------------------------
@pytest.mark.parametrize("case", CASES[:8], ids=[c[0] for c in CASES[:8]])
def test_fp32_error_bounds(case):
    name, L6, K, order, N, seed, beta = case
    sysd = make_sys(L6, K, order, N, seed, beta=beta)
    ref = run_oracle(sysd)
    _, E, vir, dx, phi = run_gpu(sysd, "fp32")
    eE = abs(E - ref[1]) / abs(ref[1])
    eF = rel_rms(dx, ref[3])
    say(f"{name} fp32: relE {eE:.2e} vir {rel(vir, ref[2]):.2e} F {eF:.2e} "
        f"phi {rel_rms(phi, ref[4]):.2e}")
    assert eE <= 1e-6
    assert eF <= 1e-5
    assert rel_rms(phi, ref[4]) <= 1e-5


@pytest.mark.parametrize("order", [4, 6])
@pytest.mark.parametrize("natoms", [40, 900])
def test_fp32_tiled_gather_sparse_tiles(order, natoms):
    """The FP32 evaluation gathers through LDS-staged tiles of 8 x 8 columns.
    With 40 atoms on a 20 x 24 x 30 grid most tiles hold a handful of atoms
    (fewer than the tile's column rows), with 900 the last wavefront of a tile is
    partly idle: forces and potentials against the oracle, and two evaluations
    on two handles bit-identical."""
    sysd = make_sys([16., 17., 18., 0.5, 0.2, -0.3], (20, 24, 30), order, natoms, 2,
                    beta=0.4)
    ref = run_oracle(sysd)
    s1, E, vir, dx, phi = run_gpu(sysd, "fp32")
    s2, E2, vir2, dx2, phi2 = run_gpu(sysd, "fp32")
    eF, eP = rel_rms(dx, ref[3]), rel_rms(phi, ref[4])
    say(f"order {order} n {natoms}: relE {abs(E - ref[1]) / abs(ref[1]):.2e} F {eF:.2e} "
        f"phi {eP:.2e}")
    assert abs(E - ref[1]) <= 1e-6 * abs(ref[1])
    assert eF <= 1e-5 and eP <= 1e-5
    assert np.array_equal(dx, dx2) and np.array_equal(phi, phi2)


# ---- determinism of the fixed-point grid --------------------------------------------------
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_fixed_point_grid_is_order_independent(precision):
    ew = _ewald()
    sysd = make_sys([20., 18., 22., 3., -2., 1.5], (24, 25, 30), 4, 3000, 12)
    perm = np.random.default_rng(0).permutation(len(sysd["w"]))
    s = ew.Sfac(sysd["L6"], np.array(sysd["K"]), 4, precision=precision, fixed=True)
    s.sfac(sysd["w"], sysd["x"])
    S1 = s.get_S().copy()
    s.sfac(sysd["w"][perm], sysd["x"][perm])
    S2 = s.get_S().copy()
    assert np.array_equal(S1.view(np.float64), S2.view(np.float64))


# ---- edge cases ----------------------------------------------------------------------------
def test_empty_and_single_atom():
    ew = _ewald()
    L6 = np.array([10., 10., 10., 0, 0, 0])
    s = ew.Sfac(L6, np.array([16, 16, 16]), 4)
    s.set_A(ew.ewald_kernel(0.4), 1.0)
    E, vir, dx, phi = s.evaluate(np.zeros(0), np.zeros((0, 3)))
    assert E == 0.0 and np.all(vir == 0.0) and dx.shape == (0, 3)
    sysd = dict(L6=L6, K=(16, 16, 16), order=4, beta=0.4, w=np.array([1.3]),
                x=np.array([[1.0, 2.0, 3.0]]))
    check_fp64("single", run_gpu(sysd), run_oracle(sysd))


def test_boundary_wrap_and_far_images():
    """Coordinates that wrap to exactly s = 1.0, sit on cell faces, or lie
    many boxes away are reduced to the periodic image."""
    L6 = np.array([12., 12., 12., 0, 0, 0])
    x = np.array([[-1e-20, 0.0, 5.0], [12.0, 6.0, 6.0], [3.0, -1e-18, 11.999999999999998],
                  [1e3 + 0.3, -2e3 + 0.7, 5e2], [6.0, 6.0, 6.0], [0.0, 0.0, 0.0]])
    w = np.array([1.0, -0.5, 0.7, -1.2, 0.3, -0.3])
    sysd = dict(L6=L6, K=(16, 16, 16), order=4, beta=0.4, w=w, x=x)
    check_fp64("boundary", run_gpu(sysd), run_oracle(sysd))
    # and the images of each atom give the same result
    shifted = dict(sysd, x=x + np.array([12.0, -24.0, 36.0]))
    ref = run_oracle(sysd)
    check_fp64("boundary_shifted", run_gpu(shifted), ref)


def test_state_machine_errors(capfd):
    ew = _ewald()
    s = ew.Sfac(np.array([10., 10., 10., 0, 0, 0]), np.array([8, 8, 8]), 4)
    assert s.en() == 0.0  # before sfac: prints, returns 0.0 (sfac.cpp:310-317)
    assert "must call sfac" in capfd.readouterr().out
    s.sfac(np.ones(2), np.ones((2, 3)))
    assert s.en() == 0.0  # no kernel set
    assert "No convolution kernel" in capfd.readouterr().out
    dx = np.zeros((2, 3))
    s.de2(2, np.ones(2), np.ones((2, 3)), dx)  # before de1
    assert "must call de1" in capfd.readouterr().out
    s.set_A(ew.ewald_f, 1.0)
    vir = np.zeros(6)
    s.de1(vir)
    assert s.de1(vir) == 0.0  # de1 twice
    assert "must call sfac" in capfd.readouterr().out


# ---- device-pointer API ------------------------------------------------------------------
def test_eval_dev_matches_host_api():
    ew = _ewald()
    sysd = make_sys([20., 18., 22., 3., -2., 1.5], (24, 25, 30), 6, 500, 13, beta=0.4)
    s = ew.Sfac(sysd["L6"], np.array(sysd["K"]), 6)
    s.set_A(ew.ewald_kernel(0.4), 1.0)
    E, vir, dx, phi = s.evaluate(sysd["w"], sysd["x"])
    dev = torch.device("cuda:0")
    w = torch.tensor(sysd["w"], device=dev)
    x = torch.tensor(sysd["x"], device=dev)
    out = torch.zeros(8, dtype=torch.float64, device=dev)
    dxd = torch.zeros((len(w), 3), dtype=torch.float64, device=dev)
    phd = torch.zeros(len(w), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    s.eval_dev(len(w), w.data_ptr(), x.data_ptr(), out.data_ptr(), dxd.data_ptr(),
               phd.data_ptr())
    s.sync()
    o = out.cpu().numpy()
    # same kernels; FP64 atomics may add in another order
    assert abs(o[0] - E) <= 1e-13 * abs(E) and rel(o[1:7], vir) < 1e-12
    assert rel_rms(dxd.cpu().numpy(), dx) < 1e-12
    assert rel_rms(phd.cpu().numpy(), phi) < 1e-12
