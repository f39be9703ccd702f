"""The C-ABI library loads and exports every symbol the public headers
declare; argument validation that needs no device.  CPU only (no compute
call reaches the GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "ewaldcorrel_amd", "libsfac.so")
HEADERS = [os.path.join(ROOT, "include", h) for h in ("sfac.hpp", "sfac_ext.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        # only the extern "C" blocks
        for block in re.findall(r'extern "C" \{(.*?)\n\}', text, re.S):
            block = re.sub(r"//[^\n]*|/\*.*?\*/", "", block, flags=re.S)
            block = re.sub(r"^\s*#[^\n]*", "", block, flags=re.M)
            for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", block):
                name = m.group(1)
                if name not in ("typedef", "double", "int", "void", "const"):
                    names.add(name)
    names.discard("rad_fn")
    return names


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libsfac.so is not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB)


def test_headers_declare_reference_abi():
    ref = {"sfac_ctor", "sfac_dtor", "sfac", "get_S", "set_L", "set_A", "en", "de1", "de2",
           "potl"}
    assert ref <= declared_functions()


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_declares_everything():
    from ewaldcorrel_amd import ewald
    sfac_p, rad_fn, lib = ewald.describe_sfac()
    for n in declared_functions():
        assert getattr(lib, n).restype is not None or n in (
            "sfac_dtor", "sfac", "set_L", "de2", "potl"), n


def test_ctor_rejects_bad_arguments_without_device():
    """Odd order and K < 1 are rejected before any device call (the
    reference prints and throws, sfac.cpp:41-51); here the ctor returns NULL
    and the binding raises."""
    from ewaldcorrel_amd import ewald
    L = np.array([10., 10., 10., 0., 0., 0.])
    with pytest.raises(RuntimeError, match="invalid order"):
        ewald.Sfac(L, np.array([8, 8, 8]), 3)
    with pytest.raises(RuntimeError, match="invalid K"):
        ewald.Sfac(L, np.array([8, 0, 8]), 4)
    with pytest.raises(ValueError):
        ewald.Sfac(L, np.array([8, 8, 8]), 4, precision="fp16")


def test_null_handle_is_an_error_not_a_crash(lib):
    lib.en.restype = ctypes.c_double
    lib.en.argtypes = [ctypes.c_void_p]
    assert lib.en(None) == 0.0
    lib.sfac_last_error.restype = ctypes.c_char_p
    assert b"NULL" in lib.sfac_last_error()


def test_oracle_not_linked_into_product():
    """The product library and binding never reference the oracle."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "ewaldcorrel_amd")):
        for fn in files:
            if fn.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in text.lower(), fn


def test_rad_fn_callback_type_is_accepted(lib):
    """set_A takes the binding's rad_fn type for every handle (one ctypes
    type shared by all Sfac objects); a NULL handle reports an error."""
    from ctypes import c_void_p, cast
    from ewaldcorrel_amd import ewald
    for _ in range(2):
        _, rad_fn, l = ewald.describe_sfac()

        def f(r2, ptr, info):
            ptr.contents.value = 0.0
            return 0.0

        assert l.set_A(None, rad_fn(f), 1.0, cast(0, c_void_p)) != 0


def test_server_create_fails_cleanly_without_a_device():
    """With no GPU the serving loop's constructor reports an error instead of
    aborting (the HIP runtime finds no device)."""
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from ewaldcorrel_amd import ewald
    with pytest.raises(RuntimeError, match="sfac_server_create failed"):
        ewald.SfacServer(slots=2)


def test_tune_records_round_trip_without_device():
    """Remembered launch shapes export and import as text (no device call);
    malformed lines are skipped."""
    from ewaldcorrel_amd import ewald
    ewald.tune_clear()
    text = ("128x128x128/4/3 8,8,2 2,8,8 8,4,4 8x8 8x8 1\n"
            "45x45x45/6/0 5,3,3 - 3,5,3 4x4x0 0x0x0 0\n"
            "192x192x384/6/3 - - - 8x8x128 8x8x96 0\n"
            "garbage line\n"
            "32x32x32/4/1 8,4,7 - - 8x8 0x0 0\n")
    assert ewald.tune_import(text) == 3
    out = ewald.tune_export()
    assert "128x128x128/4/3 8,8,2 2,8,8 8,4,4 8x8x0 8x8x0 1" in out
    assert "45x45x45/6/0 5,3,3 - 3,5,3 4x4x0 0x0x0 0" in out
    assert "192x192x384/6/3 - - - 8x8x128 8x8x96 0" in out
    assert "32x32x32" not in out
    ewald.tune_clear()
    assert ewald.tune_export() == ""


def _m4(x):
    """Closed-form cubic cardinal B-spline M_4 on [0, 4]."""
    if x < 0 or x >= 4:
        return 0.0
    if x < 1:
        return x ** 3 / 6
    if x < 2:
        return (-3 * x ** 3 + 12 * x ** 2 - 12 * x + 4) / 6
    if x < 3:
        return (3 * x ** 3 - 24 * x ** 2 + 60 * x - 44) / 6
    return (4 - x) ** 3 / 6


def test_cxx_caller_links_and_runs(tmp_path):
    """A C++ program shaped like the reference's test.cpp compiles against
    include/sfac.hpp with g++, links libsfac.so (including the reference's
    SFac(double*, int*, int) constructor symbol) and runs without a GPU."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = tmp_path / "caller"
    src = os.path.join(ROOT, "tests", "cpp", "test_caller.cpp")
    libdir = os.path.dirname(LIB)
    subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    src, "-o", str(exe), "-L", libdir, "-lsfac", f"-Wl,-rpath,{libdir}"],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True,
                         timeout=120).stdout.splitlines()
    table = [ln.split() for ln in out if ln[:1].isdigit()]
    assert len(table) == 40
    for xs, c in table:
        assert abs(float(c) - _m4(float(xs))) <= 6e-7, (xs, c)
    ds = [ln.split() for ln in out if ln.startswith("d ")]
    for _, j, b, db in ds:
        x = 0.3 + int(j)
        h = 1e-6
        assert abs(float(b) - _m4(x)) <= 1e-14
        assert abs(float(db) - (_m4(x + h) - _m4(x - h)) / (2 * h)) <= 1e-8
    joined = "\n".join(out)
    assert "ctor_odd_order NULL invalid order" in joined
    assert "ctor_bad_K NULL invalid K" in joined
    assert "cxx_ctor throws invalid order" in joined
