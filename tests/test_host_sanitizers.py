"""The host C++ of the library (sfac_api.cpp, server.cpp, tune.cpp) built
with AddressSanitizer and UndefinedBehaviorSanitizer on the host side only,
linked with the library's own compiled engine, and driven with no GPU by
tests/cpp/sanitize_driver.cpp: argument validation, every entry point on a
NULL handle, malformed and extreme tune records, short export buffers and the
serving loop's constructor failing without a device.  Any sanitizer report
fails the run (UBSan does not recover, ASan aborts)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "ewaldcorrel_amd", "csrc")
ENGINE_OBJ = os.path.join(ROOT, "ewaldcorrel_amd", "_build", "engine.hip.o")
# every sanitizer flag applies to the host compilation only
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    if os.path.exists("/dev/kfd"):
        pytest.skip("drives the library with no GPU present")
    hc = hipcc()
    if hc is None:
        pytest.skip("no hipcc")
    if not os.path.exists(ENGINE_OBJ):
        pytest.skip("the library's engine object is not built (run __graft_entry__.build())")
    objs = []
    srcs = [os.path.join(CSRC, f) for f in ("sfac_api.cpp", "server.cpp", "tune.cpp")]
    srcs.append(os.path.join(ROOT, "tests", "cpp", "sanitize_driver.cpp"))
    for src in srcs:
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        subprocess.run([hc, "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"),
                        "-I", CSRC, *SAN, "-c", src, "-o", obj],
                       check=True, capture_output=True, text=True)
        objs.append(obj)
    exe = str(tmp_path / "sanitize_driver")
    subprocess.run([hc, "--offload-arch=gfx950", *SAN, *objs, ENGINE_OBJ, "-o", exe],
                   check=True, capture_output=True, text=True)
    env = dict(os.environ)
    # the HIP runtime keeps process-lifetime allocations: leak reports are not
    # about this code
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert r.stdout.strip().endswith("ok"), r.stdout[-2000:]
