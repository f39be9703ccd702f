"""Runner of the C++ CPU restatement (oracle/sfac_cpu.cpp) -- test
infrastructure and the CPU baseline of bench.py, never the product path.

run_port(...) writes the inputs in the program's binary layout, runs the
-O3 (or -O0) build with the given OpenMP thread count and returns its JSON
report with (E, vir, dx, phi) of the last evaluation."""
from __future__ import annotations

import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")


def binary(opt="O3"):
    return os.path.join(BUILD, "sfac_cpu" if opt == "O3" else "sfac_cpu_O0")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def run_port(L6, K, order, eta, max_m, w, x, reps=1, threads=1, opt="O3", timeout=600):
    w = np.ascontiguousarray(w, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    n = len(w)
    exe = binary(opt)
    if not os.path.exists(exe):
        raise FileNotFoundError(f"{exe} is not built (make -C oracle)")
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([K[0], K[1], K[2], order, n], dtype=np.int32).tofile(f)
            np.concatenate([np.asarray(L6, float), [eta, max_m]]).astype(np.float64).tofile(f)
            w.tofile(f)
            x.tofile(f)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([exe, fin, fout, str(reps)], capture_output=True, text=True,
                           env=env, timeout=timeout, check=True)
        rep = json.loads(r.stdout.strip().splitlines()[-1])
        out = np.fromfile(fout, dtype=np.float64)
    E = out[0]
    vir = out[1:7]
    dx = out[7:7 + 3 * n].reshape(n, 3)
    phi = out[7 + 3 * n:7 + 4 * n]
    return rep, (E, vir, dx, phi)
