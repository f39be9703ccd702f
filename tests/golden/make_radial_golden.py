"""Writes radial_plot.json: input/output vectors of the reference's radial
binning, plot.py:9-32 (frobnitzem/EwaldCorrel), which no closed form or
invariant pins.  The reference's plot.main() is run here, in the container
where the reference is mounted, on a deterministic (64, 40, 33) complex array
saved as .npy (the layout run.py:47 saves: K0 x K1 x ldim, here of a
64 x 40 x 64 grid); its printed lines "k mean count" are stored with the
array's generator (numpy default_rng seed and shape) and checksums of it.

    python tests/golden/make_radial_golden.py /root/reference
"""
import contextlib
import importlib.util
import io
import json
import os
import sys
import tempfile

import numpy as np


def make_input(shape, seed):
    """The binned array: real part uniform in [0.5, 2), imaginary normal."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.5, 2.0, shape) + 1j * rng.normal(size=shape)


def main(ref_dir):
    spec = importlib.util.spec_from_file_location("ref_plot", os.path.join(ref_dir, "plot.py"))
    plot = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(plot)
    shape = (64, 40, 33)
    f = make_input(shape, 12)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "q.npy")
        np.save(p, f)
        buf = io.StringIO()
        with contextlib.redirect_stdout(buf):
            plot.main(["plot.py", p])
    rows = [[float(t) for t in ln.split()] for ln in buf.getvalue().splitlines() if ln.strip()]
    out = {
        "source": "plot.py:9-32 (main) run on the array below; L is plot.py's own 11.680933*5",
        "L": 11.680933 * 5,
        "shape": list(shape),
        "seed": 12,
        "f_real_sum": float(f.real.sum()),
        "f_real_first": f.real.ravel()[:8].tolist(),
        "rows": rows,
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "radial_plot.json")
    with open(dst, "w") as fh:
        json.dump(out, fh)
    print(f"wrote {dst}: {len(rows)} rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
