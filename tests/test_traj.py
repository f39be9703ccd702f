"""Trajectory readers (run.py:13-24 .xyz, gmx_ewald.py:65-71 .trr).  The
.trr frames here are encoded byte by byte from the published GROMACS XDR
layout (big-endian ints and reals), independently of the reader."""
import struct

import numpy as np
import pytest

from ewaldcorrel_amd.traj import NEUTRON_B, neutron_weights, read_trr, read_xyz


def test_read_xyz_frames(tmp_path):
    p = tmp_path / "w.xyz"
    p.write_text("3\nwater\nO 0.0 0.0 0.0\nH 0.9572 0 0\nH -0.24 0.9266 0.0\n"
                 "3\nsecond\nO 1 2 3\nH 4 5 6\nH 7 8 9\n")
    fr = list(read_xyz(p))
    assert len(fr) == 2
    names, x = fr[0]
    assert names == ["O", "H", "H"]
    assert np.allclose(x[1], [0.9572, 0, 0])
    assert np.array_equal(fr[1][1], np.arange(1, 10, dtype=float).reshape(3, 3))
    w = neutron_weights(names)
    assert np.array_equal(w, [NEUTRON_B["O"], NEUTRON_B["H"], NEUTRON_B["H"]])


def test_read_xyz_rejects_truncated_frame(tmp_path):
    p = tmp_path / "bad.xyz"
    p.write_text("3\nc\nO 0 0 0\nH 1 0 0\n")
    with pytest.raises(ValueError):
        list(read_xyz(p))


def trr_frame(natoms, step, t, box, x, v=None, double=False):
    r = "d" if double else "f"
    rs = 8 if double else 4
    ver = b"GMX_trn_file"
    out = struct.pack(">i", 1993)
    out += struct.pack(">ii", len(ver) + 1, len(ver)) + ver + b"\0" * ((4 - len(ver) % 4) % 4)
    xs = 3 * natoms * rs
    vs = xs if v is not None else 0
    out += struct.pack(">13i", 0, 0, 9 * rs, 0, 0, 0, 0, xs, vs, 0, natoms, step, 0)
    out += struct.pack(">" + r + r, t, 0.0)
    out += struct.pack(">" + r * 9, *np.asarray(box, float).ravel())
    out += struct.pack(">" + r * (3 * natoms), *np.asarray(x, float).ravel())
    if v is not None:
        out += struct.pack(">" + r * (3 * natoms), *np.asarray(v, float).ravel())
    return out


@pytest.mark.parametrize("double", [False, True], ids=["single", "double"])
def test_read_trr(tmp_path, double):
    rng = np.random.default_rng(0)
    box = np.diag([3.0, 3.1, 3.2])
    box[1, 0] = 0.2
    x1 = rng.uniform(0, 3, (5, 3))
    x2 = rng.uniform(0, 3, (5, 3))
    v2 = rng.normal(size=(5, 3))
    p = tmp_path / "t.trr"
    p.write_bytes(trr_frame(5, 0, 0.0, box, x1, double=double) +
                  trr_frame(5, 10, 0.02, box, x2, v2, double=double))
    fr = list(read_trr(p))
    assert [f["step"] for f in fr] == [0, 10]
    tol = 0 if double else 1e-6
    assert np.allclose(fr[0]["box"], box, atol=tol, rtol=0)
    assert np.allclose(fr[0]["x"], x1, atol=tol * 3, rtol=0)
    assert fr[0]["v"] is None
    assert np.allclose(fr[1]["v"], v2, atol=tol * 10, rtol=0)
    assert abs(fr[1]["t"] - 0.02) <= tol


def test_read_trr_rejects_bad_magic(tmp_path):
    p = tmp_path / "x.trr"
    p.write_bytes(struct.pack(">i", 1234) + b"\0" * 64)
    with pytest.raises(ValueError):
        list(read_trr(p))


@pytest.mark.parametrize("name", ["water16", "water16_d"])
def test_read_trr_golden_files(name):
    """The committed .trr files (tests/golden/make_trr_golden.py, an encoder
    of the published trn XDR layout written apart from the reader) read back
    to the stored values exactly."""
    import json
    import os
    from conftest import ROOT
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "trr_golden.json")))[name]
    frames = list(read_trr(os.path.join(ROOT, "tests", "golden", name + ".trr")))
    assert len(frames) == len(g["frames"])
    for fr, ref in zip(frames, g["frames"]):
        assert fr["step"] == ref["step"] and fr["t"] == ref["t"] and fr["lam"] == 0.0
        assert np.array_equal(fr["box"], np.array(ref["box"]))
        assert np.array_equal(fr["x"], np.array(ref["x"]))
        assert np.array_equal(fr["f"], np.array(ref["f"]))
        if ref["v"] is None:
            assert fr["v"] is None
        else:
            assert np.array_equal(fr["v"], np.array(ref["v"]))
        assert fr["vir"] is None and fr["pres"] is None
