"""Trajectory ingestion and structure-factor accumulation over frames.

The reference's S(k) drivers read coordinates (``run.py:13-24``: an .xyz
file, neutron scattering lengths as weights, ``run.py:7-11``;
``gmx_ewald.py:18,65-71``: a GROMACS .trr trajectory through the external
``xdrfile`` module) and sum ``|get_S()|^2`` over frames on the host
(``run.py:36-47``).  Here the frames are parsed in numpy and the sum stays
on the device (``Sfac.accum_S2``); only the final array is copied back.

.trr frames are decoded from the GROMACS XDR layout (big-endian): magic
1993, a version string, thirteen header ints (ir, e, box, vir, pres, top,
sym, x, v, f sizes, natoms, step, nre), t and lambda, then the blocks whose
sizes are non-zero (box 3x3, virial, pressure, x, v, f).  Single or double
precision is told by box_size / 9.
"""
from __future__ import annotations

import struct

import numpy as np

# neutron scattering lengths (fm), run.py:7-11
NEUTRON_B = {"H": -3.7390, "C": 6.6460, "O": 5.803}


def read_xyz(path):
    """Frames of an .xyz file: (names, coordinates[n, 3]) per frame.

    Each frame is a count line, a comment line and one ``name x y z`` line
    per atom; a file with a single frame is what ``run.py:13-24`` reads."""
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines):
        head = lines[i].split()
        if not head:
            i += 1
            continue
        n = int(head[0])
        names, crd = [], []
        for line in lines[i + 2:i + 2 + n]:
            tok = line.split()
            if len(tok) < 4:
                raise ValueError(f"{path}: short atom line {line!r}")
            names.append(tok[0])
            crd.append([float(t) for t in tok[1:4]])
        if len(names) != n:
            raise ValueError(f"{path}: frame at line {i + 1} has {len(names)} of {n} atoms")
        yield names, np.array(crd, dtype=np.float64)
        i += 2 + n


def neutron_weights(names):
    """Per-atom weights from ``NEUTRON_B`` (run.py:28)."""
    return np.array([NEUTRON_B[nm] for nm in names], dtype=np.float64)


def _xdr_string(buf, off):
    n = struct.unpack_from(">i", buf, off)[0]   # length incl. terminator
    m = struct.unpack_from(">i", buf, off + 4)[0]
    pad = (m + 3) // 4 * 4
    return buf[off + 8:off + 8 + m].decode(errors="replace"), off + 8 + pad, n


def read_trr(path):
    """Frames of a GROMACS .trr file: dict(step, t, lam, box[3, 3], vir,
    pres, x, v, f) (None for blocks the frame does not hold), in the file's
    units (nm).  Box rows are the box vectors (``SofL(box)`` gives L6)."""
    with open(path, "rb") as fh:
        buf = fh.read()
    off = 0
    while off < len(buf):
        magic = struct.unpack_from(">i", buf, off)[0]
        if magic != 1993:
            raise ValueError(f"{path}: bad magic {magic} at byte {off}")
        _, off, _ = _xdr_string(buf, off + 4)
        (ir, e, box, vir, pres, top, sym, xs, vs, fs, natoms, step, nre) = \
            struct.unpack_from(">13i", buf, off)
        off += 13 * 4
        if ir or e or top or sym:
            raise ValueError(f"{path}: frame with ir/e/top/sym blocks is not supported")
        if box:
            real = box // 9
        elif xs:
            real = xs // (3 * natoms)
        else:
            real = 4
        if real not in (4, 8):
            raise ValueError(f"{path}: cannot tell the precision of frame at byte {off}")
        dt = np.dtype(">f8" if real == 8 else ">f4")
        t, lam = struct.unpack_from(">dd" if real == 8 else ">ff", buf, off)
        off += 2 * real
        out = dict(step=step, t=t, lam=lam, box=None, vir=None, pres=None, x=None, v=None,
                   f=None)
        for key, size, shape in (("box", box, (3, 3)), ("vir", vir, (3, 3)),
                                 ("pres", pres, (3, 3)), ("x", xs, (natoms, 3)),
                                 ("v", vs, (natoms, 3)), ("f", fs, (natoms, 3))):
            if size:
                arr = np.frombuffer(buf, dtype=dt, count=size // real, offset=off)
                out[key] = arr.astype(np.float64).reshape(shape)
                off += size
        yield out


def accumulate_sk(sfac, frames, weights=None):
    """Sum |S_hat|^2 of every frame on the device (run.py:36-47).

    ``frames`` yields coordinates[n, 3] (or (names, coordinates) pairs, in
    which case the weights default to the neutron scattering lengths).
    Returns (sum over frames as a (K0, K1, ldim) array, number of frames)."""
    import torch
    K0, K1, K2 = (int(k) for k in sfac.K)
    H = K0 * K1 * (K2 // 2 + 1)
    acc = torch.zeros(H, dtype=torch.float64, device="cuda")
    nfr = 0
    for fr in frames:
        if isinstance(fr, tuple):
            names, x = fr
            w = weights if weights is not None else neutron_weights(names)
        else:
            x, w = fr, weights
        sfac.sfac(np.asarray(w, dtype=np.float64), np.asarray(x, dtype=np.float64))
        sfac.accum_S2(acc.data_ptr())
        nfr += 1
    sfac.sync()
    return acc.cpu().numpy().reshape(K0, K1, K2 // 2 + 1), nfr
