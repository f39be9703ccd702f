"""Writes water16.trr and water16_d.trr with the source values in
trr_golden.json: GROMACS .trr trajectories encoded field by field from the
published XDR layout of a trn frame (big-endian 4-byte ints, reals in the
file's precision), independently of the package's reader:

    int magic = 1993
    int slen = strlen + 1, then XDR string: int strlen, bytes padded to 4
        ("GMX_trn_file")
    13 ints: ir_size, e_size, box_size, vir_size, pres_size, top_size,
             sym_size, x_size, v_size, f_size, natoms, step, nre
    real t, real lambda
    then, when their sizes are non-zero: box[3][3], vir[3][3], pres[3][3],
    x[natoms][3], v[natoms][3], f[natoms][3]

water16.trr: three frames of 16 rigid waters (48 atoms, nm) in a triclinic
box, single precision, with positions, velocities and forces (the third
frame without velocities); water16_d.trr: two frames of 6 atoms in double
precision.  The JSON holds the values as the file stores them (rounded to
float32 for the single-precision file).

    python tests/golden/make_trr_golden.py
"""
import json
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def xdr_int(v):
    return struct.pack(">i", v)


def xdr_reals(vals, double):
    vals = np.asarray(vals, dtype=np.float64).ravel()
    return struct.pack(">%d%s" % (len(vals), "d" if double else "f"), *vals)


def frame_bytes(step, t, lam, box, x, v, f, double):
    rs = 8 if double else 4
    natoms = len(x)
    ver = b"GMX_trn_file"
    out = xdr_int(1993) + xdr_int(len(ver) + 1) + xdr_int(len(ver))
    out += ver + b"\0" * ((-len(ver)) % 4)
    sizes = [0, 0, 9 * rs, 0, 0, 0, 0,
             3 * natoms * rs if x is not None else 0,
             3 * natoms * rs if v is not None else 0,
             3 * natoms * rs if f is not None else 0,
             natoms, step, 0]
    out += b"".join(xdr_int(s) for s in sizes)
    out += xdr_reals([t, lam], double)
    out += xdr_reals(box, double)
    for blk in (x, v, f):
        if blk is not None:
            out += xdr_reals(blk, double)
    return out


def water(nmol, box, rng):
    centres = rng.uniform(0, 1, (nmol, 3)) @ box
    b, ang = 0.09572, np.deg2rad(104.52)
    local = np.array([[0, 0, 0], [b, 0, 0], [b * np.cos(ang), b * np.sin(ang), 0]])
    out = []
    for c in centres:
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, i, j, k = q
        R = np.array([[1 - 2 * (j * j + k * k), 2 * (i * j - k * w), 2 * (i * k + j * w)],
                      [2 * (i * j + k * w), 1 - 2 * (i * i + k * k), 2 * (j * k - i * w)],
                      [2 * (i * k - j * w), 2 * (j * k + i * w), 1 - 2 * (i * i + j * j)]])
        out.extend(c + local @ R.T)
    return np.array(out)


def as_stored(a, double):
    return None if a is None else np.asarray(a, dtype=np.float64 if double else np.float32
                                             ).astype(np.float64)


def main():
    rng = np.random.default_rng(16)
    golden = {}
    for name, double, nmol, nframes in (("water16", False, 16, 3), ("water16_d", True, 2, 2)):
        box = np.array([[1.9, 0.0, 0.0], [0.3, 2.0, 0.0], [-0.2, 0.25, 2.1]])
        data = b""
        frames = []
        for fr in range(nframes):
            x = water(nmol, box, rng)
            v = rng.normal(0, 0.3, x.shape) if fr != 2 else None
            f = rng.normal(0, 50.0, x.shape)
            step, t, lam = 500 * fr, 1.0 * fr, 0.0
            data += frame_bytes(step, t, lam, box, x, v, f, double)
            frames.append(dict(step=step, t=as_stored([t], double)[0],
                               box=as_stored(box, double).tolist(),
                               x=as_stored(x, double).tolist(),
                               v=None if v is None else as_stored(v, double).tolist(),
                               f=as_stored(f, double).tolist()))
        with open(os.path.join(HERE, name + ".trr"), "wb") as fh:
            fh.write(data)
        golden[name] = dict(double=double, frames=frames)
    with open(os.path.join(HERE, "trr_golden.json"), "w") as fh:
        json.dump(golden, fh)
    print("wrote water16.trr, water16_d.trr, trr_golden.json")


if __name__ == "__main__":
    main()
