"""Writes test_log_ref.json: the numbers stored in the reference's golden
log (frobnitzem/EwaldCorrel test.log.ref:1-20, the output of test.py:test()
for the 4-atom triclinic case), plus the inputs of that case
(test.py:66-86).  Run in a container where the reference is mounted:

    python tests/golden/make_golden.py /root/reference/test.log.ref
"""
import json
import os
import re
import sys


def floats(text):
    return [float(t) for t in re.findall(r"[-+]?\d+\.\d*(?:e[-+]?\d+)?", text)]


def main(path):
    lines = open(path).read().splitlines()
    body = "\n".join(lines)

    def block(start, stop):
        i = next(k for k, l in enumerate(lines) if l.startswith(start))
        j = next(k for k, l in enumerate(lines) if k > i and l.startswith(stop))
        return floats(" ".join(lines[i:j]).split(":", 1)[1])

    out = {
        "source": "test.log.ref (stored output of test.py:test())",
        "case": {
            "L": [9.0, 10.0, 11.0, 2.0, -0.1, 1.0],
            "K": [12, 13, 14],
            "order": 4,
            "eta": 0.5,
            "max_m": 1.0,
            "crds_frac": [[0.1, 0.5, 0.1], [0.9, 0.4, 0.1],
                          [0.1, 0.4, 0.9], [0.9, 0.5, 0.9]],
            "q": [1.0, 1.0, 1.0, 1.0],
        },
        "spl_S00": block("spl S[0,0]", "ref S[0,0]"),
        "ref_S00": block("ref S[0,0]", "Line errors"),
        "line_errors": floats(next(l for l in lines if l.startswith("Line errors"))),
        "E": floats(next(l for l in lines if l.startswith("E:")))[0],
        "E1": floats(next(l for l in lines if l.startswith("E1:")))[0],
        "spl_dx": block("spl dx", "ref dx"),
        "ref_dx": block("ref dx", "spl vir"),
        "spl_vir": block("spl vir", "ref vir"),
        "ref_vir": floats(" ".join(lines[[k for k, l in enumerate(lines)
                                           if l.startswith("ref vir")][0]:]).split(":", 1)[1]),
    }
    assert len(out["spl_S00"]) == 8 and len(out["spl_dx"]) == 12 and len(out["spl_vir"]) == 6
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "test_log_ref.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test.log.ref")
