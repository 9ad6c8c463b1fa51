"""Extracts the FD known-answer table that the reference printed in its own
Si216 example run (examples/Si216/a.log:250-272: lap(m) and coef_lap(:,m)
for Md=6) into fd_known_answer_si216.json.

Run in a container where the reference tree is mounted:
    python tests/golden/make_fd_known_answer.py /root/reference
"""
import json
import os
import re
import sys


def main(ref_root):
    path = os.path.join(ref_root, "examples", "Si216", "a.log")
    lines = open(path).read().splitlines()
    start = next(i for i, l in enumerate(lines) if re.search(r"^\s+lap\s+coef_lap\s*$", l))
    rows = {}
    for l in lines[start + 1:start + 8]:
        f = l.split()
        rows[int(f[0])] = [float(x) for x in f[1:]]
    out = {
        "source": "examples/Si216/a.log:259-266 (reference run, Md=6)",
        "Md": 6,
        "lap": [rows[m][0] for m in range(7)],
        "coef_lap0": rows[0][1],
        "coef_lap_axis1": [rows[m][1] for m in range(1, 7)],
        "printed_decimals": 15,
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fd_known_answer_si216.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
