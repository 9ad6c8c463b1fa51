"""Copies the density/potential file that the reference wrote in its own Si2
SCF example (examples/Si2_band_and_dos/band_calc_step_by_step/step1_scf/
vrho.dat1, an unversioned read_data file: sizes, LL2, then rho, Vloc, Vh,
Vxc of one spin) to si2_vrho.dat1, with the run's cell and grid from the
example's input (fort.1: AX 10.261, FCC vectors A1-A3, NGRID 12 12 12,
XCtype LDAPZ81) and the file's SHA-256 in si2_vrho.json.

Run in a container where the reference tree is mounted:
    python tests/golden/make_vrho_fixture.py /root/reference
"""
import hashlib
import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    d = os.path.join(ref_root, "examples", "Si2_band_and_dos", "band_calc_step_by_step",
                     "step1_scf")
    src = os.path.join(d, "vrho.dat1")
    dst = os.path.join(HERE, "si2_vrho.dat1")
    shutil.copyfile(src, dst)
    inp = open(os.path.join(d, "fort.1")).read()

    def field(key, n):
        m = re.search(rf"^{key}\s+(.*?)/", inp, re.M)
        return [float(x.replace("d", "e")) for x in m.group(1).split()[:n]]

    meta = {
        "source": "examples/Si2_band_and_dos/band_calc_step_by_step/step1_scf/vrho.dat1",
        "sha256": hashlib.sha256(open(dst, "rb").read()).hexdigest(),
        "ax": field("AX", 1)[0],
        "a1": field("A1", 3), "a2": field("A2", 3), "a3": field("A3", 3),
        "ngrid": [int(x) for x in field("NGRID", 3)],
        "xc": re.search(r"^XCtype\s+'(\w+)'", inp, re.M).group(1),
        "valence_electrons": 8,
    }
    json.dump(meta, open(os.path.join(HERE, "si2_vrho.json"), "w"), indent=1)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
