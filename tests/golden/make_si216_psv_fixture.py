"""Collects the data and the printed results of the reference's shipped Si216
example run (real pseudopotential, 48^3 grid, 216 atoms, 648 bands) into
fixtures under tests/golden/si216_psv/:

* Si_psv.dat.gz -- the example's PSV pseudopotential file (examples/Si216/
  Si_psv.dat), input data, gzip-compressed;
* atoms.json -- the 216 fractional positions of examples/Si216/fort.970 as
  written there (10 decimals), the cell (fort.1: AX 30.783, cubic) and grid;
* results.json -- what the run printed: read_PSV's summary (a.log:95-111),
  Init_Ggrid's cutoffs (a.log:171-174), the initial density's sum/min/max
  before and after normalisation (a.log:283-286), the Ewald energy
  (a.log:302), the converged fixed-potential Echk (a.log:392), the SCF
  trace (total and Harris energies, |rho - rho0|^2 and |Vloc - Vloc0|^2 per
  iteration, a.log:403-720) and the final SCF state of fort.99 (total
  energy, its terms, Fermi level, the 648 eigenvalues with their residual
  estimate esp_err and occupations).

Run in a container where the reference tree is mounted:
    python tests/golden/make_si216_psv_fixture.py /root/reference
"""
import gzip
import json
import os
import re
import sys


def main(ref_root):
    ex = os.path.join(ref_root, "examples", "Si216")
    out_dir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "si216_psv")
    os.makedirs(out_dir, exist_ok=True)

    with open(os.path.join(ex, "Si_psv.dat"), "rb") as f:
        raw = f.read()
    with gzip.GzipFile(os.path.join(out_dir, "Si_psv.dat.gz"), "wb", mtime=0) as g:
        g.write(raw)

    lines = open(os.path.join(ex, "fort.970")).read().splitlines()
    head = lines[0].split()
    nelem, natom = int(head[0]), int(head[1])
    frac, kind = [], []
    for l in lines[1:1 + natom]:
        f = l.split()
        kind.append(int(f[0]))
        frac.append([float(x) for x in f[1:4]])
    atoms = {"source": "examples/Si216/fort.970 (positions), fort.1 (cell, grid)",
             "AX": 30.783, "A": [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]],
             "Ngrid": [48, 48, 48], "Md": 6, "nelement": nelem, "zatom": [int(head[2])],
             "kind": kind, "frac": frac}
    json.dump(atoms, open(os.path.join(out_dir, "atoms.json"), "w"), indent=0)

    log = open(os.path.join(ex, "a.log")).read()

    def grab(pattern, cast=float, flags=0):
        m = re.search(pattern, log, flags)
        if m is None:
            raise SystemExit(f"pattern not found: {pattern}")
        return [cast(x) for x in m.groups()]

    res = {"source": "examples/Si216/a.log and fort.99 (reference run)"}
    res["psv_sum_cdd"] = grab(r"sum\(cdd\)\s*=\s*(\S+)")[0]
    res["psv_anorm"] = grab(r"anorm\s*=\s*(\S+)\s+(\S+)")
    res["psv_mr"] = grab(r"# of radial mesh points =\s*(\d+)", int)[0]
    res["psv_parloc"] = grab(r"a1,a2,a3,a4\s*=\s*(\S+)\s+(\S+)\s*\n\s*(\S+)\s+(\S+)")
    res["ecut"], res["gcut"] = grab(r"Ecut,Gcut=\s*(\S+)\s+(\S+)")
    res["ngsphere"] = grab(r"NGgrid=\s*(\d+)", int)[0]
    res["nmgl"] = grab(r"NMGL=\s*(\d+)", int)[0]
    rows = re.findall(r"^\s*864\.00000\s+(\S+)\s+(\S+)\s+(\S+)\s*$", log, re.M)
    res["initrho_before"] = [float(x) for x in rows[0]]
    res["initrho_after"] = [float(x) for x in rows[1]]
    res["eewald"] = grab(r"Eewald\(SYSTYPE=0\)=\s*(\S+)")[0]
    echk = re.findall(r"Echk,dif/tol =\s*(\S+)\s+(\S+)", log)
    res["echk_trace"] = [[float(a), float(b)] for a, b in echk]
    res["nband"] = grab(r"Nband is replaced to\s+(\d+)", int)[0]
    harris = re.findall(r"Harris Energy:\s*(\S+)", log)
    total = re.findall(r"Total Energy :\s*(\S+)", log)
    res["scf_trace"] = [[float(t), float(h)] for t, h in zip(total, harris)]
    sqv = re.findall(r"\|Vloc-Vloc0\|\^2/tol =\s*(\S+)", log)
    sqr = re.findall(r"\|rho -rho0 \|\^2/tol =\s*(\S+)", log)
    res["scf_sqerr_rho_v"] = [[float(a), float(b)] for a, b in zip(sqr, sqv)]

    f99 = open(os.path.join(ex, "fort.99")).read().splitlines()
    keys = {"Total Energy": "etot", "Harris Energy": "ehwf", "Ion-Ion": "eewald_final",
            "Local Potential": "eloc", "Ion Local Potential": "eion", "Hartree Energy": "eh",
            "Exchange-Correlation Energy": "exc", "Exchange Energy": "ex",
            "Correlation Energy": "ec", "Sum of eigenvalues": "eeig", "Fermi energy": "efermi"}
    final = {}
    for l in f99:
        for k, name in keys.items():
            if l.strip().startswith(k) and name not in final:
                rest = l.strip()[len(k):].split()
                if rest and re.match(r"^-?\d", rest[0]):
                    final[name] = float(rest[0])
    esp, err, occ = [], [], []
    for l in f99:
        f = l.split()
        if len(f) == 5 and f[0] == "1" and f[1].isdigit():
            esp.append(float(f[2]))
            err.append(float(f[3]))
            occ.append(float(f[4]))
    final["esp"], final["esp_err"], final["occ"] = esp, err, occ
    res["final"] = final
    json.dump(res, open(os.path.join(out_dir, "results.json"), "w"), indent=0)
    print(f"wrote {out_dir}: {natom} atoms, {len(esp)} eigenvalues, "
          f"{len(res['echk_trace'])} sweep lines, {len(res['scf_trace'])} scf lines")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
