"""Benchmark: seconds per CG+GS+subspace sweep, Si216 (72^3 grid, 480 bands).

One "step" is one sweep of the calc_sweep body (sweep_module.f90:147-158):
batched conjugate_gradient -> gram_schmidt -> subspace_diag over all 480
bands at fixed potential, wavefunctions resident in HBM (attached once before
the timed region).  The value is the mixed-precision (FP32 storage and
operators, FP64 small dense algebra) seconds per sweep; the FP64 sweep time,
their ratio and the eigenvalue error of the mixed run (after two FP64
refinement sweeps) against the FP64 run are reported beside it.

The roofline object is for the dominant operator of the sweep, the FP32
hamiltonian over all bands (three kernels), timed live here with HIP events on
the library's stream, against its minimal HBM bytes; each kernel is listed
with its own bytes.  The sweep roofline (sum of bytes / HBM peak + flops /
MFMA peak over the sweep's kernels), the GEMMs' MFMA utilisation and the
accuracy gate (pass/fail at 1e-6 Ha) are reported beside it.  The CPU
baseline is one sweep of the oracle's OpenMP FP64 restatement.

Usage: python bench.py [--gpus 1] [--steps 9] [--warmup 1]

HBM traffic of the hamiltonian's kernels comes from two rocprofv3 counter runs of
`python bench.py --traffic-probe` (FETCH_SIZE, then WRITE_SIZE), each scaled by
what the same counter reads for the library's plain copy of all bands; they are
summarised with `python bench.py --pmc-summary FETCH_DIR WRITE_DIR` into
profiles/hpsi_traffic_si216.json, which this script reads when present.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6, "bf16": 2500.0}    # dense MFMA, MI355X
# sweeps of the accuracy comparison (the FP64 Si216 run is 10 sweeps; the
# mixed run replaces the first 8 by mixed-precision sweeps)
NSWEEP = 10
NMIXED = 8
NCONV = 40       # sweeps of the near-convergence comparison
NSWEEP_MID = 30  # the FP64 run's drift from here to NCONV marks the bands still moving
EIG_TOL_HA = 1e-6


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hpsi_traffic_si216.json")
# the kernels of one FP32 H psi (rocprofv3 kernel-name patterns)
HPSI_KERNELS = {
    "projection": r"k_nl_project<float,",
    "nonlocal_fill": r"k_nl_fill<float,",
    "stencil": r"k_hpsi2<float, \d+, \d+, \d+, 11,",
}


def sphere_points(sysm):
    """Nat * P: grid points of all projector spheres (one channel per atom)."""
    return int(np.sum(sysm.proj.MJJ_MAP[::4]))


def covered_points(sysm):
    """Distinct grid points inside some projector sphere."""
    pr = sysm.proj
    return int(np.unique(np.concatenate([pr.JJP[l, :pr.MJJ[l]] for l in range(pr.nzlma)])).size)


def hpsi_bytes(sysm, nb, s):
    """Minimal HBM bytes of one H psi over nb bands, band block = all bands:
    psi read + H psi write, Vloc once, the projector values (4 channels) and
    an int32 index per sphere point once:  2 ML nb s + ML s + Nat P (4 s + 4)."""
    ML = sysm.grid.ML
    return 2 * ML * nb * s + ML * s + sphere_points(sysm) * (4 * s + 4)


def hpsi_flops(sysm, nb):
    """57 ML + 16 Nat P per band: stencil 55, Vloc 2, 4 projectors x (project
    + apply) x 2 per sphere point."""
    return nb * (57 * sysm.grid.ML + 16 * sphere_points(sysm))


def hpsi_kernel_bytes(sysm, nb, s):
    """Algorithmic bytes of each kernel of one H psi as this library runs it:
    projection reads psi at the covered points and the tables; the fill writes
    the nonlocal term at every point; the stencil reads psi and that term,
    writes H psi, reads Vloc."""
    ML = sysm.grid.ML
    tables = sphere_points(sysm) * (4 * s + 4)
    return {"projection": covered_points(sysm) * nb * s + tables,
            "nonlocal_fill": ML * nb * s,
            "stencil": 3 * ML * nb * s + ML * s}


def sweep_ledger(sysm, nb, s, reject_frac=0.0):
    """Algorithmic work of one fused sweep (CG with Ncg = 3 and the s-step
    preconditioner, then Cholesky + Rayleigh-Ritz): kernel -> (calls, bytes
    per call, flops per call).  A = ML nb s is one band-sized array."""
    ML = sysm.grid.ML
    A = ML * nb * s
    sym = ML * nb * (nb + 1)          # one triangle of an nb x nb product
    return {
        "hamiltonian": (4 + reject_frac, hpsi_bytes(sysm, nb, s), hpsi_flops(sysm, nb)),
        "cg_first_dots_and_gradient": (1, 5 * A, 4 * ML * nb),
        "precond_gram_pass": (3, A, 0),
        "precond_apply_pass": (3, 3 * A - A / 3, 0),     # no old direction at icg = 1
        "cg_2x2_dots": (3, 4 * A, 12 * ML * nb),
        "cg_update": (3, 7 * A, 10 * ML * nb),
        "overlap_xtx": (1, A, sym),
        "xt_hx": (1, 2 * A, sym),
        "rotation": (1, 2 * A, 2 * ML * nb * nb),
    }


def ledger_time_ms(ledger, peak_tflops):
    """Sum over kernels of bytes / HBM peak + flops / MFMA peak."""
    t = 0.0
    for calls, b, f in ledger.values():
        t += calls * (b / (HBM_PEAK_GBS * 1e9) + f / (peak_tflops * 1e12))
    return t * 1e3


def pmc_summary(fetch_dir, write_dir, out_path):
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of each FP32 hamiltonian
    kernel, scaled by what the same counters read for the plain copy of all
    bands (whose algorithmic bytes are known)."""
    import csv
    import glob
    import re
    from rsdft_amd.system import make_config

    def load(d, counter):
        vals = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] == counter:
                    vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return vals

    sysm = make_config("si216")
    ML, nb = sysm.grid.ML, sysm.nband
    copy_bytes = ML * nb * 4            # read (and write) of one FP32 copy of all bands
    res = {"kernels": {}}
    total = 0.0
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        vals = load(d, counter)
        copy = [k for k in vals if "k_copy_bands<float>" in k]
        if not copy:
            raise SystemExit(f"{counter}: copy kernel not found in {d}")
        cv = float(np.mean(vals[copy[0]])) * 1024
        scale = copy_bytes / cv
        res[counter + "_scale"] = scale
        for part, pat in HPSI_KERNELS.items():
            kern = [k for k in vals if re.search(pat, k)]
            if not kern:
                raise SystemExit(f"{counter}: {part} kernel not found in {d}")
            kv = float(np.mean(vals[kern[0]])) * 1024 * scale
            e = res["kernels"].setdefault(part, {"kernel": kern[0]})
            e[counter] = kv
            e["dispatches"] = len(vals[kern[0]])
            total += kv
    res["traffic_bytes"] = total
    res["algorithmic_bytes"] = hpsi_bytes(sysm, nb, 4)
    res["commands"] = [
        "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 90 rocprofv3 --pmc FETCH_SIZE "
        "-d <repo>/gpurun_out/pmc_fetch -o run -f csv -- python <repo>/bench.py --traffic-probe",
        "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 90 rocprofv3 --pmc WRITE_SIZE "
        "-d <repo>/gpurun_out/pmc_write -o run -f csv -- python <repo>/bench.py --traffic-probe",
        "python bench.py --pmc-summary gpurun_out/pmc_fetch gpurun_out/pmc_write",
    ]
    with open(out_path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


# MI355X: 256 CUs in 8 XCDs, 4 SIMDs per CU; GRBM_GUI_ACTIVE counts the
# cycles of every XCD, so one dispatch's duration in cycles is its sum / 8
MI355X_CUS, MI355X_XCDS, SIMDS_PER_CU = 256, 8, 4
SQ_PASSES = (("SQ_WAVES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES", "GRBM_GUI_ACTIVE"),
             ("SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT", "SQ_BUSY_CYCLES", "SQ_INSTS_SALU"))


def sq_summary(a_dir, b_dir, out_path):
    """Per-dispatch SQ counters of the sweep's kernels from the two counter
    runs of --sweep-probe, as a markdown table: VALU busy = 4 cycles per wave64
    VALU instruction over the SIMD cycles of the dispatch; LDS conflict share =
    bank-conflict cycles per CU over the dispatch's cycles."""
    import csv
    import glob
    vals = {}
    for d in (a_dir, b_dir):
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                vals.setdefault(r["Kernel_Name"], {}).setdefault(r["Counter_Name"], []).append(
                    float(r["Counter_Value"]))
    rows = []
    for k, v in vals.items():
        if not all(c in v for p in SQ_PASSES for c in p):
            continue
        m = {c: float(np.mean(x)) for c, x in v.items()}
        cyc = m["GRBM_GUI_ACTIVE"] / MI355X_XCDS
        if cyc <= 0:
            continue
        rows.append((cyc * len(v["GRBM_GUI_ACTIVE"]), k, len(v["GRBM_GUI_ACTIVE"]), cyc, m))
    rows.sort(reverse=True)
    lines = ["| kernel | dispatches | cycles | waves | VALU instr | VALU busy | LDS instr "
             "| LDS bank-conflict cycles | conflict share | SALU instr |",
             "|---|---|---|---|---|---|---|---|---|---|"]
    for _, k, n, cyc, m in rows[:16]:
        name = k.replace("void ", "").split("(")[0]
        if len(name) > 70:
            name = name[:70] + "…"
        busy = 4 * m["SQ_ACTIVE_INST_VALU"] / (MI355X_CUS * SIMDS_PER_CU * cyc)
        share = m["SQ_LDS_BANK_CONFLICT"] / MI355X_CUS / cyc
        lines.append(f"| `{name}` | {n} | {cyc / 1e6:.2f}M | {m['SQ_WAVES']:.0f} "
                     f"| {m['SQ_INSTS_VALU'] / 1e6:.1f}M | {100 * busy:.0f} % "
                     f"| {m['SQ_INSTS_LDS'] / 1e6:.1f}M | {m['SQ_LDS_BANK_CONFLICT'] / 1e6:.2f}M "
                     f"| {100 * share:.1f} % | {m['SQ_INSTS_SALU'] / 1e6:.1f}M |")
    cmds = [f"cd /tmp && export TMPDIR=/tmp && timeout -s KILL 120 rocprofv3 --pmc {' '.join(p)} "
            f"-d <repo>/gpurun_out/pmc_sq_{t} -o run -f csv -- python <repo>/bench.py --sweep-probe"
            for t, p in zip("ab", SQ_PASSES)]
    cmds.append("python bench.py --sq-summary gpurun_out/pmc_sq_a gpurun_out/pmc_sq_b")
    text = ("# SQ counters: Si216 mixed-precision sweep kernels (per dispatch)\n\n"
            "Commands (from the repository root on one MI355X; `<repo>` is that root):\n\n```\n"
            + "\n".join(cmds) + "\n```\n\n"
            "`--sweep-probe` runs one warm-up and two timed mixed-precision sweeps of Si216 "
            "(72^3 grid, 480 bands).  Cycles: GRBM_GUI_ACTIVE / 8 XCDs.  VALU busy: "
            "4 x SQ_ACTIVE_INST_VALU / (256 CUs x 4 SIMDs x cycles).  Conflict share: "
            "SQ_LDS_BANK_CONFLICT / 256 CUs / cycles.  Sorted by total cycles.  rocprofv3 "
            "may end with status 139 in the process's exit handlers after the program's "
            "last line; the counter files are complete then.\n\n"
            + "\n".join(lines) + "\n")
    with open(out_path, "w") as f:
        f.write(text)
    print(text)


def traffic_probe():
    """A few FP32 all-band applications of every hamiltonian piece and the copy
    (the counter runs profile this)."""
    from rsdft_amd.system import make_config
    from rsdft_amd.gpu import RsdftGpu, PREC_MIXED_FP32
    sysm = make_config("si216")
    x = sysm.psi0()
    dev = RsdftGpu(sysm, precision=PREC_MIXED_FP32)
    dev.attach(x)
    t = dev.test_hpsi(3)
    log(f"probe hpsi ms: {t}")
    dev.finalize()


def sweep_probe():
    """Two mixed-precision Si216 sweeps (FP32 storage, three-piece BF16 GEMMs)
    after one warm-up sweep (the counter runs of the sweep's kernels profile
    this)."""
    from rsdft_amd.system import make_config
    from rsdft_amd.gpu import RsdftGpu, PREC_MIXED_BF16
    sysm = make_config("si216")
    x = sysm.psi0()
    dev = RsdftGpu(sysm, precision=PREC_MIXED_BF16)
    dev.attach(x)
    dev.gram_schmidt(x)
    dev.sweep(x, 1)
    dev.reset_timers()
    dev.sweep(x, 2)
    log(f"probe sweeps: {dev.timers()}")
    dev.finalize()


def cpu_baseline(sysm, x0):
    """One oracle FP64 sweep (C/OpenMP restatement) from x0 on the host cores:
    CG on all bands, then Gram-Schmidt and subspace diag.  Returns the
    cpu_baseline object and the oracle's eigenvalues."""
    from oracle.oracle import Oracle
    o = Oracle(sysm)
    nb = x0.shape[0]
    t0 = time.perf_counter()
    xc, _, _ = o.conjugate_gradient(x0)
    t_cg = time.perf_counter() - t0
    log(f"cpu: CG {nb} bands {t_cg:.2f}s")
    t0 = time.perf_counter()
    q = o.gram_schmidt(xc)
    t_gs = time.perf_counter() - t0
    log(f"cpu: GS {nb} bands {t_gs:.2f}s")
    t0 = time.perf_counter()
    _, esp, _ = o.subspace_diag(q)
    t_sd = time.perf_counter() - t0
    log(f"cpu: subspace {nb} bands {t_sd:.2f}s")
    value = t_cg + t_gs + t_sd
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": value,
        "unit": "s/sweep",
        "cores": threads,
        "kind": "port",
        "sample": (f"one oracle FP64 sweep of the same workload (CG on all {nb} bands, "
                   f"Gram-Schmidt, subspace diag) from the same start; "
                   f"{value:.1f}s of CPU work on {threads} OpenMP threads"),
    }, esp


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=9)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="si216")
    ap.add_argument("--bands", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--hpsi-reps", type=int, default=20)
    ap.add_argument("--gemm", choices=["fp32", "bf16"], default="bf16",
                    help="GEMMs of the mixed path: bf16 = three-piece BF16 split (six "
                         "v_mfma_f32_16x16x32_bf16 products per block), fp32 = FP32 MFMA")
    ap.add_argument("--pc-mode", type=int, choices=[0, 1], default=1,
                    help="preconditioner: 1 s-step (default), 0 step-by-step recurrence")
    ap.add_argument("--traffic-probe", action="store_true")
    ap.add_argument("--sweep-probe", action="store_true")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
    ap.add_argument("--sq-summary", nargs=2, metavar=("PASS_A_DIR", "PASS_B_DIR"))
    args = ap.parse_args()
    if args.sq_summary:
        sq_summary(*args.sq_summary, os.path.join(ROOT, "profiles", "sweep_si216_sq_counters.md"))
        return
    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], TRAFFIC_FILE)
        return
    if args.traffic_probe:
        traffic_probe()
        return
    if args.sweep_probe:
        sweep_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the path runs on one MI355X; multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    from rsdft_amd.system import make_config
    from rsdft_amd.gpu import RsdftGpu, PREC_FP64, PREC_MIXED_FP32, PREC_MIXED_BF16

    K, W = args.steps, args.warmup
    sysm = make_config(args.workload, nband=args.bands)
    nb, ML = sysm.nband, sysm.grid.ML
    log(f"system {sysm.name}: grid {sysm.grid.N}, {nb} bands, {sysm.proj.nzlma} projectors")
    t0 = time.perf_counter()
    x = sysm.psi0()
    log(f"psi0 generated in {time.perf_counter() - t0:.1f}s")

    dev = RsdftGpu(sysm, precision=PREC_FP64)
    dev.set_precond_mode(args.pc_mode)
    dev.attach(x)
    dev.gram_schmidt(x)
    dev.sync()
    x0 = x.copy()
    log("initial Gram-Schmidt done")

    # ---- FP64 sweeps -------------------------------------------------------
    dev.sweep(x, W)
    dev.reset_timers()
    t0 = time.perf_counter()
    dev.sweep(x, K)
    t_fp64 = (time.perf_counter() - t0) / K
    tim64 = dev.timers()
    log(f"FP64: {t_fp64 * 1e3:.2f} ms/sweep  {tim64}")
    hp64 = dev.test_hpsi(args.hpsi_reps)
    gm64 = dev.test_gemm(5)
    log(f"FP64 hpsi ms: {hp64}; gemm ms {gm64}")
    dev.detach()

    # ---- mixed sweeps --------------------------------------------------------
    mode = PREC_MIXED_BF16 if args.gemm == "bf16" else PREC_MIXED_FP32
    dev.set_precision(mode)
    xm = x0.copy()
    dev.attach(xm)
    dev.sweep(xm, W)
    dev.reset_timers()
    t0 = time.perf_counter()
    dev.sweep(xm, K)
    t_mix = (time.perf_counter() - t0) / K
    timm = dev.timers()
    log(f"mixed: {t_mix * 1e3:.2f} ms/sweep  {timm}")
    hpm = dev.test_hpsi(args.hpsi_reps)
    gmm = dev.test_gemm(5)
    log(f"mixed hpsi ms: {hpm}; gemm ms {gmm}")
    dev.detach()

    # ---- accuracy: NSWEEP FP64 sweeps against NMIXED mixed + the rest FP64 ----
    def run(start, nmixed, nfp64):
        y = start.copy()
        if nmixed:
            dev.set_precision(mode)
            dev.attach(y)
            dev.sweep(y, nmixed)
            dev.sync()
            dev.detach()
        dev.set_precision(PREC_FP64)
        dev.attach(y)
        esp = dev.sweep(y, nfp64)[1]
        dev.detach()
        return esp

    esp64_1 = run(x0, 0, 1)
    esp64 = run(x0, 0, NSWEEP)
    espm = run(x0, NMIXED, NSWEEP - NMIXED)
    # sensitivity floor: FP64 from the start rounded once to FP32
    espp = run(x0.astype(np.float32).astype(np.float64), 0, NSWEEP)
    # the same comparison near convergence: NCONV FP64 sweeps against
    # NCONV - 2 mixed + 2 FP64, and the FP64 run's own drift from NSWEEP_MID
    # to NCONV sweeps (the bands still moving at NCONV)
    esp64c = run(x0, 0, NCONV)
    esp64m = run(x0, 0, NSWEEP_MID)
    espmc = run(x0, NCONV - 2, 2)
    dev.finalize()
    occ = sysm.occ

    def errs(e):
        d = np.abs(e - esp64)
        return float(np.max(d)), float(np.max(d[occ > 0])), float(abs(np.sum(occ * (e - esp64))))

    eig_err, eig_err_occ, eband_err = errs(espm)
    flo_err, flo_err_occ, flo_eband = errs(espp)
    dc = np.abs(espmc - esp64c)
    drift = np.abs(esp64c - esp64m)
    conv = {
        "protocol": f"{NCONV} FP64 sweeps vs {NCONV - 2} mixed + 2 FP64 sweeps from the same start",
        "max_eig_err_ha": float(np.max(dc)),
        "max_eig_err_occupied_ha": float(np.max(dc[occ > 0])),
        "band_energy_err_ha": float(abs(np.sum(occ * (espmc - esp64c)))),
        "bands_over_tol": int(np.sum(dc > 1e-6)),
        "fp64_drift_ha": {"from_sweeps": NSWEEP_MID, "to_sweeps": NCONV,
                          "max": float(np.max(drift)), "occupied": float(np.max(drift[occ > 0])),
                          "bands_over_tol": int(np.sum(drift > 1e-6))},
    }
    log(f"after {NCONV} sweeps: max eig err {conv['max_eig_err_ha']:.3e} Ha (occupied "
        f"{conv['max_eig_err_occupied_ha']:.3e}), band energy {conv['band_energy_err_ha']:.3e}, "
        f"{conv['bands_over_tol']} bands over 1e-6; FP64 drift {NSWEEP_MID}->{NCONV}: "
        f"{conv['fp64_drift_ha']['max']:.3e} ({conv['fp64_drift_ha']['bands_over_tol']} bands over 1e-6)")
    d = np.abs(espm - esp64)
    worst = np.argsort(d)[::-1][:5]
    log(f"after {NSWEEP} sweeps ({NMIXED} mixed + {NSWEEP - NMIXED} FP64 vs {NSWEEP} FP64): max "
        f"eig err {eig_err:.3e} Ha (occupied {eig_err_occ:.3e}), band energy {eband_err:.3e}; "
        f"worst bands {[int(i) for i in worst]} errs {[f'{d[i]:.2e}' for i in worst]}")
    log(f"FP64 from the FP32-rounded start: max {flo_err:.3e} Ha (occupied {flo_err_occ:.3e}), "
        f"band energy {flo_eband:.3e}")
    t_protocol = (NMIXED * t_mix + (NSWEEP - NMIXED) * t_fp64) / NSWEEP

    # ---- roofline: the whole FP32 H psi (its kernels, HIP events) -----------
    s32 = 4
    B = hpsi_bytes(sysm, nb, s32)
    t_h = float(hpm[0])
    achieved = B / (t_h * 1e-3) / 1e9
    kb = hpsi_kernel_bytes(sysm, nb, s32)
    kt = {"projection": float(hpm[5]), "nonlocal_fill": float(hpm[6]), "stencil": float(hpm[4])}
    per_kernel = {k: {"ms": kt[k], "algorithmic_bytes": kb[k],
                      "gbs": kb[k] / (kt[k] * 1e-3) / 1e9,
                      "frac": kb[k] / (kt[k] * 1e-3) / 1e9 / HBM_PEAK_GBS} for k in kt}
    B64 = hpsi_bytes(sysm, nb, 8)
    copy_gbs = 2 * ML * nb * s32 / (float(hpm[7]) * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        if "kernels" in tr:
            traffic = float(tr["traffic_bytes"])
            for k, e in tr["kernels"].items():
                if k in per_kernel:
                    per_kernel[k]["traffic"] = e["FETCH_SIZE"] + e["WRITE_SIZE"]

    # ---- GEMMs on MFMA -------------------------------------------------------
    sym_flops = ML * nb * (nb + 1)
    rot_flops = 2 * ML * nb * nb

    def gemm_util(ms, peak, products=1):
        """tflops: useful flops of the product; mfma_util: the matrix-core
        work issued (products x flops: six BF16 piece products per FP32
        product in the split mode) over the peak of those instructions."""
        fl = (sym_flops, sym_flops, rot_flops)
        names = ("overlap_xtx", "xt_hx", "rotation")
        return {n: {"ms": float(t), "tflops": f / (t * 1e-3) / 1e12,
                    "mfma_util": products * f / (t * 1e-3) / 1e12 / peak}
                for n, t, f in zip(names, ms, fl)}

    # ---- sweep roofline denominators (sum of bytes/BW + flops/peak) ----------
    rej = float(timm["cg_rejected_bands"]) / nb
    led32 = sweep_ledger(sysm, nb, 4, rej)
    led64 = sweep_ledger(sysm, nb, 8, float(tim64["cg_rejected_bands"]) / nb)
    # the split mode's matrix-core ceiling per useful flop: six BF16 products
    roof32 = ledger_time_ms(led32, MFMA_PEAK_TFLOPS["bf16"] / 6 if args.gemm == "bf16"
                            else MFMA_PEAK_TFLOPS["f32"])
    roof64 = ledger_time_ms(led64, MFMA_PEAK_TFLOPS["f64"])

    cpu, esp_oracle = (None, None)
    if not args.no_cpu:
        cpu, esp_oracle = cpu_baseline(sysm, x0)
        log(f"cpu baseline {cpu['value']:.2f} s/sweep; GPU FP64 first sweep vs oracle: "
            f"{np.max(np.abs(esp64_1 - esp_oracle)):.3e} Ha")

    def phases(tim):
        return {k: v / K for k, v in tim.items()
                if k not in ("hpsi_bands", "eig_refined", "cg_rejected_bands",
                             "pc_recurrence_bands")}

    out = {
        "metric": "s/sweep Si216 mixed vs FP64 (@1e-6 Ha eig err); hpsi HBM GB/s vs peak",
        "value": t_mix,
        "unit": "s/sweep",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": t_mix * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{sysm.name}: diamond Si, {sysm.grid.N[0]}^3 grid, {nb} bands, Md={sysm.Md}, "
                        f"{sysm.proj.nzlma} s+p KB projectors; CG(Ncg=3,mloop=3)+GS+subspace sweep",
            "mixed_gemm": "bf16x3 (three-piece BF16 split, six products)" if args.gemm == "bf16"
                          else "fp32",
            "fp64_s_per_sweep": t_fp64,
            "fp64_over_mixed": t_fp64 / t_mix,
            "accuracy_protocol": f"{NSWEEP} FP64 sweeps vs {NMIXED} mixed + {NSWEEP - NMIXED} FP64 "
                                 "sweeps from the same orthonormalised seeded start",
            "protocol_s_per_sweep": t_protocol,
            "fp64_over_protocol": t_fp64 / t_protocol,
            "accuracy_gate": {
                "tol_ha": EIG_TOL_HA,
                "max_eig_err_ha": eig_err,
                "max_eig_err_occupied_ha": eig_err_occ,
                "band_energy_err_ha": eband_err,
                "all_bands_pass": eig_err <= EIG_TOL_HA,
                "occupied_bands_pass": eig_err_occ <= EIG_TOL_HA,
                "band_energy_pass": eband_err <= EIG_TOL_HA,
                "fp64_rounded_start_eig_err_ha": flo_err,
                "fp64_rounded_start_eig_err_occupied_ha": flo_err_occ,
                "fp64_rounded_start_band_energy_err_ha": flo_eband,
            },
            "accuracy_near_convergence": conv,
            "fp64_first_sweep_vs_oracle_max_eig_err_ha":
                None if esp_oracle is None else float(np.max(np.abs(esp64_1 - esp_oracle))),
            "hpsi_ms_fp32": t_h,
            "hpsi_ms_fp64": float(hp64[0]),
            "hpsi_gbs_fp64": B64 / (float(hp64[0]) * 1e-3) / 1e9,
            "hpsi_gflops_fp32": hpsi_flops(sysm, nb) / (t_h * 1e-3) / 1e9,
            "hpsi_gflops_fp64": hpsi_flops(sysm, nb) / (float(hp64[0]) * 1e-3) / 1e9,
            "hpsi_parts_ms_fp32": {k: float(v) for k, v in zip(
                ("hamiltonian", "kinetic", "localpot", "nonlocal", "stencil", "projection",
                 "nl_stage", "copy"), hpm)},
            "copy_gbs_fp32": copy_gbs,
            "gemm_mixed": (gemm_util(gmm, MFMA_PEAK_TFLOPS["bf16"], 6) if args.gemm == "bf16"
                           else gemm_util(gmm, MFMA_PEAK_TFLOPS["f32"])),
            "gemm_fp64": gemm_util(gm64, MFMA_PEAK_TFLOPS["f64"]),
            "sweep_roofline": {
                "mixed_ms": roof32, "mixed_frac": roof32 / (t_mix * 1e3),
                "fp64_ms": roof64, "fp64_frac": roof64 / (t_fp64 * 1e3),
                "note": "sum over the sweep's streaming kernels and GEMMs of bytes / 8 TB/s + "
                        "flops / MFMA peak; the FP64 Cholesky and eigensolve (latency-bound) "
                        "are not in it",
            },
            "phase_ms_mixed": phases(timm),
            "hpsi_bands_per_sweep_mixed": timm["hpsi_bands"] / K,
            "pc_recurrence_bands_per_sweep_mixed": timm["pc_recurrence_bands"] / K,
            "pc_recurrence_bands_per_sweep_fp64": tim64["pc_recurrence_bands"] / K,
            "eig_refined_mixed": timm["eig_refined"],
            "phase_ms_fp64": phases(tim64),
            "hpsi_bands_per_sweep_fp64": tim64["hpsi_bands"] / K,
            "eig_refined_fp64": tim64["eig_refined"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "hamiltonian, FP32, all 480 bands: k_nl_project + k_nl_fill + "
                      "k_hpsi2<float,6,2,1,kin|loc|acc,1024> (Md=6 FD stencil + Vloc + KB nonlocal)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes": B,
            "per_kernel": per_kernel,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
