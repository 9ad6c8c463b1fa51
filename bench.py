"""Benchmark: seconds per CG+GS+subspace sweep, Si216 (72^3 grid, 480 bands).

One "step" is one sweep of the calc_sweep body (sweep_module.f90:147-158):
batched conjugate_gradient -> gram_schmidt -> subspace_diag over all 480
bands at fixed potential, wavefunctions resident in HBM (attached once before
the timed region).  The value is the mixed-precision (FP32 storage and
operators, FP64 small dense algebra) seconds per sweep; the FP64 sweep time,
their ratio and the eigenvalue error of the mixed run (after two FP64
refinement sweeps) against the FP64 run are reported beside it.

The roofline object is for the dominant kernel of the sweep, timed live here
with HIP events on the library's stream.  The CPU baseline is the oracle's
OpenMP FP64 restatement on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps 9] [--warmup 1]

HBM traffic of the roofline kernel comes from two rocprofv3 counter runs of
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
# sweeps of the accuracy comparison (the FP64 Si216 run is 10 sweeps; the
# mixed run replaces the first 8 by mixed-precision sweeps)
NSWEEP = 10


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hpsi_traffic_si216.json")


def hpsi_stencil_bytes(sysm, nb, s, acc):
    """Algorithmic HBM bytes of one all-band launch of the stencil kernel:
    psi read + H psi write per band and Vloc once; accumulating (H psi) also
    the read of the nonlocal term that the grid-order fill wrote into H psi,
    one more band array."""
    ML = sysm.grid.ML
    b = 2 * ML * nb * s + ML * s
    if acc:
        b += ML * nb * s
    return b


def nl_fill_bytes(sysm, nb, s=4):
    """Algorithmic HBM bytes of one all-band launch of the grid-order
    nonlocal fill: the term written at every grid point of every band."""
    return sysm.grid.ML * nb * s


def pmc_summary(fetch_dir, write_dir, out_path):
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of the FP32 hamiltonian stencil
    kernel, scaled by what the same counters read for the plain copy of all bands."""
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
    res = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        vals = load(d, counter)
        kern = [k for k in vals if re.search(r"k_hpsi2<float, \d+, \d+, \d+, 11,", k)]
        copy = [k for k in vals if "k_copy_bands<float>" in k]
        if not kern or not copy:
            raise SystemExit(f"{counter}: kernels not found in {d}")
        kv = float(np.mean(vals[kern[0]])) * 1024
        cv = float(np.mean(vals[copy[0]])) * 1024
        scale = copy_bytes / cv
        res[counter] = {"kernel": kern[0], "kernel_bytes_raw": kv, "copy_bytes_raw": cv,
                        "copy_bytes_algorithmic": copy_bytes, "scale": scale,
                        "kernel_bytes": kv * scale, "dispatches": len(vals[kern[0]])}
    res["traffic_bytes"] = res["FETCH_SIZE"]["kernel_bytes"] + res["WRITE_SIZE"]["kernel_bytes"]
    res["algorithmic_bytes"] = hpsi_stencil_bytes(sysm, nb, 4, True)
    res["commands"] = [
        "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 60 rocprofv3 --pmc FETCH_SIZE "
        "-d <repo>/gpurun_out/pmc_fetch -o run -f csv -- python <repo>/bench.py --traffic-probe",
        "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 60 rocprofv3 --pmc WRITE_SIZE "
        "-d <repo>/gpurun_out/pmc_write -o run -f csv -- python <repo>/bench.py --traffic-probe",
        "python bench.py --pmc-summary gpurun_out/pmc_fetch gpurun_out/pmc_write",
    ]
    with open(out_path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


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


def cpu_baseline(sysm, x0, nsample):
    """Oracle FP64 sweep on the host: CG on `nsample` bands (scaled to all
    bands) plus full-size Gram-Schmidt and subspace diagonalisation."""
    from oracle.oracle import Oracle
    o = Oracle(sysm)
    nb = x0.shape[0]
    if nsample <= 0 or nsample > nb:
        nsample = nb
    t0 = time.perf_counter()
    o.conjugate_gradient(x0[:nsample])
    t_cg = time.perf_counter() - t0
    log(f"cpu: CG {nsample} bands {t_cg:.2f}s")
    t0 = time.perf_counter()
    q = o.gram_schmidt(x0)
    t_gs = time.perf_counter() - t0
    log(f"cpu: GS {nb} bands {t_gs:.2f}s")
    t0 = time.perf_counter()
    o.subspace_diag(q)
    t_sd = time.perf_counter() - t0
    log(f"cpu: subspace {nb} bands {t_sd:.2f}s")
    value = t_cg * nb / nsample + t_gs + t_sd
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": value,
        "unit": "s/sweep",
        "cores": threads,
        "kind": "port",
        "sample": ((f"oracle FP64 sweep on all {nb} bands (CG, Gram-Schmidt, subspace diag)"
                    if nsample == nb else
                    f"oracle FP64 CG on {nsample} of {nb} bands scaled x{nb / nsample:g}, plus "
                    f"full {nb}-band Gram-Schmidt and subspace diag")
                   + f"; measured {t_cg + t_gs + t_sd:.1f}s of CPU work"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=9)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="si216")
    ap.add_argument("--bands", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="bands of the CPU CG sample (0: all bands, no scaling)")
    ap.add_argument("--hpsi-reps", type=int, default=20)
    ap.add_argument("--gemm", choices=["fp32", "bf16"], default="fp32",
                    help="GEMM precision of the mixed path")
    ap.add_argument("--traffic-probe", action="store_true")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], TRAFFIC_FILE)
        return
    if args.traffic_probe:
        traffic_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the path runs on one MI355X; multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    from rsdft_amd.system import make_config
    from rsdft_amd.gpu import RsdftGpu, PREC_FP64, PREC_MIXED_FP32, PREC_MIXED_BF16

    K, W = args.steps, args.warmup
    sysm = make_config(args.workload, nband=args.bands)
    nb = sysm.nband
    log(f"system {sysm.name}: grid {sysm.grid.N}, {nb} bands, {sysm.proj.nzlma} projectors")
    t0 = time.perf_counter()
    x = sysm.psi0()
    log(f"psi0 generated in {time.perf_counter() - t0:.1f}s")

    dev = RsdftGpu(sysm, precision=PREC_FP64)
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
    log(f"FP64: {t_fp64 * 1e3:.1f} ms/sweep  {tim64}")
    hp64 = dev.test_hpsi(args.hpsi_reps)
    log(f"FP64 hpsi ms: {hp64}")
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
    log(f"mixed: {t_mix * 1e3:.1f} ms/sweep  {timm}")
    hpm = dev.test_hpsi(args.hpsi_reps)
    log(f"mixed hpsi ms: {hpm}")
    dev.detach()

    # ---- accuracy: NSWEEP FP64 sweeps against (NSWEEP - 2) mixed + 2 FP64 ----
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

    esp64 = run(x0, 0, NSWEEP)
    espm = run(x0, NSWEEP - 2, 2)
    # sensitivity floor: FP64 from the start rounded once to FP32
    espp = run(x0.astype(np.float32).astype(np.float64), 0, NSWEEP)
    occ = sysm.occ

    def errs(e):
        d = np.abs(e - esp64)
        return float(np.max(d)), float(np.max(d[occ > 0])), float(abs(np.sum(occ * (e - esp64))))

    eig_err, eig_err_occ, eband_err = errs(espm)
    flo_err, flo_err_occ, flo_eband = errs(espp)
    d = np.abs(espm - esp64)
    worst = np.argsort(d)[::-1][:5]
    log(f"after {NSWEEP} sweeps ({NSWEEP - 2} mixed + 2 FP64 vs {NSWEEP} FP64): max eig err "
        f"{eig_err:.3e} Ha (occupied {eig_err_occ:.3e}), band energy {eband_err:.3e}; worst bands "
        f"{[int(i) for i in worst]} errs {[f'{d[i]:.2e}' for i in worst]}")
    log(f"FP64 from the FP32-rounded start: max {flo_err:.3e} Ha (occupied {flo_err_occ:.3e}), "
        f"band energy {flo_eband:.3e}")

    # ---- roofline of the dominant kernel (hpsi stencil, FP32) ---------------
    t_kernel_ms = float(hpm[4])
    B = hpsi_stencil_bytes(sysm, nb, 4, True)
    achieved = B / (t_kernel_ms * 1e-3) / 1e9
    B64 = hpsi_stencil_bytes(sysm, nb, 8, True)
    ach64 = B64 / (float(hp64[4]) * 1e-3) / 1e9
    copy_gbs = 2 * sysm.grid.ML * nb * 4 / (float(hpm[7]) * 1e-3) / 1e9
    fill_gbs = nl_fill_bytes(sysm, nb) / (float(hpm[6]) * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        # counters of the kernel timed here (the accumulating FP32 stencil) only
        if ", 11," in tr["FETCH_SIZE"]["kernel"]:
            traffic = float(tr["traffic_bytes"])

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sysm, x0, args.cpu_sample)
        log(f"cpu baseline {cpu['value']:.2f} s/sweep")

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
            "mixed_gemm": args.gemm,
            "fp64_s_per_sweep": t_fp64,
            "fp64_over_mixed": t_fp64 / t_mix,
            "accuracy_protocol": f"{NSWEEP} FP64 sweeps vs {NSWEEP - 2} mixed + 2 FP64 sweeps from the same start",
            "max_eig_err_ha": eig_err,
            "max_eig_err_occupied_ha": eig_err_occ,
            "band_energy_err_ha": eband_err,
            "fp64_rounded_start_eig_err_ha": flo_err,
            "fp64_rounded_start_eig_err_occupied_ha": flo_err_occ,
            "hpsi_ms_fp32": float(hpm[0]),
            "hpsi_ms_fp64": float(hp64[0]),
            "hpsi_stencil_gbs_fp64": ach64,
            "nl_fill_gbs_fp32": fill_gbs,
            "hpsi_parts_ms_fp32": {k: float(v) for k, v in zip(
                ("hamiltonian", "kinetic", "localpot", "nonlocal", "stencil", "projection",
                 "nl_stage", "copy"), hpm)},
            "copy_gbs_fp32": copy_gbs,
            "phase_ms_mixed": {k: v / K for k, v in timm.items() if k not in ("hpsi_bands", "eig_refined")},
            "eig_refined_mixed": timm["eig_refined"],
            "phase_ms_fp64": {k: v / K for k, v in tim64.items() if k not in ("hpsi_bands", "eig_refined")},
            "eig_refined_fp64": tim64["eig_refined"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_hpsi2<float,6,2,1,kin|loc|acc,1024> (Md=6 FD stencil + Vloc added to "
                      "the nonlocal term in H psi, all bands)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes": B,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
