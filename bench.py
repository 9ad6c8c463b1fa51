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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def hpsi_stencil_bytes(sysm, nb, s):
    """Algorithmic HBM bytes of one all-band launch of the stencil kernel
    (kinetic + local): psi read + H psi write per band and Vloc once."""
    ML = sysm.grid.ML
    return 2 * ML * nb * s + ML * s


def cpu_baseline(sysm, x0, nsample):
    """Oracle FP64 sweep on the host: CG on `nsample` bands (scaled to all
    bands) plus full-size Gram-Schmidt and subspace diagonalisation."""
    from oracle.oracle import Oracle
    o = Oracle(sysm)
    nb = x0.shape[0]
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
        "sample": (f"oracle FP64 CG on {nsample} of {nb} bands scaled x{nb / nsample:g}, plus "
                   f"full {nb}-band Gram-Schmidt and subspace diag; measured "
                   f"{t_cg + t_gs + t_sd:.1f}s of CPU work"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=9)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="si216")
    ap.add_argument("--bands", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=16)
    ap.add_argument("--hpsi-reps", type=int, default=20)
    ap.add_argument("--gemm", choices=["fp32", "bf16"], default="fp32",
                    help="GEMM precision of the mixed path")
    args = ap.parse_args()
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
    _, _ = dev.sweep(x, K)[1:]
    t_fp64 = (time.perf_counter() - t0) / K
    tim64 = dev.timers()
    log(f"FP64: {t_fp64 * 1e3:.1f} ms/sweep  {tim64}")
    _, esp64, _ = dev.sweep(x, 2)
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
    dev.sync()
    dev.detach()
    # two FP64 refinement sweeps
    dev.set_precision(PREC_FP64)
    dev.attach(xm)
    _, espm, _ = dev.sweep(xm, 2)
    dev.detach()
    occ = sysm.occ
    err = np.abs(espm - esp64)
    eig_err = float(np.max(err))
    eig_err_occ = float(np.max(err[occ > 0]))
    worst = np.argsort(err)[::-1][:5]
    eband_err = float(abs(np.sum(occ * espm) - np.sum(occ * esp64)))
    log(f"eigenvalue error after 2 FP64 refinement sweeps: {eig_err:.3e} Ha (occupied {eig_err_occ:.3e}), "
        f"band energy {eband_err:.3e}; worst bands {list(worst)} errs {[f'{err[i]:.2e}' for i in worst]}")

    # ---- roofline of the dominant kernel (fused hpsi stencil, FP32) ---------
    t_kernel_ms = float(hpm[4])
    B = hpsi_stencil_bytes(sysm, nb, 4)
    achieved = B / (t_kernel_ms * 1e-3) / 1e9
    B64 = hpsi_stencil_bytes(sysm, nb, 8)
    ach64 = B64 / (float(hp64[4]) * 1e-3) / 1e9

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
            "max_eig_err_ha": eig_err,
            "max_eig_err_occupied_ha": eig_err_occ,
            "band_energy_err_ha": eband_err,
            "hpsi_ms_fp32": float(hpm[0]),
            "hpsi_ms_fp64": float(hp64[0]),
            "hpsi_stencil_gbs_fp64": ach64,
            "phase_ms_mixed": {k: v / K for k, v in timm.items() if k not in ("hpsi_bands", "eig_refined")},
            "eig_refined_mixed": timm["eig_refined"],
            "phase_ms_fp64": {k: v / K for k, v in tim64.items() if k not in ("hpsi_bands", "eig_refined")},
            "eig_refined_fp64": tim64["eig_refined"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_hpsi<float,6,4,2> (Md=6 FD stencil + Vloc, all 480 bands)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
