#!/usr/bin/env python3
"""H+O setup throughput on one MI355X.

A step is one pass of the hot path (match + Hamiltonian/overlap setup) over
every k-point of the workload, with all per-k inputs already resident in HBM.
The default workload is the 64-atom simple-cubic supercell (a = 20.52 bohr,
lmaxapw = 10, APW order 2 + one local orbital per l <= 3, rgkmax = 8, so
ngk ~ 9340 and n ~ 10364) with 16 k-points, computed in FP64.

Prints one JSON line (see the keys below).  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("H+O setups/s at ngk≈9k, split-prec vs FP64 GPU at 1 µHa; "
          "p50/p99 latency")
FP64_PEAK_TFLOPS = 78.6       # MI355X dense FP64 matrix
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def flops(natm, M, G, nlo):
    """Algorithmic flops of one setup (upper triangles of H and O):
    natm * [8 M^2 G (aa.A) + 8 G(G+1) M (H and O) + 8 nlo M G (APW-lo)]."""
    return natm * (8.0 * M * M * G + 8.0 * G * (G + 1) * M + 8.0 * nlo * M * G)


def flops_ho(natm, M, G):
    """Algorithmic flops of the fused H/O contraction kernel for one setup."""
    return natm * 8.0 * G * (G + 1) * M


PROFILE = os.path.join(ROOT, "profiles", "ho_kernel_sc64_fp64.md")


def measured_traffic(config_index: int):
    """HBM-side bytes per ho_kernel launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by a calibration copy).
    Only the 64-atom workload was counted; other workloads report null."""
    if config_index != 3 or not os.path.exists(PROFILE):
        return None
    import re
    m = re.search(r"\|\s*traffic_bytes_per_launch\s*\|\s*([0-9.eE+-]+)", open(PROFILE).read())
    return float(m.group(1)) if m else None


def cpu_baseline(s, ik: int):
    """The oracle (C restatement + SciPy's OpenBLAS zgemm) on one k-point."""
    import oracle
    S = oracle.System(s)
    threads = None
    try:
        from threadpoolctl import threadpool_info
        infos = [i for i in threadpool_info() if i.get("internal_api") == "openblas"]
        if infos:
            threads = max(i.get("num_threads", 1) for i in infos)
    except Exception:
        pass
    t0 = time.perf_counter()
    A = oracle.match(S, ik)
    t1 = time.perf_counter()
    oracle.setup(S, ik, A)
    t2 = time.perf_counter()
    k = s.kpts[ik]
    return {
        "value": 1.0 / (t2 - t0),
        "unit": "setups/s",
        "cores": threads or os.cpu_count(),
        "kind": "port",
        "sample": (f"k-point {ik} of the workload (ngk={k.ngk}, n={k.ngk + s.nlotot}): "
                   f"match {t1 - t0:.2f} s + H/O setup {t2 - t1:.2f} s; oracle C restatement "
                   f"with SciPy OpenBLAS zgemm"),
        "cpu": platform.processor() or platform.machine(),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3, help="workload index (0..3)")
    ap.add_argument("--kpoints", type=int, default=0, help="use the first N k-points (0 = all)")
    ap.add_argument("--cpu-baseline", choices=("on", "off"), default="on")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the timed region, run one device-to-device copy of this "
                         "many MiB (a known-size kernel for calibrating HBM byte counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(k-points would shard across GPUs with no exchange; see DESIGN.md). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    from exciting_amd import lapw, synth

    spec = synth.config(args.config)
    ks = list(range(len(spec.kpoints)))
    if args.kpoints:
        ks = ks[:args.kpoints]
    log(f"building workload {spec.name} with {len(ks)} k-points")
    s = synth.build(spec, kpoints=ks)
    nk = len(s.kpts)
    ngks = [k.ngk for k in s.kpts]
    log(f"ngk min/mean/max {min(ngks)}/{np.mean(ngks):.0f}/{max(ngks)} nlotot {s.nlotot}")

    ctx = lapw.Context(0).load(s)
    dks = [lapw.DeviceKPoint(s, ik) for ik in range(nk)]
    # one apwalm and one H/O pair reused by every k-point of a step
    apwalm = dks[0].apwalm
    for d in dks[1:]:
        d.apwalm = apwalm
    nmax = max(d.rank for d in dks)
    Hbuf = torch.empty(nmax * nmax, dtype=torch.complex128, device="cuda")
    Obuf = torch.empty(nmax * nmax, dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream)
    if args.copy_calibration_mib:
        n = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(f"calibration copy of {args.copy_calibration_mib} MiB done")

    def one_setup(d):
        ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, apwalm)
        sysd = lapw.EvSystem(lapw.HermitianMatrix(d.rank, Hbuf[:d.rank * d.rank]),
                             lapw.HermitianMatrix(d.rank, Obuf[:d.rank * d.rank]))
        ctx.hamiltonandoverlapsetup(sysd, d.ngk, apwalm, d.igpig, d.vgpc, flags=0)

    for w in range(args.warmup):
        for d in dks:
            one_setup(d)
        torch.cuda.synchronize()
        log(f"warmup step {w + 1}/{args.warmup} done")

    ctx.profile(True)
    ctx.profile_reset()
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for step in range(args.steps):
        evs = []
        for d in dks:
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            one_setup(d)
            b.record(stream)
            evs.append((a, b))
        torch.cuda.synchronize()
        lat += [a.elapsed_time(b) for a, b in evs]
        log(f"step {step + 1}/{args.steps} done")
    t1 = time.perf_counter()
    prof = ctx.profile_read()
    ctx.profile(False)

    elapsed = t1 - t0
    nsetups = args.steps * nk
    M = s.maxaa
    nlo_atom = int(s.losize.max()) if s.losize.size else 0
    f_total = args.steps * sum(flops(s.natmtot, M, g, nlo_atom) for g in ngks)
    f_ho = args.steps * sum(flops_ho(s.natmtot, M, g) for g in ngks)
    ho_ms, ho_n = prof["ho"]
    achieved = (f_ho / ho_n) / (ho_ms / ho_n * 1e-3) / 1e12 if ho_n else 0.0

    out = {
        "metric": METRIC,
        "value": nsetups / elapsed,
        "unit": "setups/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{spec.name}: {s.natmtot}-atom cell, lmaxapw={s.lmaxapw}, APW order "
                         f"{s.apwordmax} + {nlo_atom} LO functions per atom, rgkmax="
                         f"{spec.rgkmax}, {nk} k-points, FP64 match + H/O setup"),
            "kpoints": nk,
            "ngk_mean": float(np.mean(ngks)),
            "n_mean": float(np.mean(ngks)) + s.nlotot,
            "natoms": s.natmtot,
            "precision": "fp64",
        },
        "gflops": f_total / elapsed / 1e9,
        "frac_fp64_peak": f_total / elapsed / 1e12 / FP64_PEAK_TFLOPS,
        "latency_ms": {"p50": float(np.percentile(lat, 50)),
                       "p99": float(np.percentile(lat, 99))},
        "kernels_ms_per_setup": {k: (v[0] / v[1] if v[1] else 0.0) for k, v in prof.items()},
        "roofline": {
            "kernel": "ho_kernel (fused upper-triangle H/O contraction + interstitial)",
            "bound": "mfma",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic(args.config),
        },
    }
    if args.cpu_baseline == "on":
        log("cpu baseline: oracle on one k-point")
        out["cpu_baseline"] = cpu_baseline(s, 0)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
