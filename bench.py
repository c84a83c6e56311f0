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
INT8_PEAK_TOPS = 5000.0       # MI355X dense int8 matrix
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


PRIMES = (251, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191)


def oz_bits(nmod, K):
    """Operand bits of the split path (the library's rule): the largest b with
    6 K 2^(2b) below half the product of the moduli."""
    import math
    lm = sum(math.log2(p) for p in PRIMES[:nmod])
    return int(min(30, math.floor((lm - 1.0 - math.log2(6.0 * K)) * 0.5)))


def gen_eigvals(H, O, n, nst):
    """Lowest nst eigenvalues of H x = e O x from the column-major device
    buffers (the tensors hold H^T): LAPACK ZHEGVX semantics ('I', upper
    triangles) on the host, as solvewithlapack does (modfvsystem.f90:691)."""
    import scipy.linalg as sl
    Hm = H[:n * n].view(n, n).transpose(0, 1).cpu().numpy()
    Om = O[:n * n].view(n, n).transpose(0, 1).cpu().numpy()
    return sl.eigh(Hm, Om, lower=False, eigvals_only=True, subset_by_index=[0, nst - 1],
                   driver="gvx"), "host LAPACK"


def accuracy_check(ctx, d, apwalm, one_setup, nmod, nst):
    """Split vs FP64 on one k-point: relative Frobenius errors of H and O and
    the largest difference among the lowest nst generalized eigenvalues."""
    import torch
    from exciting_amd import lapw
    n = d.rank
    H64 = torch.empty(n * n, dtype=torch.complex128, device="cuda")
    O64 = torch.empty_like(H64)
    Hs = torch.empty_like(H64)
    Os = torch.empty_like(H64)
    one_setup(d, 0, H64, O64)
    one_setup(d, lapw.SPLIT(nmod), Hs, Os)
    torch.cuda.synchronize()
    rh = float(torch.linalg.norm(Hs - H64) / torch.linalg.norm(H64))
    ro = float(torch.linalg.norm(Os - O64) / torch.linalg.norm(O64))
    t0 = time.perf_counter()
    w64, where = gen_eigvals(H64, O64, n, nst)
    ws, _ = gen_eigvals(Hs, Os, n, nst)
    de = float(np.abs(ws - w64).max())
    return {"kpoint": 0, "n": n, "nstates": nst, "rel_frob_H": rh, "rel_frob_O": ro,
            "max_abs_dE_Ha": de, "within_1uHa": bool(de <= 1e-6), "solver": where,
            "eig_seconds": time.perf_counter() - t0}


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
    ap.add_argument("--moduli", type=int, default=9,
                    help="split precision: number of int8 moduli (0 = FP64 only)")
    ap.add_argument("--nstates", type=int, default=400,
                    help="lowest eigenvalues compared between split and FP64 (one k-point)")
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

    def one_setup(d, flags, H=None, O=None):
        ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, apwalm)
        Hb = Hbuf if H is None else H
        Ob = Obuf if O is None else O
        sysd = lapw.EvSystem(lapw.HermitianMatrix(d.rank, Hb[:d.rank * d.rank]),
                             lapw.HermitianMatrix(d.rank, Ob[:d.rank * d.rank]))
        ctx.hamiltonandoverlapsetup(sysd, d.ngk, apwalm, d.igpig, d.vgpc, flags=flags)

    def run(flags, label):
        for w in range(args.warmup):
            for d in dks:
                one_setup(d, flags)
            torch.cuda.synchronize()
            log(f"{label}: warmup step {w + 1}/{args.warmup} done")
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
                one_setup(d, flags)
                b.record(stream)
                evs.append((a, b))
            torch.cuda.synchronize()
            lat += [a.elapsed_time(b) for a, b in evs]
            log(f"{label}: step {step + 1}/{args.steps} done")
        t1 = time.perf_counter()
        prof = ctx.profile_read()
        ctx.profile(False)
        return {"elapsed": t1 - t0, "lat": lat, "prof": prof}

    M = s.maxaa
    nlo_atom = int(s.losize.max()) if s.losize.size else 0
    f_setup = [flops(s.natmtot, M, g, nlo_atom) for g in ngks]
    f_ho = [flops_ho(s.natmtot, M, g) for g in ngks]

    def summary(r):
        el = r["elapsed"]
        return {"value": args.steps * nk / el, "ms_per_step": 1e3 * el / args.steps,
                "gflops": args.steps * sum(f_setup) / el / 1e9,
                "frac_fp64_peak": args.steps * sum(f_setup) / el / 1e12 / FP64_PEAK_TFLOPS,
                "latency_ms": {"p50": float(np.percentile(r["lat"], 50)),
                               "p99": float(np.percentile(r["lat"], 99))},
                "kernels_ms_per_setup": {k: (v[0] / v[1] if v[1] else 0.0)
                                         for k, v in r["prof"].items() if v[1]}}

    r64 = run(0, "fp64")
    fp64 = summary(r64)
    ho_ms, ho_n = r64["prof"]["ho"]
    achieved_ho = (args.steps * sum(f_ho) / ho_n) / (ho_ms / ho_n * 1e-3) / 1e12
    fp64["roofline"] = {"kernel": "ho_kernel (fused upper-triangle H/O contraction + interstitial)",
                        "bound": "mfma", "achieved": achieved_ho, "peak": FP64_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": achieved_ho / FP64_PEAK_TFLOPS,
                        "traffic": measured_traffic(args.config)}

    head = fp64
    split = None
    nmod = args.moduli
    if nmod:
        rsp = run(lapw.SPLIT(nmod), f"split({nmod})")
        split = summary(rsp)
        K = s.natmtot * M
        bits = oz_bits(nmod, K)
        gm_ms, gm_n = rsp["prof"]["oz_gemm"]
        i8_ops = args.steps * sum(6 * nmod * g * (g + 1) / 2 * K * 2.0 for g in ngks)
        achieved_i8 = (i8_ops / gm_n) / (gm_ms / gm_n * 1e-3) / 1e12
        split["roofline"] = {"kernel": f"ozgemm_kernel (int8 MFMA modular GEMMs, {nmod} moduli x "
                                       f"6 products, upper triangle)",
                             "bound": "mfma", "achieved": achieved_i8, "peak": INT8_PEAK_TOPS,
                             "unit": "TOP/s", "frac": achieved_i8 / INT8_PEAK_TOPS,
                             "traffic": None}
        split["moduli"] = nmod
        split["bits"] = bits
        split["accuracy"] = accuracy_check(ctx, dks[0], apwalm, one_setup, nmod, args.nstates)
        log(f"split accuracy {split['accuracy']}")
        head = split

    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "setups/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"i8x{nmod}+f64" if nmod else "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{spec.name}: {s.natmtot}-atom cell, lmaxapw={s.lmaxapw}, APW order "
                         f"{s.apwordmax} + {nlo_atom} LO functions per atom, rgkmax="
                         f"{spec.rgkmax}, {nk} k-points, match + H/O setup"),
            "kpoints": nk,
            "ngk_mean": float(np.mean(ngks)),
            "n_mean": float(np.mean(ngks)) + s.nlotot,
            "natoms": s.natmtot,
            "precision": (f"split: muffin-tin products by int8 Ozaki scheme II with {nmod} "
                          f"moduli (operands rounded to {oz_bits(nmod, s.natmtot * M)} bits per "
                          f"G column); match, LO blocks, interstitial FP64") if nmod else "fp64",
        },
        "gflops": head["gflops"],
        "latency_ms": head["latency_ms"],
        "kernels_ms_per_setup": head["kernels_ms_per_setup"],
        "roofline": head["roofline"],
        "fp64": fp64,
    }
    if split:
        out["speedup_vs_fp64"] = split["value"] / fp64["value"]
        out["accuracy"] = split["accuracy"]
    if args.cpu_baseline == "on":
        log("cpu baseline: oracle on one k-point")
        out["cpu_baseline"] = cpu_baseline(s, 0)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
