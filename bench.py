#!/usr/bin/env python3
"""H+O setup throughput on one MI355X.

A step is one pass of the hot path (match + Hamiltonian/overlap setup) over
every k-point of the workload, with all per-k inputs already resident in HBM.
The headline workload is the 64-atom simple-cubic supercell (a = 20.52 bohr,
lmaxapw = 10, APW order 2 + one local orbital per l <= 3, rgkmax = 8, so
ngk ~ 9340 and n ~ 10364) with 16 k-points.  It is run three ways:

  * fp64_loop: the single-k entry points called once per k-point (the
    reference's calling pattern);
  * fp64: one batched call over the 16 k-points (grouped launches, the
    persistent H/O kernel with the plan picked by the autotuner);
  * split: the same batch with the muffin-tin products in int8 Ozaki
    scheme II (the headline value).

Then, unless switched off: the 16-atom orthorhombic cell with 128 uneven
k-points as one batch, a bounded sample of the serving loop (Poisson
requests over the 2..64-atom cells, setup + eigensolve, p50/p99 latency),
and the oracle on one k-point of the headline workload as the CPU baseline.

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


def measured_traffic(profile: str, per: str = "kpoint"):
    """HBM-side bytes of one kernel from a committed rocprofv3 --pmc summary
    under profiles/ (FETCH_SIZE + WRITE_SIZE, each scaled by a calibration
    copy): the value of its ``traffic_bytes_per_<per>`` row, or None."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    import re
    m = re.search(r"\|\s*traffic_bytes_per_" + per + r"\s*\|\s*([0-9.eE+-]+)", open(path).read())
    return float(m.group(1)) if m else None


PRIMES = (251, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191)


def oz_bits(nmod, K):
    """Operand bits of the split path (the library's rule): every G column is
    scaled so that its 2-norm is below 2^c, c the largest value with 2^(2c)
    below half the product of the moduli (K does not enter)."""
    import math
    lm = sum(math.log2(p) for p in PRIMES[:nmod])
    return int(min(50, math.floor((lm - 1.0) * 0.5 - 1e-9)))


def pick_moduli(W, arg, nst):
    """Modulus count from --moduli: an integer, or 'auto' (the precision
    tuner on k-point 0, untimed).  Returns (nmod, tuning record or None)."""
    if str(arg) != "auto":
        return int(arg), None
    from exciting_amd import precision
    t0 = time.perf_counter()
    res = precision.tune_moduli(W.ctx, W.kp[0], W.dks[0].rank, W.H[0], W.O[0], nst or 400)
    res["seconds"] = time.perf_counter() - t0
    log(f"moduli tuner: {res}")
    return (res["nmod"] or 12), res


def accuracy(ctx, kp, dk, H, O, nmod, nst, kpoint=0):
    """Split vs FP64 on one k-point of a batch: relative Frobenius errors of H
    and O and the largest difference among the lowest nst generalized
    eigenvalues, both solved in FP64 on the GPU (Cholesky reduction +
    Hermitian eigensolve, the ZHEGVX problem of solvewithlapack)."""
    import torch
    from exciting_amd import lapw, solve
    n = dk.rank
    ctx.setup_batch([kp], flags=lapw.MATCH)
    torch.cuda.synchronize()
    H64, O64 = H.clone(), O.clone()
    ctx.setup_batch([kp], flags=lapw.MATCH | lapw.SPLIT(nmod))
    torch.cuda.synchronize()
    rh = float(torch.linalg.norm(H - H64) / torch.linalg.norm(H64))
    ro = float(torch.linalg.norm(O - O64) / torch.linalg.norm(O64))
    t0 = time.perf_counter()
    w64 = solve.geneig(H64, O64, n, nst).cpu().numpy()
    ws = solve.geneig(H, O, n, nst).cpu().numpy()
    de = float(np.abs(ws - w64).max())
    eig_s = time.perf_counter() - t0
    vec = eigvec_timing(H64, O64, n, sorted({40, nst}))
    del H64, O64
    return {"kpoint": kpoint, "n": n, "nstates": nst, "rel_frob_H": rh, "rel_frob_O": ro,
            "max_abs_dE_Ha": de, "within_1uHa": bool(de <= 1e-6),
            "solver": "GPU FP64: Cholesky of O, two triangular solves, Hermitian eigensolve",
            "eig_seconds": eig_s, "eigenvectors": vec}


def eigvec_timing(H, O, n, nsts):
    """solvewithlapack (eigenvalues and eigenvectors) on one FP64 H/O pair:
    seconds of the full Hermitian eigensolve and of the Chebyshev-filtered
    subspace iteration, and how far apart their eigenvalues are."""
    import torch
    from exciting_amd import solve
    # one untimed subspace solve first: the process's first one also pays
    # one-time costs (kernel and library initialisation, allocator growth)
    solve.solvewithlapack(H, O, n, nsts[0], method="subspace")
    torch.cuda.synchronize()
    out = []
    for nv in nsts:
        row = {"nstates": nv}
        w = {}
        for method in ("full", "subspace"):
            torch.cuda.synchronize()
            t = time.perf_counter()
            w[method], X = solve.solvewithlapack(H, O, n, nv, method=method)
            torch.cuda.synchronize()
            row[f"{method}_s"] = round(time.perf_counter() - t, 4)
            del X
        row["subspace_steps"] = solve.last_subspace.get("steps")
        row["subspace_fallback"] = solve.last_subspace.get("fallback")
        row["max_abs_dE_Ha"] = float((w["full"] - w["subspace"]).abs().max())
        out.append(row)
    log(f"eigenvectors: {out}")
    return out


def host_pointer_sample(ctx, s, ik: int = 0, reps: int = 2):
    """The Fortran drop-in's rate: exh_match, then
    exh_hamiltonandoverlapsetup on pageable host arrays, as the k loop calls
    match and hamiltonandoverlapsetup (seceqn.f90:57-58, seceqnfv.f90:76),
    with H and O accumulated (uploaded and downloaded) and apwalm downloaded
    by match; once with every array crossing PCIe and once with apwalm reuse
    on (the setup reads the copy match left in HBM).  Median seconds per
    k-point over reps calls after one untimed call."""
    from exciting_amd import lapw
    k = s.kpts[ik]
    lmm = (s.lmaxapw + 1) ** 2
    n = k.ngk + s.nlotot
    apwalm = np.zeros((s.ngkmax, s.apwordmax, lmm, s.natmtot), np.complex128, order="F")
    H = np.zeros((n, n), np.complex128, order="F")
    O = np.zeros((n, n), np.complex128, order="F")
    system = lapw.EvSystem(lapw.HermitianMatrix(n, H), lapw.HermitianMatrix(n, O))
    apw_bytes = 16 * k.ngk * s.apwordmax * lmm * s.natmtot
    out = {"kpoint": ik, "n": n, "arrays": "NumPy (pageable), Fortran order"}
    for reuse in (False, True):
        ctx.set_apwalm_reuse(reuse)
        u0, r0 = ctx.apwalm_transfers()
        ts = []
        for _ in range(reps + 1):
            H.fill(0)
            O.fill(0)
            t = time.perf_counter()
            ctx.match(k.ngk, k.gpc, k.tpgpc, k.sfacgp, apwalm)
            ctx.hamiltonandoverlapsetup(system, k.ngk, apwalm, k.igpig, k.vgpc)
            ts.append(time.perf_counter() - t)
        u1, r1 = ctx.apwalm_transfers()
        sec = float(np.median(ts[1:]))
        pcie = apw_bytes * (1 if reuse else 2) + 4 * 16 * n * n
        out["reuse" if reuse else "no_reuse"] = {
            "s_per_kpoint": round(sec, 4), "setups_per_s": round(1.0 / sec, 3),
            "pcie_bytes": pcie, "pcie_GBps": round(pcie / sec / 1e9, 1),
            "apwalm_uploads": u1 - u0, "apwalm_reused": r1 - r0}
    ctx.set_apwalm_reuse(False)
    log(f"host-pointer drop-in: {out}")
    return out


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


class Workload:
    """One workload resident on the GPU: library context with the cell's
    tables, per-k inputs, and per-k apwalm, H and O (a batch writes every
    k-point's own matrices)."""

    def __init__(self, index, nkpts=0, workspace_gib=0.0):
        import torch
        from exciting_amd import lapw, synth
        self.spec = synth.config(index)
        ks = list(range(len(self.spec.kpoints)))
        if nkpts:
            ks = ks[:nkpts]
        log(f"building workload {self.spec.name} with {len(ks)} k-points")
        self.s = synth.build(self.spec, kpoints=ks)
        self.nk = len(self.s.kpts)
        self.ngks = [k.ngk for k in self.s.kpts]
        log(f"  ngk min/mean/max {min(self.ngks)}/{np.mean(self.ngks):.0f}/{max(self.ngks)} "
            f"nlotot {self.s.nlotot}")
        self.ctx = lapw.Context(0).load(self.s)
        if workspace_gib:
            self.ctx.set_batch_workspace(int(workspace_gib * 2**30))
        self.dks = [lapw.DeviceKPoint(self.s, ik) for ik in range(self.nk)]
        self.H = [torch.zeros(d.rank * d.rank, dtype=torch.complex128, device="cuda")
                  for d in self.dks]
        self.O = [torch.zeros_like(h) for h in self.H]
        self.kp = [lapw.kpoint_struct(d, h, o) for d, h, o in zip(self.dks, self.H, self.O)]
        torch.cuda.synchronize()
        self.stream = torch.cuda.ExternalStream(self.ctx.stream)
        M = self.s.maxaa
        self.nlo_atom = int(self.s.losize.max()) if self.s.losize.size else 0
        self.f_setup = [flops(self.s.natmtot, M, g, self.nlo_atom) for g in self.ngks]
        self.f_ho = [flops_ho(self.s.natmtot, M, g) for g in self.ngks]
        self.K = self.s.natmtot * M

    def free(self):
        import torch
        self.ctx.close()
        self.H = self.O = self.kp = self.dks = None
        torch.cuda.empty_cache()

    def one_setup(self, ik, flags):
        from exciting_amd import lapw
        d = self.dks[ik]
        self.ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, d.apwalm)
        sysd = lapw.EvSystem(lapw.HermitianMatrix(d.rank, self.H[ik]),
                             lapw.HermitianMatrix(d.rank, self.O[ik]))
        self.ctx.hamiltonandoverlapsetup(sysd, d.ngk, d.apwalm, d.igpig, d.vgpc, flags=flags)

    def run(self, flags, label, steps, warmup, batch=True, profile=True):
        """Timed steps; per step either one batched call over every k-point or
        the single-k calls in a loop.  Returns host wall time, per-setup
        latencies (ms, HIP events on the library's stream) and the per-kernel
        device times (HIP events around every launch) when profiled."""
        import torch
        from exciting_amd import lapw

        def step(lat):
            if batch:
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record(self.stream)
                self.ctx.setup_batch(self.kp, flags=flags | lapw.MATCH)
                b.record(self.stream)
                torch.cuda.synchronize()
                if lat is not None:   # every setup of a batch completes with it
                    lat += [a.elapsed_time(b)] * self.nk
                return
            evs = []
            for ik in range(self.nk):
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record(self.stream)
                self.one_setup(ik, flags)
                b.record(self.stream)
                evs.append((a, b))
            torch.cuda.synchronize()
            if lat is not None:
                lat += [a.elapsed_time(b) for a, b in evs]

        for w in range(warmup):
            step(None)
            log(f"{label}: warmup step {w + 1}/{warmup} done")
        if profile:
            self.ctx.profile(True)
            self.ctx.profile_reset()
        lat = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            step(lat)
            log(f"{label}: step {i + 1}/{steps} done")
        t1 = time.perf_counter()
        prof = self.ctx.profile_read() if profile else {}
        if profile:
            self.ctx.profile(False)
        el = t1 - t0
        f = steps * sum(self.f_setup)
        return {"value": steps * self.nk / el, "ms_per_step": 1e3 * el / steps,
                "gflops": f / el / 1e9, "frac_fp64_peak": f / el / 1e12 / FP64_PEAK_TFLOPS,
                "latency_ms": {"p50": float(np.percentile(lat, 50)),
                               "p99": float(np.percentile(lat, 99))},
                "kernels_ms_per_setup": {k: v[0] / (steps * self.nk)
                                         for k, v in prof.items() if v[1]},
                "_prof": prof, "_steps": steps}

    def ho_roofline(self, r, kernel, profile_md, per="kpoint"):
        ms, n = r["_prof"][kernel]
        achieved = r["_steps"] * sum(self.f_ho) / (ms * 1e-3) / 1e12
        kpl = r["_steps"] * self.nk / n           # k-points per launch
        per_k = measured_traffic(profile_md, per) if profile_md else None
        return {"kernel": f"{kernel} (fused upper-triangle H/O contraction + interstitial)",
                "launches": n, "kpoints_per_launch": kpl, "avg_launch_ms": ms / n,
                "bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                "traffic": per_k * kpl if per_k else None}

    def oz_roofline(self, r, nmod, profile_md):
        ms, n = r["_prof"]["oz_gemm"]
        K = self.K
        ops = r["_steps"] * sum(6 * nmod * g * (g + 1) / 2 * K * 2.0 for g in self.ngks)
        achieved = ops / (ms * 1e-3) / 1e12
        # the FP64 work these GEMMs stand in for: the H/O contraction's flops
        fp64_eq = r["_steps"] * sum(self.f_ho) / (ms * 1e-3) / 1e12
        return {"kernel": f"ozgemm_kernel (int8 MFMA modular GEMMs, {nmod} moduli x 6 "
                          f"products, upper triangle)",
                "launches": n, "avg_launch_ms": ms / n,
                "bound": "mfma", "achieved": achieved, "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                "frac": achieved / INT8_PEAK_TOPS, "traffic": measured_traffic(profile_md),
                "fp64_equivalent_tflops": fp64_eq,
                "fp64_equivalent_vs_fp64_peak": fp64_eq / FP64_PEAK_TFLOPS}


def clean(r):
    return {k: v for k, v in r.items() if not k.startswith("_")}


def serving_sample(args):
    """Bounded sample of the serving loop: Poisson requests over the 2, 8,
    16, 32 and 64-atom cells (1..kmax k-points each, FP64 or split), each
    match + setup + GPU eigensolve of the lowest states per k-point.  The
    setups of the requests in flight run as grouped launches over several
    cells; the admission budget is capped so that requests wait for HBM,
    later ones bypass them and idle contexts are evicted."""
    from exciting_amd import serve, synth
    specs = synth.serving_cells(args.serve_kmax)
    cells = {name: synth.build(spec, kpoints=ks) for name, (spec, ks) in specs.items()}
    ex = serve.GpuExecutor(0, max_solves=args.serve_solves)
    make, close, execute = ex.make_context, ex.close_context, ex.execute
    from exciting_amd import solve as _solve
    budget = (serve.free_hbm(0) - (48 << 30)
              - args.serve_solves * _solve.WORKSPACE_BYTES)   # the solvers' hipBLAS workspaces
    if args.serve_budget_gib > 0:
        budget = min(budget, int(args.serve_budget_gib * 2**30))
    srv = serve.Server(cells, execute, make, close, budget, max_inflight=8,
                       on_finish=lambda r: log(f"serve: request {r.rid} {r.cell} "
                                               f"nk={len(r.kpoints)} nmod={r.nmod} {r.status} "
                                               f"{r.latency:.2f} s"))
    warm = [serve.Request(rid=-1 - i, cell=c, kpoints=[0], nmod=m, nstfv=args.serve_nstfv)
            for i, (c, m) in enumerate((c, m) for c in cells for m in (0, args.serve_moduli))]
    srv.run(warm)
    reqs = serve.poisson_requests(args.serve_requests, args.serve_rate, cells, seed=7,
                                  kmax=args.serve_kmax, nmod=args.serve_moduli,
                                  nstfv=args.serve_nstfv)
    warm_stats = ex.stats()
    t0 = time.perf_counter()
    srv.run(reqs)
    out = serve.summarize(reqs, time.perf_counter() - t0)
    st = ex.stats()
    launches = st["launches"] - warm_stats["launches"]
    kpts = st["kpoints"] - warm_stats["kpoints"]
    out.update({"sample": (f"{args.serve_requests} Poisson requests at {args.serve_rate}/s over "
                           f"fcc2/sc8/ortho16/sc32/sc64, 1..{args.serve_kmax} k-points each, "
                           f"half split ({args.serve_moduli} moduli), {args.serve_nstfv} states, "
                           f"one eigensolve at a time ({args.serve_solves} solver "
                           f"thread{'s' if args.serve_solves != 1 else ''}), "
                           f"admission budget "
                           f"{budget / 2**30:.0f} GiB"),
                "budget_gib": budget / 2**30,
                "peak_inflight": srv.peak_inflight, "contexts": srv.pool.created,
                "evictions": srv.pool.evicted, "bypassed": srv.bypassed,
                "hbm_waits": srv.hbm_waits, "peak_context_gib": srv.peak_bytes / 2**30,
                "setup_launches": launches,
                "kpoints_per_launch": kpts / max(launches, 1),
                "max_kpoints_per_launch": st["kpoints_per_launch"]["max"],
                "max_cells_per_launch": st["cells_per_launch"]["max"]})
    srv.pool.close_all()
    ex.close()
    return out


def interstitial_sample(ctx, s, reps=20):
    """The interstitial producers on the workload's FFT grid, inputs resident
    in HBM: gencfun (closed form over every G of the box, backward FFT, real
    part), genveffig alone ('none' relativity) and genveffig + genmeffig
    ('iora*': three products, one batched forward FFT, gathers), and the
    whole device chain gencfun -> genveffig -> set_interstitial_device.
    Per-kernel averages come from the library's launch events on the stream
    the kernels run on."""
    import torch
    n, ng = s.ngrtot, s.ngvec
    ngrid = tuple(int(x) for x in s.intgv[:, 1] - s.intgv[:, 0] + 1)
    cg = torch.zeros(n, dtype=torch.complex128, device="cuda")
    cr = torch.zeros(n, dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(1)
    vd = torch.from_numpy(-0.4 + 0.3 * rng.uniform(-1, 1, n)).cuda()
    outs = [torch.zeros(ng, dtype=torch.complex128, device="cuda") for _ in range(3)]

    def gencfun():
        ctx.gencfun(s.bvec, s.atposc, cfunig=cg, cfunir=cr)

    def veff():
        ctx.genveffig_device(vd, cr, veffig=outs[0])

    def veff_meff():
        ctx.genveffig_device(vd, cr, veffig=outs[0], meffig=outs[1], m2effig=outs[2])

    def chain():
        gencfun()
        veff()
        ctx.set_interstitial(cg, outs[0], valence_rel_none=True)

    def timed(fn):
        fn()                                    # plans and buffers
        torch.cuda.synchronize()
        ctx.profile(True)
        ctx.profile_reset()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / reps * 1e3
        prof = ctx.profile_read()
        ctx.profile(False)
        k = {name: ms / cnt for name, (ms, cnt) in prof.items() if cnt}
        return wall, k

    res = {}
    w, k = timed(gencfun)
    t = k["cfunig"] * 1e-3
    res["gencfun"] = {
        "ms": w, "kernels_ms": k,
        "cfunig_kernel": {"bound": "fp64 sincos throughput",
                          "pairs_per_s": n * s.natmtot / t,
                          "hbm_GBps": 44.0 * n / t / 1e9,
                          "bytes_per_G": 44},
    }
    w1, k1 = timed(veff)
    w3, k3 = timed(veff_meff)
    fft1 = k1["ifft"] * 1e-3
    fft3 = k3["ifft"] * 1e-3
    res["genveffig"] = {"ms": w1, "kernels_ms": k1}
    res["genveffig_genmeffig_iora"] = {"ms": w3, "kernels_ms": k3}
    res["fft"] = {"grid": ngrid, "one_pass_bytes_per_grid": 32 * n,
                  "single_GBps_one_pass": 32.0 * n / fft1 / 1e9,
                  "batched3_GBps_one_pass": 96.0 * n / fft3 / 1e9,
                  "frac_hbm_one_pass": 96.0 * n / fft3 / 8e12}
    wc, _ = timed(chain)
    res["device_chain_ms"] = wc
    torch.cuda.synchronize()
    # CPU: the oracle (C closed form with OpenMP, numpy FFTs) on the same grid
    import oracle
    t0 = time.perf_counter()
    _, cr0 = oracle.gencfun(s)
    t1 = time.perf_counter()
    v = vd.cpu().numpy()
    oracle.genveffig(s, v, cr0)
    oracle.genmeffig(s, v, cr0, True)
    t2 = time.perf_counter()
    res["cpu_baseline"] = {
        "value": 1e3 * (t2 - t0), "unit": "ms (gencfun + genveffig + genmeffig iora*)",
        "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())), "kind": "port",
        "sample": (f"the whole producer set once on the {ngrid} grid: oracle gencfunig "
                   f"(C, OpenMP) {1e3 * (t1 - t0):.0f} ms + numpy FFTs and gathers "
                   f"{1e3 * (t2 - t1):.0f} ms"),
    }
    res["gpu_ms_same_work"] = res["gencfun"]["ms"] + res["genveffig_genmeffig_iora"]["ms"]
    res["workload"] = (f"{s.natmtot} atoms, FFT grid {ngrid[0]}x{ngrid[1]}x{ngrid[2]} "
                       f"(ngrtot {n}), ngvec {ng}")
    return res


def hmlint_sample(ctx, s, lmaxvr=8, reps=10, cpu=True):
    """The muffin-tin table producer on the workload's cell, inputs resident
    in HBM: hmlint (radial integrals + Gaunt contraction into aa, loa, lolo)
    and exh_set_mt_device installing them.  Synthetic smooth potential with
    lmaxvr = 8 (the reference's default) and cubic LO radial functions."""
    import dataclasses
    import torch
    from exciting_amd import synth
    sp = synth.add_mt_potential(dataclasses.replace(s), lmaxvr=lmaxvr, seed=5)
    dev = lambda a: torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).cuda()
    v, lf = dev(sp.veffmt), dev(sp.lofr)
    mnlo = max(s.maxnlo, 1)
    outs = [torch.zeros(n, dtype=torch.complex128, device="cuda")
            for n in (s.maxaa * s.maxaa * s.natmtot, mnlo * s.maxaa * s.natmtot,
                      mnlo * mnlo * s.natmtot)]
    oalo, ololo = dev(s.oalo), dev(s.ololo)

    def run():
        ctx.hmlint(lmaxvr, v, lf, True, *outs)

    def run_install():
        run()
        ctx.set_mt(s.maxaa, s.maxnlo, *outs, oalo, ololo)

    # spin: the non-collinear alpha/beta/ab blocks (three mt_pot on B_xc
    # components) and the spin-orbit terms added to them
    bx = dev(np.concatenate([0.05 * sp.veffmt.ravel(order="F")] * 3))
    t4 = (s.maxaa * s.maxaa * s.natmtot, mnlo * s.maxaa * s.natmtot, mnlo * mnlo * s.natmtot,
          s.maxaa * mnlo * s.natmtot)
    so_tables = {k: tuple(torch.zeros(n, dtype=torch.complex128, device="cuda") for n in t4)
                 for k in ("alpha", "beta", "ab")}

    def run_spin():
        ctx.hmlint_spin(lmaxvr, bx, lf, ncmag=True)

    def run_so():
        ctx.mt_so(lmaxvr, v, lf, so_tables)

    res = {}
    for label, fn in (("hmlint", run), ("hmlint_and_set_mt_device", run_install),
                      ("spin_blocks_noncollinear", run_spin), ("mt_so", run_so)):
        fn()
        torch.cuda.synchronize()
        ctx.profile(True)
        ctx.profile_reset()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / reps * 1e3
        prof = ctx.profile_read()
        ctx.profile(False)
        res[label] = {"ms": wall, "kernels_ms": {k: ms / n for k, (ms, n) in prof.items() if n}}
    # restore the workload's own tables
    ctx.set_mt(s.maxaa, s.maxnlo, s.aa, s.loa, s.lolo, s.oalo, s.ololo, s.h1aa, s.h1loa,
               s.h1lolo)
    p = s.apwordmax
    nf = (s.lmaxmat + 1) * p + s.nlomax
    npair = nf * (nf + 1) // 2
    lmm = (lmaxvr + 1) ** 2
    nr = int(s.nrmt.max())
    flops = 2.0 * s.natmtot * npair * nr * lmm
    t_rad = res["hmlint"]["kernels_ms"]["hml_radial"] * 1e-3
    nel = s.maxaa * s.maxaa + mnlo * s.maxaa + mnlo * mnlo
    t_con = res["hmlint"]["kernels_ms"]["hml_contract"] * 1e-3
    res["radial_kernel"] = {"bound": "fp64 vector (LDS-fed)", "flops": flops,
                            "TFLOPs": flops / t_rad / 1e12,
                            "frac_fp64_peak": flops / t_rad / 78.6e12,
                            "per_atom": f"{npair} radial pairs x {lmm} lm x {nr} points"}
    res["contract_kernel"] = {"bound": "hbm (table writes)", "bytes": 16.0 * nel * s.natmtot,
                              "GBps": 16.0 * nel * s.natmtot / t_con / 1e9,
                              "frac_hbm": 16.0 * nel * s.natmtot / t_con / 8e12}
    res["tables_MB"] = 16.0 * nel * s.natmtot / 1e6
    res["workload"] = (f"{s.natmtot} atoms, lmaxmat {s.lmaxmat}, APW order {p}, "
                       f"{s.nlomax} LOs per atom, lmaxvr {lmaxvr}, {nr} radial points")
    if cpu:
        import oracle
        t0 = time.perf_counter()
        oracle.hmlint(oracle.System(sp))
        t = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": 1e3 * t, "unit": "ms per hmlint of the cell",
                               "cores": 1, "kind": "port",
                               "sample": (f"oracle hmlint (C restatement of mt_kin + mt_pot, "
                                          f"one thread) on all {s.natmtot} atoms")}
        t0 = time.perf_counter()
        oracle.mt_so(oracle.System(sp), pot=sp.veffmt)
        t = time.perf_counter() - t0
        res["mt_so_cpu_baseline"] = {"value": 1e3 * t, "unit": "ms per mt_so of the cell",
                                     "cores": 1, "kind": "port",
                                     "sample": f"oracle mt_so (C, one thread) on all {s.natmtot} atoms"}
    return res


def vnlmat_sample(W, nstfv=400, reps=3, cpu=True):
    """calc_vnlmat on k-point 0 of the workload: match + overlapsetup alone
    (O without the Hamiltonian's products) + the three products, inputs
    resident in HBM; random orthonormal-free states and a Hermitian vxnl of
    nstfv states."""
    import torch
    from exciting_amd import hybrids, lapw
    d = W.dks[0]
    n = d.rank
    g = torch.Generator(device="cuda").manual_seed(3)
    X = torch.randn(nstfv, n, dtype=torch.complex128, device="cuda", generator=g)
    V = torch.randn(nstfv, nstfv, dtype=torch.complex128, device="cuda", generator=g)
    V = V + V.conj().transpose(0, 1)
    O = W.O[0]
    hybrids.vnlmat_kpoint(W.ctx, d, X, V, O=O)            # warm-up
    torch.cuda.synchronize()
    W.ctx.profile(True)
    W.ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(reps):
        hybrids.vnlmat_kpoint(W.ctx, d, X, V, O=O)
    torch.cuda.synchronize()
    total = (time.perf_counter() - t0) / reps * 1e3
    prof = W.ctx.profile_read()
    W.ctx.profile(False)
    k = {name: ms / cnt for name, (ms, cnt) in prof.items() if cnt}
    setup_ms = sum(k.get(x, 0.0) for x in ("match", "ivgk", "pack", "a2t", "ho", "lo_border",
                                           "lolo"))
    f_o = flops_ho(W.s.natmtot, W.s.maxaa, d.ngk) / 2      # O alone: half the H/O products
    res = {"ms": total, "kernels_ms": k, "setup_ms": setup_ms,
           "ho_o_only": {"ms": k.get("ho", 0.0), "TFLOPs": f_o / (k.get("ho", 1.0) * 1e-3) / 1e12,
                         "frac_fp64_peak": f_o / (k.get("ho", 1.0) * 1e-3) / 78.6e12},
           "gemm_ms": total - setup_ms, "n": n, "nstfv": nstfv,
           "workload": f"k-point 0 of {W.spec.name}: n = {n}, nstfv = {nstfv}"}
    if cpu:
        import oracle
        S = oracle.System(W.s)
        t0 = time.perf_counter()
        A = oracle.match(S, 0)
        _, Oc = oracle.overlapsetup(S, 0, A)
        t1 = time.perf_counter()
        Xc, Vc = X.cpu().numpy().T.copy(), V.cpu().numpy()
        oracle.calc_vnlmat(Oc, Xc, Vc)
        t2 = time.perf_counter()
        res["cpu_baseline"] = {"value": 1e3 * (t2 - t0), "unit": "ms per k-point", "cores": 16,
                               "kind": "port",
                               "sample": (f"oracle match + overlapsetup {t1 - t0:.1f} s and NumPy "
                                          f"products {t2 - t1:.1f} s on k-point 0")}
    return res


def kinetic_sample(W, nstfv=400, reps=3, cpu=True, cpu_states=4):
    """kinetic_energy.f90 on k-point 0 of the workload ('none' relativity):
    muffin-tin part (kinetic radial integrals, A^T x per atom, quadratic
    forms) and interstitial part (the cyclic convolution with the step
    function as one dense G+k matrix from one FFT, one GEMM over the three
    directions), inputs resident in HBM; random states."""
    import dataclasses
    import torch
    from exciting_amd import kinetic, synth
    s = synth.add_mt_potential(dataclasses.replace(W.s), lmaxvr=8, seed=5)
    d = W.dks[0]
    W.ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, d.apwalm)
    n = d.rank
    g = torch.Generator(device="cuda").manual_seed(4)
    X = torch.randn(nstfv, n, dtype=torch.complex128, device="cuda", generator=g)
    rng = np.random.default_rng(3)
    cfunir = torch.from_numpy(0.5 + 0.5 * rng.uniform(0, 1, s.ngrtot)).cuda()
    dev = lambda a: torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).cuda()
    lofr, veffmt = dev(s.lofr), dev(s.veffmt)
    W.ctx.sys = s
    plan = kinetic.KineticPlan(s)
    kinetic.kinetic_energy(W.ctx, d, X, cfunir, lofr, True, veffmt, 8, plan=plan)   # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        kinetic.kinetic_energy(W.ctx, d, X, cfunir, lofr, True, veffmt, 8, plan=plan)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    ngk = d.ngk
    res = {"ms": ms, "nstfv": nstfv, "n": n,
           "interstitial_gemm_flops": 8.0 * ngk * ngk * 3 * nstfv,
           "grid": [int(x) for x in s.intgv[:, 1] - s.intgv[:, 0] + 1],
           "workload": f"k-point 0 of {W.spec.name}: n = {n}, nstfv = {nstfv}"}
    if cpu:
        import oracle
        S = oracle.System(s)
        A = d.apwalm.cpu().numpy().reshape(d.apwalm_shape, order="F")
        Xc = X[:cpu_states].cpu().numpy().T.copy()
        t0 = time.perf_counter()
        oracle.kinetic_energy(S, 0, Xc, A, cfunir.cpu().numpy(), s.lofr, s.veffmt, True)
        t = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": 1e3 * t * nstfv / cpu_states,
                               "unit": f"ms per k-point ({nstfv} states, scaled)",
                               "cores": 1, "kind": "port",
                               "sample": (f"oracle kinetic_energy (NumPy restatement) on "
                                          f"{cpu_states} states: {t:.1f} s, scaled to {nstfv}")}
    W.ctx.sys = W.s
    return res


def forces_sample(W, nstfv=400, reps=2, cpu=True):
    """forcek.f90 for k-point 0 of the workload (64 atoms, n = 10364) and of
    the 8-atom cell, with nstfv random states; the CPU oracle (a literal
    NumPy restatement with explicit n x n derivative matrices) is timed on
    the 8-atom cell, where it finishes in seconds."""
    import dataclasses
    import torch
    import oracle
    from exciting_amd import forces, lapw, synth

    def run(ctx, s, d, nst):
        n = d.rank
        g = torch.Generator(device="cuda").manual_seed(8)
        X = torch.randn(nst, n, dtype=torch.complex128, device="cuda", generator=g)
        e = torch.rand(nst, dtype=torch.float64, device="cuda", generator=g) - 0.5
        occ = 2.0 * torch.rand(nst, dtype=torch.float64, device="cuda", generator=g)
        ff = torch.from_numpy(np.stack([oracle.genffacg(s, i) for i in range(s.nspecies)],
                                       axis=1)).cuda()
        sg = torch.from_numpy(oracle.sfacg_of(s)).cuda()
        forces.forcek(ctx, d, X, e, occ, 1.0, ff, sg)                   # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            forces.forcek(ctx, d, X, e, occ, 1.0, ff, sg)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3, (X, e, occ, ff, sg)

    res = {}
    ms, _ = run(W.ctx, W.s, W.dks[0], nstfv)
    res["sc64_k0_ms"] = ms
    s8 = synth.build(synth.config(1), kpoints=[0])
    ctx8 = lapw.Context(0).load(s8)
    try:
        d8 = lapw.DeviceKPoint(s8, 0)
        ms8, (X, e, occ, ff, sg) = run(ctx8, s8, d8, nstfv)
        res["sc8_k0_ms"] = ms8
        if cpu:
            t0 = time.perf_counter()
            oracle.forcek(oracle.System(s8), 0, X.cpu().numpy().T.copy(), e.cpu().numpy(),
                          occ.cpu().numpy(), 1.0, ff.cpu().numpy(), sg.cpu().numpy())
            t = time.perf_counter() - t0
            res["cpu_baseline"] = {"value": 1e3 * t, "unit": "ms per k-point of the 8-atom cell",
                                   "cores": 16, "kind": "port",
                                   "sample": (f"oracle forcek (NumPy, explicit derivative "
                                              f"matrices) on the 8-atom cell's k-point 0, "
                                              f"n = {d8.rank}, {nstfv} states; GPU on the same: "
                                              f"{ms8:.1f} ms")}
    finally:
        ctx8.close()
    res["workload"] = (f"{nstfv} states; k-point 0 of sc64 (n = {W.dks[0].rank}) and of sc8 "
                       f"(n = {s8.kpts[0].ngk + s8.nlotot})")
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3, help="headline workload index (0..3)")
    ap.add_argument("--kpoints", type=int, default=0, help="use the first N k-points (0 = all)")
    ap.add_argument("--cpu-baseline", choices=("on", "off"), default="on")
    ap.add_argument("--moduli", default="auto",
                    help="split precision: number of int8 moduli, 0 = FP64 only, or 'auto': "
                         "the fewest moduli keeping the lowest eigenvalues of k-point 0 within "
                         "1/3 uHa of FP64 (checked again on the last k-point)")
    ap.add_argument("--nstates", type=int, default=400,
                    help="lowest eigenvalues compared between split and FP64 (one k-point; "
                         "0 = skip the check)")
    ap.add_argument("--tune", choices=("on", "off"), default="on",
                    help="pick the batch plan with the autotuner before timing")
    ap.add_argument("--plan", type=int, default=-1,
                    help="batch plan to use when --tune off (-1 = the library default)")
    ap.add_argument("--loop", choices=("on", "off"), default="on",
                    help="also time the single-k entry points in a loop")
    ap.add_argument("--workspace-gib", type=float, default=48.0,
                    help="batch scratch cap (48 GiB holds the 16 headline k-points in one group)")
    ap.add_argument("--uneven", choices=("on", "off"), default="on",
                    help="also run the 16-atom cell with 128 uneven k-points as one batch")
    ap.add_argument("--uneven-moduli", default="auto")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests in the serving sample (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=3.0)
    ap.add_argument("--serve-budget-gib", type=float, default=80.0,
                    help="admission budget cap of the serving sample (two 64-atom contexts "
                         "and a few small ones; 0 = free HBM minus the reserve)")
    ap.add_argument("--serve-kmax", type=int, default=4)
    ap.add_argument("--serve-nstfv", type=int, default=40)
    ap.add_argument("--serve-solves", type=int, default=1,
                    help="solver threads of the serving sample (the solves run one at a time)")
    ap.add_argument("--serve-moduli", type=int, default=9,
                    help="moduli of the split requests (9 keeps every cell of the mix within "
                         "1 uHa)")
    ap.add_argument("--interstitial", choices=("on", "off", "only"), default="on",
                    help="time the interstitial producers on the workload's grid "
                         "('only': nothing else, one JSON line with that sub-object)")
    ap.add_argument("--hmlint", choices=("on", "off", "only"), default="on",
                    help="time the muffin-tin table producer on the workload's cell "
                         "('only': nothing else)")
    ap.add_argument("--vnlmat", choices=("on", "off", "only"), default="on",
                    help="time calc_vnlmat (match + overlapsetup alone + 3 products) on "
                         "k-point 0 ('only': nothing else)")
    ap.add_argument("--kinetic", choices=("on", "off", "only"), default="on",
                    help="time kinetic_energy on k-point 0 ('only': nothing else)")
    ap.add_argument("--forces", choices=("on", "off", "only"), default="on",
                    help="time forcek on k-point 0 ('only': nothing else)")
    ap.add_argument("--host-pointer", choices=("on", "off"), default="on",
                    help="time the host-array drop-in (exh_match + exh_hamiltonandoverlapsetup "
                         "on NumPy arrays) on k-point 0, with and without apwalm reuse")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the timed region, run one device-to-device copy of this "
                         "many MiB (a known-size kernel for calibrating HBM byte counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(k-points would shard across GPUs with no exchange but a final "
              "allgather of eigenvalues). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from exciting_amd import solve
    solve.configure()          # before the first hipBLAS handle
    import torch
    from exciting_amd import lapw

    W = Workload(args.config, args.kpoints, args.workspace_gib)
    s, nk = W.s, W.nk
    if args.copy_calibration_mib:
        n = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(f"calibration copy of {args.copy_calibration_mib} MiB done")

    if args.interstitial == "only":
        print(json.dumps({"interstitial": interstitial_sample(W.ctx, s)}), flush=True)
        return
    if args.forces == "only":
        print(json.dumps({"forces": forces_sample(W, cpu=args.cpu_baseline == "on")}),
              flush=True)
        return
    if args.kinetic == "only":
        print(json.dumps({"kinetic": kinetic_sample(W, cpu=args.cpu_baseline == "on")}),
              flush=True)
        return
    if args.vnlmat == "only":
        print(json.dumps({"vnlmat": vnlmat_sample(W, cpu=args.cpu_baseline == "on")}),
              flush=True)
        return
    if args.hmlint == "only":
        print(json.dumps({"hmlint": hmlint_sample(W.ctx, s, cpu=args.cpu_baseline == "on")}),
              flush=True)
        return

    if args.plan >= 0:
        W.ctx.batch_plan = args.plan
    plan = W.ctx.batch_plan
    if args.tune == "on":
        t0 = time.perf_counter()
        plan = W.ctx.tune_batch(W.kp, flags=lapw.MATCH)
        log(f"autotuner: plan {plan} {lapw.plan_info(plan)} in {time.perf_counter() - t0:.1f} s")
    fp64_loop = None
    if args.loop == "on":
        fp64_loop = W.run(0, "fp64 loop", args.steps, args.warmup, batch=False)
    fp64 = W.run(0, "fp64 batch", args.steps, args.warmup)
    fp64["roofline"] = W.ho_roofline(fp64, "ho_grouped", f"ho_grouped_{W.spec.name}_fp64.md")
    fp64["plan"] = dict(lapw.plan_info(plan), id=plan)
    if fp64_loop:
        # the single-k kernel's traffic: its own row in the same summary
        fp64_loop["roofline"] = W.ho_roofline(fp64_loop, "ho",
                                              f"ho_grouped_{W.spec.name}_fp64.md",
                                              per="kpoint_single")

    head = fp64
    split = None
    nmod, mod_tuning = pick_moduli(W, args.moduli, args.nstates)
    if nmod:
        split = W.run(lapw.SPLIT(nmod), f"split({nmod}) batch", args.steps, args.warmup)
        split["roofline"] = W.oz_roofline(split, nmod, f"ozgemm_{W.spec.name}_split{nmod}.md")
        g = W.run(lapw.SPLIT(nmod) | lapw.GRAPH, f"split({nmod}) graph", args.steps,
                  args.warmup, profile=False)
        split["graph"] = {"value": g["value"], "ms_per_step": g["ms_per_step"]}
        split["moduli"] = nmod
        split["bits"] = oz_bits(nmod, W.K)
        split["moduli_tuning"] = mod_tuning
        if args.nstates:
            kc = nk - 1
            split["accuracy"] = accuracy(W.ctx, W.kp[kc], W.dks[kc], W.H[kc], W.O[kc], nmod,
                                         args.nstates, kpoint=kc)
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
            "workload": (f"{W.spec.name}: {s.natmtot}-atom cell, lmaxapw={s.lmaxapw}, APW order "
                         f"{s.apwordmax} + {W.nlo_atom} LO functions per atom, rgkmax="
                         f"{W.spec.rgkmax}, {nk} k-points, match + H/O setup as one batch"),
            "kpoints": nk,
            "ngk_mean": float(np.mean(W.ngks)),
            "n_mean": float(np.mean(W.ngks)) + s.nlotot,
            "natoms": s.natmtot,
            "precision": (f"split: muffin-tin products by int8 Ozaki scheme II with {nmod} "
                          f"moduli (G columns scaled to 2-norm 2^{oz_bits(nmod, W.K)} and rounded "
                          f"column); match, LO blocks, interstitial FP64") if nmod else "fp64",
        },
        "gflops": head["gflops"],
        "latency_ms": head["latency_ms"],
        "kernels_ms_per_setup": head["kernels_ms_per_setup"],
        "roofline": head["roofline"],
        "fp64": clean(fp64),
    }
    if fp64_loop:
        out["fp64_loop"] = clean(fp64_loop)
    if split:
        out["split"] = {k: split[k] for k in ("value", "ms_per_step", "graph", "moduli", "bits",
                                              "moduli_tuning")}
        out["speedup_vs_fp64"] = split["value"] / fp64["value"]
        if fp64_loop:
            out["speedup_vs_fp64_loop"] = split["value"] / fp64_loop["value"]
        if "accuracy" in split:
            out["accuracy"] = split["accuracy"]
    if args.interstitial == "on":
        out["interstitial"] = interstitial_sample(W.ctx, s)
        log(f"interstitial: {out['interstitial']}")
    if args.hmlint == "on":
        out["hmlint"] = hmlint_sample(W.ctx, s, cpu=args.cpu_baseline == "on")
        log(f"hmlint: {out['hmlint']}")
    if args.vnlmat == "on":
        out["vnlmat"] = vnlmat_sample(W, cpu=args.cpu_baseline == "on")
        log(f"vnlmat: {out['vnlmat']}")
    if args.kinetic == "on":
        out["kinetic"] = kinetic_sample(W, cpu=args.cpu_baseline == "on")
        log(f"kinetic: {out['kinetic']}")
    if args.forces == "on":
        out["forces"] = forces_sample(W, cpu=args.cpu_baseline == "on")
        log(f"forces: {out['forces']}")
    if args.host_pointer == "on":
        out["host_pointer"] = host_pointer_sample(W.ctx, s, 0)
    W.free()

    if args.uneven == "on":
        U = Workload(2, 0, 0.0)
        up = U.ctx.tune_batch(U.kp, flags=lapw.MATCH) if args.tune == "on" else U.ctx.batch_plan
        log(f"uneven: plan {up} {lapw.plan_info(up)}")
        ub = U.run(0, "uneven fp64 batch", args.steps, args.warmup)
        ug = U.run(lapw.GRAPH, "uneven fp64 graph", args.steps, args.warmup, profile=False)
        ul = U.run(0, "uneven fp64 loop", args.steps, args.warmup, batch=False)
        um, um_tuning = pick_moduli(U, args.uneven_moduli, args.nstates)
        us = U.run(lapw.SPLIT(um), f"uneven split({um}) batch", args.steps, args.warmup)
        kc = U.nk - 1
        uacc = accuracy(U.ctx, U.kp[kc], U.dks[kc], U.H[kc], U.O[kc], um, args.nstates or 400,
                        kpoint=kc)
        out["uneven"] = {
            "workload": (f"{U.spec.name}: {U.s.natmtot}-atom orthorhombic cell, {U.nk} "
                         f"k-points, ngk {min(U.ngks)}..{max(U.ngks)}, one batch per step"),
            "plan": dict(lapw.plan_info(up), id=up),
            "fp64_batch": {k: ub[k] for k in ("value", "ms_per_step", "gflops", "frac_fp64_peak")},
            "fp64_graph": {k: ug[k] for k in ("value", "ms_per_step")},
            "fp64_loop": {k: ul[k] for k in ("value", "ms_per_step", "gflops", "frac_fp64_peak")},
            "batch_vs_loop": ub["value"] / ul["value"],
            "ho_roofline": U.ho_roofline(ub, "ho_grouped", f"ho_grouped_{U.spec.name}_fp64.md"),
            "split_batch": {k: us[k] for k in ("value", "ms_per_step")},
            "split_moduli": um,
            "split_moduli_tuning": um_tuning,
            "split_vs_fp64_batch": us["value"] / ub["value"],
            "split_accuracy": uacc,
        }
        log(f"uneven: {out['uneven']}")
        U.free()

    if args.serve_requests:
        out["serving"] = serving_sample(args)
        log(f"serving: {out['serving']}")

    if args.cpu_baseline == "on":
        log("cpu baseline: oracle on one k-point")
        out["cpu_baseline"] = cpu_baseline(s, 0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
