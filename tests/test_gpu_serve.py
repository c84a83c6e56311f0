"""Serving on the GPU: each result is the same alone and in the mix.

A fixed list of setup + eigensolve requests over three cells (FP64 and split
precision, 1-4 k-points each) is served twice: one request at a time, then
all at once with up to six in flight on their own streams.  Tolerance: the H
and O matrices must be bitwise identical (SHA-1 of their bytes per k-point),
and the eigenvalues identical to 1e-12 Ha (they come from the same library
on bitwise identical inputs).  The FP64 eigenvalues are also checked against
the oracle's H and O solved with LAPACK (1e-8 Ha), the split ones within
1 uHa.
"""
import copy

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from exciting_amd import serve, synth
from tests import _cases


def _cells():
    return {"tiny": _cases.system("tiny"),
            "fcc2": _cases.system("fcc2"),
            "sc8": synth.build(synth.config(1), kpoints=[0, 1, 2, 3])}


def _requests():
    spec = [("tiny", [0], 0), ("fcc2", [0], 9), ("sc8", [0, 1], 0), ("sc8", [2], 9),
            ("fcc2", [0], 0), ("sc8", [1, 3], 9), ("tiny", [0], 9), ("sc8", [0, 1, 2, 3], 0),
            ("fcc2", [0], 8), ("sc8", [3], 0)]
    return [serve.Request(rid=i, cell=c, kpoints=ks, nmod=m, nstfv=30, arrival=0.002 * i)
            for i, (c, ks, m) in enumerate(spec)]


@pytest.fixture(scope="module")
def served():
    cells = _cells()
    make, close, execute = serve.gpu_executor(0, digests=True)
    budget = serve.free_hbm(0) - (16 << 30)
    alone = []
    srv = serve.Server(cells, execute, make, close, budget, max_inflight=1)
    for r in _requests():
        print(f"[gpu-serve] alone: request {r.rid} {r.cell} k={r.kpoints} nmod={r.nmod}",
              flush=True)
        r.arrival = 0.0
        srv.run([r])
        alone.append(r)
    srv.pool.close_all()
    print("[gpu-serve] mix of 10 requests, up to 6 in flight", flush=True)
    mix = _requests()
    srv2 = serve.Server(cells, execute, make, close, budget, max_inflight=6)
    srv2.run(mix)
    srv2.pool.close_all()
    yield cells, alone, mix, srv2


@pytest.mark.gpu
def test_every_request_succeeds_and_overlaps(served):
    cells, alone, mix, srv = served
    for r in alone + mix:
        assert r.status == "done", (r.rid, r.error)
    assert srv.peak_inflight > 1


@pytest.mark.gpu
def test_same_result_alone_and_in_the_mix(served):
    cells, alone, mix, srv = served
    nbit = 0
    for a, m in zip(alone, mix):
        assert a.rid == m.rid
        assert a.digests == m.digests, f"request {a.rid}: H/O differ between alone and mix"
        for wa, wm in zip(a.eigenvalues, m.eigenvalues):
            assert np.abs(wa - wm).max() <= 1e-12
            nbit += int(np.array_equal(wa, wm))
    print(f"  {nbit} of {sum(len(r.kpoints) for r in mix)} eigenvalue sets bitwise equal")


@pytest.mark.gpu
def test_served_eigenvalues_match_the_oracle(served):
    cells, alone, mix, srv = served
    ref = {}
    for r in mix:
        s = cells[r.cell]
        S = oracle.System(s)
        for ik, w in zip(r.kpoints, r.eigenvalues):
            if (r.cell, ik) not in ref:
                A = oracle.match(S, ik)
                H0, O0 = oracle.setup(S, ik, A)
                ref[(r.cell, ik)] = sl.eigh(H0, O0, lower=False, eigvals_only=True,
                                            subset_by_index=[0, len(w) - 1])
            de = np.abs(w - ref[(r.cell, ik)]).max()
            tol = 1e-8 if r.nmod == 0 else 1e-6
            print(f"  request {r.rid} {r.cell} k{ik} nmod={r.nmod}: max|dE| {de:.1e} Ha")
            assert de <= tol


@pytest.mark.gpu
def test_solver_workspaces_do_not_grow_with_contexts():
    """PyTorch keeps a 2 GiB hipBLAS workspace per (thread, stream) pair
    that runs a BLAS call.  The executor solves on its own fixed threads and
    streams, so building, using and closing one context after another (each
    with its own stream, each request on its own thread) adds no memory
    once the first solve has made its workspace."""
    import threading
    import torch
    cell = _cases.system("fcc2")
    make, close, execute = serve.gpu_executor(0, max_solves=2)
    held = []
    for i in range(5):
        print(f"[gpu-serve] context {i}", flush=True)
        ctx = make(cell)
        req = serve.Request(rid=i, cell="fcc2", kpoints=[0], nmod=0, nstfv=20)
        t = threading.Thread(target=lambda: execute(req, ctx))
        t.start()
        t.join()
        close(ctx)
        torch.cuda.synchronize()
        held.append(torch.cuda.memory_allocated(0))
    print(f"  PyTorch allocated after each context: {[h / 2**30 for h in held]} GiB")
    assert held[-1] <= held[0]
