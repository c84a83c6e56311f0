"""Serving on the GPU: each result is the same alone and in the mix.

A fixed list of setup + eigensolve requests over three cells (FP64 and split
precision, 1-4 k-points each) is served twice: one request at a time with
one k-point slot per context, then all at once with up to six in flight and
two slots per context, their k-point setups merged into grouped launches
over several cells (exh_setup_multi_device) and their eigensolves
queued for the solver thread (setup launches and solves take turns on the
GPU).  Tolerance: the H
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
    ex = serve.GpuExecutor(0, digests=True)
    make, close, execute = ex.make_context, ex.close_context, ex.execute
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
    alone_stats = ex.stats()
    ex.close()
    print("[gpu-serve] mix of 10 requests, up to 6 in flight", flush=True)
    mix = _requests()
    # the ten requests arrive within 20 ms: a 20 ms linger lets the setups
    # of concurrent requests of different cells share launches; two k-point
    # slots per context let a request's own k-points share them too
    ex2 = serve.GpuExecutor(0, digests=True, linger_s=0.02, slots=2)
    srv2 = serve.Server(cells, ex2.execute, ex2.make_context, ex2.close_context, budget,
                        max_inflight=6, footprint=ex2.footprint)
    srv2.run(mix)
    srv2.pool.close_all()
    mix_stats = ex2.stats()
    ex2.close()
    print(f"  batching alone: {alone_stats}\n  batching in the mix: {mix_stats}", flush=True)
    yield cells, alone, mix, srv2, alone_stats, mix_stats


@pytest.mark.gpu
def test_every_request_succeeds_and_overlaps(served):
    cells, alone, mix, srv, alone_stats, mix_stats = served
    for r in alone:
        assert r.status == "done", (r.rid, r.error)
        assert r.launches == len(r.kpoints)
    for r in mix:
        assert r.status == "done", (r.rid, r.error)
        assert 1 <= r.launches <= len(r.kpoints)
    assert srv.peak_inflight > 1


@pytest.mark.gpu
def test_requests_of_several_cells_share_launches(served):
    """In the mix the setups of concurrent requests run as grouped launches
    over several cells: fewer launches than k-points, and at least one
    launch holding k-points of two or more cells.  Alone, with one slot per
    context, every launch holds one k-point."""
    cells, alone, mix, srv, alone_stats, mix_stats = served
    nk = sum(len(r.kpoints) for r in mix)
    assert alone_stats["kpoints_per_launch"]["max"] == 1
    assert mix_stats["kpoints"] == nk
    assert mix_stats["launches"] < nk
    assert mix_stats["cells_per_launch"]["max"] >= 2


@pytest.mark.gpu
def test_same_result_alone_and_in_the_mix(served):
    cells, alone, mix, srv, _, _ = served
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
    cells, alone, mix, srv, _, _ = served
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
    from exciting_amd import solve
    cell = _cases.system("fcc2")
    fresh_pool = (0, 2) not in serve._SOLVER_POOLS
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated(0)
    held = []
    for i in range(5):
        print(f"[gpu-serve] context {i}", flush=True)
        # a new executor each time: they share the process's solver threads
        ex = serve.GpuExecutor(0, max_solves=2)
        ctx = ex.make_context(cell)
        req = serve.Request(rid=i, cell="fcc2", kpoints=[0], nmod=0, nstfv=20)
        res = {}

        def run():
            try:
                res["w"] = ex.execute(req, ctx)
            except BaseException as e:     # surfaced below, not lost in the thread
                res["e"] = e
        t = threading.Thread(target=run)
        t.start()
        t.join()
        ex.close_context(ctx)
        ex.close()
        assert "e" not in res, res.get("e")
        assert len(res["w"]) == 1 and len(res["w"][0]) == 20
        torch.cuda.synchronize()
        held.append(torch.cuda.memory_allocated(0))
    print(f"  PyTorch allocated after each context: {[h / 2**30 for h in held]} GiB "
          f"(before: {before / 2**30:.2f} GiB)")
    # the first solve on a new solver thread made a workspace, and none of
    # the later contexts and executors added one
    if fresh_pool:
        assert held[0] - before >= solve.WORKSPACE_BYTES
    assert held[-1] <= held[0]
