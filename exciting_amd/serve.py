"""Continuous-batching request loop for setup + eigensolve requests.

A request names a cell (one of a fixed set whose module state -- radial
data, muffin-tin tables, G-vector maps, interstitial functions -- stays
resident in HBM across requests), a list of its k-points, a precision
(FP64 or split with n moduli) and the number of states.  For every k-point
the server runs match, the H/O setup and the generalized eigensolve on the
GPU and returns the lowest eigenvalues.

Scheduling:
  * The device budget is the free HBM reported by hipMemGetInfo when the
    server starts, minus a reserve.  Every library context is charged its
    whole footprint (apwalm, H and O of the cell's largest k-point, setup
    scratch for FP64 and split precision, resident per-k inputs and the
    eigensolve temporaries, which PyTorch's allocator keeps cached per
    stream) for as long as it exists.
  * Requests are admitted in arrival order as soon as an execution slot is
    free and either an idle context of their cell exists or the budget has
    room for a new one; idle contexts of other cells are evicted, least
    recently used first, to make room.  A request that does not fit waits,
    and later requests that fit are admitted around it (bypass), so one large
    request does not stall the queue.  A request whose context alone exceeds
    the budget fails at once.
  * Each admitted request runs on its own context (own HIP stream and
    scratch) in its own host thread, so requests that arrive while others are
    on the GPU start at once and their kernels overlap on the device.  The
    eigensolves (PyTorch's ROCm solver, whose tridiagonal reduction is a long
    chain of small kernels) run concurrently on a fixed set of solver
    threads, each with its own stream: PyTorch keeps a 2 GiB hipBLAS
    workspace for every (thread's handle, stream) pair that ever runs a BLAS
    call, so solving on per-request threads and per-context streams would
    add one such workspace per new pair for the life of the process.
  * Results do not depend on what else runs: every setup kernel has a fixed
    reduction order and owns its output tiles (no atomics on H or O), and
    the eigensolve sees bitwise identical inputs.

Failure handling, forced from the host and tested on the CPU with a fake
executor: a queued request can be cancelled, a request larger than the
budget is refused, an executor error fails only its request (the context
goes back to the pool), and the loop drains in-flight work before it
returns.
"""
from __future__ import annotations

import argparse
import concurrent.futures
import dataclasses
import hashlib
import json
import sys
import threading
import time
from typing import Callable, Dict, List, Optional

import numpy as np


@dataclasses.dataclass
class Request:
    rid: int
    cell: str
    kpoints: List[int]
    nmod: int = 0            # 0 = FP64, else split precision with nmod moduli
    nstfv: int = 40
    arrival: float = 0.0     # seconds after the start of the run
    admitted: float = -1.0
    finished: float = -1.0
    status: str = "queued"   # queued | running | done | cancelled | failed
    error: str = ""
    eigenvalues: Optional[List[np.ndarray]] = None
    digests: Optional[List[str]] = None   # per k: hash of H and O (if asked)
    retries_setup: int = 0   # setup launches rerun after an out-of-memory error
    retries_solve: int = 0   # eigensolves rerun after an out-of-memory error
    launches: int = 0        # grouped setup launches its k-points were part of

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def context_bytes(sys_, nmod_max: int = 12, slots: int = 1) -> int:
    """HBM one context holds while it exists, for a cell (LapwSystem), with
    ``slots`` k-point slots (apwalm, H, O and eigensolve temporaries each)."""
    n = sys_.ngkmax + sys_.nlotot
    lmm = (sys_.lmaxapw + 1) ** 2
    apwalm = 16 * sys_.ngkmax * sys_.apwordmax * lmm * sys_.natmtot
    ho = 2 * 16 * n * n
    K = sys_.natmtot * sys_.maxaa
    ldb = (sys_.ngkmax + 63) // 64 * 64
    scratch = 16 * K * ldb                       # aa.A operand
    if nmod_max:
        gpad = (sys_.ngkmax + 383) // 384 * 384     # rows padded to the GEMM's panel
        kpad = (K + 63) // 64 * 64
        nt = (sys_.ngkmax + 127) // 128
        scratch += 7 * nmod_max * kpad * gpad    # int8 residues of A and aa.A
        scratch += 6 * nmod_max * nt * (nt + 1) // 2 * 128 * 128   # product residues
    perk = len(sys_.kpts) * sys_.ngkmax * (16 * sys_.natmtot + 8 * 7 + 4)
    # Cholesky factor, the two triangular solves' results, the reduced matrix
    # and its Hermitian average, and the eigensolver's copy and workspace
    solve = 8 * 16 * n * n
    return slots * (apwalm + ho + solve) + scratch + perk


class ContextPool:
    """Per-cell pools of idle library contexts with the cell's tables
    resident; tracks the bytes of all contexts (idle and in use)."""

    def __init__(self, make_context: Callable, close_context: Callable):
        self.make = make_context
        self.close = close_context
        self.idle: Dict[str, list] = {}
        self.lru: List[tuple] = []   # (last use, cell, ctx, bytes) of idle contexts
        self.bytes = 0
        self.created = 0
        self.evicted = 0

    def take_idle(self, cell: str):
        lst = self.idle.get(cell)
        if not lst:
            return None
        ctx = lst.pop()
        self.lru = [e for e in self.lru if e[2] is not ctx]
        return ctx

    def evict_one(self) -> bool:
        if not self.lru:
            return False
        self.lru.sort(key=lambda e: e[0])
        _, cell, ctx, nbytes = self.lru.pop(0)
        self.idle[cell].remove(ctx)
        self.close(ctx)
        self.bytes -= nbytes
        self.evicted += 1
        return True

    def give_back(self, cell: str, ctx, nbytes: int, now: float) -> None:
        self.idle.setdefault(cell, []).append(ctx)
        self.lru.append((now, cell, ctx, nbytes))

    def close_all(self) -> None:
        while self.evict_one():
            pass


_NEW = object()      # admission result: build a new context in the worker
_NEVER = object()    # admission result: the context alone exceeds the budget


class Server:
    """Admission + execution loop.

    ``execute(req, ctx)`` runs one request on one context and returns its
    per-k eigenvalues.  ``make_context(cell)``/``close_context(ctx)`` build
    and free a context with a cell's tables resident; ``footprint(cell)`` is
    the bytes a context of that cell holds.  Everything is injectable so the
    scheduling logic is testable without a GPU."""

    def __init__(self, cells: Dict[str, object], execute: Callable, make_context: Callable,
                 close_context: Callable, budget_bytes: int, max_inflight: int = 8,
                 footprint: Callable = context_bytes,
                 clock: Callable[[], float] = time.perf_counter,
                 on_finish: Optional[Callable[[Request], None]] = None):
        self.cells = cells
        self.execute = execute
        self.pool = ContextPool(make_context, close_context)
        self.budget = int(budget_bytes)
        self.max_inflight = max_inflight
        self.footprint = footprint
        self.clock = clock
        self.on_finish = on_finish
        self.lock = threading.Lock()
        self.inflight: Dict[int, threading.Thread] = {}
        self.cancelled = set()
        self.peak_inflight = 0
        self.peak_bytes = 0
        self.bypassed = 0        # admitted while an earlier arrival waited for HBM
        self.hbm_waits = 0       # requests that had to wait for HBM at least once
        self._waited = set()

    def cancel(self, rid: int) -> None:
        """Cancel a request that has not been admitted yet."""
        with self.lock:
            self.cancelled.add(rid)

    def _admit(self, req: Request):
        """An idle context, _NEW (bytes reserved), None (not now) or
        _NEVER.  Called with the lock held."""
        ctx = self.pool.take_idle(req.cell)
        if ctx is not None:
            return ctx
        need = self.footprint(self.cells[req.cell])
        if need > self.budget:
            return _NEVER
        idle_bytes = sum(e[3] for e in self.pool.lru)
        if self.pool.bytes - idle_bytes + need > self.budget:
            return None          # would not fit even with every idle context evicted
        while self.pool.bytes + need > self.budget:
            self.pool.evict_one()
        self.pool.bytes += need
        self.pool.created += 1
        self.peak_bytes = max(self.peak_bytes, self.pool.bytes)
        return _NEW

    def _worker(self, req: Request, t0: float, ctx):
        cell = self.cells[req.cell]
        nbytes = self.footprint(cell)
        try:
            if ctx is _NEW:
                ctx = self.pool.make(cell)
            req.eigenvalues = self.execute(req, ctx)
            req.status = "done"
        except Exception as e:  # reported per request; the loop keeps serving
            req.status = "failed"
            req.error = f"{type(e).__name__}: {e}"
        finally:
            req.finished = self.clock() - t0
            with self.lock:
                if ctx is _NEW:          # building the context failed
                    self.pool.bytes -= nbytes
                else:
                    self.pool.give_back(req.cell, ctx, nbytes, req.finished)
                self.inflight.pop(req.rid, None)
            if self.on_finish is not None:
                self.on_finish(req)

    def run(self, requests: List[Request], poll: float = 0.0005) -> List[Request]:
        """Serve ``requests`` (arrival times relative to the start) to
        completion; fills in admission and finish times and results."""
        arrivals = sorted(requests, key=lambda r: r.arrival)
        queue: List[Request] = []
        t0 = self.clock()
        nxt = 0
        while True:
            with self.lock:
                busy = bool(self.inflight)
            if nxt >= len(arrivals) and not queue and not busy:
                break
            now = self.clock() - t0
            while nxt < len(arrivals) and arrivals[nxt].arrival <= now:
                queue.append(arrivals[nxt])
                nxt += 1
            blocked = False
            for req in list(queue):
                with self.lock:
                    if req.rid in self.cancelled:
                        req.status = "cancelled"
                        req.finished = now
                        queue.remove(req)
                        continue
                    if len(self.inflight) >= self.max_inflight:
                        break
                    got = self._admit(req)
                    if got is _NEVER:
                        req.status = "failed"
                        req.error = "request larger than the device budget"
                        req.finished = now
                        queue.remove(req)
                        continue
                    if got is None:
                        blocked = True
                        if req.rid not in self._waited:
                            self._waited.add(req.rid)
                            self.hbm_waits += 1
                        continue  # bypass: later requests may fit
                    if blocked:
                        self.bypassed += 1
                    req.status = "running"
                    req.admitted = now
                    queue.remove(req)
                    th = threading.Thread(target=self._worker, args=(req, t0, got), daemon=True)
                    self.inflight[req.rid] = th
                    self.peak_inflight = max(self.peak_inflight, len(self.inflight))
                    th.start()
            wait = poll
            if nxt < len(arrivals) and not queue:
                with self.lock:
                    idle = not self.inflight
                if idle:
                    wait = max(poll, arrivals[nxt].arrival - (self.clock() - t0))
            time.sleep(min(wait, 0.05))
        return requests


# ---------------------------------------------------------------------------
# GPU executor
# ---------------------------------------------------------------------------
_SOLVER_POOLS: Dict[tuple, tuple] = {}
_SOLVER_LOCK = threading.Lock()
# Setups and eigensolves take turns on the GPU, one at a time per process.
# PyTorch's linear algebra (rocBLAS, rocSOLVER) called from several host
# threads on their own streams at once gave wrong results here: four threads
# solving the same 8-atom H and O concurrently (Cholesky, triangular solves,
# GEMMs, eigh) differed from the one-thread eigenvalues in 5 of 160 solves,
# by up to 1e-4 Ha, whether the subspace or the full eigensolve ran; with
# every library call serialized and synchronized, none of 320 did.  With
# the solves serialized but a grouped setup launch (this library's kernels,
# on the batcher's stream) allowed to run beside a solve, 2 of 30 served
# mixes still gave other eigenvalues from bitwise equal H and O; with the
# setup launches taking turns too, none of 36 did.  Each holder waits for
# its own stream before releasing the lock.
_GPU_GATE = threading.Lock()


def _solver_pool(device: int, max_solves: int):
    """The process's solver threads for one device, created once and shared by
    every executor: at most ``max_solves`` (thread, stream) pairs run BLAS,
    so PyTorch holds at most that many hipBLAS workspaces however many
    executors, contexts and request threads come and go.  The solves
    themselves run one at a time (_GPU_GATE); more threads only queue."""
    key = (device, max(1, max_solves))
    with _SOLVER_LOCK:
        if key not in _SOLVER_POOLS:
            pool = concurrent.futures.ThreadPoolExecutor(max_workers=key[1],
                                                         thread_name_prefix="exh-solve")
            _SOLVER_POOLS[key] = (pool, threading.local())
        return _SOLVER_POOLS[key]


@dataclasses.dataclass
class _Slot:
    """apwalm, H and O of one k-point of a context (CUDA tensors)."""
    apwalm: object
    H: object
    O: object


@dataclasses.dataclass
class _Setup:
    req: Request
    ctx: object
    d: object            # lapw.DeviceKPoint
    slot: _Slot
    flags: int
    fut: concurrent.futures.Future


class Batcher:
    """Merges the k-point setups of every request in flight into grouped
    launches.  Request threads submit the k-points their context's slots
    can hold (match + H/O setup into a slot's apwalm, H and O); the batcher thread takes
    whatever is pending -- up to ``max_k`` k-points, of any cells -- and
    runs it as one exh_setup_multi_device call per precision on its own
    stream, then waits for that launch.  Setups that arrive meanwhile queue
    and form the next launch, so the batch grows with the load; when the
    GPU is idle the first pending setup waits at most ``linger_s`` for others
    to join it.  Per k-point the result is bitwise what the single-context
    path gives, whatever else shares the launch."""

    def __init__(self, device: int = 0, max_k: int = 32, plan: int = -1,
                 linger_s: float = 0.002):
        import torch
        self.device = device
        self.max_k = max_k
        self.plan = plan
        self.linger_s = linger_s
        self.stream = torch.cuda.Stream(device=f"cuda:{device}")
        self.cv = threading.Condition()
        self.q: List[_Setup] = []
        self.closed = False
        self.launches = 0
        self.kpoints = 0
        self.max_kpoints = 0
        self.cells = 0          # sum over launches of distinct cells
        self.max_cells = 0
        self.requests = 0       # sum over launches of distinct requests
        self.retries = 0
        self.busy_s = 0.0
        self.thread = threading.Thread(target=self._loop, name="exh-batcher", daemon=True)
        self.thread.start()

    def submit(self, req: Request, ctx, d, slot: _Slot, flags: int) -> concurrent.futures.Future:
        """Queue match + setup of one k-point into ``slot``; the future gives
        the event that completes with it."""
        fut: concurrent.futures.Future = concurrent.futures.Future()
        with self.cv:
            if self.closed:
                raise RuntimeError("batcher closed")
            self.q.append(_Setup(req, ctx, d, slot, flags, fut))
            self.cv.notify()
        return fut

    def close(self) -> None:
        with self.cv:
            self.closed = True
            self.cv.notify()
        self.thread.join()

    def _loop(self) -> None:
        while True:
            with self.cv:
                while not self.q and not self.closed:
                    self.cv.wait()
                if not self.q:
                    return
                deadline = time.perf_counter() + self.linger_s
                while len(self.q) < self.max_k and not self.closed:
                    rem = deadline - time.perf_counter()
                    if rem <= 0:
                        break
                    self.cv.wait(rem)
                items, self.q = self.q[:self.max_k], self.q[self.max_k:]
            groups: Dict[int, List[_Setup]] = {}
            for it in items:
                groups.setdefault(it.flags, []).append(it)
            for flags, its in groups.items():
                try:
                    ev = self._launch(its, flags)
                except Exception as e:  # fails the requests of this launch only
                    for it in its:
                        it.fut.set_exception(e)
                    continue
                for it in its:
                    it.fut.set_result(ev)

    def _launch(self, its: List[_Setup], flags: int):
        import torch
        from . import lapw
        kp = [lapw.kpoint_struct(it.d, it.slot.H, it.slot.O, apwalm=it.slot.apwalm)
              for it in its]
        ctxs = [it.ctx for it in its]
        lead = ctxs[0]
        t0 = time.perf_counter()

        def run():
            if self.plan >= 0:
                lead.batch_plan = self.plan
            lapw.setup_multi(ctxs, kp, flags=flags | lapw.MATCH,
                             stream=self.stream.cuda_stream)

        with _GPU_GATE:         # no eigensolve runs beside the launch
            try:
                run()
            except lapw.ExhError as e:
                # The contexts' scratch grows on first use of a larger setup
                # (the first split one, say).  Blocks PyTorch's caching
                # allocator holds free for the eigensolves are not visible to
                # hipMalloc: release them and run the launch once more (it
                # overwrites H and O).
                if "out of memory" not in str(e):
                    raise
                for it in its:
                    it.req.retries_setup += 1
                self.retries += 1
                torch.cuda.synchronize(self.device)
                torch.cuda.empty_cache()
                run()
            ev = torch.cuda.Event()
            ev.record(self.stream)
            ev.synchronize()    # setups arriving meanwhile form the next launch
        self.busy_s += time.perf_counter() - t0
        ncell = len({id(it.ctx.sys) for it in its})
        self.launches += 1
        self.kpoints += len(its)
        self.max_kpoints = max(self.max_kpoints, len(its))
        self.cells += ncell
        self.max_cells = max(self.max_cells, ncell)
        reqs = {id(it.req): it.req for it in its}
        self.requests += len(reqs)
        for r in reqs.values():
            r.launches += 1
        return ev

    def stats(self) -> dict:
        n = max(self.launches, 1)
        return {"launches": self.launches, "kpoints": self.kpoints,
                "kpoints_per_launch": {"mean": self.kpoints / n, "max": self.max_kpoints},
                "cells_per_launch": {"mean": self.cells / n, "max": self.max_cells},
                "requests_per_launch": self.requests / n, "launch_retries": self.retries,
                "busy_s": self.busy_s}


class GpuExecutor:
    """Runs requests on ``cuda:device``: the k-point setups of all requests
    in flight go through one Batcher (grouped multi-structure launches), the
    eigensolves through the process's solver threads for the device
    (``max_solves`` threads; solves and setup launches take turns, see
    _GPU_GATE).
    ``make_context``/``close_context``/``execute`` plug into Server.  With ``digests`` each request also records a SHA-1 of
    the bytes of H and O per k-point (copies them to the host: for tests on
    small cells)."""

    def __init__(self, device: int = 0, digests: bool = False, max_solves: int = 1,
                 max_batch_k: int = 32, plan: int = -1, linger_s: float = 0.002,
                 slots: int = 1):
        from . import solve as _solve
        _solve.configure()
        self.device = device
        self.dev = f"cuda:{device}"
        self.digests = digests
        self.solvers, self._local = _solver_pool(device, max_solves)
        # k-point slots per context: with two, a request's next setup runs
        # while the previous k-point's eigensolve does, and its first two
        # setups share a launch.  One is the default: at 64 atoms two slots
        # took a 16-k-point request from 10.50 to 10.29 s (the eigensolves
        # already keep the GPU busy) for twice the context memory.
        self.slots = max(1, slots)
        self.batcher = Batcher(device, max_k=max_batch_k, plan=plan, linger_s=linger_s)

    def _solve_on_worker(self, req, ready, H, O, n, nst):
        import torch
        from . import solve as _solve
        st = getattr(self._local, "stream", None)
        if st is None:
            st = self._local.stream = torch.cuda.Stream(device=self.dev)
        st.wait_event(ready)                   # the setup of H and O is done
        with _GPU_GATE, torch.cuda.stream(st):
            try:
                w = _solve.geneig(H, O, n, nst, method="auto")
            except torch.cuda.OutOfMemoryError:
                # freed temporaries of other solves may sit in PyTorch's
                # cache in pieces too small for this one
                req.retries_solve += 1
                torch.cuda.synchronize(self.device)
                torch.cuda.empty_cache()
                w = _solve.geneig(H, O, n, nst, method="auto")
            out = w.cpu().numpy()              # waits for the solve
            del w
            st.synchronize()                   # nothing of it left running
        # the solve's n x n temporaries are free now; hand them back so that
        # they do not sit in PyTorch's cache in pieces that the admission
        # budget cannot see and hipMalloc cannot use
        torch.cuda.empty_cache()
        return out

    def make_context(self, sys_):
        import torch
        from . import lapw
        ctx = lapw.Context(self.device).load(sys_)
        ctx.kp = {}      # per-k inputs, uploaded on first use and kept
        n = sys_.ngkmax + sys_.nlotot
        lmm = (sys_.lmaxapw + 1) ** 2
        ctx.tstream = torch.cuda.ExternalStream(ctx.stream, device=self.dev)
        with torch.cuda.stream(ctx.tstream):
            ctx.slots = [_Slot(torch.zeros(sys_.natmtot * lmm * sys_.apwordmax * sys_.ngkmax,
                                           dtype=torch.complex128, device=self.dev),
                               torch.empty(n * n, dtype=torch.complex128, device=self.dev),
                               torch.empty(n * n, dtype=torch.complex128, device=self.dev))
                         for _ in range(self.slots)]
        ctx.tstream.synchronize()
        return ctx

    def _kpoint(self, ctx, ik):
        """The per-k inputs of k-point ik of the context's cell, uploaded on
        first use."""
        import torch
        from . import lapw
        d = ctx.kp.get(ik)
        if d is None:
            with torch.cuda.stream(ctx.tstream):
                d = lapw.DeviceKPoint(ctx.sys, ik, device=self.dev, with_apwalm=False)
            ctx.tstream.synchronize()     # uploaded before the batcher reads it
            ctx.kp[ik] = d
        return d

    def execute(self, req: Request, ctx):
        """The request's k-points through the context's slots: every free
        slot takes the next k-point, and the setups queued together reach the
        batcher together (one grouped launch); each eigensolve starts when its
        setup is done, and its slot takes the next k-point when it returns.
        On a failure every setup and solve already started is waited for
        before the error propagates, so no GPU work is left on the slots."""
        from . import lapw
        flags = lapw.SPLIT(req.nmod) if req.nmod else 0
        ks = list(req.kpoints)
        out, dg = [None] * len(ks), [None] * len(ks)
        free = list(ctx.slots)
        inflight = {}                  # solve future -> (index, slot)
        started = []                   # every setup and solve future
        nxt = 0
        try:
            while nxt < len(ks) or inflight:
                subs = []
                while free and nxt < len(ks):
                    slot = free.pop(0)
                    d = self._kpoint(ctx, ks[nxt])
                    fut = self.batcher.submit(req, ctx, d, slot, flags)
                    started.append(fut)
                    subs.append((nxt, slot, d, fut))
                    nxt += 1
                for i, slot, d, fut in subs:
                    ready = fut.result()
                    if self.digests:
                        nn = d.rank * d.rank
                        ready.synchronize()
                        h = hashlib.sha1(slot.H[:nn].cpu().numpy().tobytes())
                        h.update(slot.O[:nn].cpu().numpy().tobytes())
                        dg[i] = h.hexdigest()
                    sf = self.solvers.submit(self._solve_on_worker, req, ready, slot.H, slot.O,
                                             d.rank, min(req.nstfv, d.rank))
                    started.append(sf)
                    inflight[sf] = (i, slot)
                done, _ = concurrent.futures.wait(
                    list(inflight), return_when=concurrent.futures.FIRST_COMPLETED)
                for sf in done:
                    i, slot = inflight.pop(sf)
                    out[i] = sf.result()
                    free.append(slot)
        except BaseException:
            concurrent.futures.wait(started)
            raise
        if self.digests:
            req.digests = dg
        return out

    def close_context(self, ctx):
        import torch
        ctx.tstream.synchronize()
        ctx.close()
        ctx.kp.clear()
        ctx.slots = None
        torch.cuda.empty_cache()

    def stats(self) -> dict:
        return self.batcher.stats()

    def footprint(self, sys_) -> int:
        """context_bytes for this executor's slot count (Server footprint)."""
        return context_bytes(sys_, slots=self.slots)

    def close(self) -> None:
        """Stop the batcher thread (the solver threads stay for the process)."""
        self.batcher.close()


def free_hbm(device: int = 0) -> int:
    import torch
    return int(torch.cuda.mem_get_info(device)[0])


def poisson_requests(n: int, rate: float, cells: Dict[str, object], seed: int = 0,
                     kmax: int = 16, split_fraction: float = 0.5, nmod: int = 9,
                     nstfv: int = 40, start: int = 0) -> List[Request]:
    """n requests with exponential inter-arrival times (``rate`` per second),
    cells uniform over ``cells``, 1..kmax distinct k-points of the cell, split
    precision with probability ``split_fraction``."""
    rng = np.random.default_rng(seed)
    names = sorted(cells)
    t = 0.0
    out = []
    for rid in range(start, start + n):
        t += rng.exponential(1.0 / rate)
        cell = names[rng.integers(len(names))]
        nk_cell = len(cells[cell].kpts)
        nk = int(rng.integers(1, min(kmax, nk_cell) + 1))
        ks = sorted(rng.choice(nk_cell, size=nk, replace=False).tolist())
        out.append(Request(rid=rid, cell=cell, kpoints=ks,
                           nmod=nmod if rng.random() < split_fraction else 0,
                           nstfv=nstfv, arrival=t))
    return out


def _torch_allocated(peak: bool = False) -> int:
    """Bytes PyTorch's allocator holds in live blocks on GPU 0 (now, or the
    peak so far): contexts' H, O and apwalm, solve temporaries, and the
    hipBLAS workspaces of the solver threads."""
    import torch
    return int(torch.cuda.max_memory_allocated(0) if peak else torch.cuda.memory_allocated(0))


def summarize(reqs: List[Request], wall: float) -> dict:
    done = [r for r in reqs if r.status == "done"]
    lat = np.array([r.latency for r in done]) if done else np.zeros(1)
    wait = np.array([r.admitted - r.arrival for r in done]) if done else np.zeros(1)
    setups = sum(len(r.kpoints) for r in done)
    return {
        "requests": len(reqs), "done": len(done),
        "failed": sum(r.status == "failed" for r in reqs),
        "cancelled": sum(r.status == "cancelled" for r in reqs),
        "setups": setups, "wall_s": wall,
        "setups_per_s": setups / wall if wall > 0 else 0.0,
        "requests_per_s": len(done) / wall if wall > 0 else 0.0,
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "max": float(lat.max())},
        "queue_wait_s": {"p50": float(np.percentile(wait, 50)),
                         "p99": float(np.percentile(wait, 99))},
        "retries_setup": int(sum(r.retries_setup for r in reqs)),
        "retries_solve": int(sum(r.retries_solve for r in reqs)),
        "launches_per_request": (float(np.mean([r.launches for r in done])) if done else 0.0),
    }


_COUNTERS = ("contexts_created", "contexts_evicted", "bypassed", "hbm_waits")


def record_line(r: Request) -> str:
    """One request as a JSON line: what summarize() needs to pool several
    segments of a stream."""
    return json.dumps({"rid": r.rid, "cell": r.cell, "nk": len(r.kpoints), "nmod": r.nmod,
                       "arrival": r.arrival, "admitted": r.admitted, "finished": r.finished,
                       "status": r.status, "retries_setup": r.retries_setup,
                       "retries_solve": r.retries_solve, "launches": r.launches}) + "\n"


def write_records(path: str, reqs: List[Request], head: dict) -> None:
    """One line per request, then the header line (segment, wall time,
    server counters: the object that has a "segment" key)."""
    with open(path, "w") as f:
        for r in reqs:
            f.write(record_line(r))
        f.write(json.dumps(head) + "\n")


def merge_records(paths: List[str]) -> dict:
    """Pool the record files of consecutive segments of one stream: latency
    percentiles over every request, setups/s over the summed segment wall
    times, server counters summed, peaks maximised."""
    heads, reqs = [], []
    for path in paths:
        with open(path) as f:
            lines = [json.loads(x) for x in f if x.strip()]
        head = [d for d in lines if "segment" in d]
        if len(head) != 1:
            raise ValueError(f"{path}: no header line (the segment did not finish)")
        heads.append(head[0])
        for d in lines:
            if "segment" in d:
                continue
            r = Request(rid=d["rid"], cell=d["cell"], kpoints=list(range(d["nk"])),
                        nmod=d["nmod"], arrival=d["arrival"])
            r.admitted, r.finished, r.status = d["admitted"], d["finished"], d["status"]
            r.retries_setup, r.retries_solve = d["retries_setup"], d["retries_solve"]
            r.launches = d["launches"]
            reqs.append(r)
    if len({r.rid for r in reqs}) != len(reqs):
        raise ValueError("segments overlap: a request id appears twice")
    out = summarize(reqs, sum(h["wall_s"] for h in heads))
    out["segments"] = [h["segment"] for h in heads]
    for k in _COUNTERS:
        out[k] = sum(h[k] for h in heads)
    out["peak_inflight"] = max(h["peak_inflight"] for h in heads)
    out["peak_context_gib"] = max(h["peak_context_gib"] for h in heads)
    out["setup_launches"] = sum(h["batching"]["launches"] for h in heads)
    out["kpoints_per_launch"] = out["setups"] / max(out["setup_launches"], 1)
    out["max_kpoints_per_launch"] = max(h["batching"]["kpoints_per_launch"]["max"]
                                        for h in heads)
    out["max_cells_per_launch"] = max(h["batching"]["cells_per_launch"]["max"] for h in heads)
    by_cell = {}
    for r in reqs:
        if r.status == "done":
            by_cell.setdefault(r.cell, []).append(r.latency)
    out["latency_p50_by_cell_s"] = {c: float(np.percentile(v, 50))
                                    for c, v in sorted(by_cell.items())}
    return out


def main(argv=None) -> int:
    """Serve a Poisson request stream over the 2/8/16/32/64-atom cells on one
    GPU and print one JSON line with p50/p99 latency and setups/s."""
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--rate", type=float, default=2.0, help="arrivals per second")
    ap.add_argument("--kmax", type=int, default=16)
    ap.add_argument("--nstfv", type=int, default=40)
    ap.add_argument("--nmod", type=int, default=9)
    ap.add_argument("--split-fraction", type=float, default=0.5)
    ap.add_argument("--max-inflight", type=int, default=8)
    ap.add_argument("--max-solves", type=int, default=1,
                    help="solver threads (the eigensolves themselves run one at a time)")
    ap.add_argument("--max-batch-k", type=int, default=32,
                    help="k-points one grouped setup launch takes at most")
    ap.add_argument("--reserve-gib", type=float, default=48.0,
                    help="HBM left outside the admission budget (library and "
                         "allocator working room)")
    ap.add_argument("--slots", type=int, default=1,
                    help="k-point slots per context (apwalm, H and O each)")
    ap.add_argument("--cells", default="fcc2,sc8,ortho16,sc32,sc64")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--segment", default="",
                    help="START:STOP: serve only these requests of the stream (arrival "
                         "times shifted so the first arrives at 0, starting from an empty "
                         "server)")
    ap.add_argument("--records", default="",
                    help="write a header and one JSON line per request to this file")
    ap.add_argument("--merge", nargs="+", default=None, metavar="RECORDS",
                    help="pool the record files of several segments, print the summary "
                         "and exit (no GPU)")
    args = ap.parse_args(argv)
    if args.merge:
        print(json.dumps(merge_records(args.merge)), flush=True)
        return 0
    from . import solve as _solve
    _solve.configure()
    from . import synth
    t = time.perf_counter()
    specs = synth.serving_cells(args.kmax)
    cells = {}
    for name in args.cells.split(","):
        spec, ks = specs[name]
        cells[name] = synth.build(spec, kpoints=ks)
    print(f"[serve] cells built in {time.perf_counter() - t:.1f} s: "
          + ", ".join(f"{k} ngk<={c.ngkmax} nk={len(c.kpts)}" for k, c in cells.items()),
          file=sys.stderr, flush=True)
    ex = GpuExecutor(0, max_solves=args.max_solves, max_batch_k=args.max_batch_k,
                     slots=args.slots)
    make, close, execute = ex.make_context, ex.close_context, ex.execute
    budget = (free_hbm(0) - int(args.reserve_gib * 2**30)
              - args.max_solves * _solve.WORKSPACE_BYTES)   # the solvers' hipBLAS workspaces
    rec = open(args.records, "w") if args.records else None
    rec_lock = threading.Lock()

    def report(r):
        print(f"[serve] request {r.rid} {r.cell} nk={len(r.kpoints)} nmod={r.nmod} "
              f"{r.status} latency {r.latency:.3f} s {r.error}", file=sys.stderr, flush=True)
        if rec is not None and r.rid >= 0:
            with rec_lock:
                rec.write(record_line(r))
                rec.flush()

    srv = Server(cells, execute, make, close, budget, max_inflight=args.max_inflight,
                 on_finish=report, footprint=ex.footprint)
    # warm-up, untimed: one context per cell, one FP64 and one split k-point
    warm = [Request(rid=-1 - i, cell=c, kpoints=[0], nmod=m, nstfv=args.nstfv)
            for i, (c, m) in enumerate((c, m) for c in cells for m in (0, args.nmod))]
    srv.run(warm)
    bad = [r for r in warm if r.status != "done"]
    if bad:
        print(f"[serve] warm-up failed: {bad[0].error}", file=sys.stderr, flush=True)
        return 1
    print("[serve] warm-up done", file=sys.stderr, flush=True)
    reqs = poisson_requests(args.requests, args.rate, cells, seed=args.seed, kmax=args.kmax,
                            split_fraction=args.split_fraction, nmod=args.nmod,
                            nstfv=args.nstfv)
    segment = [0, len(reqs)]
    if args.segment:
        a, b = (int(x) for x in args.segment.split(":"))
        if not 0 <= a < b <= len(reqs):
            raise SystemExit(f"--segment {args.segment} outside the {len(reqs)} requests")
        segment = [a, b]
        reqs = reqs[a:b]
        t0 = reqs[0].arrival
        for r in reqs:
            r.arrival -= t0
    t = time.perf_counter()
    srv.run(reqs)
    wall = time.perf_counter() - t
    out = summarize(reqs, wall)
    out.update({"workload": "Poisson stream over fcc2/sc8/ortho16/sc32/sc64 cells",
                "rate_per_s": args.rate, "kmax": args.kmax, "nstfv": args.nstfv,
                "split_fraction": args.split_fraction, "nmod": args.nmod,
                "max_inflight": args.max_inflight, "max_solves": args.max_solves,
                "budget_gib": budget / 2**30,
                "torch_allocated_gib_at_end": _torch_allocated() / 2**30,
                "torch_peak_allocated_gib": _torch_allocated(peak=True) / 2**30,
                "peak_inflight": srv.peak_inflight, "contexts_created": srv.pool.created,
                "contexts_evicted": srv.pool.evicted, "bypassed": srv.bypassed,
                "hbm_waits": srv.hbm_waits,
                "peak_context_gib": srv.peak_bytes / 2**30, "batching": ex.stats(),
                "segment": segment})
    if rec is not None:
        rec.write(json.dumps(out) + "\n")
        rec.close()
    print(json.dumps(out), flush=True)
    srv.pool.close_all()
    ex.close()
    return 0 if out["failed"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
