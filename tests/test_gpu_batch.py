"""Batched k-points on the GPU: grouped launches, persistent H/O kernel,
plans, autotuner, HIP-graph replay.

The bar: per k-point a batch must give bitwise the same apwalm, H and O as
the single-k device entry points (exh_match_device + exh_setup_device with
the same flags), for every plan of the grouped kernel, every group size,
FP64 and split precision, accumulate and overwrite, and on every replay of a
captured graph.  The single-k path itself is pinned to the oracle at 1e-12
by test_gpu_parity; the FP64 batch is also compared with the oracle here.
"""
import numpy as np
import pytest

import oracle
from exciting_amd import lapw
from tests import _cases


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


class Case:
    def __init__(self, name):
        import torch
        self.torch = torch
        self.name = name
        self.s = _cases.system(name)
        print(f"[gpu-batch] case {name}: ngk {[k.ngk for k in self.s.kpts]}", flush=True)
        self.ctx = lapw.Context(0).load(self.s)
        self.dk = [lapw.DeviceKPoint(self.s, ik) for ik in range(len(self.s.kpts))]
        self.ref = {}

    def fresh(self, prefill=None):
        t = self.torch
        hs, os_ = [], []
        for i, d in enumerate(self.dk):
            if prefill is None:
                h = t.zeros(d.rank * d.rank, dtype=t.complex128, device="cuda")
                o = t.zeros_like(h)
            else:
                h, o = (t.from_numpy(x.ravel(order="F").copy()).cuda() for x in prefill[i])
            hs.append(h)
            os_.append(o)
        aps = [t.zeros_like(d.apwalm) for d in self.dk]
        return aps, hs, os_

    def single(self, flags, prefill=None):
        """Per-k device path: match_device + setup_device."""
        key = (flags, prefill is not None)
        if key in self.ref:
            return self.ref[key]
        aps, hs, os_ = self.fresh(prefill)
        for d, a, h, o in zip(self.dk, aps, hs, os_):
            self.ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, a)
            sysd = lapw.EvSystem(lapw.HermitianMatrix(d.rank, h), lapw.HermitianMatrix(d.rank, o))
            self.ctx.hamiltonandoverlapsetup(sysd, d.ngk, a, d.igpig, d.vgpc, flags=flags)
        self.torch.cuda.synchronize()
        out = [(a.cpu().numpy(), h.cpu().numpy(), o.cpu().numpy())
               for a, h, o in zip(aps, hs, os_)]
        self.ref[key] = out
        return out

    def batch(self, flags, prefill=None, bufs=None):
        aps, hs, os_ = bufs if bufs is not None else self.fresh(prefill)
        kp = [lapw.kpoint_struct(d, h, o, apwalm=a) for d, a, h, o in zip(self.dk, aps, hs, os_)]
        self.ctx.setup_batch(kp, flags=flags | lapw.MATCH)
        self.torch.cuda.synchronize()
        return [(a.cpu().numpy(), h.cpu().numpy(), o.cpu().numpy())
                for a, h, o in zip(aps, hs, os_)]


def same(out, ref, what=""):
    for ik, ((a, h, o), (a0, h0, o0)) in enumerate(zip(out, ref)):
        assert np.array_equal(a, a0), f"{what} k{ik}: apwalm differs"
        assert np.array_equal(h, h0), f"{what} k{ik}: H differs (max {np.abs(h - h0).max():.2e})"
        assert np.array_equal(o, o0), f"{what} k{ik}: O differs (max {np.abs(o - o0).max():.2e})"


@pytest.fixture(scope="module")
def cases():
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = Case(name)
        return cache[name]
    yield get
    for c in cache.values():
        c.ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tiny_multi", "tiny_lmaxmat_multi", "tiny_rel_multi", "sc8_4k"])
def test_batch_is_bitwise_the_single_path_for_every_plan(cases, name):
    c = cases(name)
    ref = c.single(0)
    for plan in range(lapw.plan_count()):
        c.ctx.batch_plan = plan
        same(c.batch(0), ref, f"plan {plan} {lapw.plan_info(plan)}")
    c.ctx.batch_plan = -1


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tiny_multi", "sc8_4k"])
def test_batch_accumulates_like_the_single_path(cases, name):
    c = cases(name)
    rng = np.random.default_rng(5)
    pre = []
    for d in c.dk:
        m = [rng.standard_normal((d.rank, d.rank)) + 1j * rng.standard_normal((d.rank, d.rank))
             for _ in range(2)]
        pre.append([x + x.conj().T for x in m])
    same(c.batch(lapw.ACCUMULATE, prefill=pre), c.single(lapw.ACCUMULATE, prefill=pre),
         "accumulate")


@pytest.mark.gpu
@pytest.mark.parametrize("name,nmod", [("tiny_multi", 9), ("tiny_lmaxmat_multi", 7),
                                       ("sc8_4k", 9)])
def test_batch_split_is_bitwise_the_single_path(cases, name, nmod):
    c = cases(name)
    same(c.batch(lapw.SPLIT(nmod)), c.single(lapw.SPLIT(nmod)), f"split {nmod}")


@pytest.mark.gpu
def test_batch_groups_by_workspace(cases):
    c = cases("tiny_multi")
    ref = c.single(0)
    for nbytes in (1, 2 * 1024 * 1024):   # one k-point per group, then a few
        c.ctx.set_batch_workspace(nbytes)
        same(c.batch(0), ref, f"workspace {nbytes}")
    c.ctx.set_batch_workspace(8 << 30)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tiny_multi", "sc8_4k"])
def test_batch_matches_the_oracle(cases, name):
    c = cases(name)
    S = oracle.System(c.s)
    out = c.batch(0)
    for ik, (a, h, o) in enumerate(out):
        A = oracle.match(S, ik)
        H0, O0 = oracle.setup(S, ik, A)
        n = c.dk[ik].rank
        H = h.reshape(n, n).T
        O = o.reshape(n, n).T
        eh, eo = rel(H, H0), rel(O, O0)
        print(f"  {name} k{ik}: H {eh:.1e} O {eo:.1e}")
        assert eh <= 1e-12 and eo <= 1e-12


@pytest.mark.gpu
def test_graph_replays_match_eager(cases):
    """Replays of a captured batch, with unrelated eager work and
    synchronisation between them, each equal the eager batch; after new
    tables the next call captures again and follows them."""
    c = cases("tiny_multi")
    t = c.torch
    eager = c.batch(0)
    bufs = c.fresh()
    for rep in range(3):
        for h in bufs[1] + bufs[2]:
            h.fill_(rep + 1.0)           # replays overwrite whatever is there
        x = t.randn(257, 129, device="cuda").sum()   # unrelated eager work
        t.cuda.synchronize()
        assert float(x) == float(x)
        same(c.batch(lapw.GRAPH, bufs=bufs), eager, f"replay {rep}")
    s = c.s
    aa2 = s.aa * 0.75
    c.ctx.set_mt(s.maxaa, s.maxnlo, aa2, s.loa, s.lolo, s.oalo, s.ololo, s.h1aa, s.h1loa,
                 s.h1lolo)
    try:
        # tables rewritten in place: the replay reads the new values
        new_eager = c.batch(0)
        got = c.batch(lapw.GRAPH, bufs=bufs)
        same(got, new_eager, "after set_mt in place")
        assert not np.array_equal(got[0][1], eager[0][1])
        # an h1aa table switches the a2t stage on: the graph must be recaptured
        rng = np.random.default_rng(2)
        h1aa = 0.01 * rng.standard_normal((s.apwordmax, s.apwordmax, s.lmaxapw + 1, s.natmtot))
        h1loa = 0.01 * rng.standard_normal((s.apwordmax, s.nlomax, s.natmtot))
        h1lolo = np.zeros((s.nlomax, s.nlomax, s.natmtot))
        c.ctx.set_mt(s.maxaa, s.maxnlo, s.aa, s.loa, s.lolo, s.oalo, s.ololo, h1aa, h1loa,
                     h1lolo)
        rel_eager = c.batch(0)
        same(c.batch(lapw.GRAPH, bufs=bufs), rel_eager, "after h1aa switched on")
        assert not np.array_equal(rel_eager[0][2], eager[0][2])
    finally:
        c.ctx.set_mt(s.maxaa, s.maxnlo, s.aa, s.loa, s.lolo, s.oalo, s.ololo, s.h1aa,
                     s.h1loa, s.h1lolo)
        c.ref.clear()


@pytest.mark.gpu
def test_tuner_picks_a_plan_and_keeps_results(cases):
    c = cases("sc8_4k")
    aps, hs, os_ = c.fresh()
    kp = [lapw.kpoint_struct(d, h, o, apwalm=a) for d, a, h, o in zip(c.dk, aps, hs, os_)]
    best = c.ctx.tune_batch(kp, flags=lapw.MATCH)
    print(f"  tuned plan {best}: {lapw.plan_info(best)}")
    assert 0 <= best < lapw.plan_count()
    assert c.ctx.batch_plan == best
    same(c.batch(0), c.single(0), "tuned")
    c.ctx.batch_plan = -1


@pytest.mark.gpu
def test_batch_rejects_bad_arguments(cases):
    c = cases("tiny_multi")
    aps, hs, os_ = c.fresh()
    kp = [lapw.kpoint_struct(d, h, o, apwalm=a) for d, a, h, o in zip(c.dk, aps, hs, os_)]
    bad = lapw.kpoint_struct(c.dk[0], hs[0], os_[0], apwalm=aps[0])
    bad.rank += 1
    with pytest.raises(lapw.ExhError, match="rank"):
        c.ctx.setup_batch([kp[1], bad], flags=0)
    bad = lapw.kpoint_struct(c.dk[0], hs[0], os_[0], apwalm=aps[0])
    bad.ngp = c.s.ngkmax + 1
    bad.rank = bad.ngp + c.s.nlotot
    with pytest.raises(lapw.ExhError, match="ngkmax"):
        c.ctx.setup_batch([bad], flags=0)
    with pytest.raises(lapw.ExhError, match="SPLIT"):
        c.ctx.setup_batch(kp, flags=lapw.SPLIT(3))
    with pytest.raises(lapw.ExhError):
        c.ctx.batch_plan = lapw.plan_count()


@pytest.mark.gpu
def test_graph_replays_follow_device_made_tables(cases):
    """A captured batch keeps replaying correctly when the next SCF
    iteration's tables are made and installed on the device (hmlint +
    exh_set_mt_device, gencfun/genveffig + exh_set_interstitial_device):
    same sizes, rewritten in place, so the graph reads the new values."""
    import dataclasses
    from exciting_amd import synth
    c = cases("tiny_multi")
    t = c.torch
    s = synth.add_mt_potential(dataclasses.replace(c.s), lmaxvr=4, seed=11)
    dev = lambda a: t.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).cuda()
    bufs = c.fresh()
    c.batch(lapw.GRAPH, bufs=bufs)                  # capture with the loaded tables
    mnlo = max(s.maxnlo, 1)
    outs = [t.zeros(n, dtype=t.complex128, device="cuda")
            for n in (s.maxaa * s.maxaa * s.natmtot, mnlo * s.maxaa * s.natmtot,
                      mnlo * mnlo * s.natmtot)]
    try:
        c.ctx.hmlint(s.lmaxvr, dev(s.veffmt), dev(s.lofr), True, *outs)
        c.ctx.set_mt(s.maxaa, s.maxnlo, *outs, dev(s.oalo), dev(s.ololo))
        cg = t.zeros(s.ngrtot, dtype=t.complex128, device="cuda")
        cr = t.zeros(s.ngrtot, dtype=t.float64, device="cuda")
        c.ctx.gencfun(s.bvec, s.atposc, cfunig=cg, cfunir=cr)
        vg = t.zeros(s.ngvec, dtype=t.complex128, device="cuda")
        veffir = dev(-0.4 + 0.3 * np.random.default_rng(5).uniform(-1, 1, s.ngrtot))
        c.ctx.genveffig_device(veffir, cr, veffig=vg)
        c.ctx.set_interstitial(cg, vg, valence_rel_none=True)
        eager = c.batch(0)
        replay = c.batch(lapw.GRAPH, bufs=bufs)
        same(replay, eager, "replay after device-made tables")
    finally:
        c.ctx.set_mt(c.s.maxaa, c.s.maxnlo, c.s.aa, c.s.loa, c.s.lolo, c.s.oalo, c.s.ololo,
                     c.s.h1aa, c.s.h1loa, c.s.h1lolo)
        c.ctx.set_interstitial(c.s.cfunig, c.s.veffig, c.s.meffig, c.s.m2effig,
                               c.s.valence_rel_none, c.s.h1on)
        c.ref.clear()
