"""One grouped launch over k-points of several structures
(exh_setup_multi_device).

The bar: per k-point the multi-structure batch gives bitwise the same
apwalm, H and O as the single-k device entry points on the k-point's own
context (exh_match_device + exh_setup_device, the path test_gpu_parity pins
to the oracle at 1e-12), for several plans of the grouped H/O kernel, FP64
and split precision, when the structures differ in lmaxapw, atom count,
ngkmax, APW layout (packed operand for lmaxmat < lmaxapw) and relativity
(iora* a2t stage), when a context's workspace splits its k-points over
several groups, and on graph replays.  The profile counters show that the
k-points of all structures shared one launch per stage.
"""
import numpy as np
import pytest

from exciting_amd import lapw
from tests.test_gpu_batch import Case, same

# FP64 mix: lmaxapw 3 and 4 (tiny cells, one with lmaxmat < lmaxapw and a
# packed A^T, one iora* with the a2t stage), lmaxapw 8 (fcc2) and 10 (sc8)
FP64_MIX = ("tiny_multi", "tiny_lmaxmat_multi", "tiny_rel_multi", "fcc2", "sc8_4k")
# split precision does not take iora* (h1aa)
SPLIT_MIX = ("tiny_multi", "tiny_lmaxmat_multi", "fcc2", "sc8_4k")


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


def interleave(cs):
    """(case, k index) pairs, round robin over the cases."""
    order, i = [], 0
    while any(i < len(c.dk) for c in cs):
        for c in cs:
            if i < len(c.dk):
                order.append((c, i))
        i += 1
    return order


def multi(cs, flags, stream=None):
    """All k-points of the cases, interleaved, in one exh_setup_multi_device
    call; returns per case the (apwalm, H, O) host arrays of its k-points."""
    import torch
    bufs = {c.name: c.fresh() for c in cs}
    order = interleave(cs)
    kp, ctxs = [], []
    for c, ik in order:
        aps, hs, os_ = bufs[c.name]
        kp.append(lapw.kpoint_struct(c.dk[ik], hs[ik], os_[ik], apwalm=aps[ik]))
        ctxs.append(c.ctx)
    lapw.setup_multi(ctxs, kp, flags=flags | lapw.MATCH, stream=stream)
    torch.cuda.synchronize()
    return {c.name: [(a.cpu().numpy(), h.cpu().numpy(), o.cpu().numpy())
                     for a, h, o in zip(*bufs[c.name])] for c in cs}


@pytest.mark.gpu
@pytest.mark.parametrize("plan", [-1, 5, 10, 19])
def test_multi_structure_batch_is_bitwise_the_single_path(cases, plan):
    cs = [cases(n) for n in FP64_MIX]
    lead = cs[0].ctx
    lead.batch_plan = plan
    try:
        out = multi(cs, 0)
    finally:
        lead.batch_plan = -1
    for c in cs:
        same(out[c.name], c.single(0), f"{c.name} plan {plan}")
    print(f"  plan {plan}: {sum(len(c.dk) for c in cs)} k-points of {len(cs)} structures "
          f"bitwise equal")


@pytest.mark.gpu
@pytest.mark.parametrize("nmod", [7, 9])
def test_multi_structure_split_is_bitwise_the_single_path(cases, nmod):
    cs = [cases(n) for n in SPLIT_MIX]
    out = multi(cs, lapw.SPLIT(nmod))
    for c in cs:
        same(out[c.name], c.single(lapw.SPLIT(nmod)), f"{c.name} split {nmod}")


@pytest.mark.gpu
def test_multi_structure_accumulates(cases):
    cs = [cases(n) for n in ("tiny_multi", "sc8_4k")]
    rng = np.random.default_rng(9)
    import torch
    pre = {}
    for c in cs:
        pre[c.name] = []
        for d in c.dk:
            m = [rng.standard_normal((d.rank, d.rank)) + 1j * rng.standard_normal((d.rank, d.rank))
                 for _ in range(2)]
            pre[c.name].append([x + x.conj().T for x in m])
    bufs = {c.name: c.fresh(prefill=pre[c.name]) for c in cs}
    kp, ctxs = [], []
    for c, ik in interleave(cs):
        aps, hs, os_ = bufs[c.name]
        kp.append(lapw.kpoint_struct(c.dk[ik], hs[ik], os_[ik], apwalm=aps[ik]))
        ctxs.append(c.ctx)
    lapw.setup_multi(ctxs, kp, flags=lapw.ACCUMULATE | lapw.MATCH)
    torch.cuda.synchronize()
    for c in cs:
        got = [(a.cpu().numpy(), h.cpu().numpy(), o.cpu().numpy()) for a, h, o in zip(*bufs[c.name])]
        same(got, c.single(lapw.ACCUMULATE, prefill=pre[c.name]), f"{c.name} accumulate")


@pytest.mark.gpu
def test_multi_structure_shares_one_launch_per_stage(cases):
    """With every context's k-points fitting its workspace, each stage is
    one launch for all structures (the match stage one kernel per lmaxapw
    present)."""
    cs = [cases(n) for n in ("tiny_multi", "fcc2", "sc8_4k")]
    lead = cs[0].ctx
    lead.profile(True)
    lead.profile_reset()
    try:
        multi(cs, 0)
        prof = lead.profile_read()
    finally:
        lead.profile(False)
    n = {k: v[1] for k, v in prof.items() if v[1]}
    print(f"  launches: {n}")
    assert len({c.s.lmaxapw for c in cs}) == 3
    assert n["ho_grouped"] == 1 and n["bt"] == 1 and n["ivgk"] == 1
    assert n["lo_border"] == 1 and n["lolo"] == 1
    # one profiled match stage: inside it one kernel per lmaxapw present
    assert n["match"] == 1
    assert "ho" not in n      # no single-k contraction ran


@pytest.mark.gpu
def test_multi_structure_groups_by_each_contexts_workspace(cases):
    cs = [cases(n) for n in ("tiny_multi", "sc8_4k")]
    ref = {c.name: c.single(0) for c in cs}
    cs[1].ctx.set_batch_workspace(1)        # one sc8 k-point per group
    try:
        out = multi(cs, 0)
    finally:
        cs[1].ctx.set_batch_workspace(8 << 30)
    for c in cs:
        same(out[c.name], ref[c.name], f"{c.name} grouped by workspace")


@pytest.mark.gpu
def test_multi_structure_graph_replays(cases):
    import torch
    cs = [cases(n) for n in ("tiny_multi", "fcc2")]
    eager = multi(cs, 0)
    bufs = {c.name: c.fresh() for c in cs}
    kp, ctxs = [], []
    for c, ik in interleave(cs):
        aps, hs, os_ = bufs[c.name]
        kp.append(lapw.kpoint_struct(c.dk[ik], hs[ik], os_[ik], apwalm=aps[ik]))
        ctxs.append(c.ctx)
    for rep in range(3):
        for c in cs:
            for h in bufs[c.name][1] + bufs[c.name][2]:
                h.fill_(rep + 1.0)
        x = torch.randn(257, 129, device="cuda").sum()   # unrelated eager work
        torch.cuda.synchronize()
        assert float(x) == float(x)
        lapw.setup_multi(ctxs, kp, flags=lapw.MATCH | lapw.GRAPH)
        torch.cuda.synchronize()
        for c in cs:
            got = [(a.cpu().numpy(), h.cpu().numpy(), o.cpu().numpy())
                   for a, h, o in zip(*bufs[c.name])]
            same(got, eager[c.name], f"{c.name} replay {rep}")


@pytest.mark.gpu
def test_multi_structure_rejects_bad_arguments(cases):
    cs = [cases(n) for n in ("tiny_multi", "fcc2")]
    aps, hs, os_ = cs[0].fresh()
    good = lapw.kpoint_struct(cs[0].dk[0], hs[0], os_[0], apwalm=aps[0])
    a1, h1, o1 = cs[1].fresh()
    other = lapw.kpoint_struct(cs[1].dk[0], h1[0], o1[0], apwalm=a1[0])
    bad = lapw.kpoint_struct(cs[1].dk[0], h1[0], o1[0], apwalm=a1[0])
    bad.rank += 1       # checked against its own context's nlotot
    with pytest.raises(lapw.ExhError, match="rank"):
        lapw.setup_multi([cs[0].ctx, cs[1].ctx], [good, bad], flags=0)
    with pytest.raises(lapw.ExhError, match="SPLIT"):
        lapw.setup_multi([cs[0].ctx, cs[1].ctx], [good, other], flags=lapw.SPLIT(3))
    with pytest.raises(ValueError):
        lapw.setup_multi([cs[0].ctx], [good, good], flags=0)
