"""exh_setup_device_noalloc: the single-k setup with its scratch in the
library's static per-GPU array instead of allocated buffers (for callers
that cannot allocate: graph capture, tight serving loops).

Through the C ABI, two calls in a row -- the second with a smaller basis, so
that stale scratch from the first call is in the array -- each give bitwise
the H and O of exh_setup_device on the same inputs (same kernels, same
accumulation order), overwrite and accumulate, for the uniform layout, a
non-uniform one (lmaxmat < lmaxapw: the packed operand) and iora* (the
(1+h1).A operand).  The larger call is also within 1e-12 of the oracle.
Bases beyond the array and split precision are refused.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from exciting_amd import lapw
from tests import _cases
from tests.test_gpu_edges import truncated


def log(msg):
    print(msg, flush=True)


def run(ctx, d, noalloc, flags, prefill=None):
    """One setup through the C ABI on device tensors; returns (H, O) as
    flat complex128 CUDA tensors (column-major memory)."""
    import torch
    n = d.rank
    if prefill is None:
        H = torch.full((n * n,), float("nan"), dtype=torch.complex128, device="cuda")
        O = torch.full_like(H, float("nan"))
    else:
        H, O = prefill[0].clone(), prefill[1].clone()
    fn = lapw.lib().exh_setup_device_noalloc if noalloc else lapw.lib().exh_setup_device
    rc = ctx._device_call(None, lambda st: fn(ctx._h, st, d.ngk, n, d.apwalm.data_ptr(),
                                              d.igpig.data_ptr(), d.vgpc.data_ptr(),
                                              H.data_ptr(), O.data_ptr(), flags))
    ctx._check(rc, "noalloc" if noalloc else "setup_device")
    torch.cuda.synchronize()
    return H, O


def prepared(ctx, s):
    d = lapw.DeviceKPoint(s, 0)
    ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, d.apwalm)
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("name,smaller", [("sc8_k0", 700), ("tiny_lmaxmat", 20),
                                          ("tiny_rel", 17)])
def test_noalloc_equals_setup_device_on_two_calls(name, smaller):
    import torch
    s = _cases.system(name)
    s = dataclasses.replace(s, kpts=[s.kpts[0]])
    log(f"[noalloc] case {name}: ngk {s.kpts[0].ngk}, then {smaller}")
    ctx = lapw.Context(0).load(s)
    try:
        for st in (s, truncated(s, smaller)):
            d = prepared(ctx, st)
            n = d.rank
            for flags in (0, lapw.ACCUMULATE):
                pre = None
                if flags:
                    g = torch.Generator(device="cuda").manual_seed(n)
                    pre = [torch.randn(n * n, dtype=torch.complex128, device="cuda",
                                       generator=g) for _ in range(2)]
                Hn, On = run(ctx, d, True, flags, pre)
                Hd, Od = run(ctx, d, False, flags, pre)
                log(f"  ngk {d.ngk} flags {flags}: H equal {torch.equal(Hn, Hd)}, "
                    f"O equal {torch.equal(On, Od)}")
                assert torch.equal(Hn, Hd) and torch.equal(On, Od)
            if st is s:
                S = oracle.System(st)
                H0, O0 = oracle.setup(S, 0, oracle.match(S, 0))
                Hz, Oz = run(ctx, d, True, 0)
                H = Hz.cpu().numpy().reshape(n, n).T
                O = Oz.cpu().numpy().reshape(n, n).T
                eh = np.linalg.norm(H - H0) / np.linalg.norm(H0)
                eo = np.linalg.norm(O - O0) / np.linalg.norm(O0)
                log(f"  vs oracle: H {eh:.1e} O {eo:.1e}")
                assert eh <= 1e-12 and eo <= 1e-12
    finally:
        ctx.close()


@pytest.mark.gpu
def test_noalloc_refuses_what_it_cannot_hold():
    s = _cases.system("ortho16_2k")        # 3872 x 3392 operand: 210 MB > 64 MiB
    ctx = lapw.Context(0).load(s)
    try:
        d = prepared(ctx, s)
        with pytest.raises(lapw.ExhError, match="static buffer holds"):
            run(ctx, d, True, 0)
    finally:
        ctx.close()
    s = _cases.system("tiny")
    ctx = lapw.Context(0).load(s)
    try:
        d = prepared(ctx, s)
        with pytest.raises(lapw.ExhError, match="split precision"):
            run(ctx, d, True, lapw.SPLIT(9))
    finally:
        ctx.close()
