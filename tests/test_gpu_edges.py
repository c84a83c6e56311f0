"""Sizes at the edges of the kernels' tilings, on the GPU, against the
oracle: the basis of a k-point truncated to its first m G+k vectors (any
prefix of the G+k list is a valid basis for the setup's algebra), for m
across the 64-row tiles of the FP64 kernels and the 128/256-row tiles of the
split path (1, 2, 15, 16, 17, 63, 64, 65, 127, 128, 129, 255, 256, 257), and
m = 0 (a k-point with local orbitals only).  FP64: relative Frobenius error
<= 1e-12; split (9 moduli): exact on the truncated operands is covered
elsewhere, here within 1e-12 of the FP64 path's own relative bar at 1e-7;
the batch with all these sizes at once is bitwise the single path."""
import dataclasses

import numpy as np
import pytest

import oracle
from exciting_amd import lapw
from tests import _cases

SIZES = [0, 1, 2, 15, 16, 17, 63, 64, 65, 127, 128, 129, 255, 256, 257]


def truncated(s, m):
    k = s.kpts[0]
    assert m <= k.ngk
    keep = lambda a, axis: np.where(np.arange(a.shape[axis]).reshape(
        [-1 if i == axis else 1 for i in range(a.ndim)]) < m, a, 0).astype(a.dtype)
    kt = dataclasses.replace(k, ngk=m, igpig=keep(k.igpig, 0), vgpc=keep(k.vgpc, 1),
                             gpc=keep(k.gpc, 0), tpgpc=keep(k.tpgpc, 1),
                             sfacgp=keep(k.sfacgp, 0))
    return dataclasses.replace(s, kpts=[kt])


@pytest.fixture(scope="module")
def base():
    s = _cases.system("sc8_k0")
    s = dataclasses.replace(s, kpts=[s.kpts[0]])
    ctx = lapw.Context(0).load(s)
    yield s, ctx
    ctx.close()


def device_setup(ctx, st, flags):
    import torch
    d = lapw.DeviceKPoint(st, 0)
    n = d.rank
    H = torch.zeros(max(n * n, 1), dtype=torch.complex128, device="cuda")
    O = torch.zeros_like(H)
    ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, d.apwalm)
    ctx.hamiltonandoverlapsetup(lapw.EvSystem(lapw.HermitianMatrix(n, H),
                                              lapw.HermitianMatrix(n, O)),
                                d.ngk, d.apwalm, d.igpig, d.vgpc, flags=flags)
    torch.cuda.synchronize()
    return d, H[:n * n].cpu().numpy().reshape(n, n).T, O[:n * n].cpu().numpy().reshape(n, n).T


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.gpu
def test_fp64_and_split_at_tile_edges(base):
    s, ctx = base
    for m in SIZES:
        st = truncated(s, m)
        print(f"[gpu-edges] m = {m}", flush=True)
        _, H, O = device_setup(ctx, st, 0)
        S = oracle.System(st)
        A = oracle.match(S, 0)
        H0, O0 = oracle.setup(S, 0, A)
        eh, eo = rel(H, H0), rel(O, O0)
        assert eh <= 1e-12 and eo <= 1e-12, (m, eh, eo)
        if m > 0:
            _, Hs, Os = device_setup(ctx, st, lapw.SPLIT(9))
            es, eso = rel(Hs, H0), rel(Os, O0)
            assert es <= 1e-7 and eso <= 1e-7, (m, es, eso)
        print(f"  m = {m}: fp64 H {eh:.1e} O {eo:.1e}", flush=True)


@pytest.mark.gpu
def test_batch_of_edge_sizes_is_the_single_path(base):
    import torch
    s, ctx = base
    sts = [truncated(s, m) for m in SIZES]
    dks = [lapw.DeviceKPoint(st, 0) for st in sts]
    ref = []
    for st in sts:
        _, H, O = device_setup(ctx, st, 0)
        ref.append((H, O))
    for plan in (-1, lapw.plan_count() - 1):
        ctx.batch_plan = plan
        Hs = [torch.zeros(max(d.rank ** 2, 1), dtype=torch.complex128, device="cuda")
              for d in dks]
        Os = [torch.zeros_like(h) for h in Hs]
        kp = [lapw.kpoint_struct(d, h, o) for d, h, o in zip(dks, Hs, Os)]
        ctx.setup_batch(kp, flags=lapw.MATCH)
        torch.cuda.synchronize()
        for m, d, h, o, (H, O) in zip(SIZES, dks, Hs, Os, ref):
            n = d.rank
            hb = h[:n * n].cpu().numpy().reshape(n, n).T
            ob = o[:n * n].cpu().numpy().reshape(n, n).T
            assert np.array_equal(hb, H) and np.array_equal(ob, O), (plan, m)
    ctx.batch_plan = -1
