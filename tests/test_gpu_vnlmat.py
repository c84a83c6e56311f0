"""overlapsetup alone and the hybrid non-local matrix (calc_vnlmat) on the
GPU, against the oracle.

* exh_overlapsetup(_device): O within 1e-12 (relative Frobenius) of the
  oracle's overlapsetup (overlapsetup.f90 + its hamiltonoverlapocopy_UL),
  bitwise equal to the O of the combined setup (the same tile products and
  order), for 'none', iora* (h1 terms in O) and the 8-atom cell; H is only
  mirrored, as the reference's copy does.
* calc_vnlmat (exciting_amd.hybrids): within 1e-12 of the reference's three
  products applied by NumPy to the oracle's O.
"""
import numpy as np
import pytest

import oracle
from exciting_amd import hybrids, lapw
from tests import _cases


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def ctxs():
    cache = {}

    def get(name):
        if name not in cache:
            s = _cases.system(name)
            print(f"[gpu-vnlmat] case {name}", flush=True)
            cache[name] = (s, lapw.Context(0).load(s))
        return cache[name]
    yield get
    for _, c in cache.values():
        c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tiny", "tiny_rel", "sc8_k0"])
def test_overlapsetup_alone(ctxs, name):
    import torch
    s, ctx = ctxs(name)
    d = lapw.DeviceKPoint(s, 0)
    n = d.rank
    ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, d.apwalm)
    # combined setup for reference
    H = torch.zeros(n * n, dtype=torch.complex128, device="cuda")
    O = torch.zeros_like(H)
    ctx.hamiltonandoverlapsetup(lapw.EvSystem(lapw.HermitianMatrix(n, H),
                                              lapw.HermitianMatrix(n, O)),
                                d.ngk, d.apwalm, d.igpig, d.vgpc, flags=0)
    # overlapsetup alone, onto a random O (accumulate) and a random H (mirror only)
    rng = np.random.default_rng(3)
    o0 = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    o0 = o0 + o0.conj().T
    h0 = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    Ht = torch.from_numpy(h0.ravel(order="F").copy()).cuda()
    Ot = torch.from_numpy(o0.ravel(order="F").copy()).cuda()
    ctx.overlapsetup(lapw.EvSystem(lapw.HermitianMatrix(n, Ht), lapw.HermitianMatrix(n, Ot)),
                     d.ngk, d.apwalm, d.igpig, d.vgpc, flags=lapw.ACCUMULATE)
    Oz = torch.zeros_like(O)
    ctx.overlapsetup(lapw.EvSystem(None, lapw.HermitianMatrix(n, Oz)), d.ngk, d.apwalm,
                     d.igpig, d.vgpc, flags=0)
    torch.cuda.synchronize()
    Oc = O.cpu().numpy().reshape(n, n).T
    Oa = Oz.cpu().numpy().reshape(n, n).T
    assert np.array_equal(Oa, Oc), "O alone differs from the combined setup's O"
    S = oracle.System(s)
    A = oracle.match(S, 0)
    hr, Or = oracle.overlapsetup(S, 0, A, h=h0.copy(order="F"))
    e = rel(Oa, Or)
    print(f"  {name}: O {e:.1e}")
    assert e <= 1e-12
    Oacc = Ot.cpu().numpy().reshape(n, n).T
    upper = np.triu(np.ones((n, n), bool))
    assert np.abs((Oacc - o0 - Oa)[upper]).max() <= 1e-12 * np.abs(Oa).max()
    Hm = Ht.cpu().numpy().reshape(n, n).T
    assert np.array_equal(Hm, hr)           # upper kept, lower = conj(upper)
    assert np.array_equal(np.triu(Hm), np.triu(h0))
    # host entry point
    hh, oh = h0.copy(order="F"), o0.copy(order="F")
    apw = d.apwalm.cpu().numpy().reshape(d.apwalm_shape, order="F")
    ctx.overlapsetup(lapw.EvSystem(lapw.HermitianMatrix(n, hh), lapw.HermitianMatrix(n, oh)),
                     d.ngk, apw, s.kpts[0].igpig, s.kpts[0].vgpc)
    assert rel(oh, o0 + Oa) <= 1e-13 and np.array_equal(hh, hr)


@pytest.mark.gpu
@pytest.mark.parametrize("name,nst", [("tiny", 12), ("sc8_k0", 64)])
def test_calc_vnlmat(ctxs, name, nst):
    import torch
    s, ctx = ctxs(name)
    d = lapw.DeviceKPoint(s, 0)
    n = d.rank
    rng = np.random.default_rng(7)
    X = rng.normal(size=(n, nst)) + 1j * rng.normal(size=(n, nst))
    V = rng.normal(size=(nst, nst)) + 1j * rng.normal(size=(nst, nst))
    V = V + V.conj().T
    evec = torch.from_numpy(X.T.copy()).cuda()            # row s = state s
    vx = torch.from_numpy(V).cuda()
    got = hybrids.vnlmat_kpoint(ctx, d, evec, vx)
    torch.cuda.synchronize()
    G = got.cpu().numpy().reshape(n, n).T
    S = oracle.System(s)
    A = oracle.match(S, 0)
    _, O = oracle.overlapsetup(S, 0, A)
    want = oracle.calc_vnlmat(O, X, V)
    e = rel(G, want)
    print(f"  {name}: vnlmat {e:.1e}")
    assert e <= 1e-12
    assert np.abs(G - G.conj().T).max() <= 1e-10 * np.abs(G).max()
