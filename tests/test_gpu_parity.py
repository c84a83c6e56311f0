"""HIP path vs the oracle on identical seeded inputs (through the C ABI).

Tolerances (the FP64 drop-in bar): relative Frobenius error <= 1e-12 for
apwalm, H and O; <= 1e-8 Ha for the eigenvalues of H x = e O x.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from exciting_amd import lapw
from tests import _cases

TOL_MAT = 1e-12
TOL_EIG = 1e-8

CASES = ["tiny", "tiny_lmaxmat", "tiny_p3", "tiny_nolo", "tiny_rel", "tiny_zora", "fcc2",
         "sc8_k0"]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ctxs():
    cache = {}

    def get(name):
        if name not in cache:
            s = _cases.system(name)
            print(f"[gpu-parity] case {name}: natmtot={s.natmtot} ngkmax={s.ngkmax} "
                  f"nlotot={s.nlotot}", flush=True)
            cache[name] = (s, oracle.System(s), lapw.Context(0).load(s))
        return cache[name]
    yield get
    for s, S, c in cache.values():
        c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_match(ctxs, name):
    s, S, ctx = ctxs(name)
    for ik, k in enumerate(s.kpts):
        ref = oracle.match(S, ik)
        got = np.zeros_like(ref, order="F")
        ctx.match(k.ngk, k.gpc, k.tpgpc, k.sfacgp, got)
        e = rel(got, ref)
        print(f"  match {name} k{ik}: rel {e:.2e}")
        assert e <= TOL_MAT


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_setup(ctxs, name):
    s, S, ctx = ctxs(name)
    for ik, k in enumerate(s.kpts):
        A = oracle.match(S, ik)
        H0, O0 = oracle.setup(S, ik, A)
        sysg = lapw.newsystem(k.ngk + s.nlotot)
        ctx.hamiltonandoverlapsetup(sysg, k.ngk, A, k.igpig, k.vgpc)
        H, O = sysg.hamilton.za, sysg.overlap.za
        eh, eo = rel(H, H0), rel(O, O0)
        print(f"  setup {name} k{ik}: H {eh:.2e} O {eo:.2e}")
        assert eh <= TOL_MAT and eo <= TOL_MAT
        n = H.shape[0]
        nst = min(n, 40)
        w0 = sl.eigh(H0, O0, lower=False, eigvals_only=True, subset_by_index=[0, nst - 1])
        w = sl.eigh(H, O, lower=False, eigvals_only=True, subset_by_index=[0, nst - 1])
        de = np.abs(w - w0).max()
        print(f"  eig {name} k{ik}: max |de| {de:.2e} Ha")
        assert de <= TOL_EIG


@pytest.mark.gpu
def test_accumulate_semantics(ctxs):
    """With EXH_ACCUMULATE the upper triangle and the lo rows under the APW
    block keep what was there (reference zgemm beta=1 / += semantics)."""
    s, S, ctx = ctxs("tiny")
    k = s.kpts[0]
    n = k.ngk + s.nlotot
    A = oracle.match(S, 0)
    rng = np.random.default_rng(0)
    H0 = np.asfortranarray(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    O0 = np.asfortranarray(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    Href, Oref = H0.copy(order="F"), O0.copy(order="F")
    import ctypes
    L = oracle.lib()
    ig = np.ascontiguousarray(k.igpig)
    vg = np.asfortranarray(k.vgpc)
    rc = L.orc_hamiltonandoverlapsetup(S.ptr, k.ngk, A.ctypes.data, ig.ctypes.data,
                                       vg.ctypes.data, n, Href.ctypes.data, Oref.ctypes.data,
                                       oracle.scipy_zgemm())
    assert rc == 0
    sysg = lapw.EvSystem(lapw.HermitianMatrix(n, H0.copy(order="F")),
                         lapw.HermitianMatrix(n, O0.copy(order="F")))
    ctx.hamiltonandoverlapsetup(sysg, k.ngk, A, k.igpig, k.vgpc, flags=lapw.ACCUMULATE)
    assert rel(sysg.hamilton.za, Href) <= TOL_MAT
    assert rel(sysg.overlap.za, Oref) <= TOL_MAT


@pytest.mark.gpu
def test_device_resident_path(ctxs):
    """match + setup on torch device tensors, no host copies in between."""
    import torch
    s, S, ctx = ctxs("sc8_k0")
    for ik, k in enumerate(s.kpts):
        dk = lapw.DeviceKPoint(s, ik)
        ctx.match(dk.ngk, dk.gpc, dk.tpgpc, dk.sfacgp, dk.apwalm)
        sysd = lapw.newsystem(dk.rank, device="cuda")
        ctx.hamiltonandoverlapsetup(sysd, dk.ngk, dk.apwalm, dk.igpig, dk.vgpc, flags=0)
        torch.cuda.synchronize()
        A0 = oracle.match(S, ik)
        H0, O0 = oracle.setup(S, ik, A0)
        A = dk.apwalm.cpu().numpy().reshape(dk.apwalm_shape, order="F")
        H = sysd.hamilton.za.cpu().numpy().T
        O = sysd.overlap.za.cpu().numpy().T
        print(f"  device {ik}: A {rel(A, A0):.2e} H {rel(H, H0):.2e} O {rel(O, O0):.2e}")
        assert rel(A, A0) <= TOL_MAT
        assert rel(H, H0) <= TOL_MAT and rel(O, O0) <= TOL_MAT


@pytest.mark.gpu
def test_errors_are_reported(ctxs):
    s, S, ctx = ctxs("tiny")
    k = s.kpts[0]
    sysg = lapw.newsystem(k.ngk + s.nlotot + 1)  # wrong rank
    with pytest.raises(lapw.ExhError, match="rank"):
        ctx.hamiltonandoverlapsetup(sysg, k.ngk, np.zeros((1,), np.complex128), k.igpig, k.vgpc)

