"""Interstitial producers on the GPU (exh_gencfun*, exh_genveffig*,
exh_genmeffig, exh_set_interstitial_device) against the oracle
(oracle.gencfun / genveffig / genmeffig, themselves pinned by
test_oracle_interstitial.py).

Tolerance (FP64): relative Frobenius error <= 1e-12 and largest elementwise
error <= 1e-12 of the largest reference value.  The device results are
rocFFT transforms and device sincos, so they are not bitwise the oracle's.
End to end, H and O set up from device-produced tables equal (bitwise) the
setup from the same tables uploaded from the host, and match the oracle's
setup on the oracle's tables to 1e-12.
"""
import numpy as np
import pytest

import oracle
from exciting_amd import lapw
from tests import _cases

TOL = 1e-12


def close(got, want, what):
    got, want = np.asarray(got), np.asarray(want)
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    mx = np.abs(got - want).max() / np.abs(want).max()
    print(f"  {what}: rel {rel:.1e} max {mx:.1e}")
    assert rel <= TOL and mx <= TOL, what


def veffir_of(s, seed=1):
    rng = np.random.default_rng(seed)
    return -0.4 + 0.3 * rng.uniform(-1, 1, s.ngrtot)


@pytest.fixture(scope="module")
def ctxs():
    cache = {}

    def get(name):
        if name not in cache:
            s = _cases.system(name)
            print(f"[gpu-interstitial] case {name}: ngrtot {s.ngrtot} natmtot {s.natmtot}",
                  flush=True)
            cache[name] = (s, lapw.Context(0).load(s))
        return cache[name]
    yield get
    for _, c in cache.values():
        c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tiny", "fcc2", "sc8_k0", "sc64_k0"])
def test_gencfun_matches_the_oracle(ctxs, name):
    import torch
    s, ctx = ctxs(name)
    cg0, cr0 = oracle.gencfun(s)
    cg = torch.zeros(s.ngrtot, dtype=torch.complex128, device="cuda")
    cr = torch.zeros(s.ngrtot, dtype=torch.float64, device="cuda")
    ctx.gencfun(s.bvec, s.atposc, cfunig=cg, cfunir=cr)
    torch.cuda.synchronize()
    close(cg.cpu().numpy(), cg0, f"{name} cfunig")
    close(cr.cpu().numpy(), cr0, f"{name} cfunir")
    # the host entry point runs the same kernels: bitwise equal
    hg, hr = ctx.gencfun(s.bvec, s.atposc)
    assert np.array_equal(hg, cg.cpu().numpy()) and np.array_equal(hr, cr.cpu().numpy())


@pytest.mark.gpu
def test_gencfun_damping(ctxs):
    s, ctx = ctxs("fcc2")
    cg0, cr0 = oracle.gencfun(s, cfdamp=0.7, gmaxvr=s.spec.gmaxvr)
    cg, cr = ctx.gencfun(s.bvec, s.atposc, cfdamp=0.7, gmaxvr=s.spec.gmaxvr)
    close(cg, cg0, "damped cfunig")
    close(cr, cr0, "cfunir (transformed before damping)")
    und, _ = oracle.gencfun(s)
    assert np.abs(cg - und).max() > 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tiny", "sc8_k0", "sc64_k0"])
def test_genveffig_and_genmeffig_match_the_oracle(ctxs, name):
    import torch
    s, ctx = ctxs(name)
    _, cr0 = oracle.gencfun(s)
    v = veffir_of(s)
    want = {"veffig": oracle.genveffig(s, v, cr0)}
    want["meffig"], want["m2effig"] = oracle.genmeffig(s, v, cr0, True)
    vd = torch.from_numpy(v).cuda()
    cd = torch.from_numpy(cr0).cuda()
    names = ("veffig", "meffig", "m2effig")
    for which in range(1, 8):
        outs = {n: (torch.zeros(s.ngvec, dtype=torch.complex128, device="cuda")
                    if which & (1 << i) else None) for i, n in enumerate(names)}
        ctx.genveffig_device(vd, cd, **outs)
        torch.cuda.synchronize()
        for n, t in outs.items():
            if t is not None:
                close(t.cpu().numpy(), want[n], f"{name} which={which} {n}")
    # host entry points
    close(ctx.genveffig(v, cr0), want["veffig"], f"{name} host veffig")
    m, m2 = ctx.genmeffig(v, cr0, iora_star=True)
    close(m, want["meffig"], f"{name} host meffig")
    close(m2, want["m2effig"], f"{name} host m2effig")
    m, m2 = ctx.genmeffig(v, cr0, iora_star=False)
    assert m2 is None
    close(m, want["meffig"], f"{name} host meffig (zora)")
    assert ctx.genmeffig(v, cr0, valence_rel_none=True) == (None, None)


@pytest.mark.gpu
def test_genveffig_of_a_plane_wave(ctxs):
    s, ctx = ctxs("fcc2")
    ngrid = oracle.ngrid_of(s.intgv)
    j = np.arange(s.ngrtot)
    f = np.stack([j % ngrid[0] / ngrid[0], j // ngrid[0] % ngrid[1] / ngrid[1],
                  j // (ngrid[0] * ngrid[1]) / ngrid[2]])
    g0 = 11
    veffir = np.cos(2 * np.pi * (s.ivg[:, g0].astype(float) @ f))
    got = ctx.genveffig(veffir, np.ones(s.ngrtot))
    m = -s.ivg[:, g0]
    neg = s.ivgig[m[0] - s.intgv[0, 0], m[1] - s.intgv[1, 0], m[2] - s.intgv[2, 0]] - 1
    want = np.zeros(s.ngvec, dtype=complex)
    want[g0] += 0.5
    want[neg] += 0.5
    assert np.abs(got - want).max() <= 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("name,rel", [("tiny", "none"), ("tiny_rel", "iora*"),
                                      ("sc8_k0", "none")])
def test_setup_from_device_produced_tables(ctxs, name, rel):
    """gencfun -> genveffig/genmeffig -> set_interstitial_device -> setup,
    all on the GPU, against the same tables uploaded from the host (bitwise)
    and against the oracle's producers + setup (1e-12)."""
    import torch
    s, ctx = ctxs(name)
    relnone = rel == "none"
    h1on = bool(s.h1on)
    v = veffir_of(s, 2)
    cg = torch.zeros(s.ngrtot, dtype=torch.complex128, device="cuda")
    cr = torch.zeros(s.ngrtot, dtype=torch.float64, device="cuda")
    ctx.gencfun(s.bvec, s.atposc, cfunig=cg, cfunir=cr)
    vd = torch.from_numpy(v).cuda()
    z = lambda: torch.zeros(s.ngvec, dtype=torch.complex128, device="cuda")
    vg, mg, m2g = z(), (None if relnone else z()), (z() if h1on else None)
    ctx.genveffig_device(vd, cr, veffig=vg, meffig=mg, m2effig=m2g)
    try:
        ctx.set_interstitial(cg, vg, mg, m2g, valence_rel_none=relnone, h1on=h1on)
        d = lapw.DeviceKPoint(s, 0)
        H = torch.zeros(d.rank * d.rank, dtype=torch.complex128, device="cuda")
        O = torch.zeros_like(H)
        ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, d.apwalm)
        ctx.hamiltonandoverlapsetup(lapw.EvSystem(lapw.HermitianMatrix(d.rank, H),
                                                  lapw.HermitianMatrix(d.rank, O)),
                                    d.ngk, d.apwalm, d.igpig, d.vgpc, flags=0)
        torch.cuda.synchronize()
        Hd, Od = H.cpu().numpy(), O.cpu().numpy()
        tab = [t.cpu().numpy() if t is not None else None for t in (cg, vg, mg, m2g)]
        ctx.set_interstitial(*tab, valence_rel_none=relnone, h1on=h1on)
        H.zero_()
        O.zero_()
        ctx.hamiltonandoverlapsetup(lapw.EvSystem(lapw.HermitianMatrix(d.rank, H),
                                                  lapw.HermitianMatrix(d.rank, O)),
                                    d.ngk, d.apwalm, d.igpig, d.vgpc, flags=0)
        torch.cuda.synchronize()
        assert np.array_equal(Hd, H.cpu().numpy()) and np.array_equal(Od, O.cpu().numpy())
        # oracle: its own producers, then its own setup
        cg0, cr0 = oracle.gencfun(s)
        vg0 = oracle.genveffig(s, v, cr0)
        mg0, m2g0 = (None, None) if relnone else oracle.genmeffig(s, v, cr0, rel == "iora*")
        import dataclasses
        s0 = dataclasses.replace(s, cfunig=cg0, veffig=vg0, meffig=mg0,
                                 m2effig=m2g0 if h1on else None)
        S = oracle.System(s0)
        A = oracle.match(S, 0)
        H0, O0 = oracle.setup(S, 0, A)
        n = d.rank
        close(Hd.reshape(n, n).T, H0, f"{name} H")
        close(Od.reshape(n, n).T, O0, f"{name} O")
    finally:
        ctx.set_interstitial(s.cfunig, s.veffig, s.meffig, s.m2effig, s.valence_rel_none,
                             s.h1on)


@pytest.mark.gpu
def test_producers_reject_bad_arguments(ctxs):
    import torch
    s, ctx = ctxs("tiny")
    with pytest.raises(lapw.ExhError, match="atoms"):
        ctx.gencfun(s.bvec, s.atposc[:-1])
    vd = torch.zeros(s.ngrtot, dtype=torch.float64, device="cuda")
    with pytest.raises(lapw.ExhError, match="which"):
        ctx.genveffig_device(vd, vd)
    # a G set that is not the whole FFT box
    other = lapw.Context(0)
    try:
        other.set_basis(s.natoms, s.lmaxapw, s.lmaxmat, s.apwordmax, s.apword, s.nlomax,
                        s.nlorb, s.lorbl, s.ngkmax)
        other.set_radial(s.omega, s.rmt, s.nrmt, s.nrmtmax, s.spr, s.apwfr)
        other.set_gvectors(s.ngrtot - 1, s.ngvec, s.ivg, s.intgv, s.ivgig)
        with pytest.raises(lapw.ExhError, match="FFT grid"):
            other.gencfun(s.bvec, s.atposc)
    finally:
        other.close()
