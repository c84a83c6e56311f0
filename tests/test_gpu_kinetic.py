"""Kinetic energy of the first-variational states (kinetic_energy.f90) on
the GPU (exciting_amd.kinetic: libexh's kinetic radial integrals, library
GEMMs, batched FFTs) against the oracle's restatement (pinned by
test_oracle_kinetic.py): relative error of every state's energy <= 1e-12,
for 'none' and ZORA relativity."""
import dataclasses

import numpy as np
import pytest

import oracle
from exciting_amd import kinetic, lapw, synth
from tests import _cases


@pytest.mark.gpu
@pytest.mark.parametrize("name,relnone,nst", [("tiny", True, 8), ("tiny_zora", False, 8),
                                              ("fcc2", True, 12), ("sc8_k0", True, 16)])
def test_kinetic_energy_matches_the_oracle(name, relnone, nst):
    import torch
    s = synth.add_mt_potential(dataclasses.replace(_cases.system(name)), lmaxvr=4, seed=9)
    s.valence_rel_none = relnone
    print(f"[gpu-kinetic] case {name} relnone {relnone}", flush=True)
    ctx = lapw.Context(0).load(s)
    try:
        d = lapw.DeviceKPoint(s, 0)
        n = d.rank
        ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, d.apwalm)
        rng = np.random.default_rng(2)
        X = rng.normal(size=(n, nst)) + 1j * rng.normal(size=(n, nst))
        _, cfunir = oracle.gencfun(s)
        veffir = -0.4 + 0.3 * rng.uniform(-1, 1, s.ngrtot)
        dev = lambda a: torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).cuda()
        got = kinetic.kinetic_energy(ctx, d, torch.from_numpy(X.T.copy()).cuda(), dev(cfunir),
                                     dev(s.lofr), relnone, dev(s.veffmt), s.lmaxvr,
                                     dev(veffir)).cpu().numpy()
        S = oracle.System(s)
        A = d.apwalm.cpu().numpy().reshape(d.apwalm_shape, order="F")
        want = oracle.kinetic_energy(S, 0, X, A, cfunir, s.lofr, s.veffmt, relnone, veffir)
        err = np.abs(got - want) / np.abs(want)
        print(f"  {name}: max rel {err.max():.1e}")
        assert err.max() <= 1e-12
    finally:
        ctx.close()
