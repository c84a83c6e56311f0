"""The oracle's forcek (forcek.f90) pinned by finite differences: its
derivative matrices are, per atom, minus the derivative of H and the
derivative of O with respect to that atom's position (the atom's APW and
LO functions move with it through the structure factors in apwalm, and the
interstitial step function through cfunig), so

    F_l(alpha) = - sum_j occ_j  d/dtau_l [ x_j^H (H - e_j O) x_j ]

with fixed states x_j and energies e_j.  H and O at tau +- delta come from
the oracle's setup (pinned separately) with sfacgp and the true step
function cfunig (gencfunig, pinned separately) recomputed; the muffin-tin
tables and veffig are held fixed, which leaves exactly the terms forcek
differentiates.  Central differences, delta = 1e-4 bohr."""
import dataclasses

import numpy as np
import pytest

import oracle
from exciting_amd import synth


def cell(seed=7, nk=1):
    s = synth.build(synth.tiny(seed=seed))
    s = dataclasses.replace(s, kpts=s.kpts[:nk])
    return with_positions(s, np.asarray(s.atposc, float))


def with_positions(s, pos):
    """The same cell with atoms at pos: structure factors and the step
    function follow, everything else is kept."""
    cf = oracle.gencfunig(s.ivg, s.bvec, s.omega, s.natoms, s.rmt, pos)
    kpts = []
    for k in s.kpts:
        vg = np.asarray(k.vgpc)[:, :k.ngk].T
        sf = np.zeros_like(k.sfacgp)
        t = vg @ pos.T
        sf[:k.ngk] = np.cos(t) + 1j * np.sin(t)
        kpts.append(dataclasses.replace(k, sfacgp=sf))
    return dataclasses.replace(s, atposc=pos, cfunig=cf, kpts=kpts)


def band_sum(s, X, e, occ):
    S = oracle.System(s)
    A = oracle.match(S, 0)
    H, O = oracle.setup(S, 0, A)
    return sum(occ[j] * np.real(X[:, j].conj() @ (H - e[j] * O) @ X[:, j])
               for j in range(X.shape[1]))


@pytest.mark.parametrize("ias,l", [(0, 0), (1, 2), (0, 1)])
def test_forcek_is_minus_the_derivative_of_the_band_sum(ias, l):
    s = cell()
    n = s.kpts[0].ngk + s.nlotot
    rng = np.random.default_rng(4)
    nst = 5
    X = rng.normal(size=(n, nst)) + 1j * rng.normal(size=(n, nst))
    e = rng.uniform(-1, 1, nst)
    occ = rng.uniform(0, 2, nst)
    ff = np.stack([oracle.genffacg(s, i) for i in range(s.nspecies)], axis=1)
    F = oracle.forcek(oracle.System(s), 0, X, e, occ, 1.0, ff, oracle.sfacg_of(s))
    d = 1e-4
    pos = np.asarray(s.atposc, float)
    vals = []
    for sgn in (1, -1):
        p2 = pos.copy()
        p2[ias, l] += sgn * d
        vals.append(band_sum(with_positions(s, p2), X, e, occ))
    fd = -(vals[0] - vals[1]) / (2 * d)
    print(f"  atom {ias} dir {l}: forcek {F[l, ias]:.10f}  finite difference {fd:.10f}")
    assert F[l, ias] == pytest.approx(fd, rel=1e-6, abs=1e-8 * np.abs(F).max())


def band_sum_dm(s, X, e, dm):
    """sum_ab Re[x_a^H (H - e_b O) x_b dm_ba], the second-variational form."""
    S = oracle.System(s)
    A = oracle.match(S, 0)
    H, O = oracle.setup(S, 0, A)
    tot = 0.0
    for b in range(X.shape[1]):
        col = (H - e[b] * O) @ X[:, b]
        tot += np.real(np.sum(X.conj().T @ col * dm[b, :]))
    return tot


@pytest.mark.parametrize("spiral", [None, 1])
def test_second_variational_forcek_is_minus_the_derivative(spiral):
    """tevecsv (forcek.f90:341-366): with second-variational vectors v_j the
    state sum is sum_j occ_j sum_spinor v^H ffv v = sum_ab ffv_ab dm_ba with
    dm = sum_j occ_j v_j v_j^H over the spinor blocks (or the block of one
    first-variational spin for a spin spiral), so the force is minus the
    derivative of sum_ab x_a^H (H - e_b O) x_b dm_ba."""
    s = cell()
    n = s.kpts[0].ngk + s.nlotot
    rng = np.random.default_rng(8)
    nst = 4
    X = rng.normal(size=(n, nst)) + 1j * rng.normal(size=(n, nst))
    e = rng.uniform(-1, 1, nst)
    nsv = 2 * nst
    V = np.linalg.qr(rng.normal(size=(nsv, nsv)) + 1j * rng.normal(size=(nsv, nsv)))[0]
    occ = rng.uniform(0, 1, nsv)
    blocks = [spiral] if spiral is not None else [0, 1]
    dm = sum((V[b * nst:(b + 1) * nst] * occ[None, :]) @ V[b * nst:(b + 1) * nst].conj().T
             for b in blocks)
    ff = np.stack([oracle.genffacg(s, i) for i in range(s.nspecies)], axis=1)
    F = oracle.forcek(oracle.System(s), 0, X, e, occ, 1.0, ff, oracle.sfacg_of(s), evecsv=V,
                      spiral_ispn=spiral)
    d = 1e-4
    pos = np.asarray(s.atposc, float)
    for ias, l in ((0, 0), (1, 2)):
        vals = []
        for sgn in (1, -1):
            p2 = pos.copy()
            p2[ias, l] += sgn * d
            vals.append(band_sum_dm(with_positions(s, p2), X, e, dm))
        fd = -(vals[0] - vals[1]) / (2 * d)
        print(f"  atom {ias} dir {l}: forcek {F[l, ias]:.10f}  finite difference {fd:.10f}")
        assert F[l, ias] == pytest.approx(fd, rel=1e-6, abs=1e-8 * np.abs(F).max())


def test_identity_second_variational_vectors_reduce_to_the_first_variational_sum():
    s = cell()
    n = s.kpts[0].ngk + s.nlotot
    rng = np.random.default_rng(9)
    nst = 4
    X = rng.normal(size=(n, nst)) + 1j * rng.normal(size=(n, nst))
    e = rng.uniform(-1, 1, nst)
    occ = rng.uniform(0, 2, nst)
    ff = np.stack([oracle.genffacg(s, i) for i in range(s.nspecies)], axis=1)
    S, sf = oracle.System(s), oracle.sfacg_of(s)
    a = oracle.forcek(S, 0, X, e, occ, 0.5, ff, sf)
    b = oracle.forcek(S, 0, X, e, occ, 0.5, ff, sf, evecsv=np.eye(nst), nspinor=1)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()
