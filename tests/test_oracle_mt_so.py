"""The oracle's spin-orbit table producer (mt_so.f90, the branch the
reference runs) pinned without itself:

* the radial factor (1/r) d/dr[rmtable] takes the derivative from the
  spline coefficients, which tests/golden/radial_ref.npz pins against the
  reference's own spline.f90 (test_oracle_hmlint.py);
* the tables against an independent NumPy statement: the operator L.S
  built from the Pauli matrices and the angular-momentum ladder operators
  in the basis of (spin, l, m), radial integrals as weighted sums with the
  pinned fderiv rule applied to unit vectors, laid out in the reference's
  row orders;
* alpha%alo and beta%alo equal conj(transpose(loa)), ab%alo holds the
  APW-row, LO-column elements of the same operator;
* a planted mistake (L+ in place of L- in ab) is caught.
"""
import numpy as np
import pytest

import oracle
from exciting_amd import synth

Y00 = 0.28209479177387814347
A = 0.5 * (1.0 / 137.03599911) ** 2


@pytest.fixture(autouse=True)
def no_fault():
    oracle.set_fault(0)
    yield
    oracle.set_fault(0)


def spline_weights(x):
    n = x.size
    return np.array([oracle.fderiv_int(x, np.eye(n)[j])[-1] for j in range(n)])


def l_operators(l):
    """Lz, L+, L- in the basis m = -l..l (index m + l): L+|m> =
    sqrt(l(l+1) - m(m+1)) |m+1>."""
    m = np.arange(-l, l + 1)
    lz = np.diag(m).astype(float)
    lp = np.zeros((2 * l + 1, 2 * l + 1))
    for i, mm in enumerate(m[:-1]):
        lp[i + 1, i] = np.sqrt(l * (l + 1) - mm * (mm + 1))
    return lz, lp, lp.T


def spin_orbit_blocks(l):
    """(up-up, down-down, up-down) blocks of L.S = (Lx sx + Ly sy + Lz sz)/2
    with the Pauli matrices."""
    lz, lp, lm = l_operators(l)
    lx, ly = (lp + lm) / 2, (lp - lm) / 2j
    sx = np.array([[0, 1], [1, 0]])
    sy = np.array([[0, -1j], [1j, 0]])
    sz = np.array([[1, 0], [0, -1]])
    ls = 0.5 * (np.kron(sx, lx) + np.kron(sy, ly) + np.kron(sz, lz))
    n = 2 * l + 1
    return ls[:n, :n], ls[n:, n:], ls[:n, n:]


def numpy_mt_so(s, rel_level=1):
    lmaxmat, nlo = s.lmaxmat, max(s.maxnlo, 1)
    out = {k: [np.zeros((s.maxaa, s.maxaa, s.natmtot), complex),
               np.zeros((nlo, s.maxaa, s.natmtot), complex),
               np.zeros((nlo, nlo, s.natmtot), complex),
               np.zeros((s.maxaa, nlo, s.natmtot), complex)] for k in ("alpha", "beta", "ab")}
    ias = 0
    for is_ in range(s.nspecies):
        nr = int(s.nrmt[is_])
        r = s.spr[:nr, is_]
        w = spline_weights(r)
        # basis functions of the atom: (kind, row, l, m, radial index)
        rows = []
        for l in range(lmaxmat + 1):
            for m in range(-l, l + 1):
                for io in range(s.apword[l, is_]):
                    rows.append(("apw", len([x for x in rows if x[0] == "apw"]), l, m, ("apw", l, io)))
        k = 0
        for ilo in range(s.nlorb[is_]):
            l = s.lorbl[ilo, is_]
            for m in range(-l, l + 1):
                rows.append(("lo", k, l, m, ("lo", ilo)))
                k += 1
        blocks = {l: spin_orbit_blocks(l) for l in range(lmaxmat + 1)}
        for _ in range(s.natoms[is_]):
            v = s.veffmt[0, :nr, ias] * Y00
            rm = {0: A * v, 1: 1.0 / (1.0 - A * v), 2: A / (1.0 - A * v) ** 2}[rel_level]
            xi = oracle.spline(r, rm)[0] / r

            def u(f):
                if f[0] == "apw":
                    return s.apwfr[:nr, 0, f[2], f[1], ias]
                return s.lofr[:nr, 0, f[1], ias]

            for (k1, i1, l1, m1, f1) in rows:
                for (k2, i2, l2, m2, f2) in rows:
                    if l1 != l2 or l1 == 0:
                        continue
                    R = np.sum(w * u(f1) * u(f2) * xi * r * r)
                    uu, dd, ud = blocks[l1]
                    which = {("apw", "apw"): 0, ("lo", "apw"): 1, ("lo", "lo"): 2,
                             ("apw", "lo"): 3}[(k1, k2)]
                    for key, blk in (("alpha", uu), ("beta", dd), ("ab", ud)):
                        out[key][which][i1, i2, ias] += R * blk[m1 + l1, m2 + l1]
            ias += 1
    return out


@pytest.mark.parametrize("seed,rel_level", [(7, 1), (11, 0), (5, 2)])
def test_mt_so_matches_numpy_statement(seed, rel_level):
    s = synth.add_mt_potential(synth.build(synth.tiny(seed=seed)), lmaxvr=4)
    got = oracle.mt_so(oracle.System(s), rel_level=rel_level)
    want = numpy_mt_so(s, rel_level)
    for key in ("alpha", "beta", "ab"):
        for g, w_, what in zip(got[key], want[key], ("aa", "loa", "lolo", "alo")):
            scale = max(np.abs(w_).max(), 1e-300)
            assert np.abs(g - w_).max() <= 1e-12 * scale, (key, what)
        assert np.abs(got[key][0]).max() > 0
    # alpha and beta: Hermitian blocks, alo = conj(loa^T); beta = -alpha
    for key in ("alpha", "beta"):
        aa, loa, lolo, alo = got[key]
        assert np.array_equal(alo, np.conj(np.transpose(loa, (1, 0, 2))))
        assert np.abs(aa - np.conj(np.transpose(aa, (1, 0, 2)))).max() == 0.0
    for a_, b_ in zip(got["alpha"], got["beta"]):
        assert np.array_equal(a_, -b_)


def test_mt_so_accumulates_onto_given_tables():
    s = synth.add_mt_potential(synth.build(synth.tiny(seed=7)), lmaxvr=4)
    sy = oracle.System(s)
    base = oracle.mt_so(sy)
    rng = np.random.default_rng(2)
    pre = {k: tuple(rng.normal(size=x.shape) + 1j * rng.normal(size=x.shape) for x in v)
           for k, v in base.items()}
    acc = oracle.mt_so(sy, tables=pre)
    for key in ("alpha", "beta", "ab"):
        for i in range(3):
            assert np.abs(acc[key][i] - (pre[key][i] + base[key][i])).max() <= 1e-15
        # ab%alo accumulates; alpha/beta alo are reset from loa
        if key == "ab":
            assert np.abs(acc[key][3] - (pre[key][3] + base[key][3])).max() <= 1e-15
        else:
            assert np.array_equal(acc[key][3], np.conj(np.transpose(acc[key][1], (1, 0, 2))))


def test_planted_ladder_fault_is_caught():
    s = synth.add_mt_potential(synth.build(synth.tiny(seed=7)), lmaxvr=4)
    want = numpy_mt_so(s)
    oracle.set_fault(8)
    got = oracle.mt_so(oracle.System(s))
    err = np.abs(got["ab"][0] - want["ab"][0]).max() / np.abs(want["ab"][0]).max()
    assert err > 0.1
