"""The oracle's muffin-tin table producer (hmlint.f90 -> mt_kin.f90 +
mt_pot.f90, Gaunt list of init1.f90:563-609) pinned without itself:

* spline / fderiv and wigner3j / gaunt_yyy / gaunt_yry against the
  reference's own routines (flang build) through tests/golden/radial_ref.npz;
* gaunt_yry against a quadrature of conj(Y_l1m1) R_l2m2 Y_l3m3 over the
  sphere with SciPy's Y_lm and the real harmonics of genrlm.f90;
* the whole table set against an independent NumPy statement: dense Gaunt
  tensor from that quadrature, radial integrals as weighted sums, the
  reference's row layouts built separately, contraction by einsum;
* a spherical potential gives tables diagonal in (l, m) (closed form);
* mt_pot alone and mt_kin alone (parts 2 and 1) match the NumPy
  statement with the other term off and add up to the whole.
"""
import dataclasses
import os

import numpy as np
import pytest
import scipy.special as ss

import oracle
from exciting_amd import synth

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                         "radial_ref.npz"))
Y00 = 0.28209479177387814347
ALPHA = 1.0 / 137.03599911


@pytest.fixture(autouse=True)
def no_fault():
    oracle.set_fault(0)
    yield
    oracle.set_fault(0)


@pytest.mark.parametrize("i", range(int(G["ngrids"])))
def test_spline_and_fderiv_match_reference_vectors(i):
    x, f = G[f"x{i}"], G[f"f{i}"]
    np.testing.assert_allclose(oracle.spline(x, f), G[f"cf{i}"], rtol=1e-13,
                               atol=1e-13 * np.abs(G[f"cf{i}"]).max())
    g = oracle.fderiv_int(x, f)
    np.testing.assert_allclose(g, G[f"g{i}"], rtol=1e-14, atol=1e-15 * np.abs(G[f"g{i}"]).max())


def test_wigner3j_and_gaunt_match_reference_vectors():
    L = oracle.lib()
    for row in G["w3j"]:
        j = [int(v) for v in row[:6]]
        assert L.orc_wigner3j(*j) == pytest.approx(row[6], abs=1e-15), j
    lg = int(G["lg"])
    yry, yyy = G["gaunt_yry"], G["gaunt_yyy"]
    for l1 in range(lg + 1):
        for m1 in range(-l1, l1 + 1):
            for l2 in range(lg + 1):
                for m2 in range(-l2, l2 + 1):
                    for l3 in range(lg + 1):
                        for m3 in range(-l3, l3 + 1):
                            a, b, c = l1 * (l1 + 1) + m1, l2 * (l2 + 1) + m2, l3 * (l3 + 1) + m3
                            assert abs(oracle.gaunt_yry(l1, l2, l3, m1, m2, m3) - yry[a, b, c]) \
                                <= 1e-15
                            assert abs(L.orc_gaunt_yyy(l1, l2, l3, m1, m2, m3) - yyy[a, b, c]) \
                                <= 1e-15


def ylm_table(lmax, th, ph):
    """Y_lm (Condon-Shortley, SciPy) and the real R_lm of genrlm.f90 at the
    quadrature points: arrays [lm, point]."""
    n = (lmax + 1) ** 2
    Y = np.zeros((n, th.size), dtype=complex)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            Y[l * (l + 1) + m] = ss.sph_harm_y(l, m, th, ph)
    R = np.zeros((n, th.size))
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            y = Y[l * (l + 1) + m]
            R[l * (l + 1) + m] = (np.sqrt(2) * y.imag if m < 0 else
                                  (y.real if m == 0 else np.sqrt(2) * y.real))
    return Y, R


def gaunt_by_quadrature(lmax1, lmax2):
    """G[lm1, lm2, lm3] = integral of conj(Y_lm1) R_lm2 Y_lm3, exact by
    Gauss-Legendre in cos(theta) and the trapezoid rule in phi."""
    deg = 2 * lmax1 + lmax2
    mu, w = np.polynomial.legendre.leggauss(deg // 2 + 2)
    nphi = deg + 2
    ph = 2 * np.pi * np.arange(nphi) / nphi
    th = np.repeat(np.arccos(mu), nphi)
    phs = np.tile(ph, mu.size)
    wt = np.repeat(w, nphi) * (2 * np.pi / nphi)
    Y1, _ = ylm_table(lmax1, th, phs)
    _, R2 = ylm_table(lmax2, th, phs)
    return np.einsum("ap,bp,cp,p->abc", Y1.conj(), R2, Y1, wt)


def test_gaunt_yry_matches_quadrature():
    lmax1, lmax2 = 5, 4
    Q = gaunt_by_quadrature(lmax1, lmax2)
    n1, n2 = (lmax1 + 1) ** 2, (lmax2 + 1) ** 2
    got = np.zeros_like(Q)
    for l1 in range(lmax1 + 1):
        for m1 in range(-l1, l1 + 1):
            for l2 in range(lmax2 + 1):
                for m2 in range(-l2, l2 + 1):
                    for l3 in range(lmax1 + 1):
                        for m3 in range(-l3, l3 + 1):
                            got[l1 * (l1 + 1) + m1, l2 * (l2 + 1) + m2, l3 * (l3 + 1) + m3] = \
                                oracle.gaunt_yry(l1, l2, l3, m1, m2, m3)
    assert Q.shape == (n1, n2, n1)
    np.testing.assert_allclose(got, Q, rtol=0, atol=1e-14)


def test_planted_gaunt_sign_fault_is_caught():
    Q = gaunt_by_quadrature(2, 2)
    # an imaginary coefficient: l1 = 1, m1 = 1; l2 = 2, m2 = -2; l3 = 1, m3 = -1
    assert abs(Q[3, 4, 1].imag) > 0.05
    assert abs(oracle.gaunt_yry(1, 2, 1, 1, -2, -1) - Q[3, 4, 1]) < 1e-14
    oracle.set_fault(7)
    assert abs(oracle.gaunt_yry(1, 2, 1, 1, -2, -1) - Q[3, 4, 1]) > 1e-3


def spline_weights(x):
    """w with sum(w * f) = fderiv(-1)(f)(n), from the (pinned) oracle fderiv
    applied to unit vectors: the integration rule is linear in f."""
    n = x.size
    return np.array([oracle.fderiv_int(x, np.eye(n)[j])[-1] for j in range(n)])


def numpy_hmlint(s, parts=3):
    """Independent NumPy statement of mt_kin (parts & 1) + mt_pot (parts & 2)
    for every atom."""
    lmaxmat, p, lmaxvr = s.lmaxmat, s.apwordmax, s.lmaxvr
    lmm = (lmaxmat + 1) ** 2
    Q = gaunt_by_quadrature(lmaxmat, lmaxvr)              # [lm1, lm2, lm3]
    mnlo = max(s.maxnlo, 1)
    aa = np.zeros((s.maxaa, s.maxaa, s.natmtot), dtype=complex)
    loa = np.zeros((mnlo, s.maxaa, s.natmtot), dtype=complex)
    lolo = np.zeros((mnlo, mnlo, s.natmtot), dtype=complex)
    a = 0.0 if s.valence_rel_none else 0.5 * ALPHA ** 2
    ias = 0
    for is_ in range(s.nspecies):
        nr = int(s.nrmt[is_])
        r = s.spr[:nr, is_]
        w = spline_weights(r)
        # rows: APW (l, m, io) with io fastest; LO (ilo, m)
        apw = [(l, m, io) for l in range(lmaxmat + 1) for m in range(-l, l + 1)
               for io in range(s.apword[l, is_])]
        los = [(ilo, s.lorbl[ilo, is_], m) for ilo in range(s.nlorb[is_])
               for m in range(-s.lorbl[ilo, is_], s.lorbl[ilo, is_] + 1)]
        for _ in range(s.natoms[is_]):
            V = s.veffmt[:, :nr, ias]                      # (lmmaxvr, nr)
            rm = 1.0 / (1.0 - a * V[0] * Y00)

            def fn(kind, idx, k):
                if kind == "apw":
                    l, io = idx
                    return s.apwfr[:nr, k, io, l, ias]
                return s.lofr[:nr, k, idx, ias]

            def pot(f1, f2):                               # integral per lm2
                return V @ (w * f1 * f2 * r * r)

            def kin(f1, f2, d1, d2, l):
                return np.sum(w * 0.5 * rm * (d1 * d2 * r * r + l * (l + 1) * f1 * f2))

            def block(rows, cols, out, rowfun, colfun):
                for i, (lm1, f1, d1, l1) in enumerate(rows):
                    for j, (lm3, f3, d3, l3) in enumerate(cols):
                        v = Q[lm1, :, lm3] @ pot(f1, f3) if parts & 2 else 0.0
                        if parts & 1 and lm1 == lm3:
                            v += kin(f1, f3, d1, d3, l1)
                        out[i, j] += v

            A = [(l * (l + 1) + m, fn("apw", (l, io), 0), fn("apw", (l, io), 1), l)
                 for (l, m, io) in apw]
            Lr = [(l * (l + 1) + m, fn("lo", ilo, 0), fn("lo", ilo, 1), l) for (ilo, l, m) in los]
            block(A, A, aa[:, :, ias], None, None)
            if Lr:
                block(Lr, A, loa[:, :, ias], None, None)
                block(Lr, Lr, lolo[:, :, ias], None, None)
            ias += 1
    return aa, loa, lolo


@pytest.mark.parametrize("name,relnone", [("tiny", True), ("tiny_zora", False),
                                          ("tiny_lmaxmat", True)])
def test_hmlint_matches_numpy_statement(name, relnone):
    seeds = {"tiny": 7, "tiny_zora": 12, "tiny_lmaxmat": 8}
    kw = dict(lmax=4, lmaxmat=2, natoms=(2, 1)) if name == "tiny_lmaxmat" else {}
    s = synth.add_mt_potential(synth.build(synth.tiny(seed=seeds[name], **kw)), lmaxvr=4)
    s.valence_rel_none = relnone
    got = oracle.hmlint(oracle.System(s))
    want = numpy_hmlint(s)
    for g, w_, what in zip(got, want, ("aa", "loa", "lolo")):
        err = np.abs(g - w_).max() / np.abs(w_).max()
        assert err <= 1e-12, (what, err)
    # the tables of a Hermitian operator
    for ias in range(s.natmtot):
        A = got[0][:, :, ias]
        assert np.abs(A - A.conj().T).max() <= 1e-13 * np.abs(A).max()


def test_spherical_potential_gives_lm_diagonal_tables():
    s = synth.add_mt_potential(synth.build(synth.tiny(seed=7)), lmaxvr=4)
    s.veffmt[1:] = 0.0
    aa, loa, lolo = oracle.hmlint(oracle.System(s))
    p = s.apwordmax
    lm_of_row = [l * (l + 1) + m for l in range(s.lmaxmat + 1) for m in range(-l, l + 1)
                 for _ in range(p)]
    lm = np.array(lm_of_row)
    off = lm[:, None] != lm[None, :]
    assert np.abs(aa[:len(lm), :len(lm)][off]).max() == 0.0
    assert np.abs(aa).max() > 0.1


@pytest.mark.parametrize("relnone", [True, False])
def test_hmlint_parts_split_and_match_numpy_statement(relnone):
    """mt_pot alone (parts=2) on a non-spherical function other than veffmt,
    as the spin-polarised blocks call it, and mt_kin alone (parts=1): each
    matches the NumPy statement, and the two add up to parts=3."""
    s = synth.add_mt_potential(synth.build(synth.tiny(seed=9)), lmaxvr=4)
    s.valence_rel_none = relnone
    sy = oracle.System(s)
    kin = oracle.hmlint(sy, parts=1)
    both = oracle.hmlint(sy, parts=3)
    bxc = synth.add_mt_potential(synth.build(synth.tiny(seed=9)), lmaxvr=4, seed=33).veffmt
    pot = oracle.hmlint(sy, veffmt=bxc, parts=2)
    s_b = dataclasses.replace(s, veffmt=bxc)
    for g, w_ in zip(pot, numpy_hmlint(s_b, parts=2)):
        assert np.abs(g - w_).max() <= 1e-12 * np.abs(w_).max()
    for g, w_ in zip(kin, numpy_hmlint(s, parts=1)):
        assert np.abs(g - w_).max() <= 1e-12 * np.abs(w_).max()
    pot_v = oracle.hmlint(sy, parts=2)
    for k_, p_, b_ in zip(kin, pot_v, both):
        assert np.abs(k_ + p_ - b_).max() <= 1e-13 * np.abs(b_).max()
    # mt_pot does not see the relativity flag
    other = oracle.hmlint(sy, veffmt=bxc, valence_rel_none=not relnone, parts=2)
    for a_, b_ in zip(pot, other):
        assert np.array_equal(a_, b_)
