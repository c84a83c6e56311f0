"""The oracle's interstitial producers (gencfun.f90, genveffig.f90,
genmeffig.f90, zfftifc.f90, genigfft in gengvec.f90) pinned without the
oracle itself:

* zfftifc: the reference's own zfftifc + cfftnd (flang build) through the
  committed vectors in tests/golden/fft_ref.npz;
* gencfunig: the closed form against a direct numerical Fourier integral of
  the step function over each (displaced) sphere, with no form factor or
  translation phase assumed; a sign flipped in either is caught (planted
  mistakes 5 and 6);
* G = 0, Hermitian symmetry c(-G) = conj(c(G));
* gencfun's scatter + backward transform, and genveffig's forward transform
  + gather, against direct sums over the grid/G set (the igfft map and the
  transform signs enter only the oracle side);
* a known answer: veffir = cos(G0.r) with cfunir = 1 gives veffig(+-G0) = 1/2
  and zero elsewhere.
"""
import os

import numpy as np
import pytest

import oracle
from tests import _cases

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                            "fft_ref.npz"))


@pytest.fixture(autouse=True)
def no_fault():
    oracle.set_fault(0)
    yield
    oracle.set_fault(0)


@pytest.mark.parametrize("i", range(int(GOLD["ngrids"])))
def test_zfftifc_matches_reference_vectors(i):
    g = tuple(int(x) for x in GOLD[f"grid{i}"])
    z = GOLD[f"in{i}"]
    for sgn, tag in ((-1, "fwd"), (1, "bwd")):
        want = GOLD[f"{tag}{i}"]
        got = oracle.zfftifc(g, sgn, z)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max(),
                                   err_msg=f"grid {g} sgn {sgn}")


# A general cell: non-orthogonal lattice, two species of different radii,
# atoms at positions with no symmetry, so every phase and sign matters.
AVEC = np.array([[7.0, 0.5, 0.1], [0.3, 6.5, 0.7], [0.2, -0.4, 8.0]])   # columns a1 a2 a3
BVEC = 2 * np.pi * np.linalg.inv(AVEC).T
OMEGA = abs(np.linalg.det(AVEC))
NATOMS = [1, 2]
RMT = [1.8, 2.1]
ATPOSC = np.array([[0.31, 0.22, 0.13], [3.1, 3.4, 4.2], [5.0, 1.0, 6.0]])
IVG = np.array([[0, 0, 0], [1, 0, 0], [0, -1, 0], [1, 2, -1], [-2, 1, 3], [3, -3, 2],
                [4, 0, -5], [-1, -1, -1], [2, 5, 1]]).T


def sphere_integral(G, tau, R, nr=48, nmu=48, nphi=72):
    """Integral of exp(-i G.r) over |r - tau| < R by product quadrature:
    Gauss-Legendre in the radius and in cos(theta), trapezoid in phi."""
    x, wx = np.polynomial.legendre.leggauss(nr)
    r = 0.5 * R * (x + 1)
    wr = 0.5 * R * wx * r * r
    mu, wmu = np.polynomial.legendre.leggauss(nmu)
    phi = 2 * np.pi * np.arange(nphi) / nphi
    st = np.sqrt(1 - mu * mu)
    n = np.stack([st[:, None] * np.cos(phi)[None, :], st[:, None] * np.sin(phi)[None, :],
                  np.broadcast_to(mu[:, None], (nmu, nphi))], axis=-1)   # (nmu, nphi, 3)
    pts = tau[None, None, None, :] + r[:, None, None, None] * n[None]   # (nr, nmu, nphi, 3)
    f = np.exp(-1j * (pts @ G))
    return np.einsum("r,m,rmp->", wr, wmu, f) * (2 * np.pi / nphi)


def step_by_quadrature(ivg):
    out = []
    species = np.repeat(np.arange(len(NATOMS)), NATOMS)
    for iv in ivg.T:
        G = BVEC @ iv
        c = 1.0 if not iv.any() else 0.0
        for ias, tau in enumerate(ATPOSC):
            c = c - sphere_integral(G, tau, RMT[species[ias]]) / OMEGA
        out.append(c)
    return np.array(out)


def oracle_step(ivg):
    return oracle.gencfunig(ivg, BVEC, OMEGA, NATOMS, RMT, ATPOSC)


def test_gencfunig_matches_the_fourier_integral_of_the_step_function():
    want = step_by_quadrature(IVG)
    got = oracle_step(IVG)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("fault", [5, 6])
def test_planted_form_factor_and_phase_faults_are_caught(fault):
    want = step_by_quadrature(IVG)
    oracle.set_fault(fault)
    got = oracle_step(IVG)
    assert np.abs(got - want).max() > 1e-4


def test_gencfunig_at_g0_and_hermitian_symmetry():
    got = oracle_step(np.concatenate([IVG, -IVG], axis=1))
    n = IVG.shape[1]
    fill = sum(4 * np.pi * r ** 3 / 3 * na for r, na in zip(RMT, NATOMS)) / OMEGA
    assert got[0] == pytest.approx(1 - fill, abs=1e-15)
    np.testing.assert_allclose(got[n:], np.conj(got[:n]), rtol=0, atol=1e-15)


def grid_points(ngrid, idx):
    """Fractional coordinates of FFT grid points (first index fastest)."""
    n1, n2, n3 = ngrid
    return np.stack([idx % n1 / n1, idx // n1 % n2 / n2, idx // (n1 * n2) / n3])


@pytest.mark.parametrize("name", ["tiny", "fcc2"])
def test_gencfun_real_space_is_the_direct_fourier_sum(name):
    s = _cases.system(name)
    cfunig, cfunir = oracle.gencfun(s)
    ngrid = oracle.ngrid_of(s.intgv)
    assert cfunig.size == cfunir.size == np.prod(ngrid) == s.ngrtot
    rng = np.random.default_rng(3)
    idx = rng.choice(s.ngrtot, 12, replace=False)
    f = grid_points(ngrid, idx)                                   # (3, 12)
    ph = np.exp(2j * np.pi * (s.ivg.T.astype(float) @ f))        # G.r = 2 pi i.f
    want = (cfunig[:, None] * ph).sum(axis=0)
    np.testing.assert_allclose(cfunir[idx], want.real, rtol=0, atol=1e-12)
    if all(n % 2 for n in ngrid):      # odd grids: every G has its -G in the box
        assert np.abs(want.imag).max() < 1e-12
    assert cfunir.mean() == pytest.approx(cfunig[0].real, abs=1e-13)


@pytest.mark.parametrize("name", ["tiny", "fcc2"])
def test_genveffig_and_genmeffig_are_the_direct_fourier_sums(name):
    s = _cases.system(name)
    _, cfunir = oracle.gencfun(s)
    rng = np.random.default_rng(4)
    veffir = -0.5 + rng.uniform(-1, 1, s.ngrtot)
    ngrid = oracle.ngrid_of(s.intgv)
    j = np.arange(s.ngrtot)
    f = grid_points(ngrid, j)
    ig = np.concatenate([np.arange(5), rng.choice(s.ngvec, 10, replace=False)])
    ph = np.exp(-2j * np.pi * (s.ivg[:, ig].T.astype(float) @ f))   # (nig, ngrtot)
    a2 = 0.5 * oracle.ALPHA ** 2
    for got, fn in ((oracle.genveffig(s, veffir, cfunir), veffir * cfunir),
                    (oracle.genmeffig(s, veffir, cfunir, True)[0], cfunir / (1 - veffir * a2)),
                    (oracle.genmeffig(s, veffir, cfunir, True)[1],
                     cfunir / (1 - veffir * a2) ** 2)):
        assert got.shape == (s.ngvec,)
        want = ph @ fn / s.ngrtot
        np.testing.assert_allclose(got[ig], want, rtol=0, atol=1e-13)
    assert oracle.genmeffig(s, veffir, cfunir, False)[1] is None


def test_genveffig_of_a_plane_wave():
    s = _cases.system("tiny")
    ngrid = oracle.ngrid_of(s.intgv)
    f = grid_points(ngrid, np.arange(s.ngrtot))
    g0 = 7                                                        # some G != 0
    veffir = np.cos(2 * np.pi * (s.ivg[:, g0].astype(float) @ f))
    got = oracle.genveffig(s, veffir, np.ones(s.ngrtot))
    want = np.zeros(s.ngvec, dtype=complex)
    want[g0] = 0.5
    m = -s.ivg[:, g0]
    neg = s.ivgig[m[0] - s.intgv[0, 0], m[1] - s.intgv[1, 0], m[2] - s.intgv[2, 0]] - 1
    want[neg] += 0.5
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)
