"""Pins of the oracle's special functions (CPU only).

Each routine is checked two independent ways:
  * against the reference's own source compiled by flang (oracle/_ref,
    src/src_lib/*.f90), where that library was built in this container;
  * against what the mathematics fixes: SciPy's spherical Bessel functions
    and spherical harmonics (Condon-Shortley phase), the Bessel ODE for the
    second derivative, exact polynomial data for polynom, numpy.linalg.solve
    for zgesv.
"""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.special as ss

import oracle

L = oracle.lib()
REF = oracle.ref_lib()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (no reference tree)")

XS = [0.0, 1e-9, 5e-9, 1e-7, 0.01, 0.5, 0.999, 1.0, 1.0000001, 1.5, 2.9999, 3.0, 3.0001,
      4.7, 7.999, 8.0, 8.001, 9.0, 9.99, 10.0, 10.01, 12.5, 18.0, 31.0]
LMAXES = [0, 1, 2, 3, 8, 10, 12]


def orc_sbessel(lmax, x):
    out = np.zeros(lmax + 1)
    assert L.orc_sbessel(lmax, x, out.ctypes.data) == 0
    return out


def orc_sbesseldm(m, lmax, x):
    out = np.zeros(lmax + 1)
    assert L.orc_sbesseldm(m, lmax, x, out.ctypes.data) == 0
    return out


def ref_call(name, *args):
    """Call a flang-compiled reference routine (all arguments by reference)."""
    f = getattr(REF, name + "_")
    conv = []
    for a in args:
        if isinstance(a, np.ndarray):
            conv.append(C.c_void_p(a.ctypes.data))
        elif isinstance(a, (int, np.integer)):
            conv.append(C.byref(C.c_int(int(a))))
        else:
            conv.append(C.byref(C.c_double(float(a))))
    return f, conv


@needs_ref
@pytest.mark.parametrize("lmax", LMAXES)
def test_sbessel_matches_reference_build(lmax):
    for x in XS:
        ref = np.zeros(lmax + 1)
        f, a = ref_call("sbessel", lmax, x, ref)
        f(*a)
        got = orc_sbessel(lmax, x)
        np.testing.assert_allclose(got, ref, rtol=4e-16, atol=1e-300, err_msg=f"x={x}")


@needs_ref
@pytest.mark.parametrize("m", [1, 2, 3])
@pytest.mark.parametrize("lmax", [0, 3, 8, 10])
def test_sbesseldm_matches_reference_build(m, lmax):
    for x in XS:
        ref = np.zeros(lmax + 1)
        f, a = ref_call("sbesseldm", m, lmax, x, ref)
        f(*a)
        got = orc_sbesseldm(m, lmax, x)
        np.testing.assert_allclose(got, ref, rtol=1e-15, atol=1e-300, err_msg=f"x={x}")


@pytest.mark.parametrize("lmax", LMAXES)
def test_sbessel_against_scipy(lmax):
    for x in XS:
        got = orc_sbessel(lmax, x)
        want = np.array([ss.spherical_jn(l, x) for l in range(lmax + 1)])
        scale = np.maximum(np.abs(want), 1e-300)
        # relative accuracy where the value is not an underflowing tail
        ok = np.abs(want) > 1e-250
        assert np.all(np.abs(got - want)[ok] <= 1e-13 * scale[ok] + 1e-16), (x, got, want)


@pytest.mark.parametrize("lmax", [0, 2, 8, 10])
def test_sbesseldm_first_derivative_against_scipy(lmax):
    for x in XS[1:]:
        got = orc_sbesseldm(1, lmax, x)
        want = np.array([ss.spherical_jn(l, x, derivative=True) for l in range(lmax + 1)])
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-14, err_msg=f"x={x}")


@pytest.mark.parametrize("lmax", [0, 3, 10])
def test_sbesseldm_second_derivative_from_ode(lmax):
    # j'' = -(2/x) j' - (1 - l(l+1)/x^2) j
    for x in [0.3, 0.9, 1.2, 2.5, 6.0, 9.0]:
        got = orc_sbesseldm(2, lmax, x)
        l = np.arange(lmax + 1)
        j = ss.spherical_jn(l, x)
        jp = ss.spherical_jn(l, x, derivative=True)
        want = -(2.0 / x) * jp - (1.0 - l * (l + 1) / x ** 2) * j
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-13, err_msg=f"x={x}")


def orc_genylm(lmax, theta, phi):
    out = np.zeros((lmax + 1) ** 2, dtype=np.complex128)
    tp = np.array([theta, phi])
    assert L.orc_genylm(lmax, tp.ctypes.data, out.ctypes.data) == 0
    return out


ANGLES = [(0.0, 0.0), (1e-9, 0.3), (0.4, 5.9), (1.2, 2.0), (math.pi / 2, 1.0),
          (2.7, 4.4), (math.pi, 0.0), (3.0, 6.2)]


@pytest.mark.parametrize("lmax", [0, 1, 4, 10, 12])
def test_genylm_against_scipy(lmax):
    for th, ph in ANGLES:
        got = orc_genylm(lmax, th, ph)
        want = np.array([ss.sph_harm_y(l, m, th, ph)
                         for l in range(lmax + 1) for m in range(-l, l + 1)])
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-14, err_msg=f"{th},{ph}")


@needs_ref
@pytest.mark.parametrize("lmax", [0, 3, 10])
def test_genylm_matches_reference_build(lmax):
    for th, ph in ANGLES:
        ref = np.zeros((lmax + 1) ** 2, dtype=np.complex128)
        tp = np.array([th, ph])
        f, a = ref_call("genylm", lmax, tp, ref)
        f(*a)
        np.testing.assert_allclose(orc_genylm(lmax, th, ph), ref, rtol=0, atol=1e-16)


def test_genylm_orthonormal_on_sphere():
    # Gauss-Legendre x uniform-phi quadrature integrates Y*Y' exactly for l<=6
    lmax = 6
    xg, wg = np.polynomial.legendre.leggauss(16)
    phis = np.arange(32) * (2 * math.pi / 32)
    n = (lmax + 1) ** 2
    G = np.zeros((n, n), dtype=np.complex128)
    for x, w in zip(xg, wg):
        for ph in phis:
            y = orc_genylm(lmax, math.acos(x), ph)
            G += w * (2 * math.pi / 32) * np.outer(y.conj(), y)
    np.testing.assert_allclose(G, np.eye(n), atol=1e-13)


def orc_polynom(m, xa, ya, x):
    c = np.zeros(len(xa))
    return L.orc_polynom(m, len(xa), np.ascontiguousarray(xa).ctypes.data,
                         np.ascontiguousarray(ya).ctypes.data, c.ctypes.data, x)


@pytest.mark.parametrize("np_", [1, 2, 3, 4, 5, 7])
def test_polynom_exact_on_polynomials(np_):
    rng = np.random.default_rng(np_)
    coef = rng.uniform(-1, 1, np_)
    xa = np.linspace(0.5, 2.0, np_) + rng.uniform(-0.02, 0.02, np_)
    p = np.polynomial.Polynomial(coef)
    ya = p(xa)
    x = 2.0
    for m in range(0, np_ + 1):
        want = p.deriv(m)(x) if m > 0 else p(x)
        got = orc_polynom(m, xa, ya, x)
        # divided differences lose ~one digit per derivative order
        assert abs(got - want) <= 1e-13 * 10 ** m * max(1.0, abs(want)), (m, got, want)
    want = p.integ()(x) - p.integ()(xa[0])
    assert abs(orc_polynom(-1, xa, ya, x) - want) <= 1e-9


@needs_ref
@pytest.mark.parametrize("np_", [2, 3, 4, 5, 6])
def test_polynom_matches_reference_build(np_):
    rng = np.random.default_rng(10 + np_)
    xa = np.sort(rng.uniform(1.0, 2.0, np_))
    ya = rng.uniform(-1, 1, np_)
    REF.polynom_.restype = C.c_double
    for m in range(-1, np_ + 1):
        c = np.zeros(np_)
        f, a = ref_call("polynom", m, np_, xa, ya, c, 2.0)
        ref = f(*a)
        got = orc_polynom(m, xa, ya, 2.0)
        assert got == pytest.approx(ref, rel=1e-14, abs=1e-300), m


@needs_ref
def test_sphcrd_matches_reference_build():
    rng = np.random.default_rng(3)
    vs = list(rng.normal(size=(20, 3))) + [np.zeros(3), np.array([0, 0, 2.0]),
                                          np.array([0, 0, -1.0]), np.array([1e-15, 0, 1.0])]
    for v in vs:
        v = np.ascontiguousarray(v, dtype=float)
        r1, tp1 = C.c_double(), np.zeros(2)
        REF.sphcrd_(C.c_void_p(v.ctypes.data), C.byref(r1), C.c_void_p(tp1.ctypes.data))
        r2, tp2 = C.c_double(), np.zeros(2)
        L.orc_sphcrd(v.ctypes.data, C.addressof(r2), tp2.ctypes.data)
        assert r1.value == r2.value
        np.testing.assert_array_equal(tp1, tp2)


def test_sphcrd_round_trip():
    rng = np.random.default_rng(4)
    for v in rng.normal(size=(50, 3)):
        v = np.ascontiguousarray(v)
        r, tp = C.c_double(), np.zeros(2)
        L.orc_sphcrd(v.ctypes.data, C.addressof(r), tp.ctypes.data)
        back = r.value * np.array([math.sin(tp[0]) * math.cos(tp[1]),
                                   math.sin(tp[0]) * math.sin(tp[1]), math.cos(tp[0])])
        np.testing.assert_allclose(back, v, atol=1e-14)
        assert 0.0 <= tp[1] < 2 * math.pi


@needs_ref
def test_factorials_match_reference_build():
    REF.factnm_.restype = C.c_double
    REF.factr_.restype = C.c_double
    L.orc_factnm.restype = C.c_double
    for n in range(0, 40):
        for m in (1, 2, 3):
            f, a = ref_call("factnm", n, m)
            assert L.orc_factnm(n, m) == f(*a), (n, m)
    for n in range(0, 20):
        for d in range(0, 20):
            f, a = ref_call("factr", n, d)
            assert L.orc_factr(n, d) == f(*a), (n, d)


def test_factorials_closed_form():
    for n in range(0, 20):
        assert L.orc_factnm(n, 1) == math.factorial(n)
        dbl = 1
        for k in range(n, 0, -2):
            dbl *= k
        assert L.orc_factnm(n, 2) == dbl
        for d in range(0, n + 1):
            assert L.orc_factr(n, d) == pytest.approx(math.factorial(n) / math.factorial(d))


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_zgesv_against_numpy(n):
    rng = np.random.default_rng(n)
    for trial in range(20):
        a = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)) * (trial % 2)
        if trial == 3 and n > 1:
            a[0, 0] = 0.0  # forces a pivot
        b = rng.normal(size=(n, 4)) + 1j * rng.normal(size=(n, 4))
        af, bf = np.array(a, order="F"), np.array(b, order="F")  # copies
        ipiv = np.zeros(n, dtype=np.int32)
        info = L.orc_zgesv(n, 4, af.ctypes.data, n, ipiv.ctypes.data, bf.ctypes.data, n)
        assert info == 0
        np.testing.assert_allclose(bf, np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)
