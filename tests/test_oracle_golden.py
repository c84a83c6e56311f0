"""The oracle's leaf routines against committed vectors of the reference's
own routines (src/src_lib/{sbessel,sbesseldm,genylm,polynom,sphcrd,factnm,
factr}.f90 compiled with flang; tests/golden/make_leaf_vectors.py).  Unlike
the live comparisons in test_oracle_special.py these run wherever the tests
run, with or without the reference tree."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle

L = oracle.lib()
G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "leaf_ref.npz"))
LMAX = int(G["lmax"])


def _vec(fn, *args, n, dtype=np.float64):
    out = np.zeros(n, dtype=dtype)
    assert fn(*args, out.ctypes.data) == 0
    return out


def test_sbessel_matches_reference_vectors():
    for i, x in enumerate(G["xs"]):
        got = _vec(L.orc_sbessel, LMAX, C.c_double(x), n=LMAX + 1)
        np.testing.assert_allclose(got, G["sbessel"][i], rtol=4e-16, atol=1e-300,
                                   err_msg=f"x={x}")


@pytest.mark.parametrize("m", [1, 2, 3])
def test_sbesseldm_matches_reference_vectors(m):
    for i, x in enumerate(G["xs"]):
        got = _vec(L.orc_sbesseldm, m, LMAX, C.c_double(x), n=LMAX + 1)
        np.testing.assert_allclose(got, G["sbesseldm"][m - 1, i], rtol=1e-15, atol=1e-300,
                                   err_msg=f"x={x}")


def test_genylm_matches_reference_vectors():
    for i, tp in enumerate(G["angles"]):
        tp = np.ascontiguousarray(tp)
        got = _vec(L.orc_genylm, LMAX, tp.ctypes.data, n=(LMAX + 1) ** 2, dtype=np.complex128)
        np.testing.assert_allclose(got, G["genylm"][i], rtol=0, atol=1e-16)


def test_polynom_matches_reference_vectors():
    for j, npts in enumerate(range(2, 7)):
        xa = np.ascontiguousarray(G["poly_xa"][j, :npts])
        ya = np.ascontiguousarray(G["poly_ya"][j, :npts])
        for m in range(-1, npts + 1):
            c = np.zeros(npts)
            got = L.orc_polynom(m, npts, xa.ctypes.data, ya.ctypes.data, c.ctypes.data, 2.0)
            assert got == pytest.approx(G["polynom"][j, m + 1], rel=1e-14, abs=1e-300), (npts, m)


def test_sphcrd_matches_reference_vectors():
    for v, want in zip(G["sph_v"], G["sphcrd"]):
        v = np.ascontiguousarray(v)
        r, tp = C.c_double(), np.zeros(2)
        L.orc_sphcrd(v.ctypes.data, C.addressof(r), tp.ctypes.data)
        np.testing.assert_array_equal([r.value, tp[0], tp[1]], want)


def test_factorials_match_reference_vectors():
    L.orc_factnm.restype = C.c_double
    L.orc_factr.restype = C.c_double
    for n in range(40):
        for k, m in enumerate((1, 2, 3)):
            assert L.orc_factnm(n, m) == G["factnm"][n, k], (n, m)
    for n in range(20):
        for d in range(20):
            assert L.orc_factr(n, d) == G["factr"][n, d], (n, d)


def test_planted_bessel_fault_is_caught_by_the_vectors():
    oracle.set_fault(1)
    try:
        bad = [not np.allclose(_vec(L.orc_sbessel, LMAX, C.c_double(x), n=LMAX + 1),
                               G["sbessel"][i], rtol=4e-16, atol=1e-300)
               for i, x in enumerate(G["xs"])]
    finally:
        oracle.set_fault(0)
    assert any(bad)
