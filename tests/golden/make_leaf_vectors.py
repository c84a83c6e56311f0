"""Write tests/golden/leaf_ref.npz: input/output vectors of the reference's
own leaf routines (src/src_lib/sbessel.f90, sbesseldm.f90, genylm.f90,
polynom.f90, sphcrd.f90, factnm.f90, factr.f90), computed by the library
oracle/Makefile builds from the reference sources with flang
(oracle/_ref/libref_srclib.so).

The fixture lets the oracle's restatement of these routines be checked
against the reference wherever the tests run, including where the reference
tree or a Fortran compiler is absent.  Run from the repository root:

    make -C oracle ref && python tests/golden/make_leaf_vectors.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "leaf_ref.npz")

XS = np.array([0.0, 1e-9, 5e-9, 1e-7, 0.01, 0.5, 0.999, 1.0, 1.0000001, 1.5, 2.9999, 3.0,
               3.0001, 4.7, 7.999, 8.0, 8.001, 9.0, 9.99, 10.0, 10.01, 12.5, 18.0, 31.0])
LMAX = 12
ANGLES = np.array([[0.0, 0.0], [1e-12, 0.3], [0.3, 0.0], [0.7, 1.1], [1.5707963267948966, 2.5],
                   [2.1, 4.0], [3.0, 5.9], [3.141592653589793, 1.0]])


def by_ref(*args):
    out = []
    for a in args:
        if isinstance(a, np.ndarray):
            out.append(C.c_void_p(a.ctypes.data))
        elif isinstance(a, (int, np.integer)):
            out.append(C.byref(C.c_int(int(a))))
        else:
            out.append(C.byref(C.c_double(float(a))))
    return out


def main():
    R = oracle.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` first")
    sb = np.zeros((len(XS), LMAX + 1))
    sbdm = np.zeros((3, len(XS), LMAX + 1))
    for i, x in enumerate(XS):
        row = np.zeros(LMAX + 1)
        R.sbessel_(*by_ref(LMAX, x, row))
        sb[i] = row
        for m in (1, 2, 3):
            row = np.zeros(LMAX + 1)
            R.sbesseldm_(*by_ref(m, LMAX, x, row))
            sbdm[m - 1, i] = row
    ylm = np.zeros((len(ANGLES), (LMAX + 1) ** 2), dtype=np.complex128)
    for i, tp in enumerate(ANGLES):
        row = np.zeros((LMAX + 1) ** 2, dtype=np.complex128)
        R.genylm_(*by_ref(LMAX, np.ascontiguousarray(tp), row))
        ylm[i] = row
    rng = np.random.default_rng(2024)
    poly_xa = np.sort(rng.uniform(1.0, 2.0, (5, 6)), axis=1)
    poly_ya = rng.uniform(-1.0, 1.0, (5, 6))
    R.polynom_.restype = C.c_double
    poly = np.zeros((5, 8))          # np = 2..6, m = -1..6 (m <= np)
    for j, npts in enumerate(range(2, 7)):
        for m in range(-1, npts + 1):
            c = np.zeros(npts)
            xa = np.ascontiguousarray(poly_xa[j, :npts])
            ya = np.ascontiguousarray(poly_ya[j, :npts])
            poly[j, m + 1] = R.polynom_(*by_ref(m, npts, xa, ya, c, 2.0))
    vs = np.ascontiguousarray(np.concatenate([rng.normal(size=(12, 3)),
                                              [[0, 0, 0], [0, 0, 2.0], [0, 0, -1.0]]]))
    sph = np.zeros((len(vs), 3))
    for i, v in enumerate(vs):
        r, tp = C.c_double(), np.zeros(2)
        v = np.ascontiguousarray(v)
        R.sphcrd_(C.c_void_p(v.ctypes.data), C.byref(r), C.c_void_p(tp.ctypes.data))
        sph[i] = (r.value, tp[0], tp[1])
    R.factnm_.restype = C.c_double
    R.factr_.restype = C.c_double
    factnm = np.array([[R.factnm_(*by_ref(n, m)) for m in (1, 2, 3)] for n in range(40)])
    factr = np.array([[R.factr_(*by_ref(n, d)) for d in range(20)] for n in range(20)])
    np.savez_compressed(OUT, xs=XS, lmax=LMAX, sbessel=sb, sbesseldm=sbdm, angles=ANGLES,
                        genylm=ylm, poly_xa=poly_xa, poly_ya=poly_ya, polynom=poly,
                        sph_v=vs, sphcrd=sph, factnm=factnm, factr=factr)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
