"""Write tests/golden/radial_ref.npz: input/output vectors of the reference's
own radial-integral and angular-coupling routines, computed by the library
oracle/Makefile builds from the reference sources with flang
(oracle/_ref/libref_srclib.so):

* src/src_lib/spline.f90 and src/src_lib/fderiv.f90 (m = -1, the integral
  mt_kin/mt_pot take of every radial product);
* src/math/wigner3j_symbol.f90: wigner3j, gaunt_yyy and gaunt_yry (the
  complex Gaunt coefficients behind gntnonz, init1.f90:563-609).

They pin the oracle's restatement of these routines wherever the tests run.
Run from the repository root:

    make -C oracle ref && python tests/golden/make_radial_vectors.py
"""
import ctypes as C
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "radial_ref.npz")
LG = 4          # l1, l2, l3 <= LG for the Gaunt tables
LW = 3          # j1, j2, j3 <= LW for the 3j table


def ip(v):
    return C.byref(C.c_int(int(v)))


def grids(rng):
    t = np.arange(40) / 39
    logm = 1e-5 * np.exp(t * math.log(2.0 / 1e-5))
    return [logm, np.linspace(0.0, 1.5, 12), np.sort(rng.uniform(0, 3, 9)),
            np.array([0.1, 0.4, 0.5]), np.linspace(0.2, 2.0, 250) ** 1.3]


def main():
    R = oracle.ref_lib()
    if R is None or not hasattr(R, "ref_gaunt_yry"):
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` first")
    rng = np.random.default_rng(31)
    out = {}
    gs = grids(rng)
    out["ngrids"] = len(gs)
    for i, x in enumerate(gs):
        x = np.ascontiguousarray(x)
        n = x.size
        f = np.ascontiguousarray(np.sin(3 * x) * np.exp(-x) + 0.2 * rng.normal(size=n))
        cf = np.zeros((3, n), order="F")
        R.spline_(ip(n), C.c_void_p(x.ctypes.data), ip(1), C.c_void_p(f.ctypes.data),
                  C.c_void_p(cf.ctypes.data))
        g = np.zeros(n)
        cf2 = np.zeros((3, n), order="F")
        R.fderiv_(ip(-1), ip(n), C.c_void_p(x.ctypes.data), C.c_void_p(f.ctypes.data),
                  C.c_void_p(g.ctypes.data), C.c_void_p(cf2.ctypes.data))
        out[f"x{i}"], out[f"f{i}"], out[f"cf{i}"], out[f"g{i}"] = x, f, cf, g
    w3 = []
    v = np.zeros(2)
    for j1 in range(LW + 1):
        for j2 in range(LW + 1):
            for j3 in range(LW + 1):
                for m1 in range(-j1, j1 + 1):
                    for m2 in range(-j2, j2 + 1):
                        for m3 in range(-j3, j3 + 1):
                            R.ref_wigner3j(j1, j2, j3, m1, m2, m3, C.c_void_p(v.ctypes.data))
                            w3.append((j1, j2, j3, m1, m2, m3, v[0]))
    out["w3j"] = np.array(w3)
    n = (LG + 1) ** 2
    yry = np.zeros((n, n, n), dtype=np.complex128)    # [lm1, lm2, lm3]
    yyy = np.zeros((n, n, n))
    for l1 in range(LG + 1):
        for m1 in range(-l1, l1 + 1):
            for l2 in range(LG + 1):
                for m2 in range(-l2, l2 + 1):
                    for l3 in range(LG + 1):
                        for m3 in range(-l3, l3 + 1):
                            a, b, c = l1 * (l1 + 1) + m1, l2 * (l2 + 1) + m2, l3 * (l3 + 1) + m3
                            R.ref_gaunt_yry(l1, l2, l3, m1, m2, m3, C.c_void_p(v.ctypes.data))
                            yry[a, b, c] = complex(v[0], v[1])
                            R.ref_gaunt_yyy(l1, l2, l3, m1, m2, m3, C.c_void_p(v.ctypes.data))
                            yyy[a, b, c] = v[0]
    out["lg"] = LG
    out["gaunt_yry"] = yry
    out["gaunt_yyy"] = yyy
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
