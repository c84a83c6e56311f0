"""Write tests/golden/fft_ref.npz: input/output vectors of the reference's own
3-D FFT interface, src/zfftifc.f90 built without FFTW so that it calls
src/fftlib/cfftnd.f90 (modified FFTPACK5), both compiled by the library
oracle/Makefile builds with flang (oracle/_ref/libref_srclib.so).

The vectors pin the oracle's zfftifc restatement (index order with the first
grid index fastest, exp(-i...) for sgn = -1, the 1/N scale of the forward
transform only) wherever the tests run.  Run from the repository root:

    make -C oracle ref && python tests/golden/make_fft_vectors.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_ref.npz")

# grids with radices 2, 3, 5, a unit dimension and odd lengths
GRIDS = [(4, 6, 5), (6, 10, 8), (3, 1, 7), (8, 9, 10), (2, 2, 2)]


def main():
    R = oracle.ref_lib()
    if R is None or not hasattr(R, "zfftifc_"):
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` first")
    rng = np.random.default_rng(77)
    out = {}
    for i, g in enumerate(GRIDS):
        n = int(np.prod(g))
        z = rng.normal(size=n) + 1j * rng.normal(size=n)
        out[f"grid{i}"] = np.array(g, dtype=np.int32)
        out[f"in{i}"] = z
        for sgn, tag in ((-1, "fwd"), (1, "bwd")):
            w = np.ascontiguousarray(z.copy())
            ng = np.array(g, dtype=np.int32)
            R.zfftifc_(C.byref(C.c_int(3)), C.c_void_p(ng.ctypes.data), C.byref(C.c_int(sgn)),
                       C.c_void_p(w.ctypes.data))
            out[f"{tag}{i}"] = w
    np.savez_compressed(OUT, ngrids=len(GRIDS), **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
