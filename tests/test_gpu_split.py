"""Split-precision (int8 Ozaki scheme II) H/O setup on the GPU.

1. Exactness of the modular machinery: the split result must equal the FP64
   oracle plus the effect of rounding A and aa.A to integers after scaling
   each G column to a 2-norm below 2^(c-1), computed here independently in
   NumPy.  Any error in the residues,
   the int8 GEMMs, the Karatsuba recombination or the Garner reconstruction
   shows up far above the 1e-12 tolerance.
2. Accuracy against FP64: relative Frobenius error of H and O, and the
   eigenvalues of H x = e O x within 1 uHa (the stated bar for this path).
"""
import math

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from exciting_amd import lapw
from tests import _cases, bruteforce

PRIMES = [251, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191]


def oz_bits(nmod, K):
    """c: the largest value with 2^(2c) < prod(p)/2 (Cauchy-Schwarz bound on
    Re, Im of x^H b for columns of 2-norm below 2^c)."""
    lm = sum(math.log2(p) for p in PRIMES[:nmod])
    return int(min(50, math.floor((lm - 1.0) * 0.5 - 1e-9)))


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def trunc(M, bits):
    """Round the K x G matrix M to integers with a power-of-two scale 2^s per
    column g, s = bits - e - 1 with ||M[:, g]|| < 2^e; returns
    (M_int, 2^-s)."""
    nrm = np.linalg.norm(M, axis=0)
    e = np.frexp(nrm)[1] + 1
    s = np.where(nrm > 0, bits - e, 0)
    sc = np.ldexp(1.0, s)
    return np.rint(M.real * sc) + 1j * np.rint(M.imag * sc), 1.0 / sc


def split_expected(s, k, A, H0, O0, nmod):
    """FP64 oracle + (truncated-operand product - exact product) on the APW
    block, upper triangle, mirrored."""
    ng = k.ngk
    X = np.concatenate([bruteforce.stacked_A(s, A, a, ng) for a in range(s.natmtot)], axis=0)
    B = np.concatenate([s.aa[:bruteforce.stacked_A(s, A, a, ng).shape[0],
                             :bruteforce.stacked_A(s, A, a, ng).shape[0], a]
                        @ bruteforce.stacked_A(s, A, a, ng) for a in range(s.natmtot)], axis=0)
    bits = oz_bits(nmod, X.shape[0])
    Xi, xs = trunc(X, bits)
    Bi, bs = trunc(B, bits)
    dH = (Xi.conj().T @ Bi) * np.outer(xs, bs) - X.conj().T @ B
    dO = (Xi.conj().T @ Xi) * np.outer(xs, xs) - X.conj().T @ X
    H, O = H0.copy(), O0.copy()
    H[:ng, :ng] += np.triu(dH)
    O[:ng, :ng] += np.triu(dO)
    il = np.tril_indices(H.shape[0], -1)
    H[il] = H.T[il].conj()
    O[il] = O.T[il].conj()
    return H, O, bits


@pytest.fixture(scope="module")
def ctxs():
    cache = {}

    def get(name):
        if name not in cache:
            s = _cases.system(name)
            print(f"[gpu-split] case {name}", flush=True)
            cache[name] = (s, oracle.System(s), lapw.Context(0).load(s))
        return cache[name]
    yield get
    for s, S, c in cache.values():
        c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name,nmod", [("tiny", 9), ("tiny_lmaxmat", 6), ("tiny_nolo", 12),
                                       ("fcc2", 9), ("fcc2", 7)])
def test_split_is_exact_on_truncated_operands(ctxs, name, nmod):
    s, S, ctx = ctxs(name)
    k = s.kpts[0]
    A = oracle.match(S, 0)
    H0, O0 = oracle.setup(S, 0, A)
    He, Oe, bits = split_expected(s, k, A, H0, O0, nmod)
    sysg = lapw.newsystem(k.ngk + s.nlotot)
    ctx.hamiltonandoverlapsetup(sysg, k.ngk, A, k.igpig, k.vgpc, flags=lapw.SPLIT(nmod))
    eh, eo = rel(sysg.hamilton.za, He), rel(sysg.overlap.za, Oe)
    dh, do = rel(sysg.hamilton.za, H0), rel(sysg.overlap.za, O0)
    print(f"  {name} n={nmod} b={bits}: vs truncated-exact H {eh:.1e} O {eo:.1e}; "
          f"vs FP64 H {dh:.1e} O {do:.1e}")
    assert eh <= 1e-12 and eo <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name,nmod,nst", [("fcc2", 9, 100), ("sc8_k0", 9, 200),
                                           ("sc8_k0", 8, 200)])
def test_split_accuracy_against_fp64(ctxs, name, nmod, nst):
    s, S, ctx = ctxs(name)
    for ik, k in enumerate(s.kpts):
        A = oracle.match(S, ik)
        H0, O0 = oracle.setup(S, ik, A)
        sysg = lapw.newsystem(k.ngk + s.nlotot)
        ctx.hamiltonandoverlapsetup(sysg, k.ngk, A, k.igpig, k.vgpc, flags=lapw.SPLIT(nmod))
        H, O = sysg.hamilton.za, sysg.overlap.za
        w0 = sl.eigh(H0, O0, lower=False, eigvals_only=True, subset_by_index=[0, nst - 1])
        w = sl.eigh(H, O, lower=False, eigvals_only=True, subset_by_index=[0, nst - 1])
        de = np.abs(w - w0).max()
        print(f"  {name} k{ik} n={nmod}: H {rel(H, H0):.1e} O {rel(O, O0):.1e} "
              f"max|dE| {de:.1e} Ha")
        assert rel(H, H0) < 1e-6 and rel(O, O0) < 1e-6
        assert de <= 1e-6
