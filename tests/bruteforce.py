"""Independent NumPy statements of the quantities the oracle computes, used to
check it on small inputs.  Written from the mathematics of the APW matching
conditions and the LAPW H/O matrix elements, not from the oracle's loops."""
import math

import numpy as np
import scipy.special as ss


def boundary_matrix(s, ias, l):
    """D[i, j] = d^i u_j / dr^i at R from the last four radial points (the
    apwfr tail is a cubic there)."""
    is_ = int(np.searchsorted(np.cumsum(s.natoms), ias, side="right"))
    nr = s.nrmt[is_]
    r = s.spr[nr - 4:nr, is_]
    R = s.rmt[is_]
    p = s.apword[l, is_]
    D = np.zeros((p, p))
    for j in range(p):
        poly = np.polynomial.Polynomial.fit(r - R, s.apwfr[nr - 4:nr, 0, j, l, ias], 3)
        poly = poly.convert()
        for i in range(p):
            D[i, j] = poly.deriv(i)(0.0) if i else poly(0.0)
    return D


def bessel_deriv(l, x, m):
    if m == 0:
        return ss.spherical_jn(l, x)
    if m == 1:
        return ss.spherical_jn(l, x, derivative=True)
    if m == 2:
        j = ss.spherical_jn(l, x)
        jp = ss.spherical_jn(l, x, derivative=True)
        return -(2.0 / x) * jp - (1.0 - l * (l + 1) / x ** 2) * j
    raise ValueError(m)


def rhs(s, k, ias, l, m, io):
    """b_io(G) = 4 pi i^l / sqrt(Omega) |G+k|^io j_l^(io)(|G+k| R) S(G) Y*_lm."""
    is_ = int(np.searchsorted(np.cumsum(s.natoms), ias, side="right"))
    ng = k.ngk
    q = k.gpc[:ng]
    x = q * s.rmt[is_]
    th, ph = k.tpgpc[0, :ng], k.tpgpc[1, :ng]
    y = ss.sph_harm_y(l, m, th, ph)
    return (4 * math.pi * (1j ** l) / math.sqrt(s.omega) * q ** io * bessel_deriv(l, x, io)
            * k.sfacgp[:ng, ias] * np.conj(y))


def stacked_A(s, apwalm, ias, ngp):
    """A_ias (M x ngp), rows f = (l <= lmaxmat, m, io) with io fastest."""
    is_ = int(np.searchsorted(np.cumsum(s.natoms), ias, side="right"))
    rows = []
    for l in range(s.lmaxmat + 1):
        for m in range(-l, l + 1):
            for io in range(s.apword[l, is_]):
                rows.append(apwalm[:ngp, io, l * (l + 1) + m, ias])
    return np.array(rows)


def species_of(s, ias):
    return int(np.searchsorted(np.cumsum(s.natoms), ias, side="right"))


def interstitial(s, k):
    """(H_I, O_I) upper triangles from the G-G' map (dense NumPy)."""
    ng = k.ngk
    gv = s.ivg[:, k.igpig[:ng] - 1].T                     # (ng, 3)
    d = gv[:, None, :] - gv[None, :, :]                   # G_i - G_j
    lo, hi = s.intgv[:, 0], s.intgv[:, 1]
    inb = np.all((d >= lo) & (d <= hi), axis=2)
    ig = np.zeros((ng, ng), dtype=np.int64)
    dd = d[inb]
    ig[inb] = s.ivgig[dd[:, 0] - lo[0], dd[:, 1] - lo[1], dd[:, 2] - lo[2]]
    ok = (ig > 0) & (ig <= s.ngvec) & np.triu(np.ones((ng, ng), bool))
    v = k.vgpc[:, :ng].T
    t1 = 0.5 * (v @ v.T)
    cf = np.zeros((ng, ng), complex)
    ve = np.zeros((ng, ng), complex)
    cf[ok] = s.cfunig[ig[ok] - 1]
    ve[ok] = s.veffig[ig[ok] - 1]
    kin = cf
    if not s.valence_rel_none:
        kin = np.zeros((ng, ng), complex)
        kin[ok] = s.meffig[ig[ok] - 1]
    HI = np.where(ok, ve + t1 * kin, 0)
    OI = cf.copy()
    if s.h1on:
        a2 = 0.5 / 137.03599911 ** 2
        m2 = np.zeros((ng, ng), complex)
        m2[ok] = s.m2effig[ig[ok] - 1]
        OI = np.where(ok, cf + a2 * t1 * m2, 0)
    return HI, OI


def setup(s, k, apwalm):
    """Full H and O of one k-point: muffin-tin + LO + interstitial, upper
    triangle kept, lower = conj(upper)."""
    ng = k.ngk
    n = ng + s.nlotot
    H = np.zeros((n, n), complex)
    O = np.zeros((n, n), complex)
    lob = 0
    for ias in range(s.natmtot):
        is_ = species_of(s, ias)
        A = stacked_A(s, apwalm, ias, ng)                  # M x ng
        M = A.shape[0]
        aa = s.aa[:M, :M, ias]
        H[:ng, :ng] += A.conj().T @ aa @ A
        # overlap: A^H (1 + h1) A, h1 couples io within one (l, m)
        W = np.eye(M)
        if s.h1aa is not None:
            f = 0
            for l in range(s.lmaxmat + 1):
                p = s.apword[l, is_]
                for m in range(2 * l + 1):
                    for io2 in range(p):
                        for io1 in range(p):
                            W[f + io2, f + io1] += s.h1aa[io1, io2, l, ias]
                    f += p
        O[:ng, :ng] += A.conj().T @ (W @ A)
        los = s.losize[is_]
        if los:
            rows = slice(ng + lob, ng + lob + los)
            H[rows, :ng] = s.loa[:los, :M, ias] @ A
            H[rows, rows] = s.lolo[:los, :los, ias]
            off = 0
            for ilo in range(s.nlorb[is_]):
                l = s.lorbl[ilo, is_]
                for mi in range(2 * l + 1):
                    col = ng + lob + off + mi
                    lm = l * l + mi
                    val = 0
                    for io in range(s.apword[l, is_]):
                        sv = s.oalo[io, ilo, ias] + (0 if s.h1loa is None else s.h1loa[io, ilo, ias])
                        val = val + np.conj(apwalm[:ng, io, lm, ias] * sv)
                    O[:ng, col] = val
                off2 = 0
                for ilo2 in range(s.nlorb[is_]):
                    l2 = s.lorbl[ilo2, is_]
                    if l2 == l:
                        sv = s.ololo[ilo, ilo2, ias] + (0 if s.h1lolo is None else s.h1lolo[ilo, ilo2, ias])
                        for mi in range(2 * l + 1):
                            O[ng + lob + off + mi, ng + lob + off2 + mi] += sv
                    off2 += 2 * l2 + 1
                off += 2 * l + 1
            # APW-lo H block lives in the lower-left: mirror to upper-right
            H[:ng, rows] = H[rows, :ng].conj().T
        lob += los
    HI, OI = interstitial(s, k)
    H[:ng, :ng] += HI
    O[:ng, :ng] += OI
    iu = np.triu_indices(n, 1)
    H.T[iu] = H[iu].conj()
    O.T[iu] = O[iu].conj()
    return H, O
