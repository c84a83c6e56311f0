"""ctypes front end of the oracle (test infrastructure only).

The oracle is the CPU restatement of exciting's H/O setup in oracle.c.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module; it is the checker and the CPU baseline, never the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_srclib.so")

_i = C.c_int
_d = C.c_double
_p = C.c_void_p


class OrcSystem(C.Structure):
    _fields_ = [
        ("nspecies", _i), ("natmtot", _i), ("natoms", _p),
        ("lmaxapw", _i), ("lmaxmat", _i), ("apwordmax", _i), ("apword", _p),
        ("nlomax", _i), ("nlorb", _p), ("lorbl", _p), ("ngkmax", _i),
        ("omega", _d), ("rmt", _p), ("nrmt", _p), ("nrmtmax", _i), ("spr", _p),
        ("apwfr", _p),
        ("maxaa", _i), ("maxnlo", _i), ("aa", _p), ("loa", _p), ("lolo", _p),
        ("oalo", _p), ("ololo", _p), ("h1aa", _p), ("h1loa", _p), ("h1lolo", _p),
        ("ngrtot", _i), ("ngvec", _i), ("ivg", _p), ("intgv", _i * 6),
        ("ivgig", _p), ("cfunig", _p), ("veffig", _p), ("meffig", _p),
        ("m2effig", _p), ("valence_rel_none", _i), ("h1on", _i),
    ]


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "all"], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.orc_sbessel.argtypes = [_i, _d, _p]
        L.orc_sbesseldm.argtypes = [_i, _i, _d, _p]
        L.orc_genylm.argtypes = [_i, _p, _p]
        L.orc_polynom.argtypes = [_i, _i, _p, _p, _p, _d]
        L.orc_polynom.restype = _d
        L.orc_sphcrd.argtypes = [_p, _p, _p]
        L.orc_factnm.argtypes = [_i, _i]
        L.orc_factnm.restype = _d
        L.orc_factr.argtypes = [_i, _i]
        L.orc_factr.restype = _d
        L.orc_zgesv.argtypes = [_i, _i, _p, _i, _p, _p, _i]
        L.orc_match.argtypes = [_p, _i, _p, _p, _p, _p]
        for name in ("orc_hamiltonandoverlapsetup",):
            getattr(L, name).argtypes = [_p, _i, _p, _p, _p, _i, _p, _p, _p]
        for name in ("orc_hamiltonsetup", "orc_overlapsetup"):
            getattr(L, name).argtypes = [_p, _i, _p, _p, _p, _i, _p, _p]
        for name in ("orc_hmlistln", "orc_olpistln"):
            getattr(L, name).argtypes = [_p, _i, _p, _p, _i, _p]
        L.orc_copy_ul.argtypes = [_i, _p, _p]
        L.orc_set_fault.argtypes = [_i]
        L.orc_gencfunig.argtypes = [_i, _p, _p, _d, _i, _p, _p, _p, _d, _p]
        L.orc_spline.argtypes = [_i, _p, _p, _p]
        L.orc_fderiv_int.argtypes = [_i, _p, _p, _p]
        L.orc_wigner3j.argtypes = [_i] * 6
        L.orc_wigner3j.restype = _d
        L.orc_gaunt_yyy.argtypes = [_i] * 6
        L.orc_gaunt_yyy.restype = _d
        L.orc_gaunt_yry.argtypes = [_i] * 6 + [_p]
        L.orc_hmlint.argtypes = [_p, _i, _p, _p, _i, _p, _p, _p]
        L.orc_hmlint_parts.argtypes = [_p, _i, _p, _p, _i, _p, _p, _p, _i]
        L.orc_mt_so.argtypes = [_p, _i, _p, _p, _i, _p]
        L.orc_nlotot.argtypes = [_p]
        L.orc_maxapw.argtypes = [_p]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own src_lib routines compiled by flang, or None."""
    if not os.path.exists(REF_PATH):
        return None
    R = C.CDLL(REF_PATH)
    return R


def _ptr(a):
    return None if a is None else a.ctypes.data


class System:
    """Holds an OrcSystem and keeps the numpy arrays it points at alive."""

    def __init__(self, s):
        self.s = s
        keep = []

        def arr(x, dt):
            if x is None:
                return None
            y = np.asfortranarray(x, dtype=dt)
            keep.append(y)
            return y.ctypes.data

        o = OrcSystem()
        o.nspecies = s.nspecies
        o.natmtot = s.natmtot
        o.natoms = arr(s.natoms, np.int32)
        o.lmaxapw = s.lmaxapw
        o.lmaxmat = s.lmaxmat
        o.apwordmax = s.apwordmax
        o.apword = arr(s.apword, np.int32)
        o.nlomax = s.nlomax
        o.nlorb = arr(s.nlorb, np.int32)
        o.lorbl = arr(s.lorbl, np.int32)
        o.ngkmax = s.ngkmax
        o.omega = s.omega
        o.rmt = arr(s.rmt, np.float64)
        o.nrmt = arr(s.nrmt, np.int32)
        o.nrmtmax = s.nrmtmax
        o.spr = arr(s.spr, np.float64)
        o.apwfr = arr(s.apwfr, np.float64)
        o.maxaa = s.maxaa
        o.maxnlo = s.maxnlo
        o.aa = arr(s.aa, np.complex128)
        o.loa = arr(s.loa, np.complex128)
        o.lolo = arr(s.lolo, np.complex128)
        o.oalo = arr(s.oalo, np.float64)
        o.ololo = arr(s.ololo, np.float64)
        o.h1aa = arr(s.h1aa, np.float64)
        o.h1loa = arr(s.h1loa, np.float64)
        o.h1lolo = arr(s.h1lolo, np.float64)
        o.ngrtot = s.ngrtot
        o.ngvec = s.ngvec
        o.ivg = arr(s.ivg, np.int32)
        iv = np.asfortranarray(s.intgv, dtype=np.int32).ravel(order="F")
        for k in range(6):
            o.intgv[k] = int(iv[k])
        o.ivgig = arr(s.ivgig, np.int32)
        o.cfunig = arr(s.cfunig, np.complex128)
        o.veffig = arr(s.veffig, np.complex128)
        o.meffig = arr(s.meffig, np.complex128)
        o.m2effig = arr(s.m2effig, np.complex128)
        o.valence_rel_none = 1 if s.valence_rel_none else 0
        o.h1on = 1 if s.h1on else 0
        self.o = o
        self._keep = keep

    @property
    def ptr(self):
        return C.addressof(self.o)


def scipy_zgemm():
    """Function pointer of the optimised zgemm bundled with SciPy (OpenBLAS),
    or None."""
    try:
        import scipy.linalg.cython_blas as cb
    except Exception:
        return None
    cap = cb.__pyx_capi__["zgemm"]
    get = C.pythonapi.PyCapsule_GetPointer
    get.restype = C.c_void_p
    get.argtypes = [C.py_object, C.c_char_p]
    name = C.pythonapi.PyCapsule_GetName
    name.restype = C.c_char_p
    name.argtypes = [C.py_object]
    return get(cap, name(cap))


def match(sys_, ik: int) -> np.ndarray:
    """match.f90 for k-point ik: apwalm(ngkmax, apwordmax, lmmaxapw, natmtot)."""
    s = sys_.s
    k = s.kpts[ik]
    lmm = (s.lmaxapw + 1) ** 2
    apwalm = np.zeros((s.ngkmax, s.apwordmax, lmm, s.natmtot), dtype=np.complex128, order="F")
    gpc = np.ascontiguousarray(k.gpc)
    tp = np.asfortranarray(k.tpgpc)
    sf = np.asfortranarray(k.sfacgp)
    rc = lib().orc_match(sys_.ptr, k.ngk, gpc.ctypes.data, tp.ctypes.data, sf.ctypes.data,
                         apwalm.ctypes.data)
    if rc < 0:
        raise RuntimeError(f"orc_match failed: {rc}")
    return apwalm


def setup(sys_, ik: int, apwalm: np.ndarray, fast: bool = True):
    """hamiltonsetup + overlapsetup (+UL copy) for k-point ik: (H, O)."""
    s = sys_.s
    k = s.kpts[ik]
    n = k.ngk + s.nlotot
    h = np.zeros((n, n), dtype=np.complex128, order="F")
    o = np.zeros((n, n), dtype=np.complex128, order="F")
    zg = scipy_zgemm() if fast else None
    a = np.asfortranarray(apwalm)
    ig = np.ascontiguousarray(k.igpig, dtype=np.int32)
    vg = np.asfortranarray(k.vgpc)
    rc = lib().orc_hamiltonandoverlapsetup(sys_.ptr, k.ngk, a.ctypes.data, ig.ctypes.data,
                                           vg.ctypes.data, n, h.ctypes.data, o.ctypes.data, zg)
    if rc != 0:
        raise RuntimeError(f"orc_hamiltonandoverlapsetup failed: {rc}")
    return h, o


def set_fault(code: int) -> None:
    lib().orc_set_fault(code)


# --- interstitial producers (gencfun.f90, genveffig.f90, genmeffig.f90) ---

ALPHA = 1.0 / 137.03599911   # genmeffig.f90:29


def ngrid_of(intgv) -> tuple:
    """FFT grid from the G box: intgv(:,1) = n/2 - n + 1, intgv(:,2) = n/2
    (gridsize.f90), so n = intgv(:,2) - intgv(:,1) + 1."""
    iv = np.asarray(intgv)
    return tuple(int(x) for x in iv[:, 1] - iv[:, 0] + 1)


def igfft(ngrid, ivg) -> np.ndarray:
    """gengvec.f90:129-167 (genigfft), 0-based: the FFT grid point of each G
    vector, j_d = i_d (i_d >= 0) or ngrid_d + i_d, m = j3*n1*n2 + j2*n1 + j1."""
    n = np.asarray(ngrid, dtype=np.int64)
    iv = np.asarray(ivg, dtype=np.int64)
    j = np.where(iv >= 0, iv, n[:, None] + iv)
    return j[2] * n[0] * n[1] + j[1] * n[0] + j[0]


def zfftifc(ngrid, sgn: int, z: np.ndarray) -> np.ndarray:
    """zfftifc.f90:13-110 for nd = 3: z holds the grid with the first index
    fastest; sgn -1 is the forward transform (exp(-i...)) scaled by 1/N
    (:70-77), +1 the unscaled backward one.  Restated with numpy's FFT; the
    convention is pinned against the reference's own zfftifc/cfftnd by the
    golden vectors."""
    n1, n2, n3 = ngrid
    a = np.asarray(z, dtype=np.complex128).reshape(n3, n2, n1)
    if sgn == -1:
        out = np.fft.fftn(a) * (1.0 / a.size)
    elif sgn == 1:
        out = np.fft.ifftn(a) * a.size
    else:
        raise ValueError("sgn must be -1 or 1")
    return out.reshape(-1)


def gencfunig(ivg, bvec, omega, natoms, rmt, atposc, epslat=1e-6) -> np.ndarray:
    """gencfun.f90:96-159 through the C restatement orc_gencfunig.
    atposc: (natmtot, 3) Cartesian, species-major."""
    iv = np.ascontiguousarray(np.asarray(ivg).T, dtype=np.int32)       # (ng, 3)
    bv = np.asfortranarray(np.asarray(bvec, dtype=np.float64))
    na = np.ascontiguousarray(natoms, dtype=np.int32)
    rm = np.ascontiguousarray(rmt, dtype=np.float64)
    ap = np.ascontiguousarray(atposc, dtype=np.float64)
    out = np.zeros(iv.shape[0], dtype=np.complex128)
    lib().orc_gencfunig(iv.shape[0], iv.ctypes.data, bv.ctypes.data, float(omega), len(na),
                        na.ctypes.data, rm.ctypes.data, ap.ctypes.data, float(epslat),
                        out.ctypes.data)
    return out


def g_lengths(ivg, bvec) -> np.ndarray:
    """gengvec.f90:61-70: |G| with v = i1*b1 + i2*b2 + i3*b3."""
    iv = np.asarray(ivg, dtype=np.float64)
    b = np.asarray(bvec, dtype=np.float64)
    v = iv[0][None, :] * b[:, 0:1] + iv[1][None, :] * b[:, 1:2] + iv[2][None, :] * b[:, 2:3]
    return np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])


def gencfun(s, epslat=1e-6, cfdamp=0.0, gmaxvr=12.0):
    """gencfun.f90:51-84 for a system s (synth.LapwSystem): (cfunig over all
    ngrtot G vectors, cfunir on the FFT grid)."""
    ngrid = ngrid_of(s.intgv)
    cfunig = gencfunig(s.ivg, s.bvec, s.omega, s.natoms, s.rmt, s.atposc, epslat)
    zfft = np.zeros(int(np.prod(ngrid)), dtype=np.complex128)
    zfft[igfft(ngrid, s.ivg)] = cfunig                          # :70
    cfunir = zfftifc(ngrid, 1, zfft).real.copy()                 # :73-74
    if cfdamp > 0.0:                                             # :76-82
        t1 = cfdamp / gmaxvr
        cfunig = cfunig * np.exp(-(t1 * g_lengths(s.ivg, s.bvec)) ** 2)
    return cfunig, cfunir


def _forward_gather(s, zfft):
    ngrid = ngrid_of(s.intgv)
    z = zfftifc(ngrid, -1, zfft)
    return z[igfft(ngrid, s.ivg[:, :s.ngvec])]


def genveffig(s, veffir, cfunir) -> np.ndarray:
    """genveffig.f90:31-39: veffig(1:ngvec) from veffir * cfunir."""
    return _forward_gather(s, np.asarray(veffir) * np.asarray(cfunir) + 0j)


def genmeffig(s, veffir, cfunir, iora_star: bool):
    """genmeffig.f90:33-57 for ValenceRelativity != 'none': meffig from
    cfunir / (1 - veffir*a2) and, for iora*, m2effig from
    cfunir / (1 - veffir*a2)**2, a2 = alpha^2/2."""
    a2 = 0.5 * ALPHA ** 2
    v = np.asarray(veffir, dtype=np.float64)
    c = np.asarray(cfunir, dtype=np.float64)
    meffig = _forward_gather(s, c / (1.0 - v * a2) + 0j)
    m2effig = _forward_gather(s, c / ((1.0 - v * a2) ** 2) + 0j) if iora_star else None
    return meffig, m2effig


# --- muffin-tin Hamiltonian tables (hmlint.f90 -> mt_kin.f90, mt_pot.f90) ---

def spline(x, f):
    """spline.f90: cf(3, n)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    cf = np.zeros((3, x.size), order="F")
    lib().orc_spline(x.size, x.ctypes.data, f.ctypes.data, cf.ctypes.data)
    return cf


def fderiv_int(x, f):
    """fderiv.f90 with m = -1: the running spline integral g(n)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    g = np.zeros(x.size)
    lib().orc_fderiv_int(x.size, x.ctypes.data, f.ctypes.data, g.ctypes.data)
    return g


def gaunt_yry(l1, l2, l3, m1, m2, m3) -> complex:
    out = np.zeros(2)
    lib().orc_gaunt_yry(l1, l2, l3, m1, m2, m3, out.ctypes.data)
    return complex(out[0], out[1])


def hmlint(sys_, lmaxvr=None, veffmt=None, lofr=None, valence_rel_none=None, parts=3):
    """mt_kin + mt_pot (hmlint.f90:53-55) from zero tables: (aa, loa, lolo)
    with the shapes of mt_hscf%main.  parts: 1 mt_kin only, 2 mt_pot only
    (with veffmt as the potential), 3 both."""
    s = sys_.s
    lmaxvr = s.lmaxvr if lmaxvr is None else lmaxvr
    veffmt = s.veffmt if veffmt is None else veffmt
    lofr = s.lofr if lofr is None else lofr
    relnone = s.valence_rel_none if valence_rel_none is None else valence_rel_none
    mnlo = max(s.maxnlo, 1)
    aa = np.zeros((s.maxaa, s.maxaa, s.natmtot), dtype=np.complex128, order="F")
    loa = np.zeros((mnlo, s.maxaa, s.natmtot), dtype=np.complex128, order="F")
    lolo = np.zeros((mnlo, mnlo, s.natmtot), dtype=np.complex128, order="F")
    v = np.asfortranarray(veffmt, dtype=np.float64)
    lf = np.asfortranarray(lofr, dtype=np.float64)
    rc = lib().orc_hmlint_parts(sys_.ptr, lmaxvr, v.ctypes.data, lf.ctypes.data,
                                1 if relnone else 0, aa.ctypes.data, loa.ctypes.data,
                                lolo.ctypes.data, parts)
    if rc != 0:
        raise RuntimeError(f"orc_hmlint failed: {rc}")
    return aa, loa, lolo


def mt_so(sys_, lmaxvr=None, pot=None, lofr=None, rel_level=1, tables=None):
    """mt_so.f90:11-397 (hmlint.f90:172 calls it with veffmt and level_zora
    = 1): the spin-orbit terms accumulated onto tables = {"alpha", "beta",
    "ab"}, each (aa, loa, lolo, alo), zero when not given.  Returns new
    arrays; alpha/beta alo are reset to conj(loa^T) as the reference does."""
    s = sys_.s
    lmaxvr = s.lmaxvr if lmaxvr is None else lmaxvr
    pot = s.veffmt if pot is None else pot
    lofr = s.lofr if lofr is None else lofr
    mnlo = max(s.maxnlo, 1)
    shapes = ((s.maxaa, s.maxaa, s.natmtot), (mnlo, s.maxaa, s.natmtot),
              (mnlo, mnlo, s.natmtot), (s.maxaa, mnlo, s.natmtot))
    out = {}
    for key in ("alpha", "beta", "ab"):
        src = None if tables is None else tables.get(key)
        out[key] = tuple(np.zeros(sh, dtype=np.complex128, order="F") if src is None
                         else np.array(src[i], dtype=np.complex128, order="F", copy=True)
                         for i, sh in enumerate(shapes))
    ptrs = (C.c_void_p * 12)(*[x.ctypes.data for k in ("alpha", "beta", "ab")
                                    for x in out[k]])
    v = np.asfortranarray(pot, dtype=np.float64)
    lf = np.asfortranarray(lofr, dtype=np.float64)
    rc = lib().orc_mt_so(sys_.ptr, lmaxvr, v.ctypes.data, lf.ctypes.data, rel_level, ptrs)
    if rc != 0:
        raise RuntimeError(f"orc_mt_so failed: {rc}")
    return out


def overlapsetup(sys_, ik: int, apwalm: np.ndarray, h: np.ndarray = None, fast: bool = True):
    """overlapsetup.f90:8-166 alone on zero O: the muffin-tin, lo and
    interstitial parts, then hamiltonoverlapocopy_UL (:164-166) on O and on h
    (a zero matrix unless given).  Returns (h, O)."""
    s = sys_.s
    k = s.kpts[ik]
    n = k.ngk + s.nlotot
    o = np.zeros((n, n), dtype=np.complex128, order="F")
    h = np.zeros((n, n), dtype=np.complex128, order="F") if h is None else np.asfortranarray(h)
    zg = scipy_zgemm() if fast else None
    a = np.asfortranarray(apwalm)
    ig = np.ascontiguousarray(k.igpig, dtype=np.int32)
    vg = np.asfortranarray(k.vgpc)
    rc = lib().orc_overlapsetup(sys_.ptr, k.ngk, a.ctypes.data, ig.ctypes.data, vg.ctypes.data,
                                n, o.ctypes.data, zg)
    if rc != 0:
        raise RuntimeError(f"orc_overlapsetup failed: {rc}")
    lib().orc_copy_ul(n, h.ctypes.data, o.ctypes.data)
    return h, o


def calc_vnlmat(O: np.ndarray, evec: np.ndarray, vxnl: np.ndarray) -> np.ndarray:
    """src_hybrids/calc_vnlmat.f90:54-73: with X = evec(1:n, :), temp = X^H O,
    temp1 = vxnl temp, vnlmat = temp^H temp1 (NumPy products)."""
    n = O.shape[0]
    X = evec[:n, :]
    temp = X.conj().T @ O
    temp1 = vxnl @ temp
    return temp.conj().T @ temp1


def kinetic_energy(sys_, ik, evecfv, apwalm, cfunir, lofr, veffmt, valence_rel_none=True,
                   veffir=None):
    """kinetic_energy.f90:1-230 for 'none' and ZORA: engyknst(1:nstfv).
    evecfv (nmat, nstfv), apwalm (ngkmax, apwordmax, lmmaxapw, natmtot),
    cfunir/veffir (ngrtot), lofr (nrmtmax, 2, nlomax, natmtot),
    veffmt (lmmaxvr, nrmtmax, natmtot).  Restated step by step: spline
    integrals by fderiv, the zgemm of :122-135 by NumPy, the FFTs by
    zfftifc."""
    s = sys_.s
    y00 = 0.28209479177387814347
    k = s.kpts[ik]
    ngp = k.ngk
    nst = evecfv.shape[1]
    p, lmaxmat = s.apwordmax, s.lmaxmat
    lmm = (s.lmaxapw + 1) ** 2
    e = np.zeros(nst)
    relnone = valence_rel_none
    loofs = ngp
    ias = 0
    for is_ in range(s.nspecies):
        nr = int(s.nrmt[is_])
        r = s.spr[:nr, is_]
        r2 = r * r
        r2inv = 1.0 / r2
        for _ in range(s.natoms[is_]):
            rm = 1.0 / (1.0 - 0.5 * ALPHA ** 2 * veffmt[0, :nr, ias] * y00)   # :63-65
            u = lambda io, l, c: s.apwfr[:nr, c, io, l, ias]                  # noqa: E731
            t_aa = np.zeros((p, p, lmaxmat + 1))
            for l1 in range(lmaxmat + 1):                                     # :66-121
                ang = float(l1 * (l1 + 1))
                for io1 in range(s.apword[l1, is_]):
                    for io2 in range(s.apword[l1, is_]):
                        t1 = u(io1, l1, 0) * u(io2, l1, 0)
                        t2 = u(io1, l1, 1) * u(io2, l1, 1)
                        if relnone:
                            fr = r2 * (0.5 * t2 + 0.5 * ang * t1 * r2inv)
                        else:
                            fr = (0.5 * t2 * rm + 0.5 * ang * t1 * rm * r2inv) * r2
                        t_aa[io2, io1, l1] = fderiv_int(r, fr)[-1]
            A = apwalm[:ngp, :, :, ias].reshape(ngp, p * lmm, order="F")
            zm = A.T @ evecfv[:ngp, :]                                        # :122-135
            nlo = s.nlorb[is_]
            t_alo = np.zeros((p, max(s.nlomax, 1)))
            t_lolo = np.zeros((max(s.nlomax, 1), max(s.nlomax, 1)))
            for ilo in range(nlo):                                            # :136-158
                l1 = s.lorbl[ilo, is_]
                ang = float(l1 * (l1 + 1))
                for io in range(s.apword[l1, is_] if l1 <= lmaxmat else 0):
                    t1 = u(io, l1, 0) * lofr[:nr, 0, ilo, ias]
                    t2 = u(io, l1, 1) * lofr[:nr, 1, ilo, ias]
                    fr = (0.5 * t2 * rm + 0.5 * ang * t1 * rm * r2inv) * r2
                    t_alo[io, ilo] = fderiv_int(r, fr)[-1]
            for ilo1 in range(nlo):                                           # :159-183
                for ilo2 in range(nlo):
                    l1, l3 = s.lorbl[ilo1, is_], s.lorbl[ilo2, is_]
                    if l1 != l3:
                        continue
                    ang = float(l1 * (l1 + 1))
                    t1 = lofr[:nr, 0, ilo1, ias] * lofr[:nr, 0, ilo2, ias]
                    t2 = lofr[:nr, 1, ilo1, ias] * lofr[:nr, 1, ilo2, ias]
                    fr = (0.5 * t2 * rm + 0.5 * ang * t1 * rm * r2inv) * r2
                    t_lolo[ilo1, ilo2] = fderiv_int(r, fr)[-1]
            los = int(s.losize[is_])
            for ist in range(nst):                                            # :184-199
                zvec = np.zeros(p * lmm, dtype=complex)
                zveclo = np.zeros(max(los, 1), dtype=complex)
                for l3 in range(lmaxmat + 1):
                    for m3 in range(-l3, l3 + 1):
                        lm3 = l3 * (l3 + 1) + m3
                        for io2 in range(s.apword[l3, is_]):
                            for io1 in range(s.apword[l3, is_]):
                                zvec[p * lm3 + io2] += t_aa[io1, io2, l3] * zm[p * lm3 + io1, ist]
                if3 = 0
                for ilo in range(nlo):
                    l = s.lorbl[ilo, is_]
                    for m in range(-l, l + 1):
                        lm = l * (l + 1) + m
                        for io1 in range(s.apword[l, is_]):
                            zvec[p * lm + io1] += t_alo[io1, ilo] * evecfv[loofs + if3 + m + l, ist]
                            zveclo[if3 + m + l] += t_alo[io1, ilo] * zm[p * lm + io1, ist]
                    if1 = 0
                    for ilo2 in range(nlo):
                        if s.lorbl[ilo2, is_] == l:
                            for m in range(-l, l + 1):
                                zveclo[if3 + m + l] += (t_lolo[ilo, ilo2]
                                                        * evecfv[loofs + if1 + m + l, ist])
                        if1 += 2 * s.lorbl[ilo2, is_] + 1
                    if3 += 2 * l + 1
                e[ist] += np.vdot(zm[:, ist], zvec).real
                if nlo:
                    e[ist] += np.vdot(evecfv[loofs:loofs + los, ist], zveclo[:los]).real
            loofs += los
            ias += 1
    # interstitial part (:205-234)
    ngrid = ngrid_of(s.intgv)
    idx = igfft(ngrid, s.ivg[:, k.igpig[:ngp] - 1])
    vg = np.asarray(k.vgpc)[:, :ngp]
    cf = np.asarray(cfunir, dtype=float)
    if not relnone:
        cf = cf / (1.0 - 0.5 * ALPHA * ALPHA * np.asarray(veffir))
    ntot = int(np.prod(ngrid))
    for ist in range(nst):
        for ix in range(3):
            z = np.zeros(ntot, dtype=complex)
            z[idx] = evecfv[:ngp, ist] * vg[ix]
            z = zfftifc(ngrid, 1, z) * cf
            z = zfftifc(ngrid, -1, z)
            zvec = z[idx] * vg[ix]
            e[ist] += 0.5 * np.vdot(evecfv[:ngp, ist], zvec).real
    return e


def genffacg(s, is_, cfdamp=0.0, gmaxvr=12.0, epslat=1e-6):
    """genffacg.f90:13-40: step-function form factors of species is_ for the
    first ngvec G vectors."""
    gc = g_lengths(s.ivg[:, :s.ngvec], s.bvec)
    t1 = 4 * np.pi / s.omega
    t2 = cfdamp / gmaxvr
    out = np.empty(s.ngvec)
    fin = gc > epslat
    t3 = np.exp(-(t2 * gc[fin]) ** 2) if cfdamp != 0.0 else 1.0
    t4 = gc[fin] * s.rmt[is_]
    out[fin] = t1 * t3 * (np.sin(t4) - t4 * np.cos(t4)) / gc[fin] ** 3
    out[~fin] = (t1 / 3.0) * s.rmt[is_] ** 3
    return out


def sfacg_of(s):
    """sfacg(ngvec, natmtot) = exp(i G.tau) (gensfacgp.f90 for the G set)."""
    v = np.asarray(s.bvec) @ np.asarray(s.ivg[:, :s.ngvec], dtype=float)
    return np.exp(1j * (v.T @ np.asarray(s.atposc).T))


def forcek(sys_, ik, evecfv, evalfv, occ, wkpt, ffacg, sfacg, evecsv=None, nspinor=2,
           spiral_ispn=None):
    """forcek.f90:1-404 for one k-point and one first-variational spin: the
    contribution to forceibs(3, natmtot).  Per atom H and O of the
    muffin-tin and lo border (:60-230), dlh/dlo with the folded G
    differences (:91-98, :259-285) and the three zgemm (:286-335), restated
    with NumPy.  ffacg(ngvec, nspecies), sfacg(ngvec, natmtot),
    evecfv(nmat, nstfv), occ = occsv(nstsv).  The state sum (:340-372):
    without evecsv (tevecsv false) wkpt sum_j occ_j Re ffv(j, j); with
    evecsv(nstsv, nstsv) the second-variational sum over nspinor blocks of
    nstfv rows, or over the block of first-variational spin spiral_ispn
    (0-based) for a spin spiral."""
    s = sys_.s
    k = s.kpts[ik]
    ngk = k.ngk
    n = ngk + s.nlotot
    p = s.apwordmax
    apwalm = match(sys_, ik)
    ngrid = np.asarray(ngrid_of(s.intgv))
    lo = np.asarray(s.intgv)[:, 0]
    ivk = np.asarray(s.ivg)[:, k.igpig[:ngk] - 1]                        # (3, ngk)
    iv = ivk[:, :, None] - ivk[:, None, :]
    iv = np.mod(iv - lo[:, None, None], ngrid[:, None, None]) + lo[:, None, None]
    ijgij = s.ivgig[iv[0] - lo[0], iv[1] - lo[1], iv[2] - lo[2]]       # 1-based
    assert ijgij.max() <= s.ngvec, "G differences beyond ngvec"
    vgk = np.asarray(k.vgpc)[:, :ngk]
    dpij = 0.5 * (vgk.T @ vgk)
    vgc = np.asarray(s.vgc)
    X = evecfv[:n]
    X2 = X * evalfv[None, :]
    out = np.zeros((3, s.natmtot))
    ias = 0
    lob = 0
    for is_ in range(s.nspecies):
        for _ in range(s.natoms[is_]):
            rows = [(l, m, io) for l in range(s.lmaxmat + 1) for m in range(-l, l + 1)
                    for io in range(s.apword[l, is_])]
            apwi = np.array([apwalm[:ngk, io, l * (l + 1) + m, ias] for (l, m, io) in rows])
            maxaa = s.maxaa
            apwi_full = np.zeros((maxaa, ngk), dtype=complex)
            apwi_full[:len(rows)] = apwi
            H = np.zeros((n, n), dtype=complex)
            O = np.zeros((n, n), dtype=complex)
            zm = s.aa[:, :, ias] @ apwi_full
            H[:ngk, :ngk] += apwi_full.conj().T @ zm
            los = int(s.losize[is_])
            if s.nlorb[is_]:
                r0 = ngk + lob
                H[r0:r0 + los, :ngk] += s.loa[:los, :, ias] @ apwi_full
                H[:ngk, r0:r0 + los] = H[r0:r0 + los, :ngk].conj().T
            O[:ngk, :ngk] += apwi_full.conj().T @ apwi_full
            if s.nlorb[is_]:
                c = ngk + lob
                for ilo in range(s.nlorb[is_]):
                    l = s.lorbl[ilo, is_]
                    lm1, lm2 = l * (l + 1) - l, l * (l + 1) + l
                    for io in range(s.apword[l, is_]):
                        sv = s.oalo[io, ilo, ias] + (s.h1loa[io, ilo, ias]
                                                     if s.h1loa is not None else 0.0)
                        O[:ngk, c:c + 2 * l + 1] += np.conj(apwalm[:ngk, io, lm1:lm2 + 1, ias] * sv)
                    O[c:c + 2 * l + 1, :ngk] = O[:ngk, c:c + 2 * l + 1].conj().T
                    c += 2 * l + 1
            for l in range(3):
                ig = ijgij - 1
                t1 = vgc[l, ig]
                zt1 = -ffacg[ig, is_] * np.conj(sfacg[ig, ias])
                dlh = np.zeros((n, n), dtype=complex)
                dlo = np.zeros((n, n), dtype=complex)
                dlh[:ngk, :ngk] = (dpij * zt1 + H[:ngk, :ngk]) * t1 * 1j
                dlo[:ngk, :ngk] = -(zt1 + O[:ngk, :ngk]) * t1 * 1j
                tk = vgk[l][:, None]
                dlh[:ngk, ngk:] = H[:ngk, ngk:] * tk * 1j
                dlh[ngk:, :ngk] = dlh[:ngk, ngk:].conj().T
                dlo[:ngk, ngk:] = -O[:ngk, ngk:] * tk * 1j
                dlo[ngk:, :ngk] = dlo[:ngk, ngk:].conj().T
                zmf = dlh @ X + dlo @ X2
                ffv = X.conj().T @ zmf
                nstfv = ffv.shape[0]
                if evecsv is None:
                    tot = np.sum(occ * np.real(np.diag(ffv)))
                else:
                    tot = 0.0
                    blocks = [spiral_ispn] if spiral_ispn is not None else range(nspinor)
                    for j in range(evecsv.shape[1]):
                        for b in blocks:
                            v = evecsv[b * nstfv:(b + 1) * nstfv, j]
                            tot += occ[j] * np.real(v.conj() @ (ffv @ v))
                out[l, ias] += wkpt * tot
            lob += los
            ias += 1
    return out
