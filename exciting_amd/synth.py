"""Seeded synthetic inputs for the LAPW H/O setup, in the reference's layouts.

These are the read-only inputs the exciting hot path receives from its
producers.  Each array follows the producer it stands in for:

* G-vector grid and index maps: ``gridsize.f90`` (ngrid from gmaxvr, made
  FFT-friendly by ``nfftifc.f90``), ``gengvec.f90`` (box enumeration, sort by
  length, ``ivg``/``ivgig``/``ngvec``).  Ties in |G| keep box order (the
  reference's heapsort orders ties differently; the setup takes the order as
  input, so both paths see identical arrays).
* G+k sets: ``gengpvec.f90:55-91`` (``|G+k| < gkmax`` over ``ig <= ngvec``,
  ``gkmax = rgkmax / R_MT`` from ``getgkmax.f90``), spherical coordinates from
  ``sphcrd.f90``, structure factors from ``gensfacgp.f90:44-49``.
* LO index layout: ``genidxlo.f90`` (species-major atoms, then ilo, then m).
* Radial boundary data: ``apwfr`` tails are cubics on the last grid points, so
  ``polynom`` (used by ``match.f90:141-148``) recovers boundary values and
  derivatives drawn from U(-1,1).
* Muffin-tin tables ``mt_hscf%main%{aa,loa,lolo}`` (``mod_eigensystem.F90:60-80``)
  and overlap integrals ``oalo``/``ololo`` (``olprad.f90:92-141``): random,
  Hermitian where the reference's are, with diagonally dominant spectra and
  ``ololo - oalo^T oalo`` positive so the overlap matrix is positive definite.
* ``cfunig`` is the Fourier series of a positive periodic function
  (constant plus Yukawa bumps at the atoms), so its Toeplitz form is positive
  semi-definite, with |c(G)| ~ 1/(1+G^2); ``veffig`` is Hermitian-symmetric,
  v(-G) = v(G)^*, with the same decay.

All index arrays keep 1-based Fortran VALUES (igpig, ivg box, ivgig).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

FOURPI = 12.566370614359172954


# ----------------------------------------------------------------------------
# cell / basis description
# ----------------------------------------------------------------------------
@dataclasses.dataclass
class CellSpec:
    name: str
    avec: np.ndarray                 # (3,3) columns are lattice vectors (bohr)
    positions: List[np.ndarray]      # per species: (natoms, 3) fractional coords
    rmt: List[float]                 # per species
    lmaxapw: int
    lmaxmat: int
    apword: int                      # APW order, uniform over l and species
    lo_l: List[List[int]]            # per species: l of each local orbital
    rgkmax: float
    gmaxvr: float
    kpoints: np.ndarray              # (nk, 3) lattice coords
    nrmt: int = 250
    seed: int = 0


def _fcc(a):
    return 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]]).T


def _shifted_grid(n1, n2, n3):
    g = [((i + 0.5) / n1, (j + 0.5) / n2, (k + 0.5) / n3)
         for i in range(n1) for j in range(n2) for k in range(n3)]
    return np.array(g, dtype=np.float64)


_ROCKSALT_A = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]], float)
_ROCKSALT_B = np.array([[.5, 0, 0], [0, .5, 0], [0, 0, .5], [.5, .5, .5]], float)


def config(index: int) -> CellSpec:
    """The setup workloads: 0 fcc 2-atom, 1 SC 8-atom, 2 orthorhombic 16-atom,
    3 SC 64-atom supercell, 4 SC 32-atom supercell (serving mix only)."""
    if index == 0:
        return CellSpec(
            name="fcc2", avec=_fcc(10.26),
            positions=[np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]])],
            rmt=[2.0], lmaxapw=8, lmaxmat=8, apword=2, lo_l=[[0, 1, 2]],
            rgkmax=7.0, gmaxvr=12.0, kpoints=np.array([[0.1, 0.2, 0.3]]),
            seed=1000)
    if index == 1:
        return CellSpec(
            name="sc8", avec=10.26 * np.eye(3),
            positions=[_ROCKSALT_A.copy(), _ROCKSALT_B.copy()],
            rmt=[2.0, 2.0], lmaxapw=10, lmaxmat=10, apword=2,
            lo_l=[[0, 1, 2, 3], [0, 1, 2, 3]], rgkmax=8.0, gmaxvr=12.0,
            kpoints=_shifted_grid(2, 2, 2), seed=1001)
    if index == 2:
        rng = np.random.default_rng(1002)
        avec = np.diag([10.26, 10.26, 20.52])
        counts = [6, 5, 5]
        pos = _random_positions(rng, avec, sum(counts), 4.2)
        sp, o = [], 0
        for c in counts:
            sp.append(pos[o:o + c])
            o += c
        return CellSpec(
            name="ortho16", avec=avec, positions=sp, rmt=[2.0, 2.0, 2.0],
            lmaxapw=10, lmaxmat=10, apword=2, lo_l=[[0, 1, 2, 3]] * 3,
            rgkmax=9.0, gmaxvr=12.0, kpoints=_shifted_grid(8, 4, 4), seed=1002)
    if index == 3:
        rng = np.random.default_rng(1003)
        a = 20.52
        sp = []
        for base in (_ROCKSALT_A, _ROCKSALT_B):
            pts = []
            for s in np.ndindex(2, 2, 2):
                for p in base:
                    pts.append((p + np.array(s, float)) / 2.0)
            pts = np.array(pts)
            pts = pts + rng.normal(0.0, 0.05, pts.shape) / a  # N(0, 0.05 bohr)
            sp.append(pts)
        return CellSpec(
            name="sc64", avec=a * np.eye(3), positions=sp, rmt=[2.0, 2.0],
            lmaxapw=10, lmaxmat=10, apword=2, lo_l=[[0, 1, 2, 3]] * 2,
            rgkmax=8.0, gmaxvr=12.0, kpoints=_shifted_grid(4, 2, 2), seed=1003)
    if index == 4:
        # 32-atom 2x2x1 supercell of the 8-atom rock-salt cell, jittered like
        # the 64-atom one; 16 k-points on a shifted 2x2x4 grid
        rng = np.random.default_rng(1004)
        avec = np.diag([20.52, 20.52, 10.26])
        sp = []
        for base in (_ROCKSALT_A, _ROCKSALT_B):
            pts = []
            for s in np.ndindex(2, 2, 1):
                for p in base:
                    pts.append((p + np.array(s, float)) / np.array([2.0, 2.0, 1.0]))
            pts = np.array(pts)
            pts = pts + rng.normal(0.0, 0.05, pts.shape) / np.diag(avec)
            sp.append(pts)
        return CellSpec(
            name="sc32", avec=avec, positions=sp, rmt=[2.0, 2.0],
            lmaxapw=10, lmaxmat=10, apword=2, lo_l=[[0, 1, 2, 3]] * 2,
            rgkmax=8.0, gmaxvr=12.0, kpoints=_shifted_grid(2, 2, 4), seed=1004)
    raise ValueError(f"no workload {index}")


def serving_cells(kmax: int = 16):
    """The request mix's cells, 2, 8, 16, 32 and 64 atoms, built as the setup
    workloads with at most ``kmax`` k-points each: {name: CellSpec, kpoint
    indices}."""
    out = {}
    for idx in (0, 1, 2, 4, 3):
        spec = config(idx)
        out[spec.name] = (spec, list(range(min(kmax, len(spec.kpoints)))))
    return out


def tiny(seed: int = 7, lmax: int = 3, nlo_l: Sequence[int] = (0, 1),
         rgkmax: float = 4.0, natoms: Sequence[int] = (1, 1), apword: int = 2,
         lmaxmat: Optional[int] = None, nk: int = 1) -> CellSpec:
    """A small cell for brute-force checks (few dozen G+k vectors); with
    nk > 1 the extra k-points are seeded random points of the zone, so the
    G+k sets have uneven sizes."""
    rng = np.random.default_rng(seed)
    avec = np.array([[7.0, 0.3, 0.1], [0.2, 7.5, 0.4], [0.0, 0.5, 8.0]]).T
    pos = _random_positions(rng, avec, sum(natoms), 3.8)
    sp, o = [], 0
    for c in natoms:
        sp.append(pos[o:o + c])
        o += c
    return CellSpec(
        name=f"tiny{seed}", avec=avec, positions=sp, rmt=[1.8] * len(natoms),
        lmaxapw=lmax, lmaxmat=lmax if lmaxmat is None else lmaxmat,
        apword=apword, lo_l=[list(nlo_l)] * len(natoms), rgkmax=rgkmax,
        gmaxvr=2.0 * rgkmax / 1.8 + 1.0,
        kpoints=np.concatenate([np.array([[0.13, -0.21, 0.37]]),
                                rng.uniform(-0.5, 0.5, (nk - 1, 3))]),
        nrmt=120, seed=seed)


def _random_positions(rng, avec, n, dmin):
    pts = []
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > 200000:
            raise RuntimeError("could not place atoms")
        p = rng.random(3)
        ok = True
        for q in pts:
            d = p - q
            d -= np.round(d)
            if np.linalg.norm(avec @ d) < dmin:
                ok = False
                break
        if ok:
            pts.append(p)
    return np.array(pts)


# ----------------------------------------------------------------------------
# reference producers
# ----------------------------------------------------------------------------
def nfftifc(n: int) -> int:
    """nfftifc.f90: next n whose only prime factors are 2, 3, 5."""
    while True:
        i = n
        for p in (2, 3, 5):
            while i % p == 0:
                i //= p
        if i == 1:
            return n
        n += 1


def sphcrd(v: np.ndarray):
    """sphcrd.f90:13-50, vectorised over rows of v (n,3)."""
    eps = 1.0e-14
    r = np.sqrt(v[:, 0] ** 2 + v[:, 1] ** 2 + v[:, 2] ** 2)
    tp = np.zeros((v.shape[0], 2))
    big = r > eps
    t1 = np.where(big, v[:, 2] / np.where(big, r, 1.0), 0.0)
    t1 = np.clip(t1, -1.0, 1.0)
    tp[:, 0] = np.where(big, np.arccos(t1), 0.0)
    has = big & ((np.abs(v[:, 0]) > eps) | (np.abs(v[:, 1]) > eps))
    ph = np.arctan2(v[:, 1], v[:, 0])
    ph = np.where(ph < 0.0, ph + 2.0 * math.pi, ph)
    tp[:, 1] = np.where(has, ph, 0.0)
    return r, tp


@dataclasses.dataclass
class KPoint:
    vkl: np.ndarray
    ngk: int
    igpig: np.ndarray     # (ngkmax,) int32, 1-based
    vgpc: np.ndarray      # (3, ngkmax) F-order
    gpc: np.ndarray       # (ngkmax,)
    tpgpc: np.ndarray     # (2, ngkmax) F-order
    sfacgp: np.ndarray    # (ngkmax, natmtot) complex F-order


@dataclasses.dataclass
class LapwSystem:
    """Everything the drop-in reads: module state plus per-k inputs."""
    spec: CellSpec
    nspecies: int
    natoms: np.ndarray          # int32 (nspecies,)
    natmtot: int
    atposc: np.ndarray          # (natmtot, 3) cartesian
    omega: float
    bvec: np.ndarray
    lmaxapw: int
    lmaxmat: int
    apwordmax: int
    apword: np.ndarray          # int32 (lmaxapw+1, nspecies) F-order
    nlomax: int
    nlorb: np.ndarray           # int32 (nspecies,)
    lorbl: np.ndarray           # int32 (nlomax, nspecies) F-order
    maxaa: int
    maxnlo: int
    losize: np.ndarray          # (nspecies,)
    nlotot: int
    rmt: np.ndarray
    nrmt: np.ndarray            # int32
    nrmtmax: int
    spr: np.ndarray             # (nrmtmax, nspecies) F
    apwfr: np.ndarray           # (nrmtmax, 2, apwordmax, lmaxapw+1, natmtot) F
    aa: np.ndarray              # (maxaa, maxaa, natmtot) complex F
    loa: np.ndarray             # (maxnlo, maxaa, natmtot) complex F
    lolo: np.ndarray            # (maxnlo, maxnlo, natmtot) complex F
    oalo: np.ndarray            # (apwordmax, nlomax, natmtot) F
    ololo: np.ndarray           # (nlomax, nlomax, natmtot) F
    ngrid: np.ndarray
    ngrtot: int
    ngvec: int
    ivg: np.ndarray             # (3, ngrtot) int32 F
    intgv: np.ndarray           # (3, 2) int32 F
    ivgig: np.ndarray           # box (n1, n2, n3) int32 F, 1-based values
    vgc: np.ndarray             # (3, ngrtot) F
    gc: np.ndarray
    cfunig: np.ndarray          # (ngrtot,) complex
    veffig: np.ndarray          # (ngvec,) complex
    gkmax: float
    ngkmax: int
    kpts: List[KPoint]
    h1aa: Optional[np.ndarray] = None
    h1loa: Optional[np.ndarray] = None
    h1lolo: Optional[np.ndarray] = None
    meffig: Optional[np.ndarray] = None
    m2effig: Optional[np.ndarray] = None
    valence_rel_none: bool = True
    h1on: bool = False
    # muffin-tin potential and LO radial functions (inputs of hmlint)
    lmaxvr: int = 0
    veffmt: Optional[np.ndarray] = None   # (lmmaxvr, nrmtmax, natmtot) F
    lofr: Optional[np.ndarray] = None     # (nrmtmax, 2, nlomax, natmtot) F

    def nmat(self, ik: int) -> int:
        return self.kpts[ik].ngk + self.nlotot


def _lattice(spec: CellSpec):
    avec = np.asarray(spec.avec, dtype=np.float64)
    bvec = 2.0 * math.pi * np.linalg.inv(avec).T
    omega = abs(np.linalg.det(avec))
    return avec, bvec, omega


def gvectors(avec, bvec, gmaxvr):
    """gridsize.f90 + gengvec.f90 (sort by length; ties in box order)."""
    ngrid = np.array([nfftifc(int(gmaxvr * np.linalg.norm(avec[:, i]) / math.pi) + 1)
                      for i in range(3)], dtype=np.int64)
    lo = ngrid // 2 - ngrid + 1
    hi = ngrid // 2
    i1, i2, i3 = np.meshgrid(np.arange(lo[0], hi[0] + 1), np.arange(lo[1], hi[1] + 1),
                             np.arange(lo[2], hi[2] + 1), indexing="ij")
    # enumeration order of gengvec: i1 outer, i3 inner
    iv = np.stack([i1.ravel(), i2.ravel(), i3.ravel()], axis=0)
    v = bvec @ iv.astype(np.float64)
    gc = np.sqrt((v ** 2).sum(axis=0))
    order = np.argsort(gc, kind="stable")
    iv = iv[:, order]
    gc = gc[order]
    v = v[:, order]
    ngrtot = iv.shape[1]
    shape = tuple(int(x) for x in (hi - lo + 1))
    ivgig = np.zeros(shape, dtype=np.int32, order="F")
    ivgig[iv[0] - lo[0], iv[1] - lo[1], iv[2] - lo[2]] = np.arange(1, ngrtot + 1, dtype=np.int32)
    below = np.nonzero(gc < gmaxvr)[0]
    ngvec = int(below[-1] + 1) if below.size else 1
    intgv = np.asfortranarray(np.stack([lo, hi], axis=1).astype(np.int32))
    return (ngrid, ngrtot, ngvec, np.asfortranarray(iv.astype(np.int32)), intgv, ivgig,
            np.asfortranarray(v), gc)


def gkvectors(vkl, bvec, vgc, ngvec, gkmax):
    """gengpvec.f90:55-91 for one k-point: (igpig 1-based, vgpc (ngp,3))."""
    vkc = bvec @ np.asarray(vkl, float)
    v = vgc[:, :ngvec].T + vkc[None, :]
    t2 = v[:, 0] ** 2 + v[:, 1] ** 2 + v[:, 2] ** 2
    sel = np.nonzero(t2 < gkmax ** 2)[0]
    return sel.astype(np.int32) + 1, v[sel]


def _herm_dd(rng, n, off, dlo, dhi):
    """Random Hermitian, diagonally dominant with positive diagonal."""
    x = (rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))) * off
    h = 0.5 * (x + x.conj().T)
    np.fill_diagonal(h, 0.0)
    d = rng.uniform(dlo, dhi, n) + np.abs(h).sum(axis=1)
    return h + np.diag(d)


def build(spec: CellSpec, kpoints: Optional[Sequence[int]] = None) -> LapwSystem:
    """Construct all inputs for a cell; ``kpoints`` selects a subset of its
    k-points (default: all)."""
    rng = np.random.default_rng(spec.seed)
    avec, bvec, omega = _lattice(spec)
    nspecies = len(spec.positions)
    natoms = np.array([len(p) for p in spec.positions], dtype=np.int32)
    natmtot = int(natoms.sum())
    atposc = np.concatenate([(avec @ np.asarray(p, float).T).T for p in spec.positions], axis=0)
    L = spec.lmaxapw
    p = spec.apword
    apword = np.full((L + 1, nspecies), p, dtype=np.int32, order="F")
    nlorb = np.array([len(x) for x in spec.lo_l], dtype=np.int32)
    nlomax = max(1, int(nlorb.max()))
    lorbl = np.zeros((nlomax, nspecies), dtype=np.int32, order="F")
    for s, ls in enumerate(spec.lo_l):
        lorbl[:len(ls), s] = ls
    losize = np.array([sum(2 * l + 1 for l in ls) for ls in spec.lo_l], dtype=np.int32)
    maxnlo = int(losize.max()) if losize.size else 0
    maxaa = max(sum((2 * l + 1) * p for l in range(spec.lmaxmat + 1)) for _ in range(nspecies))
    nlotot = int(sum(natoms[s] * losize[s] for s in range(nspecies)))

    # radial mesh and apwfr tails (cubic on the outer points)
    nrmt = np.full(nspecies, spec.nrmt, dtype=np.int32)
    nrmtmax = int(nrmt.max())
    spr = np.zeros((nrmtmax, nspecies), order="F")
    for s in range(nspecies):
        rmin = 1.0e-5
        t = np.arange(spec.nrmt) / (spec.nrmt - 1)
        r = rmin * np.exp(t * math.log(spec.rmt[s] / rmin))
        r[-1] = spec.rmt[s]
        spr[:spec.nrmt, s] = r
    apwfr = np.zeros((nrmtmax, 2, p, L + 1, natmtot), order="F")
    ias = 0
    for s in range(nspecies):
        R = spec.rmt[s]
        dr = spr[:, s] - R
        for _ in range(natoms[s]):
            for l in range(L + 1):
                while True:
                    d = rng.uniform(-1.0, 1.0, (p, p))  # d[io1, io2] = u_io2^(io1)(R)
                    if p == 1 or abs(np.linalg.det(d[:p, :p])) > 0.15:
                        break
                for io in range(p):
                    c = np.zeros(4)
                    c[:min(p, 4)] = d[:min(p, 4), io] / np.array([1.0, 1.0, 2.0, 6.0])[:min(p, 4)]
                    if p < 4:
                        c[p:] = rng.uniform(-0.5, 0.5, 4 - p)
                    apwfr[:, 0, io, l, ias] = c[0] + dr * (c[1] + dr * (c[2] + dr * c[3]))
                    apwfr[:, 1, io, l, ias] = c[1] + dr * (2 * c[2] + dr * 3 * c[3])
            ias += 1

    # muffin-tin tables
    aa = np.zeros((maxaa, maxaa, natmtot), dtype=np.complex128, order="F")
    loa = np.zeros((max(maxnlo, 1), maxaa, natmtot), dtype=np.complex128, order="F")
    lolo = np.zeros((max(maxnlo, 1), max(maxnlo, 1), natmtot), dtype=np.complex128, order="F")
    oalo = np.zeros((p, nlomax, natmtot), order="F")
    ololo = np.zeros((nlomax, nlomax, natmtot), order="F")
    ias = 0
    for s in range(nspecies):
        for _ in range(natoms[s]):
            aa[:, :, ias] = _herm_dd(rng, maxaa, 0.02, 0.5, 2.0)
            ns = losize[s]
            if ns:
                loa[:ns, :, ias] = 0.05 * (rng.uniform(-1, 1, (ns, maxaa))
                                           + 1j * rng.uniform(-1, 1, (ns, maxaa)))
                lolo[:ns, :ns, ias] = _herm_dd(rng, ns, 0.05, 0.5, 1.5)
                nl = nlorb[s]
                oalo[:, :nl, ias] = rng.uniform(-0.5, 0.5, (p, nl))
                w = oalo[:, :nl, ias]
                x = rng.uniform(-0.1, 0.1, (nl, nl))
                ololo[:nl, :nl, ias] = w.T @ w + 0.5 * (x + x.T) * 0.2 + np.diag(rng.uniform(0.8, 1.2, nl))
            ias += 1

    # G vectors and interstitial functions
    ngrid, ngrtot, ngvec, ivg, intgv, ivgig, vgc, gc = gvectors(avec, bvec, spec.gmaxvr)
    phase = np.exp(-1j * (vgc.T @ atposc.T))          # (ngrtot, natmtot)
    cf = (0.3 / natmtot) * phase.sum(axis=1) / (1.0 + gc ** 2)
    cf[gc < 1e-12] += 0.6
    cfunig = np.ascontiguousarray(cf)
    lo, hi = intgv[:, 0], intgv[:, 1]
    m = -ivg
    has = np.all((m >= lo[:, None]) & (m <= hi[:, None]), axis=0)
    neg = np.zeros(ngrtot, dtype=np.int64)
    neg[has] = ivgig[m[0, has] - lo[0], m[1, has] - lo[1], m[2, has] - lo[2]] - 1
    raw = rng.normal(size=ngrtot) + 1j * rng.normal(size=ngrtot)
    v = raw.copy()
    # v(-G) = v(G)^*; vectors whose negative falls outside the box keep raw
    v[has] = 0.5 * (raw[has] + np.conj(raw[neg[has]]))
    v = -0.3 * v / (1.0 + gc ** 2)
    veffig = np.ascontiguousarray(v[:ngvec])

    # k-points
    gkmax = spec.rgkmax / 2.0  # getgkmax.f90 with the default isgkmax = -1
    ks = range(len(spec.kpoints)) if kpoints is None else kpoints
    sets = []
    for ik in ks:
        igpig, vg = gkvectors(spec.kpoints[ik], bvec, vgc, ngvec, gkmax)
        sets.append((spec.kpoints[ik], igpig, vg))
    ngkmax = max(len(s[1]) for s in sets)
    kpts = []
    for vkl, igpig, vg in sets:
        ngk = len(igpig)
        ig = np.zeros(ngkmax, dtype=np.int32)
        ig[:ngk] = igpig
        vgpc = np.zeros((3, ngkmax), order="F")
        vgpc[:, :ngk] = vg.T
        r, tp = sphcrd(vg)
        gpc = np.zeros(ngkmax)
        gpc[:ngk] = r
        tpgpc = np.zeros((2, ngkmax), order="F")
        tpgpc[:, :ngk] = tp.T
        sf = np.zeros((ngkmax, natmtot), dtype=np.complex128, order="F")
        t1 = vg @ atposc.T
        sf[:ngk, :] = np.cos(t1) + 1j * np.sin(t1)
        kpts.append(KPoint(np.asarray(vkl, float), ngk, ig, vgpc, gpc, tpgpc, sf))

    return LapwSystem(
        spec=spec, nspecies=nspecies, natoms=natoms, natmtot=natmtot, atposc=atposc,
        omega=omega, bvec=bvec, lmaxapw=L, lmaxmat=spec.lmaxmat, apwordmax=p,
        apword=apword, nlomax=nlomax, nlorb=nlorb, lorbl=lorbl, maxaa=maxaa,
        maxnlo=maxnlo, losize=losize, nlotot=nlotot, rmt=np.array(spec.rmt, float),
        nrmt=nrmt, nrmtmax=nrmtmax, spr=spr, apwfr=apwfr, aa=aa, loa=loa, lolo=lolo,
        oalo=oalo, ololo=ololo, ngrid=ngrid, ngrtot=ngrtot, ngvec=ngvec, ivg=ivg,
        intgv=intgv, ivgig=ivgig, vgc=vgc, gc=gc, cfunig=cfunig, veffig=veffig,
        gkmax=gkmax, ngkmax=ngkmax, kpts=kpts)


def add_relativity(sys: LapwSystem, seed: int = 5, iora: bool = True) -> LapwSystem:
    """Attach the non-'none' ValenceRelativity inputs (meffig, and for 'iora*'
    m2effig and the h1aa/h1loa/h1lolo corrections) with random values."""
    rng = np.random.default_rng(seed)
    n = sys.ngvec
    sys.meffig = sys.cfunig[:n] * (1.0 + 0.1 * rng.uniform(-1, 1, n))
    sys.valence_rel_none = False
    if iora:
        p, L = sys.apwordmax, sys.lmaxapw
        sys.m2effig = 0.05 * (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n))
        sys.h1aa = np.asfortranarray(0.01 * rng.uniform(-1, 1, (p, p, L + 1, sys.natmtot)))
        sys.h1loa = np.asfortranarray(0.01 * rng.uniform(-1, 1, (p, sys.nlomax, sys.natmtot)))
        sys.h1lolo = np.asfortranarray(0.01 * rng.uniform(-1, 1, (sys.nlomax, sys.nlomax, sys.natmtot)))
        sys.h1on = True
    return sys


def add_mt_potential(sys: LapwSystem, lmaxvr: int = 8, seed: int = 21) -> LapwSystem:
    """Attach the inputs of hmlint: a smooth real muffin-tin potential
    veffmt(lm, r) = c0 + c1 (r/R) + c2 (r/R)^2 per (lm, atom) with a deep
    spherical part, and LO radial functions lofr with their derivatives
    (cubics in r - R like the APW functions)."""
    rng = np.random.default_rng(seed)
    lmm = (lmaxvr + 1) ** 2
    nr, na = sys.nrmtmax, sys.natmtot
    veffmt = np.zeros((lmm, nr, na), order="F")
    lofr = np.zeros((nr, 2, max(sys.nlomax, 1), na), order="F")
    ias = 0
    for is_ in range(sys.nspecies):
        R = sys.rmt[is_]
        n = sys.nrmt[is_]
        x = sys.spr[:n, is_] / R
        dr = sys.spr[:n, is_] - R
        for _ in range(sys.natoms[is_]):
            c = rng.uniform(-1, 1, (lmm, 3)) * 0.3
            c[0] = [-6.0, 2.5, 1.0]
            c /= (1.0 + np.sqrt(np.arange(lmm)))[:, None]
            veffmt[:, :n, ias] = c[:, 0:1] + x[None, :] * (c[:, 1:2] + x[None, :] * c[:, 2:3])
            for ilo in range(sys.nlorb[is_]):
                k = rng.uniform(-0.5, 0.5, 4)
                lofr[:n, 0, ilo, ias] = k[0] + dr * (k[1] + dr * (k[2] + dr * k[3]))
                lofr[:n, 1, ilo, ias] = k[1] + dr * (2 * k[2] + dr * 3 * k[3])
            ias += 1
    sys.lmaxvr = lmaxvr
    sys.veffmt = veffmt
    sys.lofr = lofr
    return sys
