"""Host-side mirror of exciting's H/O-setup interface over the libexh C ABI.

The reference entry points and their Python counterparts:

  match(ngp, gpc, tpgpc, sfacgp, apwalm)                src/match.f90:13
      -> Context.match(ngp, gpc, tpgpc, sfacgp, apwalm)
  hamiltonandoverlapsetup(system, ngp, apwalm, igpig, vgpc)
      src/src_eigensystem/hamiltonandoverlapsetup.f90:8 (semantics of
      hamiltonsetup.f90 + overlapsetup.f90, called at seceqnfv.f90:76-77)
      -> Context.hamiltonandoverlapsetup(system, ngp, apwalm, igpig, vgpc)
  newsystem(system, packed, rank)        src/src_eigensystem/modfvsystem.f90:92
      -> newsystem(rank, device=...)

Arrays are accepted either as NumPy arrays in the reference's column-major
(Fortran) layout, which go through the host-pointer entry points, or as
torch CUDA tensors, which go through the device-resident entry points
without host copies.  A torch tensor holding a Fortran array ``a(n1,...,nk)``
has shape ``(nk, ..., n1)`` in row-major order; ``fortran_view`` converts.

Errors raise ``ExhError`` carrying the library's message (the reference
stops the program; the C ABI returns a code).  There is no CPU fallback:
without a built libexh.so every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import synth

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libexh.so")

ACCUMULATE = 1
Y00 = 0.28209479177387814       # 1 / sqrt(4 pi)
# ge alpha / 4 of hmlint.f90:45, with physical_constants.f90:53,56
GA4 = 2.00231930436256 * (1.0 / 137.035999084) / 4.0

_p = C.c_void_p
_i = C.c_int
_d = C.c_double


class ExhError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libexh.so (built by ``python -m exciting_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ExhError(f"{LIB_PATH} is not built; run `python -m exciting_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.exh_create.argtypes = [_i, C.POINTER(_p)]
    L.exh_destroy.argtypes = [_p]
    L.exh_destroy.restype = None
    L.exh_error.argtypes = [_p]
    L.exh_error.restype = C.c_char_p
    L.exh_stream.argtypes = [_p]
    L.exh_stream.restype = _p
    L.exh_set_basis.argtypes = [_p, _i, _p, _i, _i, _i, _p, _i, _p, _p, _i]
    L.exh_set_radial.argtypes = [_p, _d, _p, _p, _i, _p, _p]
    L.exh_set_mt.argtypes = [_p, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p]
    L.exh_set_gvectors.argtypes = [_p, _i, _i, _p, _p, _p]
    L.exh_set_interstitial.argtypes = [_p, _p, _p, _p, _p, _i, _i]
    L.exh_match.argtypes = [_p, _i, _p, _p, _p, _p]
    L.exh_match_device.argtypes = [_p, _p, _i, _p, _p, _p, _p]
    L.exh_hamiltonandoverlapsetup.argtypes = [_p, _i, _i, _p, _p, _p, _p, _p]
    L.exh_hamiltonandoverlapsetup_ex.argtypes = [_p, _i, _i, _p, _p, _p, _p, _p, _i]
    L.exh_setup_device.argtypes = [_p, _p, _i, _i, _p, _p, _p, _p, _p, _i]
    L.exh_setup_device_noalloc.argtypes = [_p, _p, _i, _i, _p, _p, _p, _p, _p, _i]
    L.exh_set_apwalm_reuse.argtypes = [_p, _i]
    _ll = C.c_longlong
    L.exh_zgemm_cn_device.argtypes = [_p, _i, _i, _i, _i, _p, _ll, _p, _ll, _d, _d, _p, _ll,
                                      _d, _p, _ll, _p, _ll, _i, _p]
    L.exh_apwalm_transfers.argtypes = [_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.exh_profile_enable.argtypes = [_p, _i]
    L.exh_profile_read.argtypes = [_p, _i, C.POINTER(_d), C.POINTER(_i)]
    L.exh_profile_reset.argtypes = [_p]
    L.exh_setup_batch_device.argtypes = [_p, _p, _i, _p, _i]
    L.exh_setup_multi_device.argtypes = [C.POINTER(_p), _p, _i, _p, _i]
    L.exh_batch_plan_count.argtypes = []
    L.exh_batch_plan_info.argtypes = [_i] + [C.POINTER(_i)] * 5
    L.exh_set_batch_plan.argtypes = [_p, _i]
    L.exh_get_batch_plan.argtypes = [_p]
    L.exh_tune_batch.argtypes = [_p, _p, _i, _p, _i, C.POINTER(_i)]
    L.exh_set_batch_workspace.argtypes = [_p, C.c_size_t]
    L.exh_set_interstitial_device.argtypes = [_p, _p, _p, _p, _p, _p, _i, _i]
    L.exh_gencfun.argtypes = [_p, _p, _p, _i, _d, _d, _d, _p, _p]
    L.exh_gencfun_device.argtypes = [_p, _p, _p, _p, _i, _d, _d, _d, _p, _p]
    L.exh_genveffig.argtypes = [_p, _p, _p, _p]
    L.exh_genmeffig.argtypes = [_p, _p, _p, _i, _i, _p, _p]
    L.exh_genveffig_device.argtypes = [_p, _p, _p, _p, _i, _p, _p, _p]
    L.exh_set_mt_device.argtypes = [_p, _p, _i, _i] + [_p] * 8
    L.exh_hmlint.argtypes = [_p, _i, _p, _p, _i, _p, _p, _p, _i]
    L.exh_hmlint_device.argtypes = [_p, _p, _i, _p, _p, _i, _p, _p, _p, _i]
    L.exh_overlapsetup.argtypes = [_p, _i, _i, _p, _p, _p, _p, _p]
    L.exh_overlapsetup_device.argtypes = [_p, _p, _i, _i, _p, _p, _p, _p, _p, _i]
    L.exh_mt_kinetic_device.argtypes = [_p, _p, _i, _p, _p, _i, _p]
    L.exh_mt_pot.argtypes = [_p, _i, _p, _p, _p, _p, _p, _i]
    L.exh_mt_pot_device.argtypes = [_p, _p, _i, _p, _p, _p, _p, _p, _i]
    L.exh_mt_so.argtypes = [_p, _i, _p, _p, _i] + [_p] * 12 + [_i]
    L.exh_mt_so_device.argtypes = [_p, _p, _i, _p, _p, _i] + [_p] * 12 + [_i]
    _lib = L
    return L


class KPoint(C.Structure):
    """exh_kpoint: one k-point of a batch (device pointers)."""
    _fields_ = [("ngp", _i), ("rank", _i), ("gpc", _p), ("tpgpc", _p), ("sfacgp", _p),
                ("apwalm", _p), ("igpig", _p), ("vgpc", _p), ("h", _p), ("o", _p)]


EXPORTS = (
    "exh_create", "exh_destroy", "exh_error", "exh_stream", "exh_set_basis",
    "exh_set_radial", "exh_set_mt", "exh_set_gvectors", "exh_set_interstitial",
    "exh_match", "exh_match_device", "exh_hamiltonandoverlapsetup",
    "exh_hamiltonandoverlapsetup_ex", "exh_setup_device", "exh_profile_enable",
    "exh_profile_read", "exh_profile_reset", "exh_setup_batch_device",
    "exh_batch_plan_count", "exh_batch_plan_info", "exh_set_batch_plan",
    "exh_get_batch_plan", "exh_tune_batch", "exh_set_batch_workspace",
    "exh_set_interstitial_device", "exh_gencfun", "exh_gencfun_device", "exh_genveffig",
    "exh_genmeffig", "exh_genveffig_device", "exh_set_mt_device", "exh_hmlint",
    "exh_hmlint_device", "exh_overlapsetup", "exh_overlapsetup_device",
    "exh_mt_kinetic_device", "exh_mt_pot", "exh_mt_pot_device", "exh_mt_so",
    "exh_mt_so_device", "exh_setup_device_noalloc", "exh_set_apwalm_reuse",
    "exh_apwalm_transfers", "exh_setup_multi_device", "exh_zgemm_cn_device",
)

KERNELS = ("match", "ivgk", "bt", "a2t", "ho", "lo_border", "lolo", "pack", "oz_residues",
           "oz_gemm", "oz_combine", "ho_grouped", "batch_graph", "cfunig", "ifft",
           "iproducts", "igather", "hml_radial", "hml_contract", "mt_so")

MATCH = 2   # batch: run match first
GRAPH = 4   # batch: capture in a HIP graph, replay on identical calls
VEFFIG, MEFFIG, M2EFFIG = 1, 2, 4   # exh_genveffig_device outputs
KIN_MASS_APW, KIN_MASS_LO = 1, 2     # exh_mt_kinetic_device mass factors


def plan_count() -> int:
    return lib().exh_batch_plan_count()


def plan_info(plan: int) -> dict:
    """Tile rows, columns, K slab, persistence and XCD ordering of a batch
    plan of the grouped H/O kernel."""
    v = [_i() for _ in range(5)]
    rc = lib().exh_batch_plan_info(plan, *[C.byref(x) for x in v])
    if rc != 0:
        raise ExhError(f"no batch plan {plan}")
    return dict(zip(("tile_rows", "tile_cols", "k_slab", "persistent", "xcd_order"),
                    (x.value for x in v)))


def SPLIT(nmod: int) -> int:
    """Flag selecting the split-precision (int8 Ozaki scheme II) muffin-tin
    contraction with nmod moduli (6..12)."""
    return nmod << 8


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _host(a, dtype):
    if a is None:
        return None
    return np.asfortranarray(a, dtype=dtype)


def _ptr(a):
    if a is None:
        return None
    if _is_torch(a):
        return a.data_ptr()
    return a.ctypes.data


def fortran_view(t, shape):
    """View a row-major torch tensor as the Fortran array of ``shape``:
    returns a tensor indexed [i1, ..., ik] (a permuted view, no copy)."""
    rev = t.reshape(tuple(reversed(shape)))
    return rev.permute(*reversed(range(len(shape))))


class HermitianMatrix:
    """modfvsystem.f90:15-26 (dense storage only: packed=False)."""

    def __init__(self, rank: int, za):
        self.rank = rank
        self.packed = False
        self.za = za


class EvSystem:
    """modfvsystem.f90:27-29: the Hamiltonian and overlap of one k-point."""

    def __init__(self, hamilton: HermitianMatrix, overlap: HermitianMatrix):
        self.hamilton = hamilton
        self.overlap = overlap


def newsystem(rank: int, device=None) -> EvSystem:
    """modfvsystem.f90:92-101: zero-filled H and O.  ``device`` None gives
    NumPy Fortran arrays, otherwise torch tensors on that device whose
    row-major storage is the column-major matrix (i.e. they hold H^T)."""
    if device is None:
        mk = lambda: np.zeros((rank, rank), dtype=np.complex128, order="F")
    else:
        import torch
        mk = lambda: torch.zeros((rank, rank), dtype=torch.complex128, device=device)
    return EvSystem(HermitianMatrix(rank, mk()), HermitianMatrix(rank, mk()))


class Context:
    """One libexh context (one GPU, one stream) holding the module state."""

    def __init__(self, device: int = 0):
        h = _p()
        rc = lib().exh_create(device, C.byref(h))
        if rc != 0:
            raise ExhError(f"exh_create failed with code {rc}")
        self._h = h
        self.device = device
        self.sys: Optional[synth.LapwSystem] = None
        self.nlotot = 0
        self.ngkmax = 0
        self.natoms: list = []
        self.maxaa = self.maxnlo = 0
        self.ngrtot = self.ngvec = 0

    def close(self):
        self._tstream = None
        if getattr(self, "_h", None):
            lib().exh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = lib().exh_error(self._h).decode(errors="replace")
            raise ExhError(f"{what}: code {rc}: {msg}")

    @property
    def stream(self) -> int:
        return lib().exh_stream(self._h)

    def _device_call(self, stream, launch):
        """Run launch(stream handle) for a _device entry point.  With an
        explicit stream the caller orders the work.  Without one the work
        runs on the context's stream, ordered after everything already queued
        on PyTorch's current stream (which produced the inputs) and before
        anything queued there afterwards (which reads the outputs)."""
        if stream is not None:
            return launch(stream)
        import torch
        if getattr(self, "_tstream", None) is None:
            self._tstream = torch.cuda.ExternalStream(self.stream,
                                                      device=f"cuda:{self.device}")
        cur = torch.cuda.current_stream(self.device)
        self._tstream.wait_stream(cur)
        rc = launch(self.stream)
        cur.wait_stream(self._tstream)
        return rc

    # ---- host-array residency -------------------------------------------
    def set_apwalm_reuse(self, on: bool = True):
        """Let host-array setups reuse the apwalm the last host match left in
        HBM when called with the same array (exh_set_apwalm_reuse)."""
        self._check(lib().exh_set_apwalm_reuse(self._h, 1 if on else 0), "exh_set_apwalm_reuse")

    def apwalm_transfers(self):
        """(uploads, reused): host apwalm uploads made and skipped so far."""
        u, r = C.c_longlong(), C.c_longlong()
        self._check(lib().exh_apwalm_transfers(self._h, C.byref(u), C.byref(r)),
                    "exh_apwalm_transfers")
        return u.value, r.value

    # ---- instrumentation ------------------------------------------------
    def profile(self, on: bool = True):
        self._check(lib().exh_profile_enable(self._h, 1 if on else 0), "exh_profile_enable")

    def profile_reset(self):
        self._check(lib().exh_profile_reset(self._h), "exh_profile_reset")

    def profile_read(self):
        """{kernel name: (total device ms, launches)} since the last reset."""
        out = {}
        for k, name in enumerate(KERNELS):
            ms, n = _d(), _i()
            self._check(lib().exh_profile_read(self._h, k, C.byref(ms), C.byref(n)),
                        "exh_profile_read")
            out[name] = (ms.value, n.value)
        return out

    # ---- module state ---------------------------------------------------
    def set_basis(self, natoms, lmaxapw, lmaxmat, apwordmax, apword, nlomax, nlorb,
                  lorbl, ngkmax):
        natoms = _host(natoms, np.int32)
        apword = _host(apword, np.int32)
        nlorb = _host(nlorb, np.int32)
        lorbl = _host(lorbl, np.int32)
        self._check(lib().exh_set_basis(self._h, len(natoms), _ptr(natoms), lmaxapw, lmaxmat,
                                        apwordmax, _ptr(apword), nlomax, _ptr(nlorb),
                                        _ptr(lorbl), ngkmax), "exh_set_basis")
        self.ngkmax = ngkmax
        self.natoms = [int(x) for x in natoms]
        # maxaa (MaxAPWs) and maxnlo (largest losize) as the library derives them
        ap = np.asarray(apword).reshape((lmaxapw + 1, len(self.natoms)), order="F")
        lb = np.asarray(lorbl).reshape((max(nlomax, 1), len(self.natoms)), order="F")
        self.maxaa = max(sum((2 * l + 1) * int(ap[l, i]) for l in range(lmaxmat + 1))
                         for i in range(len(self.natoms)))
        self.maxnlo = max([sum(2 * int(lb[j, i]) + 1 for j in range(int(nlorb[i])))
                           for i in range(len(self.natoms))] + [0])

    def set_radial(self, omega, rmt, nrmt, nrmtmax, spr, apwfr):
        rmt, nrmt = _host(rmt, np.float64), _host(nrmt, np.int32)
        spr, apwfr = _host(spr, np.float64), _host(apwfr, np.float64)
        self._check(lib().exh_set_radial(self._h, omega, _ptr(rmt), _ptr(nrmt), nrmtmax,
                                         _ptr(spr), _ptr(apwfr)), "exh_set_radial")

    def set_mt(self, maxaa, maxnlo, aa, loa, lolo, oalo, ololo, h1aa=None, h1loa=None,
               h1lolo=None, stream=None):
        """Muffin-tin tables from host arrays, or from CUDA tensors through
        exh_set_mt_device (no host copy)."""
        if _is_torch(aa):
            rc = self._device_call(stream, lambda st: lib().exh_set_mt_device(
                self._h, st, maxaa, maxnlo, *[_ptr(a) for a in (aa, loa, lolo, oalo, ololo,
                                                                  h1aa, h1loa, h1lolo)]))
            self._check(rc, "exh_set_mt_device")
            return
        arrs = [_host(aa, np.complex128), _host(loa, np.complex128),
                _host(lolo, np.complex128), _host(oalo, np.float64), _host(ololo, np.float64),
                _host(h1aa, np.float64), _host(h1loa, np.float64), _host(h1lolo, np.float64)]
        self._check(lib().exh_set_mt(self._h, maxaa, maxnlo, *[_ptr(a) for a in arrs]),
                    "exh_set_mt")

    def set_gvectors(self, ngrtot, ngvec, ivg, intgv, ivgig):
        ivg, intgv, ivgig = _host(ivg, np.int32), _host(intgv, np.int32), _host(ivgig, np.int32)
        self._check(lib().exh_set_gvectors(self._h, ngrtot, ngvec, _ptr(ivg), _ptr(intgv),
                                           _ptr(ivgig)), "exh_set_gvectors")
        self.ngrtot, self.ngvec = int(ngrtot), int(ngvec)

    def set_interstitial(self, cfunig, veffig, meffig=None, m2effig=None,
                         valence_rel_none=True, h1on=False, stream=None):
        """Interstitial tables from host arrays, or from CUDA tensors through
        exh_set_interstitial_device (no host copy)."""
        if _is_torch(cfunig):
            rc = self._device_call(stream, lambda st: lib().exh_set_interstitial_device(
                self._h, st, _ptr(cfunig), _ptr(veffig), _ptr(meffig), _ptr(m2effig),
                1 if valence_rel_none else 0, 1 if h1on else 0))
            self._check(rc, "exh_set_interstitial_device")
            return
        a = [_host(x, np.complex128) for x in (cfunig, veffig, meffig, m2effig)]
        self._check(lib().exh_set_interstitial(self._h, *[_ptr(x) for x in a],
                                               1 if valence_rel_none else 0,
                                               1 if h1on else 0), "exh_set_interstitial")

    # ---- muffin-tin table producer --------------------------------------
    def hmlint(self, lmaxvr, veffmt, lofr, valence_rel_none=True, aa=None, loa=None, lolo=None,
               flags: int = 0, stream=None):
        """hmlint.f90:53-55 (mt_kin + mt_pot): the muffin-tin tables
        (aa, loa, lolo).  With CUDA tensors (veffmt, lofr and the three
        outputs, Fortran layouts flattened) the device entry point fills the
        outputs; otherwise NumPy arrays are returned (accumulated into the
        given ones with flags=ACCUMULATE)."""
        return self._mt_tables(True, lmaxvr, veffmt, lofr, valence_rel_none, aa, loa, lolo,
                               flags, stream)

    def mt_pot(self, lmaxvr, pot, lofr, aa=None, loa=None, lolo=None, flags: int = 0,
               stream=None):
        """mt_pot.f90:14-323 alone: the same tables from any real muffin-tin
        function pot (lmmaxvr, nrmtmax, natmtot), without the kinetic part
        (the spin-polarised blocks of hmlint.f90:58-173).  Same calling
        convention as hmlint."""
        return self._mt_tables(False, lmaxvr, pot, lofr, True, aa, loa, lolo, flags, stream)

    def hmlint_spin(self, lmaxvr, bxcmt, lofr, ncmag=False, bext=None):
        """The spin-polarised blocks of hmlint (hmlint.f90:58-173, without
        mt_so) from CUDA float64 tensors: bxcmt(lmmaxvr, nrmtmax, natmtot,
        ndmag) flattened in Fortran order (ndmag 1 collinear, 3 non-collinear),
        lofr as for hmlint, and bext (natmtot, 3) = bfcmt of the atom +
        bfieldc, the external fields added to the l = 0 channel as
        ga4 b / y00 (ga4 = ge alpha / 4).  Returns {"alpha", "beta"[, "ab"]},
        each a tuple (aa, loa, lolo) of CUDA complex128 tensors with the
        shapes of the spinless tables:
          collinear:      alpha = T(bxc_1 + f_z), beta = -alpha;
          non-collinear:  alpha = T(bxc_3 + f_z), beta = -alpha,
                          ab = -i T(bxc_2 + f_y) + T(bxc_1 + f_x),
        T being mt_pot.  bxcmt is not modified."""
        import torch
        na, nr = sum(self.natoms), self.sys.nrmtmax
        lmm = (lmaxvr + 1) ** 2
        ndmag = 3 if ncmag else 1
        plane = lmm * nr * na
        if not _is_torch(bxcmt) or bxcmt.numel() != plane * ndmag \
                or bxcmt.dtype != torch.float64:
            raise ExhError(f"bxcmt must be a CUDA float64 tensor of {plane * ndmag} elements "
                           f"(lmmaxvr {lmm}, nrmtmax {nr}, natmtot {na}, ndmag {ndmag})")
        b = np.zeros((na, 3)) if bext is None else np.asarray(bext, dtype=np.float64)
        if b.shape != (na, 3):
            raise ExhError(f"bext must be ({na}, 3)")
        mnlo = max(self.maxnlo, 1)
        sizes = (self.maxaa * self.maxaa * na, mnlo * self.maxaa * na, mnlo * mnlo * na)
        dev = bxcmt.device

        def component(c, axis):
            v = bxcmt[c * plane:(c + 1) * plane].clone()
            shift = torch.from_numpy(GA4 * b[:, axis] / Y00).to(dev)
            v.view(na, nr, lmm)[:, :, 0] += shift[:, None]
            return v

        def tables():
            # mt_pot with flags 0 writes every element of aa (and of loa and
            # lolo when there are LOs; without them those stay as allocated)
            new = torch.empty if self.maxnlo > 0 else torch.zeros
            return tuple(new(n, dtype=torch.complex128, device=dev) for n in sizes)

        alpha = tables()
        self.mt_pot(lmaxvr, component(2 if ncmag else 0, 2), lofr, *alpha, flags=0)
        out = {"alpha": alpha, "beta": tuple(-t for t in alpha)}
        if ncmag:
            ab = tables()
            self.mt_pot(lmaxvr, component(1, 1), lofr, *ab, flags=0)
            for t in ab:
                t.mul_(-1j)
            self.mt_pot(lmaxvr, component(0, 0), lofr, *ab, flags=ACCUMULATE)
            out["ab"] = ab
        return out

    def mt_so(self, lmaxvr, pot, lofr, tables=None, rel_level=1, flags: int = ACCUMULATE,
              stream=None):
        """mt_so.f90:11-397: the spin-orbit terms added to the alpha, beta and
        ab blocks, tables = {"alpha": (aa, loa, lolo, alo), "beta": ...,
        "ab": ...}.  With CUDA tensors (pot, lofr and every table, Fortran
        layouts flattened) the device entry point updates the tables in
        place; otherwise NumPy arrays are returned (the given ones updated,
        missing ones created as zeros).  rel_level 1 is level_zora, the
        level hmlint.f90:172 uses.  flags=0 zeroes the tables first."""
        keys = ("alpha", "beta", "ab")
        if _is_torch(pot):
            ptrs = [_ptr(x) for k in keys for x in tables[k]]
            rc = self._device_call(stream, lambda st: lib().exh_mt_so_device(
                self._h, st, lmaxvr, _ptr(pot), _ptr(lofr), rel_level, *ptrs, flags))
            self._check(rc, "exh_mt_so_device")
            return tables
        na, mnlo, maxaa = sum(self.natoms), max(self.maxnlo, 1), self.maxaa
        shapes = ((maxaa, maxaa, na), (mnlo, maxaa, na), (mnlo, mnlo, na), (maxaa, mnlo, na))
        out = {}
        for k in keys:
            given = None if tables is None else tables.get(k)
            arrs = []
            for i, shape in enumerate(shapes):
                x = None if given is None else given[i]
                if x is None:
                    x = np.zeros(shape, dtype=np.complex128, order="F")
                elif not (isinstance(x, np.ndarray) and x.flags.f_contiguous
                          and x.dtype == np.complex128 and x.shape == shape):
                    raise ExhError(f"mt_so {k} tables must be Fortran-ordered complex128 "
                                   f"arrays of shapes {shapes}")
                arrs.append(x)
            out[k] = tuple(arrs)
        v = _host(pot, np.float64)
        lf = _host(lofr, np.float64) if lofr is not None else None
        self._check(lib().exh_mt_so(self._h, lmaxvr, _ptr(v), _ptr(lf), rel_level,
                                    *[_ptr(x) for k in keys for x in out[k]], flags),
                    "exh_mt_so")
        return out

    def _mt_tables(self, kin, lmaxvr, veffmt, lofr, valence_rel_none, aa, loa, lolo, flags,
                   stream):
        name = "exh_hmlint" if kin else "exh_mt_pot"
        if _is_torch(veffmt):
            if kin:
                fn = lambda st: lib().exh_hmlint_device(
                    self._h, st, lmaxvr, _ptr(veffmt), _ptr(lofr),
                    1 if valence_rel_none else 0, _ptr(aa), _ptr(loa), _ptr(lolo), flags)
            else:
                fn = lambda st: lib().exh_mt_pot_device(
                    self._h, st, lmaxvr, _ptr(veffmt), _ptr(lofr), _ptr(aa), _ptr(loa),
                    _ptr(lolo), flags)
            self._check(self._device_call(stream, fn), name + "_device")
            return aa, loa, lolo
        v = _host(veffmt, np.float64)
        lf = _host(lofr, np.float64) if lofr is not None else None
        na = sum(self.natoms)
        maxaa, mnlo = self.maxaa, max(self.maxnlo, 1)
        outs = []
        for x, shape in ((aa, (maxaa, maxaa, na)), (loa, (mnlo, maxaa, na)),
                         (lolo, (mnlo, mnlo, na))):
            if x is None:
                x = np.zeros(shape, dtype=np.complex128, order="F")
            elif not (isinstance(x, np.ndarray) and x.flags.f_contiguous
                      and x.dtype == np.complex128):
                raise ExhError("hmlint tables must be Fortran-ordered complex128 arrays")
            outs.append(x)
        if kin:
            rc = lib().exh_hmlint(self._h, lmaxvr, _ptr(v), _ptr(lf),
                                  1 if valence_rel_none else 0, *[_ptr(x) for x in outs], flags)
        else:
            rc = lib().exh_mt_pot(self._h, lmaxvr, _ptr(v), _ptr(lf),
                                  *[_ptr(x) for x in outs], flags)
        self._check(rc, name)
        return tuple(outs)

    def mt_kinetic(self, lofr, mass_flags=0, veffmt=None, lmaxvr=0, stream=None):
        """The kinetic radial integrals of every atom (exh_mt_kinetic_device):
        a CUDA float64 tensor (natmtot, nf, nf), nf = (lmaxmat+1)*apwordmax
        + nlomax, from CUDA tensors lofr (and veffmt with a mass flag,
        KIN_MASS_APW | KIN_MASS_LO)."""
        import torch
        nf = (self.sys.lmaxmat + 1) * self.sys.apwordmax + self.sys.nlomax
        out = torch.zeros((sum(self.natoms), nf, nf), dtype=torch.float64,
                          device=f"cuda:{self.device}")
        rc = self._device_call(stream, lambda st: lib().exh_mt_kinetic_device(
            self._h, st, lmaxvr, _ptr(veffmt), _ptr(lofr), mass_flags, _ptr(out)))
        self._check(rc, "exh_mt_kinetic_device")
        return out

    # ---- interstitial producers ----------------------------------------
    def _atposc(self, atposc):
        """(natmtot, 3) Cartesian positions (species-major) -> the reference's
        atposc(3, maxatoms, nspecies) layout and maxatoms."""
        a = np.asarray(atposc, dtype=np.float64).reshape(-1, 3)
        if a.shape[0] != sum(self.natoms):
            raise ExhError(f"atposc holds {a.shape[0]} atoms, the basis {sum(self.natoms)}")
        maxatoms = max(self.natoms) if self.natoms else 1
        out = np.zeros((3, maxatoms, len(self.natoms)), order="F")
        ias = 0
        for is_, na in enumerate(self.natoms):
            out[:, :na, is_] = a[ias:ias + na].T
            ias += na
        return out, maxatoms

    def gencfun(self, bvec, atposc, epslat=1e-6, cfdamp=0.0, gmaxvr=12.0, cfunig=None,
                cfunir=None, stream=None):
        """gencfun.f90:13-85: the smooth step function.  bvec (3,3) with the
        reciprocal vectors as columns, atposc (natmtot, 3).  With CUDA
        tensors cfunig (complex128, ngrtot) and cfunir (float64, ngrtot) the
        device entry point fills them on the GPU; otherwise NumPy arrays are
        returned.  Returns (cfunig, cfunir)."""
        bv = np.asfortranarray(bvec, dtype=np.float64)
        ap, maxatoms = self._atposc(atposc)
        if _is_torch(cfunig):
            rc = self._device_call(stream, lambda st: lib().exh_gencfun_device(
                self._h, st, _ptr(bv), _ptr(ap), maxatoms, epslat, cfdamp, gmaxvr,
                _ptr(cfunig), _ptr(cfunir)))
            self._check(rc, "exh_gencfun_device")
            return cfunig, cfunir
        cg = np.zeros(self.ngrtot, dtype=np.complex128)
        cr = np.zeros(self.ngrtot, dtype=np.float64)
        self._check(lib().exh_gencfun(self._h, _ptr(bv), _ptr(ap), maxatoms, epslat, cfdamp,
                                      gmaxvr, _ptr(cg), _ptr(cr)), "exh_gencfun")
        return cg, cr

    def genveffig(self, veffir, cfunir):
        """genveffig.f90:13-44 from host arrays: veffig(ngvec)."""
        v, c = _host(veffir, np.float64), _host(cfunir, np.float64)
        out = np.zeros(self.ngvec, dtype=np.complex128)
        self._check(lib().exh_genveffig(self._h, _ptr(v), _ptr(c), _ptr(out)), "exh_genveffig")
        return out

    def genmeffig(self, veffir, cfunir, valence_rel_none=False, iora_star=False):
        """genmeffig.f90:13-61 from host arrays: (meffig, m2effig); None for
        what the relativity setting does not produce."""
        if valence_rel_none:
            return None, None
        v, c = _host(veffir, np.float64), _host(cfunir, np.float64)
        m = np.zeros(self.ngvec, dtype=np.complex128)
        m2 = np.zeros(self.ngvec, dtype=np.complex128) if iora_star else None
        self._check(lib().exh_genmeffig(self._h, _ptr(v), _ptr(c), 0, 1 if iora_star else 0,
                                        _ptr(m), _ptr(m2)), "exh_genmeffig")
        return m, m2

    def genveffig_device(self, veffir, cfunir, veffig=None, meffig=None, m2effig=None,
                         stream=None):
        """Device genveffig/genmeffig: fills whichever of the CUDA tensors
        veffig, meffig, m2effig (complex128, ngvec) are given, from CUDA
        tensors veffir and cfunir (float64, ngrtot), in one batched FFT."""
        which = ((VEFFIG if veffig is not None else 0) | (MEFFIG if meffig is not None else 0)
                 | (M2EFFIG if m2effig is not None else 0))
        rc = self._device_call(stream, lambda st: lib().exh_genveffig_device(
            self._h, st, _ptr(veffir), _ptr(cfunir), which, _ptr(veffig), _ptr(meffig),
            _ptr(m2effig)))
        self._check(rc, "exh_genveffig_device")

    def load(self, s: synth.LapwSystem) -> "Context":
        """Install every module-state table of a LapwSystem."""
        self.set_basis(s.natoms, s.lmaxapw, s.lmaxmat, s.apwordmax, s.apword, s.nlomax,
                       s.nlorb, s.lorbl, s.ngkmax)
        self.set_radial(s.omega, s.rmt, s.nrmt, s.nrmtmax, s.spr, s.apwfr)
        self.set_mt(s.maxaa, s.maxnlo, s.aa, s.loa, s.lolo, s.oalo, s.ololo, s.h1aa,
                    s.h1loa, s.h1lolo)
        self.set_gvectors(s.ngrtot, s.ngvec, s.ivg, s.intgv, s.ivgig)
        self.set_interstitial(s.cfunig, s.veffig, s.meffig, s.m2effig, s.valence_rel_none,
                              s.h1on)
        self.sys = s
        self.nlotot = s.nlotot
        return self

    # ---- hot path -------------------------------------------------------
    def match(self, ngp, gpc, tpgpc, sfacgp, apwalm, stream=None):
        """match.f90:13-194: fills apwalm (NumPy Fortran array or CUDA tensor)."""
        if _is_torch(apwalm):
            rc = self._device_call(stream, lambda st: lib().exh_match_device(
                self._h, st, ngp, _ptr(gpc), _ptr(tpgpc), _ptr(sfacgp), _ptr(apwalm)))
            self._check(rc, "exh_match_device")
            return apwalm
        if not (isinstance(apwalm, np.ndarray) and apwalm.flags.f_contiguous
                and apwalm.dtype == np.complex128):
            raise ExhError("apwalm must be a Fortran-ordered complex128 array")
        gpc, tpgpc = _host(gpc, np.float64), _host(tpgpc, np.float64)
        sfacgp = _host(sfacgp, np.complex128)
        self._check(lib().exh_match(self._h, ngp, _ptr(gpc), _ptr(tpgpc), _ptr(sfacgp),
                                    _ptr(apwalm)), "exh_match")
        return apwalm

    def hamiltonandoverlapsetup(self, system: EvSystem, ngp, apwalm, igpig, vgpc,
                                flags: int = ACCUMULATE, stream=None):
        """hamiltonsetup + overlapsetup: accumulates into system%hamilton%za
        and system%overlap%za (pass flags=0 to overwrite)."""
        h, o = system.hamilton.za, system.overlap.za
        rank = system.hamilton.rank
        if _is_torch(h):
            rc = self._device_call(stream, lambda st: lib().exh_setup_device(
                self._h, st, ngp, rank, _ptr(apwalm), _ptr(igpig), _ptr(vgpc), _ptr(h), _ptr(o),
                flags))
            self._check(rc, "exh_setup_device")
            return system
        for m in (h, o):
            if not (isinstance(m, np.ndarray) and m.flags.f_contiguous
                    and m.dtype == np.complex128 and m.shape == (rank, rank)):
                raise ExhError("H and O must be Fortran-ordered complex128 rank x rank arrays")
        apwalm = _host(apwalm, np.complex128)
        igpig, vgpc = _host(igpig, np.int32), _host(vgpc, np.float64)
        self._check(lib().exh_hamiltonandoverlapsetup_ex(self._h, ngp, rank, _ptr(apwalm),
                                                         _ptr(igpig), _ptr(vgpc), _ptr(h),
                                                         _ptr(o), flags),
                    "exh_hamiltonandoverlapsetup")
        return system

    def overlapsetup(self, system: EvSystem, ngp, apwalm, igpig, vgpc,
                     flags: int = ACCUMULATE, stream=None):
        """overlapsetup.f90 alone: accumulates system%overlap%za and mirrors
        both matrices' upper triangles into their lower ones (the reference's
        hamiltonoverlapocopy_UL); H itself is not computed.  system.hamilton
        may be None."""
        h = system.hamilton.za if system.hamilton is not None else None
        o, rank = system.overlap.za, system.overlap.rank
        if _is_torch(o):
            rc = self._device_call(stream, lambda st: lib().exh_overlapsetup_device(
                self._h, st, ngp, rank, _ptr(apwalm), _ptr(igpig), _ptr(vgpc), _ptr(h), _ptr(o),
                flags))
            self._check(rc, "exh_overlapsetup_device")
            return system
        for m in (h, o):
            if m is not None and not (isinstance(m, np.ndarray) and m.flags.f_contiguous
                                      and m.dtype == np.complex128 and m.shape == (rank, rank)):
                raise ExhError("H and O must be Fortran-ordered complex128 rank x rank arrays")
        if flags != ACCUMULATE:
            raise ExhError("the host overlapsetup accumulates, as the reference does")
        apwalm = _host(apwalm, np.complex128)
        igpig, vgpc = _host(igpig, np.int32), _host(vgpc, np.float64)
        self._check(lib().exh_overlapsetup(self._h, ngp, rank, _ptr(apwalm), _ptr(igpig),
                                           _ptr(vgpc), _ptr(h), _ptr(o)), "exh_overlapsetup")
        return system

    # ---- batched k-points -------------------------------------------------
    @staticmethod
    def _kpoint_array(kpts):
        arr = (KPoint * len(kpts))(*kpts)
        return arr

    def setup_batch(self, kpts, flags: int = ACCUMULATE, stream=None):
        """Grouped match (with MATCH) + H/O setup over a list of KPoint
        structures (device pointers; see ``kpoint_struct``)."""
        arr = self._kpoint_array(kpts)
        self._check(self._device_call(stream, lambda st: lib().exh_setup_batch_device(
            self._h, st, len(kpts), arr, flags)), "exh_setup_batch_device")
        return arr

    def tune_batch(self, kpts, flags: int = 0, stream=None) -> int:
        """Time every batch plan on these k-points and keep the fastest.  H
        and O are scratch; on return they hold the non-accumulated result."""
        arr = self._kpoint_array(kpts)
        best = _i(-1)
        self._check(self._device_call(stream, lambda st: lib().exh_tune_batch(
            self._h, st, len(kpts), arr, flags, C.byref(best))), "exh_tune_batch")
        return best.value

    @property
    def batch_plan(self) -> int:
        return lib().exh_get_batch_plan(self._h)

    @batch_plan.setter
    def batch_plan(self, plan: int):
        self._check(lib().exh_set_batch_plan(self._h, plan), "exh_set_batch_plan")

    def set_batch_workspace(self, nbytes: int):
        self._check(lib().exh_set_batch_workspace(self._h, int(nbytes)),
                    "exh_set_batch_workspace")


def setup_multi(ctxs, kpts, flags: int = ACCUMULATE, stream=None):
    """One grouped match (with MATCH) + H/O setup over k-points of several
    structures: ``kpts[i]`` (a KPoint) belongs to ``ctxs[i]`` (a Context).
    Runs on ``stream``, or on the first context's stream ordered after
    PyTorch's current stream as the single-context calls do."""
    if len(ctxs) != len(kpts) or not ctxs:
        raise ValueError("setup_multi: one context per k-point, at least one")
    lead = ctxs[0]
    arr = (KPoint * len(kpts))(*kpts)
    handles = (_p * len(ctxs))(*[c._h for c in ctxs])
    lead._check(lead._device_call(stream, lambda st: lib().exh_setup_multi_device(
        handles, st, len(kpts), arr, flags)), "exh_setup_multi_device")
    return arr


def zgemm_cn(a, x, alpha: float = 1.0, beta: float = 0.0, y0=None, gamma: float = 0.0, z=None,
             split: int = 0, stream=None, hermitian: bool = False):
    """exh_zgemm_cn_device on complex128 CUDA tensors: returns
    alpha a^H x + beta y0 + gamma z (nout x m) for a (k x nout) and x (k x m).
    With hermitian=True, ``a`` is a Hermitian matrix C and the result is
    alpha C x + ...; C is read in place whether it is stored row-major or
    column-major (no copy of C).  ``split`` chunks of K (0: enough for about
    4096 output tiles x chunks, with chunks of at least 64 of K: the Gram
    matrices of a block, 56 x 56 over K = 10364, get 162 chunks)."""
    import torch
    if a.dtype != torch.complex128 or x.dtype != torch.complex128 or not a.is_cuda:
        raise ExhError("zgemm_cn: complex128 CUDA tensors expected")
    if any(t is not None and t.device != a.device for t in (x, y0, z)):
        raise ExhError("zgemm_cn: all tensors must be on the device of a")
    conj_a = 1
    if hermitian and not a.is_contiguous() and a.t().is_contiguous():
        a, conj_a = a.t(), 0      # row-major C^T: C x = (C^T)^T x
    a = a.contiguous()
    x = x.contiguous()
    k, nout = a.shape
    if x.shape[0] != k:
        raise ExhError("zgemm_cn: a and x disagree on k")
    m = x.shape[1]
    tiles = -(-nout // 64) * -(-m // 64)
    split = split or max(1, min(4096, -(-4096 // max(tiles, 1)), -(-k // 64)))
    out = torch.empty((nout, m), dtype=torch.complex128, device=a.device)
    work = torch.empty(split * nout * m, dtype=torch.complex128, device=a.device)
    for t in (y0, z):
        if t is not None and (t.shape != (nout, m) or t.dtype != torch.complex128):
            raise ExhError("zgemm_cn: y0 and z must be complex128 nout x m")
    y0 = y0.contiguous() if y0 is not None else None
    z = z.contiguous() if z is not None else None
    st = stream if stream is not None else torch.cuda.current_stream(a.device).cuda_stream
    rc = lib().exh_zgemm_cn_device(st, conj_a, nout, m, k, _ptr(a), nout, _ptr(x), m, alpha,
                                   beta, _ptr(y0) if y0 is not None else None, m, gamma,
                                   _ptr(z) if z is not None else None, m, _ptr(out), m, split,
                                   _ptr(work))
    if rc != 0:
        raise ExhError(f"exh_zgemm_cn_device failed ({rc})")
    return out


def kpoint_struct(d: "DeviceKPoint", h, o, apwalm=None) -> KPoint:
    """exh_kpoint for a DeviceKPoint and its H, O tensors."""
    a = d.apwalm if apwalm is None else apwalm
    return KPoint(d.ngk, d.rank, _ptr(d.gpc), _ptr(d.tpgpc), _ptr(d.sfacgp), _ptr(a),
                  _ptr(d.igpig), _ptr(d.vgpc), _ptr(h), _ptr(o))


class DeviceKPoint:
    """Per-k inputs of one k-point resident in HBM (torch tensors)."""

    def __init__(self, s: synth.LapwSystem, ik: int, device="cuda", with_apwalm: bool = True):
        import torch
        k = s.kpts[ik]
        self.ngk = k.ngk
        self.rank = k.ngk + s.nlotot
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(np.asarray(a, order="F").ravel(order="F"))).to(device=device, dtype=dt)
        self.gpc = t(k.gpc, torch.float64)
        self.tpgpc = t(k.tpgpc, torch.float64)
        self.sfacgp = t(k.sfacgp, torch.complex128)
        self.igpig = t(k.igpig, torch.int32)
        self.vgpc = t(k.vgpc, torch.float64)
        lmm = (s.lmaxapw + 1) ** 2
        self.apwalm = torch.zeros(s.natmtot * lmm * s.apwordmax * s.ngkmax,
                                  dtype=torch.complex128, device=device) if with_apwalm else None
        self.apwalm_shape = (s.ngkmax, s.apwordmax, lmm, s.natmtot)
