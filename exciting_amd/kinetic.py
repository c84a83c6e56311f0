"""Kinetic energy of the first-variational states, kinetic_energy.f90:1-230
('none' and ZORA relativity; the reference stops for iora*), on the GPU.

Per k-point and state x (evecfv(:, ist)) the reference adds
  * a muffin-tin part per atom: zm = A^T x (A = apwalm of the atom,
    kinetic_energy.f90:122-135), then zm^H (t_aa zm + t_alo x_lo) +
    x_lo^H (t_alo^T zm + t_lolo x_lo) with the kinetic radial integrals
    t_aa, t_alo, t_lolo (:94-121, :136-183; diagonal in (l, m)).  As in the
    reference, t_alo and t_lolo carry the ZORA mass factor whatever the
    relativity setting, t_aa only for ZORA;
  * an interstitial part (:208-234): for each Cartesian direction, the
    backward FFT of x_G (G+k)_d placed on the grid, times cfunir (divided by
    1 - alpha^2/2 veffir for ZORA), forward FFT, gathered at the G+k and
    weighted by (G+k)_d; 0.5 x^H of that.

Here the integrals come from libexh (exh_mt_kinetic_device, the same
kernel as hmlint's kinetic block), zm is one library GEMM per atom batch,
the per-atom quadratic forms are batched matrix products, and the
interstitial part is one batched 3-D FFT pair over all states per direction
(torch.fft, i.e. hipFFT), all in FP64 with the states resident in HBM.
"""
from __future__ import annotations

import numpy as np

ALPHA = 1.0 / 137.03599911


def _grid_index(ivg_cols, ngrid):
    """genigfft (gengvec.f90:129-167), 0-based, for integer G columns."""
    n = np.asarray(ngrid, dtype=np.int64)[:, None]
    iv = np.asarray(ivg_cols, dtype=np.int64)
    j = np.where(iv >= 0, iv, n + iv)
    return j[2] * n[0, 0] * n[1, 0] + j[1] * n[0, 0] + j[0]


class KineticPlan:
    """Index maps of one structure: per atom, the rows (lm, io) of zm and the
    LO rows, and the kinetic integral each pair of them takes."""

    def __init__(self, s):
        p, lmaxmat, lmaxapw = s.apwordmax, s.lmaxmat, s.lmaxapw
        lmm = (lmaxapw + 1) ** 2
        nfa = (lmaxmat + 1) * p
        self.lmm, self.p = lmm, p
        self.maxlo = int(max(s.losize)) if len(s.losize) else 0
        dim = lmm * p + self.maxlo
        self.dim = dim
        # per species: f-index of every row of u = [zm (lm, io) ; x_lo (ilo, m)],
        # the row's (l, m), or -1
        self.rowf, self.rowlm = [], []
        for is_ in range(s.nspecies):
            f = np.full(dim, -1, np.int64)
            lmr = np.full(dim, -1, np.int64)
            for l in range(lmaxmat + 1):
                for m in range(-l, l + 1):
                    lm = l * (l + 1) + m
                    for io in range(s.apword[l, is_]):
                        f[lm * p + io] = l * p + io
                        lmr[lm * p + io] = lm
            r = lmm * p
            for ilo in range(s.nlorb[is_]):
                l = s.lorbl[ilo, is_]
                for m in range(-l, l + 1):
                    f[r] = nfa + ilo
                    lmr[r] = l * (l + 1) + m
                    r += 1
            self.rowf.append(f)
            self.rowlm.append(lmr)
        self.species = np.repeat(np.arange(s.nspecies), s.natoms)
        self.lobase = np.concatenate([[0], np.cumsum([s.losize[i] for i in self.species])])[:-1]


def kinetic_energy(ctx, d, evec, cfunir, lofr, valence_rel_none=True, veffmt=None, lmaxvr=0,
                   veffir=None, plan=None):
    """kinetic_energy.f90 for one k-point: engyknst(1:nstfv) as a float64
    CUDA tensor.  d: lapw.DeviceKPoint whose apwalm holds match's result;
    evec: (nstfv, nmat) complex128 CUDA tensor, row = state (the memory of
    evecfv(nmat, nstfv)); cfunir, veffir, lofr, veffmt: CUDA tensors in the
    reference's layouts."""
    import torch
    s = ctx.sys
    if plan is None:
        plan = KineticPlan(s)
    dev = evec.device
    ngp, n = d.ngk, d.rank
    nst = evec.shape[0]
    natm = s.natmtot
    # kinetic_energy.f90 applies the ZORA mass factor to t_alo and t_lolo even
    # for 'none' relativity (:139-167), to t_aa only for ZORA (:94-121)
    from . import lapw
    mass = lapw.KIN_MASS_LO | (0 if valence_rel_none else lapw.KIN_MASS_APW)
    T = ctx.mt_kinetic(lofr, mass, veffmt, lmaxvr)                     # (natm, nf, nf)
    X = evec[:, :ngp].transpose(0, 1).contiguous()                     # (ngp, nst)
    # muffin-tin part
    p, lmm = plan.p, plan.lmm
    A = d.apwalm.view(natm, lmm * p, s.ngkmax)[:, :, :ngp]             # rows (lm, io)
    zm = torch.matmul(A, X)                                            # (natm, lmm p, nst)
    u = torch.zeros((natm, plan.dim, nst), dtype=torch.complex128, device=dev)
    u[:, :lmm * p] = zm
    xlo = evec[:, ngp:].transpose(0, 1)                                # (nlotot, nst)
    M = torch.zeros((natm, plan.dim, plan.dim), dtype=torch.float64, device=dev)
    for ias in range(natm):
        is_ = plan.species[ias]
        los = int(s.losize[is_])
        if los:
            b = int(plan.lobase[ias])
            u[ias, lmm * p:lmm * p + los] = xlo[b:b + los]
        f = torch.as_tensor(plan.rowf[is_], device=dev)
        lmr = torch.as_tensor(plan.rowlm[is_], device=dev)
        ok = f >= 0
        same = ok[:, None] & ok[None, :] & (lmr[:, None] == lmr[None, :])
        fi = f.clamp(min=0)
        M[ias] = torch.where(same, T[ias][fi[:, None], fi[None, :]], torch.zeros((), device=dev,
                                                                               dtype=torch.float64))
    Mu = torch.matmul(M.to(torch.complex128), u)
    e_mt = (u.conj() * Mu).sum(dim=(0, 1)).real
    # interstitial part
    ngrid = tuple(int(x) for x in s.intgv[:, 1] - s.intgv[:, 0] + 1)
    igp = d.igpig.cpu().numpy()[:ngp].astype(np.int64) - 1
    grid_idx = torch.as_tensor(_grid_index(s.ivg[:, igp], ngrid), device=dev)
    cfun = cfunir
    if not valence_rel_none:
        cfun = cfunir / (1.0 - 0.5 * ALPHA * ALPHA * veffir)
    cf = cfun.view(ngrid[2], ngrid[1], ngrid[0])
    vg = d.vgpc.view(-1, 3)[:ngp]                                      # (ngp, 3)
    ntot = ngrid[0] * ngrid[1] * ngrid[2]
    e_ist = torch.zeros(nst, dtype=torch.float64, device=dev)
    for dxyz in range(3):
        z = torch.zeros((nst, ntot), dtype=torch.complex128, device=dev)
        z[:, grid_idx] = (X * vg[:, dxyz:dxyz + 1]).transpose(0, 1)
        z = z.view(nst, ngrid[2], ngrid[1], ngrid[0])
        z = torch.fft.ifftn(z, dim=(1, 2, 3), norm="forward")         # unscaled backward
        z = z * cf
        z = torch.fft.fftn(z, dim=(1, 2, 3), norm="forward")          # forward with 1/N
        y = z.reshape(nst, ntot)[:, grid_idx] * vg[:, dxyz]
        e_ist += 0.5 * (evec[:, :ngp].conj() * y).sum(dim=1).real
    return e_mt + e_ist
